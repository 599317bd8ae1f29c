"""ctypes binding of oracle/pftcc_cpu.cpp (TEST INFRASTRUCTURE ONLY: tests and the CPU baseline).

cpu_search_fp32 restates the reference CPU path (prep_ref4corr + calc_corrs_over_k,
simple_polarft_corrcalc.f90:961-985, 1132-1162) in FP32 with OpenMP over particles;
gencorrs_direct64 is a brute-force FP64 oracle for larger parity cases.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libpftcc_cpu.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run make -C oracle")
        L = C.CDLL(LIB)
        P = C.c_void_p
        L.cpu_memoize.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P, C.c_int]
        L.cpu_search_fp32.argtypes = [C.c_int] * 6 + [P] * 6 + [P] * 5 + [C.c_int]
        L.cpu_gencorrs_direct64.argtypes = [C.c_int] * 5 + [P] * 4 + [C.c_int, P, C.c_int]
        L.cpu_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def memoize(ptcls, nthreads):
    """PA [np, nk, n/2+1], PB [np, nk, n] of memoize_fft (:864-877)."""
    ptcls = np.ascontiguousarray(ptcls, np.complex64)
    npt, nk, n = ptcls.shape
    pa = np.zeros((npt, nk, n // 2 + 1), np.complex64)
    pb = np.zeros((npt, nk, n), np.complex64)
    lib().cpu_memoize(npt, nk, n, _p(ptcls), _p(pa), _p(pb), nthreads)
    return pa, pb


def search_fp32(refs, ctf, sqsums, pa, pb, shmat, kcorr, nthreads, want_all=False):
    """Exhaustive greedy search of the reference algorithm; returns (cc, iref, ishift, irot[, cc_all])."""
    refs = np.ascontiguousarray(refs, np.complex64)
    nr, nk, n = refs.shape
    npt = pa.shape[0]
    ns = 1 if shmat is None else shmat.shape[0]
    cc = np.zeros(npt, np.float32)
    ir = np.zeros(npt, np.int32)
    ish = np.zeros(npt, np.int32)
    rot = np.zeros(npt, np.int32)
    allc = np.zeros((npt, nr, ns, 2 * n), np.float32) if want_all else None
    ctf = None if ctf is None else np.ascontiguousarray(ctf, np.float32)
    shm = None if shmat is None else np.ascontiguousarray(shmat, np.complex64)
    lib().cpu_search_fp32(nr, npt, ns, nk, kcorr, n, _p(refs), _p(ctf), _p(np.ascontiguousarray(sqsums, np.float32)),
                          _p(pa), _p(pb), _p(shm), _p(cc), _p(ir), _p(ish), _p(rot), _p(allc), nthreads)
    return (cc, ir, ish, rot, allc) if want_all else (cc, ir, ish, rot)


def gencorrs_direct64(refs, ptcls, ctf, shmat, kcorr, sq_norm_all, nthreads):
    """FP64 brute force: cc[nr, np, 2n] (refs x particles, one shift)."""
    refs = np.ascontiguousarray(refs, np.complex64)
    ptcls = np.ascontiguousarray(ptcls, np.complex64)
    nr, nk, n = refs.shape
    npt = ptcls.shape[0]
    out = np.zeros((nr, npt, 2 * n), np.float64)
    ctf = None if ctf is None else np.ascontiguousarray(ctf, np.float32)
    shm = None if shmat is None else np.ascontiguousarray(shmat, np.complex128)
    lib().cpu_gencorrs_direct64(nr, npt, nk, kcorr, n, _p(refs), _p(ptcls), _p(ctf), _p(shm), int(sq_norm_all),
                                _p(out), nthreads)
    return out
