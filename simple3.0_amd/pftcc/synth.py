"""Seeded synthetic workloads: Gaussian-blob phantom projections as polar Fourier transforms.

Generation runs on the GPU (libpftcc_synth.so, same HIP runtime as the correlation library) so
that benchmark inputs are created in HBM and handed to the library with the *_dev setters.
Parameters follow the benchmark configurations: phantom of 50 random 3-D Gaussians (sigma
U(3,8) px, centres uniform in a ball of radius 40 px, amplitudes U(0.5,1)); references are
projections along uniform random directions; particles are a random reference rotated by a
uniform in-plane angle, shifted, multiplied by |CTF| (300 kV, Cs 2.7 mm, amplitude contrast
0.1, astigmatism <= 0.1 um) and given white noise at the requested SNR.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import LIBDIR, PftccError

_SYNTH = None


def synth_lib():
    global _SYNTH
    if _SYNTH is None:
        path = os.path.join(LIBDIR, "libpftcc_synth.so")
        if not os.path.exists(path):
            raise PftccError(f"{path} missing: run python simple3.0_amd/build.py")
        lib = C.CDLL(path)
        lib.synth_refs.restype = C.c_int
        lib.synth_refs.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_void_p]
        lib.synth_ptcls.restype = C.c_int
        lib.synth_ptcls.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                    C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
        lib.synth_malloc.restype = C.c_void_p
        lib.synth_malloc.argtypes = [C.c_size_t]
        lib.synth_free.argtypes = [C.c_void_p]
        lib.synth_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        lib.synth_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        lib.synth_param_size.restype = C.c_int
        lib.synth_event_create.restype = C.c_void_p
        lib.synth_event_record.argtypes = [C.c_void_p, C.c_void_p]
        lib.synth_event_elapsed_ms.restype = C.c_float
        lib.synth_event_elapsed_ms.argtypes = [C.c_void_p, C.c_void_p]
        lib.synth_device_sync.restype = C.c_int
        lib.synth_copy.restype = C.c_int
        lib.synth_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        _SYNTH = lib
    return _SYNTH


class DeviceBuffer:
    """hipMalloc'd memory owned by the synth library (freed on close)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = synth_lib().synth_malloc(self.nbytes)
        if not self.ptr:
            raise PftccError(f"device allocation of {nbytes} bytes failed")

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        synth_lib().synth_h2d(b.ptr, a.ctypes.data, a.nbytes)
        return b

    def to_array(self, dtype, shape):
        out = np.empty(shape, dtype)
        synth_lib().synth_d2h(out.ctypes.data, self.ptr, out.nbytes)
        return out

    def close(self):
        if self.ptr:
            synth_lib().synth_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# particle parameter record, matches PtclParam in csrc/synth.hip (13 x 4 bytes)
PTCL_DTYPE = np.dtype([("ref", np.int32), ("psi", np.float32), ("sx", np.float32), ("sy", np.float32),
                       ("wl", np.float32), ("cs", np.float32), ("ampc", np.float32), ("dfx", np.float32),
                       ("dfy", np.float32), ("angast", np.float32), ("noise_sd", np.float32),
                       ("seed", np.uint32)])


@dataclass
class CtfP:
    smpd: float
    kv: float
    cs: float
    fraca: float
    dfx: float
    dfy: float
    angast: float
    phshift: float = 0.0
    phaseplate: bool = False


@dataclass
class Workload:
    box: int
    smpd: float
    kfromto: tuple
    pftsz: int
    nrefs: int
    nptcls: int
    snr: float
    seed: int
    dfrange: tuple = (1.0, 3.0)
    astig: float = 0.1
    max_shift: float = 5.0
    nblobs: int = 50
    uniform_dirs: bool = False
    # filled by generate
    blobs: np.ndarray = field(default=None, repr=False)
    uvs: np.ndarray = field(default=None, repr=False)
    prm: np.ndarray = field(default=None, repr=False)
    ctfparams: list = field(default=None, repr=False)

    @property
    def nk(self):
        return self.kfromto[1] - self.kfromto[0] + 1


def phantom(rng, nblobs=50, radius=40.0):
    """50 random 3-D Gaussians: centres uniform in a ball, sigma U(3,8), amplitude U(0.5,1)."""
    out = np.zeros((nblobs, 5), np.float32)
    for b in range(nblobs):
        while True:
            c = rng.uniform(-radius, radius, 3)
            if np.dot(c, c) <= radius * radius:
                break
        out[b, :3] = c
        out[b, 3] = rng.uniform(3.0, 8.0)
        out[b, 4] = rng.uniform(0.5, 1.0)
    return out


def spiral_directions(count):
    """Near-uniform directions on the sphere (golden-angle spiral): [count, 3]."""
    i = np.arange(count, dtype=np.float64) + 0.5
    z = 1.0 - 2.0 * i / count
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def projection_axes(rng, count, uniform=False):
    """Orthonormal (u, v) of planes perpendicular to the projection directions: [count, 6].

    Directions are uniform random, or an even spiral sampling of the sphere when uniform is set
    (4684 directions are ~3 degrees apart)."""
    if uniform:
        d = spiral_directions(count)
    else:
        d = rng.standard_normal((count, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    helper = np.where(np.abs(d[:, 2:3]) < 0.9, np.array([[0, 0, 1.0]]), np.array([[1.0, 0, 0]]))
    u = np.cross(d, helper)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    v = np.cross(d, u)
    return np.concatenate([u, v], axis=1).astype(np.float32)


def ctf_pixel_constants(p: CtfP):
    """The reference CTF model's constants in pixel units (simple_ctf.f90:54-81, 522-525)."""
    a = 1000.0 * p.kv
    wl = 12.26 / math.sqrt(a + 0.9784 * a * a / 1e6) / p.smpd
    return dict(wl=wl, cs=p.cs * 1e7 / p.smpd, ampc=math.atan(p.fraca / math.sqrt(1 - p.fraca ** 2)),
                dfx=p.dfx * 1e4 / p.smpd, dfy=p.dfy * 1e4 / p.smpd, angast=p.angast / 180.0 * math.pi)


def make_params(w: Workload):
    rng = np.random.default_rng(w.seed)
    w.blobs = phantom(rng, w.nblobs)
    w.uvs = projection_axes(rng, w.nrefs, w.uniform_dirs)
    prm = np.zeros(w.nptcls, PTCL_DTYPE)
    prm["ref"] = rng.integers(0, w.nrefs, w.nptcls)
    prm["psi"] = rng.uniform(0, 2 * np.pi, w.nptcls)
    prm["sx"] = rng.uniform(-w.max_shift, w.max_shift, w.nptcls)
    prm["sy"] = rng.uniform(-w.max_shift, w.max_shift, w.nptcls)
    df = rng.uniform(*w.dfrange, w.nptcls)
    ast = rng.uniform(0, w.astig, w.nptcls)
    ang = rng.uniform(0, 180.0, w.nptcls)
    cps = []
    for i in range(w.nptcls):
        p = CtfP(w.smpd, 300.0, 2.7, 0.1, float(df[i] + ast[i] / 2), float(df[i] - ast[i] / 2), float(ang[i]))
        cps.append(p)
        c = ctf_pixel_constants(p)
        for key, val in c.items():
            prm[key][i] = val
    prm["seed"] = rng.integers(1, 2 ** 31, w.nptcls, dtype=np.uint32)
    w.prm = prm
    w.ctfparams = cps
    return w


def gen_refs(w: Workload, dbuf: DeviceBuffer):
    lib = synth_lib()
    rc = lib.synth_refs(w.nblobs, w.blobs.ctypes.data, w.nrefs, w.uvs.ctypes.data, w.nk, w.pftsz,
                        w.kfromto[0], w.box, C.c_void_p(dbuf.ptr))
    if rc:
        raise PftccError("synth_refs failed")


def gen_ptcls(w: Workload, first: int, count: int, dbuf: DeviceBuffer, with_ctf=True):
    lib = synth_lib()
    assert lib.synth_param_size() == PTCL_DTYPE.itemsize
    sub = np.ascontiguousarray(w.prm[first:first + count])
    rc = lib.synth_ptcls(w.nblobs, w.blobs.ctypes.data, w.nrefs, w.uvs.ctypes.data, count, sub.ctypes.data, w.nk,
                         w.pftsz, w.kfromto[0], w.box, int(with_ctf), C.c_void_p(dbuf.ptr))
    if rc:
        raise PftccError("synth_ptcls failed")


def set_noise(w: Workload, refs_host: np.ndarray, mean_ctf2: float = 0.5):
    """White noise of per-sample variance = mean signal power / SNR (signal = |CTF| * reference)."""
    ps = float(np.mean(np.abs(refs_host.astype(np.complex128)) ** 2)) * mean_ctf2
    var = ps / w.snr
    w.prm["noise_sd"] = np.float32(math.sqrt(var / 2.0))


def load_into(pftcc, w: Workload, chunk: int = 10000, with_ctf=True, progress=None):
    """Generates refs and particles on the device and loads them into a PolarFTCorrCalc."""
    ring = w.nk * w.pftsz * 8
    rbuf = DeviceBuffer(w.nrefs * ring)
    gen_refs(w, rbuf)
    refs_host = rbuf.to_array(np.complex64, (w.nrefs, w.nk, w.pftsz))
    set_noise(w, refs_host)
    pftcc.set_ref_pfts_dev(1, w.nrefs, rbuf.ptr, True)
    pftcc.set_ref_pfts_dev(1, w.nrefs, rbuf.ptr, False)
    rbuf.close()
    pbuf = DeviceBuffer(min(chunk, w.nptcls) * ring)
    for first in range(0, w.nptcls, chunk):
        cnt = min(chunk, w.nptcls - first)
        gen_ptcls(w, first, cnt, pbuf, with_ctf)
        pftcc.set_ptcl_pfts_dev(first + 1, cnt, pbuf.ptr)
        if progress:
            progress(first + cnt)
    pbuf.close()
    if with_ctf:
        pftcc.create_polar_absctfmats(w.ctfparams)
    pftcc.memoize_ffts()
    return refs_host
