"""Writes the golden fixtures of the correlation search (tests/golden/*.npz).

The reference ships no golden vectors for this path and cannot be built here (FFTW3 and the
simple_parameters stack are absent), so the expected values are produced by the FP64 oracle's
brute-force routes, which do not share code with the FFT routes they check: the direct O(R K n)
sum over every rotation (corrk_all_rots_direct) for gencorrs, the direct single-rotation sums for
gencorr(_grad)_for_rot_8, the closed-form CTF and shift phases.  Inputs are seeded and stored
with the outputs, so the fixtures pin (1) the oracle across changes and (2) the device results
without recomputing the oracle.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pftcc_oracle as O  # noqa: E402
from cases import make_case  # noqa: E402

CASES = {
    # name: (seed, n, kfrom, kto, ldim, nrefs, nptcls, astig)
    "small_n48": (901, 48, 2, 14, 48, 3, 4, 0.1),
    "box128_n192": (902, 192, 2, 12, 128, 2, 2, 0.1),
}
SHIFTS = np.array([[0.75, -1.25], [-2.0, 0.5]], np.float64)


def build(name):
    seed, n, kfrom, kto, ldim, nrefs, nptcls, astig = CASES[name]
    pf, cps = make_case(seed=seed, n=n, kfrom=kfrom, kto=kto, ctf=True, astig=astig, nrefs=nrefs,
                        nptcls=nptcls, ldim=ldim, eo_mixed=True)
    ctfp = np.array([[p.smpd, p.kv, p.cs, p.fraca, p.dfx, p.dfy, p.angast, p.phshift, float(p.phaseplate)]
                     for p in cps], np.float64)
    nrots = 2 * n
    # gencorrs: unshifted and at each shift, by the direct sum over all rotations
    cc = np.zeros((nptcls, nrefs, 1 + len(SHIFTS), nrots))
    for i in range(nptcls):
        for r in range(nrefs):
            cc[i, r, 0] = pf.gencorrs(r, i, None, method="direct")
            for s, sh in enumerate(SHIFTS):
                cc[i, r, s + 1] = pf.gencorrs(r, i, tuple(sh), method="direct")
    # FP64 objective + shift gradient at a few (ref, ptcl, shift, rot) tuples
    rng = np.random.default_rng(seed + 1)
    nq = 6
    q_ref = rng.integers(0, nrefs, nq)
    q_ptcl = rng.integers(0, nptcls, nq)
    q_rot = rng.integers(1, nrots + 1, nq)
    q_sh = rng.uniform(-3, 3, (nq, 2))
    q_f = np.zeros(nq)
    q_g = np.zeros((nq, 2))
    for t in range(nq):
        q_f[t], q_g[t] = pf.gencorr_grad_for_rot_8(q_ref[t], q_ptcl[t], q_sh[t], q_rot[t])
    ctf64 = np.stack([O.ctf_abs_mat(pf.grid, p) for p in cps])
    return dict(
        n=np.int32(n), kfrom=np.int32(kfrom), kto=np.int32(kto), ldim=np.int32(ldim),
        refs_even=pf.refs_even, refs_odd=pf.refs_odd, ptcls=pf.ptcls, iseven=pf.iseven.astype(np.int8),
        ctfparams=ctfp, ctf_fp32=pf.ctfmats.astype(np.float32), ctf64=ctf64, shifts=SHIFTS,
        gencorrs=cc, q_ref=q_ref.astype(np.int32), q_ptcl=q_ptcl.astype(np.int32), q_rot=q_rot.astype(np.int32),
        q_sh=q_sh, q_f=q_f, q_grad=q_g,
        shmat=O.shmat_closed(pf.grid, tuple(SHIFTS[0])),
        magic_ring2=np.arange(10, 200, 7, dtype=np.int32),
        magic_pftsz=np.array([O.magic_pftsz(int(r)) for r in range(10, 200, 7)], np.int32),
    )


def main():
    for name in CASES:
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **build(name))
        print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
