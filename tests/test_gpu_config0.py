"""The CPU-reference configuration end to end: the device search against the reference algorithm
restated in C++ (FP32 FFT route, OpenMP), on the full shape.

100 references (projections of a Gaussian-blob phantom along uniform random directions), 1000
particles (random reference, in-plane angle, +-5 px shift, astigmatic |CTF|, noise at SNR 0.05),
box 128 at 1.3 A/px, kfromto [2,40] (pftsz 192), exhaustive gencorrs over references x rotations.
Best (reference, rotation) must be identical wherever the CPU's own best two differ by more than
FP32 resolution (1e-5 of the peak); where they do not, the device's choice must score within that
resolution of the CPU's best on the CPU's own correlations.
"""
import os

import numpy as np
import pytest

import cpu_port
from gpu_helpers import say
import pftcc
from pftcc import Params, PolarFTCorrCalc, synth

pytestmark = pytest.mark.gpu


def test_config0_exhaustive_search_matches_cpu_reference():
    say("config0: 100 refs x 1000 ptcls, box 128, GPU vs CPU reference algorithm")
    box, smpd, kfromto = 128, 1.3, (2, 40)
    pftsz = pftcc.magic_pftsz(62)
    w = synth.Workload(box=box, smpd=smpd, kfromto=kfromto, pftsz=pftsz, nrefs=100, nptcls=1000, snr=0.05,
                       seed=20251206, dfrange=(1.0, 3.0), astig=0.1, max_shift=5.0)
    synth.make_params(w)
    lib = PolarFTCorrCalc().new(w.nrefs, (1, w.nptcls), Params(boxmatch=box, kfromto=kfromto, pftsz=pftsz, ctf=True))
    synth.load_into(lib, w, chunk=1000)
    ids = np.arange(1, w.nptcls + 1)
    gcc, gir, _, grot = lib.search(ids)
    refs = np.stack([lib.get_ref_pft(r + 1, True) for r in range(w.nrefs)])
    pt = np.stack([lib.get_ptcl_pft(i) for i in ids])
    ctf = np.stack([lib.get_ctfmat(i) for i in ids])
    sq = np.array([lib.get_sqsum_ptcl(i) for i in ids], np.float32)
    lib.kill()
    nthr = min(16, os.cpu_count() or 1)
    pa, pb = cpu_port.memoize(pt, nthr)
    ccc, cir, _, crot, allc = cpu_port.search_fp32(refs, ctf, sq, pa, pb, None, w.nk, nthr, want_all=True)
    same = (gir == cir) & (grot == crot)
    for i in np.where(~same)[0]:
        v = allc[i, :, 0]
        top = np.sort(v.ravel())
        assert top[-1] - top[-2] <= 1e-5 * top[-1], (i, top[-1] - top[-2])
        assert v[gir[i] - 1, grot[i] - 1] >= top[-1] - 1e-5 * top[-1]
    say(f"  identical assignments: {np.mean(same):.4f}; max |dcc| {np.max(np.abs(gcc - ccc)):.2e}")
    assert np.max(np.abs(gcc - ccc)) < 1e-5
