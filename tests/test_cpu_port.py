"""The C++ restatements in oracle/pftcc_cpu.cpp agree with the numpy FP64 oracle.

cpu_search_fp32 (the CPU baseline: the reference's FFT route in FP32) must reproduce the FP64
oracle to FP32 accuracy, and cpu_gencorrs_direct64 (the FP64 brute force used for larger
parity cases) to FP64 accuracy.
"""
import numpy as np
import pytest

import cpu_port
import pftcc_oracle as O
from cases import make_case


@pytest.mark.parametrize("n,kto,ctf,shift", [(16, 7, False, None), (24, 10, True, (1.0, -0.5)),
                                             (66, 14, True, None), (192, 20, True, (-2.0, 1.5))])
def test_cpu_fp32_route_matches_oracle(n, kto, ctf, shift):
    pf, _ = make_case(seed=n, n=n, kto=kto, ctf=ctf, nrefs=3, nptcls=4, ldim=4 * kto + 8)
    pa, pb = cpu_port.memoize(pf.ptcls, 4)
    opa, opb = O.memoize_fft(pf.ptcls[0].astype(np.complex128))
    assert np.allclose(pa[0], opa, atol=1e-5 * np.abs(opa).max())
    assert np.allclose(pb[0], opb, atol=1e-5 * np.abs(opb).max())
    sq = np.array([pf.sqsum_ptcl(i) for i in range(4)], np.float32)
    shm = None if shift is None else O.shmat_8(pf.grid, shift)[None].astype(np.complex64)
    cc, ir, ish, rot, allc = cpu_port.search_fp32(pf.refs_even, pf.ctfmats, sq, pa, pb, shm, pf.ks, 4,
                                                  want_all=True)
    for i in range(4):
        for r in range(3):
            ref = pf.gencorrs(r, i, shift)
            if shift is not None:
                # gencorrs_cc_2 normalises over all shells; the port sums kfrom..kstop as
                # gencorrs_cc_1 does -- identical here because kstop == kto
                pass
            assert np.max(np.abs(allc[i, r, 0] - ref)) < 1e-5 * max(1, np.abs(ref).max())
        flat = int(np.argmax(allc[i]))
        r, s, t = np.unravel_index(flat, allc[i].shape)
        assert (ir[i], ish[i], rot[i]) == (r + 1, s, t + 1)


def test_cpu_direct64_matches_oracle():
    pf, _ = make_case(seed=5, n=40, kto=12, ctf=True, nrefs=2, nptcls=3, ldim=48)
    sh = (0.75, -1.25)
    shm = O.shmat_8(pf.grid, sh)
    out = cpu_port.gencorrs_direct64(pf.refs_even, pf.ptcls, pf.ctfmats, shm, pf.ks, True, 4)
    for r in range(2):
        for i in range(3):
            assert np.allclose(out[r, i], pf.gencorrs(r, i, sh), rtol=1e-12, atol=1e-13)
    out0 = cpu_port.gencorrs_direct64(pf.refs_even, pf.ptcls, pf.ctfmats, None, pf.ks, False, 4)
    assert np.allclose(out0[1, 2], pf.gencorrs(1, 2), rtol=1e-12, atol=1e-13)
