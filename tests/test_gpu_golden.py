"""The gfx950 library against the committed golden fixtures (tests/golden/), through the C ABI.

The library computes its own |CTF| from the stored CTF parameters (FP32, within the argument
error of the FP64 model the fixture used), so correlations, FP64 objectives and gradients are
compared at the FP32 tolerance of the plain form (1e-5 relative to the largest value).
"""
import os

import numpy as np
import pytest

from gpu_helpers import peak_rel_err, say
from pftcc import Params, PolarFTCorrCalc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def lib_from(g):
    n, kfrom, kto, ldim = int(g["n"]), int(g["kfrom"]), int(g["kto"]), int(g["ldim"])
    nrefs, nptcls = g["refs_even"].shape[0], g["ptcls"].shape[0]
    eo = np.where(g["iseven"] != 0, 0, 1).astype(np.int32)
    lib = PolarFTCorrCalc().new(nrefs, (1, nptcls), Params(boxmatch=ldim, kfromto=(kfrom, kto), kstop=kto,
                                                          pftsz=n, ctf=True), eoarr=eo)
    lib.set_ref_pfts(1, g["refs_even"], True)
    lib.set_ref_pfts(1, g["refs_odd"], False)
    lib.set_ptcl_pfts(1, g["ptcls"])
    from pftcc_oracle import CtfParams
    cps = [CtfParams(smpd=c[0], kv=c[1], cs=c[2], fraca=c[3], dfx=c[4], dfy=c[5], angast=c[6], phshift=c[7],
                     phaseplate=bool(c[8])) for c in g["ctfparams"]]
    lib.create_polar_absctfmats(cps)
    lib.memoize_ffts()
    return lib


@pytest.mark.parametrize("name", ["small_n48", "box128_n192"])
def test_library_reproduces_golden(name):
    say(f"golden fixture {name}")
    g = dict(np.load(os.path.join(GOLDEN, f"{name}.npz")))
    lib = lib_from(g)
    cc = g["gencorrs"]
    npt, nr = cc.shape[:2]
    for i in range(npt):
        for r in range(nr):
            assert peak_rel_err(lib.gencorrs(r + 1, i + 1), cc[i, r, 0]) < 1e-5
            for s, sh in enumerate(g["shifts"]):
                assert peak_rel_err(lib.gencorrs(r + 1, i + 1, sh.astype(np.float32)), cc[i, r, s + 1]) < 1e-5
    # FP64 objective and gradient: the device's FP32 |CTF| differs from the fixture's FP64 model
    # by the FP32 error of sin(chi) (up to ~1e-6 where chi is large), so compare relative to the
    # largest value, as for the correlations
    f, gr = lib.grad_batch(g["q_ref"] + 1, g["q_ptcl"] + 1, g["q_sh"], g["q_rot"])
    assert np.max(np.abs(f - g["q_f"])) <= 1e-5 * np.max(np.abs(g["q_f"]))
    assert np.max(np.abs(gr[:, :2] - g["q_grad"])) <= 1e-5 * np.max(np.abs(g["q_grad"]))
    # searches: plain (unshifted and shifted) and the rescored precision form give the fixture's argmax
    ids = np.arange(1, npt + 1)
    pcc, pir, _, prot = lib.search(ids)
    mcc, mir, mrot = lib.search_mfma(ids, ntop=nr)
    for i in range(npt):
        v = cc[i, :, 0]
        r_, t_ = np.unravel_index(int(np.argmax(v)), v.shape)
        top = np.sort(v.ravel())
        if top[-1] - top[-2] < 1e-5 * top[-1]:
            continue          # a tie at FP32 resolution: either answer is right
        assert (pir[i], prot[i]) == (r_ + 1, t_ + 1)
        assert (mir[i], mrot[i]) == (r_ + 1, t_ + 1)
        assert abs(mcc[i] - cc[i, r_, 0, t_]) < 1e-6
    lib.kill()
