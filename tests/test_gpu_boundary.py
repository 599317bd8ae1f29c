"""The rest of the polarft_corrcalc call surface through the C ABI, against the FP64 oracle.

* objective-specific forms gencorr_cc_* / gencorr_euclid_* (:1645-2148): the objective is fixed
  instead of read from cc_objfun; within 1e-9 relative of the oracle (the FP64 kernel);
* orientation derivatives set_dref_fcomp / set_dref_pft + gencorr_cont_grad_cc_for_rot_8 and
  gencorr_cont_shift_grad_cc_for_rot_8 (:459, :1735-1847), including the reference's use of
  particle slot iref in the T1 terms: within 1e-6 relative (the denominator takes the FP32
  memoized particle norm, as the reference's sqsums_ptcls); the Euclid cont forms are the
  reference's stubs (all zero);
* calc_ptcl_pspec (:897), assign_pinds (:706) with a particle mask, cp_refs (:491).
"""
import numpy as np
import pytest

import pftcc_oracle as O
from cases import make_case, random_pfts
from gpu_helpers import say, to_lib
from pftcc import Params, PftccError, PolarFTCorrCalc

pytestmark = pytest.mark.gpu


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-30)


@pytest.mark.parametrize("euclid_handle", [False, True])
def test_objective_specific_forms(euclid_handle):
    say(f"objective-specific forms, euclid handle={euclid_handle}")
    pf, cps = make_case(seed=201, n=64, kfrom=2, kto=20, ctf=True, nrefs=3, nptcls=4, ldim=64, euclid=True,
                        eo_mixed=True)
    sigma2 = pf.sigma2
    pf.euclid = euclid_handle
    lib = to_lib(pf, cps)
    if not euclid_handle:
        lib.assign_sigma2_noise(sigma2)
    for (r, p, irot, sh) in [(1, 1, 1, (0.0, 0.0)), (2, 3, 40, (0.75, -1.5)), (3, 4, 65, (-2.0, 1.0)),
                             (1, 2, 100, (1.5, 0.5))]:
        for euclid in (False, True):
            ref = pf.gencorr_obj_for_rot_8(r - 1, p - 1, sh, irot, euclid)
            fref, gref = pf.gencorr_obj_grad_for_rot_8(r - 1, p - 1, sh, irot, euclid)
            if euclid:
                v8 = lib.gencorr_euclid_for_rot_8(r, p, sh, irot)
                v4 = lib.gencorr_euclid_for_rot(r, p, sh, irot)
                f, g = lib.gencorr_euclid_grad_for_rot_8(r, p, sh, irot)
                go = lib.gencorr_euclid_grad_only_for_rot_8(r, p, sh, irot)
            else:
                v8 = lib.gencorr_cc_for_rot_8(r, p, sh, irot)
                v4 = lib.gencorr_cc_for_rot(r, p, sh, irot)
                f, g = lib.gencorr_cc_grad_for_rot_8(r, p, sh, irot)
                go = lib.gencorr_cc_grad_only_for_rot_8(r, p, sh, irot)
            assert rel(v8, ref) < 1e-9, (euclid, v8, ref)
            assert rel(v4, ref) < 1e-6
            assert rel(f, fref) < 1e-9
            scale = max(np.max(np.abs(gref)), 1e-30)
            assert np.max(np.abs(g - gref)) / scale < 1e-9
            assert np.array_equal(g, go)
            # the dispatching form agrees with the fixed form of the handle's own objective
            if euclid == euclid_handle:
                assert lib.gencorr_for_rot_8(r, p, sh, irot) == v8
    lib.kill()


def _cont_lib(seed, n, kto, ldim, match_filt=False):
    pf, cps = make_case(seed=seed, n=n, kfrom=2, kto=kto, ctf=True, nrefs=4, nptcls=6, ldim=ldim, eo_mixed=True,
                        match_filt=match_filt)
    lib = to_lib(pf, cps, nthr=4)
    rng = np.random.default_rng(seed + 7)
    drefs = {}
    for r in range(1, 5):
        for even in (True, False):
            d = random_pfts(rng, 3, pf.ks, pf.pftsz)
            drefs[(r, even)] = d
            lib.set_dref_pft(r, d, even)
    return pf, lib, drefs


@pytest.mark.parametrize("match_filt", [False, True])
def test_cont_grad_matches_oracle(match_filt):
    say(f"gencorr_cont_* vs oracle, match_filt={match_filt}")
    pf, lib, drefs = _cont_lib(203, 64, 20, 64, match_filt)
    for (r, p, irot, sh) in [(1, 1, 1, (0.0, 0.0)), (2, 5, 30, (0.75, -1.5)), (3, 2, 65, (-2.0, 1.0)),
                             (4, 6, 100, (1.5, 0.5)), (4, 3, 64, (0.1, 0.2))]:
        d = drefs[(r, bool(pf.iseven[p - 1]))]
        cc_ref, dcc_ref = pf.cont_grad_cc_for_rot_8(r - 1, p - 1, sh, irot, d)
        cc, dcc = lib.gencorr_cont_grad_cc_for_rot_8(r, p, sh, irot)
        assert rel(cc, cc_ref) < 1e-6, (cc, cc_ref)
        sc = max(np.max(np.abs(dcc_ref)), 1e-30)
        assert np.max(np.abs(dcc - dcc_ref)) / sc < 1e-6, (dcc, dcc_ref)
        f_ref, g_ref = pf.cont_shift_grad_cc_for_rot_8(r - 1, p - 1, sh, irot, d)
        f, g = lib.gencorr_cont_shift_grad_cc_for_rot_8(r, p, sh, irot)
        assert f == cc and np.array_equal(g[:3], dcc)
        sc = max(np.max(np.abs(g_ref)), 1e-30)
        assert np.max(np.abs(g - g_ref)) / sc < 1e-6, (g, g_ref)
    # the Euclid cont forms are the reference's stubs
    cc, dcc = lib.gencorr_cont_grad_euclid_for_rot_8(1, 1, (0.5, 0.5), 3)
    assert cc == 0.0 and not np.any(dcc)
    f, g = lib.gencorr_cont_shift_grad_euclid_for_rot_8(1, 1, (0.5, 0.5), 3)
    assert f == 0.0 and not np.any(g)
    lib.kill()


def test_set_dref_fcomp_equals_block_setter():
    say("set_dref_fcomp vs set_dref_pft")
    pf, lib, drefs = _cont_lib(205, 16, 7, 32)
    r, p, irot, sh = 2, 2, 5, (0.5, -0.5)
    even = bool(pf.iseven[p - 1])
    want = lib.gencorr_cont_grad_cc_for_rot_8(r, p, sh, irot)
    lib.set_dref_pft(r, np.zeros_like(drefs[(r, even)]), even)
    assert not np.array_equal(lib.gencorr_cont_grad_cc_for_rot_8(r, p, sh, irot)[1], want[1])
    d = drefs[(r, even)]
    for kk in range(pf.ks):
        for j in range(pf.pftsz):
            lib.set_dref_fcomp(r, j + 1, pf.kfrom + kk, d[:, kk, j], even)
    got = lib.gencorr_cont_grad_cc_for_rot_8(r, p, sh, irot)
    assert got[0] == want[0] and np.array_equal(got[1], want[1])
    with pytest.raises(PftccError):
        lib.set_dref_fcomp(5, 1, 2, d[:, 0, 0], even)        # slots 1..nthr = 4
    with pytest.raises(PftccError):
        lib.set_dref_fcomp(1, pf.pftsz + 1, 2, d[:, 0, 0], even)
    lib.kill()


def test_pspec_pinds_cp_refs():
    say("calc_ptcl_pspec, assign_pinds, cp_refs")
    pf, cps = make_case(seed=207, n=64, kfrom=2, kto=20, kstop=18, ctf=False, nrefs=3, nptcls=5, ldim=64)
    mask = np.array([1, 0, 1, 1, 0, 1, 1], np.int8)        # 7 global ids 11..17, 5 held
    prm = Params(boxmatch=64, kfromto=(2, 20), kstop=18, pftsz=64, ctf=False)
    lib = PolarFTCorrCalc().new(3, (11, 17), prm, ptcl_mask=mask)
    held = [11 + i for i in range(7) if mask[i]]
    for li, g in enumerate(held):
        lib.set_ptcl_pft(g, pf.ptcls[li])
    pfrom, pinds = lib.assign_pinds()
    assert pfrom == 11 and list(pinds) == [1, 0, 2, 3, 0, 4, 5]
    for li, g in enumerate(held):
        ps = lib.calc_ptcl_pspec(g)
        ref = pf.calc_ptcl_pspec(li)
        assert np.max(np.abs(ps - ref) / ref) < 1e-6
    # cp_refs: references of another handle of the same shape
    lib.set_ref_pfts(1, pf.refs_even, True)
    lib.set_ref_pfts(1, pf.refs_odd, False)
    lib.memoize_ffts()
    other = PolarFTCorrCalc().new(3, (11, 17), prm, ptcl_mask=mask)
    other.cp_refs(lib)
    for r in range(1, 4):
        for even in (True, False):
            assert np.array_equal(other.get_ref_pft(r, even), lib.get_ref_pft(r, even))
    bad = PolarFTCorrCalc().new(2, (11, 17), prm, ptcl_mask=mask)
    with pytest.raises(PftccError):
        bad.cp_refs(lib)
    for h in (bad, other, lib):
        h.kill()
