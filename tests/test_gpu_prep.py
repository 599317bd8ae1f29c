"""GPU parity of the batched particle preparation (pftcc_prep_imgs[_dev], prepimg4align of
simple_strategy2D3D_common.f90:316-366) against the oracle restatement (oracle/prep_oracle.py).

Tolerance: the device transforms in FP32 where the oracle uses FP64, and evaluates cos/sin/atan2
in FP32 with a few ulps of difference, so each prepared transform agrees within 2e-5 of its
largest component.  CTF parameters are drawn so that no component lies within 2e-3 (box 64) or
1e-4 (box 200) of a CTF zero (the phase flip's sign is then the same on both sides), and the
inputs are fixed, so the edge medians of zero_background take the same quickselect path on both
sides.
"""
import numpy as np
import pytest

import pftcc_oracle as O
import polarizer_oracle as PZ
import prep_oracle as P
from gpu_helpers import say
from pftcc import Params, PolarFTCorrCalc, PftccError

pytestmark = pytest.mark.gpu

TOL = 2e-5


def _particles(rng, count, B):
    yy, xx = np.mgrid[0:B, 0:B].astype(np.float64)
    out = []
    for _ in range(count):
        img = rng.standard_normal((B, B)) * 0.5 + 0.3
        for _b in range(5):
            cx, cy = rng.uniform(-B / 6, B / 6, 2)
            s = rng.uniform(1.5, B / 20)
            img += 2.0 * np.exp(-((xx - B / 2 - cx) ** 2 + (yy - B / 2 - cy) ** 2) / (2 * s * s))
        out.append(img.astype(np.float32))
    return np.stack(out)


def _ctfs(rng, count, B):
    """CTF parameters whose values stay away from zero on the B x (B/2+1) grid: 2e-3 at box 64,
    1e-4 at larger boxes (both far above the few-ulp differences of the FP32 evaluations)."""
    margin = 2e-3 if B <= 64 else 1e-4
    out = []
    for _ in range(2000):
        p = O.CtfParams(smpd=1.3, kv=300.0, cs=2.7, fraca=0.1, dfx=float(rng.uniform(0.8, 3.0)),
                        dfy=float(rng.uniform(0.8, 3.0)), angast=float(rng.uniform(0, 180)))
        if np.abs(P.ctf_values_fp64(B, p)).min() > margin:
            out.append(p)
            if len(out) == count:
                return out
    raise AssertionError("no CTF parameters clear of the zeros")


def _handle(M, kto=None):
    n = 48 if M <= 64 else 308
    kto = kto or min(M // 2 - 2, 40)
    prm = Params(boxmatch=M, kfromto=(2, kto), kstop=kto, pftsz=n, ctf=False)
    return PolarFTCorrCalc().new(2, (1, 16), prm)


def _check(got, imgs, shifts, cfg, ctfs):
    for b in range(imgs.shape[0]):
        ref = P.prepimg4align(imgs[b], shifts[b], cfg, None if ctfs is None else ctfs[b])
        err = np.abs(got[b] - ref).max() / np.abs(ref).max()
        assert err <= TOL, (b, err)


@pytest.mark.parametrize("case", ["plain", "clip_shift_flip", "inner", "focus", "griddev"])
def test_prep_matches_oracle(case):
    B, M = (64, 48) if case == "clip_shift_flip" else (64, 64)
    say(f"prep parity: {case}, box {B}, boxmatch {M}")
    rng = np.random.default_rng({"plain": 11, "clip_shift_flip": 12, "inner": 13, "focus": 14, "griddev": 15}[case])
    count = 6
    imgs = _particles(rng, count, B)
    shifts = np.zeros((count, 2), np.float32)
    ctfs = None
    kw = dict(msk=20.0)
    cfg = P.PrepCfg(box=B, boxmatch=M, msk=20.0)
    if case == "clip_shift_flip":
        shifts = rng.uniform(-4, 4, (count, 2)).astype(np.float32)
        shifts[0] = (0.0005, -0.0002)            # below SHTHRESH: no shift applied
        shifts[1] = (3.0, 0.0)
        ctfs = _ctfs(rng, count, B)
        kw = dict(msk=17.0, ctfflag=1, ctfparams=ctfs)
        cfg = P.PrepCfg(box=B, boxmatch=M, msk=17.0, ctfflag=P.CTFFLAG_YES)
    elif case == "inner":
        kw = dict(msk=20.0, inner=8.0, width=4.0)
        cfg = P.PrepCfg(box=B, boxmatch=M, msk=20.0, inner=8.0, width=4.0)
    elif case == "focus":
        kw = dict(msk=20.0, focusmsk=14.0)
        cfg = P.PrepCfg(box=B, boxmatch=M, msk=20.0, focusmsk=14.0)
    elif case == "griddev":
        kw = dict(msk=20.0, griddev=True)
        cfg = P.PrepCfg(box=B, boxmatch=M, msk=20.0, griddev=True)
    lib = _handle(M)
    got = lib.prep_imgs(imgs, shifts, box=B, **kw)
    lib.kill()
    _check(got, imgs, shifts, cfg, ctfs)


def test_prep_production_box_and_polarize_chain():
    """Box 200 (radices 4 2 5 5), shifts and phase flip, then polarize the prepared transforms
    into particle slots: both stages against the oracle."""
    B = M = 200
    say("prep parity: box 200 with polarize")
    rng = np.random.default_rng(200)
    count = 4
    imgs = _particles(rng, count, B)
    shifts = rng.uniform(-5, 5, (count, 2)).astype(np.float32)
    ctfs = _ctfs(rng, count, B)
    cfg = P.PrepCfg(box=B, boxmatch=M, msk=80.0, ctfflag=P.CTFFLAG_YES)
    lib = _handle(M, kto=60)
    got = lib.prep_imgs(imgs, shifts, box=B, msk=80.0, ctfflag=1, ctfparams=ctfs)
    _check(got, imgs, shifts, cfg, ctfs)
    lib.init_polarizer()
    lib.polarize(np.arange(1, count + 1), got, isptcl=True)
    g = O.make_grid(308, 2, 60, M)
    tab = PZ.polarizer_tables(g.polar_x, g.polar_y, 308, M)
    for b in range(count):
        ref = PZ.polarize_vec(tab, P.prepimg4align(imgs[b], shifts[b], cfg, ctfs[b]))
        pft = lib.get_ptcl_pft(b + 1)
        assert np.abs(pft - ref).max() <= TOL * np.abs(ref).max(), b
    lib.kill()


def test_prep_dev_matches_host():
    from pftcc.synth import DeviceBuffer
    B, M = 64, 48
    say("prep: device-resident images")
    rng = np.random.default_rng(7)
    count = 5
    imgs = _particles(rng, count, B)
    shifts = rng.uniform(-3, 3, (count, 2)).astype(np.float32)
    lib = _handle(M)
    host = lib.prep_imgs(imgs, shifts, box=B, msk=18.0)
    d_in = DeviceBuffer.from_array(imgs)
    d_out = DeviceBuffer(count * M * (M // 2 + 1) * 8)
    lib.prep_imgs_dev(count, d_in.ptr, shifts, B, 18.0, d_out.ptr)
    got = d_out.to_array(np.complex64, (count, M, M // 2 + 1))
    d_in.close()
    d_out.close()
    np.testing.assert_array_equal(got, host)
    lib.kill()


def test_prep_errors():
    say("prep: error paths")
    lib = _handle(64)
    imgs = np.zeros((1, 64, 64), np.float32)
    with pytest.raises(PftccError, match="differs from the handle"):
        lib.prep_imgs(imgs, [[0, 0]], box=64, msk=20.0, boxmatch=48)
    with pytest.raises(PftccError, match="even"):
        lib.prep_imgs(np.zeros((1, 63, 63), np.float32), [[0, 0]], box=63, msk=20.0)
    with pytest.raises(PftccError, match="CTF parameters"):
        lib.prep_imgs(imgs, [[0, 0]], box=64, msk=20.0, ctfflag=1)
    with pytest.raises(PftccError, match="mask radius"):
        lib.prep_imgs(imgs, [[0, 0]], box=64, msk=0.0)
    lib.kill()

