"""GPU parity of the batched gradient shift search (pftcc_shsrch_grad_batch) against the oracle's
restatement of pftcc_shsrch_grad%minimize (simple_pftcc_shsrch_grad.f90:254-338) on the same
inputs, pair by pair.

Both sides evaluate the same FP64 objective (gencorr_grad_for_rot_8) and run the same L-BFGS-B
decisions, so the trajectories coincide; the oracle's in-plane callback uses the FP64
restatement of gencorrs where the device uses the FP32 search kernel, which can only change the
chosen rotation at a near tie (top-2 gap below 1e-5 of the peak; none occurs on these inputs).
Tolerances: rotation exact; cc within 1e-5; shift within 1e-4 px.
"""
import numpy as np
import pytest

import lbfgsb_oracle as L
import pftcc_oracle as O
from cases import make_case
from gpu_helpers import say, to_lib
from pftcc import PftccError

pytestmark = pytest.mark.gpu

TRS = 3.0


def _planted_case(seed, ctf, objective="cc", n=48, kto=16, ldim=64, nptcls=5):
    pf, cps = make_case(seed=seed, n=n, kfrom=2, kto=kto, nrefs=3, nptcls=nptcls, ctf=ctf, ldim=ldim,
                        euclid=objective == "euclid", match_filt=objective == "match_filt")
    rng = np.random.default_rng(seed)
    n = pf.pftsz
    # particles 0..2: reference p rotated by a whole number of steps and shifted by a sub-pixel
    # vector, plus noise; particles 3, 4 stay random
    for p in range(3):
        t = rng.uniform(-2.0, 2.0, 2)
        rho = int(rng.integers(0, 2 * n))
        q = pf.refs_even[p].astype(np.complex128) * O.shmat_8(pf.grid, -t)
        full = np.roll(O.full_ring(q), rho, axis=1)[:, :n]
        noise = 0.05 * np.abs(q).mean() * (rng.standard_normal(full.shape) + 1j * rng.standard_normal(full.shape))
        pf.ptcls[p] = (full + noise).astype(np.complex64)
    return pf, cps


def _oracle_pair(pf, r, p, opt_angle, irot, coarse=False):
    def fdf(sh, rot):
        return pf.gencorr_grad_for_rot_8(r, p, sh, rot)

    def gencorrs(sh32):
        return pf.gencorrs(r, p, tuple(float(v) for v in sh32), method="fft")
    return L.shsrch_minimize(fdf, gencorrs, pf.grid.angtab_deg, trs=TRS, irot=irot, opt_angle=opt_angle,
                             coarse_init=coarse)


@pytest.mark.parametrize("opt_angle,ctf,objective,coarse",
                         [(True, False, "cc", False), (True, True, "cc", False), (False, True, "cc", False),
                          (True, True, "euclid", False), (True, True, "match_filt", False),
                          (True, True, "cc", True), (False, True, "cc", True)])
def test_shift_search_matches_oracle(opt_angle, ctf, objective, coarse):
    say(f"shift search parity: opt_angle {opt_angle}, ctf {ctf}, objective {objective}, coarse grid {coarse}")
    pf, cps = _planted_case(31 + int(ctf), ctf, objective)
    lib = to_lib(pf, cps)
    nr, npt = 3, 5
    pairs = [(r, p) for p in range(npt) for r in range(nr)]
    irots = []
    for r, p in pairs:
        irots.append(int(np.argmax(pf.gencorrs(r, p, (0.0, 0.0), method="fft"))) + 1)
    cxy, rot = lib.shsrch_grad_batch([r + 1 for r, _ in pairs], [p + 1 for _, p in pairs], trs=TRS, maxits=60,
                                     opt_angle=opt_angle, irots=irots, coarse_init=coarse)
    lib.kill()
    found = 0
    for i, (r, p) in enumerate(pairs):
        ocxy, orot = _oracle_pair(pf, r, p, opt_angle, irots[i], coarse)
        assert rot[i] == orot, (r, p, rot[i], orot)
        if orot == 0:
            continue
        found += 1
        assert abs(cxy[i, 0] - ocxy[0]) <= 1e-5 * max(1.0, abs(ocxy[0])), (r, p, cxy[i], ocxy)
        np.testing.assert_allclose(cxy[i, 1:], ocxy[1:], rtol=0, atol=1e-4, err_msg=f"pair {(r, p)}")
    assert found >= 3
    # the planted pairs: high cc at the planted particle's own reference (planted without the
    # |CTF| the correlation applies to the reference, hence below 1 with ctf)
    if objective == "cc":
        for p in range(3):
            i = pairs.index((p, p))
            assert rot[i] > 0 and cxy[i, 0] > (0.8 if ctf else 0.95)


def test_shift_search_ring_length_600():
    """The in-plane callback's compile-time kernel at ring length 600 (box 384; transform radices
    4, 2, 3, 5, 5): planted pairs against the oracle's L-BFGS-B restatement."""
    say("shift search parity at n 600")
    pf, cps = _planted_case(61, True, n=600, kto=20, ldim=384, nptcls=3)
    lib = to_lib(pf, cps)
    pairs = [(r, p) for p in range(3) for r in range(3)]
    irots = [int(np.argmax(pf.gencorrs(r, p, (0.0, 0.0), method="fft"))) + 1 for r, p in pairs]
    cxy, rot = lib.shsrch_grad_batch([r + 1 for r, _ in pairs], [p + 1 for _, p in pairs], trs=TRS, maxits=60,
                                     opt_angle=True, irots=irots)
    lib.kill()
    found = 0
    for i, (r, p) in enumerate(pairs):
        ocxy, orot = _oracle_pair(pf, r, p, True, irots[i])
        assert rot[i] == orot, (r, p, rot[i], orot)
        if orot == 0:
            continue
        found += 1
        assert abs(cxy[i, 0] - ocxy[0]) <= 1e-5 * max(1.0, abs(ocxy[0])), (r, p, cxy[i], ocxy)
        np.testing.assert_allclose(cxy[i, 1:], ocxy[1:], rtol=0, atol=1e-4, err_msg=f"pair {(r, p)}")
    assert found >= 3
    # the planted pairs are found (planted without the |CTF| the correlation applies to the
    # reference: 0.76 at this shape)
    for p in range(3):
        i = pairs.index((p, p))
        assert rot[i] > 0 and cxy[i, 0] > 0.6


def test_shift_search_batch_is_order_independent():
    """Results are per pair: a permuted, repeated batch gives the same answers pair by pair."""
    say("shift search: permuted batch")
    pf, cps = _planted_case(41, True)
    lib = to_lib(pf, cps)
    refs = np.array([1, 2, 3, 1, 2, 3, 1, 2], np.int32)
    ptcls = np.array([1, 2, 3, 4, 5, 1, 2, 3], np.int32)
    c1, r1 = lib.shsrch_grad_batch(refs, ptcls, trs=TRS)
    perm = np.random.default_rng(0).permutation(refs.size)
    reps = np.concatenate([perm, perm])
    c2, r2 = lib.shsrch_grad_batch(refs[reps], ptcls[reps], trs=TRS)
    lib.kill()
    np.testing.assert_array_equal(r2[: refs.size], r1[perm])
    np.testing.assert_array_equal(r2[refs.size:], r1[perm])
    np.testing.assert_array_equal(c2[: refs.size], c1[perm])


def test_shift_search_errors():
    say("shift search: error paths")
    pf, cps = make_case(seed=2, n=32, kfrom=2, kto=10, nrefs=2, nptcls=2, ctf=False, ldim=32)
    lib = to_lib(pf, cps, memoize=False)
    with pytest.raises(PftccError, match="memoize"):
        lib.shsrch_grad_batch([1], [1], trs=2.0)
    # without the rotation callback only the FP64 objective is needed (no memoized spectra)
    lib.shsrch_grad_batch([1], [1], trs=2.0, opt_angle=False, irots=[1])
    lib.memoize_ffts()
    with pytest.raises(PftccError, match="empty"):
        lib.shsrch_grad_batch([1], [1], lims=(1.0, -1.0, -1.0, 1.0))
    with pytest.raises(PftccError, match="irot"):
        lib.shsrch_grad_batch([1], [1], trs=2.0, opt_angle=False, irots=[0])
    with pytest.raises(PftccError):
        lib.shsrch_grad_batch([3], [1], trs=2.0)
    with pytest.raises(PftccError):
        lib.shsrch_grad_batch([1], [7], trs=2.0)
    cxy, rot = lib.shsrch_grad_batch([], [], trs=2.0)
    assert cxy.shape == (0, 3)
    lib.kill()
