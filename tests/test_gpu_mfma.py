"""Parity of the precision (MFMA, split-BF16) form with the FP64 oracle.

The form is exact for ring-constant |CTF| (no CTF, or no astigmatism); with astigmatism it uses the
particle's ring-mean |CTF| per shell, and is then checked against the oracle fed that same
ring-mean CTF.  Tolerance: |cc_mfma - cc_oracle| <= 1e-4 relative to the correlation peak (split
BF16 keeps ~16 mantissa bits per operand; the dropped lo*lo term is 2^-16 relative), best
(reference, rotation) identical wherever the oracle's top-2 gap exceeds 10x the measured error.
"""
import copy

import numpy as np
import pytest

import pftcc_oracle as O
from cases import make_case
from gpu_helpers import peak_rel_err, say, to_lib, top2_gap
from pftcc import (PFTCC_PREC_BF16, PFTCC_PREC_FP8, PFTCC_PREC_FP16, PFTCC_PREC_SPLIT_BF16, PFTCC_PREC_SPLIT_FP16,
                   PftccError)

pytestmark = pytest.mark.gpu

MFMA_TOL = 1e-4
FP16_TOL = 1e-3


def oracle_cube(pf, ptcl_idx, ref_idx):
    """[np, nr, nrots] FP64 correlations of the oracle (FFT route, unshifted)."""
    return np.stack([np.stack([pf.gencorrs(r, p, None, method="fft") for r in ref_idx]) for p in ptcl_idx])


def ring_mean_case(pf):
    """The oracle case with each particle's |CTF| replaced by its per-shell ring mean."""
    m = pf.ctfmats.astype(np.float64).mean(axis=-1, keepdims=True)
    pf.ctfmats = np.broadcast_to(m, pf.ctfmats.shape).astype(np.float32).copy()
    return pf


def check_cube(got, ref):
    errs = []
    for p in range(ref.shape[0]):
        e = peak_rel_err(got[p], ref[p])
        errs.append(e)
        assert e < MFMA_TOL, (p, e)
        if top2_gap(ref[p]) > 10 * e * np.max(np.abs(ref[p])):
            assert int(np.argmax(got[p])) == int(np.argmax(ref[p]))
    return max(errs)


@pytest.mark.parametrize("n,kfrom,kto,ldim,ctf,astig", [
    (40, 2, 12, 32, False, 0.0),
    (64, 2, 20, 64, True, 0.0),
    (192, 2, 40, 128, True, 0.0),
    (308, 2, 60, 200, False, 0.0),
    (400, 2, 80, 256, True, 0.0),
    (600, 2, 40, 384, True, 0.0),     # radices 4, 2, 3, 5, 5 in the inverse stage
])
def test_mfma_exact_for_ring_constant_ctf(n, kfrom, kto, ldim, ctf, astig):
    say(f"mfma vs oracle n={n} K=[{kfrom},{kto}] ctf={ctf} astig={astig}")
    pf, cps = make_case(seed=300 + n, n=n, kfrom=kfrom, kto=kto, ctf=ctf, astig=astig, nrefs=5, nptcls=7,
                        ldim=ldim, eo_mixed=True)
    lib = to_lib(pf, cps)
    ids = np.arange(1, 8)
    got = lib.gencorrs_mfma_batch(ids)
    ref = oracle_cube(pf, range(7), range(5))
    e = check_cube(got, ref)
    say(f"  split-bf16 max peak-relative error {e:.2e}")
    lo = lib.gencorrs_mfma_batch(ids, prec=PFTCC_PREC_BF16)
    say(f"  plain bf16 max peak-relative error {max(peak_rel_err(lo[p], ref[p]) for p in range(7)):.2e}")
    # split FP16 (about 22 significant bits per operand): the split-BF16 bar
    e16s = check_cube(lib.gencorrs_mfma_batch(ids, prec=PFTCC_PREC_SPLIT_FP16), ref)
    say(f"  split-fp16 max peak-relative error {e16s:.2e}")
    # FP16, one product (11 significant bits per operand, rounding errors of ~2^-12 that average
    # over the 2 K terms of each angular frequency and the 2n of the inverse transform)
    h = lib.gencorrs_mfma_batch(ids, prec=PFTCC_PREC_FP16)
    e16 = max(peak_rel_err(h[p], ref[p]) for p in range(7))
    say(f"  fp16 max peak-relative error {e16:.2e}")
    assert e16 < FP16_TOL, e16
    lib.kill()


def test_mfma_astigmatic_uses_ring_mean_ctf():
    say("mfma astigmatic ctf = oracle with ring-mean ctf")
    pf, cps = make_case(seed=41, n=192, kfrom=2, kto=40, ctf=True, astig=0.3, nrefs=4, nptcls=6, ldim=128)
    lib = to_lib(pf, cps)
    got = lib.gencorrs_mfma_batch(np.arange(1, 7))
    exact = oracle_cube(pf, range(6), range(4))
    ref = oracle_cube(ring_mean_case(pf), range(6), range(4))
    e = check_cube(got, ref)
    say(f"  vs ring-mean oracle {e:.2e}; vs exact astigmatic oracle "
        f"{max(peak_rel_err(got[p], exact[p]) for p in range(6)):.2e}")
    lib.kill()


def test_mfma_subsets_order_and_search():
    """Particle / reference subsets in any order, eo classes mixed, and search_mfma = argmax."""
    say("mfma subsets + search_mfma")
    pf, cps = make_case(seed=42, n=64, kfrom=2, kto=20, ctf=True, astig=0.0, nrefs=9, nptcls=11, ldim=64,
                        eo_mixed=True)
    lib = to_lib(pf, cps)
    pids = np.array([7, 2, 11, 1, 5], np.int32)
    rids = np.array([9, 3, 4, 1], np.int32)
    got = lib.gencorrs_mfma_batch(pids, rids)
    ref = oracle_cube(pf, pids - 1, rids - 1)
    check_cube(got, ref)
    cc, ir, rot = lib.search_mfma(pids, rids)
    for i in range(len(pids)):
        flat = int(np.argmax(got[i]))
        r, t = np.unravel_index(flat, got[i].shape)
        assert (ir[i], rot[i]) == (rids[r], t + 1)
        assert cc[i] == got[i][r, t]
    # a second search with the same reference list reuses the packed operand
    cc2, ir2, rot2 = lib.search_mfma(pids[::-1], rids)
    assert np.array_equal(cc2, cc[::-1]) and np.array_equal(ir2, ir[::-1]) and np.array_equal(rot2, rot[::-1])
    # changing a reference invalidates it
    lib.set_ref_pft(rids[0], np.zeros_like(pf.refs_even[0]), True)
    lib.set_ref_pft(rids[0], np.zeros_like(pf.refs_odd[0]), False)
    pf.refs_even[rids[0] - 1] = 0
    pf.refs_odd[rids[0] - 1] = 0
    cc3, ir3, _ = lib.search_mfma(pids, rids[1:])
    assert not np.any(ir3 == rids[0])
    lib.kill()


def test_mfma_many_particles_chunks():
    """More particles than one chunk (256) and more references than one 64-row tile."""
    say("mfma 600 ptcls x 70 refs, n=40")
    pf, cps = make_case(seed=43, n=40, kfrom=2, kto=12, ctf=True, astig=0.0, nrefs=70, nptcls=600, ldim=32,
                        eo_mixed=True)
    lib = to_lib(pf, cps)
    pids = np.arange(1, 601)
    cc, ir, rot = lib.search_mfma(pids)
    sel = np.array([0, 1, 255, 256, 257, 511, 512, 599])
    got = lib.gencorrs_mfma_batch(pids[sel])
    ref = oracle_cube(pf, sel, range(70))
    check_cube(got, ref)
    for i, p in enumerate(sel):
        r, t = np.unravel_index(int(np.argmax(got[i])), got[i].shape)
        assert (ir[p], rot[p]) == (r + 1, t + 1)
    lib.kill()


def test_mfma_rejects_unsupported():
    say("mfma errors")
    pf, cps = make_case(seed=44, n=40, kto=12, ctf=False, nrefs=2, nptcls=2, euclid=True)
    lib = to_lib(pf, cps)
    with pytest.raises(PftccError):
        lib.search_mfma(np.arange(1, 3))
    lib.kill()
    pf, cps = make_case(seed=44, n=40, kto=12, ctf=False, nrefs=2, nptcls=2)
    lib = to_lib(pf, cps)
    with pytest.raises(PftccError):
        lib.search_mfma(np.arange(1, 3), prec=5)
    with pytest.raises(PftccError):
        lib.search_mfma(np.array([5], np.int32))
    with pytest.raises(PftccError):
        lib.search_mfma(np.arange(1, 3), np.array([3], np.int32))
    lib.kill()


@pytest.mark.parametrize("astig,prec", [(0.0, PFTCC_PREC_SPLIT_BF16), (0.3, PFTCC_PREC_SPLIT_BF16),
                                        (0.3, PFTCC_PREC_FP16), (0.3, PFTCC_PREC_SPLIT_FP16)])
def test_mfma_rescored_search_is_exact(astig, prec):
    """ntop > 0: the candidates are rescored in FP64 with the particle's own |CTF|, so the result is
    the oracle's exact argmax (astigmatic CTF included) and its cc the oracle's FP64 value."""
    say(f"mfma rescored search astig={astig} prec={prec}")
    pf, cps = make_case(seed=45, n=192, kfrom=2, kto=40, ctf=True, astig=astig, nrefs=12, nptcls=9, ldim=128,
                        eo_mixed=True)
    # plant rotated, CTF-multiplied copies of references so that each particle has a clear winner
    rng = np.random.default_rng(46)
    for i in range(9):
        r = int(rng.integers(0, 12))
        rho = int(rng.integers(0, 2 * pf.pftsz))
        src = pf.refs_even[r] if pf.iseven[i] else pf.refs_odd[r]
        rolled = np.roll(O.full_ring(src.astype(np.complex128)), rho, axis=-1)[:, :pf.pftsz]
        noise = 0.3 * (rng.standard_normal(rolled.shape) + 1j * rng.standard_normal(rolled.shape))
        pf.ptcls[i] = (rolled + noise * np.abs(rolled).mean()).astype(np.complex64)
    lib = to_lib(pf, cps)
    ids = np.arange(1, 10)
    ref = oracle_cube(pf, range(9), range(12))
    cc, ir, rot = lib.search_mfma(ids, ntop=4, prec=prec)
    pcc, pir, _, prot = lib.search(ids)
    for i in range(9):
        r, t = np.unravel_index(int(np.argmax(ref[i])), ref[i].shape)
        assert (ir[i], rot[i]) == (r + 1, t + 1)
        assert abs(cc[i] - ref[i][r, t]) <= 1e-7
        assert (pir[i], prot[i]) == (ir[i], rot[i])
    raw_cc, raw_ir, raw_rot = lib.search_mfma(ids, ntop=0, prec=prec)
    say(f"  raw vs exact max |dcc| {np.max(np.abs(raw_cc - cc)):.2e}")
    with pytest.raises(PftccError):
        lib.search_mfma(ids, ntop=257)
    lib.kill()


def test_fp8_coarse_pass_error_and_rescoring():
    """FP8 e4m3 contraction: a coarse score (e4m3 keeps 3 mantissa bits, ~3.6% rms per operand
    element, so the score error is a few percent of the peak on noise-only data); with ntop
    candidates rescored in FP64 the search returns the oracle's exact argmax and cc."""
    say("fp8 coarse pass + rescoring")
    pf, cps = make_case(seed=47, n=192, kfrom=2, kto=40, ctf=True, astig=0.1, nrefs=24, nptcls=10, ldim=128,
                        eo_mixed=True)
    ids = np.arange(1, 11)
    lib = to_lib(pf, cps)
    got = lib.gencorrs_mfma_batch(ids, prec=PFTCC_PREC_FP8)
    ref = oracle_cube(ring_mean_case(copy.deepcopy(pf)), range(10), range(24))
    errs = [peak_rel_err(got[p], ref[p]) for p in range(10)]
    say(f"  fp8 peak-relative error vs ring-mean oracle: max {max(errs):.2e}")
    assert max(errs) < 0.15
    lib.kill()
    # planted winners: exact after rescoring
    rng = np.random.default_rng(48)
    for i in range(10):
        r = int(rng.integers(0, 24))
        rho = int(rng.integers(0, 2 * pf.pftsz))
        src = pf.refs_even[r] if pf.iseven[i] else pf.refs_odd[r]
        rolled = np.roll(O.full_ring(src.astype(np.complex128)), rho, axis=-1)[:, :pf.pftsz]
        noise = 1.0 * (rng.standard_normal(rolled.shape) + 1j * rng.standard_normal(rolled.shape))
        pf.ptcls[i] = (rolled * pf.ctfmats[i] + noise * np.abs(rolled).mean()).astype(np.complex64)
    lib = to_lib(pf, cps)
    exact = oracle_cube(pf, range(10), range(24))
    cc, ir, rot = lib.search_mfma(ids, prec=PFTCC_PREC_FP8, ntop=16)
    for i in range(10):
        r, t = np.unravel_index(int(np.argmax(exact[i])), exact[i].shape)
        assert (ir[i], rot[i]) == (r + 1, t + 1)
        assert abs(cc[i] - exact[i][r, t]) <= 1e-7
    lib.kill()


@pytest.mark.parametrize("n,kto,ldim", [(66, 20, 48), (192, 70, 160), (600, 20, 384)])
def test_fp8_spectra_paths(n, kto, ldim):
    """The FP8 contraction's other paths: at ring length 66 (no compile-time inverse stage) the
    spectra stay float2 and each wavefront takes 8 consecutive m (K' = 128); with 69 shells
    (K' = 256, two k' steps) one m per wavefront, with FP16 spectra at ring length 192; FP16
    spectra and 8 m per wavefront at ring length 600 (inverse-stage radices 4, 2, 3, 5, 5).  Coarse
    correlations against the ring-mean oracle; with rescoring, the FP64 oracle's exact argmax and
    cc for planted winners."""
    say(f"fp8 spectra paths n={n} K=[2,{kto}]")
    nr, npt = 12, 5
    pf, cps = make_case(seed=400 + n, n=n, kfrom=2, kto=kto, ctf=True, astig=0.1, nrefs=nr, nptcls=npt, ldim=ldim,
                        eo_mixed=True)
    rng = np.random.default_rng(401 + n)
    for i in range(npt):
        r = int(rng.integers(0, nr))
        rho = int(rng.integers(0, 2 * pf.pftsz))
        src = pf.refs_even[r] if pf.iseven[i] else pf.refs_odd[r]
        rolled = np.roll(O.full_ring(src.astype(np.complex128)), rho, axis=-1)[:, :pf.pftsz]
        noise = 1.0 * (rng.standard_normal(rolled.shape) + 1j * rng.standard_normal(rolled.shape))
        pf.ptcls[i] = (rolled * pf.ctfmats[i] + noise * np.abs(rolled).mean()).astype(np.complex64)
    ids = np.arange(1, npt + 1)
    lib = to_lib(pf, cps)
    coarse = lib.gencorrs_mfma_batch(ids, prec=PFTCC_PREC_FP8).astype(np.float64)
    ref = oracle_cube(ring_mean_case(copy.deepcopy(pf)), range(npt), range(nr))
    errs = [peak_rel_err(coarse[p], ref[p]) for p in range(npt)]
    say(f"  fp8 peak-relative error vs ring-mean oracle: max {max(errs):.2e}")
    assert max(errs) < 0.15
    exact = oracle_cube(pf, range(npt), range(nr))
    cc, ir, rot = lib.search_mfma(ids, prec=PFTCC_PREC_FP8, ntop=nr)
    for i in range(npt):
        r, t = np.unravel_index(int(np.argmax(exact[i])), exact[i].shape)
        assert (ir[i], rot[i]) == (r + 1, t + 1)
        assert abs(cc[i] - exact[i][r, t]) <= 1e-7
    lib.kill()


def _fp8_keys(cube, n2):
    """Per (particle, reference) best key of the coarse pass: its largest value, ties to the lowest
    rotation; references then ordered by key (value descending, ties to the lowest reference)."""
    np_, nr = cube.shape[:2]
    rot = np.argmax(cube, axis=-1)                  # first maximum = lowest rotation
    val = np.take_along_axis(cube, rot[..., None], axis=-1)[..., 0]
    order = [sorted(range(nr), key=lambda r: (-val[p, r], r)) for p in range(np_)]
    return rot, order


def test_fp8_coarse_pass_c4_shape_topt_and_rescoring():
    """FP8 e4m3 coarse pass at the large coarse-to-fine shape (box 256, kfromto [2,64]: pftsz 400,
    K 63) against the ring-mean oracle, and the rescoring of the top-T references for T from 1 to
    256 (beyond nrefs): from the device's own coarse correlations and the FP64 oracle the result is
    predicted exactly for every T (the T references with the largest coarse maxima, each rescored
    in FP64 at its coarse best rotation and both neighbours, the best winning, ties to the lowest
    (reference, rotation)); its FP64 cc can only grow with T."""
    say("fp8 coarse pass, n 400, K 63: top-T selection and FP64 rescoring")
    nr, npt = 120, 6
    pf, cps = make_case(seed=57, n=400, kfrom=2, kto=64, ctf=True, astig=0.1, nrefs=nr, nptcls=npt, ldim=256,
                        eo_mixed=True)
    ids = np.arange(1, npt + 1)
    lib = to_lib(pf, cps)
    coarse = lib.gencorrs_mfma_batch(ids, prec=PFTCC_PREC_FP8).astype(np.float64)
    ref = oracle_cube(ring_mean_case(copy.deepcopy(pf)), range(npt), range(nr))
    errs = [peak_rel_err(coarse[p], ref[p]) for p in range(npt)]
    say(f"  fp8 peak-relative error vs ring-mean oracle: max {max(errs):.2e}")
    assert max(errs) < 0.15
    exact = oracle_cube(pf, range(npt), range(nr))
    n2 = 2 * pf.pftsz
    rot, order = _fp8_keys(coarse.astype(np.float32), n2)
    prev = np.full(npt, -np.inf)
    for T in (1, 4, 16, 64, 100, 120, 256):
        cc, ir, rt = lib.search_mfma(ids, prec=PFTCC_PREC_FP8, ntop=T)
        for p in range(npt):
            best, bkey = -np.inf, None
            for r in order[p][:T]:
                for t in ((rot[p, r] - 1) % n2, rot[p, r], (rot[p, r] + 1) % n2):
                    v = exact[p, r, t]
                    if v > best or (v == best and (r, t) < bkey):
                        best, bkey = v, (r, t)
            assert (ir[p], rt[p]) == (bkey[0] + 1, bkey[1] + 1), (T, p, ir[p], rt[p], bkey)
            assert abs(cc[p] - best) <= 1e-7, (T, p, cc[p], best)
            assert best >= prev[p]
            prev[p] = best
    with pytest.raises(PftccError):
        lib.search_mfma(ids, prec=PFTCC_PREC_FP8, ntop=257)
    lib.kill()


def test_mfma_with_shifts_matches_oracle_and_rescoring():
    """Shifted precision form (gencorrs_2): each (reference, shift) is one operand row and the
    reference norm runs over all shells.  Ring-constant |CTF|: within the split-BF16 tolerance of
    the oracle; with astigmatism and rescoring: the oracle's exact (ref, shift, rot) and cc."""
    say("mfma with shifts")
    shifts = np.array([[0.0, 0.0], [1.5, -0.5], [-2.0, 1.0]], np.float32)
    pf, cps = make_case(seed=49, n=64, kfrom=2, kto=20, kstop=18, ctf=True, astig=0.0, nrefs=4, nptcls=5, ldim=64,
                        eo_mixed=True)
    lib = to_lib(pf, cps)
    got = lib.gencorrs_mfma_batch(np.arange(1, 6), shifts=shifts)
    assert got.shape == (5, 4, 3, 128)
    for i in range(5):
        ref = np.stack([np.stack([pf.gencorrs(r, i, tuple(sh), method="fft") for sh in shifts]) for r in range(4)])
        e = peak_rel_err(got[i], ref)
        assert e < MFMA_TOL, (i, e)
    lib.kill()
    # astigmatic CTF + planted shifted winners: rescoring returns the exact argmax
    pf, cps = make_case(seed=50, n=64, kfrom=2, kto=20, ctf=True, astig=0.3, nrefs=6, nptcls=6, ldim=64, eo_mixed=True)
    rng = np.random.default_rng(51)
    for i in range(6):
        r = int(rng.integers(0, 6))
        s = shifts[int(rng.integers(0, 3))]
        rho = int(rng.integers(0, 2 * pf.pftsz))
        src = pf.refs_even[r] if pf.iseven[i] else pf.refs_odd[r]
        sh = O.shmat_closed(pf.grid, tuple(s.astype(np.float64)))
        full = O.full_ring(src.astype(np.complex128) * sh)
        rolled = np.roll(full, rho, axis=-1)[:, :pf.pftsz]
        noise = 0.5 * (rng.standard_normal(rolled.shape) + 1j * rng.standard_normal(rolled.shape))
        pf.ptcls[i] = (rolled + noise * np.abs(rolled).mean()).astype(np.complex64)
    lib = to_lib(pf, cps)
    cc, ir, ish, rot = lib.search_mfma(np.arange(1, 7), shifts=shifts, ntop=4)
    pcc, pir, pish, prot = lib.search(np.arange(1, 7), None, shifts)
    for i in range(6):
        ref = np.stack([np.stack([pf.gencorrs(r, i, tuple(sh)) for sh in shifts]) for r in range(6)])
        r_, s_, t_ = np.unravel_index(int(np.argmax(ref)), ref.shape)
        assert (ir[i], ish[i], rot[i]) == (r_ + 1, s_, t_ + 1)
        assert abs(cc[i] - ref[r_, s_, t_]) <= 1e-6
        assert (pir[i], pish[i], prot[i]) == (ir[i], ish[i], rot[i])
    lib.kill()
