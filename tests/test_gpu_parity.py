"""Parity of the gfx950 library (through the C ABI) with the FP64 oracle on identical inputs.

Tolerances: FP32 correlations within 1e-5 of the FP64 oracle relative to the correlation peak
(cc is normalised, |cc| <= 1); best rotation identical wherever the oracle's top-2 gap exceeds
10x the measured error; FP64 objectives and gradients within 1e-9 relative (identical FP32 inputs,
FP64 arithmetic on both sides).
"""
import numpy as np
import pytest

import pftcc_oracle as O
from cases import make_case
from gpu_helpers import peak_rel_err, say, to_lib, top2_gap
from pftcc import PftccError

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5


@pytest.mark.parametrize("n,kto,kstop,ctf,eo_mixed", [
    (16, 7, None, False, False),
    (16, 9, 8, True, True),
    (24, 10, None, True, False),
    (40, 12, 11, True, True),
    (66, 14, None, True, False),
])
def test_gencorrs_small(n, kto, kstop, ctf, eo_mixed):
    say(f"gencorrs small n={n} kto={kto} kstop={kstop} ctf={ctf}")
    pf, cps = make_case(seed=100 + n, n=n, kto=kto, kstop=kstop, ctf=ctf, nrefs=3, nptcls=4,
                        eo_mixed=eo_mixed)
    lib = to_lib(pf, cps)
    for iref in range(3):
        for i in range(4):
            got = lib.gencorrs(iref + 1, i + 1)
            ref = pf.gencorrs(iref, i)
            assert peak_rel_err(got, ref) < FP32_TOL
            for sh in ((0.5, -1.25), (-2.0, 0.75)):
                got = lib.gencorrs(iref + 1, i + 1, sh)
                ref = pf.gencorrs(iref, i, sh)
                assert peak_rel_err(got, ref) < FP32_TOL
    lib.kill()


def test_ctfmats_match_fp64_model():
    say("ctfmats vs FP64 CTF model")
    pf, cps = make_case(seed=7, n=192, kfrom=2, kto=40, ctf=True, nrefs=1, nptcls=6, ldim=128)
    lib = to_lib(pf, cps, memoize=False)
    for i, p in enumerate(cps):
        got = lib.get_ctfmat(i + 1).astype(np.float64)
        ref = O.ctf_abs_mat(pf.grid, p)
        # FP32 evaluation of sin(chi): error grows with |chi| (a few ulp of the argument)
        tol = 8 * 6e-8 * O.ctf_chi_abs(pf.grid, p) + 1e-6
        assert np.all(np.abs(got - ref) <= tol)
    lib.kill()


def test_memo_equals_reference_memo():
    """W[2t] = 2 PA[t] and W[2t+1] = 2 PB[t] with PA, PB the reference's memoize_fft spectra."""
    say("memo vs memoize_fft")
    pf, cps = make_case(seed=8, n=24, kto=8, ctf=False, nrefs=1, nptcls=3)
    lib = to_lib(pf, cps)
    for i in range(3):
        W = lib.get_memo(i + 1).astype(np.complex128)
        pa, pb = O.memoize_fft(pf.ptcls[i].astype(np.complex128))
        h = pf.pftsz // 2
        scale = np.max(np.abs(W))
        assert np.max(np.abs(W[:, 0::2] - 2 * pa[:, :h + 1])) < 2e-6 * scale
        assert np.max(np.abs(W[:, 1::2] - 2 * pb[:, :h])) < 2e-6 * scale
        assert abs(lib.get_sqsum_ptcl(i + 1) - pf.sqsum_ptcl(i)) < 1e-6 * pf.sqsum_ptcl(i)
    lib.kill()


@pytest.mark.parametrize("n,kfrom,kto,ldim", [(192, 2, 40, 128), (308, 2, 60, 200), (400, 2, 80, 256)])
def test_gencorrs_config_shapes(n, kfrom, kto, ldim):
    """The three ring/shell shapes of the benchmark configurations (box 128/200/256)."""
    say(f"gencorrs config shape n={n} K=[{kfrom},{kto}]")
    pf, cps = make_case(seed=n, n=n, kfrom=kfrom, kto=kto, ctf=True, nrefs=2, nptcls=2, ldim=ldim)
    lib = to_lib(pf, cps)
    errs = []
    for iref in range(2):
        for i in range(2):
            for sh in (None, (1.5, -2.5)):
                got = lib.gencorrs(iref + 1, i + 1, sh)
                ref = pf.gencorrs(iref, i, sh, method="fft")
                e = peak_rel_err(got, ref)
                errs.append(e)
                assert e < FP32_TOL
                if top2_gap(ref) > 10 * e * np.max(np.abs(ref)):
                    assert int(np.argmax(got)) == int(np.argmax(ref))
    say(f"  max peak-relative error {max(errs):.2e}")
    lib.kill()


def test_planted_rotation_and_shift_on_gpu():
    say("planted rotation + shift")
    pf, _ = make_case(seed=9, n=64, kto=20, ctf=False, nrefs=1, nptcls=1, ldim=64)
    full = O.full_ring(pf.refs_even[0].astype(np.complex128))
    rho0, s0 = 23, (1.0, -0.5)
    rolled = np.roll(full, rho0, axis=-1)[:, :64]
    # the particle carries the phase of s0 in the frame of the rotated reference
    sh = O.shmat_closed(pf.grid, s0)
    sh_full = np.concatenate([sh, np.conj(sh)], axis=-1)
    pf.ptcls[0] = (rolled * np.roll(sh_full, rho0, axis=-1)[:, :64]).astype(np.complex64)
    lib = to_lib(pf, None)
    cc = lib.gencorrs(1, 1, s0)
    assert int(np.argmax(cc)) == rho0
    assert abs(cc[rho0] - 1.0) < 2e-6
    lib.kill()


def test_batch_equals_per_pair_and_search():
    say("gencorrs_batch / search vs per-pair gencorrs")
    pf, cps = make_case(seed=11, n=40, kto=12, ctf=True, nrefs=5, nptcls=6, eo_mixed=True)
    lib = to_lib(pf, cps)
    shifts = np.array([[0.0, 0.0], [1.0, -0.5], [-1.5, 2.0]], np.float32)
    ids = np.arange(1, 7)
    full = lib.gencorrs_batch(ids, None, shifts)
    for p in range(6):
        for r in range(5):
            for s in range(3):
                one = lib.gencorrs(r + 1, p + 1, shifts[s])
                assert np.array_equal(full[p, r, s], one)
    cc, ir, ish, rot = lib.search(ids, None, shifts)
    for p in range(6):
        v = full[p]
        flat = int(np.argmax(v))          # first maximum in (ref, shift, rot) order
        r, s, t = np.unravel_index(flat, v.shape)
        assert (ir[p], ish[p], rot[p]) == (r + 1, s, t + 1)
        assert cc[p] == v[r, s, t]
    # unshifted search
    full0 = lib.gencorrs_batch(ids)
    cc0, ir0, _, rot0 = lib.search(ids)
    for p in range(6):
        r, _, t = np.unravel_index(int(np.argmax(full0[p])), full0[p].shape)
        assert (ir0[p], rot0[p]) == (r + 1, t + 1)
    lib.kill()


@pytest.mark.parametrize("match_filt", [False, True])
def test_for_rot_8_and_gradients(match_filt):
    say(f"gencorr_for_rot_8 / grad match_filt={match_filt}")
    pf, cps = make_case(seed=13, n=48, kto=14, ctf=True, nrefs=3, nptcls=3, match_filt=match_filt,
                        eo_mixed=True)
    lib = to_lib(pf, cps)
    for (r, i, sh, irot) in [(0, 0, (0.3, -0.8), 1), (1, 2, (-2.1, 1.7), 49), (2, 1, (1.0, 1.0), 96),
                             (2, 2, (0.0, 0.0), 50)]:
        f = lib.gencorr_for_rot_8(r + 1, i + 1, sh, irot)
        fo = pf.gencorr_for_rot_8(r, i, sh, irot)
        assert abs(f - fo) <= 1e-9 * max(1.0, abs(fo)) + (1e-7 if match_filt else 0.0)
        f2, g = lib.gencorr_grad_for_rot_8(r + 1, i + 1, sh, irot)
        fo2, go = pf.gencorr_grad_for_rot_8(r, i, sh, irot)
        assert abs(f2 - f) < 1e-14
        tol = 1e-7 if match_filt else 1e-9
        assert np.all(np.abs(g - go) <= tol * np.maximum(1.0, np.abs(go)))
        g2 = lib.gencorr_grad_only_for_rot_8(r + 1, i + 1, sh, irot)
        assert np.array_equal(g2, g)
        f32 = lib.gencorr_for_rot(r + 1, i + 1, np.float32(sh), irot)
        assert abs(f32 - fo) < 1e-6
    lib.kill()


def test_grad_batch_with_angle_derivative():
    say("grad_batch (shift + in-plane angle)")
    pf, cps = make_case(seed=17, n=32, kto=12, ctf=True, nrefs=4, nptcls=3)
    lib = to_lib(pf, cps)
    rng = np.random.default_rng(3)
    npairs = 12
    refs = rng.integers(1, 5, npairs)
    pts = rng.integers(1, 4, npairs)
    shs = rng.uniform(-4, 4, (npairs, 2))
    rots = rng.integers(1, 65, npairs)
    f, g = lib.grad_batch(refs, pts, shs, rots)
    for t in range(npairs):
        fo, go = pf.gencorr_grad_for_rot_8(refs[t] - 1, pts[t] - 1, shs[t], rots[t])
        assert abs(f[t] - fo) <= 1e-9 * max(1, abs(fo))
        assert np.all(np.abs(g[t, :2] - go) <= 1e-9 * np.maximum(1, np.abs(go)))
        psi = (rots[t] - 1) * np.pi / pf.pftsz
        ga = pf.grad_angle(refs[t] - 1, pts[t] - 1, shs[t], psi)
        # the angular derivative rings are FP32 (spectral derivative of FP32 rings)
        assert abs(g[t, 2] - ga) <= 2e-5 * max(1.0, abs(ga))
    lib.kill()


def test_euclid_objective():
    say("objfun euclid")
    pf, cps = make_case(seed=19, n=32, kto=10, ctf=True, nrefs=2, nptcls=2, euclid=True)
    lib = to_lib(pf, cps)
    for r in range(2):
        for i in range(2):
            got = lib.gencorrs(r + 1, i + 1)
            ref = pf.gencorrs_euclid(r, i, None, method="fft")
            assert peak_rel_err(got, ref) < FP32_TOL
            got = lib.gencorrs(r + 1, i + 1, (0.5, -1.0))
            ref = pf.gencorrs_euclid(r, i, (0.5, -1.0), method="fft")
            assert peak_rel_err(got, ref) < FP32_TOL
            f, g = lib.gencorr_grad_for_rot_8(r + 1, i + 1, (0.25, 0.5), 7)
            fo, go = pf.euclid_grad_for_rot_8(r, i, (0.25, 0.5), 7)
            assert abs(f - fo) <= 1e-9 * max(1, abs(fo))
            assert np.all(np.abs(g - go) <= 1e-8 * np.maximum(1, np.abs(go)))
            sc, ps, rs = lib.gencorr_sigma_contrib(r + 1, i + 1, (0.25, 0.5), 7)
            so = pf.sigma_contrib(r, i, (0.25, 0.5), 7)
            assert np.allclose(sc, so, rtol=2e-5)
    lib.kill()


def test_specscore_and_corrmat():
    say("specscore / calc_roinv_corrmat")
    pf, cps = make_case(seed=23, n=32, kto=12, ctf=True, nrefs=4, nptcls=4)
    lib = to_lib(pf, cps)
    for r, i, irot in [(0, 1, 5), (3, 2, 40)]:
        s = lib.specscore(r + 1, i + 1, irot)
        assert abs(s - pf.specscore(r, i, irot)) < 1e-5
    cm = lib.calc_roinv_corrmat()
    for r in range(4):
        for p in range(4):
            if r < p:
                assert abs(cm[r, p] - pf.gencorrs(r, p).max()) < FP32_TOL
                assert cm[p, r] == cm[r, p]
            elif r == p:
                assert cm[r, p] == 1.0
    lib.kill()


def test_setters_getters_and_errors():
    say("setters / getters / error paths")
    pf, cps = make_case(seed=29, n=16, kto=8, ctf=False, nrefs=3, nptcls=4)
    lib = to_lib(pf, None, pfrom=11)
    assert lib.get_nrots() == 32 and lib.get_pftsz() == 16 and lib.get_nrefs() == 3
    assert lib.get_nptcls() == 4
    assert np.array_equal(lib.get_ptcl_pft(12), pf.ptcls[1])
    assert np.array_equal(lib.get_ref_pft(2, False), pf.refs_odd[1])
    assert abs(lib.get_rot(9) - 90.0) < 1e-4
    assert lib.get_roind(90.2) == 9 and lib.get_roind(359.9) == 1
    xy = lib.get_coord(9, 5)
    assert np.allclose(xy, [5.0, 0.0], atol=1e-5)
    with pytest.raises(PftccError):
        lib.get_rot(33)
    with pytest.raises(PftccError):
        lib.gencorrs(4, 11)
    with pytest.raises(PftccError):
        lib.gencorrs(1, 10)
    # eo handling: odd particles correlate with the odd references
    lib.set_eo(12, False)
    assert not lib.ptcl_iseven(12)
    pf.iseven[1] = False
    assert peak_rel_err(lib.gencorrs(2, 12), pf.gencorrs(1, 1)) < FP32_TOL
    lib.swap_ptclsevenodd()
    pf.iseven = ~pf.iseven
    assert peak_rel_err(lib.gencorrs(3, 13), pf.gencorrs(2, 2)) < FP32_TOL
    lib.set_eos(np.zeros(4, np.int32))
    pf.iseven[:] = True
    lib.cp_even2odd_ref(1)
    assert np.array_equal(lib.get_ref_pft(1, False), pf.refs_even[0])
    # a changed particle invalidates the memo until memoize_ffts
    lib.cp_even_ref2ptcl(1, 11)
    with pytest.raises(PftccError):
        lib.gencorrs(1, 11)
    lib.memoize_ffts()
    cc = lib.gencorrs(1, 11)
    assert abs(cc[0] - 1.0) < 2e-6 and int(np.argmax(cc)) == 0
    # reallocate_ptcls re-indexes the particle slots (:509-555)
    lib.reallocate_ptcls(2, np.array([40, 37], np.int32))
    assert lib.get_nptcls() == 2
    lib.set_ptcl_pft(40, pf.ptcls[0])
    lib.set_ptcl_pft(37, pf.ptcls[3])
    lib.memoize_ffts()
    assert peak_rel_err(lib.gencorrs(1, 37), pf.gencorrs(0, 3)) < FP32_TOL
    lib.kill()


def test_grid_matches_oracle():
    """The host-built FP32 polar grid (new :261-274) equals the oracle's correctly rounded grid
    (at most rare 1-ulp differences from libm's sinf/cosf)."""
    say("polar grid vs oracle")
    pf, _ = make_case(seed=31, n=308, kfrom=2, kto=60, ctf=False, nrefs=1, nptcls=1, ldim=200)
    lib = to_lib(pf, None, memoize=False)
    g = pf.grid
    mism = 0
    for k in (2, 17, 60):
        for irot in range(1, g.nrots + 1):
            xy = lib.get_coord(irot, k)
            ref = np.array([g.polar_x[k - 2, irot - 1], g.polar_y[k - 2, irot - 1]], np.float32)
            ulp = np.abs(xy.view(np.int32) - ref.view(np.int32))
            assert np.all(ulp <= 1)
            mism += int(np.sum(ulp))
    assert mism <= 0.01 * 3 * 2 * g.nrots
    for irot in (1, 2, 155, 616):
        assert lib.get_rot(irot) == g.angtab_deg[irot - 1]
    lib.kill()


@pytest.mark.parametrize("n,kto,ldim", [(16, 8, 32), (48, 14, 48), (192, 30, 128), (308, 30, 200)])
def test_fused_and_generic_kernels_agree(n, kto, ldim, monkeypatch):
    """The compile-time-size kernel and the runtime-size kernel compute the same correlations
    (different radix orders, so agreement is to FP32 rounding, not bitwise)."""
    say(f"fused vs generic kernel n={n}")
    pf, cps = make_case(seed=200 + n, n=n, kto=kto, ctf=True, nrefs=3, nptcls=3, ldim=ldim)
    shifts = np.array([[0.0, 0.0], [0.5, -1.0]], np.float32)
    lib = to_lib(pf, cps)
    fused = lib.gencorrs_batch(np.arange(1, 4), None, shifts)
    lib.kill()
    monkeypatch.setenv("PFTCC_GENERIC", "1")
    lib = to_lib(pf, cps)
    gen = lib.gencorrs_batch(np.arange(1, 4), None, shifts)
    lib.kill()
    assert np.max(np.abs(fused - gen)) < 2e-6
    for p in range(3):
        for r in range(3):
            for s in range(2):
                ref = pf.gencorrs(r, p, tuple(shifts[s]) if s else None, method="fft")
                assert peak_rel_err(fused[p, r, s], ref) < FP32_TOL


@pytest.mark.parametrize("pssnr,ctf", [(False, True), (False, False), (True, True)])
def test_prep_matchfilt_matches_oracle(pssnr, ctf):
    """prep_matchfilt on the device against the oracle restatement (FP32 on the device, FP64 in the
    oracle: particle rings within 2e-6 relative per shell); the filtered particles are then searched
    with the shell-normalised references exactly as the oracle's gencorrs."""
    say(f"prep_matchfilt pssnr={pssnr} ctf={ctf}")
    pf, cps = make_case(seed=91, n=48, kfrom=2, kto=16, kstop=14, ctf=ctf, nrefs=3, nptcls=5, ldim=48,
                        match_filt=True, eo_mixed=True)
    pf.pssnr = pssnr
    lib = to_lib(pf, cps, memoize=False)
    ids = [1, 2, 3, 4, 5]
    refs = [2, 0, 3, 1, 1]
    rots = [1, 17, 49, 96, 60]
    lib.prep_matchfilt(ids[0], refs[0], rots[0])
    lib.prep_matchfilt_batch(ids[1:], refs[1:], rots[1:])
    for t, i in enumerate(ids):
        want = pf.prep_matchfilt(i - 1, refs[t] - 1, rots[t]) if refs[t] else pf.ptcls[i - 1].astype(np.complex128)
        got = lib.get_ptcl_pft(i).astype(np.complex128)
        scale = np.max(np.abs(want), axis=-1, keepdims=True)
        assert np.max(np.abs(got - want) / scale) < 2e-6, t
        pf.ptcls[i - 1] = got.astype(np.complex64)      # the oracle continues from the device's particles
        assert abs(lib.get_sqsum_ptcl(i) - pf.sqsum_ptcl(i - 1)) <= 1e-6 * pf.sqsum_ptcl(i - 1)
    for i in range(5):
        for r in range(3):
            assert peak_rel_err(lib.gencorrs(r + 1, i + 1), pf.gencorrs(r, i)) < FP32_TOL
    with pytest.raises(PftccError):
        lib.prep_matchfilt_batch([1, 1], [1, 2], [1, 1])
    lib.kill()


@pytest.mark.parametrize("mode", ["euclid", "match_filt"])
def test_fused_prime_factor_objectives(mode):
    """Ring length 308 runs the prime-factor shell transform of the fused kernel (308 = 4 x 11 x 7).
    Its Euclid branch (shell weights staged with the particle spectrum, the offset sums of pass 1)
    and its shell-normalised branch (per-pair reference scales in pass 1) against the FP64 oracle,
    unshifted and shifted, over a batch whose last tile is partial and whose particles fall in both
    eo classes."""
    say(f"fused prime-factor kernel n=308 {mode}")
    pf, cps = make_case(seed=57, n=308, kto=24, ctf=True, nrefs=4, nptcls=5, ldim=200, eo_mixed=True,
                        euclid=(mode == "euclid"), match_filt=(mode == "match_filt"))
    lib = to_lib(pf, cps)
    shifts = np.array([[0.0, 0.0], [1.5, -0.5], [-2.0, 1.0]], np.float32)
    got = lib.gencorrs_batch(np.arange(1, 6), None, shifts)
    for p in range(5):
        for r in range(4):
            for s in range(3):
                sh = tuple(shifts[s]) if s else None
                ref = pf.gencorrs_euclid(r, p, sh, method="fft") if mode == "euclid" else \
                    pf.gencorrs(r, p, sh, method="fft")
                assert peak_rel_err(got[p, r, s], ref) < FP32_TOL, (p, r, s)
    for p, r in [(0, 0), (4, 3)]:
        ref = pf.gencorrs_euclid(r, p, None, method="fft") if mode == "euclid" else pf.gencorrs(r, p, method="fft")
        assert peak_rel_err(lib.gencorrs(r + 1, p + 1), ref) < FP32_TOL
    lib.kill()


@pytest.mark.parametrize("mode", ["cc", "euclid"])
@pytest.mark.parametrize("n,kto,ldim", [(192, 30, 128), (240, 30, 160), (400, 30, 256), (500, 30, 320),
                                        (600, 30, 384)])
def test_per_pair_kernel_reference_prefetch(n, kto, ldim, mode):
    """At ring lengths 192, 240, 400, 500 and 600 (radices 4, 2, 3, 5, 5) the per-pair variant of the fused kernel loads its pass-1
    reference values one shell ahead (at 400 with the middle passes' twiddle tables).  Batches of 20 particles x 7 references (the last tile
    partial, the particle groups of 16 wrapping), unshifted and with two shifts, against the FP64
    oracle; the search returns the batch's first maximum."""
    say(f"per-pair kernel with reference prefetch n={n} {mode}")
    np_, nr = 20, 7
    pf, cps = make_case(seed=300 + n, n=n, kto=kto, ctf=True, nrefs=nr, nptcls=np_, ldim=ldim, eo_mixed=True,
                        euclid=(mode == "euclid"))
    lib = to_lib(pf, cps)
    ids = np.arange(1, np_ + 1)
    shifts = np.array([[0.0, 0.0], [1.5, -2.0]], np.float32)
    for shv in (None, shifts):
        got = lib.gencorrs_batch(ids, None, shv)
        for p in range(np_):
            for r in range(nr):
                for s in range(got.shape[2]):
                    sh = tuple(shv[s]) if (shv is not None and s) else None
                    ref = pf.gencorrs_euclid(r, p, sh, method="fft") if mode == "euclid" else \
                        pf.gencorrs(r, p, sh, method="fft")
                    assert peak_rel_err(got[p, r, s], ref) < FP32_TOL, (p, r, s)
        cc, ir, ish, rot = lib.search(ids, None, shv)
        for p in range(np_):
            r, s, t = np.unravel_index(int(np.argmax(got[p])), got[p].shape)
            assert (ir[p], ish[p], rot[p]) == (r + 1, s, t + 1)
            assert cc[p] == got[p, r, s, t]
    lib.kill()
