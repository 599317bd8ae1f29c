"""Pins the FP64 oracle (oracle/pftcc_oracle.py) without the reference binary.

The reference ships no golden vectors for this path, so each oracle function is checked
against something a plausible mistake would break: an independent formulation of the same
quantity (FFT route vs brute-force direct sums, Euclid via correlations vs via differences),
a closed form (planted rotations and shifts give cc = 1, amplitude-contrast-only CTF equals
the amplitude fraction, relativistic electron wavelength), hand-derived integer tables, or
central finite differences.
"""
import math

import numpy as np
import pytest

import pftcc_oracle as O
from cases import make_case, random_pfts


# --------------------------------------------------------------------------------- a3 / a2 ---
@pytest.mark.parametrize("box,pftsz", [(128, 192), (200, 308), (256, 400), (64, 100)])
def test_magic_pftsz_known_boxes(box, pftsz):
    """Hand-derived: ring2 = round2even(box/2-2); pi*ring2 -> nearest FFT-friendly size.

    box 128: ring2 62, pi*62 = 194.8 -> 192 (|192-195|=3 < |200-195|=5).
    box 200: ring2 98, pi*98 = 307.9 -> 308.  box 256: ring2 126, pi*126 = 395.8 -> 400.
    box 64: ring2 30, pi*30 = 94.2 -> tie between 88 and 100 resolves to the larger, 100.
    """
    assert O.magic_pftsz(O.default_ring2(box)) == pftsz


def test_round2even_hand_values():
    assert O.round2even(194.78) == 194      # nint 195 is odd; 194 is closer
    assert O.round2even(3.0) == 2           # tie 2 vs 4 goes down (<=)
    assert O.round2even(5.5) == 6           # nint 6 even
    assert O.round2even(62.0) == 62


def test_boxmatch_hand_values():
    # box 256 with msk 62: 2*(62+10) = 144 < 256 -> nearest magic box to 144 is 144
    assert O.find_boxmatch(256, 62.0) == 144
    assert O.find_boxmatch(128, 62.0) == 128   # 128 <= 144 keeps the box


def test_grid_closed_form():
    g = O.make_grid(192, 2, 40, 128)
    rho = np.arange(g.nrots)
    ang = 2 * np.pi * rho / g.nrots
    ks = g.ks[:, None]
    assert np.allclose(g.polar_x, ks * np.sin(ang)[None, :], atol=4e-5 * 40)
    assert np.allclose(g.polar_y, -ks * np.cos(ang)[None, :], atol=4e-5 * 40)
    assert np.allclose(g.angtab_deg, rho * 360.0 / g.nrots, atol=2e-4)
    # argtransf = polar * pi/(ldim/2)
    assert np.allclose(g.argtransf_x, g.polar_x[:, :192].astype(np.float64) * np.pi / 64, rtol=1e-12)
    for irot in (1, 2, 97, 384):
        assert O.get_roind(g, float(g.angtab_deg[irot - 1])) == irot
    assert O.get_roind(g, 359.99) == 1


def test_pxls_p_shell_hand_and_area():
    # h,k in {1,2}: (1,1)->1, (1,2),(2,1)->2, (2,2)->3 (outside [1,2])
    assert list(O.pxls_p_shell(1, 2)) == [1.0, 2.0]
    p = O.pxls_p_shell(2, 60)
    k = np.arange(2, 61)
    # a quarter ring of radius k holds about pi*k/2 lattice points
    sel = k >= 20
    assert np.all(np.abs(p[sel] / (np.pi * k[sel] / 2) - 1) < 0.15)


# -------------------------------------------------------------------------------------- a6 ---
def test_ctf_wavelength_300kv():
    """Relativistic electron wavelength at 300 kV is 0.019687 A (textbook, 12.2643/sqrt(V(1+
    0.97848e-6 V))); the reference's constant 12.26 makes it 3.1e-4 smaller.  Dropping the
    relativistic term would be 14 % off, so rtol 5e-4 pins the formula."""
    g = O.make_grid(16, 2, 7, 32)
    # Recover wl from a pure-defocus CTF: chi = pi*wl*s2*df when Cs = fraca = 0.
    p = O.CtfParams(smpd=1.0, kv=300.0, cs=0.0, fraca=0.0, dfx=1e-4, dfy=1e-4, angast=0.0)
    v = O.ctf_abs_mat(g, p)
    s2 = (g.polar_x[:, :16].astype(float) / 32) ** 2 + (g.polar_y[:, :16].astype(float) / 32) ** 2
    df = 1e-4 * 1e4
    wl = np.arcsin(v) / (np.pi * s2 * df)
    assert np.allclose(wl, 0.019687, rtol=5e-4)


def test_ctf_amplitude_contrast_only():
    """df = Cs = 0: |sin(atan(fa/sqrt(1-fa^2)))| = fa everywhere."""
    g = O.make_grid(24, 2, 9, 40)
    p = O.CtfParams(smpd=1.2, kv=200.0, cs=0.0, fraca=0.07, dfx=0.0, dfy=0.0, angast=33.0)
    assert np.allclose(O.ctf_abs_mat(g, p), 0.07, atol=1e-12)


def test_ctf_phase_plate_quarter_wave():
    g = O.make_grid(24, 2, 9, 40)
    p = O.CtfParams(smpd=1.2, kv=200.0, cs=0.0, fraca=0.0, dfx=0.0, dfy=0.0, angast=0.0,
                    phshift=np.pi / 2, phaseplate=True)
    assert np.allclose(O.ctf_abs_mat(g, p), 1.0, atol=1e-12)


def test_ctf_isotropic_is_ring_constant_and_astig_axes():
    g = O.make_grid(32, 2, 12, 48)
    iso = O.CtfParams(smpd=1.3, kv=300, cs=2.7, fraca=0.1, dfx=2.0, dfy=2.0, angast=17.0)
    v = O.ctf_abs_mat(g, iso)
    # the FP32 grid radius varies by ~1 ulp around a ring; |dCTF| <= |d chi| ~ 2 eps32 chi
    tol = 4 * 6e-8 * O.ctf_chi_abs(g, iso).max()
    assert np.allclose(v, v[:, :1], atol=tol)
    # astigmatic, angast = 0: along theta = pi/2 (x axis) df = dfx; along theta = 0 (y axis) dfy
    ast = O.CtfParams(smpd=1.3, kv=300, cs=2.7, fraca=0.1, dfx=2.5, dfy=1.5, angast=0.0)
    va = O.ctf_abs_mat(g, ast)
    vx = O.ctf_abs_mat(g, O.CtfParams(1.3, 300, 2.7, 0.1, 2.5, 2.5, 0.0))
    vy = O.ctf_abs_mat(g, O.CtfParams(1.3, 300, 2.7, 0.1, 1.5, 1.5, 0.0))
    assert np.allclose(va[:, 16], vx[:, 16], atol=1e-5)   # j = n/2 -> theta = 90 deg
    assert np.allclose(va[:, 0], vy[:, 0], atol=1e-5)     # j = 0   -> theta = 0


# ------------------------------------------------------------------------------------- a11 ---
def test_shmat_recurrence_matches_closed_form():
    """De Moivre: shmat(:,k) = exp(i(arg_kfrom + (k-kfrom) arg_1)) exactly (to FP64 rounding);
    against exp(i*argtransf) it differs only by the FP32 rounding of polar = sin*k."""
    g = O.make_grid(64, 2, 30, 100)
    for sh in ([1.5, -2.25], [0.0, 3.0], [-4.0, 0.5]):
        arg0 = g.argtransf_x[0] * sh[0] + g.argtransf_y[0] * sh[1]
        arg1 = g.shellone_x * sh[0] + g.shellone_y * sh[1]
        steps = np.arange(g.nk)[:, None]
        exact = np.exp(1j * (arg0[None, :] + steps * arg1[None, :]))
        assert np.max(np.abs(O.shmat_8(g, sh) - exact)) < 1e-12
        argmax = np.max(np.abs(g.argtransf_x) * abs(sh[0]) + np.abs(g.argtransf_y) * abs(sh[1]))
        assert np.max(np.abs(O.shmat_8(g, sh) - O.shmat_closed(g, sh))) < 4 * 6e-8 * argmax


# -------------------------------------------------------------------------- a9 / a10 / a12 ---
@pytest.mark.parametrize("n", [8, 12, 16, 24, 40])
def test_fourcase_equals_modular_index(n):
    rng = np.random.default_rng(n)
    q = random_pfts(rng, 1, 3, n)[0].astype(np.complex128)
    p = random_pfts(rng, 1, 3, n)[0]
    for irot in range(1, 2 * n + 1):
        assert np.allclose(O.corrk_for_rot(q, p, irot), O.corrk_for_rot_fourcase(q, p, irot),
                           rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n,ctf,shift", [(8, False, None), (16, True, None), (24, True, (1.0, -0.5)),
                                         (40, True, (-2.0, 1.25)), (192, True, None)])
def test_fft_route_equals_direct_sum(n, ctf, shift):
    """calc_corrs_over_k's FFT identity (4 conj(A)PA / 4 conj(B)PB, c2r_2n, /(2 nrots)) equals
    the brute-force rotational sum; fixes the scaling and the conjugation convention."""
    kto = 7 if n < 100 else 20
    pf, _ = make_case(seed=n, n=n, kto=kto, ctf=ctf, nrefs=2, nptcls=2, ldim=4 * kto)
    for iref in range(2):
        for i in range(2):
            a = pf.gencorrs(iref, i, shift, method="direct")
            b = pf.gencorrs(iref, i, shift, method="fft")
            assert np.allclose(a, b, rtol=1e-11, atol=1e-13)


def _plant_rotation(ref_half, rho0):
    """Particle = reference rotated so that P[j + rho0] = ref_full[j]."""
    n = ref_half.shape[-1]
    full = O.full_ring(ref_half.astype(np.complex128))
    rolled = np.roll(full, rho0, axis=-1)
    return rolled[..., :n].astype(np.complex64)


@pytest.mark.parametrize("rho0", [0, 1, 5, 16, 17, 31])
def test_planted_rotation_gives_unit_correlation(rho0):
    pf, _ = make_case(seed=3, n=16, kto=9, ctf=False, nrefs=1, nptcls=1)
    pf.ptcls[0] = _plant_rotation(pf.refs_even[0], rho0)
    cc = pf.gencorrs(0, 0)
    assert abs(cc[rho0] - 1.0) < 1e-6
    assert int(np.argmax(cc)) == rho0
    assert np.all(np.abs(cc) <= 1 + 1e-12)                      # Cauchy-Schwarz


def test_rotation_equivariance():
    pf, _ = make_case(seed=5, n=24, kto=8, ctf=False, nrefs=1, nptcls=1)
    cc0 = pf.gencorrs(0, 0)
    base = pf.ptcls[0].copy()
    full = O.full_ring(base.astype(np.complex128))
    for d in (1, 7, 30):
        pf.ptcls[0] = np.roll(full, d, axis=-1)[..., :24].astype(np.complex64)
        assert np.allclose(pf.gencorrs(0, 0), np.roll(cc0, d), atol=1e-6)
    pf.ptcls[0] = base


def test_shift_recovery_invariant():
    """simple_test_shiftsrch's invariant: a particle carrying the phase of shift s0 is matched
    best by gencorrs with shvec = s0 (the shift is applied to the reference, :1529)."""
    pf, _ = make_case(seed=7, n=32, kto=12, ctf=False, nrefs=1, nptcls=1, ldim=40)
    s0 = (1.0, -1.0)
    pf.ptcls[0] = (pf.refs_even[0].astype(np.complex128) *
                   O.shmat_closed(pf.grid, s0)).astype(np.complex64)
    best = max(((pf.gencorrs(0, 0, (x, y)).max(), (x, y))
                for x in (-2.0, -1.0, 0.0, 1.0, 2.0) for y in (-2.0, -1.0, 0.0, 1.0, 2.0)))
    assert best[1] == s0 and abs(best[0] - 1.0) < 1e-5


# ------------------------------------------------------------------------------- a13 / a14 ---
@pytest.mark.parametrize("match_filt", [False, True])
def test_for_rot_8_consistent_with_gencorrs(match_filt):
    """With match_filt (unit weights, memoised S_p) and optlp = 1, gencorr_for_rot_8 equals the
    shifted gencorrs at irot when kstop == kto (both normalise over the same shells)."""
    pf, _ = make_case(seed=11, n=16, kto=8, ctf=True, nrefs=2, nptcls=2, match_filt=match_filt)
    if match_filt:
        pf.ref_optlp[:] = 1.0
    sh = (0.7, -1.3)
    cc = pf.gencorrs(1, 1, sh)
    for irot in (1, 5, 17, 30):
        v = pf.gencorr_for_rot_8(1, 1, sh, irot)
        if match_filt:
            # shellnorm with optlp=1 rescales each shell; recompute gencorrs with it applied
            q = pf.prep_ref(1, 1, sh)
            c = O.corrk_for_rot(q, pf.ptcls[1], irot).sum()
            ref = c / math.sqrt(np.sum(np.abs(q) ** 2) * pf.sqsum_ptcl(1))
            assert abs(v - ref) < 1e-12
        else:
            assert np.isfinite(v) and abs(v) <= 1.0
    # the planted case: weights k give exactly 1 as well
    pf2, _ = make_case(seed=12, n=16, kto=8, ctf=False, nrefs=1, nptcls=1)
    pf2.ptcls[0] = _plant_rotation(pf2.refs_even[0], 9)
    assert abs(pf2.gencorr_for_rot_8(0, 0, (0.0, 0.0), 10) - 1.0) < 1e-6


@pytest.mark.parametrize("match_filt", [False, True])
def test_shift_gradient_matches_central_fd(match_filt):
    pf, _ = make_case(seed=21, n=24, kto=10, ctf=True, nrefs=2, nptcls=2, match_filt=match_filt)
    h = 1e-3
    for (iref, i, sh, irot) in [(0, 0, (0.3, -0.8), 4), (1, 1, (-2.1, 1.7), 37)]:
        f, g = pf.gencorr_grad_for_rot_8(iref, i, sh, irot)
        assert abs(f - pf.gencorr_for_rot_8(iref, i, sh, irot)) < 1e-13
        for d in range(2):
            sp = list(sh); sm = list(sh)
            sp[d] += h; sm[d] -= h
            fd = (pf.gencorr_for_rot_8(iref, i, sp, irot) -
                  pf.gencorr_for_rot_8(iref, i, sm, irot)) / (2 * h)
            assert abs(fd - g[d]) <= 1e-6 * max(1.0, abs(g[d]))


# ------------------------------------------------------------------------------------- a15 ---
def test_angle_interpolant_and_derivative():
    pf, _ = make_case(seed=31, n=16, kto=8, ctf=True, nrefs=1, nptcls=1)
    sh = (0.4, 0.2)
    for irot in (1, 6, 20):
        psi = (irot - 1) * np.pi / 16
        assert abs(pf.corr_at_angle(0, 0, sh, psi) - pf.gencorr_for_rot_8(0, 0, sh, irot)) < 1e-12
    h = 1e-5
    for psi in (0.0, 0.37, 2.0, 5.5):
        fd = (pf.corr_at_angle(0, 0, sh, psi + h) - pf.corr_at_angle(0, 0, sh, psi - h)) / (2 * h)
        assert abs(fd - pf.grad_angle(0, 0, sh, psi)) < 1e-7


# ------------------------------------------------------------------------------------- a16 ---
def test_euclid_two_formulations_agree():
    """gencorrs_euclid (via correlations and square sums) equals calc_euclid_for_rot_8's direct
    differences once the latter's extra factor k is divided out through sigma2."""
    pf, _ = make_case(seed=41, n=16, kto=8, ctf=True, nrefs=2, nptcls=2, euclid=True)
    e_fft = pf.gencorrs_euclid(1, 0, None, method="fft")
    e_dir = pf.gencorrs_euclid(1, 0, None, method="direct")
    assert np.allclose(e_fft, e_dir, rtol=1e-11, atol=1e-12)
    s2 = pf.sigma2.copy()
    pf.sigma2 = (s2 * np.arange(pf.kfrom, pf.grid.kto + 1)[None, :]).astype(np.float64)
    for irot in (1, 9, 17, 32):
        assert abs(pf.euclid_for_rot_8(1, 0, (0.0, 0.0), irot) -
                   e_dir[irot - 1] * 1.0) < 1e-9 * max(1, abs(e_dir[irot - 1]))
    pf.sigma2 = s2


def test_euclid_gradient_fd():
    pf, _ = make_case(seed=43, n=16, kto=8, ctf=True, nrefs=1, nptcls=1, euclid=True)
    sh, irot, h = (0.5, -0.25), 7, 1e-4
    f, g = pf.euclid_grad_for_rot_8(0, 0, sh, irot)
    for d in range(2):
        sp = list(sh); sm = list(sh)
        sp[d] += h; sm[d] -= h
        fd = (pf.euclid_for_rot_8(0, 0, sp, irot) - pf.euclid_for_rot_8(0, 0, sm, irot)) / (2 * h)
        assert abs(fd - g[d]) <= 1e-6 * max(1.0, abs(g[d]))


def test_sigma_contrib_zero_for_identical():
    pf, _ = make_case(seed=44, n=16, kto=8, ctf=False, nrefs=1, nptcls=1, euclid=True)
    pf.ptcls[0] = pf.refs_even[0]
    assert np.max(np.abs(pf.sigma_contrib(0, 0, (0.0, 0.0), 1))) < 1e-10


# ------------------------------------------------------------------------------------- a18 ---
def test_specscore_planted_and_bounds():
    pf, _ = make_case(seed=51, n=16, kto=9, ctf=False, nrefs=1, nptcls=1)
    pf.ptcls[0] = _plant_rotation(pf.refs_even[0], 3)
    assert abs(pf.specscore(0, 0, 4) - 1.0) < 1e-6
    pf2, _ = make_case(seed=52, n=16, kto=9, ctf=True, nrefs=1, nptcls=1)
    frc = pf2.genfrc(0, 0, 5)
    assert np.all(np.abs(frc) <= 1 + 1e-12)
    assert pf2.specscore(0, 0, 5) >= 0.0


# ------------------------------------------------- parity hazards: closed-form pins -----------
def _single_bin_case(n=16, kfrom=2, kto=8, ldim=32):
    """A case with no CTF whose reference holds one non-zero value on each of two shells."""
    pf, _ = make_case(seed=61, n=n, kfrom=kfrom, kto=kto, ctf=False, nrefs=1, nptcls=1, ldim=ldim)
    pf.refs_even[:] = 0
    return pf


def _phase(k, j, n, ldim, d):
    """Closed-form shift phase of ring position j on shell k: (x dx + y dy) pi/(ldim/2) with
    x = k sin(pi j/n), y = -k cos(pi j/n) (new :261-317, gen_shmat_8 :1098-1120)."""
    th = math.pi * j / n
    return (k * math.sin(th) * d[0] - k * math.cos(th) * d[1]) * math.pi / (ldim / 2)


@pytest.mark.parametrize("s", [(0.0, 0.0), (0.4, -0.7), (-1.3, 0.9)])
def test_k_weighted_objective_and_gradient_closed_form(s):
    """gencorr_cc_for_rot_8 / gencorr_cc_grad_for_rot_8 without match filter weight shell k by k
    in the numerator and both denominators (:1717-1727, :1902-1924).  With the reference non-zero
    only at (k1, j1) = a and (k2, j2) = b, and the particle equal to the reference shifted by s0 on
    k1 and to its negative on k2, at irot = 1 and shift s (d = s - s0):
        f = (k1 |a|^2 cos(phi1(d)) - k2 |b|^2 cos(phi2(d))) / (k1 |a|^2 + k2 |b|^2)
        df/dsx = (-k1 |a|^2 x1 sin(phi1(d)) + k2 |b|^2 x2 sin(phi2(d))) / (k1 |a|^2 + k2 |b|^2)
    (x = k sin(theta) pi/(ldim/2), y likewise).  Unit weights, a sign or a dropped weight in either
    denominator changes f at d = 0 already: (k1|a|^2 - k2|b|^2)/(k1|a|^2 + k2|b|^2)."""
    n, kfrom, ldim = 16, 2, 32
    pf = _single_bin_case(n=n, kfrom=kfrom, ldim=ldim)
    k1, j1, a = 3, 5, 0.8 - 0.3j
    k2, j2, b = 6, 11, -0.2 + 1.1j
    pf.refs_even[0, k1 - kfrom, j1] = a
    pf.refs_even[0, k2 - kfrom, j2] = b
    s0 = (0.6, -0.2)
    ph = O.shmat_closed(pf.grid, s0)
    p = np.zeros_like(pf.refs_even[0], dtype=np.complex128)
    p[k1 - kfrom, j1] = a * ph[k1 - kfrom, j1]
    p[k2 - kfrom, j2] = -b * ph[k2 - kfrom, j2]
    pf.ptcls[0] = p.astype(np.complex64)
    # the closed form on the FP32-rounded values the oracle correlates
    A2 = abs(complex(pf.refs_even[0, k1 - kfrom, j1])) ** 2
    B2 = abs(complex(pf.refs_even[0, k2 - kfrom, j2])) ** 2
    d = (s[0] - s0[0], s[1] - s0[1])
    p1, p2 = _phase(k1, j1, n, ldim, d), _phase(k2, j2, n, ldim, d)
    den = k1 * A2 + k2 * B2
    f_cf = (k1 * A2 * math.cos(p1) - k2 * B2 * math.cos(p2)) / den
    g_cf = []
    for ax in (0, 1):
        e = [0.0, 0.0]
        e[ax] = 1.0
        x1, x2 = _phase(k1, j1, n, ldim, e), _phase(k2, j2, n, ldim, e)
        g_cf.append((-k1 * A2 * x1 * math.sin(p1) + k2 * B2 * x2 * math.sin(p2)) / den)
    # FP32 inputs (particle phase, polar grid) leave ~1e-7 relative
    f = pf.gencorr_for_rot_8(0, 0, s, 1)
    assert abs(f - f_cf) < 2e-6
    f2, g = pf.gencorr_grad_for_rot_8(0, 0, s, 1)
    assert abs(f2 - f_cf) < 2e-6
    for ax in range(2):
        assert abs(g[ax] - g_cf[ax]) < 2e-6 * max(1.0, abs(g_cf[ax]))
    if s == (0.0, 0.0):
        # at d = -s0 the closed form is not the unweighted value
        f_unw = (A2 * math.cos(p1) - B2 * math.cos(p2)) / (A2 + B2)
        assert abs(f - f_unw) > 1e-2


def test_shifted_normaliser_uses_all_shells_closed_form():
    """gencorrs_cc_2 normalises by S_q over ALL shells kfrom..kto (:1531), gencorrs_cc_1 over
    kfrom..kstop (prep_ref4corr :984).  At shift (0, 0) the numerators agree, so with a reference of
    constant magnitude a_k on shell k (no CTF) the ratio is, at every rotation,
        gencorrs(shvec = (0, 0)) / gencorrs(no shift) = sqrt(sum_{k<=kstop} a_k^2 / sum_{k<=kto} a_k^2).
    Normalising the shifted form over kfrom..kstop as well would give 1."""
    n, kfrom, kto, kstop = 16, 2, 10, 6
    pf, _ = make_case(seed=62, n=n, kfrom=kfrom, kto=kto, kstop=kstop, ctf=False, nrefs=1, nptcls=1)
    rng = np.random.default_rng(3)
    amp = np.linspace(1.0, 0.3, kto - kfrom + 1)
    pf.refs_even[0] = (amp[:, None] * np.exp(2j * np.pi * rng.random((kto - kfrom + 1, n)))).astype(np.complex64)
    a2 = np.abs(pf.refs_even[0].astype(np.complex128)) ** 2        # FP32-rounded magnitudes
    ratio_cf = math.sqrt(a2[:kstop - kfrom + 1].sum() / a2.sum())
    assert abs(ratio_cf - math.sqrt((amp[:kstop - kfrom + 1] ** 2).sum() / (amp ** 2).sum())) < 1e-6
    c1 = pf.gencorrs(0, 0, None)
    c2 = pf.gencorrs(0, 0, (0.0, 0.0))
    sel = np.abs(c1) > 1e-3 * np.max(np.abs(c1))
    np.testing.assert_allclose(c2[sel] / c1[sel], ratio_cf, rtol=1e-12)
    assert abs(ratio_cf - 1.0) > 0.05


def test_gencorrs_shifts_equals_per_shift():
    """The vectorised shift list form equals gencorrs_cc_2 shift by shift (FFT and direct routes)."""
    pf, _ = make_case(seed=63, n=24, kfrom=2, kto=9, kstop=8, ctf=True, nrefs=2, nptcls=2, eo_mixed=True)
    shifts = np.array([[0.0, 0.0], [0.5, -1.0], [-2.25, 1.5]])
    for r in range(2):
        for i in range(2):
            many = pf.gencorrs_shifts(r, i, shifts)
            for s in range(3):
                one = pf.gencorrs(r, i, tuple(shifts[s]), method="direct")
                assert np.max(np.abs(many[s] - one)) < 1e-12


@pytest.mark.parametrize("irot", [1, 5, 16, 17, 20, 32])
def test_sigma_contrib_identity_and_planted_energy(irot):
    """gencorr_sigma_contrib (:2150-2188) = 0.5/n sum_j |q_k[j] - P_k[j + rho]|^2 on every shell
    kfrom..kto, with q = ref x |CTF| x shmat (no shell normalisation).
    (1) Identity: 0.5/n (S_q,k + S_p,k - 2 C_k(rho)), C_k from the reference's FFT route
        (memoize_fft + calc_k_corrs), at irot = 1, <= n, n + 1 and > n + 1.
    (2) Planted offset: a particle equal to the reference rotated by rho plus an offset e of
        magnitude c_k on shell k (as a Hermitian full ring) gives exactly 0.5 c_k^2 on shell k: the
        0.5/n scale and the rotation index are both fixed by it."""
    n = 16
    pf, _ = make_case(seed=64, n=n, kfrom=2, kto=8, kstop=7, ctf=True, nrefs=1, nptcls=1, euclid=True)
    sh = (0.5, -0.25)
    q = pf.prep_ref(0, 0, sh, shellnorm=False)
    p = pf.ptcls[0].astype(np.complex128)
    pa, pb = O.memoize_fft(p)
    ck = O.corrs_over_k_fft(q, pa, pb)[:, irot - 1]
    ident = 0.5 / n * (np.sum(np.abs(q) ** 2, -1) + np.sum(np.abs(p) ** 2, -1) - 2.0 * ck)
    np.testing.assert_allclose(pf.sigma_contrib(0, 0, sh, irot), ident, rtol=1e-10, atol=1e-12)
    # planted offset (no CTF so that q is the shifted reference itself)
    pf.ctfmats = None
    q = pf.prep_ref(0, 0, sh, shellnorm=False)
    rng = np.random.default_rng(irot)
    c = np.linspace(0.05, 0.4, pf.grid.nk)
    e = c[:, None] * np.exp(2j * np.pi * rng.random((pf.grid.nk, n)))
    planted = np.roll(O.full_ring(q), irot - 1, axis=-1)[:, :n] + e
    pf.ptcls[0] = planted.astype(np.complex64)
    # the FP32 rounding of the planted particle moves the result by ~1e-8
    np.testing.assert_allclose(pf.sigma_contrib(0, 0, sh, irot), 0.5 * c ** 2, rtol=2e-6, atol=1e-8)


def test_median_nocopy_positions_and_selec():
    """median_nocopy (simple_math.f90:2111-2130).  With an exact order statistic in place of selec
    it is np.median for both parities (positions n/2 and n/2 + 1 averaged, nint(n/2.) for odd n);
    with the reference's selec each call stays within one rank of the exact order statistic."""
    import prep_oracle
    rng = np.random.default_rng(5)
    orig = prep_oracle.selec
    try:
        prep_oracle.selec = lambda k, a: np.float32(np.sort(a)[k - 1])
        for size in (1, 2, 5, 8, 39, 40):
            x = rng.standard_normal(size).astype(np.float32)
            assert abs(float(O.median_nocopy(x.copy())) - float(np.median(x))) <= 1e-7
    finally:
        prep_oracle.selec = orig
    for size in (7, 8, 39, 59):
        for _ in range(50):
            x = rng.standard_normal(size).astype(np.float32)
            s = np.sort(x)
            v = float(O.median_nocopy(x.copy()))
            lo = s[max(0, size // 2 - 2)] if size % 2 == 0 else s[max(0, size // 2 - 1)]
            hi = s[min(size - 1, size // 2 + 1)]
            assert lo <= v <= hi


def test_gencorrs_refs_fft_equals_per_pair():
    """The many-references form equals gencorrs_cc_1 pair by pair (direct sums), with and without
    CTF and with kstop < kto."""
    pf, _ = make_case(seed=65, n=24, kfrom=2, kto=9, kstop=7, ctf=True, nrefs=5, nptcls=2)
    for i in range(2):
        for ctf in (True, False):
            c = pf.ctfmats[i] if ctf else None
            many = O.gencorrs_refs_fft(pf.refs_even, pf.ptcls[i], c, pf.ks, chunk=2)
            saved = pf.ctfmats
            if not ctf:
                pf.ctfmats = None
            for r in range(5):
                assert np.max(np.abs(many[r] - pf.gencorrs(r, i, None, method="direct"))) < 1e-12
            pf.ctfmats = saved


# -------------------------------------------- orientation derivatives (gencorr_cont_*) ---
def _cont_case(seed):
    """kstop = kto, so the unweighted cont objective equals the shifted gencorrs normalisation."""
    pf, _ = make_case(seed=seed, n=16, kfrom=2, kto=7, ctf=True, nrefs=3, nptcls=4, ldim=32)
    pf.refs_even = pf.refs_even.astype(np.complex128)
    pf.refs_odd = pf.refs_odd.astype(np.complex128)
    pf.ptcls = pf.ptcls.astype(np.complex128)
    rng = np.random.default_rng(seed + 1)
    drefs = random_pfts(rng, 3, pf.ks, pf.pftsz).astype(np.complex128)
    return pf, drefs


def _cc_moved(pf, iref, i, sh, irot, dref, eps):
    refs = pf.refs_even if pf.iseven[i] else pf.refs_odd
    saved = refs[iref].copy()
    refs[iref][:pf.ks] += eps * dref
    try:
        return pf.gencorrs(iref, i, sh)[irot - 1]
    finally:
        refs[iref][:] = saved


@pytest.mark.parametrize("irot", [1, 6, 17, 25])
def test_cont_grad_equals_fd_of_reference_motion(irot):
    """gencorr_cont_grad_cc_for_rot_8 (:1735-1786) with the T1 particle equal to the scored one:
    cc equals the (unweighted, shifted) gencorrs value at irot, and dcc_d is the derivative of
    that cc when the reference moves along dref_d (central differences, eps 1e-4); the shift
    terms of the shift form are the derivatives over the shift (h 1e-5 px).  A dropped T2 term,
    a wrong sign or a CTF/shift factor missing on dq breaks these."""
    pf, drefs = _cont_case(101)
    i = iref = 1
    sh = (0.7, -1.1)
    cc, dcc = pf.cont_grad_cc_for_rot_8(iref, i, sh, irot, drefs)
    assert abs(cc - pf.gencorrs(iref, i, sh)[irot - 1]) < 1e-12
    eps = 1e-4
    for d in range(3):
        fd = (_cc_moved(pf, iref, i, sh, irot, drefs[d], eps) - _cc_moved(pf, iref, i, sh, irot, drefs[d], -eps)) / (2 * eps)
        assert abs(dcc[d] - fd) <= 1e-7 * max(1.0, abs(fd)), (d, dcc[d], fd)
    f, g = pf.cont_shift_grad_cc_for_rot_8(iref, i, sh, irot, drefs)
    assert f == cc and np.array_equal(g[:3], dcc)
    h = 1e-5
    for d in range(2):
        sp = list(sh)
        sm = list(sh)
        sp[d] += h
        sm[d] -= h
        fd = (pf.gencorrs(iref, i, tuple(sp))[irot - 1] - pf.gencorrs(iref, i, tuple(sm))[irot - 1]) / (2 * h)
        assert abs(g[3 + d] - fd) <= 1e-6 * max(1.0, abs(fd)), (d, g[3 + d], fd)


def test_cont_t1_reads_particle_slot_iref():
    """The reference passes iref as the particle of the T1 terms (:1781, :1836).  With particle
    slot iref holding a copy of the scored particle the derivative is the true one; with other
    contents it is not, and it follows that slot."""
    pf, drefs = _cont_case(103)
    i, iref, irot, sh = 0, 2, 9, (0.3, 0.4)
    pf.iseven[:] = True
    fd = [(_cc_moved(pf, iref, i, sh, irot, drefs[d], 1e-4) - _cc_moved(pf, iref, i, sh, irot, drefs[d], -1e-4)) / 2e-4
          for d in range(3)]
    _, d_other = pf.cont_grad_cc_for_rot_8(iref, i, sh, irot, drefs)
    assert np.max(np.abs(d_other - fd)) > 1e-3
    pf.ptcls[iref] = pf.ptcls[i]
    _, d_same = pf.cont_grad_cc_for_rot_8(iref, i, sh, irot, drefs)
    assert np.max(np.abs(d_same - fd)) < 1e-7 * max(1.0, np.max(np.abs(fd)))


def test_calc_ptcl_pspec_sums_to_sqsum():
    """calc_ptcl_pspec (:897-904) times pftsz, summed over kfrom..kstop, is memoize_sqsum_ptcl's
    particle norm (:841-845); a full-ring (2 pftsz) or a per-rotation mean breaks it."""
    pf, _ = make_case(seed=105, n=16, kto=9, kstop=7, nrefs=2, nptcls=3)
    for i in range(3):
        ps = pf.calc_ptcl_pspec(i)
        assert ps.shape == (pf.grid.nk,)
        assert abs(np.sum(ps[:pf.ks]) * pf.pftsz - pf.sqsum_ptcl(i)) < 1e-9 * pf.sqsum_ptcl(i)


def test_objective_specific_forms_ignore_cc_objfun():
    """gencorr_cc_* / gencorr_euclid_* (:1645-2148) fix the objective that gencorr_for_rot_8
    otherwise takes from cc_objfun (:1632-1643)."""
    pf, _ = make_case(seed=107, n=16, kto=7, nrefs=2, nptcls=3, euclid=True)
    e = pf.gencorr_for_rot_8(1, 2, (0.5, -0.2), 5)
    assert pf.gencorr_obj_for_rot_8(1, 2, (0.5, -0.2), 5, euclid=True) == e
    c = pf.gencorr_obj_for_rot_8(1, 2, (0.5, -0.2), 5, euclid=False)
    assert pf.euclid and c != e and abs(c) <= 1.0
    pf.euclid = False
    assert pf.gencorr_for_rot_8(1, 2, (0.5, -0.2), 5) == c
    f, g = pf.gencorr_grad_for_rot_8(1, 2, (0.5, -0.2), 5)
    pf.euclid = True
    f2, g2 = pf.gencorr_obj_grad_for_rot_8(1, 2, (0.5, -0.2), 5, euclid=False)
    assert f2 == f and np.array_equal(g, g2)
