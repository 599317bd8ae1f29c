"""Pins the polarizer oracle (oracle/polarizer_oracle.py) against what the mathematics fixes.

* bessi0: the published error bound of the Abramowitz & Stegun polynomials against scipy's I0;
* the Kaiser-Bessel window: closed forms (beta, support, symmetry, values at 0 and +-Whalf);
* window and cyclic-address helpers: Fortran nint / wrap rules on hand cases;
* get_fcomp2D: equality with the full complex FFT of a real image for every logical address;
* polarize: a brute-force direct sum (scipy I0, full FFT array, independently placed window),
  partition of unity, exactness for linear transforms at integer coordinates, and agreement with
  the analytic transform of an off-centre Gaussian image (axis roles and phase sign).
"""
import numpy as np
import pytest
import scipy.special as sp

import pftcc_oracle as O
import polarizer_oracle as P


def test_bessi0_within_published_bound():
    x = np.concatenate([np.linspace(0, 3.75, 1501), np.linspace(3.75, 40, 3001)]).astype(np.float32)
    ref = sp.i0(x.astype(np.float64))
    rel = np.abs(P.bessi0(x) / ref - 1.0)
    # A&S 9.8.1: |eps| < 1.6e-7 (relative to I0); 9.8.2: |eps| < 1.9e-7 (relative to the
    # scaled polynomial, i.e. to I0 sqrt(x) e^-x)
    assert rel[x < 3.75].max() < 1.6e-7
    assert rel[x >= 3.75].max() < 2.0e-7
    # a wrong coefficient would be visible: perturbing one by 1e-4 breaks the bound
    bad = list(P._AS_981)
    bad[3] += 1e-4
    ys = (x[x < 3.75].astype(np.float64) / 3.75) ** 2
    assert np.abs(P._horner(bad, ys) / ref[x < 3.75] - 1).max() > 1e-6


def test_kb_params_closed_forms():
    kb = P.kb_params()
    assert (float(kb.whalf), float(kb.W), float(kb.beta)) == (1.5, 3.0, pytest.approx(7.4, rel=1e-7))
    assert P.wdim(kb) == 3
    # above half-width 1.5: beta = pi sqrt((W/alpha)^2 (alpha - 1/2)^2 - 0.8) (Beatty et al. 2005)
    kb2 = P.kb_params(2.0, 2.0)
    W, a = 4.0, 2.0
    assert float(kb2.beta) == pytest.approx(np.pi * np.sqrt((W / a) ** 2 * (a - 0.5) ** 2 - 0.8), rel=1e-6)
    assert P.wdim(kb2) == 5


def test_apod_closed_forms():
    kb = P.kb_params()
    x = np.linspace(-1.5, 1.5, 601).astype(np.float32)
    a = P.apod(x, kb)
    np.testing.assert_array_equal(a, a[::-1])                       # symmetric
    assert float(P.apod(np.float32(0.0), kb)) == pytest.approx(sp.i0(7.4) / 3.0, rel=2e-7)
    assert float(P.apod(np.float32(1.5), kb)) == pytest.approx(1.0 / 3.0, rel=1e-7)   # I0(0) / W
    assert float(P.apod(np.float32(1.5001), kb)) == 0.0             # support |x| <= Whalf
    ref = sp.i0(7.4 * np.sqrt(np.maximum(1 - (2 * x.astype(np.float64) / 3.0) ** 2, 0))) / 3.0
    assert np.abs(a / ref - 1).max() < 1e-6
    assert np.all(np.diff(a[300:]) < 0)                            # decreasing away from 0


def test_window_and_cyclic_helpers():
    assert P.sqwin_1d(2.5, 1.5) == (2, 4)        # nint(2.5) = 3
    assert P.sqwin_1d(-2.5, 1.5) == (-4, -2)     # nint(-2.5) = -3 (half away from zero)
    assert P.sqwin_1d(0.49, 1.5) == (-1, 1)
    lims = (-4, 4)
    assert [P.cyci_1d(lims, i) for i in (-6, -5, -4, 0, 4, 5, 6)] == [3, 4, -4, 0, 4, -4, -3]
    assert P.fourier_lims(64) == ((-32, 32), (-32, 31))


def test_get_fcomp2D_matches_full_fft():
    rng = np.random.default_rng(3)
    L = 16
    img = rng.standard_normal((L, L))
    cmat = P.image_fft(img)
    full = np.fft.fft2(np.fft.fftshift(img)) / (L * L)
    (hlo, hhi), (klo, khi) = P.fourier_lims(L)
    for h in range(hlo, hhi + 1):
        for k in range(klo, khi + 1):
            assert P.get_fcomp2D(cmat, h, k) == pytest.approx(complex(full[k % L, h % L]), abs=1e-6)


def _grid(n=24, kfrom=2, kto=12, ldim=32):
    return O.make_grid(n, kfrom, kto, ldim)


def _brute_force(grid, cmat_full, L):
    """Independent gridding: window placed with numpy rounding, weights from scipy's I0, values
    from the full complex FFT array (no Friedel rule)."""
    n, nk = grid.pftsz, grid.kto - grid.kfrom + 1
    out = np.zeros((nk, n), np.complex128)
    kb = lambda d: sp.i0(7.4 * np.sqrt(max(0.0, 1 - (2 * d / 3.0) ** 2))) / 3.0 if abs(d) <= 1.5 else 0.0
    for kk in range(nk):
        for i in range(n):
            x = float(grid.polar_x[kk, i])
            y = float(grid.polar_y[kk, i])
            cx = int(np.sign(x) * np.floor(abs(x) + 0.5))
            cy = int(np.sign(y) * np.floor(abs(y) + 0.5))
            acc, ws = 0j, 0.0
            for hh in range(cx - 1, cx + 2):
                for kq in range(cy - 1, cy + 2):
                    w = kb(hh - x) * kb(kq - y)
                    acc += w * cmat_full[kq % L, hh % L]
                    ws += w
            out[kk, i] = acc / ws
    return out


def test_polarize_matches_brute_force():
    rng = np.random.default_rng(5)
    g = _grid()
    L = g.ldim
    img = rng.standard_normal((L, L))
    cmat = P.image_fft(img)
    full = np.fft.fft2(np.fft.fftshift(img)) / (L * L)
    tab = P.polarizer_tables(g.polar_x, g.polar_y, g.pftsz, L)
    got = P.polarize(tab, cmat)
    ref = _brute_force(g, full, L)
    assert np.abs(got - ref).max() < 2e-6 * np.abs(ref).max()
    np.testing.assert_allclose(P.polarize_vec(tab, cmat), got, rtol=0, atol=1e-12 * np.abs(got).max())


def test_polarize_partition_of_unity_and_griddev():
    g = _grid()
    L = g.ldim
    cmat = np.full((L, L // 2 + 1), 2.5 + 0j, np.complex64)
    tab = P.polarizer_tables(g.polar_x, g.polar_y, g.pftsz, L)
    np.testing.assert_allclose(P.polarize_vec(tab, cmat), 2.5, rtol=2e-7)
    raw = P.polarizer_tables(g.polar_x, g.polar_y, g.pftsz, L, griddev=True)
    s = raw.weights.astype(np.float64).sum(axis=2)
    np.testing.assert_allclose(P.polarize_vec(raw, cmat).real, 2.5 * s, rtol=1e-6)
    assert s.min() > 1.0          # the un-normalised KB window does not sum to one


def test_polarize_linear_exact_at_integer_coordinates():
    # F(h, k) = c + i (a h + b k) is consistent with the Friedel rule; at theta = 0 the polar point
    # sits at the integer location (0, -k), the window is symmetric and the interpolant is exact.
    g = _grid()
    L = g.ldim
    a, b, c = 0.3, -0.7, 1.25
    cmat = np.zeros((L, L // 2 + 1), np.complex64)
    for kp in range(L):
        k = kp if kp < L // 2 else kp - L
        for h in range(L // 2 + 1):
            cmat[kp, h] = c + 1j * (a * h + b * k)
    tab = P.polarizer_tables(g.polar_x, g.polar_y, g.pftsz, L)
    got = P.polarize(tab, cmat)
    ks = np.arange(g.kfrom, g.kto + 1)
    np.testing.assert_allclose(got[:, 0], c + 1j * b * (-ks), rtol=0, atol=2e-6)


def test_polarize_gaussian_image_against_analytic_transform():
    # off-centre Gaussian: its transform is 2 pi s^2 exp(-2 pi^2 s^2 |f|^2) exp(-2 pi i f.c) / L^2,
    # f = (x, y) / L in the polar grid's (h, k) frame; a swapped axis or a flipped phase sign
    # breaks the agreement.
    L, s = 64, 2.5
    cx, cy = 5.0, -3.0                     # offset from the phase origin (pixel L/2)
    yy, xx = np.mgrid[0:L, 0:L].astype(np.float64)
    img = np.exp(-((xx - L / 2 - cx) ** 2 + (yy - L / 2 - cy) ** 2) / (2 * s * s))
    # without the instrument-function division the window's roll-off shows (several %)
    raw = P.image_fft(img)
    ins = P.instrfun_img(L)
    cmat = P.image_fft(img / (ins / ins[L // 2, L // 2]))   # scale-free: unit gain at the centre
    g = O.make_grid(48, 2, 14, L)
    tab = P.polarizer_tables(g.polar_x, g.polar_y, g.pftsz, L)
    got = P.polarize_vec(tab, cmat)
    fx = g.polar_x[:, :g.pftsz].astype(np.float64) / L
    fy = g.polar_y[:, :g.pftsz].astype(np.float64) / L
    ana = 2 * np.pi * s * s * np.exp(-2 * np.pi ** 2 * s * s * (fx ** 2 + fy ** 2)) \
        * np.exp(-2j * np.pi * (fx * cx + fy * cy)) / (L * L)
    err = np.abs(got - ana).max() / np.abs(ana).max()
    err_raw = np.abs(P.polarize_vec(tab, raw) - ana).max() / np.abs(ana).max()
    assert err < 0.01, err            # 0.4 % measured: the 3 x 3 window's residual error
    assert err_raw > 5 * err          # 5.8 % without the instrument-function division
    swapped = 2 * np.pi * s * s * np.exp(-2 * np.pi ** 2 * s * s * (fx ** 2 + fy ** 2)) \
        * np.exp(-2j * np.pi * (fy * cx + fx * cy)) / (L * L)
    assert np.abs(got - swapped).max() / np.abs(ana).max() > 0.3
