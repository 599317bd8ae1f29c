"""Pinning of the oracle's prep_matchfilt and Savitzky-Golay smoother by closed forms.

* weights: the 7-point quartic smoother is (5, -30, 75, 131, 75, -30, 5)/231 (classical table);
  a transposed window or a wrong order fails;
* the filter reproduces any quartic polynomial exactly at interior points, drops (does not
  renormalise) window points left of the first sample, and keeps the last 3 samples;
* prep_matchfilt without pssnr leaves each shell k <= kstop with power ssnr/(ssnr+1), where
  ssnr = P_ref / fit(P_diff) from independently recomputed spectra; shells beyond kstop are
  untouched; iref < 0 returns the particle unchanged;
* with pssnr the shell power is k <|p|^2 (1 + ssnr |CTF|^2)> / <|p|^2>.
"""
import math

import numpy as np

import pftcc_oracle as O
from cases import make_case


def test_savgol_weights_closed_form():
    c = O.savgol_coefficients()
    want = np.array([5, -30, 75, 131, 75, -30, 5], np.float64) / 231.0
    got = np.array([c[o] for o in range(-3, 4)])
    assert np.max(np.abs(got - want)) < 1e-14


def test_savgol_reproduces_quartics_and_border_rules():
    rng = np.random.default_rng(5)
    x = np.arange(20, dtype=np.float64)
    coef = rng.standard_normal(5)
    y = sum(cf * x ** d for d, cf in enumerate(coef))
    f = O.savitzky_golay_filter(y)
    assert np.max(np.abs(f[3:17] - y[3:17])) < 1e-9 * np.max(np.abs(y))
    assert np.array_equal(f[17:], y[17:])
    c = O.savgol_coefficients()
    # sample 0: only the window points at or right of it
    assert abs(f[0] - sum(c[o] * y[o] for o in range(0, 4))) < 1e-12 * np.max(np.abs(y))
    assert abs(f[1] - sum(c[o] * y[1 + o] for o in range(-1, 4))) < 1e-12 * np.max(np.abs(y))


def _power(v):
    return np.sum(np.abs(v) ** 2, axis=-1) / v.shape[-1]


def test_matchfilt_shell_power_and_untouched_parts():
    pf, _ = make_case(seed=81, n=40, kfrom=2, kto=16, kstop=14, ctf=True, nrefs=3, nptcls=3, match_filt=True)
    i, iref, irot = 1, 2, 17
    out = pf.prep_matchfilt(i, iref, irot)
    ks = pf.ks
    # recompute the spectra the filter uses, independently of the oracle's loop
    n = pf.pftsz
    opt = pf.ref_optlp[iref].astype(np.float64)
    p = pf.ptcls[i].astype(np.complex128)
    r = pf.ref(iref, i) * pf.ctf(i)
    pn = p[:ks] * (opt / np.sqrt(_power(p[:ks])))[:, None]
    rn = r[:ks] * (opt / np.sqrt(_power(r[:ks])))[:, None]
    P = np.concatenate([pn, np.conj(pn)], axis=-1)
    rho = irot - 1
    rolled = np.stack([P[:, (j + rho) % (2 * n)] for j in range(n)], axis=-1)
    diff = _power(rn - rolled)
    fit = O.savitzky_golay_filter(diff)
    fit = np.where(fit <= 0, diff, fit)
    fit[:3] = diff[:3]
    ssnr = _power(rn) / fit
    assert np.allclose(_power(out[:ks]), ssnr / (ssnr + 1.0), rtol=1e-12)
    # phases are kept, shells beyond kstop untouched
    assert np.allclose(out[:ks] / np.abs(out[:ks]), p[:ks] / np.abs(p[:ks]), atol=1e-12)
    assert np.array_equal(out[ks:], p[ks:])
    assert np.array_equal(pf.prep_matchfilt(i, -1, irot), p)


def test_matchfilt_pssnr_shell_power():
    pf, _ = make_case(seed=82, n=32, kfrom=2, kto=12, ctf=True, nrefs=2, nptcls=2, match_filt=True)
    pf.pssnr = True
    out = pf.prep_matchfilt(0, 1, 1)
    p = pf.ptcls[0].astype(np.complex128)
    c = pf.ctf(0)
    opt = pf.ref_optlp[1].astype(np.float64)
    for kk in range(pf.ks):
        k = pf.kfrom + kk
        want = k * np.mean(np.abs(p[kk]) ** 2 * (1 + opt[kk] * c[kk] ** 2)) / np.mean(np.abs(p[kk]) ** 2)
        assert math.isclose(_power(out[kk]), want, rel_tol=1e-12)
