"""Pins the oracle of prepimg4align (oracle/prep_oracle.py) by order statistics, closed forms and
transform identities that do not restate its formulas."""
import numpy as np
import pytest

import pftcc_oracle as O
import polarizer_oracle as PZ
import prep_oracle as P

f32 = np.float32


def _selec_nr(k, arr):
    """Numerical Recipes' select with the textbook final test (ir == l + 1): the exact k-th order
    statistic.  Differs from the reference's selec only in that one comparison."""
    a = list(arr)
    l, ir = 1, len(a)
    while True:
        if ir <= l + 1:
            if ir == l + 1 and a[ir - 1] < a[l - 1]:
                a[l - 1], a[ir - 1] = a[ir - 1], a[l - 1]
            return a[k - 1]
        mid = (l + ir) // 2
        a[mid - 1], a[l] = a[l], a[mid - 1]
        if a[l - 1] > a[ir - 1]:
            a[l - 1], a[ir - 1] = a[ir - 1], a[l - 1]
        if a[l] > a[ir - 1]:
            a[l], a[ir - 1] = a[ir - 1], a[l]
        if a[l - 1] > a[l]:
            a[l - 1], a[l] = a[l], a[l - 1]
        i, j, piv = l + 1, ir, a[l]
        while True:
            i += 1
            while a[i - 1] < piv:
                i += 1
            j -= 1
            while a[j - 1] > piv:
                j -= 1
            if j < i:
                break
            a[i - 1], a[j - 1] = a[j - 1], a[i - 1]
        a[l] = a[j - 1]
        a[j - 1] = piv
        if j >= k:
            ir = j - 1
        if j <= k:
            l = i


def test_selec_against_order_statistics():
    rng = np.random.default_rng(0)
    differ = 0
    for trial in range(400):
        n = int(rng.integers(3, 220))
        x = rng.standard_normal(n).astype(np.float32)
        k = n // 2 if trial % 2 else int(rng.integers(1, n + 1))
        exact = np.sort(x)[k - 1]
        # the textbook variant is the exact order statistic
        assert _selec_nr(k, x) == exact
        arr = x.copy()
        v = P.selec(k, arr)
        # the array is permuted in place, never altered
        np.testing.assert_array_equal(np.sort(arr), np.sort(x))
        # the reference's skipped two-element swap can only leave a neighbouring rank at k
        srt = np.sort(x)
        ranks = [srt[r] for r in (k - 2, k - 1, k) if 0 <= r < n]
        assert v in ranks
        differ += v != exact
    # the quirk is real: some tails are left unsorted
    assert differ > 0


def test_selec_quirk_case():
    # [5, 4, 3, 2, 1], k = 2 -> the final segment [l, ir] = [1, 2] is sorted (ir - 1 == 1); a
    # segment [2, 3] stays unsorted: k = 3 of [1, 5, 4, 3, 2] ends with a two-element tail at 3-4
    arr = np.array([1, 2, 3, 5, 4], np.float32)
    assert P.selec(4, arr.copy()) in (4.0, 5.0)
    assert P.selec(2, np.array([5, 4, 3, 2, 1], np.float32)) == 2.0


def test_disc_mask_closed_form():
    B, r = 64, 20.0
    m = P.disc_mask(B, r)
    c = B // 2                          # 0-based centre of cendist (ldim/2 + 1, 1-based)
    assert m[c, c]
    for d in range(0, 30):
        assert m[c, c + d] == (d <= r) and m[c + (d if c + d < B else 0), c] == (d <= r or c + d >= B)
    assert m.sum() == sum(1 for y in range(B) for x in range(B) if (x - c) ** 2 + (y - c) ** 2 <= r * r)


def test_noise_norm_background_statistics():
    rng = np.random.default_rng(1)
    B = 64
    lm = P.disc_mask(B, 20.0)
    img = (3.0 + 2.5 * rng.standard_normal((B, B))).astype(np.float32)
    img[lm] += 10.0
    x, sdev, ave, used = P.noise_norm(img, lm)
    bg = x[~lm].astype(np.float64)
    assert abs(bg.mean()) < 1e-6
    assert abs(bg.std(ddof=1) - 1.0) < 1e-5
    assert abs(float(ave) - img[~lm].astype(np.float64).mean()) < 1e-5
    assert abs(float(sdev) - img[~lm].astype(np.float64).std(ddof=1)) < 1e-4


def test_fft_round_trip_and_shift_against_roll():
    rng = np.random.default_rng(2)
    B = 48
    img = rng.standard_normal((B, B)).astype(np.float32)
    c = PZ.image_fft(img)
    np.testing.assert_allclose(P.ifft_image(c), img, atol=1e-5)
    for v in ((3.0, -5.0), (-7.0, 2.0), (0.0, 11.0)):
        g = P.ifft_image(P.shift2Dserial(c, v, B))
        # G = F exp(+i 2 pi (h vx + k vy) / B)  <=>  g(x) = f(x + v)
        ref = np.roll(img, (-int(v[1]), -int(v[0])), axis=(0, 1))
        np.testing.assert_allclose(g, ref, atol=2e-5)


def test_ctf_flip_sign_and_zero_crossing():
    B = 64
    # no astigmatism, Cs = 0, no amplitude contrast: ctf = sin(pi wl df s^2), first zero at
    # s^2 = 1 / (wl df)
    p = O.CtfParams(smpd=1.0, kv=300.0, cs=0.0, fraca=0.0, dfx=1.0, dfy=1.0, angast=0.0)
    t = P.ctf_values(B, p)
    wl = 12.26 / np.sqrt(300e3 + 0.9784 * (300e3) ** 2 / 1e6)
    s2z = 1.0 / (wl * 1.0e4)
    h = np.arange(B // 2 + 1) / B
    row = t[0]                          # k = 0
    for hh, v in zip(h, row):
        if hh * hh < s2z * 0.98 and hh > 0:
            assert v > 0
        if s2z * 1.02 < hh * hh < 2 * s2z * 0.98:
            assert v < 0
    rng = np.random.default_rng(3)
    c = (rng.standard_normal((B, B // 2 + 1)) + 1j * rng.standard_normal((B, B // 2 + 1))).astype(np.complex64)
    fl = P.ctf_flip(c, B, p)
    np.testing.assert_array_equal(np.abs(fl), np.abs(c))
    np.testing.assert_array_equal(P.ctf_flip(fl, B, p), c)
    assert np.all((fl == c) == ~np.signbit(t))
    np.testing.assert_allclose(t, P.ctf_values_fp64(B, p), atol=2e-4)


def test_clip_centre():
    B, M = 64, 40
    img = np.zeros((B, B), np.float32)
    img[B // 2, B // 2] = 1.0
    cl = P.clip(img, M)
    assert cl.shape == (M, M) and cl[M // 2, M // 2] == 1.0 and cl.sum() == 1.0


def test_soft_mask_closed_form():
    M, r = 64, 20.0
    w = P.soft_mask_weights(M, r)
    minlen = min(int(np.floor(2 * (r + 6) + 0.5)), M)
    maxrad = minlen // 2
    width = 2.0 * (maxrad - r)
    for i in range(1, M // 2 + 1):
        for j in range(1, M // 2 + 1):
            x, y = -M / 2 + (i - 1), -M / 2 + (j - 1)
            rad = np.hypot(x, y)
            if rad >= maxrad:
                e = 0.0
            elif rad >= maxrad - width:
                e = (np.cos((rad - (maxrad - width)) / width * np.pi) + 1) / 2
            else:
                e = 1.0
            if e > 0.9999:
                e = 1.0
            for (a, b) in ((i, j), (M + 1 - i, j), (i, M + 1 - j), (M + 1 - i, M + 1 - j)):
                assert abs(w[b - 1, a - 1] - e) < 2e-6
    wi = P.soft_mask_weights(M, r, inner=8.0, width=4.0)
    c = M // 2
    assert wi[c, c] == 0.0 and wi[c, c + 12] == w[c, c + 12]


def test_zero_background_constant_edges():
    rng = np.random.default_rng(4)
    M = 32
    img = rng.standard_normal((M, M)).astype(np.float32)
    img[0, :] = img[-1, :] = img[:, 0] = img[:, -1] = f32(2.5)
    med = P.zero_background(img)
    assert med == f32(2.5)
    assert np.all(img[0, :] == 0) and np.all(img[:, -1] == 0)


def test_prep_identity_settings_reduce_to_fft():
    """No shift, no CTF, boxmatch = box, a mask whose weights are 1 over the image's support and
    edges already at zero: prepimg4align = fft of the noise-normalised image minus its edge median."""
    rng = np.random.default_rng(5)
    B = 64
    yy, xx = np.mgrid[0:B, 0:B]
    blob = np.exp(-((xx - B / 2) ** 2 + (yy - B / 2) ** 2) / (2 * 4.0 ** 2))
    img = (blob + 0.01 * rng.standard_normal((B, B))).astype(np.float32)
    cfg = P.PrepCfg(box=B, boxmatch=B, msk=12.0)
    out, steps = P.prepimg4align(img, (0.0, 0.0), cfg, return_steps=True)
    x, *_ = P.noise_norm(img, P.disc_mask(B, 12.0))
    r = x.copy()
    P.zero_background(r)
    w = P.soft_mask_weights(B, 12.0)
    ref = PZ.image_fft(np.where(w != 1.0, w * r, r).astype(np.float32))
    np.testing.assert_allclose(out, ref, atol=2e-6 * np.abs(ref).max())
    # a shifted copy: integer shifts move the clipped image by -shift before masking
    out2, s2 = P.prepimg4align(img, (2.0, -3.0), cfg, return_steps=True)
    x_ro = np.roll(x, (-3, 2), axis=(0, 1))
    r2 = P.ifft_image(PZ.image_fft(x_ro))
    P.zero_background(r2)
    np.testing.assert_allclose(s2["clipped_masked"], np.where(w != 1.0, w * r2, r2), atol=2e-5)


@pytest.mark.parametrize("B,M", [(64, 64), (64, 48)])
def test_prep_output_shape_and_ratio(B, M):
    rng = np.random.default_rng(6)
    img = rng.standard_normal((B, B)).astype(np.float32)
    out = P.prepimg4align(img, (0.0, 0.0), P.PrepCfg(box=B, boxmatch=M, msk=M / 2 - 8))
    assert out.shape == (M, M // 2 + 1)
    assert np.isfinite(out).all()
