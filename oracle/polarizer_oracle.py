"""Oracle of SIMPLE's image polarizer: Kaiser-Bessel gridding of a Cartesian 2-D Fourier transform
onto the polar grid of the correlation search (TEST INFRASTRUCTURE ONLY).

Restates, in numpy, ``init_polarizer`` and ``polarize`` (``src/main/simple_polarizer.f90:41-151``),
the Kaiser-Bessel window (``src/ops/simple_kbinterpol.f90:92-141,229-251``), the window and
cyclic-address helpers (``src/utils/simple_math.f90:746-757,778-809``), the Fourier-address rules
of the image object (``src/main/simple_ftiter.f90:69-120,224-231``,
``src/main/simple_image.f90:1894-1907``) and its forward FFT convention
(``src/main/simple_image.f90:5773-5782,6016-6114``).  Each function cites the lines it follows.

The reference computes the weights in FP32 around an FP64 Bessel polynomial; this restatement
does the same with numpy float32 scalars, so weights agree with the reference's to FP32 rounding.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this file; the product (the HIP library) never calls into it.

Parity status: every function is pinned by ``tests/test_oracle_polarizer.py``:
* ``bessi0`` against ``scipy.special.i0`` within the published error bound of the polynomial
  (Abramowitz & Stegun 9.8.1-9.8.2, |eps| < 1.6e-7 and 1.9e-7);
* ``kb_params`` / ``apod`` by closed forms (beta = 7.4 for half-width <= 1.5, the Beatty formula
  above it, apod symmetric with support |x| <= half-width and apod(0) = I0(beta)/W);
* ``get_fcomp2D`` against the full complex FFT of a real image (Friedel mates and wrap);
* ``polarize`` by brute force (an independent direct sum with scipy's I0), by the partition of
  unity (a constant transform stays constant), by exactness for linear transforms at integer
  polar coordinates, and end to end against the analytic transform of a Gaussian image.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

f32 = np.float32

KBWINSZ = 1.5            # simple_defs.f90:131 (window half-width of the 2-D polarizer)
KBALPHA = float(np.sqrt(f32(2.0)))   # simple_defs.f90:132


# ---------------------------------------------------------------------------------------------
# Kaiser-Bessel window
# ---------------------------------------------------------------------------------------------
# Abramowitz & Stegun 9.8.1 (|x| <= 3.75, t = x/3.75) and 9.8.2 (x >= 3.75) polynomial
# coefficients, the published table the reference's bessi0 evaluates (simple_kbinterpol.f90:229-251)
_AS_981 = (1.0, 3.5156229, 3.0899424, 1.2067492, 0.2659732, 0.0360768, 0.0045813)
_AS_982 = (0.39894228, 0.01328592, 0.00225319, -0.00157565, 0.00916281, -0.02057706, 0.02635537,
           -0.01647633, 0.00392377)


def _horner(coefs, y):
    acc = coefs[-1]
    for c in coefs[-2::-1]:
        acc = c + y * acc
    return acc


def bessi0(x):
    """bessi0, simple_kbinterpol.f90:229-251: I0 of an FP32 argument (assumed >= 0), evaluated
    in FP64 with the A&S polynomials."""
    ax = np.asarray(x, dtype=np.float32).astype(np.float64)
    small = ax < 3.75
    out = np.empty_like(ax)
    ys = (ax[small] / 3.75) ** 2
    out[small] = _horner(_AS_981, ys)
    al = ax[~small]
    yl = 3.75 / al
    out[~small] = _horner(_AS_982, yl) * np.exp(al) / np.sqrt(al)
    return out if out.shape else float(out)


@dataclass(frozen=True)
class KB:
    """kbinterpol%new, simple_kbinterpol.f90:92-109 (all fields FP32)."""
    whalf: np.float32
    alpha: np.float32
    W: np.float32
    beta: np.float32
    twooW: np.float32
    oneoW: np.float32


def kb_params(whalf: float = KBWINSZ, alpha: float = KBALPHA) -> KB:
    wh, al = f32(whalf), f32(alpha)
    W = f32(2.0) * wh
    if wh <= f32(1.5):
        beta = f32(7.4)
    else:
        beta = f32(np.pi) * np.sqrt(f32((W ** f32(2.0) / al ** f32(2.0)) * (al - f32(0.5)) ** f32(2.0) - f32(0.8)))
        beta = f32(beta)
    return KB(wh, al, W, beta, f32(2.0) / W, f32(1.0) / W)


def apod(x, kb: KB):
    """kbinterpol%apod, simple_kbinterpol.f90:130-141: zero outside |x| <= Whalf, else
    (1/W) I0(beta sqrt(1 - (2x/W)^2)), FP32 argument, FP64 Bessel, FP32 result."""
    x = np.asarray(x, dtype=np.float32)
    arg = kb.twooW * x
    arg = f32(1.0) - arg * arg
    inside = np.abs(x) <= kb.whalf
    b = kb.beta * np.sqrt(np.where(inside, arg, f32(0.0)).astype(np.float32))
    r = (np.float64(kb.oneoW) * bessi0(b.astype(np.float32))).astype(np.float32)
    return np.where(inside, r, f32(0.0)).astype(np.float32)


def instr(x, kb: KB):
    """kbinterpol%instr, simple_kbinterpol.f90:211-225: the window's transform,
    sinh(a)/a with a = sqrt(beta^2 - (pi W x)^2) for |x| < beta/(pi W) - TINY^2, else 1 (sinh(a)/a
    in FP64 here; the reference evaluates it with a Hart & Cheney rational approximation)."""
    x = np.asarray(x, dtype=np.float64)
    beta, W = float(kb.beta), float(kb.W)
    thresh = beta / (np.pi * W) - 1e-20
    a2 = beta * beta - (np.pi * W * x) ** 2
    a = np.sqrt(np.maximum(a2, 0.0))
    r = np.where(a < 1e-10, 1.0, np.sinh(np.maximum(a, 1e-10)) / np.maximum(a, 1e-10))
    return np.where(np.abs(x) < thresh, r, 1.0)


def instrfun_img(ldim: int, kb: KB | None = None) -> np.ndarray:
    """gen_instrfun_img, simple_gridding.f90:18-55, 2-D: w(i) w(j), w(i) = instr((i - center) / ldim),
    center = ldim/2 + 1 (1-based); the real-space image is divided by it before fwd_ft
    (div_by_instrfun, simple_polarizer.f90:106-112)."""
    kb = kb or kb_params()
    i = np.arange(1, ldim + 1)
    w = instr((i - (ldim // 2 + 1)) / ldim, kb)
    return np.outer(w, w)


def sqwin_1d(x: float, winsz: float):
    """sqwin_1d_1, simple_math.f90:778-787: [nint(x) - iw, nint(x) + iw], iw = ceiling(winsz - 0.5)."""
    c = int(np.floor(abs(float(x)) + 0.5)) * (1 if x >= 0 else -1)    # Fortran nint
    iw = int(np.ceil(f32(winsz) - f32(0.5)))
    return c - iw, c + iw


def wdim(kb: KB) -> int:
    """kbinterpol%get_wdim, simple_kbinterpol.f90:121-127."""
    a, b = sqwin_1d(0.0, float(kb.whalf))
    return b - a + 1


def cyci_1d(lims, i: int) -> int:
    """cyci_1d, simple_math.f90:746-757 (cyclic index into [lims[0], lims[1]])."""
    if i > lims[1]:
        return lims[0] + (i - lims[1]) - 1
    if i < lims[0]:
        return lims[1] - (lims[0] - i) + 1
    return i


def fourier_lims(ldim: int):
    """loop_lims(3) of a square 2-D image (simple_ftiter.f90:69-120, 224-231): logical h in
    [-ldim/2, ldim/2] (odd: [-(ldim-1)/2, (ldim-1)/2]) and k in [-ldim/2, ldim/2 - 1]."""
    if ldim % 2 == 0:
        return (-ldim // 2, ldim // 2), (-ldim // 2, ldim // 2 - 1)
    h = (ldim - 1) // 2
    return (-h, h), (-h, h)


# ---------------------------------------------------------------------------------------------
# The image object's Fourier transform
# ---------------------------------------------------------------------------------------------
def image_fft(img: np.ndarray) -> np.ndarray:
    """fwd_ft, simple_image.f90:5773-5782, of a real square image of even size given as
    img[y, x] (the C view of Fortran rmat(x, y)): quadrant swap to the phase origin
    (shift_phorig, simple_image.f90:6016-6114), r2c, 1/N.  Returns cmat as a C array
    [k_phys][h_phys] = Fortran cmat(h_phys, k_phys), shape [ldim][ldim/2 + 1], complex64."""
    ny, nx = img.shape
    if nx % 2 or ny % 2:
        raise ValueError("even dimensions only")
    c = np.fft.rfft2(np.fft.fftshift(img.astype(np.float64))) / float(nx * ny)
    return c.astype(np.complex64)


def get_fcomp2D(cmat: np.ndarray, h: int, k: int) -> complex:
    """get_fcomp2D, simple_image.f90:1894-1907: logical (h, k) -> cmat, Friedel mate (conjugate)
    for h < 0; cmat is the C view [k_phys][h_phys] (0-based here)."""
    ldim2 = cmat.shape[0]
    if h >= 0:
        p1 = h
        p2 = k + (ldim2 if k < 0 else 0)
        return complex(cmat[p2, p1])
    p1 = -h
    p2 = -k + (ldim2 if -k < 0 else 0)
    return complex(np.conj(cmat[p2, p1]))


# ---------------------------------------------------------------------------------------------
# init_polarizer / polarize
# ---------------------------------------------------------------------------------------------
@dataclass
class PolarTables:
    """polcyc1_mat, polcyc2_mat, polweights_mat of init_polarizer (simple_polarizer.f90:41-92),
    as C arrays over (k - kfrom, i - 1)."""
    cyc1: np.ndarray      # int [nk, pftsz, wdim]  logical h of window column l
    cyc2: np.ndarray      # int [nk, pftsz, wdim]  logical k of window row m
    weights: np.ndarray   # f32 [nk, pftsz, wdim*wdim]  Fortran reshape order: l fastest
    wdim: int


def polarizer_tables(polar_x: np.ndarray, polar_y: np.ndarray, pftsz: int, ldim: int,
                     griddev: bool = False, kb: KB | None = None) -> PolarTables:
    """init_polarizer, simple_polarizer.f90:41-92.  polar_x/polar_y: FP32 [nk, nrots] grid
    (get_coord(i, k) = (polar(i, k), polar(i + nrots, k)), i <= pftsz).  For each polar point:
    window = sqwin around nint(loc), w(l, m) = apod(x_l - loc_x) * apod(y_m - loc_y) built as
    the reference does (rows then columns, FP32), cyclic logical addresses, and w / sum(w) unless
    griddev."""
    kb = kb or kb_params()
    wd = wdim(kb)
    limh, limk = fourier_lims(ldim)
    nk = polar_x.shape[0]
    cyc1 = np.zeros((nk, pftsz, wd), np.int32)
    cyc2 = np.zeros((nk, pftsz, wd), np.int32)
    wts = np.zeros((nk, pftsz, wd * wd), np.float32)
    for kk in range(nk):
        for i in range(pftsz):
            lx = f32(polar_x[kk, i])
            ly = f32(polar_y[kk, i])
            w1 = sqwin_1d(float(lx), float(kb.whalf))[0]
            w2 = sqwin_1d(float(ly), float(kb.whalf))[0]
            w = np.ones((wd, wd), np.float32)       # w(l, m): l along h, m along k
            for l in range(wd):
                w[l, :] = w[l, :] * apod(f32(w1 + l) - lx, kb)
                w[:, l] = w[:, l] * apod(f32(w2 + l) - ly, kb)
                cyc1[kk, i, l] = cyci_1d(limh, w1 + l)
                cyc2[kk, i, l] = cyci_1d(limk, w2 + l)
            flat = w.reshape(-1, order="F")          # reshape(w, [wlen]): l fastest
            if not griddev:
                s = f32(0.0)
                for v in flat:                        # Fortran sum, in order, FP32
                    s = f32(s + v)
                flat = (flat / s).astype(np.float32)
            wts[kk, i] = flat
    return PolarTables(cyc1, cyc2, wts, wd)


def polarize(tab: PolarTables, cmat: np.ndarray, mask=None, dtype=np.complex128) -> np.ndarray:
    """polarize, simple_polarizer.f90:120-151: pft(i, k) = sum_{l, m} w(l, m) F(h_l, k_m) over
    the window (Fortran order, l fastest), shells with mask(k) false zeroed.  Returns
    pft[k - kfrom, i - 1] (complex, FP64 accumulation by default)."""
    nk, n, wd = tab.cyc1.shape
    out = np.zeros((nk, n), dtype)
    for kk in range(nk):
        if mask is not None and not mask[kk]:
            continue
        for i in range(n):
            acc = 0j
            idx = 0
            for m in range(wd):
                for l in range(wd):
                    acc += float(tab.weights[kk, i, idx]) * get_fcomp2D(cmat, int(tab.cyc1[kk, i, l]),
                                                                         int(tab.cyc2[kk, i, m]))
                    idx += 1
            out[kk, i] = acc
    return out


def polarize_vec(tab: PolarTables, cmat: np.ndarray, mask=None) -> np.ndarray:
    """polarize with vectorised gathers (same formula and summation terms as ``polarize``), for
    larger test cases."""
    nk, n, wd = tab.cyc1.shape
    ldim2 = cmat.shape[0]
    h = tab.cyc1[:, :, None, :].repeat(wd, axis=2)          # [nk, n, m, l]
    k = tab.cyc2[:, :, :, None].repeat(wd, axis=3)
    neg = h < 0
    hh = np.where(neg, -h, h)
    kk_ = np.where(neg, -k, k)
    kk_ = np.where(kk_ < 0, kk_ + ldim2, kk_)
    vals = cmat.astype(np.complex128)[kk_, hh]
    vals = np.where(neg, np.conj(vals), vals)
    w = tab.weights.reshape(nk, n, wd, wd).astype(np.float64)   # [nk, n, m, l] (l fastest)
    out = (w * vals).sum(axis=(2, 3))
    if mask is not None:
        out[~np.asarray(mask, bool)] = 0
    return out


def touched_components(tab: PolarTables) -> int:
    """Number of distinct Cartesian components (after Friedel folding) that a polarize reads."""
    nk, n, wd = tab.cyc1.shape
    h = tab.cyc1[:, :, None, :].repeat(wd, axis=2).ravel()
    k = tab.cyc2[:, :, :, None].repeat(wd, axis=3).ravel()
    neg = h < 0
    return len(set(zip(np.where(neg, -h, h).tolist(), np.where(neg, -k, k).tolist())))
