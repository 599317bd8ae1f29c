"""Oracle of SIMPLE's particle preparation for alignment, prepimg4align (TEST INFRASTRUCTURE ONLY).

Restates, in numpy, ``prepimg4align`` (``src/main/simple_strategy2D3D_common.f90:316-366``) for a
2-D particle image and the image-object routines it calls:

* ``noise_norm`` (``src/main/simple_image.f90:6623-6640``) with the background mask ``lmsk`` of
  ``disc`` / ``cendist`` (``simple_image.f90:497-515, 2719-2748``; ``simple_builder.f90:237``);
* ``fft`` / ``ifft`` (``fwd_ft`` / ``bwd_ft``, ``simple_image.f90:5773-5791``, quadrant swap
  ``shift_phorig`` ``:6016-6114``);
* ``shift2Dserial`` (``simple_image.f90:6149-6174``, shift constant ``:438-447``);
* the CTF phase flip ``ctf%apply_serial(img, 'flip', ...)`` (``src/main/simple_ctf.f90:312-403``,
  ``eval`` ``:147-163``);
* ``clip`` in real space (``simple_image.f90:6471-6516``);
* the soft mask ``mask(msk, 'soft'[, inner, width])`` (``simple_image.f90:6208-6300``) with
  ``zero_background`` (``:5176-5203``), the median of ``selec`` (``src/utils/simple_math.f90:
  2204-2261``) and ``cosedge`` / ``cosedge_inner`` (``simple_math.f90:1089-1103, 1128-1140``);
* ``div_by_instrfun`` (``src/main/simple_polarizer.f90:111-116``) when griddev is set.

The reference works in FP32 (FFTW single precision).  This restatement keeps the FP32 scalar
formulas where their rounding decides a branch (mask weights, distances, the median) and does
the transforms in FP64, so its outputs agree with the reference to FP32 rounding of the
transforms.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this file; the product (the HIP library) never calls into it.

Parity status: every function is pinned by ``tests/test_oracle_prep.py``: ``selec`` against
order statistics (and the reference's skipped final swap on a traced case), ``noise_norm`` by
closed forms of the background mean and the corrected two-pass variance, the transforms by
``image_fft`` round trips, ``shift2Dserial`` by integer shifts against ``np.roll``, the flip by
the first zero of a Cs-free CTF and the FP64 evaluation, ``clip`` by the centre pixel, the mask weights by their
closed form at chosen radii, and the whole chain by invariants (identity settings reduce it to
``image_fft`` of the noise-normalised image).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

import polarizer_oracle as PZ

f32 = np.float32
PI32 = f32(np.arccos(f32(-1.0)))          # simple_defs.f90:20 (PI = acos(-1.), single precision)
TINY = f32(1e-10)                          # simple_defs.f90:28
COSMSKHALFWIDTH = f32(6.0)                 # simple_defs.f90:130
SHTHRESH = f32(0.001)                      # simple_strategy2D3D_common.f90:31

CTFFLAG_NO, CTFFLAG_YES, CTFFLAG_FLIP = 0, 1, 2


@dataclass
class PrepCfg:
    box: int                 # input image size
    boxmatch: int            # output transform size (<= box, even)
    msk: float               # mask radius in pixels (params%msk)
    inner: float = 0.0       # l_innermsk: inner radius (0: no inner mask)
    width: float = 10.0      # inner mask width (params%width)
    focusmsk: float = 0.0    # l_focusmsk: soft-mask radius instead of msk (0: unused)
    ctfflag: int = CTFFLAG_NO
    griddev: bool = False


# ---------------------------------------------------------------------------------------------
# selec (simple_math.f90:2204-2261): the reference's quickselect, restated step for step.  Its
# final two-element test compares ir - 1 with 1 (not ir with l + 1), so a two-element tail away
# from the start is left unsorted; the returned value is whatever arr(k) holds at that point.
# The array is permuted in place, as the reference's intent(inout) argument is.
# ---------------------------------------------------------------------------------------------
def selec(k: int, arr: np.ndarray) -> np.float32:
    a_ = arr                      # 0-based view; indices below are the reference's 1-based ones
    n = a_.size
    l, ir = 1, n

    def g(i):
        return a_[i - 1]

    def s(i, v):
        a_[i - 1] = v

    while True:
        if ir - l <= 1:
            if ir - 1 == 1:
                if g(ir) < g(l):
                    t = g(l)
                    s(l, g(ir))
                    s(ir, t)
            return f32(g(k))
        mid = (l + ir) // 2
        t = g(mid)
        s(mid, g(l + 1))
        s(l + 1, t)
        if g(l) > g(ir):
            t = g(l); s(l, g(ir)); s(ir, t)
        if g(l + 1) > g(ir):
            t = g(l + 1); s(l + 1, g(ir)); s(ir, t)
        if g(l) > g(l + 1):
            t = g(l); s(l, g(l + 1)); s(l + 1, t)
        i, j = l + 1, ir
        a = g(l + 1)
        while True:
            i += 1
            while g(i) < a:
                i += 1
            j -= 1
            while g(j) > a:
                j -= 1
            if j < i:
                break
            t = g(i); s(i, g(j)); s(j, t)
        s(l + 1, g(j))
        s(j, a)
        if j >= k:
            ir = j - 1
        if j <= k:
            l = i


# ---------------------------------------------------------------------------------------------
# real-space helpers (images are C arrays img[y, x] = Fortran rmat(x, y))
# ---------------------------------------------------------------------------------------------
def disc_mask(ldim: int, radius: float) -> np.ndarray:
    """disc_2 (simple_image.f90:497-515) via cendist (:2719-2748): lmsk = sqrt((i-c)^2 + (j-c)^2)
    <= radius with c = ldim/2 + 1 (1-based, FP32).  The squared distance is an exact integer, so
    the FP64 square root rounded to FP32 is the FP32 square root."""
    c = ldim // 2 + 1
    i = np.arange(1, ldim + 1)
    d2 = (i[None, :] - c) ** 2 + (i[:, None] - c) ** 2
    dist = np.sqrt(d2.astype(np.float64)).astype(np.float32)
    return dist <= f32(radius)


def noise_norm(img: np.ndarray, lmsk: np.ndarray):
    """noise_norm (simple_image.f90:6623-6640): background mean removed (if |ave| > TINY), then
    divided by the background standard deviation of the corrected two-pass formula (if var > 0).
    Sums in FP64 (the reference: FP32).  Returns (image, sdev_noise, ave, sdev) in FP32."""
    x = img.astype(np.float32)
    bg = ~lmsk
    npix = int(bg.sum())
    ave = f32(np.float32(x[bg].astype(np.float64).sum()) / f32(npix))
    if abs(ave) > TINY:
        x = (x - ave).astype(np.float32)
    else:
        ave = f32(0.0)
    ep = f32(x[bg].astype(np.float64).sum())
    var = f32(np.square(x[bg].astype(np.float64)).sum())
    var = f32((var - ep * ep / f32(npix)) / (f32(npix) - f32(1.0)))
    sdev = f32(0.0)
    used = f32(1.0)
    if np.isfinite(var):
        sdev = f32(np.sqrt(var))
        if var > 0:
            x = (x / sdev).astype(np.float32)
            used = sdev
    return x, sdev, ave, used


def ifft_image(cmat: np.ndarray) -> np.ndarray:
    """bwd_ft (simple_image.f90:5784-5791): unnormalised c2r (FFTW backward) then the quadrant
    swap.  The c2r treats the imaginary parts of the h = 0 and h = ldim/2 columns as zero after
    the transform along y (numpy's irfft2 does the same)."""
    ny = cmat.shape[0]
    nx = 2 * (cmat.shape[1] - 1)
    r = np.fft.irfft2(cmat.astype(np.complex128), s=(ny, nx)) * float(nx * ny)
    return np.fft.fftshift(r).astype(np.float32)


def shift_consts(ldim: int) -> np.float32:
    """shconst (simple_image.f90:438-447) for an even dimension: PI / real(ldim/2.), FP32."""
    return f32(PI32 / f32(f32(ldim) / f32(2.0)))


def shift2Dserial(cmat: np.ndarray, shvec, ldim: int) -> np.ndarray:
    """shift2Dserial_1 (simple_image.f90:6149-6174): cmat(h, k) *= exp(i (h sh1 + k sh2)) with
    sh = real(shvec * shconst, dp), the phase built in FP64 as (ck hcos - sk hsin, ck hsin +
    sk hcos) and cast to complex(sp); loop_lims(2) covers every stored component."""
    sc = shift_consts(ldim)
    sh = (np.asarray(shvec, np.float32) * sc).astype(np.float64)
    nh = cmat.shape[1]
    h = np.arange(nh, dtype=np.float64)
    kp = np.arange(cmat.shape[0])
    k = np.where(kp < ldim // 2, kp, kp - ldim).astype(np.float64)
    hc, hs = np.cos(h * sh[0]), np.sin(h * sh[0])
    ck, sk = np.cos(k * sh[1]), np.sin(k * sh[1])
    re = ck[:, None] * hc[None, :] - sk[:, None] * hs[None, :]
    im = ck[:, None] * hs[None, :] + sk[:, None] * hc[None, :]
    ph = (re + 1j * im).astype(np.complex64)
    return (cmat.astype(np.complex64) * ph).astype(np.complex64)


def ctf_values(ldim: int, p) -> np.ndarray:
    """ctf%eval at every stored component of apply_serial's loops (simple_ctf.f90:312-403,
    147-163, 185-189) in the reference's real(sp) arithmetic: hinv = real(h) / ldim, kinv likewise,
    s2 = hinv^2 + kinv^2, ang = atan2(real(k), real(h)), value = sin(chi(s2, ang) + ampc) with the
    constants of the ctf constructor (:54-81, 522-525) and the phase shift passed as given."""
    nh = ldim // 2 + 1
    h = np.arange(nh).astype(np.float32)
    kp = np.arange(ldim)
    k = np.where(kp < ldim // 2, kp, kp - ldim).astype(np.float32)
    H, K = np.meshgrid(h, k)
    inv = f32(1.0) / f32(ldim)
    hinv, kinv = (H * inv).astype(np.float32), (K * inv).astype(np.float32)
    s2 = (hinv * hinv + kinv * kinv).astype(np.float32)
    ang = np.arctan2(K, H).astype(np.float32)
    a = f32(1000.0) * f32(p.kv)
    wl = f32(f32(12.26) / f32(np.sqrt(f32(a + f32(f32(0.9784) * f32(a * a)) / f32(1.0e6)))))
    wl = f32(wl / f32(p.smpd))
    cs = f32(f32(f32(p.cs) * f32(1.0e7)) / f32(p.smpd))
    dfx = f32(f32(f32(p.dfx) * f32(1.0e4)) / f32(p.smpd))
    dfy = f32(f32(f32(p.dfy) * f32(1.0e4)) / f32(p.smpd))
    angast = f32(f32(f32(p.angast) / f32(180.0)) * PI32)
    fa = f32(p.fraca)
    ampc = f32(np.arctan(f32(fa / f32(np.sqrt(f32(f32(1.0) - fa * fa))))))
    df = (f32(0.5) * (dfx + dfy + np.cos(f32(2.0) * (ang - angast)).astype(np.float32) * (dfx - dfy))).astype(np.float32)
    chi = (PI32 * wl * s2 * (df - f32(0.5) * wl * wl * s2 * cs) + f32(getattr(p, "phshift", 0.0))).astype(np.float32)
    return np.sin((chi + ampc).astype(np.float32)).astype(np.float32)


def ctf_values_fp64(ldim: int, p) -> np.ndarray:
    """The same CTF in FP64 (to keep test inputs away from zero crossings)."""
    nh = ldim // 2 + 1
    h = np.arange(nh, dtype=np.float64)
    kp = np.arange(ldim)
    k = np.where(kp < ldim // 2, kp, kp - ldim).astype(np.float64)
    H, K = np.meshgrid(h, k)
    s2 = (H / ldim) ** 2 + (K / ldim) ** 2
    ang = np.arctan2(K, H)
    wl = 12.26 / np.sqrt(1000.0 * p.kv + 0.9784 * (1000.0 * p.kv) ** 2 / 1.0e6) / p.smpd
    cs = p.cs * 1.0e7 / p.smpd
    dfx, dfy = p.dfx * 1.0e4 / p.smpd, p.dfy * 1.0e4 / p.smpd
    angast = np.deg2rad(p.angast)
    ampc = np.arctan(p.fraca / np.sqrt(1.0 - p.fraca * p.fraca))
    df = 0.5 * (dfx + dfy + np.cos(2.0 * (ang - angast)) * (dfx - dfy))
    return np.sin(np.pi * wl * s2 * (df - 0.5 * wl * wl * s2 * cs) + getattr(p, "phshift", 0.0) + ampc)


def ctf_flip(cmat: np.ndarray, ldim: int, p) -> np.ndarray:
    """apply_serial(img, 'flip') (simple_ctf.f90:345-355): cmat *= sign(1., tval)."""
    t = ctf_values(ldim, p)
    return (cmat * np.where(np.signbit(t), -1.0, 1.0)).astype(np.complex64)


def clip(img: np.ndarray, ldim_out: int) -> np.ndarray:
    """clip, real space (simple_image.f90:6499-6511): starts = (ldim_in - ldim_out)/2 + 1."""
    s = (img.shape[0] - ldim_out) // 2
    return img[s:s + ldim_out, s:s + ldim_out].copy()


def zero_background(img: np.ndarray) -> np.float32:
    """zero_background (simple_image.f90:5176-5203), 2-D, in place: med = mean of selec(ldim/2)
    over the edges x = 1, x = ldim, y = 1, y = ldim, in that order (selec permutes each edge in
    the image, so later edges see the earlier permutations at the corners); subtract med if
    |med| > TINY.  Returns med."""
    n = img.shape[0]
    k = n // 2
    vals = []
    for sl in ((slice(None), 0), (slice(None), n - 1), (0, slice(None)), (n - 1, slice(None))):
        edge = img[sl].copy()
        vals.append(selec(k, edge))
        img[sl] = edge
    med = f32(f32(f32(f32(vals[0] + vals[1]) + vals[2]) + vals[3]) / f32(4.0))
    if abs(med) > TINY:
        img -= med
    return med


def cosedge(x, y, box: int, mskrad) -> np.float32:
    """cosedge_1 (simple_math.f90:1089-1103), FP32."""
    x, y, mskrad = f32(x), f32(y), f32(mskrad)
    maxrad = f32(box // 2)
    rad = f32(np.sqrt(f32(x * x + y * y)))
    width = f32(f32(2.0) * f32(maxrad - mskrad))
    if rad >= maxrad:
        return f32(0.0)
    if rad >= f32(maxrad - width):
        return f32(f32(np.cos(f32(f32(f32(rad - f32(maxrad - width)) / width) * PI32)) + f32(1.0)) / f32(2.0))
    return f32(1.0)


def cosedge_inner(x, y, width, mskrad) -> np.float32:
    """cosedge_inner_1 (simple_math.f90:1128-1140), FP32."""
    x, y, width, mskrad = f32(x), f32(y), f32(width), f32(mskrad)
    rad = f32(np.sqrt(f32(x * x + y * y)))
    if rad < f32(mskrad - width):
        return f32(0.0)
    if rad > mskrad:
        return f32(1.0)
    return f32(f32(f32(1.0) + f32(np.cos(f32(f32(PI32 * f32(mskrad - rad)) / width)))) / f32(2.0))


def soft_mask_weights(ldim: int, mskrad: float, inner: float = 0.0, width: float = 10.0) -> np.ndarray:
    """The weights mask(mskrad, 'soft'[, inner, width]) multiplies by (simple_image.f90:6208-6286),
    2-D: e(i, j) for i, j <= ldim/2 from cis(i) = -ldim/2 + (i - 1), mirrored to ir = ldim + 1 - i
    and jr; minlen = min(nint(2 (mskrad + COSMSKHALFWIDTH)), ldim).  Entries > 0.9999 are left
    untouched by the reference; they are returned as 1 here (the caller skips them)."""
    mskrad = f32(mskrad)
    minlen = min(int(np.floor(float(f32(f32(2.0) * f32(mskrad + COSMSKHALFWIDTH))) + 0.5)), ldim)
    w = np.ones((ldim, ldim), np.float32)
    half = ldim // 2
    for i in range(1, half + 1):
        ci = f32(f32(-f32(ldim) / f32(2.0)) + f32(i - 1))
        for j in range(1, half + 1):
            cj = f32(f32(-f32(ldim) / f32(2.0)) + f32(j - 1))
            e = cosedge(ci, cj, minlen, mskrad)
            if inner > 0.0:
                e = f32(e * cosedge_inner(ci, cj, width, inner))
            if e > f32(0.9999):
                continue
            ir, jr = ldim + 1 - i, ldim + 1 - j
            for (a, b) in ((i, j), (i, jr), (ir, j), (ir, jr)):
                w[b - 1, a - 1] = e      # img[y, x] = rmat(x, y)
    return w


def apply_soft_mask(img: np.ndarray, w: np.ndarray) -> np.ndarray:
    """rmat = e * rmat where the weight is <= 0.9999 (weights of 1 leave the pixel untouched)."""
    out = img.copy()
    m = w != f32(1.0)
    out[m] = (w[m] * img[m]).astype(np.float32)
    return out


def instrfun_weights(ldim: int) -> np.ndarray:
    """gen_instrfun_img (simple_gridding.f90:18-55) with the polarizer's window: w(i) w(j) in FP32."""
    kb = PZ.kb_params()
    i = np.arange(1, ldim + 1)
    w = PZ.instr((i - (ldim // 2 + 1)) / ldim, kb).astype(np.float32)
    return (w[:, None] * w[None, :]).astype(np.float32)


# ---------------------------------------------------------------------------------------------
# prepimg4align
# ---------------------------------------------------------------------------------------------
def prepimg4align(img: np.ndarray, shvec, cfg: PrepCfg, ctfp=None, return_steps: bool = False):
    """prepimg4align (simple_strategy2D3D_common.f90:316-366) for one 2-D particle image given as
    img[y, x] (box x box, FP32).  shvec = the particle's (x, y); ctfp: CtfParams of the a6
    restatement (pftcc_oracle.CtfParams) when cfg.ctfflag == CTFFLAG_YES.  Returns the transform
    of the boxmatch image as cmat [k_phys][h_phys] (complex64), ready for polarize."""
    B, M = cfg.box, cfg.boxmatch
    lmsk = disc_mask(B, cfg.msk)
    x, _, _, _ = noise_norm(img, lmsk)                                         # :334
    c = PZ.image_fft(x)                                                        # :336
    sx, sy = f32(shvec[0]), f32(shvec[1])
    if abs(sx) > SHTHRESH or abs(sy) > SHTHRESH:                               # :338
        c = shift2Dserial(c, (-sx, -sy), B)
    if cfg.ctfflag == CTFFLAG_YES:                                             # :339-348
        c = ctf_flip(c, B, ctfp)
    r = ifft_image(c)                                                          # :350
    r = clip(r, M)                                                             # :352
    zero_background(r)                                                         # :354-360 (mask 'soft')
    if cfg.inner > 0.0:
        w = soft_mask_weights(M, cfg.msk, cfg.inner, cfg.width)
    elif cfg.focusmsk > 0.0:
        w = soft_mask_weights(M, cfg.focusmsk)
    else:
        w = soft_mask_weights(M, cfg.msk)
    r = apply_soft_mask(r, w)
    if cfg.griddev:                                                            # :361-363
        r = (r / instrfun_weights(M)).astype(np.float32)
    out = PZ.image_fft(r).astype(np.complex128)                                # :365
    out = out * float(f32(f32(M * M) / f32(B * B)))                            # :366
    if return_steps:
        return out.astype(np.complex64), {"normalised": x, "clipped_masked": r}
    return out.astype(np.complex64)
