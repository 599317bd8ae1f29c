"""FP64 oracle of SIMPLE's polar-FT correlation search (TEST INFRASTRUCTURE ONLY).

This module restates, in numpy float64, the arithmetic of
``src/main/simple_polarft_corrcalc.f90`` (module ``simple_polarft_corrcalc``) and the
helpers it calls (``simple_ctf.f90``, ``simple_magic_boxes.f90``, ``simple_math.f90``).
Each function cites the reference lines it follows.

It is the checker, never the product: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product path (the HIP library behind
``include/pftcc.h``) never calls into this file.

Conventions
-----------
* Arrays use the C view of the reference's Fortran layout: a polar FT is
  ``pft[k - kfrom, j]`` (complex), j = 0..pftsz-1 the half-ring angle index, k the shell.
  Fortran ``pft(pftsz, kfrom:kto)`` column-major has exactly this memory order.
* ``irot`` is 1-based as in the reference; ``rho = irot - 1`` is the 0-based rotation.
* Inputs that the reference holds in FP32 (polar FTs, CTF matrices, polar grid) are taken
  as FP32 values promoted to FP64, which is how the reference's ``_8`` routines see them
  (``simple_polarft_corrcalc.f90:1713,1221``).

Parity status (each function is pinned by tests/test_oracle_*.py; see the test docstrings):
every function here is pinned by at least one of: two independent formulations that must
agree, a closed form, brute force on tiny inputs, or central finite differences.
``prep_matchfilt`` (``:987-1095``) and its Savitzky-Golay smoother (``simple_math.f90:3511-3588``)
are pinned by closed forms: the 7-point quartic weights (5, -30, 75, 131, 75, -30, 5)/231, exact
reproduction of quartic polynomials inside the window, the border rules, the per-shell power of
the filtered particle (ssnr/(ssnr+1), or k (1 + ssnr <|CTF|^2>) with pssnr) and the identity for
iref = 0.  The reference's FP32 rounding inside ``prep_matchfilt`` (FP32 ``ludcmp`` weights, FP32
scaling) is not reproduced bit for bit.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import math
from dataclasses import dataclass, field

import numpy as np

f32 = np.float32

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.sinf.restype = ctypes.c_float
_libm.sinf.argtypes = [ctypes.c_float]
_libm.cosf.restype = ctypes.c_float
_libm.cosf.argtypes = [ctypes.c_float]
_libm_sinf = _libm.sinf
_libm_cosf = _libm.cosf

# ---------------------------------------------------------------------------------------------
# a3: FFT-friendly ring length.  Table of sizes from simple_magic_boxes.f90:14-19 (a data table
# that has to be identical for the derived pftsz to be identical).
# ---------------------------------------------------------------------------------------------
BOXSIZES = (
    32, 36, 40, 48, 52, 56, 64, 66, 70, 72, 80, 84, 88, 100, 104, 108, 112, 120, 128, 130, 132,
    140, 144, 150, 160, 162, 168, 176, 180, 182, 192, 200, 208, 216, 220, 224, 240, 256, 264, 288,
    300, 308, 320, 324, 336, 338, 352, 364, 384, 400, 420, 432, 448, 450, 462, 480, 486, 500, 504,
    512, 520, 528, 546, 560, 576, 588, 600, 640, 648, 650, 660, 672, 686, 700, 702, 704, 720, 726,
    728, 750, 768, 770, 784, 800, 810, 840, 882, 896, 910, 924, 936, 972, 980, 1008, 1014, 1020,
    1024, 1296, 1536, 1728, 1944, 2048, 2304, 2592, 3072, 3200, 3456, 3888, 4096, 4608, 5000, 5184,
    6144, 6250, 6400, 6912, 7776, 8192, 9216, 10240, 12288, 12500)

PI32 = f32(math.acos(-1.0))          # simple_defs.f90:20  PI = acos(-1.)  (real(sp))
TWOPI32 = f32(2.0) * PI32            # simple_defs.f90:23
DPI = math.acos(-1.0)                # simple_defs.f90:21


def fortran_nint(x: float) -> int:
    """Fortran NINT: round half away from zero."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def round2even(val: float) -> int:
    """simple_math.f90:391-405 (round2even), evaluated on a real(sp) argument."""
    val = float(f32(val))
    rounded = fortran_nint(val)
    if rounded % 2 == 0:
        return rounded
    if abs(float(rounded - 1) - val) <= abs(float(rounded + 1) - val):
        return rounded - 1
    return rounded + 1


def _locate(arr, x: int) -> int:
    """simple_math.f90 locate_2 (bisection); returns a 1-based index."""
    n = len(arr)
    jl, ju = 0, n + 1
    while ju - jl > 1:
        jm = (ju + jl) // 2
        if (arr[n - 1] >= arr[0]) == (x >= arr[jm - 1]):
            jl = jm
        else:
            ju = jm
    if x == arr[0]:
        return 1
    if x == arr[n - 1]:
        return n - 1
    return jl


def _find(arr, x: int) -> int:
    """simple_math.f90:2069-2084 (find_2): closest element, ties towards the larger entry."""
    n = len(arr)
    j = max(1, _locate(arr, x))
    dist1 = arr[j - 1] - x
    if j < n:
        dist2 = arr[j] - x
        if abs(dist1) >= abs(dist2):
            j = j + 1
    return j


def find_magic_box(trial_box: int) -> int:
    """simple_magic_boxes.f90:53-62 (find_magic_box)."""
    best = BOXSIZES[_find(BOXSIZES, trial_box) - 1]
    if best == BOXSIZES[-1] and trial_box - best > 16:
        best = round2even(float(trial_box))
    return best


def magic_pftsz(ring2: int) -> int:
    """simple_magic_boxes.f90:25-37 (magic_pftsz).  a = PI*real(ring2) in real(sp)."""
    a = float(PI32 * f32(ring2))
    pftsz_old = round2even(a)
    pftsz = find_magic_box(fortran_nint(a))
    if abs(pftsz - pftsz_old) / float(pftsz_old) > 0.1:
        return pftsz_old
    return pftsz


def default_ring2(box: int) -> int:
    """simple_parameters.f90:1182-1189: ring2 = round2even(real(box/2-2)) when msk/ring2 unset."""
    return round2even(float(box // 2 - 2))


def find_boxmatch(box: int, msk: float) -> int:
    """simple_magic_boxes.f90:66-75 (find_boxmatch)."""
    if box > 2 * (fortran_nint(msk) + 10):
        return min(find_magic_box(2 * (fortran_nint(msk) + 10)), box)
    return box


# ---------------------------------------------------------------------------------------------
# a2: polar grid and shift argument transfer (new, simple_polarft_corrcalc.f90:251-317)
# ---------------------------------------------------------------------------------------------
@dataclass
class Grid:
    pftsz: int
    kfrom: int
    kto: int
    ldim: int
    angtab_rad: np.ndarray = field(repr=False)   # f32 [nrots]
    angtab_deg: np.ndarray = field(repr=False)   # f32 [nrots]
    polar_x: np.ndarray = field(repr=False)      # f32 [nk, nrots]   polar(irot,k)
    polar_y: np.ndarray = field(repr=False)      # f32 [nk, nrots]   polar(irot+nrots,k)
    argtransf_x: np.ndarray = field(repr=False)  # f64 [nk, pftsz]
    argtransf_y: np.ndarray = field(repr=False)  # f64 [nk, pftsz]
    shellone_x: np.ndarray = field(repr=False)   # f64 [pftsz]
    shellone_y: np.ndarray = field(repr=False)   # f64 [pftsz]

    @property
    def nrots(self) -> int:
        return 2 * self.pftsz

    @property
    def nk(self) -> int:
        return self.kto - self.kfrom + 1

    @property
    def ks(self) -> np.ndarray:
        return np.arange(self.kfrom, self.kto + 1)


def make_grid(pftsz: int, kfrom: int, kto: int, ldim: int) -> Grid:
    """new(), simple_polarft_corrcalc.f90:261-317.

    angtab(irot) = real(irot-1)*ang, ang = twopi/real(nrots) (real(sp)); polar x = sin(angtab)*k,
    y = -cos(angtab)*k (real(sp)); argtransf = real(polar,dp) * DPI/real(ldim/2,dp).
    """
    nrots = 2 * pftsz
    ang = f32(TWOPI32 / f32(nrots))
    irot = np.arange(nrots, dtype=np.float32)
    angtab = (irot * ang).astype(np.float32)
    # sin/cos of the real(sp) angle through libm's sinf/cosf, which the reference's gfortran
    # build calls for real(sp) sin/cos (numpy's float32 sin rounds differently)
    s = np.array([_libm_sinf(float(a)) for a in angtab], dtype=np.float32)
    c = np.array([_libm_cosf(float(a)) for a in angtab], dtype=np.float32)
    ks = np.arange(kfrom, kto + 1, dtype=np.float32)
    polar_x = (s[None, :] * ks[:, None]).astype(np.float32)
    polar_y = (-c[None, :] * ks[:, None]).astype(np.float32)
    A = DPI / float(ldim // 2)
    argtransf_x = polar_x[:, :pftsz].astype(np.float64) * A
    argtransf_y = polar_y[:, :pftsz].astype(np.float64) * A
    shellone_x = s[:pftsz].astype(np.float64) * A
    shellone_y = (-c[:pftsz]).astype(np.float64) * A
    angtab_deg = ((angtab / PI32) * f32(180.0)).astype(np.float32)  # rad2deg_1, simple_math.f90
    return Grid(pftsz, kfrom, kto, ldim, angtab, angtab_deg, polar_x, polar_y,
                argtransf_x, argtransf_y, shellone_x, shellone_y)


def get_roind(grid: Grid, rot: float) -> int:
    """get_roind, simple_polarft_corrcalc.f90:635-645 (1-based index of the closest angle)."""
    dists = np.abs(grid.angtab_deg - f32(rot)).astype(np.float32)
    dists = np.where(dists > 180.0, f32(360.0) - dists, dists)
    return int(np.argmin(dists)) + 1


def pxls_p_shell(kfrom: int, kto: int) -> np.ndarray:
    """setup_pxls_p_shell, simple_polarft_corrcalc.f90:879-893 (positive quadrant h,k>=1)."""
    out = np.zeros(kto - kfrom + 1)
    for h in range(1, kto + 1):
        for k in range(1, kto + 1):
            sh = fortran_nint(float(f32(math.sqrt(float(h * h + k * k)))))
            if kfrom <= sh <= kto:
                out[sh - kfrom] += 1.0
    return out


# ---------------------------------------------------------------------------------------------
# a6: |CTF| matrices in the reference frame
# ---------------------------------------------------------------------------------------------
@dataclass
class CtfParams:
    smpd: float
    kv: float
    cs: float
    fraca: float
    dfx: float
    dfy: float
    angast: float
    phshift: float = 0.0
    phaseplate: bool = False


def ctf_abs_mat(grid: Grid, p: CtfParams) -> np.ndarray:
    """create_polar_absctfmats (simple_polarft_corrcalc.f90:906-959) for one particle, in FP64.

    ctf constructor/init/eval (simple_ctf.f90:54-81, 88-98, 147-163, 185-189, 522-525):
      wl = 12.26/sqrt(1000kV + 0.9784(1000kV)^2/1e6)/smpd, Cs = Cs*1e7/smpd,
      df = dfx,dfy *1e4/smpd, angast in radians, amp-contrast phase = atan(fa/sqrt(1-fa^2)),
      chi = pi*wl*s2*(df(ang) - 0.5*wl^2*s2*Cs) + phshift,  value = |sin(chi + ampphase)|,
      df(ang) = 0.5*(dfx + dfy + cos(2(ang - angast))*(dfx - dfy)),
      s2 = (x/ldim)^2 + (y/ldim)^2, ang = atan2(y, x) at the ref-frame polar grid.
    Returns float64 [nk, pftsz] (evaluated on the FP32 grid promoted to FP64).
    """
    x = grid.polar_x[:, :grid.pftsz].astype(np.float64)
    y = grid.polar_y[:, :grid.pftsz].astype(np.float64)
    inv = 1.0 / float(grid.ldim)
    s2 = (x * inv) ** 2 + (y * inv) ** 2
    ang = np.arctan2(y, x)
    kv = p.kv
    wl = 12.26 / math.sqrt(1000.0 * kv + 0.9784 * (1000.0 * kv) ** 2 / 1.0e6) / p.smpd
    cs = p.cs * 1.0e7 / p.smpd
    dfx = p.dfx * 1.0e4 / p.smpd
    dfy = p.dfy * 1.0e4 / p.smpd
    angast = p.angast / 180.0 * DPI
    ampc = math.atan(p.fraca / math.sqrt(1.0 - p.fraca * p.fraca))
    df = 0.5 * (dfx + dfy + np.cos(2.0 * (ang - angast)) * (dfx - dfy))
    phsh = p.phshift if p.phaseplate else 0.0
    chi = DPI * wl * s2 * (df - 0.5 * wl * wl * s2 * cs) + phsh
    return np.abs(np.sin(chi + ampc))


def ctf_chi_abs(grid: Grid, p: CtfParams) -> np.ndarray:
    """|argument| of the sine in ctf_abs_mat (used for FP32 error bounds in tests)."""
    x = grid.polar_x[:, :grid.pftsz].astype(np.float64)
    y = grid.polar_y[:, :grid.pftsz].astype(np.float64)
    s2 = (x / grid.ldim) ** 2 + (y / grid.ldim) ** 2
    kv = p.kv
    wl = 12.26 / math.sqrt(1000.0 * kv + 0.9784 * (1000.0 * kv) ** 2 / 1.0e6) / p.smpd
    dfmax = max(p.dfx, p.dfy) * 1.0e4 / p.smpd
    cs = p.cs * 1.0e7 / p.smpd
    return DPI * wl * s2 * (dfmax + 0.5 * wl * wl * s2 * cs) + abs(p.phshift) + 1.0


# ---------------------------------------------------------------------------------------------
# a11: shift phase matrix (gen_shmat_8, :1098-1120; gen_shmat :1123-1130)
# ---------------------------------------------------------------------------------------------
def shmat_8(grid: Grid, shvec) -> np.ndarray:
    """gen_shmat_8: first shell analytic, then de Moivre recurrence with the k=1 phase.

    Returns complex128 [nk, pftsz].
    """
    sx, sy = float(shvec[0]), float(shvec[1])
    out = np.empty((grid.nk, grid.pftsz), dtype=np.complex128)
    arg0 = grid.argtransf_x[0] * sx + grid.argtransf_y[0] * sy
    out[0] = np.cos(arg0) + 1j * np.sin(arg0)
    arg1 = grid.shellone_x * sx + grid.shellone_y * sy
    step = np.cos(arg1) + 1j * np.sin(arg1)
    for i in range(1, grid.nk):
        out[i] = out[i - 1] * step
    return out


def shmat_8_many(grid: Grid, shifts) -> np.ndarray:
    """gen_shmat_8 (:1098-1120) for a list of shifts at once: the same first-shell phase and de
    Moivre recurrence over k, vectorised over the shifts.  shifts [ns, 2] -> complex128 [ns, nk, n].
    """
    sh = np.asarray(shifts, dtype=np.float64).reshape(-1, 2)
    sx, sy = sh[:, 0:1], sh[:, 1:2]
    out = np.empty((sh.shape[0], grid.nk, grid.pftsz), dtype=np.complex128)
    arg0 = grid.argtransf_x[0][None, :] * sx + grid.argtransf_y[0][None, :] * sy
    out[:, 0] = np.cos(arg0) + 1j * np.sin(arg0)
    arg1 = grid.shellone_x[None, :] * sx + grid.shellone_y[None, :] * sy
    step = np.cos(arg1) + 1j * np.sin(arg1)
    for i in range(1, grid.nk):
        out[:, i] = out[:, i - 1] * step
    return out


def shmat_closed(grid: Grid, shvec) -> np.ndarray:
    """Closed form exp(i*(ax*sx + ay*sy)) that gen_shmat_8's recurrence approximates."""
    arg = grid.argtransf_x * float(shvec[0]) + grid.argtransf_y * float(shvec[1])
    return np.exp(1j * arg)


# ---------------------------------------------------------------------------------------------
# Helpers: full ring and the rotational correlation primitives
# ---------------------------------------------------------------------------------------------
def full_ring(p: np.ndarray) -> np.ndarray:
    """Hermitian full ring P[j+n] = conj(P[j]) along the last axis (Friedel symmetry)."""
    return np.concatenate([p, np.conj(p)], axis=-1)


def corrk_for_rot(q: np.ndarray, p: np.ndarray, irot: int) -> np.ndarray:
    """calc_corrk_for_rot(_8) (:1336-1390) for every shell at once.

    Re sum_j q[k,j] conj(P[k, j+rho]) with the four-case split written as one modular index.
    q, p: [nk, n] -> [nk].
    """
    n = q.shape[-1]
    rho = irot - 1
    P = full_ring(p.astype(np.complex128))
    idx = (np.arange(n) + rho) % (2 * n)
    return np.real(np.sum(q.astype(np.complex128) * np.conj(P[..., idx]), axis=-1))


def corrk_for_rot_fourcase(q: np.ndarray, p: np.ndarray, irot: int) -> np.ndarray:
    """The four explicit branches of calc_corrk_for_rot_8 (:1362-1390), shell by shell.

    Kept separately so tests can check the modular-index form against the reference's case
    split (irot==1, irot<=pftsz, irot==pftsz+1, irot>pftsz+1).
    """
    n = q.shape[-1]
    q = q.astype(np.complex128)
    p = p.astype(np.complex128)
    rot = irot - n if irot >= n + 1 else irot
    out = np.empty(q.shape[0])
    for k in range(q.shape[0]):
        if irot == 1:
            tmp = np.sum(q[k] * np.conj(p[k]))
        elif irot <= n:
            tmp = np.sum(q[k, 0:n - rot + 1] * np.conj(p[k, rot - 1:n]))
            tmp += np.sum(q[k, n - rot + 1:n] * p[k, 0:rot - 1])
        elif irot == n + 1:
            tmp = np.sum(q[k] * p[k])
        else:
            tmp = np.sum(q[k, 0:n - rot + 1] * p[k, rot - 1:n])
            tmp += np.sum(q[k, n - rot + 1:n] * np.conj(p[k, 0:rot - 1]))
        out[k] = tmp.real
    return out


def corrk_all_rots_direct(q: np.ndarray, p: np.ndarray) -> np.ndarray:
    """Brute force O(R*K*n): C_k(rho) for every rho; returns [nk, 2n]."""
    n = q.shape[-1]
    P = full_ring(p.astype(np.complex128))
    q = q.astype(np.complex128)
    out = np.empty((q.shape[0], 2 * n))
    j = np.arange(n)
    for rho in range(2 * n):
        out[:, rho] = np.real(np.sum(q * np.conj(P[:, (j + rho) % (2 * n)]), axis=-1))
    return out


# ---------------------------------------------------------------------------------------------
# a7 / a9: the reference's FFT route (memoize_fft :864-877, calc_corrs_over_k :1132-1162)
# ---------------------------------------------------------------------------------------------
def fft_factors(n: int) -> np.ndarray:
    """new(), :410-413: fft_factors(j) = exp(-i*pi*(j-1)/n)."""
    return np.exp(-1j * DPI * np.arange(n) / n)


def memoize_fft(p: np.ndarray):
    """memoize_fft (:864-877): PA = r2c_n(Re p) (n/2+1 bins), PB = c2c_n(Im p * fft_factors).

    p: [nk, n] -> (PA [nk, n/2+1], PB [nk, n]) complex128.
    """
    n = p.shape[-1]
    pa = np.fft.rfft(np.real(p).astype(np.float64), axis=-1)
    pb = np.fft.fft(np.imag(p).astype(np.float64) * fft_factors(n), axis=-1)
    return pa, pb


def corrs_over_k_fft(q: np.ndarray, pa: np.ndarray, pb: np.ndarray) -> np.ndarray:
    """calc_k_corrs (:1164-1188) for each shell: [nk, 2n] per-shell correlation vectors.

    product(2t) = 4 conj(A_t) PA_t, product(2t+1) = 4 conj(B_t) PB_t for t = 0..n/2,
    c2r of length 2n (FFTW's unnormalised backward transform), divided by 2*nrots.
    """
    n = q.shape[-1]
    a = np.fft.rfft(np.real(q).astype(np.float64), axis=-1)
    b = np.fft.fft(np.imag(q).astype(np.float64) * fft_factors(n), axis=-1)
    h = n // 2
    prod = np.zeros((q.shape[0], 2 * n + 2), dtype=np.complex128)
    prod[:, 0:2 * h + 1:2] = 4.0 * np.conj(a[:, :h + 1]) * pa[:, :h + 1]
    prod[:, 1:2 * h + 2:2] = 4.0 * np.conj(b[:, :h + 1]) * pb[:, :h + 1]
    # c2r_{2n} reads bins 0..n; numpy's irfft drops Im of bins 0 and n as FFTW does.
    back = np.fft.irfft(prod[:, :n + 1], n=2 * n, axis=-1) * (2 * n)
    return back / (2.0 * 2 * n)


def gencorrs_refs_fft(refs: np.ndarray, ptcl: np.ndarray, ctf, kcorr: int, chunk: int = 256) -> np.ndarray:
    """gencorrs_cc_1 (:1486-1500) of one particle against many references, FP64, through the
    reference's FFT route (prep_ref4corr :961-985 without shell normalisation, memoize_fft
    :864-877, calc_corrs_over_k :1132-1162), vectorised over the references.

    refs [nr, nk, n] complex, ptcl [nk, n] complex, ctf [nk, n] real (the particle's |CTF|) or
    None -> cc [nr, 2n].  S_q and S_p over the first kcorr shells (kfrom..kstop)."""
    n = refs.shape[-1]
    h = n // 2
    p = ptcl[:kcorr].astype(np.complex128)
    pa, pb = memoize_fft(p)
    sp = float(np.sum(np.abs(p) ** 2))
    c = None if ctf is None else np.asarray(ctf, np.float64)[:kcorr]
    ff = fft_factors(n)
    out = np.empty((refs.shape[0], 2 * n))
    for r0 in range(0, refs.shape[0], chunk):
        q = refs[r0:r0 + chunk, :kcorr].astype(np.complex128)
        if c is not None:
            q = q * c[None]
        sq = np.sum(np.abs(q) ** 2, axis=(1, 2))
        a = np.fft.rfft(np.real(q), axis=-1)
        b = np.fft.fft(np.imag(q) * ff, axis=-1)
        prod = np.zeros(q.shape[:2] + (2 * n + 2,), dtype=np.complex128)
        prod[..., 0:2 * h + 1:2] = 4.0 * np.conj(a[..., :h + 1]) * pa[None, :, :h + 1]
        prod[..., 1:2 * h + 2:2] = 4.0 * np.conj(b[..., :h + 1]) * pb[None, :, :h + 1]
        back = np.fft.irfft(prod[..., :n + 1].sum(axis=1), n=2 * n, axis=-1) * (2 * n) / (2.0 * 2 * n)
        out[r0:r0 + chunk] = back / np.sqrt(sq * sp)[:, None]
    return out


# ---------------------------------------------------------------------------------------------
# a17: Savitzky-Golay smoothing used by prep_matchfilt (simple_math.f90:3511-3588)
# ---------------------------------------------------------------------------------------------
SG_NL, SG_NR, SG_M = 3, 3, 4


def savgol_coefficients(nl: int = SG_NL, nr: int = SG_NR, m: int = SG_M) -> dict:
    """Smoothing weights c[o], o = -nl..nr, of the least-squares polynomial of order m fitted
    to the window and evaluated at its centre (savgol with ld = 0, :3546-3587): solve the normal
    equations A a = e_0 with A[i][j] = sum_o o^(i+j) and c[o] = sum_i a_i o^i.  FP64."""
    offs = np.arange(-nl, nr + 1, dtype=np.float64)
    A = np.array([[np.sum(offs ** (i + j)) for j in range(m + 1)] for i in range(m + 1)])
    e = np.zeros(m + 1)
    e[0] = 1.0
    a = np.linalg.solve(A, e)
    c = np.array([np.sum(a * o ** np.arange(m + 1)) for o in offs])
    return {int(o): float(w) for o, w in zip(offs, c)}


def savitzky_golay_filter(y: np.ndarray) -> np.ndarray:
    """SavitzkyGolay_filter (:3511-3542): y_i <- sum_o c[o] y_{i+o} for i = 1..n-nr, skipping the
    window points left of the first sample (no renormalisation); the last nr samples keep their
    values."""
    c = savgol_coefficients()
    n = len(y)
    out = np.array(y, dtype=np.float64)
    for i in range(n - SG_NR):
        acc = 0.0
        for o, w in c.items():
            if i + o >= 0:
                acc += w * float(y[i + o])
        out[i] = acc
    return out


# ---------------------------------------------------------------------------------------------
# Object state (the parts of type polarft_corrcalc, :68-198, that the math needs)
# ---------------------------------------------------------------------------------------------
@dataclass
class Pftcc:
    """Host FP64 model of a polarft_corrcalc instance (:68-198) with params_glob fields."""
    grid: Grid
    kstop: int
    refs_even: np.ndarray            # complex64 [nrefs, nk, n]
    refs_odd: np.ndarray             # complex64 [nrefs, nk, n]
    ptcls: np.ndarray                # complex64 [nptcls, nk, n]
    ctfmats: np.ndarray | None       # float32 [nptcls, nk, n] or None (ctf=no)
    iseven: np.ndarray               # bool [nptcls]
    match_filt: bool = False
    pssnr: bool = False
    ref_optlp: np.ndarray | None = None   # float32 [nrefs, kstop-kfrom+1]
    clsfrcs: bool = False
    sigma2: np.ndarray | None = None      # float32 [nptcls, nk]  sigma2_noise(k, iptcl)
    euclid: bool = False

    @property
    def pftsz(self) -> int:
        return self.grid.pftsz

    @property
    def nrots(self) -> int:
        return self.grid.nrots

    @property
    def kfrom(self) -> int:
        return self.grid.kfrom

    @property
    def ks(self) -> int:
        """Number of shells kfrom..kstop."""
        return self.kstop - self.grid.kfrom + 1

    def sqsum_ptcl(self, i: int) -> float:
        """memoize_sqsum_ptcl (:841-845): sum |p|^2 over kfrom..kstop."""
        p = self.ptcls[i, :self.ks].astype(np.complex128)
        return float(np.sum(np.abs(p) ** 2))

    # -- reference preparation ----------------------------------------------------------------
    def ref(self, iref: int, i: int) -> np.ndarray:
        """Even/odd reference selected by the particle's eo flag (:968-972)."""
        src = self.refs_even if self.iseven[i] else self.refs_odd
        return src[iref].astype(np.complex128)

    def shellnorm(self, q: np.ndarray, iref: int, i: int) -> np.ndarray:
        """shellnorm_and_filter_ref(_8) (:756-805); l_clsfrcs indexes optlp by the particle."""
        if not (self.match_filt or self.pssnr):
            return q
        idx = i if self.clsfrcs else iref
        q = q.copy()
        n = self.pftsz
        for kk in range(self.ks):
            k = self.kfrom + kk
            pw = float(np.sum(np.abs(q[kk]) ** 2)) / n
            if self.pssnr:
                w = math.sqrt(float(self.ref_optlp[idx, kk]) * k)
                if pw > 1e-12:
                    w /= math.sqrt(pw)
            else:
                w = float(self.ref_optlp[idx, kk])
                if pw > 1e-12:
                    w /= math.sqrt(pw)
            q[kk] *= w
        return q

    def prep_matchfilt(self, i: int, iref: int, irot: int) -> np.ndarray:
        """prep_matchfilt (:987-1095): the particle rings after the match filter, [nk, n] complex128.

        iref < 0 (the reference's iref = 0) or match_filt off: unchanged.  pssnr: per shell
        p <- p sqrt(k)/sqrt(pw_p) sqrt(1 + ssnr |CTF|^2) (ssnr = ref_optlp).  Otherwise: particle
        and reference (x |CTF|) are each scaled to power optlp^2 per shell, the power of their
        difference at rotation irot is smoothed over shells (Savitzky-Golay, the first 3 shells and
        non-positive values kept raw), and the particle is scaled to power ssnr/(ssnr+1) with
        ssnr = pw_ref / pw_diff.  Shells beyond kstop are untouched.  FP64.
        """
        p = self.ptcls[i].astype(np.complex128).copy()
        if iref < 0 or not self.match_filt:
            return p
        n = self.pftsz
        ks = self.ks
        opt = self.ref_optlp[iref].astype(np.float64)
        c = self.ctf(i)
        if self.pssnr:
            for kk in range(ks):
                k = self.kfrom + kk
                pw = float(np.sum(np.abs(p[kk]) ** 2)) / n
                w = math.sqrt(k)
                if pw > 1e-12:
                    w /= math.sqrt(pw)
                p[kk] = p[kk] * w * np.sqrt(1.0 + opt[kk] * c[kk] ** 2)
            return p
        r = self.ref(iref, i) * c
        for kk in range(ks):
            pw = float(np.sum(np.abs(p[kk]) ** 2)) / n
            p[kk] = p[kk] * opt[kk] / math.sqrt(pw) if pw > 1e-12 else p[kk] * opt[kk]
            pr = float(np.sum(np.abs(r[kk]) ** 2)) / n
            r[kk] = r[kk] * opt[kk] / math.sqrt(pr) if pr > 1e-12 else r[kk] * opt[kk]
        P = full_ring(p[:ks])
        idx = (np.arange(n) + (irot - 1)) % (2 * n)
        diff = np.sum(np.abs(r[:ks] - P[:, idx]) ** 2, axis=-1) / n
        fit = savitzky_golay_filter(diff)
        fit = np.where(fit <= 0.0, diff, fit)
        fit[:3] = diff[:3]
        for kk in range(ks):
            pw = float(np.sum(np.abs(p[kk]) ** 2)) / n
            w = 1.0 / math.sqrt(pw) if pw > 1e-12 else 1.0
            pr = float(np.sum(np.abs(r[kk]) ** 2)) / n
            if fit[kk] > 1e-12:
                ssnr = pr / fit[kk]
                w *= math.sqrt(ssnr / (ssnr + 1.0))
            p[kk] = w * p[kk]
        return p

    def ctf(self, i: int) -> np.ndarray:
        if self.ctfmats is None:
            return np.ones((self.grid.nk, self.pftsz))
        return self.ctfmats[i].astype(np.float64)

    def prep_ref(self, iref: int, i: int, shvec=None, shellnorm=True) -> np.ndarray:
        """q = shellnorm(ref) * |CTF_i| [* shmat(shvec)] (prep_ref4corr :961-985, gencorrs_cc_2
        :1515-1530, gencorr_cc_for_rot_8 :1700-1712)."""
        q = self.ref(iref, i)
        if shellnorm:
            q = self.shellnorm(q, iref, i)
        q = q * self.ctf(i)
        if shvec is not None:
            q = q * shmat_8(self.grid, shvec)
        return q

    # -- a10: gencorrs -----------------------------------------------------------------------
    def gencorrs(self, iref: int, i: int, shvec=None, method: str = "direct") -> np.ndarray:
        """gencorrs_cc_1 (:1486-1500) / gencorrs_cc_2 (:1502-1534): cc[nrots] (FP64).

        Unshifted: S_q over kfrom..kstop (prep_ref4corr :984).  Shifted: S_q over ALL shells
        kfrom..kto (:1531).  Correlations over kfrom..kstop.
        """
        if self.euclid:
            return self.gencorrs_euclid(iref, i, shvec, method)
        q = self.prep_ref(iref, i, shvec)
        ks = self.ks
        if shvec is None:
            sq = float(np.sum(np.abs(q[:ks]) ** 2))
        else:
            sq = float(np.sum(np.abs(q) ** 2))
        p = self.ptcls[i]
        if method == "direct":
            c = corrk_all_rots_direct(q[:ks], p[:ks]).sum(axis=0)
        else:
            pa, pb = memoize_fft(p[:ks].astype(np.complex128))
            c = corrs_over_k_fft(q[:ks], pa, pb).sum(axis=0)
        return c / math.sqrt(sq * self.sqsum_ptcl(i))

    def gencorrs_shifts(self, iref: int, i: int, shifts) -> np.ndarray:
        """gencorrs_cc_2 (:1502-1534) for a list of shifts at once through the FFT route
        (memoize_fft + calc_corrs_over_k): cc[ns, nrots] (FP64).  The same arithmetic as
        gencorrs(iref, i, shvec, method="fft") for each shift, vectorised over the shifts (the
        particle spectra are taken once); the reference norm runs over all shells (:1531)."""
        if self.euclid or self.match_filt or self.pssnr:
            raise NotImplementedError("gencorrs_shifts restates the CC objective without match filter")
        ks = self.ks
        n = self.pftsz
        sh = shmat_8_many(self.grid, shifts)                       # [ns, nk, n]
        q = (self.ref(iref, i) * self.ctf(i))[None] * sh           # [ns, nk, n]
        sq = np.sum(np.abs(q) ** 2, axis=(1, 2))                   # all shells
        pa, pb = memoize_fft(self.ptcls[i, :ks].astype(np.complex128))
        qk = q[:, :ks]
        a = np.fft.rfft(np.real(qk), axis=-1)
        b = np.fft.fft(np.imag(qk) * fft_factors(n), axis=-1)
        h = n // 2
        prod = np.zeros(qk.shape[:2] + (2 * n + 2,), dtype=np.complex128)
        prod[..., 0:2 * h + 1:2] = 4.0 * np.conj(a[..., :h + 1]) * pa[None, :, :h + 1]
        prod[..., 1:2 * h + 2:2] = 4.0 * np.conj(b[..., :h + 1]) * pb[None, :, :h + 1]
        # c2r_2n of bins 0..n per shell, summed over shells (taken as one c2r of the shell sum:
        # the transform is linear)
        back = np.fft.irfft(prod[..., :n + 1].sum(axis=1), n=2 * n, axis=-1) * (2 * n) / (2.0 * 2 * n)
        return back / np.sqrt(sq * self.sqsum_ptcl(i))[:, None]

    def gencorrs_per_shell(self, iref: int, i: int, shvec=None) -> np.ndarray:
        """Per-shell C_k(rho) of the prepared reference: [ks, nrots] (calc_k_corrs :1164)."""
        q = self.prep_ref(iref, i, shvec)
        return corrk_all_rots_direct(q[:self.ks], self.ptcls[i, :self.ks])

    # -- a16: Euclid ---------------------------------------------------------------------------
    def gencorrs_euclid(self, iref: int, i: int, shvec=None, method: str = "direct"):
        """gencorrs_euclid_1 (:1536-1556) / _2 (:1558-1592).

        euclid(rho) = sum_k pxls(k)/n * (2 C_k(rho) - |p_k|^2 - |q_k|^2) / (2 sigma2(k,iptcl)).
        The unshifted form shell-normalises the reference (via prep_ref4corr); the shifted
        form does not (:1573-1577).  No factor k here (unlike calc_euclid_for_rot :1304).
        """
        q = self.prep_ref(iref, i, shvec, shellnorm=shvec is None)
        ks = self.ks
        p = self.ptcls[i, :ks].astype(np.complex128)
        if method == "direct":
            ck = corrk_all_rots_direct(q[:ks], p)
        else:
            pa, pb = memoize_fft(p)
            ck = corrs_over_k_fft(q[:ks], pa, pb)
        pxl = pxls_p_shell(self.kfrom, self.grid.kto)[:ks]
        s2 = self.sigma2[i, :ks].astype(np.float64)
        sp = np.sum(np.abs(p) ** 2, axis=-1)
        sqr = np.sum(np.abs(q[:ks]) ** 2, axis=-1)
        w = pxl / self.pftsz / (2.0 * s2)
        return np.sum(w[:, None] * (2.0 * ck - sp[:, None] - sqr[:, None]), axis=0)

    def euclid_for_rot_8(self, iref: int, i: int, shvec, irot: int) -> float:
        """gencorr_euclid_for_rot_8 (:2002-2019) -> calc_euclid_for_rot_8 (:1307-1334):
        -sum_k k * |q_k - P_k(rot)|^2 / (2 sigma2) * pxls(k)/n (no shell normalisation)."""
        q = self.prep_ref(iref, i, shvec, shellnorm=False)
        ks = self.ks
        n = self.pftsz
        P = full_ring(self.ptcls[i, :ks].astype(np.complex128))
        idx = (np.arange(n) + irot - 1) % (2 * n)
        d = np.sum(np.abs(q[:ks] - P[:, idx]) ** 2, axis=-1)
        k = np.arange(self.kfrom, self.kstop + 1, dtype=np.float64)
        pxl = pxls_p_shell(self.kfrom, self.grid.kto)[:ks]
        s2 = self.sigma2[i, :ks].astype(np.float64)
        return float(-np.sum(k * d / (2.0 * s2) * pxl / n))

    def euclid_grad_for_rot_8(self, iref: int, i: int, shvec, irot: int):
        """gencorr_euclid_grad_for_rot_8 (:2071-2109): f and d/dshift."""
        f = self.euclid_for_rot_8(iref, i, shvec, irot)
        q = self.prep_ref(iref, i, shvec, shellnorm=False)
        ks = self.ks
        k = np.arange(self.kfrom, self.kstop + 1, dtype=np.float64)
        pxl = pxls_p_shell(self.kfrom, self.grid.kto)[:ks]
        s2 = self.sigma2[i, :ks].astype(np.float64)
        grad = np.zeros(2)
        for d, ax in enumerate((self.grid.argtransf_x, self.grid.argtransf_y)):
            qt = q * 1j * ax
            diff = corrk_for_rot(qt[:ks], self.ptcls[i, :ks], irot) - \
                np.real(np.sum(qt[:ks] * np.conj(q[:ks]), axis=-1))
            grad[d] = np.sum(k * diff / s2 * pxl / self.pftsz)
        return f, grad

    def sigma_contrib(self, iref: int, i: int, shvec, irot: int) -> np.ndarray:
        """gencorr_sigma_contrib (:2150-2188): 0.5*|q_k - P_k(rot)|^2/n over kfrom..kto."""
        q = self.prep_ref(iref, i, shvec, shellnorm=False)
        n = self.pftsz
        P = full_ring(self.ptcls[i].astype(np.complex128))
        idx = (np.arange(n) + irot - 1) % (2 * n)
        return 0.5 * np.sum(np.abs(q - P[:, idx]) ** 2, axis=-1) / n

    # -- a13 / a14: single-rotation FP64 objective and shift gradient ----------------------------
    def _weights(self):
        ks = self.ks
        if self.match_filt:
            return np.ones(ks)
        return np.arange(self.kfrom, self.kstop + 1, dtype=np.float64)

    def _denom(self, q: np.ndarray, i: int) -> float:
        ks = self.ks
        if self.match_filt:
            return math.sqrt(float(np.sum(np.abs(q[:ks]) ** 2)) * self.sqsum_ptcl(i))
        w = self._weights()
        sq = float(np.sum(w * np.sum(np.abs(q[:ks]) ** 2, axis=-1)))
        p = self.ptcls[i, :ks].astype(np.complex128)
        sp = float(np.sum(w * np.sum(np.abs(p) ** 2, axis=-1)))
        return math.sqrt(sq * sp)

    def gencorr_for_rot_8(self, iref: int, i: int, shvec, irot: int) -> float:
        """gencorr_cc_for_rot_8 (:1690-1733).  match_filt: unweighted with memoised S_p;
        otherwise k-weighted numerator and denominators (:1717-1727)."""
        if self.euclid:
            return self.euclid_for_rot_8(iref, i, shvec, irot)
        q = self.prep_ref(iref, i, shvec)
        ks = self.ks
        c = corrk_for_rot(q[:ks], self.ptcls[i, :ks], irot)
        return float(np.sum(self._weights() * c)) / self._denom(q, i)

    def gencorr_grad_for_rot_8(self, iref: int, i: int, shvec, irot: int):
        """gencorr_cc_grad_for_rot_8 (:1863-1925): f and (d f/d sx, d f/d sy).

        grad_d = sum_k w_k Re sum_j (i*argtransf_d*q)[j] conj(P[j+rho]) / denom, the denominator
        being shift invariant.
        """
        if self.euclid:
            return self.euclid_grad_for_rot_8(iref, i, shvec, irot)
        q = self.prep_ref(iref, i, shvec)
        ks = self.ks
        w = self._weights()
        den = self._denom(q, i)
        p = self.ptcls[i, :ks]
        f = float(np.sum(w * corrk_for_rot(q[:ks], p, irot))) / den
        g = np.empty(2)
        for d, ax in enumerate((self.grid.argtransf_x, self.grid.argtransf_y)):
            qt = q * 1j * ax
            g[d] = float(np.sum(w * corrk_for_rot(qt[:ks], p, irot))) / den
        return f, g

    def gencorr_obj_for_rot_8(self, iref: int, i: int, shvec, irot: int, euclid: bool) -> float:
        """gencorr_cc_for_rot_8 (:1690-1733) or gencorr_euclid_for_rot_8 (:2026-2048), the
        objective fixed instead of taken from cc_objfun (gencorr_for_rot_8's dispatch, :1632-1643)."""
        saved = self.euclid
        self.euclid = euclid
        try:
            return self.gencorr_for_rot_8(iref, i, shvec, irot)
        finally:
            self.euclid = saved

    def gencorr_obj_grad_for_rot_8(self, iref: int, i: int, shvec, irot: int, euclid: bool):
        """gencorr_cc_grad_for_rot_8 (:1863-1925) or gencorr_euclid_grad_for_rot_8 (:2071-2109)."""
        saved = self.euclid
        self.euclid = euclid
        try:
            return self.gencorr_grad_for_rot_8(iref, i, shvec, irot)
        finally:
            self.euclid = saved

    # -- orientation derivatives (continuous refinement) ------------------------------------------
    def cont_terms(self, iref: int, i: int, shvec, irot: int, drefs: np.ndarray, it1: int | None = None):
        """The sums of gencorr_cont_grad_cc_for_rot_8 (:1735-1786) and
        gencorr_cont_shift_grad_cc_for_rot_8 (:1788-1847), FP64.

        q = shellnorm(ref) |CTF| shmat_8 (:1760-1775; the shell filter is applied to q only),
        dq_d = dref_d |CTF| shmat_8 for the three orientation derivatives drefs [3, ks, n]
        (pfts_drefs of the particle's eo class), over shells kfrom..kstop:
        y = sum_k corrk(q, P_i), S_q = sum |q|^2, T1_d = sum_k corrk(dq_d, P_it1),
        T2_d = Re sum dq_d conj(q) (calc_T1_T2_for_rot_8, :1246-1278), and the shift terms
        g_x, g_y = sum_k corrk(i a_x q, P_i), likewise a_y (:1841-1846).  The reference passes iref as
        the particle index of the T1 terms (:1781, :1836), so it1 defaults to iref.
        """
        it1 = iref if it1 is None else it1
        ks = self.ks
        q = self.prep_ref(iref, i, shvec)[:ks]
        sh = shmat_8(self.grid, shvec)[:ks]
        c = self.ctf(i)[:ks]
        p = self.ptcls[i, :ks]
        pt = self.ptcls[it1, :ks]
        y = float(np.sum(corrk_for_rot(q, p, irot)))
        sq = float(np.sum(np.abs(q) ** 2))
        t1 = np.zeros(3)
        t2 = np.zeros(3)
        for d in range(3):
            dq = drefs[d].astype(np.complex128) * c * sh
            t1[d] = float(np.sum(corrk_for_rot(dq, pt, irot)))
            t2[d] = float(np.real(np.sum(dq * np.conj(q))))
        gs = np.zeros(2)
        for d, ax in enumerate((self.grid.argtransf_x, self.grid.argtransf_y)):
            gs[d] = float(np.sum(corrk_for_rot(q * 1j * ax[:ks], p, irot)))
        return y, sq, t1, t2, gs

    def cont_grad_cc_for_rot_8(self, iref: int, i: int, shvec, irot: int, drefs: np.ndarray, it1=None):
        """gencorr_cont_grad_cc_for_rot_8 (:1735-1786): cc = y / sqrt(S_q S_p) and
        dcc_d = (T1_d - y T2_d / S_q) / sqrt(S_q S_p), S_p the memoized particle norm."""
        y, sq, t1, t2, _ = self.cont_terms(iref, i, shvec, irot, drefs, it1)
        den = math.sqrt(sq * self.sqsum_ptcl(i))
        return y / den, (t1 - y * t2 / sq) / den

    def cont_shift_grad_cc_for_rot_8(self, iref: int, i: int, shvec, irot: int, drefs: np.ndarray, it1=None):
        """gencorr_cont_shift_grad_cc_for_rot_8 (:1788-1847): f and grad[5] (three orientation
        derivatives, then d/dshift_x, d/dshift_y)."""
        y, sq, t1, t2, gs = self.cont_terms(iref, i, shvec, irot, drefs, it1)
        den = math.sqrt(sq * self.sqsum_ptcl(i))
        return y / den, np.concatenate([(t1 - y * t2 / sq) / den, gs / den])

    def calc_ptcl_pspec(self, i: int) -> np.ndarray:
        """calc_ptcl_pspec (:897-904): sum_j |p_k[j]|^2 / pftsz over kfrom..kto."""
        return np.sum(np.abs(self.ptcls[i].astype(np.complex128)) ** 2, axis=-1) / self.pftsz

    # -- a15: in-plane angle derivative (continuous rotation psi, radians) ----------------------
    def corr_at_angle(self, iref: int, i: int, shvec, psi: float) -> float:
        """Band-limited rotational interpolant of the gencorr_for_rot_8 objective.

        cc(psi) = sum_k w_k Re sum_j q_k[j] conj(P_k(theta_j + psi)) / denom, where P_k(.) is
        the trigonometric interpolant of the full 2n-sample ring (Nyquist term as a cosine).
        At psi = rho*pi/n it equals gencorr_for_rot_8(irot = rho+1).
        """
        q = self.prep_ref(iref, i, shvec)
        ks = self.ks
        w = self._weights()
        Pr = rotate_ring_continuous(self.ptcls[i, :ks].astype(np.complex128), psi)
        c = np.real(np.sum(q[:ks] * np.conj(Pr), axis=-1))
        return float(np.sum(w * c)) / self._denom(q, i)

    def grad_angle(self, iref: int, i: int, shvec, psi: float) -> float:
        """d cc(psi)/d psi (radians) of corr_at_angle, via the spectral derivative."""
        q = self.prep_ref(iref, i, shvec)
        ks = self.ks
        w = self._weights()
        dP = rotate_ring_continuous(self.ptcls[i, :ks].astype(np.complex128), psi, deriv=True)
        c = np.real(np.sum(q[:ks] * np.conj(dP), axis=-1))
        return float(np.sum(w * c)) / self._denom(q, i)

    # -- a18: FRC based spectral score --------------------------------------------------------
    def genfrc(self, iref: int, i: int, irot: int, shvec=None) -> np.ndarray:
        """genfrc (:1417-1440) / calc_frc (:1442-1484): per-shell normalised correlation."""
        q = self.prep_ref(iref, i, shvec)
        ks = self.ks
        p = self.ptcls[i, :ks].astype(np.complex128)
        ck = corrk_for_rot(q[:ks], p, irot)
        den = np.sqrt(np.sum(np.abs(p) ** 2, -1) * np.sum(np.abs(q[:ks]) ** 2, -1))
        return np.where(den < 1e-12, 0.0, ck / np.where(den < 1e-12, 1.0, den))

    def specscore(self, iref: int, i: int, irot: int, shvec=None) -> float:
        """specscore_1/2 (:2190-2205): max(0, median_nocopy(frc)) on the FP32 FRC values."""
        frc = self.genfrc(iref, i, irot, shvec).astype(np.float32)
        return max(0.0, float(median_nocopy(frc)))


def median_nocopy(arr: np.ndarray) -> np.float32:
    """median_nocopy (simple_math.f90:2111-2130): pos1 = n/2, pos2 = pos1 + 1 for even n (both
    selected by two successive selec calls on the same, already permuted array, then averaged in
    FP32), pos1 = pos2 = nint(n/2.) for odd n.  selec (:2204-2261) is the reference's quickselect
    with its final two-element test, restated in prep_oracle.  arr is permuted in place."""
    from prep_oracle import selec
    a = arr.astype(np.float32) if arr.dtype != np.float32 else arr
    n = a.size
    if n % 2 == 0:
        p1, p2 = n // 2, n // 2 + 1
    else:
        p1 = p2 = fortran_nint(float(f32(n) / f32(2.0)))
    if p1 == p2:
        return selec(p1, a)
    v1 = selec(p1, a)
    v2 = selec(p2, a)
    return f32(f32(v1 + v2) / f32(2.0))


def rotate_ring_continuous(p: np.ndarray, psi: float, deriv: bool = False) -> np.ndarray:
    """Samples P(theta_j + psi), j = 0..n-1, of the trigonometric interpolant of the full ring.

    p: [nk, n] half ring.  With N = 2n and F = DFT_N(full ring), the interpolant is
    P(theta) = (1/N) sum_{|m|<n} F[m] e^{i m theta} + (1/N) F[n] cos(n theta), theta = 2 pi t/N.
    deriv=True returns dP/dtheta instead.
    """
    n = p.shape[-1]
    N = 2 * n
    P = full_ring(p)
    F = np.fft.fft(P, axis=-1)
    m = np.fft.fftfreq(N, d=1.0 / N)      # 0..n-1, -n..-1 ; index n holds m = -n (Nyquist)
    theta = np.arange(n) * (2 * DPI / N) + psi
    E = np.exp(1j * np.outer(m, theta))   # [N, n]
    nyq = n
    coef = F.copy()
    if deriv:
        coef = coef * (1j * m)[None, :]
        coef[:, nyq] = 0.0
        out = coef @ E / N
        out += (F[:, nyq:nyq + 1] * (-n * np.sin(n * theta))[None, :]) / N
        return out
    coef[:, nyq] = 0.0
    out = coef @ E / N
    out += (F[:, nyq:nyq + 1] * np.cos(n * theta)[None, :]) / N
    return out
