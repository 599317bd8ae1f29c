// polarizer.hpp - Kaiser-Bessel gridding of Cartesian 2-D Fourier transforms onto the polar grid of
// the correlation search: init_polarizer and polarize (simple_polarizer.f90:41-151).
//
// Each polar point (i, k), i < pftsz, takes the 3 x 3 Cartesian components around
// (nint(x), nint(y)) of its coordinates (x, y) = get_coord(i, k), weighted by the separable
// Kaiser-Bessel window w(l, m) = apod(h_l - x) apod(k_m - y) (half-width 1.5, beta 7.4), the
// weights divided by their sum unless griddev.  Components come from the image's half-plane
// transform with the Friedel rule of get_fcomp2D (simple_image.f90:1894-1907).
//
// Layout: the tables are SoA over the n * nk polar points, point index p = kk * n + i:
//   wts[t][p], t = l + 3 m (the reference's reshape order, l fastest);  hk[c][p], c < 3 the
//   logical h of window column l = c, c >= 3 the logical k of window row m = c - 3.
// An image's transform is its cmat in Fortran order, i.e. the C array [ldim][ldim/2 + 1] of
// complex (k_phys major), which is what the reference's image object holds after fwd_ft.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pftcc_kernels.hpp"

namespace pfk {

constexpr int PZ_WDIM = 3;                   // sqwin of half-width 1.5: nint(x) - 1 .. nint(x) + 1
constexpr int PZ_WLEN = PZ_WDIM * PZ_WDIM;

// kbinterpol%new (simple_kbinterpol.f90:92-109) for the 2-D polarizer window, FP32 fields
struct KbWin {
  float whalf, W, beta, twooW, oneoW;
};

// I0 for x >= 0 from the Abramowitz & Stegun polynomials 9.8.1 (x < 3.75) and 9.8.2, FP64 accumulation
// of an FP32 argument (bessi0, simple_kbinterpol.f90:229-251).
__device__ inline double kb_bessi0(float xf) {
  const double x = (double)xf;
  if (x < 3.75) {
    double t = x / 3.75;
    t *= t;
    return 1.0 + t * (3.5156229 + t * (3.0899424 + t * (1.2067492 + t * (0.2659732 + t * (0.0360768 + t * 0.0045813)))));
  }
  const double t = 3.75 / x;
  const double p = 0.39894228 +
                   t * (0.01328592 +
                        t * (0.00225319 +
                             t * (-0.00157565 +
                                  t * (0.00916281 + t * (-0.02057706 + t * (0.02635537 + t * (-0.01647633 + t * 0.00392377)))))));
  return p * exp(x) / sqrt(x);
}

// kbinterpol%apod (simple_kbinterpol.f90:130-141): (1/W) I0(beta sqrt(1 - (2x/W)^2)) for |x| <= Whalf
__device__ inline float kb_apod(const KbWin& kb, float x) {
  if (fabsf(x) > kb.whalf) return 0.f;
  float arg = kb.twooW * x;
  arg = 1.f - arg * arg;
  return (float)((double)kb.oneoW * kb_bessi0(kb.beta * sqrtf(arg)));
}

// Fortran nint (half away from zero) of an FP32 value
__device__ inline int f_nint(float x) { return (int)(x >= 0.f ? floorf(x + 0.5f) : -floorf(-x + 0.5f)); }

// cyci_1d (simple_math.f90:746-757)
__device__ inline int cyci(int lo, int hi, int i) {
  if (i > hi) return lo + (i - hi) - 1;
  if (i < lo) return hi - (lo - i) + 1;
  return i;
}

// init_polarizer (simple_polarizer.f90:61-89): one thread per polar point p = kk * n + i.
// polar_x/polar_y: the handle's FP32 grid [nk][nrots]; hlo..hhi, klo..khi the logical Fourier
// limits of the image (loop_lims(3)).
__global__ void k_polar_tables(const float* __restrict__ polar_x, const float* __restrict__ polar_y, int nk, int n,
                               int nrots, KbWin kb, int hlo, int hhi, int klo, int khi, int griddev,
                               float* __restrict__ wts, int* __restrict__ hk) {
  const int np = nk * n;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int kk = p / n, i = p - kk * n;
  const float lx = polar_x[(size_t)kk * nrots + i];
  const float ly = polar_y[(size_t)kk * nrots + i];
  const int w1 = f_nint(lx) - 1, w2 = f_nint(ly) - 1;
  float w[PZ_WDIM][PZ_WDIM];                   // w[l][m]
#pragma unroll
  for (int l = 0; l < PZ_WDIM; ++l)
#pragma unroll
    for (int m = 0; m < PZ_WDIM; ++m) w[l][m] = 1.f;
  // as the reference builds it: row l scaled by the h factor, then column l by the k factor
#pragma unroll
  for (int l = 0; l < PZ_WDIM; ++l) {
    const float ax = kb_apod(kb, (float)(w1 + l) - lx);
#pragma unroll
    for (int m = 0; m < PZ_WDIM; ++m) w[l][m] *= ax;
    const float ay = kb_apod(kb, (float)(w2 + l) - ly);
#pragma unroll
    for (int m = 0; m < PZ_WDIM; ++m) w[m][l] *= ay;
    hk[(size_t)l * np + p] = cyci(hlo, hhi, w1 + l);
    hk[(size_t)(PZ_WDIM + l) * np + p] = cyci(klo, khi, w2 + l);
  }
  float s = 1.f;
  if (!griddev) {
    s = 0.f;
#pragma unroll
    for (int m = 0; m < PZ_WDIM; ++m)
#pragma unroll
      for (int l = 0; l < PZ_WDIM; ++l) s += w[l][m];
  }
#pragma unroll
  for (int m = 0; m < PZ_WDIM; ++m)
#pragma unroll
    for (int l = 0; l < PZ_WDIM; ++l) wts[(size_t)(l + PZ_WDIM * m) * np + p] = griddev ? w[l][m] : w[l][m] / s;
}

// polarize (simple_polarizer.f90:120-151), batched: thread (point p, image group y) reads the
// point's table once and grids PZ_IMGS consecutive images with it.  cmats: [nimg][ldim][ldim/2 + 1]
// complex; out: the handle's ring block, image b written at slot[b] * nk * n; mask: [nk] shell
// flags or null.  FP32 accumulation in the reference's order (l fastest, then m).
// 1-D grid of (point blocks) x (image groups) in XCD-aware order: the point blocks of one image
// group run on one XCD and read its images from that XCD's L2.
constexpr int PZ_IMGS = 8;

__global__ void __launch_bounds__(256) k_polarize(const float2* __restrict__ cmats, int ldim, int nimg,
                                                  const float* __restrict__ wts, const int* __restrict__ hk,
                                                  int nk, int n, const int* __restrict__ slot,
                                                  const int8_t* __restrict__ mask, float2* __restrict__ out) {
  const int np = nk * n;
  const int gx = (np + (int)blockDim.x - 1) / (int)blockDim.x;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int by = t / gx, bx = t - by * gx;
  const int p = bx * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int b0 = by * PZ_IMGS;
  const int kk = p / n;
  const bool zero = mask && !mask[kk];
  const int hc = ldim / 2 + 1;
  // the 9 addresses of the window (Friedel-folded, image-relative) and their weights
  int addr[PZ_WLEN];
  bool conj[PZ_WLEN];
  float w[PZ_WLEN];
#pragma unroll
  for (int m = 0; m < PZ_WDIM; ++m) {
#pragma unroll
    for (int l = 0; l < PZ_WDIM; ++l) {
      const int h = hk[(size_t)l * np + p], k = hk[(size_t)(PZ_WDIM + m) * np + p];
      // get_fcomp2D: h >= 0 reads (h, k); h < 0 reads the conjugate of (-h, -k)
      const bool neg = h < 0;
      const int ph = neg ? -h : h;
      int pk = neg ? -k : k;
      pk += pk < 0 ? ldim : 0;
      const int t = l + PZ_WDIM * m;
      addr[t] = pk * hc + ph;
      conj[t] = neg;
      w[t] = wts[(size_t)t * np + p];
    }
  }
  const size_t isz = (size_t)ldim * hc;
#pragma unroll
  for (int i = 0; i < PZ_IMGS; ++i) {
    const int b = b0 + i;
    if (b >= nimg) break;
    float2* dst = out + (size_t)slot[b] * np + p;
    if (zero) {
      *dst = make_float2(0.f, 0.f);
      continue;
    }
    const float2* F = cmats + (size_t)b * isz;
    float re = 0.f, im = 0.f;
#pragma unroll
    for (int t = 0; t < PZ_WLEN; ++t) {
      const float2 v = F[addr[t]];
      re = fmaf(w[t], v.x, re);
      im = fmaf(w[t], conj[t] ? -v.y : v.y, im);
    }
    *dst = make_float2(re, im);
  }
}

}  // namespace pfk
