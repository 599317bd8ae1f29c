// prep.hpp - particle preparation for alignment (prepimg4align, simple_strategy2D3D_common.f90:
// 316-366) for batches of 2-D images, ahead of the polarizer.
//
// Per image (box B in, boxmatch M out; images are C arrays [y][x] = Fortran rmat(x, y), transforms
// [k_phys][h] = Fortran cmat(h, k_phys)):
//   noise_norm over the background of the disc of radius msk (simple_image.f90:6623-6640)
//   -> fft (quadrant swap, r2c, 1/N; :5773-5782) -> shift2Dserial([-x, -y]) (:6149-6174)
//   -> CTF phase flip (simple_ctf.f90:345-355) -> ifft (c2r, quadrant swap; :5784-5791)
//   -> real-space clip to M (:6499-6511) -> soft mask with zero_background (:6208-6286, 5176-5203)
//   -> [griddev: divide by the instrument function] -> fft -> x M^2/B^2.
// The 2-D transforms run as batched 1-D Stockham FFTs in LDS (fft_lds.hpp): rows as pairs of real
// rows packed into one complex transform, columns as complex transforms of CP columns per
// work-group.  Every kernel streams the batch through HBM once.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "corr_fused.hpp"
#include "fft_lds.hpp"
#include "pftcc_kernels.hpp"

namespace pfk {

#ifndef PFTCC_PREP_RP
#define PFTCC_PREP_RP 4
#endif
#ifndef PFTCC_PREP_CP
#define PFTCC_PREP_CP 8
#endif
constexpr int PREP_RP = PFTCC_PREP_RP;     // complex row transforms (pairs of real rows) per work-group
constexpr int PREP_CP = PFTCC_PREP_CP;     // columns per work-group

// Forward transforms of F length-L rows in LDS (bufA, stride L).  LC > 0: the compile-time
// in-place Stockham passes of the correlation kernel (radix order 4, then the other factors in
// descending order); LC = 0: the runtime-radix passes, ping-ponging with bufB.  Returns the buffer
// that holds the result.
template <int LC, int F>
__device__ __forceinline__ float2* prep_fft(float2* bufA, float2* bufB, int L, const FftPlan& plan,
                                            const float2* tw2, int tid, int nt) {
  if constexpr (LC > 0) {
    lds_passes<LC, F, 256, 0, Fp<LC>::np, LC>(bufA, tw2, tid);
    return bufA;
  } else {
    return fft_forward_oop(bufA, bufB, F, L, L, plan, tw2, tid, nt);
  }
}
constexpr float PREP_TINY = 1e-10f;          // simple_defs.f90:28
constexpr float PREP_SHTHRESH = 0.001f;      // simple_strategy2D3D_common.f90:31

// noise_norm statistics of one image: (mean to subtract, standard deviation to divide by); the
// identities (0, 1) where the reference skips the step (|ave| <= TINY, var not > 0).
__global__ void __launch_bounds__(256) k_prep_stats(const float* __restrict__ imgs, int B, float msk,
                                                    float2* __restrict__ stats) {
  const int img = blockIdx.x, tid = threadIdx.x;
  const float* x = imgs + (size_t)img * B * B;
  const int c = B / 2 + 1;                   // cendist centre, 1-based (:2725)
  __shared__ double red[2][4];
  __shared__ float ave_s;
  // background: cendist > msk.  (i - c)^2 + (j - c)^2 is an exact integer, so the FP64 root
  // rounded to FP32 is the FP32 root.
  double s = 0.0, cnt = 0.0;
  for (int p = tid; p < B * B; p += blockDim.x) {
    const int ix = p % B, iy = p / B;
    const int dx = ix + 1 - c, dy = iy + 1 - c;
    const float dist = (float)sqrt((double)(dx * dx + dy * dy));
    if (!(dist <= msk)) {
      s += (double)x[p];
      cnt += 1.0;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o);
    cnt += __shfl_xor(cnt, o);
  }
  if ((tid & 63) == 0) {
    red[0][tid >> 6] = s;
    red[1][tid >> 6] = cnt;
  }
  __syncthreads();
  const int nw = blockDim.x >> 6;
  if (tid == 0) {
    double S = 0.0, N = 0.0;
    for (int w = 0; w < nw; ++w) {
      S += red[0][w];
      N += red[1][w];
    }
    const float ave = (float)S / (float)N;
    ave_s = fabsf(ave) > PREP_TINY ? ave : 0.0f;
    red[1][0] = N;
  }
  __syncthreads();
  const float ave = ave_s;
  const double npix_d = red[1][0];
  __syncthreads();
  double ep = 0.0, vr = 0.0;
  for (int p = tid; p < B * B; p += blockDim.x) {
    const int ix = p % B, iy = p / B;
    const int dx = ix + 1 - c, dy = iy + 1 - c;
    const float dist = (float)sqrt((double)(dx * dx + dy * dy));
    if (!(dist <= msk)) {
      const float v = x[p] - ave;
      ep += (double)v;
      vr += (double)v * (double)v;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    ep += __shfl_xor(ep, o);
    vr += __shfl_xor(vr, o);
  }
  if ((tid & 63) == 0) {
    red[0][tid >> 6] = ep;
    red[1][tid >> 6] = vr;
  }
  __syncthreads();
  if (tid == 0) {
    double E = 0.0, V = 0.0;
    for (int w = 0; w < nw; ++w) {
      E += red[0][w];
      V += red[1][w];
    }
    const float np = (float)npix_d;
    const float e32 = (float)E;
    const float var = ((float)V - e32 * e32 / np) / (np - 1.0f);   // corrected two-pass formula
    float sd = 1.0f;
    if (isfinite(var) && var > 0.0f) sd = (float)sqrt((double)var);
    stats[img] = make_float2(ave, sd);
  }
}

// Forward real transforms of the rows of the quadrant-swapped image: two real rows per complex
// transform, split into their half spectra.  src [count][L][L]; stats (noise_norm) or null;
// dst [count][L][L/2+1].  Grid (ceil(L/2 / RP), count).
// LC > 0: compile-time length (L = LC).
template <int LC>
__global__ void __launch_bounds__(256) k_prep_rows_fwd(const float* __restrict__ src, int Lrt,
                                                       const float2* __restrict__ stats,
                                                       const float2* __restrict__ tw2g, FftPlan plan,
                                                       float2* __restrict__ dst) {
  const int L = LC > 0 ? LC : Lrt;
  extern __shared__ float2 smem[];
  float2* bufA = smem;
  float2* bufB = bufA + PREP_RP * L;
  float2* tw2 = bufB + PREP_RP * L;
  const int img = blockIdx.y, tid = threadIdx.x, nt = blockDim.x;
  const int p0 = blockIdx.x * PREP_RP;
  const int hl = L / 2, nh = L / 2 + 1;
  const float* x = src + (size_t)img * L * L;
  float ave = 0.0f, sd = 1.0f;
  if (stats) {
    const float2 st = stats[img];
    ave = st.x;
    sd = st.y;
  }
  for (int i = tid; i < 2 * L; i += nt) tw2[i] = tw2g[i];
  for (int e = tid; e < PREP_RP * L; e += nt) {
    const int r = e / L, xx = e - r * L;
    const int p = p0 + r;
    float2 z = make_float2(0.f, 0.f);
    if (p < hl) {
      const int sx = (xx + hl) % L;
      const int y0 = (2 * p + hl) % L, y1 = (2 * p + 1 + hl) % L;
      float v0 = x[(size_t)y0 * L + sx], v1 = x[(size_t)y1 * L + sx];
      v0 = (v0 - ave) / sd;      // rmat - ave, then rmat / sdev (exact identities when skipped)
      v1 = (v1 - ave) / sd;
      z = make_float2(v0, v1);
    }
    bufA[e] = z;
  }
  __syncthreads();
  const float2* res = prep_fft<LC, PREP_RP>(bufA, bufB, L, plan, tw2, tid, nt);
  float2* out = dst + (size_t)img * L * nh;
  for (int e = tid; e < PREP_RP * nh; e += nt) {
    const int r = e / nh, m = e - r * nh;
    const int p = p0 + r;
    if (p >= hl) continue;
    const float2 Z = res[r * L + m];
    const float2 Zc = cconj(res[r * L + (L - m) % L]);
    const float2 a = make_float2(0.5f * (Z.x + Zc.x), 0.5f * (Z.y + Zc.y));
    const float2 d = csub(Z, Zc);
    const float2 b = make_float2(0.5f * d.y, -0.5f * d.x);      // (Z - conj Z[L-m]) / (2i)
    out[(size_t)(2 * p) * nh + m] = a;
    out[(size_t)(2 * p + 1) * nh + m] = b;
  }
}

// Complex transforms along y of CP columns of [count][L][L/2+1].  inverse: unnormalised
// backward transform.  Forward post-processing (post): 1 = the first transform of prepimg4align
// (1/N, shift2Dserial, CTF flip), 2 = the final one (1/N, x ratio), 0 = none.
struct PrepColArgs {
  float2* data;
  int L;
  int inverse;
  int post;
  const float2* tw2g;
  FftPlan plan;
  const float* shifts;        // [count][2] particle (x, y) or null
  float shconst;              // PI / (L / 2), FP32
  const KCtfConst* ctf;       // [count] or null: phase flip
  float ratio;                // M^2 / B^2, FP32
};

__device__ __forceinline__ float prep_ctf_value(const KCtfConst& c, int h, int k, float invl) {
  // apply_serial 'flip' loop body (simple_ctf.f90:345-355) with eval (:147-163, 185-189)
  const float hinv = (float)h * invl, kinv = (float)k * invl;
  const float s2 = hinv * hinv + kinv * kinv;
  const float ang = atan2f((float)k, (float)h);
  const float df = 0.5f * (c.dfx + c.dfy + cosf(2.0f * (ang - c.angast)) * (c.dfx - c.dfy));
  const float pi32 = 3.14159274101257324f;
  const float arg = pi32 * c.wl * s2 * (df - 0.5f * c.wl * c.wl * s2 * c.cs) + c.phshift;
  return sinf(arg + c.ampc);
}

template <int LC>
__global__ void __launch_bounds__(256) k_prep_cols(PrepColArgs a) {
  extern __shared__ float2 smem[];
  const int L = LC > 0 ? LC : a.L, nh = L / 2 + 1;
  float2* bufA = smem;
  float2* bufB = bufA + PREP_CP * L;
  float2* tw2 = bufB + PREP_CP * L;
  const int img = blockIdx.y, tid = threadIdx.x, nt = blockDim.x;
  const int h0 = blockIdx.x * PREP_CP;
  float2* d = a.data + (size_t)img * L * nh;
  for (int i = tid; i < 2 * L; i += nt) tw2[i] = a.tw2g[i];
  for (int e = tid; e < PREP_CP * L; e += nt) {
    const int y = e / PREP_CP, c = e - y * PREP_CP;
    float2 v = make_float2(0.f, 0.f);
    if (h0 + c < nh) v = d[(size_t)y * nh + h0 + c];
    bufA[c * L + y] = a.inverse ? cconj(v) : v;
  }
  __syncthreads();
  const float2* res = prep_fft<LC, PREP_CP>(bufA, bufB, L, a.plan, tw2, tid, nt);
  // per-image shift (shift2Dserial([-x, -y]) only above SHTHRESH, :338)
  bool doshift = false;
  double sh1 = 0.0, sh2 = 0.0;
  if (a.post == 1 && a.shifts) {
    const float sx = a.shifts[2 * img], sy = a.shifts[2 * img + 1];
    doshift = fabsf(sx) > PREP_SHTHRESH || fabsf(sy) > PREP_SHTHRESH;
    sh1 = (double)(-sx * a.shconst);   // real(shvec * shconst, dp)
    sh2 = (double)(-sy * a.shconst);
  }
  const float nn = (float)L * (float)L;
  const float invl = 1.0f / (float)L;
  for (int e = tid; e < PREP_CP * L; e += nt) {
    const int y = e / PREP_CP, c = e - y * PREP_CP;
    const int h = h0 + c;
    if (h >= nh) continue;
    float2 v = res[c * L + y];
    if (a.inverse) {
      v = cconj(v);
    } else if (a.post) {
      v = make_float2(v.x / nn, v.y / nn);      // cmat / real(product(ldim)) (:5779)
      if (a.post == 1) {
        const int k = y < L / 2 ? y : y - L;
        if (doshift) {
          double hs, hc, sk, ck;
          sincos((double)h * sh1, &hs, &hc);
          sincos((double)k * sh2, &sk, &ck);
          const float2 ph = make_float2((float)(ck * hc - sk * hs), (float)(ck * hs + sk * hc));
          v = cmul(v, ph);
        }
        if (a.ctf) {
          const float t = prep_ctf_value(a.ctf[img], h, k, invl);
          if (signbit(t)) v = make_float2(-v.x, -v.y);       // sign(1., tval)
        }
      } else {
        v = make_float2(v.x * a.ratio, v.y * a.ratio);      // mul(boxmatch^2 / box^2) (:366)
      }
    }
    d[(size_t)y * nh + h] = v;
  }
}

// Inverse real transforms of the rows (c2r, imaginary parts of h = 0 and h = L/2 ignored), the
// quadrant swap and the real-space clip to M: out[y'][x'] = row (y' + s + B/2) % B at
// (x' + s + B/2) % B, s = (B - M)/2.  Two output rows per complex transform.  Grid
// (ceil(M/2 / RP), count).
template <int LC>
__global__ void __launch_bounds__(256) k_prep_rows_inv(const float2* __restrict__ data, int Brt, int M,
                                                       const float2* __restrict__ tw2g, FftPlan plan,
                                                       float* __restrict__ out) {
  const int B = LC > 0 ? LC : Brt;
  extern __shared__ float2 smem[];
  float2* bufA = smem;
  float2* bufB = bufA + PREP_RP * B;
  float2* tw2 = bufB + PREP_RP * B;
  const int img = blockIdx.y, tid = threadIdx.x, nt = blockDim.x;
  const int p0 = blockIdx.x * PREP_RP;
  const int nh = B / 2 + 1, hb = B / 2, s = (B - M) / 2;
  const float2* X = data + (size_t)img * B * nh;
  for (int i = tid; i < 2 * B; i += nt) tw2[i] = tw2g[i];
  for (int e = tid; e < PREP_RP * B; e += nt) {
    const int r = e / B, m = e - r * B;
    const int p = p0 + r;
    float2 z = make_float2(0.f, 0.f);
    if (p < M / 2) {
      const int yy0 = (2 * p + s + hb) % B, yy1 = (2 * p + 1 + s + hb) % B;
      float2 a, b;
      if (m <= hb) {
        a = X[(size_t)yy0 * nh + m];
        b = X[(size_t)yy1 * nh + m];
        if (m == 0 || m == hb) {
          a.y = 0.f;
          b.y = 0.f;
        }
      } else {
        a = cconj(X[(size_t)yy0 * nh + (B - m)]);
        b = cconj(X[(size_t)yy1 * nh + (B - m)]);
      }
      z = make_float2(a.x - b.y, a.y + b.x);     // A + i B
      z = cconj(z);                              // inverse via conj(FFT(conj))
    }
    bufA[e] = z;
  }
  __syncthreads();
  const float2* res = prep_fft<LC, PREP_RP>(bufA, bufB, B, plan, tw2, tid, nt);
  float* o = out + (size_t)img * M * M;
  for (int e = tid; e < PREP_RP * 2 * M; e += nt) {
    const int r = e / (2 * M), rem = e - r * (2 * M);
    const int t = rem / M, xo = rem - t * M;
    const int p = p0 + r;
    if (p >= M / 2) continue;
    const float2 v = res[r * B + (xo + s + hb) % B];
    o[(size_t)(2 * p + t) * M + xo] = t == 0 ? v.x : -v.y;    // conj: real row A, then row B
  }
}

// selec (simple_math.f90:2204-2261), restated step for step on an LDS array (1-based indices as
// the reference's; arr permuted in place).  Its final test compares ir - 1 with 1.
__host__ __device__ float prep_selec(int k, int n, float* arr) {
  int l = 1, ir = n;
  auto A = [&](int i) -> float& { return arr[i - 1]; };
  for (;;) {
    if (ir - l <= 1) {
      if (ir - 1 == 1 && A(ir) < A(l)) {
        const float t = A(l);
        A(l) = A(ir);
        A(ir) = t;
      }
      return A(k);
    }
    const int mid = (l + ir) / 2;
    float t = A(mid);
    A(mid) = A(l + 1);
    A(l + 1) = t;
    if (A(l) > A(ir)) { t = A(l); A(l) = A(ir); A(ir) = t; }
    if (A(l + 1) > A(ir)) { t = A(l + 1); A(l + 1) = A(ir); A(ir) = t; }
    if (A(l) > A(l + 1)) { t = A(l); A(l) = A(l + 1); A(l + 1) = t; }
    int i = l + 1, j = ir;
    const float a = A(l + 1);
    for (;;) {
      do { ++i; } while (A(i) < a);
      do { --j; } while (A(j) > a);
      if (j < i) break;
      t = A(i);
      A(i) = A(j);
      A(j) = t;
    }
    A(l + 1) = A(j);
    A(j) = a;
    if (j >= k) ir = j - 1;
    if (j <= k) l = i;
  }
}

// cosedge_1 (simple_math.f90:1089-1103) and cosedge_inner_1 (:1128-1140), FP32.
__device__ __forceinline__ float prep_cosedge(float x, float y, int box, float mskrad) {
  const float pi32 = 3.14159274101257324f;
  const float maxrad = (float)(box / 2);
  const float rad = (float)sqrt((double)(x * x + y * y));
  const float width = 2.0f * (maxrad - mskrad);
  if (rad >= maxrad) return 0.0f;
  if (rad >= maxrad - width) return (cosf(((rad - (maxrad - width)) / width) * pi32) + 1.0f) / 2.0f;
  return 1.0f;
}
__device__ __forceinline__ float prep_cosedge_inner(float x, float y, float width, float mskrad) {
  const float pi32 = 3.14159274101257324f;
  const float rad = (float)sqrt((double)(x * x + y * y));
  if (rad < mskrad - width) return 0.0f;
  if (rad > mskrad) return 1.0f;
  return (1.0f + cosf(pi32 * (mskrad - rad) / width)) / 2.0f;
}

// zero_background + mask(mskrad, 'soft'[, inner, width]) (+ div_by_instrfun), in place on one
// M x M image per work-group.  LDS: the four edges, in the reference's order of selec calls.
__global__ void __launch_bounds__(256) k_prep_mask(float* __restrict__ imgs, int M, float mskrad, float inner,
                                                   float width, int minlen, const float* __restrict__ instr_w) {
  extern __shared__ float edge[];            // [4][M]
  __shared__ float med_s;
  const int img = blockIdx.x, tid = threadIdx.x;
  float* x = imgs + (size_t)img * M * M;
  float* e0 = edge;            // rmat(1, :)   : x = 0, all y
  float* e1 = edge + M;        // rmat(M, :)   : x = M-1
  float* e2 = edge + 2 * M;    // rmat(:, 1)   : y = 0, all x
  float* e3 = edge + 3 * M;    // rmat(:, M)   : y = M-1
  for (int i = tid; i < M; i += blockDim.x) {
    e0[i] = x[(size_t)i * M];
    e1[i] = x[(size_t)i * M + M - 1];
    e2[i] = x[i];
    e3[i] = x[(size_t)(M - 1) * M + i];
  }
  __syncthreads();
  if (tid == 0) {
    const int k = M / 2;
    const float v1 = prep_selec(k, M, e0);
    const float v2 = prep_selec(k, M, e1);
    e2[0] = e0[0];             // corners as the first two calls left them
    e2[M - 1] = e1[0];
    const float v3 = prep_selec(k, M, e2);
    e3[0] = e0[M - 1];
    e3[M - 1] = e1[M - 1];
    const float v4 = prep_selec(k, M, e3);
    med_s = (v1 + v2 + v3 + v4) / 4.0f;
  }
  __syncthreads();
  const float med = med_s;
  const bool sub = fabsf(med) > PREP_TINY;
  const float half = -(float)M / 2.0f;
  for (int p = tid; p < M * M; p += blockDim.x) {
    const int ix = p % M, iy = p / M;
    float v = x[p];
    if (sub) v = v - med;
    // mirrored loop indices of mask (:6262-6275): i, j <= M/2
    const int i = ix < M / 2 ? ix + 1 : M - ix;
    const int j = iy < M / 2 ? iy + 1 : M - iy;
    const float ci = half + (float)(i - 1), cj = half + (float)(j - 1);
    float e = prep_cosedge(ci, cj, minlen, mskrad);
    if (inner > 0.0f) e = e * prep_cosedge_inner(ci, cj, width, inner);
    if (!(e > 0.9999f)) v = e * v;
    if (instr_w) v = v / (instr_w[ix] * instr_w[iy]);
    x[p] = v;
  }
}

}  // namespace pfk
