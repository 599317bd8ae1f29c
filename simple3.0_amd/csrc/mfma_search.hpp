// mfma_search.hpp - the precision form: per-angular-frequency contraction on MFMA.
//
// With a ring-constant |CTF| (ctf = no, no astigmatism, or its ring mean c_p,k), the rotational
// correlation of every (reference r, particle p) pair is
//   Z[r,p,m] = sum_k conj(V_r[k,m]) * c_p,k * W_p[k,m],   m = 0..n,
//   cc[r,p,rho] = irfft_2n(Z[r,p,:])[rho] / (4n) / sqrt(sum_k c_p,k^2 S_r[k] * S_p),
// with V_r, W_p the length-2n real FFTs of the rings (the particle memo, and the same transform of
// the references).  For each m this is one dense complex GEMM over the shells, which runs on
// MFMA in split BF16: x = hi + lo (hi = bf16(x), lo = bf16(x - hi)), products hi*hi + hi*lo +
// lo*hi accumulated in FP32 (the lo*lo term, 2^-16 relative, is dropped).  Real formulation with
// k' = 2k + {re, im}: Zr = A1 . B, Zi = A2 . B, A1 = (Re V, Im V), A2 = (-Im V, Re V), B = (Re W', Im W').
// The spectra Z (FP32) go through HBM to the inverse-FFT kernel, which normalises with FP64
// denominators and reduces each particle to its best (reference, rotation).  Z is stored
// particle-group major, [r][p / 16][m][p % 16], so that one inverse-FFT work-group (one reference,
// 16 particles) reads one contiguous block of (n + 1) x 16 complex values.
#pragma once
#include "corr_fused.hpp"

namespace pfk {

constexpr int ZG = 16;   // particles per Z group (the MFMA tile's column count)

__host__ __device__ __forceinline__ size_t zidx(int r, int p, int m, int n, int cols_pad) {
  return ((size_t)(r * (cols_pad / ZG) + p / ZG) * (n + 1) + m) * ZG + (p % ZG);
}
// FP16 spectra (4 bytes per value): groups of ZG16 particles, so that one (row, m) run fills a
// 128-byte line
#ifndef PFTCC_ZG16
#define PFTCC_ZG16 32
#endif
constexpr int ZG16 = PFTCC_ZG16;
__host__ __device__ __forceinline__ size_t zidx16(int r, int p, int m, int n, int cols_pad) {
  return ((size_t)(r * (cols_pad / ZG16) + p / ZG16) * (n + 1) + m) * ZG16 + (p % ZG16);
}

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// One 16x16x32 MFMA step on packed 16-bit operands (8 per lane): BF16, or FP16 when H.  Both
// formats keep the sign in bit 15 of each element, so rot_neg below serves both.
template <bool H>
__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
  if constexpr (H)
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0,
                                                  0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Ring mean of |CTF| per (particle, shell) and its square in FP64.
__global__ void k_ring_mean_ctf(const float* __restrict__ ctf, int np, int nk, int n, float* __restrict__ cbar,
                                double* __restrict__ cbar2) {
  const int item = blockIdx.x * blockDim.y + threadIdx.y;   // (p, kk)
  if (item >= np * nk) return;
  const float* c = ctf + (size_t)item * n;
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) s += c[j];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (threadIdx.x == 0) {
    const double m = s / n;
    cbar[item] = (float)m;
    cbar2[item] = m * m;
  }
}

// FP64 per-shell sums of squares of the reference rings: sr[r][kk], kk < kcount.
__global__ void k_ring_sq64(const float2* __restrict__ rings, int count, int nk, int n, int kcount,
                            double* __restrict__ out) {
  const int item = blockIdx.x * blockDim.y + threadIdx.y;   // (r, kk)
  if (item >= count * kcount) return;
  const int r = item / kcount, kk = item % kcount;
  const float2* v = rings + ((size_t)r * nk + kk) * n;
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) s += (double)v[j].x * v[j].x + (double)v[j].y * v[j].y;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (threadIdx.x == 0) out[item] = s;
}

__device__ __forceinline__ unsigned short bf16_bits(float x) {
  const __bf16 h = (__bf16)x;   // round to nearest even
  return *reinterpret_cast<const unsigned short*>(&h);
}
__device__ __forceinline__ float bf16_float(unsigned short b) {
  return __uint_as_float((unsigned int)b << 16);
}
__device__ __forceinline__ unsigned short f16_bits(float x) {
  return __builtin_bit_cast(unsigned short, (_Float16)x);   // round to nearest even
}
__device__ __forceinline__ float f16_float(unsigned short b) {
  return (float)__builtin_bit_cast(_Float16, b);
}

// Power of two g with max * g <= lim (1 when the operand is all zero); max is the bit pattern of
// a non-negative float.
__device__ __forceinline__ float pow2_gain(const unsigned int* gmax, float lim) {
  const float m = __uint_as_float(*gmax);
  if (!(m > 0.f)) return 1.0f;
  int e;
  (void)frexpf(lim / m, &e);   // lim / m = f 2^e, f in [0.5, 1)
  return ldexpf(1.0f, e - 1);
}
constexpr float kF16Lim = 32768.0f;   // FP16 operands: largest component <= 2^15 (max normal 65504)

// Split-BF16 operand layout, in 16-byte vectors (8 k' each): [m][row / 16][k' / 32][row % 16][4],
// so that the 64 lanes of one MFMA operand load (16 rows x 32 k') read one contiguous kilobyte.
__host__ __device__ __forceinline__ size_t split_vec(int m, int row, int vk, int rows_pad, int KP) {
  return ((((size_t)m * (rows_pad / 16) + row / 16) * (KP / 32) + vk / 4) * 16 + (row % 16)) * 4 + (vk % 4);
}

// Packs spectra spec[row][kk][m] (float2, m in [0, n]) times an optional per-(row, kk) scale into
// split 16-bit operands hi/lo (k' = 2 kk + {0: re, 1: im}, zero past 2 kcorr; layout split_vec):
// BF16, or FP16 when H, where the values are also multiplied by the shell scale shellscale[kk]
// and the operand's power-of-two gain (largest component <= 2^15, gmax from k_absmax_scaled).
// One thread per (m, row, complex k); rows beyond count and k' beyond 2 kcorr are written as zero.
template <bool H>
__global__ void k_pack_split(const float2* __restrict__ spec, const float* __restrict__ scale,
                             const float* __restrict__ shellscale, const unsigned int* __restrict__ gmax, int count,
                             int rows_pad, int nk, int kcorr, int n, int KP, unsigned int* __restrict__ hi,
                             unsigned int* __restrict__ lo) {
  const size_t KW = KP / 2;                                  // 32-bit words (one complex) per row
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t tot = (size_t)(n + 1) * rows_pad * KW;
  if (idx >= tot) return;
  const int kk = (int)(idx % KW);
  const int row = (int)((idx / KW) % rows_pad);
  const int m = (int)(idx / (KW * rows_pad));
  float re = 0.f, im = 0.f;
  if (row < count && kk < kcorr) {
    const float2 v = spec[((size_t)row * nk + kk) * (n + 1) + m];
    float s = scale ? scale[(size_t)row * nk + kk] : 1.0f;
    if (H) s *= shellscale[kk] * pow2_gain(gmax, kF16Lim);
    re = v.x * s;
    im = v.y * s;
  }
  unsigned short hr, hiim, lr, li;
  if (H) {
    hr = f16_bits(re);
    hiim = f16_bits(im);
    lr = f16_bits(re - f16_float(hr));
    li = f16_bits(im - f16_float(hiim));
  } else {
    hr = bf16_bits(re);
    hiim = bf16_bits(im);
    lr = bf16_bits(re - bf16_float(hr));
    li = bf16_bits(im - bf16_float(hiim));
  }
  const size_t w = split_vec(m, row, kk / 4, rows_pad, KP) * 4 + (kk % 4);
  hi[w] = (unsigned int)hr | ((unsigned int)hiim << 16);
  lo[w] = (unsigned int)lr | ((unsigned int)li << 16);
}

// A2 fragment from A1: per complex (re, im) word -> (-im, re).
__device__ __forceinline__ bf16x8 rot_neg(bf16x8 a) {
  union U {
    bf16x8 v;
    unsigned int w[4];
  } u;
  u.v = a;
#pragma unroll
  for (int i = 0; i < 4; ++i) u.w[i] = ((u.w[i] >> 16) | (u.w[i] << 16)) ^ 0x00008000u;
  return u.v;
}

// Z[r][p][m] = sum_k' A[m][r][k'] (x) B[m][p][k'] for one 64 x 64 (ref x particle) tile per wave.
// split = 3: hi.hi + hi.lo + lo.hi; split = 1: hi.hi only.  Operands are BF16, or FP16 when H (then
// Z is multiplied back by 1 / (gA gB), the operands' power-of-two gains).
// Tile coordinates of a 1-D launch of gx * gy * (n+1) work-groups: the XCD-aware order puts the
// particle tiles that share one reference tile (and m) on one XCD, next to each other in time.
struct ZTile {
  int p0, r0, m;
};
__device__ __forceinline__ ZTile ztile(int gx, int gy, int tp, int tr) {
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  ZTile z;
  z.p0 = (t % gx) * tp;
  z.r0 = ((t / gx) % gy) * tr;
  z.m = t / (gx * gy);
  return z;
}

#ifndef PFTCC_ZGEMM_TR
#define PFTCC_ZGEMM_TR 4
#endif
#ifndef PFTCC_ZGEMM_TP
#define PFTCC_ZGEMM_TP 4
#endif
constexpr int ZGEMM_TR = PFTCC_ZGEMM_TR;   // 16-row reference blocks per wavefront tile
constexpr int ZGEMM_TP = PFTCC_ZGEMM_TP;   // 16-column particle blocks per wavefront tile

// One (16 TR) x (16 TP) (reference x particle) tile per wavefront, 32 k' per MFMA step.
template <int SPLIT, int TR, int TP, bool H>
__global__ void __launch_bounds__(64) k_zgemm(const bf16x8* __restrict__ Ahi, const bf16x8* __restrict__ Alo,
                                              const bf16x8* __restrict__ Bhi, const bf16x8* __restrict__ Blo,
                                              int rows_pad, int cols_pad, int KP, int n, float2* __restrict__ Z,
                                              const unsigned int* __restrict__ gmaxA,
                                              const unsigned int* __restrict__ gmaxB) {
  const int l = threadIdx.x;
  const ZTile zt = ztile(cols_pad / (16 * TP), rows_pad / (16 * TR), 16 * TP, 16 * TR);
  const int p0 = zt.p0, r0 = zt.r0, m = zt.m;
  const int KV = KP / 8;                                  // bf16x8 vectors per row
  f32x4 zr[TR][TP], zi[TR][TP];
#pragma unroll
  for (int i = 0; i < TR; ++i)
#pragma unroll
    for (int j = 0; j < TP; ++j) {
      zr[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
      zi[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  const int lr = l % 16, lk = l / 16;
  for (int ks = 0; ks < KV; ks += 4) {                    // 32 k' per step, lane takes 8 of them
    bf16x8 ah[TR], al[TR], bh[TP], bl[TP];
#pragma unroll
    for (int i = 0; i < TR; ++i) {
      const size_t ao = split_vec(m, r0 + i * 16 + lr, ks + lk, rows_pad, KP);
      ah[i] = Ahi[ao];
      if (SPLIT == 3) al[i] = Alo[ao];
    }
#pragma unroll
    for (int j = 0; j < TP; ++j) {
      const size_t bo = split_vec(m, p0 + j * 16 + lr, ks + lk, cols_pad, KP);
      bh[j] = Bhi[bo];
      if (SPLIT == 3) bl[j] = Blo[bo];
    }
#pragma unroll
    for (int i = 0; i < TR; ++i) {
      const bf16x8 a2h = rot_neg(ah[i]);
      bf16x8 a2l;
      if (SPLIT == 3) a2l = rot_neg(al[i]);
#pragma unroll
      for (int j = 0; j < TP; ++j) {
        if (SPLIT == 3) {
          zr[i][j] = mfma16<H>(al[i], bh[j], zr[i][j]);
          zr[i][j] = mfma16<H>(ah[i], bl[j], zr[i][j]);
          zi[i][j] = mfma16<H>(a2l, bh[j], zi[i][j]);
          zi[i][j] = mfma16<H>(a2h, bl[j], zi[i][j]);
        }
        zr[i][j] = mfma16<H>(ah[i], bh[j], zr[i][j]);
        zi[i][j] = mfma16<H>(a2h, bh[j], zi[i][j]);
      }
    }
  }
  const float inv = H ? 1.0f / (pow2_gain(gmaxA, kF16Lim) * pow2_gain(gmaxB, kF16Lim)) : 1.0f;
  // D layout: lane l, element e -> row 4 (l/16) + e, column l % 16
#pragma unroll
  for (int i = 0; i < TR; ++i)
#pragma unroll
    for (int j = 0; j < TP; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = r0 + i * 16 + 4 * lk + e;
        const int p = p0 + j * 16 + lr;
        Z[zidx(r, p, m, n, cols_pad)] = make_float2(zr[i][j][e] * inv, zi[i][j][e] * inv);
      }
}

#ifndef PFTCC_ZGEMM_AS_TP
#define PFTCC_ZGEMM_AS_TP 2   // 32-column tiles: 252 registers, 2 wavefronts per SIMD (64 columns: 1)
#endif

// A-stationary variant: one wavefront holds the reference operand of a 16 TR-row slab at one m
// (all NSTEP k' steps) in registers and sweeps every 16 TP-column particle tile of the chunk, so the
// large operand is read once per launch and only the particle operand (prefetched one step ahead)
// is loaded in the inner loop.  The XCD-aware order keeps the slabs of one m on one XCD (the
// particle operand of that m stays in its L2).
template <int SPLIT, int TR, int NSTEP, bool H>
#ifndef PFTCC_ZGEMM_AS_W
#define PFTCC_ZGEMM_AS_W 1     // minimum wavefronts per SIMD imposed on the register allocation (tuning builds)
#endif
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PFTCC_ZGEMM_AS_W)))
k_zgemm_as(const bf16x8* __restrict__ Ahi, const bf16x8* __restrict__ Alo,
                                                 const bf16x8* __restrict__ Bhi, const bf16x8* __restrict__ Blo,
                                                 int rows_pad, int cols_pad, int n, float2* __restrict__ Z,
                                                 const unsigned int* __restrict__ gmaxA,
                                                 const unsigned int* __restrict__ gmaxB) {
  constexpr int TP = PFTCC_ZGEMM_AS_TP;        // 16-column particle blocks per tile
  constexpr int KP = NSTEP * 32;
  const int l = threadIdx.x;
  const int gr = rows_pad / (16 * TR);
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int r0 = (t % gr) * 16 * TR, m = t / gr;
  const int lr = l % 16, lk = l / 16;
  bf16x8 ah[NSTEP][TR], al[NSTEP][TR];
#pragma unroll
  for (int st = 0; st < NSTEP; ++st)
#pragma unroll
    for (int i = 0; i < TR; ++i) {
      const size_t ao = split_vec(m, r0 + i * 16 + lr, 4 * st + lk, rows_pad, KP);
      ah[st][i] = Ahi[ao];
      if (SPLIT == 3) al[st][i] = Alo[ao];
    }
  bf16x8 bh[TP], bl[TP], nh[TP], nl[TP];
  auto load_b = [&](int p0, int st, bf16x8 (&h)[TP], bf16x8 (&lo)[TP]) {
#pragma unroll
    for (int j = 0; j < TP; ++j) {
      const size_t bo = split_vec(m, p0 + j * 16 + lr, 4 * st + lk, cols_pad, KP);
      h[j] = Bhi[bo];
      if (SPLIT == 3) lo[j] = Blo[bo];
    }
  };
  const int ntile = cols_pad / (16 * TP);
  const float inv = H ? 1.0f / (pow2_gain(gmaxA, kF16Lim) * pow2_gain(gmaxB, kF16Lim)) : 1.0f;
  load_b(0, 0, bh, bl);
  for (int tile = 0; tile < ntile; ++tile) {
    const int p0 = tile * 16 * TP;
    f32x4 zr[TR][TP], zi[TR][TP];
#pragma unroll
    for (int i = 0; i < TR; ++i)
#pragma unroll
      for (int j = 0; j < TP; ++j) {
        zr[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        zi[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
    for (int st = 0; st < NSTEP; ++st) {
      // prefetch the next step (or the next tile's first step)
      const bool more = st + 1 < NSTEP || tile + 1 < ntile;
      if (more) load_b(st + 1 < NSTEP ? p0 : p0 + 16 * TP, st + 1 < NSTEP ? st + 1 : 0, nh, nl);
#pragma unroll
      for (int i = 0; i < TR; ++i) {
        const bf16x8 a2h = rot_neg(ah[st][i]);
        bf16x8 a2l;
        if (SPLIT == 3) a2l = rot_neg(al[st][i]);
#pragma unroll
        for (int j = 0; j < TP; ++j) {
          if (SPLIT == 3) {
            zr[i][j] = mfma16<H>(al[st][i], bh[j], zr[i][j]);
            zr[i][j] = mfma16<H>(ah[st][i], bl[j], zr[i][j]);
            zi[i][j] = mfma16<H>(a2l, bh[j], zi[i][j]);
            zi[i][j] = mfma16<H>(a2h, bl[j], zi[i][j]);
          }
          zr[i][j] = mfma16<H>(ah[st][i], bh[j], zr[i][j]);
          zi[i][j] = mfma16<H>(a2h, bh[j], zi[i][j]);
        }
      }
      if (more) {
#pragma unroll
        for (int j = 0; j < TP; ++j) {
          bh[j] = nh[j];
          if (SPLIT == 3) bl[j] = nl[j];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < TR; ++i)
#pragma unroll
      for (int j = 0; j < TP; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = r0 + i * 16 + 4 * lk + e;
          const int p = p0 + j * 16 + lr;
          Z[zidx(r, p, m, n, cols_pad)] = make_float2(zr[i][j][e] * inv, zi[i][j][e] * inv);
        }
  }
}

struct ZifftArgs {
  int n, nrefs, np, rows_pad, cols_pad, kcorr;
  int ns;                      // shifts per reference: row r belongs to reference r / ns
  int knorm;                   // shells in the reference norm (kcorr unshifted, nk shifted)
  const float2* Z;             // zidx layout
  const double* sr;            // [nrefs / ns][knorm] reference shell sums of squares
  const double* cbar2;         // [np_chunk][nk] ring-mean |CTF|^2 of the chunk's particles (or null)
  int nk;
  const double* sp;            // [np_chunk] particle sums of squares (FP64)
  const int* rlist;            // key id of each row (global reference * ns + shift; null = row index)
  const int* outpos;           // [np_chunk] output position of each chunk particle
  float* cc_out;               // [npos][nrefs][2n] or null
  unsigned long long* best;    // [npos] or null
  unsigned long long* prbest;  // [npos][nrefs] best key of each (particle, reference), or null
  const float2* tw2g;
  FftPlan plan;
  // FP16 spectra (FP8 coarse pass): Z16 holds Z gA gB zs as FP16 pairs in the zidx16 layout, with
  // gA, gB the operands' power-of-two gains (from gmaxA, gmaxB) and zs a power of two
  const unsigned int* Z16;
  const unsigned int* gmaxA;
  const unsigned int* gmaxB;
  double zs;
};

// Inverse real FFT, FP64-normalised correlations and argmax for F pairs (one reference row,
// F consecutive particles) per work-group.  Dynamic LDS: X[F][n+1], buf A/B [F][n], tw2[2n].
template <int F>
__global__ void __launch_bounds__(256) k_zifft(ZifftArgs a) {
  extern __shared__ float2 smem[];
  const int n = a.n, tid = threadIdx.x, nt = blockDim.x;
  float2* X = smem;                        // [F][n+1], reused as the FFT's second buffer
  float2* bufA = X + F * (n + 1);
  float2* bufB = X;
  float2* tw2 = bufA + F * n;
  __shared__ double den[F];
  __shared__ unsigned long long red[256];
  const int r = blockIdx.y;               // reference row
  const int p0 = blockIdx.x * F;
  for (int i = tid; i < 2 * n; i += nt) tw2[i] = a.tw2g[i];
  // gather Z[m][r][p0 .. p0+F-1]
  for (int idx = tid; idx < (n + 1) * F; idx += nt) {
    const int m = idx / F, f = idx % F;
    X[f * (n + 1) + m] = a.Z[zidx(r, p0 + f, m, n, a.cols_pad)];
  }
  if (tid < F) {
    const int p = p0 + tid;
    double d = 0.0;
    if (p < a.np) {
      double sq = 0.0;
      for (int kk = 0; kk < a.knorm; ++kk)
        sq += a.sr[(size_t)(r / a.ns) * a.knorm + kk] * (a.cbar2 ? a.cbar2[(size_t)p * a.nk + kk] : 1.0);
      d = sqrt(sq * a.sp[p]);
    }
    den[tid] = d;
  }
  __syncthreads();
  for (int idx = tid; idx < F * n; idx += nt) {
    const int f = idx / n, m = idx % n;
    float2 x = X[f * (n + 1) + m];
    float2 xn = X[f * (n + 1) + (n - m)];
    if (m == 0) {
      x.y = 0.f;
      xn.y = 0.f;
    }
    bufA[f * n + m] = unsplit_bin_conj(x, xn, cconj(tw2[m]));
  }
  __syncthreads();
  float2* res = fft_forward_oop(bufA, bufB, F, n, n, a.plan, tw2, tid, nt);
  const double inv4n = 1.0 / (4.0 * n);
  unsigned long long bestkey[F];
#pragma unroll
  for (int f = 0; f < F; ++f) bestkey[f] = 0ull;
  const int iref = a.rlist ? a.rlist[r] : r;
  for (int idx = tid; idx < F * n; idx += nt) {
    const int f = idx / n, j = idx % n;
    const int p = p0 + f;
    if (p >= a.np || r >= a.nrefs) continue;
    const float2 y = res[f * n + j];
    const double sc = inv4n / den[f];
    const float x0 = (float)(y.x * sc), x1 = (float)(-y.y * sc);
    if (a.cc_out) {
      float2* o = reinterpret_cast<float2*>(a.cc_out + ((size_t)a.outpos[p] * a.nrefs + r) * (2 * n));
      o[j] = make_float2(x0, x1);
    }
    if (a.best || a.prbest) {
      const unsigned int lin0 = (unsigned int)iref * (unsigned int)(2 * n) + 2u * j;
      const unsigned long long k0 = ((unsigned long long)ordered_bits(x0) << 32) | (0xFFFFFFFFu - lin0);
      const unsigned long long k1 = ((unsigned long long)ordered_bits(x1) << 32) | (0xFFFFFFFFu - lin0 - 1u);
#pragma unroll
      for (int g = 0; g < F; ++g)
        if (g == f) {
          bestkey[g] = k0 > bestkey[g] ? k0 : bestkey[g];
          bestkey[g] = k1 > bestkey[g] ? k1 : bestkey[g];
        }
    }
  }
  if (a.best || a.prbest) {
#pragma unroll
    for (int f = 0; f < F; ++f) {
      red[tid] = bestkey[f];
      __syncthreads();
      for (int w = nt / 2; w > 0; w >>= 1) {
        if (tid < w && red[tid + w] > red[tid]) red[tid] = red[tid + w];
        __syncthreads();
      }
      if (tid == 0 && p0 + f < a.np && r < a.nrefs) {
        if (a.best) atomicMax(&a.best[a.outpos[p0 + f]], red[0]);
        if (a.prbest) a.prbest[(size_t)a.outpos[p0 + f] * a.nrefs + r] = red[0];
      }
      __syncthreads();
    }
  }
}

// The inverse stage with its first and last passes in registers.
//  * pass 1 (radix 4, from global): a thread takes the mirrored butterflies g and N/4 - g, whose
//    inputs Y[m] = f(Z[m], Z[N-m]) need exactly the same 8 spectrum values;
//  * middle passes in LDS;
//  * the last pass leaves its outputs in registers, where they are normalised, written (cc_out)
//    and reduced to per-particle best keys (LDS 64-bit max, order independent).
// Four consecutive float2 as two 16-byte LDS stores (o 16-byte aligned).
__device__ __forceinline__ void store4(float2* o, const float2 (&y)[4]) {
  float4* d = reinterpret_cast<float4*>(o);
  d[0] = make_float4(y[0].x, y[0].y, y[1].x, y[1].y);
  d[1] = make_float4(y[2].x, y[2].y, y[3].x, y[3].y);
}

typedef _Float16 half2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 h2_to_f2(unsigned int w) {
  const half2v h = __builtin_bit_cast(half2v, w);
  return make_float2((float)h.x, (float)h.y);
}

template <int N, int F, int NT, bool Z16>
__global__ void __launch_bounds__(NT) k_zifft_v2(ZifftArgs a) {
  static_assert(ZG % F == 0, "F divides the Z group");
  static_assert(Fp<N>::fusable, "first radix must be 4");
  constexpr int TPF = NT / F;
  static_assert(TPF * F == NT && TPF <= 64 && (TPF & (TPF - 1)) == 0, "TPF is a power of two <= 64");
  constexpr int NP = Fp<N>::np;
  constexpr int Q = N / 4;
  constexpr int G1 = Q / 2 + 1;                  // mirrored pass-1 groups per transform
  constexpr int T1 = F * G1;
  constexpr int TPT1 = (T1 + NT - 1) / NT;
  constexpr int RL = Fp<N>::R(NP - 1);
  constexpr int NsL = N / RL;
  extern __shared__ __align__(16) float2 smem[];
  float2* buf = smem;                    // [F][N]
  float2* tw2 = buf + F * N;             // [2N]
  __shared__ double den[F];
  const int tid = threadIdx.x;
  // 1-D launch of gx * nrefs work-groups; XCD-aware order keeps the work-groups that read one Z
  // group (consecutive particle blocks of one reference) on one XCD
  const int gx = (a.np + F - 1) / F;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int r = t / gx;
  const int p0 = (t % gx) * F;
  for (int i = tid; i < 2 * N; i += NT) tw2[i] = a.tw2g[i];
  // ---- pass 1: loads first (they do not depend on LDS), then the butterflies ----
  static_assert(ZG16 % F == 0, "F divides the FP16 Z group");
  const float2* Zb = a.Z + zidx(r, p0, 0, N, a.cols_pad);
  const unsigned int* Zh = a.Z16 + (Z16 ? zidx16(r, p0, 0, N, a.cols_pad) : 0);
  float2 za[TPT1][4], zb[TPT1][4], zn[TPT1];
#pragma unroll
  for (int c = 0; c < TPT1; ++c) {
    const int task = tid + c * NT;
    if (!(T1 % NT == 0 || task < T1)) continue;
    const int g = task / F, f = task % F;
    // element offsets below (N + 1) ZG16 from the wave-uniform block base: 32-bit
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if constexpr (Z16) {
        za[c][t] = h2_to_f2(*reinterpret_cast<const unsigned int*>(
            reinterpret_cast<const char*>(Zh) + ((unsigned int)((g + t * Q) * ZG16 + f) << 2)));
        zb[c][t] = h2_to_f2(*reinterpret_cast<const unsigned int*>(
            reinterpret_cast<const char*>(Zh) + ((unsigned int)(((t + 1) * Q - g) * ZG16 + f) << 2)));
      } else {
        za[c][t] = ld8(Zb, (unsigned int)((g + t * Q) * ZG + f));
        zb[c][t] = ld8(Zb, (unsigned int)(((t + 1) * Q - g) * ZG + f));
      }
    }
    zn[c] = zb[c][3];   // Z[N - g]; for g = 0 this is Z[N]
  }
  // FP64 denominators (their loads overlap the spectra's)
  {
    const int f = tid / TPF, u = tid % TPF;
    const int p = p0 + f;
    double s = 0.0;
    if (p < a.np)
      for (int kk = u; kk < a.knorm; kk += TPF)
        s += a.sr[(size_t)(r / a.ns) * a.knorm + kk] * (a.cbar2 ? a.cbar2[(size_t)p * a.nk + kk] : 1.0);
#pragma unroll
    for (int o = TPF / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, TPF);
    if (u == 0) den[f] = p < a.np ? sqrt(s * a.sp[p]) : 1.0;
  }
  __syncthreads();   // tw2
#pragma unroll
  for (int c = 0; c < TPT1; ++c) {
    const int task = tid + c * NT;
    if (!(T1 % NT == 0 || task < T1)) continue;
    const int g = task / F, f = task % F;
    float2* o = buf + f * N;
    float2 y[4];
    if (g == 0) {
      float2 x0 = za[c][0], xn = zn[c];
      x0.y = 0.f;
      xn.y = 0.f;
      y[0] = unsplit_bin_conj(x0, xn, cconj(tw2[0]));
      y[1] = unsplit_bin_conj(za[c][1], za[c][3], cconj(tw2[Q]));
      y[2] = unsplit_bin_conj(za[c][2], za[c][2], cconj(tw2[2 * Q]));
      y[3] = unsplit_bin_conj(za[c][3], za[c][1], cconj(tw2[3 * Q]));
      Dft<4>::run(y);
      store4(o, y);
    } else {
#pragma unroll
      for (int t = 0; t < 4; ++t) y[t] = unsplit_bin_conj(za[c][t], zb[c][3 - t], cconj(tw2[g + t * Q]));
      Dft<4>::run(y);
      store4(o + 4 * g, y);
      if (2 * g != Q) {
#pragma unroll
        for (int t = 0; t < 4; ++t) y[t] = unsplit_bin_conj(zb[c][t], za[c][3 - t], cconj(tw2[(t + 1) * Q - g]));
        Dft<4>::run(y);
        store4(o + 4 * (Q - g), y);
      }
    }
  }
  __syncthreads();
  lds_passes<N, F, NT, 1, NP - 1>(buf, tw2, tid);
  // ---- last pass in registers, then normalisation, output and best keys ----
  // Each wavefront owns whole transforms (f = wave, wave + NW, ...), so a pair's best key is a
  // wavefront reduction with no LDS atomics.
  const double inv4n = 1.0 / (4.0 * N);
  const int iref = a.rlist ? a.rlist[r] : r;
  const bool want_key = a.best || a.prbest;
  constexpr int NW = NT / 64;
  const int wave = tid >> 6, lane = tid & 63;
  for (int f = wave; f < F; f += NW) {
    const int p = p0 + f;
    const bool live = p < a.np;
    // FP64 normalisation factor, applied as one FP32 multiply per value (FP16 spectra: with the
    // powers of two that scaled them)
    float sc;
    if constexpr (Z16)
      sc = live ? (float)(inv4n / den[f] / ((double)pow2_gain(a.gmaxA, 448.0f) * pow2_gain(a.gmaxB, 448.0f) * a.zs))
                : 0.f;
    else
      sc = live ? (float)(inv4n / den[f]) : 0.f;
    float bv = -INFINITY;      // this lane's best value and its rotation index (ties: lowest)
    int bi = -1;
    // the lane's butterflies j = lane, lane + 64, ... (JB of them), transformed first; their outputs
    // jj = j + t NsL are then visited t-major, i.e. in increasing rotation order, so that a strict
    // comparison keeps the lowest index among equal values
    constexpr int JB = (NsL + 63) / 64;
    float2 v[JB][RL];
#pragma unroll
    for (int jb = 0; jb < JB; ++jb) {
      const int j = jb * 64 + lane;
      if (j >= NsL) continue;
#pragma unroll
      for (int t = 0; t < RL; ++t) v[jb][t] = buf[f * N + j + t * NsL];
#pragma unroll
      for (int t = 1; t < RL; ++t) v[jb][t] = cmul(v[jb][t], tw2[2 * j * t]);
      Dft<RL>::run(v[jb]);
    }
    if (live) {
#pragma unroll
      for (int t = 0; t < RL; ++t)
#pragma unroll
        for (int jb = 0; jb < JB; ++jb) {
          const int j = jb * 64 + lane;
          if (j >= NsL) continue;
          const int jj = j + t * NsL;                   // output index: rotations 2 jj, 2 jj + 1
          const float x0 = v[jb][t].x * sc, x1 = -v[jb][t].y * sc;
          if (a.cc_out) {
            float2* oo = reinterpret_cast<float2*>(a.cc_out + ((size_t)a.outpos[p] * a.nrefs + r) * (2 * N));
            oo[jj] = make_float2(x0, x1);
          }
          if (x0 > bv) {
            bv = x0;
            bi = 2 * jj;
          }
          if (x1 > bv) {
            bv = x1;
            bi = 2 * jj + 1;
          }
        }
    }
    if (want_key && live) {
      unsigned long long bk = 0ull;
      if (bi >= 0)
        bk = ((unsigned long long)ordered_bits(bv) << 32) |
             (0xFFFFFFFFu - ((unsigned int)iref * (unsigned int)(2 * N) + (unsigned int)bi));
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(bk, o);
        bk = w > bk ? w : bk;
      }
      if (lane == 0) {
        if (a.best) atomicMax(&a.best[a.outpos[p]], bk);
        if (a.prbest) a.prbest[(size_t)a.outpos[p] * a.nrefs + r] = bk;
      }
    }
  }
}

#ifndef PFTCC_ZIFFT_F
#define PFTCC_ZIFFT_F 4
#endif
#ifndef PFTCC_ZIFFT_NT
#define PFTCC_ZIFFT_NT 256
#endif

// Transforms per work-group: 8 at N = 308, whose 11-point pass has 28 butterflies per transform
// (4 transforms fill 112 of 256 lanes; 8 ran the inverse stage 12.5 % faster, 16 slower), 4
// elsewhere (N = 400: 8 and 16 measured slower).
template <int N>
struct ZifftV2 {
  static constexpr int F0 = (N == 308 && PFTCC_ZIFFT_F == 4) ? 8 : PFTCC_ZIFFT_F;
  static constexpr int F = (sizeof(float2) * (F0 * N + 2 * N) <= 65536) ? F0
                         : (sizeof(float2) * (F0 / 2 * N + 2 * N) <= 65536) ? F0 / 2 : F0 / 4;
  // at most one wavefront per transform
  static constexpr int NT = PFTCC_ZIFFT_NT < 64 * F ? PFTCC_ZIFFT_NT : 64 * F;
  static constexpr size_t lds = sizeof(float2) * ((size_t)F * N + 2 * N);
};

// ---------------------------------------------------------------- FP8 e4m3 coarse pass ---
// Operands are scaled per shell by a power of two, V'_k = a_k V_k and W'_k = W_k / a_k, so the
// contraction sum_k conj(V_k) W_k is unchanged before rounding (a_k balances the magnitudes of the
// two operands shell by shell), and then by a global power of two per operand that maps its largest
// component to at most 448, the e4m3 maximum.  Z is scaled back by the two global factors.
typedef int i32x8 __attribute__((ext_vector_type(8)));

// Per-shell sums of |spec|^2 over listed rows (list null: rows 0..count-1) and m; partial sums
// [kk][gridDim.y] in a fixed order.
__global__ void __launch_bounds__(256) k_shell_ms(const float2* __restrict__ spec, const int* __restrict__ list,
                                                  int count, int nk, int n, double* __restrict__ part) {
  __shared__ double red[256];
  const int kk = blockIdx.x, chunk = blockIdx.y, nch = gridDim.y;
  const int per = (count + nch - 1) / nch;
  const int r0 = chunk * per, r1 = min(count, r0 + per);
  double s = 0.0;
  const int len = n + 1;
  for (int e = threadIdx.x; e < (r1 - r0) * len; e += 256) {
    const int rl = r0 + e / len, m = e % len;
    const int row = list ? list[rl] : rl;
    const float2 v = spec[((size_t)row * nk + kk) * len + m];
    s += (double)v.x * v.x + (double)v.y * v.y;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)kk * nch + chunk] = red[0];
}

// Largest |component| of spec[row][kk][m] * rowscale[row][kk] * shellscale[kk] (kk < kcorr), as
// the bits of a non-negative float (atomicMax on the bit pattern orders non-negative floats).
// Grid-stride loop, 256 threads, one atomic per work-group.
__global__ void k_absmax_scaled(const float2* __restrict__ spec, const float* __restrict__ rowscale,
                                const float* __restrict__ shellscale, int count, int nk, int n, int kcorr,
                                unsigned int* __restrict__ out) {
  const size_t tot = (size_t)count * kcorr * (n + 1);
  float mx = 0.f;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (size_t)gridDim.x * blockDim.x) {
    const int m = (int)(idx % (n + 1));
    const int kk = (int)((idx / (n + 1)) % kcorr);
    const int row = (int)(idx / ((size_t)(n + 1) * kcorr));
    const float2 v = spec[((size_t)row * nk + kk) * (n + 1) + m];
    const float sc = (rowscale ? rowscale[(size_t)row * nk + kk] : 1.0f) * shellscale[kk];
    mx = fmaxf(mx, fmaxf(fabsf(v.x * sc), fabsf(v.y * sc)));
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  __shared__ float wmax[4];
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float b = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
    if (b > 0.f) atomicMax(out, __float_as_uint(b));
  }
}

// Power of two g with max * g <= 448 (the inverse stage's FP16-spectra scale uses the same gain).
__device__ __forceinline__ float fp8_gain(const unsigned int* gmax) { return pow2_gain(gmax, 448.0f); }

// e4m3 operand tiles out[m][row][k'] (bytes; k' = 2 kk + {re, im}, zero past 2 kcorr), one 32-bit
// word (two complex values) per thread.
__global__ void k_pack_fp8(const float2* __restrict__ spec, const float* __restrict__ rowscale,
                           const float* __restrict__ shellscale, const unsigned int* __restrict__ gmax, int count,
                           int rows_pad, int nk, int kcorr, int n, int KP8, unsigned int* __restrict__ out) {
  const size_t KW = KP8 / 4;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t tot = (size_t)(n + 1) * rows_pad * KW;
  if (idx >= tot) return;
  const int w = (int)(idx % KW);
  const int row = (int)((idx / KW) % rows_pad);
  const int m = (int)(idx / (KW * rows_pad));
  const float g = fp8_gain(gmax);
  float v[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int kk = 2 * w + h;
    if (row < count && kk < kcorr) {
      const float2 x = spec[((size_t)row * nk + kk) * (n + 1) + m];
      const float sc = (rowscale ? rowscale[(size_t)row * nk + kk] : 1.0f) * shellscale[kk] * g;
      v[2 * h] = x.x * sc;
      v[2 * h + 1] = x.y * sc;
    }
  }
  unsigned int word = (unsigned int)__builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
  word = (unsigned int)__builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], (int)word, true);
  out[idx] = word;
}

// (re, im) byte pairs -> (-im, re); a zero keeps a clear sign bit.
__device__ __forceinline__ unsigned int rot_neg8_word(unsigned int w) {
  const unsigned int sw = ((w & 0x00FF00FFu) << 8) | ((w >> 8) & 0x00FF00FFu);
  unsigned int flip = 0u;
  if (sw & 0x0000007Fu) flip |= 0x00000080u;
  if (sw & 0x007F0000u) flip |= 0x00800000u;
  return sw ^ flip;
}

__device__ __forceinline__ i32x8 rot_neg8(i32x8 a) {
  i32x8 r;
#pragma unroll
  for (int i = 0; i < 8; ++i) r[i] = (int)rot_neg8_word((unsigned int)a[i]);
  return r;
}

// FP8 contraction: Z[r][p][m] = (A1 . B, A2 . B) / (gA gB) for one 64 x 64 tile per wave, 128 k'
// per MFMA (16x16x128, e4m3 x e4m3, FP32 accumulation).
// Z16: Z gA gB zs as FP16 pairs (round to nearest) in the zidx16 layout, |e4m3 products| <= 448^2
// bounding |Z gA gB| by KP8 448^2 so that zs can be chosen to exclude overflow.
// Tile: 64 rows x 16 ZGEMM8_TJ columns per wavefront.
#ifndef PFTCC_ZGEMM8_TJ
#define PFTCC_ZGEMM8_TJ 4
#endif
#ifndef PFTCC_ZGEMM8_W
#define PFTCC_ZGEMM8_W 1       // minimum wavefronts per SIMD imposed on the register allocation
#endif
#ifndef PFTCC_ZGEMM8_MT
#define PFTCC_ZGEMM8_MT 8      // consecutive m per wavefront when KP8 = 128 (next operands loaded ahead)
#endif
constexpr int ZGEMM8_TJ = PFTCC_ZGEMM8_TJ;
constexpr int ZGEMM8_MT = PFTCC_ZGEMM8_MT;
template <bool Z16, int MT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PFTCC_ZGEMM8_W)))
k_zgemm8(const i32x8* __restrict__ A, const i32x8* __restrict__ B, int rows_pad, int cols_pad, int KP8, int n,
         float2* __restrict__ Z, const unsigned int* __restrict__ gmaxA, const unsigned int* __restrict__ gmaxB,
         float zs) {
  constexpr int TJ = ZGEMM8_TJ;
  const int l = threadIdx.x;
  const ZTile zt = ztile(cols_pad / (16 * TJ), rows_pad / 64, 16 * TJ, 64);
  const int p0 = zt.p0, r0 = zt.r0;
  const int KV = KP8 / 32;                                 // i32x8 vectors per row
  const int lr = l % 16, lk = l / 16;
  const float inv = Z16 ? zs : 1.0f / (fp8_gain(gmaxA) * fp8_gain(gmaxB));
  auto load_ops = [&](int m, int ks, i32x8 (&a)[4], i32x8 (&b)[TJ]) {
    const size_t abase = ((size_t)m * rows_pad + r0) * KV;
    const size_t bbase = ((size_t)m * cols_pad + p0) * KV;
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = A[abase + (size_t)(i * 16 + lr) * KV + ks + lk];
#pragma unroll
    for (int j = 0; j < TJ; ++j) b[j] = B[bbase + (size_t)(j * 16 + lr) * KV + ks + lk];
  };
  auto mma = [&](const i32x8 (&a)[4], const i32x8 (&b)[TJ], f32x4 (&zr)[4][TJ], f32x4 (&zi)[4][TJ]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const i32x8 a2 = rot_neg8(a[i]);
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
        zr[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a[i], b[j], zr[i][j], 0, 0, 0, 0, 0, 0);
        zi[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a2, b[j], zi[i][j], 0, 0, 0, 0, 0, 0);
      }
    }
  };
  auto zero = [&](f32x4 (&zr)[4][TJ], f32x4 (&zi)[4][TJ]) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
        zr[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        zi[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
      }
  };
  auto store = [&](int m, const f32x4 (&zr)[4][TJ], const f32x4 (&zi)[4][TJ]) {
    if constexpr (Z16) {
      // element (r0 + dr, p0 + dp, m) of zidx16 lies at zidx16(r0, p0, m) + dr RS + (dp / 32) (n + 1) 32
      // + dp % 32 (p0 is a multiple of 32): a wave-uniform base and 32-bit offsets below 64 RS
      static_assert(16 * TJ % ZG16 == 0 || ZG16 % (16 * TJ) == 0, "column tile and FP16 group");
      char* const zb = reinterpret_cast<char*>(reinterpret_cast<unsigned int*>(Z) + zidx16(r0, p0, m, n, cols_pad));
      const unsigned int RS = (unsigned int)cols_pad * (unsigned int)(n + 1);
      const unsigned int lo = 4u * lk * RS + (unsigned int)lr;
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const unsigned int dp = 16u * j;
            const unsigned int cst = (16u * i + e) * RS + (dp / ZG16) * (unsigned int)(n + 1) * ZG16 + dp % ZG16;
            const half2v h = {(_Float16)(zr[i][j][e] * inv), (_Float16)(zi[i][j][e] * inv)};
            *reinterpret_cast<unsigned int*>(zb + ((lo + cst) << 2)) = __builtin_bit_cast(unsigned int, h);
          }
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int r = r0 + i * 16 + 4 * lk + e;
            const int p = p0 + j * 16 + lr;
            Z[zidx(r, p, m, n, cols_pad)] = make_float2(zr[i][j][e] * inv, zi[i][j][e] * inv);
          }
    }
  };
  if constexpr (MT == 1) {
    const int m = zt.m;
    f32x4 zr[4][TJ], zi[4][TJ];
    zero(zr, zi);
    for (int ks = 0; ks < KV; ks += 4) {
      i32x8 a[4], b[TJ];
      load_ops(m, ks, a, b);
      mma(a, b, zr, zi);
    }
    store(m, zr, zi);
  } else {
    // one k' step per m (KP8 = 128, checked by the host); MT consecutive m per wavefront, the next
    // m's operands loaded before the current m's products
    const int m0 = zt.m * MT;
    i32x8 a[4], b[TJ];
    load_ops(m0, 0, a, b);
    for (int mm = 0; mm < MT; ++mm) {
      const int m = m0 + mm;
      if (m > n) break;
      const bool more = mm + 1 < MT && m + 1 <= n;
      i32x8 an[4], bn[TJ];
      if (more) load_ops(m + 1, 0, an, bn);
      f32x4 zr[4][TJ], zi[4][TJ];
      zero(zr, zi);
      mma(a, b, zr, zi);
      store(m, zr, zi);
      if (more) {
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = an[i];
#pragma unroll
        for (int j = 0; j < TJ; ++j) b[j] = bn[j];
      }
    }
  }
}

// Shifted reference rings out[r * ns + s][kk][j] = rings[r][kk][j] * shtab[s][kk][j].
__global__ void k_shift_rings(const float2* __restrict__ rings, const float2* __restrict__ shtab, int nr, int ns,
                              int nk, int n, float2* __restrict__ out) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t ring = (size_t)nk * n;
  if (idx >= (size_t)nr * ns * ring) return;
  const size_t e = idx % ring;
  const size_t row = idx / ring;
  const int r = (int)(row / ns), s = (int)(row % ns);
  out[idx] = cmul(rings[(size_t)r * ring + e], shtab[(size_t)s * ring + e]);
}

// Gathers rows of a [N][nk][len] float2 array by a list into [count][nk][len].
__global__ void k_gather_rows(const float2* __restrict__ src, const int* __restrict__ list, int count, size_t rowlen,
                              float2* __restrict__ dst) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)count * rowlen) return;
  const size_t row = idx / rowlen, e = idx % rowlen;
  dst[idx] = src[(size_t)list[row] * rowlen + e];
}

// Ring mean of |CTF| for listed particles: cbar [count][nk] (float), cbar2 (double).
__global__ void k_ring_mean_ctf_list(const float* __restrict__ ctf, const int* __restrict__ list, int count, int nk,
                                     int n, float* __restrict__ cbar, double* __restrict__ cbar2) {
  const int item = blockIdx.x * blockDim.y + threadIdx.y;   // (p, kk)
  if (item >= count * nk) return;
  const int p = item / nk, kk = item % nk;
  const float* c = ctf + ((size_t)list[p] * nk + kk) * n;
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) s += c[j];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (threadIdx.x == 0) {
    const double m = s / n;
    cbar[item] = (float)m;
    cbar2[item] = m * m;
  }
}

// FP64 sum |p|^2 over shells kk < kcount for listed particles (one wave per particle).
__global__ void k_ring_sq64_list(const float2* __restrict__ rings, const int* __restrict__ list, int count, int nk,
                                 int n, int kcount, double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.y + threadIdx.y;
  if (p >= count) return;
  const float2* v = rings + (size_t)list[p] * nk * n;
  double s = 0.0;
  for (int e = threadIdx.x; e < kcount * n; e += blockDim.x) s += (double)v[e].x * v[e].x + (double)v[e].y * v[e].y;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (threadIdx.x == 0) out[p] = s;
}

// The T largest (particle, reference) keys of each particle, in descending order (0 when fewer
// than T references exist).  One work-group per particle.  Keys are distinct (their low words encode
// distinct (reference, rotation) indices), so the T-th largest is found by a most-significant-digit
// radix select: per 8-bit digit, a histogram of the keys that share the digits chosen so far picks
// the bucket holding the T-th largest; it stops when that bucket is taken whole (at the latest at
// the last digit, where a bucket holds one key).  The selected keys (those at or above the chosen
// prefix) are compacted into LDS and sorted by a bitonic network.  A few passes over the row
// instead of one pass per candidate.
constexpr int TOPT_MAX = 256;
__global__ void __launch_bounds__(256) k_topt(const unsigned long long* __restrict__ prbest, int np, int nrefs, int T,
                                              unsigned long long* __restrict__ cand) {
  __shared__ unsigned int hist[256];
  __shared__ unsigned long long sel[TOPT_MAX];
  __shared__ int s_bin;
  __shared__ unsigned int s_need, s_bucket, s_count;
  const int i = blockIdx.x, tid = threadIdx.x;
  if (i >= np) return;
  const unsigned long long* row = prbest + (size_t)i * nrefs;
  unsigned long long prefix = 0ull, mask = 0ull;
  unsigned int need = (unsigned int)T;
  bool take_all = false;
  for (int d = 7; d >= 0; --d) {
    const int sh = 8 * d;
    hist[tid] = 0u;
    __syncthreads();
    for (int r = tid; r < nrefs; r += 256) {
      const unsigned long long k = row[r];
      if (k != 0ull && (k & mask) == prefix) atomicAdd(&hist[(k >> sh) & 255ull], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      unsigned int above = 0u;
      int b = 255;
      for (; b >= 0; --b) {
        if (above + hist[b] >= need) break;
        above += hist[b];
      }
      s_bin = b;
      s_need = b >= 0 ? need - above : 0u;
      s_bucket = b >= 0 ? hist[b] : 0u;
    }
    __syncthreads();
    if (s_bin < 0) {        // fewer than need keys share the prefix: only at the first digit
      take_all = true;
      break;
    }
    need = s_need;
    prefix |= (unsigned long long)s_bin << sh;
    mask |= 255ull << sh;
    if (s_bucket == need) break;
  }
  if (tid == 0) s_count = 0u;
  __syncthreads();
  for (int r = tid; r < nrefs; r += 256) {
    const unsigned long long k = row[r];
    if (k != 0ull && (take_all || (k & mask) >= prefix)) {
      const unsigned int idx = atomicAdd(&s_count, 1u);
      if (idx < (unsigned int)TOPT_MAX) sel[idx] = k;
    }
  }
  __syncthreads();
  const int cnt = (int)min(s_count, (unsigned int)T);
  int P = 1;
  while (P < cnt) P <<= 1;
  if (tid >= cnt && tid < P) sel[tid] = 0ull;
  __syncthreads();
  // bitonic sort of sel[0, P) into descending order
  for (int kk = 2; kk <= P; kk <<= 1)
    for (int j = kk >> 1; j > 0; j >>= 1) {
      const int x = tid ^ j;
      if (tid < P && x > tid) {
        const unsigned long long a = sel[tid], b = sel[x];
        const bool desc = (tid & kk) == 0;
        if (desc ? a < b : a > b) {
          sel[tid] = b;
          sel[x] = a;
        }
      }
      __syncthreads();
    }
  for (int t = tid; t < T; t += 256) cand[(size_t)i * T + t] = t < cnt ? sel[t] : 0ull;
}

struct RescoreArgs {
  int n, nk, kcorr, T;
  int ns;                      // shifts per reference in the keys (1 when unshifted)
  int knorm;                   // shells of the reference norm
  const double* shifts;        // [ns][2] or null (unshifted)
  const double* argx;          // [nk][n] shift-phase arguments (argtransf)
  const double* argy;
  const float2* refs_even;
  const float2* refs_odd;
  const float2* ptcls;         // [nptcls][nk][n]
  const float* ctf;            // [nptcls][nk][n] or null
  const int8_t* iseven;
  const int* plist;            // [np] local particle index of each output position
  const unsigned long long* cand;   // [np][T] keys (iref * 2n + rho in the low word)
  int np;
  float* best_cc;
  int* best_iref;
  int* best_irot;
  int* best_ishift;            // or null
  double* best_cc64;           // or null
  const int* order;            // [np] output position taken by the t-th work-group (XCD order), or null
};

// Exact rescoring of the candidates in FP64: for each candidate (reference, shift, rho) the
// rotations rho - 1, rho, rho + 1 are evaluated as in gencorrs: q = ref * |CTF| [* e^{i phi}] with
// the particle's own angle-dependent |CTF| and phi = argx sx + argy sy, cc = sum_k<kcorr sum_j
// Re q[k,j] conj(P[k,(j+rho) mod 2n]) / sqrt(sum_k<knorm |q|^2 * sum_k<kcorr |p|^2) (knorm = all
// shells when shifted, gencorrs_cc_2).  The best cc wins, ties to the lowest (reference, shift,
// rotation).
// One work-group per particle, one wavefront per candidate (T > NW candidates take several passes).
// The particle's shell k (ring and |CTF|) is staged in LDS and shared by every wavefront, so a
// candidate's only global reads are its reference ring.  While shell k is computed, the reference
// values of shell k + 1 (JB per lane) and the particle's shell k + 1 are loaded into registers; the
// latter is written to the other LDS buffer after the arithmetic, so one barrier per shell
// separates the buffers' writers and readers.  Lane sums are reduced over the wavefront in a fixed
// order, so a candidate's value does not depend on its slot.  JB = ceil(n / 64) positions per lane.
// SH: the candidates carry shifts (gencorrs_cc_2), whose phases are evaluated per element.
// Processing order of the rescoring: output positions counting-sorted by the key row (reference x
// shift) of their best candidate, the first of k_topt's descending list (row nb - 1 when there is
// none).  The work-groups that run together on one XCD then take particles whose candidates share
// reference rings, which stay in its L2.  The order within a row follows atomics and is not fixed;
// each position's result does not depend on it.
__global__ void k_order_hist(const unsigned long long* __restrict__ cand, int np, int T, int n2, int nb,
                             int* __restrict__ bin, int* __restrict__ hist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const unsigned long long key = cand[(size_t)i * T];
  int b = nb - 1;
  if (key) b = min(nb - 1, (int)((0xFFFFFFFFu - (unsigned int)(key & 0xFFFFFFFFull)) / (unsigned int)n2));
  bin[i] = b;
  atomicAdd(&hist[b], 1);
}

// Exclusive prefix sum of hist[0, nb) in place: one work-group of 1024 threads, each a contiguous
// chunk.
__global__ void __launch_bounds__(1024) k_order_scan(int* __restrict__ hist, int nb) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b0 = min(nb, t * per), b1 = min(nb, b0 + per);
  int s = 0;
  for (int b = b0; b < b1; ++b) s += hist[b];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t ? part[t - 1] : 0;
  for (int b = b0; b < b1; ++b) {
    const int c = hist[b];
    hist[b] = run;
    run += c;
  }
}

__global__ void k_order_scatter(const int* __restrict__ bin, int np, int* __restrict__ offs, int* __restrict__ order) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  order[atomicAdd(&offs[bin[i]], 1)] = i;
}

constexpr int RESCORE_MAX = TOPT_MAX;
constexpr int RESCORE_WAVES = 16;
template <bool SH, int JB>
__global__ void __launch_bounds__(64 * RESCORE_WAVES) k_rescore64(RescoreArgs a) {
  __shared__ double ccs[RESCORE_MAX * 3];          // numerators, then cc
  __shared__ unsigned int lins[RESCORE_MAX * 3];
  __shared__ double sqs[RESCORE_MAX];              // S_q of each candidate
  __shared__ double spw[RESCORE_WAVES];
  extern __shared__ __align__(16) float2 pk[];    // 2 x [n] particle ring, then 2 x [n] |CTF|
  const int n = a.n, n2 = 2 * n;
  float* ck = reinterpret_cast<float*>(pk + 2 * n);
  const int i = a.order ? a.order[xcd_remap(blockIdx.x, gridDim.x)] : blockIdx.x;
  const int nw = blockDim.x >> 6;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ip = a.plist[i];
  const float2* pt = a.ptcls + (size_t)ip * a.nk * n;
  const float* ct = a.ctf ? a.ctf + (size_t)ip * a.nk * n : nullptr;
  const float2* refs = a.iseven[ip] ? a.refs_even : a.refs_odd;
  const int kn = a.knorm;
  double sp = 0.0;                                 // sum |p|^2 over kcorr shells (first pass only)
  for (int c0 = 0; c0 < a.T; c0 += nw) {
    const int c = c0 + w;
    const unsigned long long key = c < a.T ? a.cand[(size_t)i * a.T + c] : 0ull;
    const bool live = key != 0ull;
    const unsigned int lin = live ? 0xFFFFFFFFu - (unsigned int)(key & 0xFFFFFFFFull) : 0u;
    const unsigned int row = lin / (unsigned)n2;
    const int iref = (int)(row / (unsigned)a.ns), ish = (int)(row % (unsigned)a.ns);
    const int rho = (int)(lin % (unsigned)n2);
    const int rr[3] = {(rho + n2 - 1) % n2, rho, (rho + 1) % n2};
    const float2* ref = refs + (size_t)iref * a.nk * n;
    const double sx = SH && live ? a.shifts[2 * ish] : 0.0, sy = SH && live ? a.shifts[2 * ish + 1] : 0.0;
    double num0 = 0.0, num1 = 0.0, num2 = 0.0, sq = 0.0;
    // prologue: shell 0 staged, its reference values in registers
    float2 rcur[JB], rnext[JB];
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
      const float2 v = pt[e];
      pk[e] = v;
      ck[e] = ct ? ct[e] : 1.f;
      if (c0 == 0 && a.kcorr > 0) sp += (double)v.x * v.x + (double)v.y * v.y;
    }
#pragma unroll
    for (int u = 0; u < JB; ++u) {
      const int j = lane + 64 * u;
      rnext[u] = (live && j < n) ? ref[j] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    for (int k = 0; k < kn; ++k) {
      const int b = k & 1;
      const float2* pb = pk + b * n;
      const float* cb = ck + b * n;
#pragma unroll
      for (int u = 0; u < JB; ++u) rcur[u] = rnext[u];
      // loads of shell k + 1: reference values and the particle's staging values
      const bool more = k + 1 < kn;
      float2 pnext = make_float2(0.f, 0.f);
      float cnext = 1.f;
      const int e1 = threadIdx.x;
      if (more) {
#pragma unroll
        for (int u = 0; u < JB; ++u) {
          const int j = lane + 64 * u;
          if (live && j < n) rnext[u] = ref[(size_t)(k + 1) * n + j];
        }
        if (e1 < n) {
          pnext = pt[(size_t)(k + 1) * n + e1];
          if (ct) cnext = ct[(size_t)(k + 1) * n + e1];
        }
      }
      if (live) {
        const bool corr = k < a.kcorr;
#pragma unroll
        for (int u = 0; u < JB; ++u) {
          const int j = lane + 64 * u;
          if (j >= n) continue;
          const double cv = (double)cb[j];
          const float2 rv = rcur[u];
          double qx = rv.x * cv, qy = rv.y * cv;
          if (SH) {
            const size_t e = (size_t)k * n + j;
            double sn, cs;
            sincos(a.argx[e] * sx + a.argy[e] * sy, &sn, &cs);
            const double tx = qx * cs - qy * sn;
            qy = qx * sn + qy * cs;
            qx = tx;
          }
          sq += qx * qx + qy * qy;
          if (!corr) continue;
          double nm[3];
#pragma unroll
          for (int t = 0; t < 3; ++t) {
            int idx = j + rr[t];
            if (idx >= n2) idx -= n2;
            float2 P;
            if (idx < n) {
              P = pb[idx];
            } else {
              P = pb[idx - n];
              P.y = -P.y;
            }
            nm[t] = qx * P.x + qy * P.y;
          }
          num0 += nm[0];
          num1 += nm[1];
          num2 += nm[2];
        }
      }
      if (more && e1 < n) {
        pk[(1 - b) * n + e1] = pnext;
        ck[(1 - b) * n + e1] = cnext;
        if (c0 == 0 && k + 1 < a.kcorr) sp += (double)pnext.x * pnext.x + (double)pnext.y * pnext.y;
      }
      __syncthreads();
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      num0 += __shfl_xor(num0, o);
      num1 += __shfl_xor(num1, o);
      num2 += __shfl_xor(num2, o);
      sq += __shfl_xor(sq, o);
    }
    if (lane == 0 && c < a.T) {
      const double nums[3] = {num0, num1, num2};
      for (int t = 0; t < 3; ++t) {
        ccs[c * 3 + t] = live ? nums[t] : -INFINITY;      // divided by the norms once S_p is known
        lins[c * 3 + t] = live ? row * (unsigned int)n2 + (unsigned int)rr[t] : 0xFFFFFFFFu;
      }
      sqs[c] = sq;
    }
    __syncthreads();                               // the next pass restages shell 0
  }
  // S_p: the staging threads' partial sums, reduced in a fixed order
  for (int o = 32; o > 0; o >>= 1) sp += __shfl_xor(sp, o);
  if (lane == 0) spw[w] = sp;
  __syncthreads();
  if (threadIdx.x == 0) {
    double spt = 0.0;
    for (int v = 0; v < nw; ++v) spt += spw[v];
    double bc = -INFINITY;
    unsigned int bl = 0xFFFFFFFFu;
    for (int c = 0; c < a.T; ++c) {
      if (lins[c * 3] == 0xFFFFFFFFu) continue;
      const double d = sqrt(sqs[c] * spt);
      for (int t = 0; t < 3; ++t) {
        const double v = ccs[c * 3 + t] / d;
        if (v > bc || (v == bc && lins[c * 3 + t] < bl)) {
          bc = v;
          bl = lins[c * 3 + t];
        }
      }
    }
    if (bl == 0xFFFFFFFFu) {
      a.best_cc[i] = __uint_as_float(0x7FC00000u);
      a.best_iref[i] = 0;
      a.best_irot[i] = 0;
      if (a.best_ishift) a.best_ishift[i] = 0;
    } else {
      const unsigned int rw = bl / (unsigned)n2;
      a.best_cc[i] = (float)bc;
      a.best_iref[i] = (int)(rw / (unsigned)a.ns) + 1;
      a.best_irot[i] = (int)(bl % (unsigned)n2) + 1;
      if (a.best_ishift) a.best_ishift[i] = (int)(rw % (unsigned)a.ns);
    }
    if (a.best_cc64) a.best_cc64[i] = bc;
  }
}

// Decodes keys whose low word encodes iref * nrots + rho.
__global__ void k_best_decode_rot(const unsigned long long* __restrict__ best, int np, int nrots,
                                  float* __restrict__ cc, int* __restrict__ iref, int* __restrict__ irot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const unsigned long long k = best[i];
  if (k == 0ull) {
    cc[i] = __uint_as_float(0x7FC00000u);
    iref[i] = 0;
    irot[i] = 0;
    return;
  }
  const unsigned int hi = (unsigned int)(k >> 32);
  const unsigned int bits = (hi & 0x80000000u) ? (hi & 0x7FFFFFFFu) : ~hi;
  const unsigned int lin = 0xFFFFFFFFu - (unsigned int)(k & 0xFFFFFFFFull);
  cc[i] = __uint_as_float(bits);
  iref[i] = (int)(lin / (unsigned)nrots) + 1;
  irot[i] = (int)(lin % (unsigned)nrots) + 1;
}

}  // namespace pfk
