// pftcc_kernels.hpp - gfx950 kernels of the polar-FT correlation search.
//
// Data layout in HBM (C order, one GPU):
//   refs_even/odd  float2 [nrefs][nk][n]      reference rings (pfts_refs_even/odd, :83-84)
//   ptcls          float2 [nptcls][nk][n]     particle rings  (pfts_ptcls, :87)
//   ctf            float  [nptcls][nk][n]     |CTF| in the reference frame (ctfmats, :77)
//   memo           float2 [nptcls][nk][n+1]   real FFT of length 2n of w = Re P + Im P over the
//                                             full ring (replaces the fftdat_ptcls memo, :89)
//   sqsums         float  [nptcls]            sum |p|^2 over kfrom..kstop (sqsums_ptcls, :74)
//   sqsums_k       float  [nptcls][nk]        per-shell sum |p|^2
// with n = pftsz, nk = kto - kfrom + 1 and k the shell index kfrom + kk.
//
// The rotational correlation of one (reference, particle[, shift]) pair is computed as in
// calc_corrs_over_k (:1132-1162) but with one real FFT of length 2n per shell instead of the
// reference's r2c_n + c2c_n pair:
//   q = ref * |CTF| [* shift phase]; v[t] = Re q[t] + Im q[t], v[t+n] = Re q[t] - Im q[t];
//   V = rfft_2n(v) (bins 0..n);  Z += conj(V) * W_k;  cc = irfft_2n(Z) / (4n) / sqrt(Sq Sp).
// Re Q is n-periodic and Im Q n-antiperiodic, so V holds the reference's 2A (even bins) and
// 2B (odd bins) and conj(V)W equals the reference's 4 conj(A)PA / 4 conj(B)PB product.
// rfft_2n is one complex FFT of length n of z[u] = v[2u] + i v[2u+1] plus the usual split.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_lds.hpp"

namespace pfk {

struct KCtfConst {      // per-particle constants in the reference's FP32 order (simple_ctf.f90)
  float wl;             // kV2wl()/smpd                          (:54-70, 522-525)
  float cs;             // Cs*1e7/smpd
  float ampc;           // atan(fraca/sqrt(1-fraca^2))
  float dfx, dfy;       // df*1e4/smpd                            (:73-81)
  float angast;         // deg2rad(angast)
  float phshift;        // add_phshift (0 unless phase plate)
  int32_t active;       // pinds(iptcl) > 0
};

// --------------------------------------------------------------------------- small kernels ---
__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u32(unsigned int* p, int n, unsigned int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// memoize_sqsum_ptcl (:841-845) and per-shell sums; one block per particle.
__global__ void k_sqsums(const float2* __restrict__ ptcls, int nk, int n, int kcount,
                         float* __restrict__ sqsums, float* __restrict__ sqsums_k, int first, int count,
                         const int* __restrict__ list) {
  if (blockIdx.x >= count) return;
  const int ip = list ? list[blockIdx.x] : first + blockIdx.x;
  __shared__ double red[256];
  double tot = 0.0;
  for (int kk = 0; kk < nk; ++kk) {
    const float2* p = ptcls + ((size_t)ip * nk + kk) * n;
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      float2 v = p[j];
      s += (double)v.x * v.x + (double)v.y * v.y;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (sqsums_k) sqsums_k[(size_t)ip * nk + kk] = (float)red[0];
      if (kk < kcount) tot += red[0];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) sqsums[ip] = (float)tot;
}

// create_polar_absctfmats (:906-959): |ctf%eval(spaFreqSq, ang[, phshift])| evaluated in FP32 in
// the reference's operation order; spafreq and angmat are the host-built FP32 tables of :927-936.
__global__ void k_ctfmats(const KCtfConst* __restrict__ cst, const float* __restrict__ spafreq,
                          const float* __restrict__ angmat, int nk, int n, int nptcls,
                          float* __restrict__ ctf) {
  const size_t per = (size_t)nk * n;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= per * nptcls) return;
  const int ip = (int)(idx / per);
  const size_t e = idx - (size_t)ip * per;
  const KCtfConst c = cst[ip];
  if (!c.active) return;
  const float s2 = spafreq[e];
  const float ang = angmat[e];
  // eval_df (:185-189)
  const float df = 0.5f * (c.dfx + c.dfy + cosf(2.0f * (ang - c.angast)) * (c.dfx - c.dfy));
  // evalPhSh (:88-98): PI*wl*spaFreqSq*(df - 0.5*wl*wl*spaFreqSq*Cs) + add_phshift
  const float pi32 = 3.14159274101257324f;
  const float arg = pi32 * c.wl * s2 * (df - 0.5f * c.wl * c.wl * s2 * c.cs) + c.phshift;
  ctf[idx] = fabsf(sinf(arg + c.ampc));
}

// gen_shmat_8 (:1098-1120) cast to single precision as gen_shmat does (:1123-1130).
// One thread per (shift, j): first shell analytic, then de Moivre over the shells.
__global__ void k_shift_table(const double* __restrict__ shifts /*[ns][2]*/, int ns,
                              const double* __restrict__ argx /*[nk][n]*/,
                              const double* __restrict__ argy, const double* __restrict__ one_x,
                              const double* __restrict__ one_y, int nk, int n,
                              float2* __restrict__ out /*[ns][nk][n]*/) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ns * n) return;
  const int s = idx / n;
  const int j = idx - s * n;
  const double sx = shifts[2 * s], sy = shifts[2 * s + 1];
  double sa, ca, ss, cs;
  sincos(argx[j] * sx + argy[j] * sy, &sa, &ca);
  sincos(one_x[j] * sx + one_y[j] * sy, &ss, &cs);
  double re = ca, im = sa;
  for (int kk = 0; kk < nk; ++kk) {
    out[((size_t)s * nk + kk) * n + j] = make_float2((float)re, (float)im);
    const double r2 = re * cs - im * ss;
    const double i2 = re * ss + im * cs;
    re = r2;
    im = i2;
  }
}

// Per (eo, ref, shell) 1/sqrt(mean |r_k|^2) used by shellnorm_and_filter_ref(_8) (:756-805);
// out[(eo * nrefs + iref) * kcount + kk] = mean power (double).
__global__ void k_ref_power(const float2* __restrict__ refs_even, const float2* __restrict__ refs_odd,
                            int nrefs, int nk, int n, int kcount, double* __restrict__ pw) {
  const int item = blockIdx.x;   // (eo, iref, kk)
  const int kk = item % kcount;
  const int iref = (item / kcount) % nrefs;
  const int eo = item / (kcount * nrefs);
  const float2* r = (eo ? refs_odd : refs_even) + ((size_t)iref * nk + kk) * n;
  __shared__ double red[256];
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    float2 v = r[j];
    s += (double)v.x * v.x + (double)v.y * v.y;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) pw[item] = red[0] / (double)n;
}

// ------------------------------------------------------------------ ring -> packed FFT input ---
// For ring values a (at t) and b (at t+1), t = 2u even, writes z[u] and z[u + n/2] of the
// packed length-n complex input whose FFT yields rfft_2n of v = (Re + Im | Re - Im).
__device__ __forceinline__ void pack_pair(float2* __restrict__ z, int u, int half, float2 a, float2 b) {
  z[u] = make_float2(a.x + a.y, b.x + b.y);
  z[u + half] = make_float2(a.x - a.y, b.x - b.y);
}

// Split of the packed FFT Zc (length n) into the real FFT V of length 2n, bin m in [0, n].
__device__ __forceinline__ float2 rfft_split(const float2* __restrict__ zc, int n, int m,
                                             const float2* __restrict__ tw2) {
  const float2 A = zc[m == n ? 0 : m];
  const float2 B = zc[(n - m) % n];
  const float2 sum = make_float2(0.5f * (A.x + B.x), 0.5f * (A.y - B.y));   // (A + conj B)/2
  const float2 dif = make_float2(0.5f * (A.x - B.x), 0.5f * (A.y + B.y));   // (A - conj B)/2
  const float2 w = tw2[m];                                                  // exp(-i pi m / n)
  const float2 wd = cmul(w, dif);
  // V = sum - i * wd
  return make_float2(sum.x + wd.y, sum.y - wd.x);
}

// --------------------------------------------------------------------- particle memoization ---
// memoize_ffts (:848-877): W[ip][kk][m] = rfft_2n(w)[m], w = (Re p + Im p | Re p - Im p).
// One block handles F (particle, shell) items.  Dynamic LDS: 2*F*n + 2n float2.
__global__ void k_memoize(const float2* __restrict__ ptcls, int first, int count, int nk, int n,
                          FftPlan plan, const float2* __restrict__ tw2g, float2* __restrict__ memo, int F,
                          const int* __restrict__ list) {
  extern __shared__ float2 smem[];
  float2* bufA = smem;
  float2* bufB = bufA + (size_t)F * n;
  float2* tw2 = bufB + (size_t)F * n;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int i = tid; i < 2 * n; i += nt) tw2[i] = tw2g[i];
  const int nitems = count * nk;
  const int item0 = blockIdx.x * F;
  const int half = n / 2;
  for (int idx = tid; idx < F * half; idx += nt) {
    const int f = idx / half, u = idx - f * half;
    const int item = item0 + f;
    float2 a = make_float2(0.f, 0.f), b = a;
    if (item < nitems) {
      const int ip = (list ? list[item / nk] : first + item / nk), kk = item % nk;
      const float2* p = ptcls + ((size_t)ip * nk + kk) * n;
      a = p[2 * u];
      b = p[2 * u + 1];
    }
    pack_pair(bufA + (size_t)f * n, u, half, a, b);
  }
  __syncthreads();
  float2* res = fft_forward_oop(bufA, bufB, F, n, n, plan, tw2, tid, nt);
  for (int idx = tid; idx < F * (n + 1); idx += nt) {
    const int f = idx / (n + 1), m = idx - f * (n + 1);
    const int item = item0 + f;
    if (item >= nitems) continue;
    const int ip = (list ? list[item / nk] : first + item / nk), kk = item % nk;
    memo[((size_t)ip * nk + kk) * (n + 1) + m] = rfft_split(res + (size_t)f * n, n, m, tw2);
  }
}

// --------------------------------------------------------- the rotational correlation kernel ---
struct CorrArgs {
  int n, nk, kcorr, knorm, nrefs;
  int mode;                    // 0: CC (normalised), 1: Euclid (gencorrs_euclid)
  FftPlan plan;
  const float2* refs_even;
  const float2* refs_odd;
  const float2* memo;
  const float* ctf;            // null when ctf = no
  const float* sqsums;         // S_p over kfrom..kstop
  const float* sqsums_k;       // per shell S_p (Euclid)
  const int8_t* iseven;
  const float2* shtab;         // [ns][nk][n] or null (unshifted)
  int ns;
  // shellnorm_and_filter_ref (:756-781): q_k *= ofac[row][kk] * refinv[eo][iref][kk], row = iref,
  // or the particle's optlp row when l_clsfrcs (optrow[ip]).  Null refinv disables it.
  const float* refinv;         // [2][nrefs][kcorr]  1/sqrt(mean |r_k|^2) (1 if power <= 1e-12)
  const float* ofac;           // [nrows][kcorr]     optlp (match_filt) or sqrt(optlp*k) (pssnr)
  const int* optrow;           // [nptcls] or null
  const float* shellw;         // [nptcls][nk] Euclid weights pxls(k)/n/(2 sigma2), null for CC
  const int* plist;            // local particle indices (0-based)
  int np;
  const int* rlist;            // 0-based reference indices or null (identity)
  int nr;
  int rlist_ld;                // 0: one list for all particles; else particle slot pi uses rlist[pi * rlist_ld ..]
  float* cc_out;               // [np][nr][ns][2n] or null
  unsigned long long* best;    // [np] packed (cc, -lin) or null
  const float2* tw2g;          // exp(-i pi a/n), a in [0, 2n)
  int F;
  int ntiles;
  // item mode (item_shift != null): block b correlates particle plist[b] with reference rlist[b]
  // at its own FP32 shift item_shift[b] (F = 1, one pair per block); the shift phases follow
  // gen_shmat_8 (:1098-1120) from argtransf of the first shell (argx/argy row 0) and of shell
  // one (onex/oney), in FP64, rounded to FP32 as gen_shmat does.
  const float* item_shift;     // [np][2]
  const double* argx;          // [nk][n]
  const double* argy;
  const double* onex;          // [n]
  const double* oney;
};

__device__ __forceinline__ unsigned int ordered_bits(float v) {
  unsigned int b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// One block: particle plist[blockIdx.x / ntiles] against F consecutive (ref, shift) pairs.
// Dynamic LDS (float2 units): bufA[F n] bufB[F n] acc[F (n+1)] Wk[n+1] tw2[2n], then floats
// ctfk[n] sqacc[F n/2] offacc[F n/2].  Every sum has a fixed order (no atomics), so a pair's
// result does not depend on F or on the launch it belongs to.
__global__ void __launch_bounds__(256) k_corr_generic(CorrArgs a) {
  extern __shared__ float2 smem[];
  const int n = a.n, F = a.F, half = n / 2;
  float2* bufA = smem;
  float2* bufB = bufA + (size_t)F * n;
  float2* acc = bufB + (size_t)F * n;
  float2* Wk = acc + (size_t)F * (n + 1);
  float2* tw2 = Wk + (n + 1);
  float* ctfk = reinterpret_cast<float*>(tw2 + 2 * n);
  float* sqacc = ctfk + n;                  // [F][half] per-slot sum |q|^2 over shells
  float* offacc = sqacc + (size_t)F * half; // [F][half] Euclid: sum_k w_k |q_k|^2
  __shared__ float sq[64], off[64];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int pi = blockIdx.x / a.ntiles;
  const int tile = blockIdx.x - pi * a.ntiles;
  const int ip = a.plist[pi];
  const bool even = a.iseven[ip] != 0;
  const float2* refs = even ? a.refs_even : a.refs_odd;
  const int npairs = a.nr * a.ns;
  const bool items = a.item_shift != nullptr;
  // item mode: running shift phase of each angle (current shell) and its shell-to-shell step
  // (after offacc, 16-byte aligned; present only in item mode)
  float* const fbase = reinterpret_cast<float*>(smem);
  const size_t shoff = ((size_t)(offacc + (size_t)F * half - fbase) + 3) & ~(size_t)3;
  double2* shph = reinterpret_cast<double2*>(fbase + shoff);
  double2* shst = shph + n;
  if (items) {
    const double sx = (double)a.item_shift[2 * pi], sy = (double)a.item_shift[2 * pi + 1];
    for (int j = tid; j < n; j += nt) {
      double sn, cs;
      sincos(a.argx[j] * sx + a.argy[j] * sy, &sn, &cs);
      shph[j] = make_double2(cs, sn);
      sincos(a.onex[j] * sx + a.oney[j] * sy, &sn, &cs);
      shst[j] = make_double2(cs, sn);
    }
  }

  for (int i = tid; i < 2 * n; i += nt) tw2[i] = a.tw2g[i];
  for (int i = tid; i < F * (n + 1); i += nt) acc[i] = make_float2(0.f, 0.f);
  for (int i = tid; i < F * half; i += nt) { sqacc[i] = 0.f; offacc[i] = 0.f; }

  const int kloop = a.kcorr > a.knorm ? a.kcorr : a.knorm;
  for (int kk = 0; kk < kloop; ++kk) {
    __syncthreads();
    const bool docorr = kk < a.kcorr;
    const bool donorm = kk < a.knorm;
    for (int j = tid; j < n; j += nt)
      ctfk[j] = a.ctf ? a.ctf[((size_t)ip * a.nk + kk) * n + j] : 1.0f;
    if (docorr)
      for (int m = tid; m <= n; m += nt) Wk[m] = a.memo[((size_t)ip * a.nk + kk) * (n + 1) + m];
    const float wk = (a.mode == 1 && docorr) ? a.shellw[(size_t)ip * a.nk + kk] : 1.0f;
    __syncthreads();
    for (int idx = tid; idx < F * half; idx += nt) {
      const int f = idx / half, u = idx - f * half;
      const int P = tile * F + f;
      float2 q0 = make_float2(0.f, 0.f), q1 = q0;
      if (P < npairs) {
        const int ir = P / a.ns, is = P - (P / a.ns) * a.ns;
        const int iref = items ? a.rlist[pi] : (a.rlist ? a.rlist[(size_t)pi * a.rlist_ld + ir] : ir);
        const float2* r = refs + ((size_t)iref * a.nk + kk) * n;
        const int t = 2 * u;
        float2 r0 = r[t], r1 = r[t + 1];
        q0 = cscale(r0, ctfk[t]);
        q1 = cscale(r1, ctfk[t + 1]);
        if (items) {
          // this thread owns angles t, t+1 for every shell: use, then advance, their phases
          const double2 p0 = shph[t], p1 = shph[t + 1];
          q0 = cmul(q0, make_float2((float)p0.x, (float)p0.y));
          q1 = cmul(q1, make_float2((float)p1.x, (float)p1.y));
          const double2 s0 = shst[t], s1 = shst[t + 1];
          shph[t] = make_double2(p0.x * s0.x - p0.y * s0.y, p0.x * s0.y + p0.y * s0.x);
          shph[t + 1] = make_double2(p1.x * s1.x - p1.y * s1.y, p1.x * s1.y + p1.y * s1.x);
        } else if (a.shtab) {
          const float2* sh = a.shtab + ((size_t)is * a.nk + kk) * n;
          q0 = cmul(q0, sh[t]);
          q1 = cmul(q1, sh[t + 1]);
        }
        if (a.refinv && docorr) {
          const int row = a.optrow ? a.optrow[ip] : iref;
          const float s = a.ofac[(size_t)row * a.kcorr + kk] *
                          a.refinv[((size_t)(even ? 0 : 1) * a.nrefs + iref) * a.kcorr + kk];
          q0 = cscale(q0, s);
          q1 = cscale(q1, s);
        }
        const float e2 = q0.x * q0.x + q0.y * q0.y + q1.x * q1.x + q1.y * q1.y;
        if (donorm) sqacc[idx] += e2;                       // slot idx = f*half + u, one owner
        if (a.mode == 1 && docorr) offacc[idx] = fmaf(wk, e2, offacc[idx]);
      }
      if (docorr) pack_pair(bufA + (size_t)f * n, u, half, q0, q1);
    }
    if (!docorr) continue;
    __syncthreads();
    float2* res = fft_forward_oop(bufA, bufB, F, n, n, a.plan, tw2, tid, nt);
    for (int idx = tid; idx < F * (n + 1); idx += nt) {
      const int f = idx / (n + 1), m = idx - f * (n + 1);
      const float2 V = rfft_split(res + (size_t)f * n, n, m, tw2);
      const float2 prod = cmulcj(V, Wk[m]);
      float2& z = acc[idx];
      z.x = fmaf(wk, prod.x, z.x);
      z.y = fmaf(wk, prod.y, z.y);
    }
  }
  __syncthreads();
  if (tid < F) {
    float s0 = 0.f, s1 = 0.f;
    for (int u = 0; u < half; ++u) {
      s0 += sqacc[tid * half + u];
      s1 += offacc[tid * half + u];
    }
    sq[tid] = s0;
    off[tid] = s1;
  }
  // inverse real FFT of length 2n from Z[0..n]: y = IDFT_n(Y), x[2j] = Re y, x[2j+1] = Im y,
  // Y[m] = X[m] + conj(X[n-m]) + i e^{i pi m/n} (X[m] - conj(X[n-m])), via conj(FFT(conj(Y))).
  for (int idx = tid; idx < F * n; idx += nt) {
    const int f = idx / n, m = idx - f * n;
    float2 X = acc[(size_t)f * (n + 1) + m];
    float2 Xn = acc[(size_t)f * (n + 1) + (n - m)];
    if (m == 0) { X.y = 0.f; Xn.y = 0.f; }
    const float2 s = make_float2(X.x + Xn.x, X.y - Xn.y);
    const float2 d = make_float2(X.x - Xn.x, X.y + Xn.y);
    const float2 e = cconj(tw2[m]);                 // exp(+i pi m/n)
    const float2 ed = cmul(e, d);
    const float2 Y = make_float2(s.x - ed.y, s.y + ed.x);   // s + i*ed
    bufA[(size_t)f * n + m] = cconj(Y);
  }
  __syncthreads();
  float2* res = fft_forward_oop(bufA, bufB, F, n, n, a.plan, tw2, tid, nt);
  // per-pair scale
  __shared__ float pscal;
  if (tid == 0) {
    if (a.mode == 1) {
      float s = 0.f;
      for (int kk = 0; kk < a.kcorr; ++kk)
        s += a.shellw[(size_t)ip * a.nk + kk] * a.sqsums_k[(size_t)ip * a.nk + kk];
      pscal = s;
    } else {
      pscal = a.sqsums[ip];
    }
  }
  __syncthreads();
  const float inv4n = 1.0f / (4.0f * (float)n);
  unsigned long long bestkey = 0ull;
  for (int idx = tid; idx < F * n; idx += nt) {
    const int f = idx / n, j = idx - f * n;
    const int P = tile * F + f;
    if (P >= npairs) continue;
    const float2 y = res[(size_t)f * n + j];
    float x0 = y.x, x1 = -y.y;
    if (a.mode == 1) {
      const float o = off[f] + pscal;
      x0 = 2.f * x0 * inv4n - o;
      x1 = 2.f * x1 * inv4n - o;
    } else {
      const float sc = inv4n / sqrtf(sq[f] * pscal);
      x0 *= sc;
      x1 *= sc;
    }
    if (a.cc_out) {
      float2* o = reinterpret_cast<float2*>(a.cc_out + ((size_t)pi * npairs + P) * (2 * n));
      o[j] = make_float2(x0, x1);
    }
    if (a.best) {
      const unsigned int lin0 = (unsigned int)P * (unsigned int)(2 * n) + 2u * j;
      const unsigned long long k0 = ((unsigned long long)ordered_bits(x0) << 32) | (0xFFFFFFFFu - lin0);
      const unsigned long long k1 = ((unsigned long long)ordered_bits(x1) << 32) | (0xFFFFFFFFu - lin0 - 1u);
      bestkey = k0 > bestkey ? k0 : bestkey;
      bestkey = k1 > bestkey ? k1 : bestkey;
    }
  }
  if (a.best) {
    __shared__ unsigned long long red[256];
    red[tid] = bestkey;
    __syncthreads();
    for (int w = nt / 2; w > 0; w >>= 1) {
      if (tid < w && red[tid + w] > red[tid]) red[tid] = red[tid + w];
      __syncthreads();
    }
    if (tid == 0) atomicMax(&a.best[pi], red[0]);
  }
}

// Decodes the packed best keys of k_corr_* into (cc, iref, ishift, irot).
__global__ void k_best_decode(const unsigned long long* __restrict__ best, int np, int nr, int ns, int nrots,
                              const int* __restrict__ rlist, float* __restrict__ cc, int* __restrict__ iref,
                              int* __restrict__ ishift, int* __restrict__ irot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const unsigned long long k = best[i];
  const unsigned int hi = (unsigned int)(k >> 32);
  const unsigned int bits = (hi & 0x80000000u) ? (hi & 0x7FFFFFFFu) : ~hi;
  const unsigned int lin = 0xFFFFFFFFu - (unsigned int)(k & 0xFFFFFFFFull);
  const unsigned int rho = lin % (unsigned)nrots;
  const unsigned int P = lin / (unsigned)nrots;
  const unsigned int ir = P / (unsigned)ns, is = P % (unsigned)ns;
  if (k == 0ull || ir >= (unsigned)nr) {   // never written: no evaluation reached this particle
    cc[i] = __uint_as_float(0x7FC00000u);
    iref[i] = 0;
    ishift[i] = 0;
    irot[i] = 0;
    return;
  }
  cc[i] = __uint_as_float(bits);
  iref[i] = (rlist ? rlist[ir] : (int)ir) + 1;
  ishift[i] = (int)is;
  irot[i] = (int)rho + 1;
}

// ------------------------------------------------- single-rotation FP64 objective and gradient ---
struct Rot8Args {
  int n, nk, kcorr, kfrom, nrefs;
  int mode;                    // 0: CC k-weighted (match_filt = no); 1: CC match_filt; 2: Euclid
  const float2* refs_even;
  const float2* refs_odd;
  const float2* ptcls;
  const float2* dptcls;        // angular derivative rings [nptcls][nk][n] or null
  const float* ctf;
  const float* sqsums;         // memoised S_p (match_filt mode)
  const int8_t* iseven;
  const double* refinv;        // as CorrArgs, in double (shellnorm_and_filter_ref_8, :783-805)
  const double* ofac;
  const int* optrow;
  const float* sn32;           // [n] sin(theta_j), FP32: argtransf(j, k) = (double)(sn32[j] * k) * arg_a
  const float* ncn32;          // [n] -cos(theta_j), FP32, for the y component
  double arg_a;                // pi / (ldim / 2)
  const double* sqp8;          // [nptcls][nk] FP64 shell sums of |P|^2 (k-weighted CC)
  const double* argx;          // [nk][n]; the first shell's row starts the phase recurrence
  const double* argy;
  const double* one_x;         // [n]
  const double* one_y;
  const double* sigma2w;       // Euclid: [nptcls][nk] pxls(k)/n/sigma2(k,iptcl)
  const int* irefs;            // 0-based
  const int* ips;              // local 0-based
  const double* shifts;        // [npairs][2]
  const int* rhos;             // 0-based rotation
  int npairs;
  double* f;                   // [npairs]
  double* grad;                // [npairs][3] or null
};

// One block of 256 threads per (ref, particle, shift, rotation) item; threads own angles j and
// run the shells in order so the shift phase follows gen_shmat_8's recurrence (:1098-1120).
// Work-groups are dispatched round-robin over the 8 XCDs (each with its own L2): block b runs on
// XCD b % 8.  This gives each XCD a contiguous range of logical items, so that consecutive items
// (e.g. the pairs of one particle) share an L2.  A bijection of [0, nblocks).
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  constexpr int NX = 8;
  const int per = nblocks / NX, rem = nblocks % NX;
  const int x = b % NX, q = b / NX;
  return x * per + (x < rem ? x : rem) + q;
}

// The operands of one shell at ring position j, loaded ahead of their use (raw FP32 values, so
// that several shells fit in registers).
struct Rot8Shell {
  float2 rv, pv, dv;           // reference q_k[j], particle P_k[t], dP_k[t]
  float cf;                    // |CTF|
  double fac;                  // filter factors (match_filt / pssnr; uniform over the work-group)
  double s2w;                  // Euclid shell weight (uniform)
};

template <int MODE, bool NORM>
__device__ __forceinline__ void rot8_load(const Rot8Args a, Rot8Shell& s, int kk, int j, int t, int ip, int iref,
                                          bool even, const float2* ref, const float2* ptc, const float2* dpt,
                                          const float* ctf) {
  const int n = a.n;
  s.rv = ref[(size_t)kk * n + j];
  s.pv = ptc[(size_t)kk * n + t];
  s.dv = dpt ? dpt[(size_t)kk * n + t] : make_float2(0.f, 0.f);
  s.cf = ctf ? ctf[(size_t)kk * n + j] : 1.f;
  s.fac = 1.0;
  if (NORM) {
    const int row = a.optrow ? a.optrow[ip] : iref;
    s.fac = a.ofac[(size_t)row * a.kcorr + kk] * a.refinv[((size_t)(even ? 0 : 1) * a.nrefs + iref) * a.kcorr + kk];
  }
  s.s2w = MODE == 2 ? a.sigma2w[(size_t)ip * a.nk + kk] : 0.0;
}

// The FP64 sums of one thread.
struct Rot8Acc {
  double num = 0, gx = 0, gy = 0, gpsi = 0, sqref = 0, sqptcl = 0;
};

// One shell's contributions; advances the shift phase to the next shell (gen_shmat_8's recurrence).
template <int MODE, bool NORM>
__device__ __forceinline__ void rot8_step(const Rot8Args a, const Rot8Shell& s, int kk, bool cj, bool angle,
                                          float snj, float ncnj, Rot8Acc& acc, double& shr, double& shi, double cs,
                                          double ss) {
  const double w = (MODE == 0 || MODE == 2) ? (double)(a.kfrom + kk) : 1.0;
  double c = (double)s.cf;
  if (NORM) c *= s.fac;
  // argtransf (:311-317) from the FP32 polar coordinate k sin(theta_j), -k cos(theta_j)
  const float kf = (float)(a.kfrom + kk);
  const double ax = (double)(snj * kf) * a.arg_a;
  const double ay = (double)(ncnj * kf) * a.arg_a;
  // q = ref * ctf * shmat   ((pft_ref*ctfmats)*shmat, :1709)
  const double ar = (double)s.rv.x * c, ai = (double)s.rv.y * c;
  const double qr = ar * shr - ai * shi;
  const double qi = ar * shi + ai * shr;
  const double pr = s.pv.x, pim = cj ? -(double)s.pv.y : (double)s.pv.y;
  if (MODE == 2) {
    const double dr = qr - pr, di = qi - pim;
    acc.num -= w * 0.5 * (dr * dr + di * di) * s.s2w;
    // corrk(i*a*q) - Re sum (i a q) conj(q);  Re((i a q) conj P) = -a Im(q conj P)
    const double imqp = qi * pr - qr * pim;
    acc.gx += w * (-ax * imqp) * s.s2w;
    acc.gy += w * (-ay * imqp) * s.s2w;
  } else {
    const double re = qr * pr + qi * pim;            // Re(q conj P)
    const double im = qi * pr - qr * pim;            // Im(q conj P)
    acc.num += w * re;
    acc.gx += w * (-ax * im);
    acc.gy += w * (-ay * im);
    if (angle) {
      const double dr = s.dv.x, di = cj ? -(double)s.dv.y : (double)s.dv.y;
      acc.gpsi += w * (qr * dr + qi * di);
    }
    acc.sqref += w * (qr * qr + qi * qi);
  }
  const double r2 = shr * cs - shi * ss;
  const double i2 = shr * ss + shi * cs;
  shr = r2;
  shi = i2;
}

#ifndef PFTCC_ROT8_AHEAD
#define PFTCC_ROT8_AHEAD 1     // shells of loads in flight ahead of the arithmetic in k_rot8
#endif
#ifndef PFTCC_ROT8_W
#define PFTCC_ROT8_W 7         // wavefronts per SIMD the register allocation of k_rot8 must allow
#endif

// FP64 sums of |P|^2 per (particle, shell): one wavefront per ring.
__global__ void __launch_bounds__(256) k_shell_sq8(const float2* __restrict__ rings, int nitems, int n,
                                                   double* __restrict__ out) {
  const int it = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (it >= nitems) return;
  const float2* v = rings + (size_t)it * n;
  double s = 0.0;
  for (int j = threadIdx.x & 63; j < n; j += 64) s += (double)v[j].x * v[j].x + (double)v[j].y * v[j].y;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) out[it] = s;
}

// Threads own angles j and run the shells in order.  The loads of shell k + D are issued before
// the arithmetic of shell k (a software pipeline over a ring of D + 1 shells), so a thread always
// has loads in flight.  MODE: 0 CC k-weighted, 1 CC match_filt, 2 Euclid (equal to a.mode; a
// template argument so that each form's sums stay in registers).  NORM: shell normalisation
// (match_filt / pssnr filter factors per shell), so that the other forms carry no factor.
template <int D, int MODE, bool NORM>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(PFTCC_ROT8_W))) k_rot8(Rot8Args a) {
  const int item = xcd_remap(blockIdx.x, gridDim.x);
  if (item >= a.npairs) return;
  const int n = a.n;
  const int iref = a.irefs[item];
  const int ip = a.ips[item];
  const int rho = a.rhos[item];
  const double sx = a.shifts[2 * item], sy = a.shifts[2 * item + 1];
  const bool even = a.iseven[ip] != 0;
  const float2* ref = (even ? a.refs_even : a.refs_odd) + (size_t)iref * a.nk * n;
  const float2* ptc = a.ptcls + (size_t)ip * a.nk * n;
  const float2* dpt = a.dptcls ? a.dptcls + (size_t)ip * a.nk * n : nullptr;
  const float* ctf = a.ctf ? a.ctf + (size_t)ip * a.nk * n : nullptr;
  const bool angle = dpt != nullptr;
  Rot8Acc acc;
  // k-weighted particle norm from the per-shell FP64 sums, one shell per thread
  if (MODE == 0)
    for (int kk = threadIdx.x; kk < a.kcorr; kk += blockDim.x)
      acc.sqptcl += (double)(a.kfrom + kk) * a.sqp8[(size_t)ip * a.nk + kk];
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const float snj = a.sn32[j], ncnj = a.ncn32[j];
    double sa, ca, ss, cs;
    sincos(a.argx[j] * sx + a.argy[j] * sy, &sa, &ca);
    sincos(a.one_x[j] * sx + a.one_y[j] * sy, &ss, &cs);
    double shr = ca, shi = sa;
    int t = j + rho;           // index of P on the full ring (mod 2n)
    bool cj = false;
    if (t >= 2 * n) t -= 2 * n;
    if (t >= n) { t -= n; cj = true; }
    // a ring of D + 1 shells: shell m lives in slot m % (D + 1); the loads of shell m + D are
    // issued before the arithmetic of shell m
    const int kc = a.kcorr;
    Rot8Shell s[D + 1];
#pragma unroll
    for (int u = 0; u < D; ++u)
      if (u < kc) rot8_load<MODE, NORM>(a, s[u], u, j, t, ip, iref, even, ref, ptc, dpt, ctf);
    int kk = 0;
    for (; kk + D + 1 <= kc; kk += D + 1) {
#pragma unroll
      for (int u = 0; u <= D; ++u) {
        if (kk + u + D < kc) rot8_load<MODE, NORM>(a, s[(u + D) % (D + 1)], kk + u + D, j, t, ip, iref, even, ref, ptc, dpt, ctf);
        rot8_step<MODE, NORM>(a, s[u], kk + u, cj, angle, snj, ncnj, acc, shr, shi, cs, ss);
      }
    }
#pragma unroll
    for (int u = 0; u <= D; ++u)
      if (kk + u < kc) {
        if (kk + u + D < kc) rot8_load<MODE, NORM>(a, s[(u + D) % (D + 1)], kk + u + D, j, t, ip, iref, even, ref, ptc, dpt, ctf);
        rot8_step<MODE, NORM>(a, s[u], kk + u, cj, angle, snj, ncnj, acc, shr, shi, cs, ss);
      }
  }
  // wavefront sums, then the waves' partial sums in a fixed order
  __shared__ double red[6][16];
  const double accv[6] = {acc.num, acc.gx, acc.gy, acc.gpsi, acc.sqref, acc.sqptcl};
#pragma unroll
  for (int v = 0; v < 6; ++v) {
    double x = accv[v];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if ((threadIdx.x & 63) == 0) red[v][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int v = 0; v < 6; ++v)
      for (int w = 1; w < nw; ++w) red[v][0] += red[v][w];
    double f, g0, g1, g2;
    if (a.mode == 2) {
      f = red[0][0];
      g0 = red[1][0];
      g1 = red[2][0];
      g2 = 0.0;
    } else {
      double den;
      if (a.mode == 1) den = sqrt(red[4][0] * (double)a.sqsums[ip]);
      else den = sqrt(red[4][0] * red[5][0]);
      f = red[0][0] / den;
      g0 = red[1][0] / den;
      g1 = red[2][0] / den;
      g2 = red[3][0] / den;
    }
    a.f[item] = f;
    if (a.grad) {
      a.grad[3 * item] = g0;
      a.grad[3 * item + 1] = g1;
      a.grad[3 * item + 2] = g2;
    }
  }
}

// Orientation-derivative objective of continuous refinement: gencorr_cont_grad_cc_for_rot_8 and
// gencorr_cont_shift_grad_cc_for_rot_8 (:1735-1847).  q = ref x shellnorm x |CTF| x shmat_8 and the
// three derivative rings dq_d = dref_d x |CTF| x shmat_8 (the shell normalisation is applied to q
// only, as the reference does), over shells kfrom..kstop, in FP64:
//   y   = sum_k Re sum_j q conj(P_ip rotated),           S_q = sum_k sum_j |q|^2,
//   T1d = sum_k Re sum_j dq_d conj(P_it1 rotated),       T2d = sum_k Re sum_j dq_d conj(q),
//   gx  = Re sum (i a_x q) conj(P_ip rotated),            gy likewise with a_y,
// where the T1 particle is it1, which the reference takes from its iref argument
// (calc_T1_T2_for_rot_8(pft_ref, pft_dref, iref, ...), :1781, :1836).  out[kCont8Sums] = {y, S_q,
// T1[3], T2[3], gx, gy} per item, reduced in a fixed order.  One work-group per item.
constexpr int kCont8Sums = 10;
struct Cont8Args {
  int n, nk, kcorr, kfrom, nrefs, ndslots;
  const float2* refs_even;
  const float2* refs_odd;
  const float2* drefs_even;    // [ndslots][3][kcorr][n]
  const float2* drefs_odd;
  const float2* ptcls;
  const float* ctf;
  const int8_t* iseven;
  const double* refinv;        // shell normalisation (match_filt / pssnr), as Rot8Args, or null
  const double* ofac;
  const int* optrow;
  const float* sn32;
  const float* ncn32;
  double arg_a;
  const double* argx;          // [nk][n]; the first shell's row starts the phase recurrence
  const double* argy;
  const double* one_x;
  const double* one_y;
  const int* irefs;            // 0-based reference (and derivative slot)
  const int* ips;              // local 0-based particle
  const int* it1s;             // local 0-based particle of the T1 terms
  const double* shifts;        // [nitems][2]
  const int* rhos;             // 0-based rotation
  int nitems;
  double* out;                 // [nitems][kCont8Sums]
};

// Rotated particle value P[(j + rho) mod 2n] on the half ring (conjugated past n).
__device__ __forceinline__ void rot_ptcl(const float2* ring, int n, int j, int rho, double& pr, double& pi) {
  int t = j + rho;
  bool cj = false;
  if (t >= 2 * n) t -= 2 * n;
  if (t >= n) { t -= n; cj = true; }
  const float2 v = ring[t];
  pr = v.x;
  pi = cj ? -(double)v.y : (double)v.y;
}

__global__ void __launch_bounds__(256) k_cont8(Cont8Args a) {
  const int item = blockIdx.x;
  if (item >= a.nitems) return;
  const int n = a.n;
  const int iref = a.irefs[item], ip = a.ips[item], it1 = a.it1s[item], rho = a.rhos[item];
  const double sx = a.shifts[2 * item], sy = a.shifts[2 * item + 1];
  const bool even = a.iseven[ip] != 0;
  const float2* ref = (even ? a.refs_even : a.refs_odd) + (size_t)iref * a.nk * n;
  const float2* dref = (even ? a.drefs_even : a.drefs_odd) + (size_t)iref * 3 * a.kcorr * n;
  const float2* ptc = a.ptcls + (size_t)ip * a.nk * n;
  const float2* pt1 = a.ptcls + (size_t)it1 * a.nk * n;
  const float* ctf = a.ctf ? a.ctf + (size_t)ip * a.nk * n : nullptr;
  double acc[kCont8Sums] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double sa, ca, ss, cs;
    sincos(a.argx[j] * sx + a.argy[j] * sy, &sa, &ca);
    sincos(a.one_x[j] * sx + a.one_y[j] * sy, &ss, &cs);
    double shr = ca, shi = sa;
    for (int kk = 0; kk < a.kcorr; ++kk) {
      const double c = ctf ? (double)ctf[(size_t)kk * n + j] : 1.0;
      double fac = 1.0;
      if (a.refinv) {
        const int row = a.optrow ? a.optrow[ip] : iref;
        fac = a.ofac[(size_t)row * a.kcorr + kk] * a.refinv[((size_t)(even ? 0 : 1) * a.nrefs + iref) * a.kcorr + kk];
      }
      const float2 rv = ref[(size_t)kk * n + j];
      const double ar = (double)rv.x * fac * c, ai = (double)rv.y * fac * c;
      const double qr = ar * shr - ai * shi, qi = ar * shi + ai * shr;
      double pr, pim, tr, tim;
      rot_ptcl(ptc + (size_t)kk * n, n, j, rho, pr, pim);
      rot_ptcl(pt1 + (size_t)kk * n, n, j, rho, tr, tim);
      acc[0] += qr * pr + qi * pim;
      acc[1] += qr * qr + qi * qi;
      const double im = qi * pr - qr * pim;          // Im(q conj P); Re((i a q) conj P) = -a Im(q conj P)
      const float kf = (float)(a.kfrom + kk);
      acc[8] += -((double)(a.sn32[j] * kf) * a.arg_a) * im;
      acc[9] += -((double)(a.ncn32[j] * kf) * a.arg_a) * im;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float2 dv = dref[((size_t)d * a.kcorr + kk) * n + j];
        const double br = (double)dv.x * c, bi = (double)dv.y * c;
        const double dr = br * shr - bi * shi, di = br * shi + bi * shr;
        acc[2 + d] += dr * tr + di * tim;
        acc[5 + d] += dr * qr + di * qi;
      }
      const double r2 = shr * cs - shi * ss, i2 = shr * ss + shi * cs;
      shr = r2;
      shi = i2;
    }
  }
  __shared__ double red[kCont8Sums][4];
#pragma unroll
  for (int v = 0; v < kCont8Sums; ++v) {
    double x = acc[v];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if ((threadIdx.x & 63) == 0) red[v][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < kCont8Sums) {
    const int nw = (blockDim.x + 63) >> 6;
    double x = red[threadIdx.x][0];
    for (int w = 1; w < nw; ++w) x += red[threadIdx.x][w];
    a.out[(size_t)item * kCont8Sums + threadIdx.x] = x;
  }
}

}  // namespace pfk
