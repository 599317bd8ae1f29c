// matchfilt.hpp - prep_matchfilt (simple_polarft_corrcalc.f90:987-1095) on the device.
//
// One work-group (4 wavefronts) per listed particle.  Wavefront w takes shells w, w+4, ...: it
// forms the per-shell powers (FP64 sums of FP32 values, cast to FP32 as the reference's
// real(sum(csq_fast(dcmplx(.)))/pftsz)), the normalised particle and reference (x |CTF|) and the
// power of their difference at the particle's rotation.  One thread then smooths the difference
// spectrum over shells (Savitzky-Golay, 7 points, order 4, SavitzkyGolay_filter in
// simple_math.f90:3511-3542: window points left of the first shell are skipped, the last 3 shells
// keep their values; then non-positive fits and the first 3 shells take the raw values) and sets
// the final shell weights.  Finally every thread rescales the particle rings in place.  With
// pssnr the weight is sqrt(k)/sqrt(pw) sqrt(1 + ssnr |CTF|^2) per sample.  Shells beyond kstop
// are untouched.  FP32 arithmetic in the reference's operation order.
#pragma once
#include "pftcc_kernels.hpp"

namespace pfk {

constexpr int MF_MAXK = 512;

struct MatchFiltArgs {
  int n, nk, kcorr, kfrom, nrefs;
  int pssnr;
  float2* ptcls;               // [nptcls][nk][n], rescaled in place
  const float2* refs_even;
  const float2* refs_odd;
  const float* ctf;            // [nptcls][nk][n] or null
  const float* optlp;          // [nrefs][kcorr]
  const int8_t* iseven;
  const int* plist;            // local particle index per item
  const int* rlist;            // 0-based reference per item (< 0: no filtering)
  const int* rhos;             // 0-based rotation per item
  int count;
};

// 7-point quartic smoothing weights (5, -30, 75, 131, 75, -30, 5) / 231, offsets -3..3
__device__ __forceinline__ float sg_weight(int o) {
  switch (o) {
    case -3:
    case 3: return 5.0f / 231.0f;
    case -2:
    case 2: return -30.0f / 231.0f;
    case -1:
    case 1: return 75.0f / 231.0f;
    default: return 131.0f / 231.0f;
  }
}

__global__ void __launch_bounds__(256) k_prep_matchfilt(MatchFiltArgs a) {
  __shared__ float s_pwp[MF_MAXK], s_pwp2[MF_MAXK], s_pwr2[MF_MAXK], s_diff[MF_MAXK], s_w[MF_MAXK];
  const int item = blockIdx.x;
  if (item >= a.count) return;
  const int iref = a.rlist[item];
  if (iref < 0) return;                         // memoize only
  const int ip = a.plist[item];
  const int n = a.n, rho = a.rhos[item];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float2* pt = a.ptcls + (size_t)ip * a.nk * n;
  const float* ct = a.ctf ? a.ctf + (size_t)ip * a.nk * n : nullptr;
  const float* opt = a.optlp + (size_t)iref * a.kcorr;
  if (a.pssnr) {
    for (int kk = wv; kk < a.kcorr; kk += 4) {
      double s = 0.0;
      for (int j = lane; j < n; j += 64) {
        const float2 v = pt[(size_t)kk * n + j];
        s += (double)v.x * v.x + (double)v.y * v.y;
      }
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
      const float pw = (float)(s / n);
      const float ssnr = opt[kk];
      float w = sqrtf((float)(a.kfrom + kk));
      if (pw > 1.e-12f) w = w / sqrtf(pw);
      for (int j = lane; j < n; j += 64) {
        float2 v = pt[(size_t)kk * n + j];
        const float c = ct ? ct[(size_t)kk * n + j] : 0.f;
        const float f = ct ? sqrtf(1.f + ssnr * (c * c)) : sqrtf(1.f + ssnr);
        v.x = (v.x * w) * f;
        v.y = (v.y * w) * f;
        pt[(size_t)kk * n + j] = v;
      }
    }
    return;
  }
  const float2* ref = (a.iseven[ip] ? a.refs_even : a.refs_odd) + (size_t)iref * a.nk * n;
  // ---- per shell: powers of the particle and the reference, normalisation, difference ----
  for (int kk = wv; kk < a.kcorr; kk += 4) {
    const float2* p = pt + (size_t)kk * n;
    const float2* r = ref + (size_t)kk * n;
    const float* c = ct ? ct + (size_t)kk * n : nullptr;
    double sp = 0.0, sr = 0.0;
    for (int j = lane; j < n; j += 64) {
      const float2 v = p[j];
      sp += (double)v.x * v.x + (double)v.y * v.y;
      float2 q = r[j];
      if (c) {
        q.x *= c[j];
        q.y *= c[j];
      }
      sr += (double)q.x * q.x + (double)q.y * q.y;
    }
    for (int o = 32; o > 0; o >>= 1) {
      sp += __shfl_xor(sp, o);
      sr += __shfl_xor(sr, o);
    }
    const float pwp = (float)(sp / n), pwr = (float)(sr / n);
    const float o_k = opt[kk];
    const bool np_ = pwp > 1.e-12f, nr_ = pwr > 1.e-12f;
    const float dp_ = np_ ? sqrtf(pwp) : 1.f, dr_ = nr_ ? sqrtf(pwr) : 1.f;
    double sd = 0.0, sp2 = 0.0, sr2 = 0.0;
    for (int j = lane; j < n; j += 64) {
      float2 q = r[j];
      if (c) {
        q.x *= c[j];
        q.y *= c[j];
      }
      float2 rn = make_float2(q.x * o_k, q.y * o_k);
      if (nr_) rn = make_float2(rn.x / dr_, rn.y / dr_);
      int t = j + rho;
      if (t >= 2 * n) t -= 2 * n;
      const bool cj = t >= n;
      if (cj) t -= n;
      float2 pv = p[t];
      if (cj) pv.y = -pv.y;
      float2 pn = make_float2(pv.x * o_k, pv.y * o_k);
      if (np_) pn = make_float2(pn.x / dp_, pn.y / dp_);
      const float dx = rn.x - pn.x, dy = rn.y - pn.y;
      sd += (double)dx * dx + (double)dy * dy;
      float2 pj = p[j];
      pj = make_float2(pj.x * o_k, pj.y * o_k);
      if (np_) pj = make_float2(pj.x / dp_, pj.y / dp_);
      sp2 += (double)pj.x * pj.x + (double)pj.y * pj.y;
      sr2 += (double)rn.x * rn.x + (double)rn.y * rn.y;
    }
    for (int o = 32; o > 0; o >>= 1) {
      sd += __shfl_xor(sd, o);
      sp2 += __shfl_xor(sp2, o);
      sr2 += __shfl_xor(sr2, o);
    }
    if (lane == 0) {
      s_pwp[kk] = pwp;
      s_diff[kk] = (float)(sd / n);
      s_pwp2[kk] = (float)(sp2 / n);
      s_pwr2[kk] = (float)(sr2 / n);
    }
  }
  __syncthreads();
  // ---- smoothing of the difference spectrum and the final shell weights ----
  if (threadIdx.x == 0) {
    const int K = a.kcorr;
    for (int i = 0; i < K; ++i) {
      float fit = s_diff[i];
      if (i < K - 3) {
        float acc = 0.f;
        for (int o = -3; o <= 3; ++o)
          if (i + o >= 0) acc += sg_weight(o) * s_diff[i + o];
        fit = acc;
      }
      if (fit <= 0.f || i < 3) fit = s_diff[i];
      float w = 1.f;
      if (s_pwp2[i] > 1.e-12f) w = 1.f / sqrtf(s_pwp2[i]);
      if (fit > 1.e-12f) {
        const float ssnr = s_pwr2[i] / fit;
        w = w * sqrtf(ssnr / (ssnr + 1.f));
      }
      s_w[i] = w;
    }
  }
  __syncthreads();
  // ---- rescale the particle: p <- w_k ((p optlp_k) / sqrt(pw_k)) ----
  for (int e = threadIdx.x; e < a.kcorr * n; e += blockDim.x) {
    const int kk = e / n;
    float2 v = pt[e];
    const float o_k = opt[kk];
    v = make_float2(v.x * o_k, v.y * o_k);
    if (s_pwp[kk] > 1.e-12f) {
      const float d = sqrtf(s_pwp[kk]);
      v = make_float2(v.x / d, v.y / d);
    }
    pt[e] = make_float2(s_w[kk] * v.x, s_w[kk] * v.y);
  }
}

}  // namespace pfk
