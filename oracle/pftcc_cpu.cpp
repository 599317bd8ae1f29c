// pftcc_cpu.cpp - CPU restatement of SIMPLE's gencorrs path (TEST INFRASTRUCTURE ONLY).
//
// Two things live here, both used only by tests/ and by bench.py's cpu_baseline leg:
//  1. cpu_search_fp32: the reference CPU algorithm in FP32 with OpenMP over particles, as
//     simple_strategy3D_matcher.f90:204-205 runs it (static schedule), per pair
//       prep_ref4corr (:961-985)   q = ref * |CTF|, sqsum_ref over kfrom..kstop
//       calc_corrs_over_k (:1132-1162), per shell:
//         A = r2c_n(Re q), B = c2c_n(Im q * fft_factors), product(2t) = 4 conj(A) PA,
//         product(2t+1) = 4 conj(B) PB, backtransf = c2r_2n(product), corrs += backtransf
//       cc = corrs / (2 nrots) / sqrt(sqsum_ref * sqsums_ptcl)            (:1499)
//     then maxloc over rotations and the greedy maximum over references.  FFTW3 is not
//     installed here, so the transforms are this file's own mixed-radix Stockham FFT.
//  2. cpu_gencorrs_direct64: brute-force FP64 direct sums of the same correlation (the
//     four-case sum of calc_corr_for_rot_8, :1218-1243, for every rotation), used as the
//     oracle at sizes where numpy would be slow.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

#ifdef _OPENMP
#include <omp.h>
#endif

using cf = std::complex<float>;
using cd = std::complex<double>;

namespace {

struct Plan {
  int n = 0;
  std::vector<int> radix;
  std::vector<cf> tw;   // exp(-2 pi i a / n)
};

Plan make_plan(int n) {
  Plan p;
  p.n = n;
  int m = n;
  while (m % 4 == 0) { p.radix.push_back(4); m /= 4; }
  while (m % 2 == 0) { p.radix.push_back(2); m /= 2; }
  for (int q : {3, 5, 7, 11, 13, 17, 19, 23}) while (m % q == 0) { p.radix.push_back(q); m /= q; }
  if (m > 1) p.radix.push_back(m);   // generic radix (O(m^2))
  p.tw.resize(n);
  for (int a = 0; a < n; ++a) {
    const double ang = -2.0 * M_PI * a / n;
    p.tw[a] = cf((float)std::cos(ang), (float)std::sin(ang));
  }
  return p;
}

// generic DFT of R points with roots from the length-n table (stride n/R)
inline void dft_generic(cf* v, int R, const Plan& p, cf* tmp) {
  const int st = p.n / R;
  for (int k = 0; k < R; ++k) {
    cf s = 0;
    for (int t = 0; t < R; ++t) s += v[t] * p.tw[((size_t)k * t % R) * st];
    tmp[k] = s;
  }
  for (int k = 0; k < R; ++k) v[k] = tmp[k];
}

inline void dft4(cf* v) {
  const cf t0 = v[0] + v[2], t1 = v[0] - v[2], t2 = v[1] + v[3], t3 = v[1] - v[3];
  v[0] = t0 + t2;
  v[2] = t0 - t2;
  v[1] = cf(t1.real() + t3.imag(), t1.imag() - t3.real());
  v[3] = cf(t1.real() - t3.imag(), t1.imag() + t3.real());
}

// forward complex FFT (exp(-)), Stockham; x is overwritten with the result, w is scratch
void fft(const Plan& p, cf* x, cf* w) {
  const int n = p.n;
  int Ns = 1;
  cf* src = x;
  cf* dst = w;
  cf v[32], tmp[32];
  for (int R : p.radix) {
    const int nb = n / R;
    const int tstep = n / (Ns * R);
    for (int j = 0; j < nb; ++j) {
      const int jm = j % Ns;
      for (int t = 0; t < R; ++t) v[t] = src[j + t * nb];
      if (Ns > 1)
        for (int t = 1; t < R; ++t) v[t] *= p.tw[(size_t)jm * t * tstep % n];
      if (R == 4) dft4(v);
      else if (R == 2) { const cf a = v[0] + v[1]; v[1] = v[0] - v[1]; v[0] = a; }
      else dft_generic(v, R, p, tmp);
      const int base = (j / Ns) * Ns * R + jm;
      for (int t = 0; t < R; ++t) dst[base + t * Ns] = v[t];
    }
    Ns *= R;
    std::swap(src, dst);
  }
  if (src != x) std::memcpy(x, src, sizeof(cf) * n);
}

// Per-thread transforms of the reference: r2c_n, c2c_n, c2r_2n (FFTW conventions, unnormalised)
struct Transforms {
  int n;
  Plan ph, pn;          // n/2 and n
  std::vector<cf> a, b, c, d;
  explicit Transforms(int n_) : n(n_), ph(make_plan(n_ / 2)), pn(make_plan(n_)), a(2 * n_ + 2), b(2 * n_ + 2),
                                c(2 * n_ + 2), d(2 * n_ + 2) {}
  // out[0..n/2] = sum_j x[j] exp(-2 pi i j k / n), x real of length n (n even)
  void r2c(const float* x, cf* out) {
    const int h = n / 2;
    for (int j = 0; j < h; ++j) a[j] = cf(x[2 * j], x[2 * j + 1]);
    fft(ph, a.data(), b.data());
    for (int k = 0; k <= h; ++k) {
      const cf A = a[k % h], B = std::conj(a[(h - k) % h]);
      const cf e = 0.5f * (A + B), o = cf(0.f, -0.5f) * (A - B);
      const double ang = -2.0 * M_PI * k / n;
      out[k] = e + cf((float)std::cos(ang), (float)std::sin(ang)) * o;
    }
  }
  void c2c(const cf* x, cf* out) {
    for (int j = 0; j < n; ++j) a[j] = x[j];
    fft(pn, a.data(), b.data());
    for (int j = 0; j < n; ++j) out[j] = a[j];
  }
  // out[t] = sum_{m=0}^{2n-1} X[m] exp(+2 pi i m t / 2n), X Hermitian from X[0..n]
  void c2r2n(const cf* X, float* out) {
    for (int m = 0; m < n; ++m) {
      cf Xm = X[m], Xn = X[n - m];
      if (m == 0) { Xm = cf(Xm.real(), 0.f); Xn = cf(Xn.real(), 0.f); }
      const cf s = Xm + std::conj(Xn), df = Xm - std::conj(Xn);
      const double ang = M_PI * m / n;
      const cf e((float)std::cos(ang), (float)std::sin(ang));
      c[m] = std::conj(s + cf(0.f, 1.f) * e * df);
    }
    fft(pn, c.data(), d.data());
    for (int j = 0; j < n; ++j) {
      out[2 * j] = c[j].real();
      out[2 * j + 1] = -c[j].imag();
    }
  }
};

// ------------------------------------------------------------------------------------------
// Eight transforms at once: the shells of one (reference, particle) pair in groups of eight, one
// shell per lane of an 8-wide float vector (AVX2), structure-of-arrays complex values.  The same
// Stockham passes as above with every twiddle precomputed per pass (no trigonometry per call).
// ------------------------------------------------------------------------------------------
typedef float v8 __attribute__((vector_size(32)));
struct vc {
  v8 re, im;
};
inline vc vadd(const vc& a, const vc& b) { return {a.re + b.re, a.im + b.im}; }
inline vc vsub(const vc& a, const vc& b) { return {a.re - b.re, a.im - b.im}; }
inline vc vmul(const vc& a, cf w) {   // by a scalar twiddle
  const float c = w.real(), s = w.imag();
  return {a.re * c - a.im * s, a.re * s + a.im * c};
}
inline vc vmulc(const vc& a, const vc& b) {   // conj(a) * b, lane by lane
  return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}

struct VPlan {
  int n = 0;
  std::vector<int> radix;
  std::vector<std::vector<cf>> ptw;   // per pass: [jm][t] = W_n^(jm t tstep)
  std::vector<std::vector<cf>> dft;   // per pass: [k][t] = exp(-2 pi i k t / R) for generic radices
};

VPlan make_vplan(int n) {
  VPlan p;
  p.n = n;
  const Plan sp = make_plan(n);
  p.radix = sp.radix;
  int Ns = 1;
  for (int R : p.radix) {
    const int tstep = n / (Ns * R);
    std::vector<cf> tw((size_t)Ns * R);
    for (int jm = 0; jm < Ns; ++jm)
      for (int t = 0; t < R; ++t) {
        const double ang = -2.0 * M_PI * (double)((size_t)jm * t * tstep % n) / n;
        tw[(size_t)jm * R + t] = cf((float)std::cos(ang), (float)std::sin(ang));
      }
    p.ptw.push_back(tw);
    std::vector<cf> d((size_t)R * R);
    for (int k = 0; k < R; ++k)
      for (int t = 0; t < R; ++t) {
        const double ang = -2.0 * M_PI * (double)((k * t) % R) / R;
        d[(size_t)k * R + t] = cf((float)std::cos(ang), (float)std::sin(ang));
      }
    p.dft.push_back(d);
    Ns *= R;
  }
  return p;
}

// forward complex FFT (exp(-)) of 8 lanes; x overwritten, w scratch (n each)
void vfft(const VPlan& p, vc* x, vc* w) {
  const int n = p.n;
  int Ns = 1;
  vc* src = x;
  vc* dst = w;
  vc v[32], o[32];
  for (size_t ip = 0; ip < p.radix.size(); ++ip) {
    const int R = p.radix[ip];
    const int nb = n / R;
    const cf* tw = p.ptw[ip].data();
    const cf* D = p.dft[ip].data();
    for (int j = 0; j < nb; ++j) {
      const int jm = j % Ns;
      for (int t = 0; t < R; ++t) v[t] = src[j + t * nb];
      if (Ns > 1)
        for (int t = 1; t < R; ++t) v[t] = vmul(v[t], tw[(size_t)jm * R + t]);
      const int base = (j / Ns) * Ns * R + jm;
      if (R == 4) {
        const vc t0 = vadd(v[0], v[2]), t1 = vsub(v[0], v[2]), t2 = vadd(v[1], v[3]), t3 = vsub(v[1], v[3]);
        dst[base] = vadd(t0, t2);
        dst[base + 2 * Ns] = vsub(t0, t2);
        dst[base + Ns] = {t1.re + t3.im, t1.im - t3.re};
        dst[base + 3 * Ns] = {t1.re - t3.im, t1.im + t3.re};
      } else if (R == 2) {
        dst[base] = vadd(v[0], v[1]);
        dst[base + Ns] = vsub(v[0], v[1]);
      } else {
        // generic radix: o[k] = sum_t v[t] D[k][t]; k = 0 is the plain sum
        o[0] = v[0];
        for (int t = 1; t < R; ++t) o[0] = vadd(o[0], v[t]);
        for (int k = 1; k < R; ++k) {
          vc s = v[0];
          for (int t = 1; t < R; ++t) s = vadd(s, vmul(v[t], D[(size_t)k * R + t]));
          o[k] = s;
        }
        for (int k = 0; k < R; ++k) dst[base + k * Ns] = o[k];
      }
    }
    Ns *= R;
    std::swap(src, dst);
  }
  if (src != x) std::memcpy(static_cast<void*>(x), src, sizeof(vc) * n);
}

// r2c_n, c2c_n and c2r_2n of the reference (FFTW conventions, unnormalised) on 8 lanes.
struct Transforms8 {
  int n;
  VPlan ph, pn;
  std::vector<vc> a, b;
  std::vector<cf> er;   // exp(-2 pi i k / n), k = 0..n/2 (the r2c split)
  std::vector<cf> ec;   // exp(+i pi m / n), m = 0..n-1 (the c2r pre-twiddle)
  explicit Transforms8(int n_) : n(n_), ph(make_vplan(n_ / 2)), pn(make_vplan(n_)), a(n_ + 2), b(n_ + 2),
                                 er(n_ / 2 + 1), ec(n_) {
    for (int k = 0; k <= n / 2; ++k) {
      const double ang = -2.0 * M_PI * k / n;
      er[k] = cf((float)std::cos(ang), (float)std::sin(ang));
    }
    for (int m = 0; m < n; ++m) {
      const double ang = M_PI * m / n;
      ec[m] = cf((float)std::cos(ang), (float)std::sin(ang));
    }
  }
  // out[0..n/2] = sum_j x[j] exp(-2 pi i j k / n), x real
  void r2c(const v8* x, vc* out) {
    const int h = n / 2;
    for (int j = 0; j < h; ++j) a[j] = {x[2 * j], x[2 * j + 1]};
    vfft(ph, a.data(), b.data());
    for (int k = 0; k <= h; ++k) {
      const vc A = a[k % h], Bc = a[(h - k) % h];            // B = conj(Bc)
      const vc e = {0.5f * (A.re + Bc.re), 0.5f * (A.im - Bc.im)};
      const vc d = {A.re - Bc.re, A.im + Bc.im};              // A - B
      const vc o = {0.5f * d.im, -0.5f * d.re};               // -i/2 (A - B)
      out[k] = vadd(e, vmul(o, er[k]));
    }
  }
  void c2c(vc* x) { vfft(pn, x, b.data()); }
  // out[t] = sum_{m=0}^{2n-1} X[m] exp(+2 pi i m t / 2n), X Hermitian from X[0..n]
  void c2r2n(const vc* X, v8* out) {
    const v8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int m = 0; m < n; ++m) {
      vc Xm = X[m], Xn = X[n - m];
      if (m == 0) {
        Xm.im = zero;
        Xn.im = zero;
      }
      const vc s = {Xm.re + Xn.re, Xm.im - Xn.im};          // Xm + conj(Xn)
      const vc df = {Xm.re - Xn.re, Xm.im + Xn.im};         // Xm - conj(Xn)
      const vc ed = vmul(df, ec[m]);
      const vc t = {s.re - ed.im, s.im + ed.re};             // s + i e df
      a[m] = {t.re, -t.im};                                   // conj
    }
    vfft(pn, a.data(), b.data());
    for (int j = 0; j < n; ++j) {
      out[2 * j] = a[j].re;
      out[2 * j + 1] = -a[j].im;
    }
  }
};

}  // namespace

extern "C" {

// memoize_fft (:864-877): pa[np][nk][n/2+1], pb[np][nk][n]
void cpu_memoize(int np, int nk, int n, const float* ptcls, float* pa, float* pb, int nthreads) {
  const int h = n / 2;
#pragma omp parallel num_threads(nthreads)
  {
    Transforms T(n);
    std::vector<float> re(n);
    std::vector<cf> im(n);
#pragma omp for schedule(static)
    for (int i = 0; i < np; ++i)
      for (int kk = 0; kk < nk; ++kk) {
        const cf* p = reinterpret_cast<const cf*>(ptcls) + ((size_t)i * nk + kk) * n;
        for (int j = 0; j < n; ++j) {
          re[j] = p[j].real();
          const double ang = -M_PI * j / n;
          im[j] = p[j].imag() * cf((float)std::cos(ang), (float)std::sin(ang));
        }
        T.r2c(re.data(), reinterpret_cast<cf*>(pa) + ((size_t)i * nk + kk) * (h + 1));
        T.c2c(im.data(), reinterpret_cast<cf*>(pb) + ((size_t)i * nk + kk) * n);
      }
  }
}

// Exhaustive FP32 search of the reference algorithm.
//  refs [nr][nk][n] complex, ctf [np][nk][n] (or null), sqsums_ptcl [np], memo pa/pb,
//  shmat [ns][nk][n] complex FP32 (or null with ns = 1): for every particle the best
//  (ref, shift, irot) by greedy maxloc.  cc_all (optional) [np][nr][ns][2n].
void cpu_search_fp32(int nr, int np, int ns, int nk, int kcorr, int n, const float* refs, const float* ctf,
                     const float* sqsums_ptcl, const float* pa, const float* pb, const float* shmat,
                     float* best_cc, int32_t* best_ref, int32_t* best_sh, int32_t* best_rot, float* cc_all,
                     int nthreads) {
  const int h = n / 2, R = 2 * n;
  const cf* R_ = reinterpret_cast<const cf*>(refs);
  const cf* PA = reinterpret_cast<const cf*>(pa);
  const cf* PB = reinterpret_cast<const cf*>(pb);
  const cf* SH = reinterpret_cast<const cf*>(shmat);
  const int ng = (kcorr + 7) / 8;   // groups of eight shells
#pragma omp parallel num_threads(nthreads)
  {
    Transforms8 T(n);
    std::vector<cf> q((size_t)nk * n);
    std::vector<v8> re(n), back(R), corrs(R);
    std::vector<vc> A(h + 1), B(n + 2), prod(n + 2);
    std::vector<vc> pav((size_t)ng * (h + 1)), pbv((size_t)ng * n);   // the particle's memo, 8 shells per vector
    std::vector<cf> fftf(n);
    for (int j = 0; j < n; ++j) {
      const double ang = -M_PI * j / n;
      fftf[j] = cf((float)std::cos(ang), (float)std::sin(ang));
    }
    const v8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma omp for schedule(static)
    for (int i = 0; i < np; ++i) {
      for (int g = 0; g < ng; ++g) {
        for (int t = 0; t <= h; ++t) {
          vc v = {zero, zero};
          for (int l = 0; l < 8 && 8 * g + l < kcorr; ++l) {
            const cf x = PA[((size_t)i * nk + 8 * g + l) * (h + 1) + t];
            v.re[l] = x.real();
            v.im[l] = x.imag();
          }
          pav[(size_t)g * (h + 1) + t] = v;
        }
        for (int t = 0; t < n; ++t) {
          vc v = {zero, zero};
          for (int l = 0; l < 8 && 8 * g + l < kcorr; ++l) {
            const cf x = PB[((size_t)i * nk + 8 * g + l) * n + t];
            v.re[l] = x.real();
            v.im[l] = x.imag();
          }
          pbv[(size_t)g * n + t] = v;
        }
      }
      float bc = -INFINITY;
      int br = 0, bs = 0, bt = 0;
      for (int r = 0; r < nr; ++r)
        for (int s = 0; s < ns; ++s) {
          // prep_ref4corr: copy, |CTF|, (shift), square sum over kfrom..kstop
          float sqsum = 0.f;
          for (int kk = 0; kk < kcorr; ++kk)
            for (int j = 0; j < n; ++j) {
              cf v = R_[((size_t)r * nk + kk) * n + j];
              if (ctf) v *= ctf[((size_t)i * nk + kk) * n + j];
              if (SH) v *= SH[((size_t)s * nk + kk) * n + j];
              q[(size_t)kk * n + j] = v;
              sqsum += v.real() * v.real() + v.imag() * v.imag();
            }
          // calc_corrs_over_k, eight shells per pass (one per lane); the lanes are summed at the end
          std::fill(corrs.begin(), corrs.end(), zero);
          for (int g = 0; g < ng; ++g) {
            const int nl = std::min(8, kcorr - 8 * g);
            for (int j = 0; j < n; ++j) {
              v8 xr = zero, yr = zero, yi = zero;
              for (int l = 0; l < nl; ++l) {
                const cf x = q[(size_t)(8 * g + l) * n + j];
                xr[l] = x.real();
                yr[l] = x.imag() * fftf[j].real();
                yi[l] = x.imag() * fftf[j].imag();
              }
              re[j] = xr;
              B[j] = {yr, yi};
            }
            T.r2c(re.data(), A.data());
            T.c2c(B.data());
            const vc* pak = &pav[(size_t)g * (h + 1)];
            const vc* pbk = &pbv[(size_t)g * n];
            for (int t = 0; t <= h; ++t) {
              const vc a4 = vmulc(A[t], pak[t]), b4 = vmulc(B[t], pbk[t]);
              prod[2 * t] = {4.f * a4.re, 4.f * a4.im};
              if (2 * t + 1 <= n) prod[2 * t + 1] = {4.f * b4.re, 4.f * b4.im};
            }
            T.c2r2n(prod.data(), back.data());
            for (int t = 0; t < R; ++t) corrs[t] += back[t];
          }
          const float scale = 1.0f / (float)(2 * R);
          const float den = std::sqrt(sqsum * sqsums_ptcl[i]);
          float* out = cc_all ? cc_all + (((size_t)i * nr + r) * ns + s) * R : nullptr;
          for (int t = 0; t < R; ++t) {
            float csum = 0.f;
            for (int l = 0; l < 8; ++l) csum += corrs[t][l];
            const float c = csum * scale / den;
            if (out) out[t] = c;
            if (c > bc) {
              bc = c;
              br = r;
              bs = s;
              bt = t;
            }
          }
        }
      best_cc[i] = bc;
      best_ref[i] = br + 1;
      best_sh[i] = bs;
      best_rot[i] = bt + 1;
    }
  }
}

// FP64 brute-force correlations: cc[nr*np][2n] for q = ref*ctf(*shmat) and the particle ring,
// sq_norm_all = 1 normalises the reference over all nk shells (gencorrs_cc_2), else kcorr.
void cpu_gencorrs_direct64(int nr, int np, int nk, int kcorr, int n, const float* refs, const float* ptcls,
                           const float* ctf, const double* shmat, int sq_norm_all, double* cc, int nthreads) {
  const int R = 2 * n;
  const cf* Rf = reinterpret_cast<const cf*>(refs);
  const cf* Pf = reinterpret_cast<const cf*>(ptcls);
  const cd* SH = reinterpret_cast<const cd*>(shmat);
#pragma omp parallel for collapse(2) schedule(dynamic) num_threads(nthreads)
  for (int r = 0; r < nr; ++r)
    for (int i = 0; i < np; ++i) {
      std::vector<cd> q((size_t)nk * n), P((size_t)nk * R);
      double sq = 0, sp = 0;
      for (int kk = 0; kk < nk; ++kk)
        for (int j = 0; j < n; ++j) {
          cd v = cd(Rf[((size_t)r * nk + kk) * n + j]);
          if (ctf) v *= (double)ctf[((size_t)i * nk + kk) * n + j];
          if (SH) v *= SH[(size_t)kk * n + j];
          q[(size_t)kk * n + j] = v;
          if (sq_norm_all || kk < kcorr) sq += std::norm(v);
          const cd p = cd(Pf[((size_t)i * nk + kk) * n + j]);
          P[(size_t)kk * R + j] = p;
          P[(size_t)kk * R + j + n] = std::conj(p);
          if (kk < kcorr) sp += std::norm(p);
        }
      const double den = std::sqrt(sq * sp);
      double* out = cc + ((size_t)r * np + i) * R;
      for (int rho = 0; rho < R; ++rho) {
        double s = 0;
        for (int kk = 0; kk < kcorr; ++kk) {
          const cd* qk = &q[(size_t)kk * n];
          const cd* pk = &P[(size_t)kk * R];
          for (int j = 0; j < n; ++j) {
            const int t = (j + rho) % R;
            s += qk[j].real() * pk[t].real() + qk[j].imag() * pk[t].imag();
          }
        }
        out[rho] = s / den;
      }
    }
}

int cpu_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
