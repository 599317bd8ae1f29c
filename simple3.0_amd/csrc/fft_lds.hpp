// fft_lds.hpp - batched mixed-radix complex FFTs of the angular (ring) axis, staged in LDS.
//
// The reference transforms every ring with FFTW (r2c_n, c2c_n, c2r_2n; simple_polarft_corrcalc.f90
// :397-404, 874-875, 1146-1156).  Here a work-group runs F transforms of length n at once with a
// Stockham auto-sort formulation (natural order in and out, one LDS round trip per pass).
// Radices 2, 3, 4, 5, 7, 11, 13 cover every FFT-friendly ring length of simple_magic_boxes.f90.
// Twiddles come from a table tw2[a] = exp(-i*pi*a/n), a in [0, 2n), computed on the host in
// double precision, so tw2[2a] is the n-point root exp(-2*pi*i*a/n).
#pragma once
#include <hip/hip_runtime.h>

namespace pfk {

// ---------------------------------------------------------------- constexpr trigonometry ----
constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double cx_reduce(double x) {
  // reduce to [-pi, pi]
  const double twopi = 2.0 * kPi;
  long long q = (long long)(x / twopi);
  x -= (double)q * twopi;
  if (x > kPi) x -= twopi;
  if (x < -kPi) x += twopi;
  return x;
}
constexpr double cx_sin(double x) {
  x = cx_reduce(x);
  double term = x, sum = x;
  for (int i = 1; i < 30; ++i) {
    term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos(double x) {
  x = cx_reduce(x);
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 30; ++i) {
    term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
    sum += term;
  }
  return sum;
}

// ------------------------------------------------------------------------- complex helpers ----
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ float2 cmulcj(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// --------------------------------------------------------------------------- plan of radices --
struct FftPlan {
  int np;        // number of passes, -1 if n has an unsupported prime factor
  int r[16];     // radices in pass order
};

__host__ __device__ constexpr FftPlan make_plan(int n) {
  FftPlan p{0, {0}};
  if (n <= 0) { p.np = -1; return p; }
  int m = n;
  while (m % 4 == 0) { p.r[p.np++] = 4; m /= 4; }
  while (m % 2 == 0) { p.r[p.np++] = 2; m /= 2; }
  const int primes[5] = {3, 5, 7, 11, 13};
  for (int i = 0; i < 5; ++i) {
    while (m % primes[i] == 0) { p.r[p.np++] = primes[i]; m /= primes[i]; }
  }
  if (m != 1) p.np = -1;
  return p;
}

// --------------------------------------------------------------------------- radix kernels ----
template <int R>
struct RootTable {
  float c[R];
  float s[R];
};
template <int R>
constexpr RootTable<R> make_roots() {
  RootTable<R> t{};
  for (int i = 0; i < R; ++i) {
    t.c[i] = (float)cx_cos(2.0 * kPi * i / R);
    t.s[i] = (float)cx_sin(2.0 * kPi * i / R);
  }
  return t;
}

// In-place forward DFT of R points, exp(-2*pi*i*k*t/R).
template <int R>
struct Dft {
  // odd primes: pairwise symmetric form with compile-time cos/sin constants
  static __device__ __forceinline__ void run(float2 (&v)[R]) {
    static_assert(R % 2 == 1, "odd prime radix expected");
    constexpr int H = (R - 1) / 2;
    float2 a[H], b[H];
#pragma unroll
    for (int t = 1; t <= H; ++t) {
      a[t - 1] = cadd(v[t], v[R - t]);
      b[t - 1] = csub(v[t], v[R - t]);
    }
    float2 y0 = v[0];
#pragma unroll
    for (int t = 0; t < H; ++t) y0 = cadd(y0, a[t]);
    float2 out[R];
    out[0] = y0;
    constexpr RootTable<R> rt = make_roots<R>();
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      float ax = v[0].x, ay = v[0].y, sx = 0.f, sy = 0.f;
#pragma unroll
      for (int t = 1; t <= H; ++t) {
        const float c = rt.c[(k * t) % R];
        const float s = rt.s[(k * t) % R];
        ax = fmaf(a[t - 1].x, c, ax);
        ay = fmaf(a[t - 1].y, c, ay);
        sx = fmaf(b[t - 1].y, s, sx);
        sy = fmaf(b[t - 1].x, s, sy);
      }
      out[k] = make_float2(ax + sx, ay - sy);
      out[R - k] = make_float2(ax - sx, ay + sy);
    }
#pragma unroll
    for (int t = 0; t < R; ++t) v[t] = out[t];
  }
};

template <>
struct Dft<2> {
  static __device__ __forceinline__ void run(float2 (&v)[2]) {
    float2 a = cadd(v[0], v[1]);
    v[1] = csub(v[0], v[1]);
    v[0] = a;
  }
};

template <>
struct Dft<4> {
  static __device__ __forceinline__ void run(float2 (&v)[4]) {
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    // y1 = t1 - i t3 ; y3 = t1 + i t3
    v[1] = make_float2(t1.x + t3.y, t1.y - t3.x);
    v[3] = make_float2(t1.x - t3.y, t1.y + t3.x);
  }
};

// ------------------------------------------------------------------- Stockham passes in LDS ----
// buf holds F transforms of length n at stride ld.  One pass of radix R with Ns = product of the
// radices of the previous passes.  Butterfly j (0 <= j < n/R) of transform f reads
// buf[f][j + t*n/R], applies twiddles exp(-2 pi i (j mod Ns) t / (Ns R)) and writes
// buf[f][(j/Ns)*Ns*R + (j mod Ns) + t*Ns].  Out of place (src -> dst).
template <int R>
__device__ __forceinline__ void stockham_pass_oop(const float2* __restrict__ src, float2* __restrict__ dst,
                                                  int F, int n, int ld, int Ns,
                                                  const float2* __restrict__ tw2, int tid, int nt) {
  const int nb = n / R;
  const int tstep = n / (Ns * R);   // n-point root exponent per unit of (j mod Ns)*t
  for (int task = tid; task < F * nb; task += nt) {
    const int f = task / nb;
    const int j = task - f * nb;
    const int jm = j % Ns;
    float2 v[R];
#pragma unroll
    for (int t = 0; t < R; ++t) v[t] = src[f * ld + j + t * nb];
    if (Ns > 1) {
#pragma unroll
      for (int t = 1; t < R; ++t) {
        const int a = (jm * t * tstep) % n;
        v[t] = cmul(v[t], tw2[2 * a]);
      }
    }
    Dft<R>::run(v);
    const int base = (j / Ns) * Ns * R + jm;
#pragma unroll
    for (int t = 0; t < R; ++t) dst[f * ld + base + t * Ns] = v[t];
  }
}

// Runs a whole forward FFT plan out of place, ping-ponging between a and b.  Returns the buffer
// that holds the result (a or b).  Every thread of the block must call it.
__device__ inline float2* fft_forward_oop(float2* a, float2* b, int F, int n, int ld, const FftPlan& plan,
                                          const float2* tw2, int tid, int nt) {
  int Ns = 1;
  float2* src = a;
  float2* dst = b;
  for (int p = 0; p < plan.np; ++p) {
    const int R = plan.r[p];
    switch (R) {
      case 2: stockham_pass_oop<2>(src, dst, F, n, ld, Ns, tw2, tid, nt); break;
      case 3: stockham_pass_oop<3>(src, dst, F, n, ld, Ns, tw2, tid, nt); break;
      case 4: stockham_pass_oop<4>(src, dst, F, n, ld, Ns, tw2, tid, nt); break;
      case 5: stockham_pass_oop<5>(src, dst, F, n, ld, Ns, tw2, tid, nt); break;
      case 7: stockham_pass_oop<7>(src, dst, F, n, ld, Ns, tw2, tid, nt); break;
      case 11: stockham_pass_oop<11>(src, dst, F, n, ld, Ns, tw2, tid, nt); break;
      case 13: stockham_pass_oop<13>(src, dst, F, n, ld, Ns, tw2, tid, nt); break;
      default: break;
    }
    __syncthreads();
    Ns *= R;
    float2* t = src;
    src = dst;
    dst = t;
  }
  return src;
}

}  // namespace pfk
