// corr_fused.hpp - compile-time-size correlation kernel (the hot path of the plain FP32 search).
//
// Same arithmetic as k_corr_generic (pftcc_kernels.hpp), organised for gfx950:
//  * the ring length N and the pair tile F are template constants, so every index, radix and
//    loop bound folds at compile time;
//  * pass 1 (radix 4) builds its inputs straight from the reference, CTF and shift rings:
//    z[j], z[j+N/4], z[j+N/2], z[j+3N/4] all come from q[2j..2j+1] and q[2j+N/2..2j+N/2+1];
//  * the last pass runs the mirrored butterflies j and Ns-j in one thread, so the real-FFT split
//    V[m] = f(Zc[m], Zc[N-m]) and the product conj(V) W_k accumulate in registers;
//  * the spectra Z of the F pairs stay in VGPRs across all shells; the pre-processing of the
//    final inverse transform (X[m], X[N-m] -> Y[m]) is also done from those registers.
// Deterministic: a pair's arithmetic does not depend on the tile it belongs to.
#pragma once
#include "pftcc_kernels.hpp"

namespace pfk {

// radix order for the fused kernel: one 4 first, the remaining factors in descending order
__host__ __device__ constexpr FftPlan make_fused_plan(int n) {
  FftPlan p = make_plan(n);
  if (p.np < 2 || n % 4 != 0) return p;
  FftPlan q{0, {0}};
  q.r[q.np++] = 4;
  int rest[16] = {0};
  int nr = 0;
  bool dropped = false;
  for (int i = 0; i < p.np; ++i) {
    if (!dropped && p.r[i] == 4) {
      dropped = true;
      continue;
    }
    rest[nr++] = p.r[i];
  }
  for (int i = 0; i < nr; ++i)
    for (int k = i + 1; k < nr; ++k)
      if (rest[k] > rest[i]) {
        const int t = rest[i];
        rest[i] = rest[k];
        rest[k] = t;
      }
  // a radix-2 last pass would leave N / 4 + 1 mirrored groups per transform (one transform per
  // work-group at n 600): where a radix 2 remains, it goes first among the middle passes and the
  // others follow in increasing order, so that the last pass has the largest radix
  bool has2 = false;
  for (int i = 0; i < nr; ++i) has2 = has2 || rest[i] == 2;
  if (!has2) {
    for (int i = 0; i < nr; ++i) q.r[q.np++] = rest[i];
  } else {
    for (int i = nr - 1; i >= 0; --i) q.r[q.np++] = rest[i];
  }
  return q;
}

template <int N>
struct Fp {
  static constexpr FftPlan plan = make_fused_plan(N);
  static constexpr int np = plan.np;
  static constexpr int R(int i) { return plan.r[i]; }
  static constexpr int Ns(int i) {
    int s = 1;
    for (int k = 0; k < i; ++k) s *= plan.r[k];
    return s;
  }
  static constexpr bool fusable = np >= 2 && plan.r[0] == 4 && N % 4 == 0;
};

// Offset of middle pass P's twiddle table [t - 1][jm] (t = 1..R-1, jm < Ns) in a block holding the
// tables of passes 1, 2, ...: entry (t, jm) = tw2[2 jm t N / (Ns R)], read with stride 1 across
// lanes where tw2's index strides land on few banks.
template <int N>
constexpr int twm_off(int P) {
  int off = 0;
  for (int q = 1; q < P; ++q) off += (Fp<N>::R(q) - 1) * Fp<N>::Ns(q);
  return off;
}

// In-place Stockham pass P of F transforms held in LDS (transform f at buf + f S); each thread
// stages its butterflies' inputs in registers between two barriers.  TM: twiddles from the
// middle-pass tables twm (twm_off layout) instead of tw2.
template <int N, int F, int NT, int P, int S = N, bool TM = false>
__device__ __forceinline__ void lds_pass(float2* __restrict__ buf, const float2* __restrict__ tw2, int tid,
                                         const float2* __restrict__ twm = nullptr) {
  constexpr int R = Fp<N>::R(P);
  constexpr int Ns = Fp<N>::Ns(P);
  constexpr int nb = N / R;
  constexpr int TASKS = F * nb;
  constexpr int TPT = (TASKS + NT - 1) / NT;
  constexpr int tstep = N / (Ns * R);
  float2 v[TPT][R];
#pragma unroll
  for (int c = 0; c < TPT; ++c) {
    const int task = tid + c * NT;
    if (TASKS % NT == 0 || task < TASKS) {
      const int f = task / nb, j = task % nb;
#pragma unroll
      for (int t = 0; t < R; ++t) v[c][t] = buf[f * S + j + t * nb];
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < TPT; ++c) {
    const int task = tid + c * NT;
    if (TASKS % NT == 0 || task < TASKS) {
      const int f = task / nb, j = task % nb;
      const int jm = j % Ns;
      if constexpr (Ns > 1) {
#pragma unroll
        for (int t = 1; t < R; ++t)
          v[c][t] = cmul(v[c][t], TM ? twm[twm_off<N>(P) + (t - 1) * Ns + jm] : tw2[2 * jm * t * tstep]);
      }
      Dft<R>::run(v[c]);
      const int base = (j / Ns) * Ns * R + jm;
#pragma unroll
      for (int t = 0; t < R; ++t) buf[f * S + base + t * Ns] = v[c][t];
    }
  }
  __syncthreads();
}

template <int N, int F, int NT, int P, int PEND, int S = N, bool TM = false>
__device__ __forceinline__ void lds_passes(float2* buf, const float2* tw2, int tid, const float2* twm = nullptr) {
  if constexpr (P < PEND) {
    lds_pass<N, F, NT, P, S, TM>(buf, tw2, tid, twm);
    lds_passes<N, F, NT, P + 1, PEND, S, TM>(buf, tw2, tid, twm);
  }
}

template <int N, int NT>
constexpr int fused_F() {
  constexpr int RL = Fp<N>::R(Fp<N>::np - 1);
  constexpr int G = (N / RL) / 2 + 1;
  int F = NT / G;
  if (F < 1) F = 1;
  if (F > 32) F = 32;
  return F;
}

// Pair tile and register budget measured fastest on MI355X for the ring lengths of the benchmark
// shapes: n = 308, 9 pairs at >= 3 wavefronts per SIMD (167 VGPRs, no spills, against 2
// wavefronts at ~190 VGPRs unconstrained); n = 192, 5 pairs; n = 400, 3 pairs with the reference
// prefetch (128 VGPRs, 4 wavefronts; 240 against 254 ms without it and 265 ms with 4 pairs, which
// drop to 3 wavefronts with it; 2 pairs 318 ms, 5 to 10 pairs slower); n = 500 (box 320), 4 pairs
// (150 VGPRs, 3 wavefronts: 202.8 ms per 300-particle search against 217.6 / 242.8 / 265.3 ms
// with 3 / 5 / 6); n = 240 (box 160), 4 pairs (96 VGPRs: 143.5 ms per 1000-particle search against
// 148.1 / 156.4 / 165.2 ms with 6 / 5 / 8); n = 600 (box 384), 2 pairs (281.2 ms per 200-particle
// search against 288.9 / 305.4 ms with 4 / 3).  fused_F elsewhere.
template <int N, int NT>
constexpr int fused_F_tuned() {
  if (NT == 256 && (N == 192 || N == 308 || N == 400)) return N == 192 ? 5 : N == 308 ? 9 : 3;
  if (NT == 256 && (N == 240 || N == 500)) return 4;
  if (NT == 256 && N == 600) return 2;
  return fused_F<N, NT>();
}
#ifndef PFTCC_FUSED_GROUP
#define PFTCC_FUSED_GROUP 16   // > 0: work-groups ordered tile-major over groups of this many particles, per XCD
#endif
#ifndef PFTCC_FUSED_PREFETCH
#define PFTCC_FUSED_PREFETCH 1 // the next shell's staged rows (|CTF|, reference rows, spectrum) loaded ahead
#endif
#ifndef PFTCC_FUSED_PREFETCH_REF
#define PFTCC_FUSED_PREFETCH_REF 1   // per-pair variant: pass-1 reference values loaded a shell ahead
#endif                               // (1: where no register budget is imposed, 2: always)

// Tile of the staged-reference variant.  Its pass 1 holds no per-pair reference pointers or norm
// sums: at N = 308 that leaves room for 11 pairs without the staging prefetch, and for 9 pairs
// with it (160 VGPRs; 11 pairs then spill), which was measured faster.
template <int N, int NT>
constexpr int fused_F_rst_tuned() {
  if (NT == 256 && N == 308) return PFTCC_FUSED_PREFETCH ? 9 : 11;
  if (NT == 256 && N == 400) return 4;
  if (NT == 256 && (N == 240 || N == 600)) return fused_F<N, NT>();   // not measured: the generic tile
  // n 500: 4 pairs as in the per-pair variant (742 against 781 ms for a 100-particle x 9-shift
  // search with 5)
  return fused_F_tuned<N, NT>();
}
template <int N, int NT>
constexpr int fused_W_tuned() {
  return (NT == 256 && N == 308) ? 3 : 1;
}
// The per-pair variant's reference prefetch holds 8 more registers per pass-1 task, so it is used
// where no register budget is imposed on the tile (at N = 308 it spills).
template <int N, int NT>
constexpr bool fused_prefetch_ref() {
  return PFTCC_FUSED_PREFETCH_REF == 2 || (PFTCC_FUSED_PREFETCH_REF == 1 && fused_W_tuned<N, NT>() == 1);
}

#ifndef PFTCC_FUSED_TWL
#define PFTCC_FUSED_TWL 1      // last-pass twiddles from a [t][g] table (conflict-free reads)
#endif
#ifndef PFTCC_FUSED_PFA
#define PFTCC_FUSED_PFA 1      // prime-factor shell transform where N = 4 A B allows it
#endif
#ifndef PFTCC_PFA_SA
#define PFTCC_PFA_SA 1         // prime-factor LDS layout strides (tuning builds; 0: the defaults)
#endif
#ifndef PFTCC_PFA_SB
#define PFTCC_PFA_SB 0
#endif
#ifndef PFTCC_PFA_QS
#define PFTCC_PFA_QS 0
#endif

__host__ __device__ constexpr int pfa_modinv(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 0;
}

// Prime-factor (Good-Thomas) plan of the shell transform, for N = 4 A B with A > B distinct odd
// primes (fused radix order 4, A, B) and N/4 = 1 (mod 4).  Input index n = CRT(n1, nA, nB)
// (n = n1 mod 4, nA mod A, nB mod B), output index m = (N/4 k1 + N/A kA + N/B kB) mod N: the three
// sub-transforms are plain DFTs of 4, A and B points with no twiddle factors between them.
// LDS layout of one transform: position k1 QS + a + b SB (a = nA or kA, b = nB or kB), with rows
// padded to SB = A + 2 and QS = B SB; every pass works in place on its own butterfly's positions.
// The padding spreads the scattered pass-1 stores (column (2j mod A, 2j mod B) of butterfly j)
// over the LDS banks; passes A and B read rows and columns at unit and SB strides.
template <int N>
struct PfaPlan {
  static constexpr bool ok = Fp<N>::np == 3 && Fp<N>::R(0) == 4 && Fp<N>::R(1) % 2 == 1 &&
                             Fp<N>::R(2) % 2 == 1 && Fp<N>::R(1) != Fp<N>::R(2) && (N / 4) % 4 == 1;
  static constexpr int A = ok ? Fp<N>::R(1) : 1;
  static constexpr int B = ok ? Fp<N>::R(2) : 1;
  static constexpr int Q = N / 4;
  static constexpr int SA = PFTCC_PFA_SA;     // stride of a
  static constexpr int SB = PFTCC_PFA_SB > 0 ? PFTCC_PFA_SB : A + 2;   // stride of b
  static constexpr int QS = PFTCC_PFA_QS > 0 ? PFTCC_PFA_QS : B * SB;  // stride of k1
  static_assert(!ok || (((SA == 1 && SB >= A) || (SB == 1 && SA >= B)) && QS >= (A - 1) * SA + (B - 1) * SB + 1),
                "prime-factor layout: distinct positions within a row");
  static constexpr int IA = ok ? pfa_modinv((N / A) % A, A) : 0;   // kA = m IA mod A
  static constexpr int IB = ok ? pfa_modinv((N / B) % B, B) : 0;   // kB = m IB mod B
  // LDS position of output bin m (k1 = m mod 4 since N/4 = 1 mod 4); the Nyquist bin N after them
  static __device__ __forceinline__ int pos(int m) {
    if (m >= N) return 4 * QS;
    return (m & 3) * QS + ((m * IA) % A) * SA + ((m * IB) % B) * SB;
  }
  // Butterfly g of the last pass, 0 <= g < 2 A + 1: (k1, kA) and its mirror (-k1, -kA), whose
  // outputs hold the bins N - m.  g = 0: (0, 0), g = 2: (2, 0) are their own mirrors; g = 1:
  // (1, 0) with (3, 0); g = 3 + i: k1 = i / H, kA = 1 + i % H with (-k1, A - kA), H = (A - 1) / 2.
  static __device__ __forceinline__ void rep(int g, int& k1, int& ka) {
    constexpr int H = (A - 1) / 2;
    if (g < 3) {
      k1 = g;
      ka = 0;
    } else {
      k1 = (g - 3) / H;
      ka = 1 + (g - 3) % H;
    }
  }
};

// Pass A of the prime-factor transform: DFT of A points along a for every (k1, nB), in place.
template <int N, int F, int NT, int S>
__device__ __forceinline__ void pfa_pass_a(float2* __restrict__ buf, int tid) {
  constexpr int A = PfaPlan<N>::A, B = PfaPlan<N>::B, QS = PfaPlan<N>::QS, SA = PfaPlan<N>::SA, SB = PfaPlan<N>::SB;
  constexpr int TASKS = F * 4 * B;
  constexpr int TPT = (TASKS + NT - 1) / NT;
#pragma unroll
  for (int c = 0; c < TPT; ++c) {
    const int task = tid + c * NT;
    if (TASKS % NT == 0 || task < TASKS) {
      const int f = task / (4 * B), r = task % (4 * B);
      float2* p = buf + f * S + (r / B) * QS + (r % B) * SB;
      float2 v[A];
#pragma unroll
      for (int t = 0; t < A; ++t) v[t] = p[t * SA];
      Dft<A>::run(v);
#pragma unroll
      for (int t = 0; t < A; ++t) p[t * SA] = v[t];
    }
  }
}
#ifndef PFTCC_FUSED_P1V4
#define PFTCC_FUSED_P1V4 1     // pass-1 outputs stored as two 16-byte writes
#endif
#ifndef PFTCC_FUSED_W12
#define PFTCC_FUSED_W12 1      // split folded into two staged spectra W1, W2 (see k_corr_fused)
#endif
#ifndef PFTCC_FUSED_PAD
#define PFTCC_FUSED_PAD 0      // extra float2 between the transforms of a tile (even: 16-byte rows)
#endif

// Prime-factor shell transform in the fused kernel (no twiddles between the passes).
template <int N>
constexpr bool fused_pfa() {
  return PFTCC_FUSED_PFA && PFTCC_FUSED_W12 && PfaPlan<N>::ok;
}

// LDS row stride of one transform of the fused kernel.
template <int N>
constexpr int fused_stride() {
  return fused_pfa<N>() ? 4 * PfaPlan<N>::QS : N + PFTCC_FUSED_PAD;
}

// Length of each staged particle spectrum (W1, W2) in LDS: N + 1 bins, in the prime-factor layout
// with its padding.
template <int N>
constexpr int fused_wlen() {
  return fused_pfa<N>() ? 4 * PfaPlan<N>::QS + 1 : N + 1;
}

// Entries of the last pass's twiddle table: (RL - 1) x NsL.
template <int N>
constexpr int fused_twl_len() {
  constexpr int RL = Fp<N>::R(Fp<N>::np - 1);
  return PFTCC_FUSED_TWL ? (RL - 1) * (N / RL) : 0;
}

// Entries of the middle passes' twiddle tables (Stockham passes 1 .. np - 2), 0 when unused.
// Measured at n 400 (radix-5 pass with Ns = 20: tw2's index stride 8 jm t put 16 lanes on two
// banks): 239.9 -> 223.0 ms per 600-particle search; n 192 not faster (192.3 against 193.7 ms).
#ifndef PFTCC_FUSED_TWM
#define PFTCC_FUSED_TWM 1      // middle-pass twiddles from [t][jm] tables (1: at n 400, 2: wherever Stockham)
#endif
template <int N>
constexpr bool twm_on() {
  return PFTCC_FUSED_TWM == 2 || (PFTCC_FUSED_TWM == 1 && N == 400);
}
template <int N>
constexpr int fused_twm_len() {
  return (twm_on<N>() && !fused_pfa<N>()) ? twm_off<N>(Fp<N>::np - 1) : 0;
}

template <int N, int F>
constexpr size_t fused_lds_bytes() {
  // buf[F N] + Wk[N+1] (W1, W2: 2 (N+1)) + tw2[2N] + twL float2;  ctfk[N] + sq/off slots 2 F N/4 floats
  return sizeof(float2) * ((size_t)F * fused_stride<N>() + (PFTCC_FUSED_W12 ? 2 : 1) * fused_wlen<N>() + 2 * N +
                           fused_twl_len<N>() + fused_twm_len<N>() + 2 * N) +
         sizeof(float) * ((size_t)N + (size_t)F * (N / 4) * 2);
}

// Loads from an array at a uniform base by a 32-bit byte offset, which lets the load take the
// base from scalar registers: ld16 reads elements elem and elem + 1 of a float2 array.
__device__ __forceinline__ float4 ld16(const float2* __restrict__ base, unsigned int elem) {
  return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(base) + (elem << 3));
}
__device__ __forceinline__ float2 ld8(const float2* __restrict__ base, unsigned int elem) {
  return *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(base) + (elem << 3));
}
__device__ __forceinline__ float ldf(const float* __restrict__ base, unsigned int elem) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + (elem << 2));
}

// Value select (a conditional on array lvalues would select an address and force the array
// out of registers).
__device__ __forceinline__ float2 vsel(bool c, float2 a, float2 b) {
  return make_float2(c ? a.x : b.x, c ? a.y : b.y);
}

// Staged particle spectra of bin position wp: W1 at wp and W2 at WL + wp, or (I: the prime-factor
// order) both at float4 position wp.
template <bool I, int WL>
__device__ __forceinline__ void wk_store(float2* Wk, int wp, float2 w1, float2 w2) {
  if constexpr (I) {
    reinterpret_cast<float4*>(Wk)[wp] = make_float4(w1.x, w1.y, w2.x, w2.y);
  } else {
    Wk[wp] = w1;
    Wk[WL + wp] = w2;
  }
}

// acc += conj(Z) w1 + Zn w2: the split and the product conj(V) W of one bin with the split's
// m-dependent factors folded into w1 = conj(A') W / 2 and w2 = conj(B') W / 2.
__device__ __forceinline__ void acc_w12(float2& acc, float2 Z, float2 Zn, float2 w1, float2 w2) {
  acc.x = fmaf(Z.x, w1.x, acc.x);
  acc.x = fmaf(Z.y, w1.y, acc.x);
  acc.y = fmaf(Z.x, w1.y, acc.y);
  acc.y = fmaf(-Z.y, w1.x, acc.y);
  acc.x = fmaf(Zn.x, w2.x, acc.x);
  acc.x = fmaf(-Zn.y, w2.y, acc.x);
  acc.y = fmaf(Zn.x, w2.y, acc.y);
  acc.y = fmaf(Zn.y, w2.x, acc.y);
}

// 2 V[m] of the real FFT of length 2N from the packed FFT values A = Zc[m], B = Zc[N-m].  The
// factor 1/2 of the split is applied to the particle spectrum instead (W_k / 2 when staged):
// scaling by powers of two is exact, so conj(2V) (W/2) rounds exactly as conj(V) W.
__device__ __forceinline__ float2 split_bin2(float2 A, float2 B, float2 w) {
  const float2 sum = make_float2(A.x + B.x, A.y - B.y);
  const float2 dif = make_float2(A.x - B.x, A.y + B.y);
  const float2 wd = cmul(w, dif);
  return make_float2(sum.x + wd.y, sum.y - wd.x);
}

// Y[m] of the inverse real FFT from X[m] and X[N-m]; e = exp(+i pi m / N).  Returns conj(Y).
__device__ __forceinline__ float2 unsplit_bin_conj(float2 X, float2 Xn, float2 e) {
  const float2 s = make_float2(X.x + Xn.x, X.y - Xn.y);
  const float2 d = make_float2(X.x - Xn.x, X.y + Xn.y);
  const float2 ed = cmul(e, d);
  return make_float2(s.x - ed.y, -(s.y + ed.x));
}

// W: minimum wavefronts per SIMD the register allocation must allow (1: no constraint).
// RST: the tile's pairs use at most two references (shifts >= F per reference), CC objective, no
// shell normalisation: the CTF-weighted reference rings of the shell are staged once per
// work-group in LDS and each reference's norm is summed from them (|ref x CTF x shmat|^2 =
// |ref x CTF|^2), instead of per pair.
template <int N, int F, int NT, int W, bool RST>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(W))) k_corr_fused(CorrArgs a) {
  static_assert(Fp<N>::fusable, "first radix must be 4");
  static_assert(NT % 64 == 0, "whole wavefronts");
  constexpr int NP = Fp<N>::np;
  constexpr int RL = Fp<N>::R(NP - 1);      // last radix
  constexpr int NsL = N / RL;               // product of the other radices
  constexpr int G = NsL / 2 + 1;            // mirrored butterfly groups of the last pass
  constexpr int TL = F * G;
  constexpr int TPTL = (TL + NT - 1) / NT;
  constexpr int NB = 2 * RL + 1;            // bins per group: pair 2 RL; group 0: RL + Nyquist
  constexpr int Q = N / 4;                  // pass-1 butterflies per transform
  constexpr int T1 = F * Q;
  constexpr int TPT1 = (T1 + NT - 1) / NT;
  // prime-factor shell transform (no twiddles between the passes); Stockham passes otherwise
  constexpr bool PFA = fused_pfa<N>();
  constexpr int PA = PfaPlan<N>::A, PB = PfaPlan<N>::B, QS = PfaPlan<N>::QS, SA = PfaPlan<N>::SA, SB = PfaPlan<N>::SB;
  constexpr int WL = fused_wlen<N>();        // W1 at Wk[0, WL), W2 at Wk[WL, 2 WL); Nyquist bin at WL - 1
  static_assert(!PFA || (G == 2 * PA + 1 && RL == PB), "prime-factor plan");

  extern __shared__ __align__(16) float2 smem[];
  float2* buf = smem;                        // [F][N]
  constexpr int S = fused_stride<N>();       // transform stride in buf
  static_assert(S % 2 == 0, "16-byte aligned rows");
  float2* Wk = buf + F * S;                  // [N+1]  W_k / 2, or W1 [N+1] then W2 [N+1]
  float2* tw2 = Wk + (PFTCC_FUSED_W12 ? 2 : 1) * WL;   // [2N]   exp(-i pi a / N)
  // prime-factor order: W1 and W2 of a bin interleaved, one 16-byte read per bin in the last pass
  const float4* Wk4 = reinterpret_cast<const float4*>(Wk);
  static_assert(!PFA || (F * S) % 2 == 0, "16-byte aligned staged spectra");
  float2* twL = tw2 + 2 * N;                 // [RL-1][NsL]  twL[(t-1) NsL + g] = tw2[2 g t]
  float2* twM = twL + fused_twl_len<N>();    // middle-pass twiddle tables (Stockham sizes)
  float2* rck = twM + fused_twm_len<N>();    // [2][N] ref x |CTF| of the tile's references (RST)
  float* ctfk = reinterpret_cast<float*>(rck + 2 * N);   // [N]
  float* sqs = ctfk + N;                     // [F][Q]
  float* offs = sqs + F * Q;                 // [F][Q]
  __shared__ float sq[F], off[F];
  __shared__ float pscal;
  __shared__ double rsq[2][NT / 64];
  __shared__ unsigned long long red[(NT + 63) / 64];

  const int tid = threadIdx.x;
  int pi, tile;
  if constexpr (PFTCC_FUSED_GROUP > 0) {
    // each XCD takes a contiguous range of the logical order; within it, the tiles of a group of
    // particles are interleaved, so that the work-groups resident together share reference and
    // shift rows in their L2
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int grp = t / (PFTCC_FUSED_GROUP * a.ntiles);
    const int base = grp * PFTCC_FUSED_GROUP;
    const int gsz = min(PFTCC_FUSED_GROUP, a.np - base);
    const int r = t - grp * PFTCC_FUSED_GROUP * a.ntiles;
    tile = r / gsz;
    pi = base + (r - tile * gsz);
  } else {
    pi = blockIdx.x / a.ntiles;
    tile = blockIdx.x - pi * a.ntiles;
  }
  const int ip = a.plist[pi];
  const bool even = a.iseven[ip] != 0;
  const float2* refs = even ? a.refs_even : a.refs_odd;
  const int npairs = a.nr * a.ns;
  const int nk = a.nk;
  // the Euclid objective (never in the staged-reference variant, which runs the CC objective)
  const bool EUC = !RST && a.mode == 1;

  for (int i = tid; i < 2 * N; i += NT) tw2[i] = a.tw2g[i];
  if constexpr (PFTCC_FUSED_TWL && !PFA)
    for (int i = tid; i < (RL - 1) * NsL; i += NT) {
      const int t = i / NsL + 1, g = i - (t - 1) * NsL;
      twL[i] = a.tw2g[2 * g * t];
    }
  constexpr bool TWM = fused_twm_len<N>() > 0;
  if constexpr (TWM)
    for (int q = 1; q < NP - 1; ++q) {
      const int R = Fp<N>::R(q), ns = Fp<N>::Ns(q), tstep = N / (ns * R);
      for (int i = tid; i < (R - 1) * ns; i += NT) {
        const int t = i / ns + 1, jm = i - (t - 1) * ns;
        twM[twm_off<N>(q) + i] = a.tw2g[2 * jm * t * tstep];
      }
    }
  // prime-factor index tables, in the last-pass twiddle table's place (read once per shell, so
  // that no per-lane index is held in registers across the shell loop):
  //   p1tab[j]: pass-1 butterfly j: LDS column jb (bits 0-7), offset ob of its second q pair
  //             (bits 8-17), sign of z[b] (bit 30) and of z[b + Q] (bit 31)
  //   p3tab[g]: last-pass butterfly g: positions o1 (bits 0-8) and o2 (bits 9-17), own mirror (bit 30)
  //   wptab[m]: LDS position of spectrum bin m, 0 <= m <= N
  int* const p1tab = reinterpret_cast<int*>(twL);
  int* const p3tab = p1tab + Q;
  int* const wptab = p3tab + G;
  static_assert(!PFA || (size_t)(Q + G + N + 1) * sizeof(int) <= sizeof(float2) * fused_twl_len<N>(),
                "prime-factor tables fit the twiddle table's space");
  static_assert(!PFA || (2 * N < 1024 && 4 * QS < 512 && QS < 256), "prime-factor table fields");
  if constexpr (PFA) {
    for (int j = tid; j < Q; j += NT) {
      const bool lo = 2 * j < Q;
      const int ob = lo ? 4 * j + 2 * Q : 4 * j - 2 * Q;
      const int jb = ((2 * j) % PA) * SA + ((2 * j) % PB) * SB;
      p1tab[j] = jb | (ob << 8) | ((j & 1) << 30) | ((int)((j & 1) != (lo ? 0 : 1)) << 31);
    }
    for (int g = tid; g < G; g += NT) {
      int k1, ka;
      PfaPlan<N>::rep(g, k1, ka);
      const int o1 = k1 * QS + ka * SA;
      const int o2 = ((4 - k1) & 3) * QS + (ka == 0 ? 0 : PA - ka) * SA;
      p3tab[g] = o1 | (o2 << 9) | ((int)(ka == 0 && (k1 & 1) == 0) << 30);
    }
    for (int m = tid; m <= N; m += NT) wptab[m] = PfaPlan<N>::pos(m);
  }

  // RST: the tile's references and their rows
  int rr0 = 0;
  const float2* rs0 = refs;
  const float2* rs1 = refs;
  if constexpr (RST) {
    rr0 = (tile * F) / a.ns;
    const int rr1 = min(rr0 + 1, a.nr - 1);
    // uniform across the work-group: row numbers in scalar registers
    const int row0 = __builtin_amdgcn_readfirstlane(a.rlist ? a.rlist[rr0] : rr0);
    const int row1 = __builtin_amdgcn_readfirstlane(a.rlist ? a.rlist[rr1] : rr1);
    rs0 = refs + (size_t)row0 * nk * N;
    rs1 = refs + (size_t)row1 * nk * N;
  }
  double rsum0 = 0.0, rsum1 = 0.0;     // this thread's share of the two references' norms
  // pass-1 slots: pair, reference and shift rows are fixed for all shells
  float sqp[TPT1], offp[TPT1];
  // element offsets (32 bits) of the pair's reference and shift rows from the uniform bases refs
  // and a.shtab; the shell's rows follow at + kk N
  unsigned int roff[TPT1], soff[TPT1];
  int slotref[TPT1];
#pragma unroll
  for (int c = 0; c < TPT1; ++c) {
    const int task = tid + c * NT;
    const int f = task / Q;
    const int P = tile * F + f;
    const bool slotok = task < T1 && P < npairs;
    sqp[c] = 0.f;
    offp[c] = 0.f;
    roff[c] = 0u;
    soff[c] = 0u;
    slotref[c] = 0;
    if (slotok) {
      const int ir = P / a.ns, is = P % a.ns;
      const int iref = a.rlist ? a.rlist[(size_t)pi * a.rlist_ld + ir] : ir;
      slotref[c] = iref;
      roff[c] = (unsigned int)iref * (unsigned int)(nk * N);
      soff[c] = (unsigned int)is * (unsigned int)(nk * N);
    }
  }

  float2 acc[TPTL][NB];
#pragma unroll
  for (int c = 0; c < TPTL; ++c)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[c][b] = make_float2(0.f, 0.f);

  const int kloop = a.kcorr > a.knorm ? a.kcorr : a.knorm;
  // per-pair variant: the reference values of this thread's pass-1 butterflies for the next shell,
  // loaded while the current shell runs (element offsets oa, ob of the butterfly's q pairs)
  constexpr bool PRF = !RST && fused_prefetch_ref<N, NT>();
  float4 pra[PRF ? TPT1 : 1], prb[PRF ? TPT1 : 1];
  auto p1_offsets = [&](int j, int& oa, int& ob) {
    if constexpr (PFA) {
      oa = 4 * j;
      ob = 2 * j < Q ? 4 * j + 2 * Q : 4 * j - 2 * Q;
    } else {
      oa = 2 * j;
      ob = 2 * j + N / 2;
    }
  };
  if constexpr (PRF)
    if (kloop > 0)
#pragma unroll
      for (int c = 0; c < TPT1; ++c) {
        const int task = tid + c * NT;
        if (!(T1 % NT == 0 || task < T1)) continue;
        int oa, ob;
        p1_offsets(task % Q, oa, ob);
        pra[c] = ld16(refs, roff[c] + oa);
        prb[c] = ld16(refs, roff[c] + ob);
      }
  // the staging operands of shell kk + 1, loaded while shell kk runs
  constexpr int JP = (N + NT - 1) / NT, MP = (N + 1 + NT - 1) / NT;
  float pcv[JP];
  float2 pr0[JP], pr1[JP], pwm[MP];
  // the particle's |CTF| and memoized spectra: uniform bases, 32-bit offsets within the particle
  const float* ctfp = a.ctf ? a.ctf + (size_t)ip * nk * N : nullptr;
  const float2* memop = a.memo + (size_t)ip * nk * (N + 1);
  auto fetch_stage = [&](int k) {
#pragma unroll
    for (int u = 0; u < JP; ++u) {
      const int j = tid + u * NT;
      if (j < N) {
        pcv[u] = ctfp ? ldf(ctfp, (unsigned int)(k * N + j)) : 1.0f;
        if constexpr (RST) {
          pr0[u] = ld8(rs0, (unsigned int)(k * N + j));
          pr1[u] = ld8(rs1, (unsigned int)(k * N + j));
        }
      }
    }
    if (k < a.kcorr)
#pragma unroll
      for (int u = 0; u < MP; ++u) {
        const int m = tid + u * NT;
        if (m <= N) pwm[u] = ld8(memop, (unsigned int)(k * (N + 1) + m));
      }
  };
  if constexpr (PFTCC_FUSED_PREFETCH)
    if (kloop > 0) fetch_stage(0);
  for (int kk = 0; kk < kloop; ++kk) {
    const bool docorr = kk < a.kcorr;
    // the staged-reference variant runs for shifted CC searches, whose norms cover every shell
    // (knorm = nk = kloop)
    const bool donorm = RST || kk < a.knorm;
    __syncthreads();   // the previous shell's last pass has finished with buf and Wk
    const size_t prow = (size_t)ip * nk + kk;
    if constexpr (PFTCC_FUSED_PREFETCH) {
      // the prefetched operands go to LDS, then the next shell's loads are issued
      const float wkp = (EUC && docorr) ? a.shellw[prow] : 1.0f;
#pragma unroll
      for (int u = 0; u < JP; ++u) {
        const int j = tid + u * NT;
        if (j < N) {
          if constexpr (RST) {
            const float cv = pcv[u];
            const float2 q0 = make_float2(pr0[u].x * cv, pr0[u].y * cv), q1 = make_float2(pr1[u].x * cv, pr1[u].y * cv);
            rck[j] = q0;
            rck[N + j] = q1;
            if (donorm) {
              rsum0 += (double)(q0.x * q0.x + q0.y * q0.y);
              rsum1 += (double)(q1.x * q1.x + q1.y * q1.y);
            }
          } else {
            ctfk[j] = pcv[u];
          }
        }
      }
      if (docorr)
#pragma unroll
        for (int u = 0; u < MP; ++u) {
          const int m = tid + u * NT;
          if (m <= N) {
            float2 wh = make_float2(0.5f * pwm[u].x, 0.5f * pwm[u].y);
            if constexpr (PFTCC_FUSED_W12) {
              if (EUC) wh = make_float2(wkp * wh.x, wkp * wh.y);
              const float2 t = tw2[m];
              const int wp = PFA ? wptab[m] : m;
              wk_store<PFA, WL>(Wk, wp, cmul(make_float2(1.f + t.y, t.x), wh), cmul(make_float2(1.f - t.y, -t.x), wh));
            } else {
              Wk[m] = wh;
            }
          }
        }
      if (kk + 1 < kloop) fetch_stage(kk + 1);
    } else if constexpr (RST) {
      for (int j = tid; j < N; j += NT) {
        const float cv = a.ctf ? a.ctf[prow * N + j] : 1.0f;
        const float2 r0 = rs0[(size_t)kk * N + j], r1 = rs1[(size_t)kk * N + j];
        const float2 q0 = make_float2(r0.x * cv, r0.y * cv), q1 = make_float2(r1.x * cv, r1.y * cv);
        rck[j] = q0;
        rck[N + j] = q1;
        if (donorm) {
          rsum0 += (double)(q0.x * q0.x + q0.y * q0.y);
          rsum1 += (double)(q1.x * q1.x + q1.y * q1.y);
        }
      }
    } else {
      for (int j = tid; j < N; j += NT) ctfk[j] = a.ctf ? a.ctf[prow * N + j] : 1.0f;
    }
    const float wk = (EUC && docorr) ? a.shellw[prow] : 1.0f;
    if (!PFTCC_FUSED_PREFETCH && docorr)
      for (int m = tid; m <= N; m += NT) {
        const float2 w = a.memo[prow * (N + 1) + m];
        float2 wh = make_float2(0.5f * w.x, 0.5f * w.y);
        if constexpr (PFTCC_FUSED_W12) {
          // conj(V) W = conj(Z[m]) conj(A) W + Z[N-m] conj(B) W with 2V = A' Z[m] + B' conj(Z[N-m]),
          // A' = 1 - i w, B' = 1 + i w, w = exp(-i pi m / N): both factors depend on m only, so they
          // are applied to the staged spectrum (with the Euclid shell weight) once per shell
          if (EUC) wh = make_float2(wk * wh.x, wk * wh.y);
          const float2 t = tw2[m];
          const int wp = PFA ? wptab[m] : m;
          // conj(A') W / 2 and conj(B') W / 2
          wk_store<PFA, WL>(Wk, wp, cmul(make_float2(1.f + t.y, t.x), wh), cmul(make_float2(1.f - t.y, -t.x), wh));
        } else {
          Wk[m] = wh;
        }
      }
    __syncthreads();
    // ---- pass 1: z from q = ref * |CTF| [* shift] [* shell scale]; radix-4 DFT (Ns = 1) ----
#pragma unroll
    for (int c = 0; c < TPT1; ++c) {
      const int task = tid + c * NT;
      if (!(T1 % NT == 0 || task < T1)) continue;
      const int f = task / Q, j = task % Q;
      float2 y[4];
      // q pairs of the butterfly: Stockham z[j], z[j + Q] (+ their mirrors z[j + 2Q], z[j + 3Q]);
      // prime-factor: z[b], z[b + Q] with b = 2j or 2j + 2Q, i.e. q at 4j and 4j +- 2Q
      const int p1 = PFA ? p1tab[j] : 0;
      const int oa = PFA ? 4 * j : 2 * j;
      const int ob = PFA ? (p1 >> 8) & 1023 : 2 * j + N / 2;
      // a slot past the tile's last pair reads valid rows (reference 0, shift 0, the second staged
      // row) and its results are never read
      {
        float2 q0, q1, q2, q3;
        if constexpr (RST) {
          const float2* rc = rck + min(((tile * F + f) / a.ns) - rr0, 1) * N;
          const float4 ra = *reinterpret_cast<const float4*>(rc + oa);
          const float4 rb = *reinterpret_cast<const float4*>(rc + ob);
          q0 = make_float2(ra.x, ra.y);
          q1 = make_float2(ra.z, ra.w);
          q2 = make_float2(rb.x, rb.y);
          q3 = make_float2(rb.z, rb.w);
        } else {
          float4 ra, rb;
          if constexpr (PRF) {
            ra = pra[c];
            rb = prb[c];
            if (kk + 1 < kloop) {
              const unsigned int r = roff[c] + (unsigned int)((kk + 1) * N);
              pra[c] = ld16(refs, r + oa);
              prb[c] = ld16(refs, r + ob);
            }
          } else {
            const unsigned int r = roff[c] + (unsigned int)(kk * N);
            ra = ld16(refs, r + oa);
            rb = ld16(refs, r + ob);
          }
          const float2 c01 = *reinterpret_cast<const float2*>(ctfk + oa);
          const float2 c23 = *reinterpret_cast<const float2*>(ctfk + ob);
          q0 = make_float2(ra.x * c01.x, ra.y * c01.x);
          q1 = make_float2(ra.z * c01.y, ra.w * c01.y);
          q2 = make_float2(rb.x * c23.x, rb.y * c23.x);
          q3 = make_float2(rb.z * c23.y, rb.w * c23.y);
        }
        if (RST || a.shtab) {   // the staged-reference variant runs for shifted searches only
          const unsigned int r = soff[c] + (unsigned int)(kk * N);
          const float4 sa = ld16(a.shtab, r + oa);
          const float4 sb = ld16(a.shtab, r + ob);
          q0 = cmul(q0, make_float2(sa.x, sa.y));
          q1 = cmul(q1, make_float2(sa.z, sa.w));
          q2 = cmul(q2, make_float2(sb.x, sb.y));
          q3 = cmul(q3, make_float2(sb.z, sb.w));
        }
        if (!RST && a.refinv && docorr) {
          const int iref = slotref[c];
          const int row = a.optrow ? a.optrow[ip] : iref;
          const float sc = a.ofac[(size_t)row * a.kcorr + kk] *
                           a.refinv[((size_t)(even ? 0 : 1) * a.nrefs + iref) * a.kcorr + kk];
          q0 = cscale(q0, sc);
          q1 = cscale(q1, sc);
          q2 = cscale(q2, sc);
          q3 = cscale(q3, sc);
        }
        if constexpr (!RST) {
          const float e2 = q0.x * q0.x + q0.y * q0.y + q1.x * q1.x + q1.y * q1.y + q2.x * q2.x + q2.y * q2.y +
                           q3.x * q3.x + q3.y * q3.y;
          if (donorm) sqp[c] += e2;
          if (EUC && docorr) offp[c] = fmaf(wk, e2, offp[c]);
        }
        if constexpr (PFA) {
          // z[b + Q n1], n1 = 0..3: signs +-1 from b >= 2Q (j odd) and b + Q >= 2Q (2j >= Q)
          const float sa = __int_as_float((int)(((unsigned)p1 << 1) & 0x80000000u) | 0x3f800000);
          const float sb = __int_as_float((int)((unsigned)p1 & 0x80000000u) | 0x3f800000);
          y[0] = make_float2(fmaf(sa, q0.y, q0.x), fmaf(sa, q1.y, q1.x));
          y[1] = make_float2(fmaf(sb, q2.y, q2.x), fmaf(sb, q3.y, q3.x));
          y[2] = make_float2(fmaf(-sa, q0.y, q0.x), fmaf(-sa, q1.y, q1.x));
          y[3] = make_float2(fmaf(-sb, q2.y, q2.x), fmaf(-sb, q3.y, q3.x));
        } else {
          // z[j], z[j+Q], z[j+2Q], z[j+3Q]
          y[0] = make_float2(q0.x + q0.y, q1.x + q1.y);
          y[1] = make_float2(q2.x + q2.y, q3.x + q3.y);
          y[2] = make_float2(q0.x - q0.y, q1.x - q1.y);
          y[3] = make_float2(q2.x - q2.y, q3.x - q3.y);
        }
      }
      if (docorr) {
        Dft<4>::run(y);
        if constexpr (PFA) {
          // outputs k1 at k1 Q + (nA, nB) of b: nA = 2j mod A, nB = 2j mod B
          const int jb = p1 & 255;
#pragma unroll
          for (int t = 0; t < 4; ++t) buf[f * S + t * QS + jb] = y[t];
        } else if constexpr (PFTCC_FUSED_P1V4) {
          float4* dst = reinterpret_cast<float4*>(buf + f * S + 4 * j);   // 16-byte aligned: S even
          dst[0] = make_float4(y[0].x, y[0].y, y[1].x, y[1].y);
          dst[1] = make_float4(y[2].x, y[2].y, y[3].x, y[3].y);
        } else {
#pragma unroll
          for (int t = 0; t < 4; ++t) buf[f * S + 4 * j + t] = y[t];
        }
      }
    }
    if (!docorr) continue;
    __syncthreads();
    // ---- middle passes in LDS ----
    if constexpr (PFA) {
      pfa_pass_a<N, F, NT, S>(buf, tid);
      __syncthreads();
    } else {
      lds_passes<N, F, NT, 1, NP - 1, S, TWM>(buf, tw2, tid, twM);
    }
    // ---- last pass (radix RL) on mirrored butterflies + split + conj(V) W_k ----
    if constexpr (PFA) {
#pragma unroll
      for (int c = 0; c < TPTL; ++c) {
        const int task = tid + c * NT;
        if (!(TL % NT == 0 || task < TL)) continue;
        const int f = task / G, g = task % G;
        const int p3 = p3tab[g];
        const bool self = (p3 >> 30) & 1;
        const int o1 = p3 & 511;              // positions of the bins m(k1, kA, t)
        const int o2 = (p3 >> 9) & 511;       // and of m(-k1, -kA, t)
        // a butterfly that is its own mirror has o2 = o1, so y2 = y1 there: the bin N - m of y1[t]
        // is y2[-t] in every case, and only the bins of y1 are accumulated.
        float2 y1[RL], y2[RL];
#pragma unroll
        for (int t = 0; t < RL; ++t) y1[t] = buf[f * S + o1 + t * SB];
        Dft<RL>::run(y1);
#pragma unroll
        for (int t = 0; t < RL; ++t) y2[t] = buf[f * S + o2 + t * SB];
        Dft<RL>::run(y2);
        // Wk is staged in the same order.  Each bin is accumulated together with its mirror, so
        // that the staged values are used as read.
#pragma unroll
        for (int t = 0; t < RL; ++t) {
          const int tn = (RL - t) % RL;
          const float4 wa = Wk4[o1 + t * SB], wb = Wk4[o2 + tn * SB];
          acc_w12(acc[c][t], y1[t], y2[tn], make_float2(wa.x, wa.y), make_float2(wa.z, wa.w));
          if (!self) acc_w12(acc[c][RL + tn], y2[tn], y1[t], make_float2(wb.x, wb.y), make_float2(wb.z, wb.w));
        }
        if (g == 0) {
          const float4 wn = Wk4[WL - 1];
          acc_w12(acc[c][2 * RL], y1[0], y1[0], make_float2(wn.x, wn.y), make_float2(wn.z, wn.w));
        }
      }
      continue;
    }
#pragma unroll
    for (int c = 0; c < TPTL; ++c) {
      const int task = tid + c * NT;
      if (!(TL % NT == 0 || task < TL)) continue;
      const int f = task / G, g = task % G;
      const bool pair = g != 0 && 2 * g != NsL;
      const int j2 = NsL - g;
      float2 y1[RL], y2[RL];
#pragma unroll
      for (int t = 0; t < RL; ++t) y1[t] = buf[f * S + g + t * NsL];
#pragma unroll
      for (int t = 1; t < RL; ++t) y1[t] = cmul(y1[t], PFTCC_FUSED_TWL ? twL[(t - 1) * NsL + g] : tw2[2 * g * t]);
      Dft<RL>::run(y1);
      if (pair) {
#pragma unroll
        for (int t = 0; t < RL; ++t) y2[t] = buf[f * S + j2 + t * NsL];
#pragma unroll
        for (int t = 1; t < RL; ++t) y2[t] = cmul(y2[t], PFTCC_FUSED_TWL ? twL[(t - 1) * NsL + j2] : tw2[2 * j2 * t]);
        Dft<RL>::run(y2);
      } else {
#pragma unroll
        for (int t = 0; t < RL; ++t) y2[t] = make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int t = 0; t < RL; ++t) {
        const int m = g + t * NsL;
        const float2 b0 = y1[(RL - t) % RL], b1 = y2[RL - 1 - t], b2 = y1[RL - 1 - t];
        const float2 B = vsel(g == 0, b0, vsel(pair, b1, b2));
        if constexpr (PFTCC_FUSED_W12) {
          acc_w12(acc[c][t], y1[t], B, Wk[m], Wk[WL + m]);
        } else {
          const float2 V = split_bin2(y1[t], B, tw2[m]);
          const float2 p = cmulcj(V, Wk[m]);
          acc[c][t].x = fmaf(wk, p.x, acc[c][t].x);
          acc[c][t].y = fmaf(wk, p.y, acc[c][t].y);
        }
      }
      if (pair) {
#pragma unroll
        for (int t = 0; t < RL; ++t) {
          const int m = j2 + t * NsL;
          if constexpr (PFTCC_FUSED_W12) {
            acc_w12(acc[c][RL + t], y2[t], y1[RL - 1 - t], Wk[m], Wk[WL + m]);
          } else {
            const float2 V = split_bin2(y2[t], y1[RL - 1 - t], tw2[m]);
            const float2 p = cmulcj(V, Wk[m]);
            acc[c][RL + t].x = fmaf(wk, p.x, acc[c][RL + t].x);
            acc[c][RL + t].y = fmaf(wk, p.y, acc[c][RL + t].y);
          }
        }
      }
      if (g == 0) {
        if constexpr (PFTCC_FUSED_W12) {
          acc_w12(acc[c][2 * RL], y1[0], y1[0], Wk[N], Wk[WL + N]);
        } else {
          const float2 V = split_bin2(y1[0], y1[0], tw2[N]);
          const float2 p = cmulcj(V, Wk[N]);
          acc[c][2 * RL].x = fmaf(wk, p.x, acc[c][2 * RL].x);
          acc[c][2 * RL].y = fmaf(wk, p.y, acc[c][2 * RL].y);
        }
      }
    }
  }
  __syncthreads();
  // ---- normalisation sums (fixed order) ----
#pragma unroll
  for (int c = 0; c < TPT1; ++c) {
    const int task = tid + c * NT;
    if (task < T1) {
      sqs[task] = sqp[c];
      offs[task] = offp[c];
    }
  }
  // ---- Y = pre-processed inverse input, conj'd, from the register spectra ----
  if constexpr (PFA) {
#pragma unroll
    for (int c = 0; c < TPTL; ++c) {
      const int task = tid + c * NT;
      if (!(TL % NT == 0 || task < TL)) continue;
      const int f = task / G, g = task % G;
      int k1, ka;
      PfaPlan<N>::rep(g, k1, ka);
      const bool self = ka == 0 && (k1 & 1) == 0;
      const int mb1 = (Q * k1 + (N / PA) * ka) % N;
      const int mb2 = (Q * ((4 - k1) & 3) + (N / PA) * (ka == 0 ? 0 : PA - ka)) % N;
#pragma unroll
      for (int t = 0; t < RL; ++t) {
        int m = mb1 + (N / RL) * t;
        if (m >= N) m -= N;
        float2 X = acc[c][t];
        float2 Xn = vsel(self, vsel(g == 0 && t == 0, acc[c][2 * RL], acc[c][(RL - t) % RL]),
                         acc[c][RL + (RL - t) % RL]);
        if (m == 0) {
          X.y = 0.f;
          Xn.y = 0.f;
        }
        buf[f * S + m] = unsplit_bin_conj(X, Xn, cconj(tw2[m]));
      }
      if (!self) {
#pragma unroll
        for (int t = 0; t < RL; ++t) {
          int m = mb2 + (N / RL) * t;
          if (m >= N) m -= N;
          buf[f * S + m] = unsplit_bin_conj(acc[c][RL + t], acc[c][(RL - t) % RL], cconj(tw2[m]));
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < TPTL; ++c) {
    if constexpr (PFA) break;
    const int task = tid + c * NT;
    if (!(TL % NT == 0 || task < TL)) continue;
    const int f = task / G, g = task % G;
    const bool pair = g != 0 && 2 * g != NsL;
    const int j2 = NsL - g;
#pragma unroll
    for (int t = 0; t < RL; ++t) {
      const int m = g + t * NsL;
      float2 X = acc[c][t];
      const float2 x0 = (t == 0) ? acc[c][2 * RL] : acc[c][(RL - t) % RL];
      const float2 x1 = acc[c][2 * RL - 1 - t], x2 = acc[c][RL - 1 - t];
      float2 Xn = vsel(g == 0, x0, vsel(pair, x1, x2));
      if (m == 0) {
        X.y = 0.f;
        Xn.y = 0.f;
      }
      buf[f * S + m] = unsplit_bin_conj(X, Xn, cconj(tw2[m]));
    }
    if (pair) {
#pragma unroll
      for (int t = 0; t < RL; ++t) {
        const int m = j2 + t * NsL;
        buf[f * S + m] = unsplit_bin_conj(acc[c][RL + t], acc[c][RL - 1 - t], cconj(tw2[m]));
      }
    }
  }
  __syncthreads();
  if constexpr (RST) {
    // the two references' norms: wavefront sums, then the wavefronts in order
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      rsum0 += __shfl_xor(rsum0, o);
      rsum1 += __shfl_xor(rsum1, o);
    }
    if ((tid & 63) == 0) {
      rsq[0][tid >> 6] = rsum0;
      rsq[1][tid >> 6] = rsum1;
    }
    __syncthreads();
    if (tid < F) {
      double s0 = 0.0;
      const int row = ((tile * F + tid) / a.ns) - rr0;
      for (int w = 0; w < NT / 64; ++w) s0 += rsq[row < 1 ? 0 : 1][w];
      sq[tid] = (float)s0;
      off[tid] = 0.f;
    }
  } else if (tid < F) {
    float s0 = 0.f, s1 = 0.f;
    for (int u = 0; u < Q; ++u) {
      s0 += sqs[tid * Q + u];
      s1 += offs[tid * Q + u];
    }
    sq[tid] = s0;
    off[tid] = s1;
  }
  if (tid == 0) {
    if (a.mode == 1) {
      float s = 0.f;
      for (int kk = 0; kk < a.kcorr; ++kk)
        s += a.shellw[(size_t)ip * nk + kk] * a.sqsums_k[(size_t)ip * nk + kk];
      pscal = s;
    } else {
      pscal = a.sqsums[ip];
    }
  }
  // ---- inverse FFT of every pair (all passes in LDS) ----
  lds_passes<N, F, NT, 0, NP, S>(buf, tw2, tid);
  const float inv4n = 1.0f / (4.0f * (float)N);
  unsigned long long bestkey = 0ull;
  for (int idx = tid; idx < F * N; idx += NT) {
    const int f = idx / N, j = idx % N;
    const int P = tile * F + f;
    if (P >= npairs) continue;
    const float2 y = buf[f * S + j];
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
      float2* o = reinterpret_cast<float2*>(a.cc_out + ((size_t)pi * npairs + P) * (2 * N));
      o[j] = make_float2(x0, x1);
    }
    if (a.best) {
      const unsigned int lin0 = (unsigned int)P * (unsigned int)(2 * N) + 2u * j;
      const unsigned long long k0 = ((unsigned long long)ordered_bits(x0) << 32) | (0xFFFFFFFFu - lin0);
      const unsigned long long k1 = ((unsigned long long)ordered_bits(x1) << 32) | (0xFFFFFFFFu - lin0 - 1u);
      bestkey = k0 > bestkey ? k0 : bestkey;
      bestkey = k1 > bestkey ? k1 : bestkey;
    }
  }
  if (a.best) {
    // wavefront maxima, then the wavefronts' in a fixed order (any NT that is a multiple of 64)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long v = __shfl_xor(bestkey, o);
      bestkey = v > bestkey ? v : bestkey;
    }
    if ((tid & 63) == 0) red[tid >> 6] = bestkey;
    __syncthreads();
    if (tid == 0) {
      unsigned long long m = red[0];
      for (int w = 1; w < NT / 64; ++w) m = red[w] > m ? red[w] : m;
      atomicMax(&a.best[pi], m);
    }
  }
}

}  // namespace pfk

namespace pfk {

// ------------------------------------------------------------------- item mode (callback) ---
// The shift search's in-plane callback: every item is its own (reference, particle, shift)
// triple, gencorrs(iref, iptcl, shvec) (:1502-1534) + maxloc.  Same arithmetic as
// k_corr_fused, with per-item particle data: the particle's |CTF| and memoized spectra are read
// from global memory in pass 1 and the last pass, and each item's shift phases follow
// gen_shmat_8's FP64 recurrence over the shells (:1098-1120) in LDS, rounded to FP32 as
// gen_shmat does.  F items per work-group; the best (cc, rotation) key of item P goes to best[P].
#ifndef PFTCC_ITEMS_FMAX
#define PFTCC_ITEMS_FMAX 4
#endif
#ifndef PFTCC_ITEMS_W
#define PFTCC_ITEMS_W 1        // minimum wavefronts per SIMD imposed on the register allocation
#endif
#ifndef PFTCC_ITEMS_PREFETCH
#define PFTCC_ITEMS_PREFETCH 1 // pass 1 loads the next shell's reference and |CTF| values ahead
#endif
#ifndef PFTCC_ITEMS_PREFETCH_W
#define PFTCC_ITEMS_PREFETCH_W 1   // the last pass loads the next shell's memoized spectra ahead
#endif

template <int N>
constexpr int items_F() {
  // LDS: buf[F N] + tw2[2N] + twL float2, sq/off slots 2 F N/4 floats, phases 2 F N double2
  for (int F = PFTCC_ITEMS_FMAX; F > 1; F /= 2)
    if (sizeof(float2) * ((size_t)F * N + 2 * N + fused_twl_len<N>()) + sizeof(float) * 2 * F * (N / 4) + 16 +
            sizeof(double2) * 2 * (size_t)F * N <= 65536)
      return F;
  return 1;
}
template <int N, int F>
constexpr size_t items_lds_bytes() {
  return sizeof(float2) * ((size_t)F * N + 2 * N + fused_twl_len<N>()) + sizeof(float) * 2 * F * (N / 4) + 16 +
         sizeof(double2) * 2 * (size_t)F * N;
}

template <int N, int F, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(PFTCC_ITEMS_W))) k_corr_fused_items(CorrArgs a) {
  static_assert(Fp<N>::fusable, "first radix must be 4");
  constexpr int NP = Fp<N>::np;
  constexpr int RL = Fp<N>::R(NP - 1);
  constexpr int NsL = N / RL;
  constexpr int G = NsL / 2 + 1;
  constexpr int TL = F * G;
  constexpr int TPTL = (TL + NT - 1) / NT;
  constexpr int NB = 2 * RL + 1;
  constexpr int Q = N / 4;
  constexpr int T1 = F * Q;
  constexpr int TPT1 = (T1 + NT - 1) / NT;

  extern __shared__ __align__(16) float2 smem[];
  float2* buf = smem;                        // [F][N]
  float2* tw2 = buf + F * N;                 // [2N]
  float2* twL = tw2 + 2 * N;                 // [RL-1][NsL]
  float* sqs = reinterpret_cast<float*>(twL + fused_twl_len<N>());   // [F][Q]
  float* offs = sqs + F * Q;                                          // [F][Q]
  float* const fbase = reinterpret_cast<float*>(smem);
  const size_t shoff = ((size_t)(offs + F * Q - fbase) + 3) & ~(size_t)3;    // 16-byte aligned
  double2* shph = reinterpret_cast<double2*>(fbase + shoff);                 // [F][N] current phase
  double2* shst = shph + F * N;                                             // [F][N] shell step
  __shared__ float sq[F], off[F], pscal[F];
  __shared__ unsigned long long bestf[F];

  const int tid = threadIdx.x;
  const int item0 = blockIdx.x * F;
  const int nk = a.nk;

  for (int i = tid; i < 2 * N; i += NT) tw2[i] = a.tw2g[i];
  if constexpr (PFTCC_FUSED_TWL)
    for (int i = tid; i < (RL - 1) * NsL; i += NT) {
      const int t = i / NsL + 1, g = i - (t - 1) * NsL;
      twL[i] = a.tw2g[2 * g * t];
    }
  for (int i = tid; i < F * N; i += NT) {
    const int f = i / N, j = i - f * N;
    const int P = item0 + f;
    double2 ph = make_double2(1.0, 0.0), st = make_double2(1.0, 0.0);
    if (P < a.np) {
      const double sx = (double)a.item_shift[2 * P], sy = (double)a.item_shift[2 * P + 1];
      double sn, cs;
      sincos(a.argx[j] * sx + a.argy[j] * sy, &sn, &cs);
      ph = make_double2(cs, sn);
      sincos(a.onex[j] * sx + a.oney[j] * sy, &sn, &cs);
      st = make_double2(cs, sn);
    }
    shph[i] = ph;
    shst[i] = st;
  }
  if (tid < F) bestf[tid] = 0ull;

  // pass-1 slots: item, particle and reference rows
  float sqp[TPT1], offp[TPT1];
  const float2* rrow[TPT1];
  const float* crow[TPT1];
  int slotref[TPT1], slotip[TPT1];
  bool slotok[TPT1];
#pragma unroll
  for (int c = 0; c < TPT1; ++c) {
    const int task = tid + c * NT;
    const int f = task / Q;
    const int P = item0 + f;
    slotok[c] = task < T1 && P < a.np;
    sqp[c] = 0.f;
    offp[c] = 0.f;
    rrow[c] = a.refs_even;
    crow[c] = nullptr;
    slotref[c] = 0;
    slotip[c] = 0;
    if (slotok[c]) {
      const int ip = a.plist[P], iref = a.rlist[P];
      slotip[c] = ip;
      slotref[c] = iref;
      rrow[c] = (a.iseven[ip] ? a.refs_even : a.refs_odd) + (size_t)iref * nk * N;
      crow[c] = a.ctf ? a.ctf + (size_t)ip * nk * N : nullptr;
    }
  }
  // last-pass tasks: the item's memoized spectra
  int lip[TPTL];
  bool lok[TPTL];
#pragma unroll
  for (int c = 0; c < TPTL; ++c) {
    const int task = tid + c * NT;
    const int P = item0 + task / G;
    lok[c] = task < TL && P < a.np;
    lip[c] = lok[c] ? a.plist[P] : 0;
  }

  float2 acc[TPTL][NB];
#pragma unroll
  for (int c = 0; c < TPTL; ++c)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[c][b] = make_float2(0.f, 0.f);

  const int kloop = a.kcorr > a.knorm ? a.kcorr : a.knorm;
  // pass-1 operands of the next shell, loaded while the current shell's passes run
  float4 pra[TPT1], prb[TPT1];
  float2 pc01[TPT1], pc23[TPT1];
  auto fetch1 = [&](int c, int k) {
    const int j = (tid + c * NT) % Q;
    const float2* r = rrow[c] + (size_t)k * N;
    pra[c] = *reinterpret_cast<const float4*>(r + 2 * j);
    prb[c] = *reinterpret_cast<const float4*>(r + 2 * j + N / 2);
    if (crow[c]) {
      pc01[c] = *reinterpret_cast<const float2*>(crow[c] + (size_t)k * N + 2 * j);
      pc23[c] = *reinterpret_cast<const float2*>(crow[c] + (size_t)k * N + 2 * j + N / 2);
    } else {
      pc01[c] = pc23[c] = make_float2(1.f, 1.f);
    }
  };
  if constexpr (PFTCC_ITEMS_PREFETCH) {
#pragma unroll
    for (int c = 0; c < TPT1; ++c)
      if (slotok[c] && kloop > 0) fetch1(c, 0);
  }
  // last-pass spectra of the next corr shell (bins g + t NsL, NsL - g + t NsL and N of task c)
  float2 pw[TPTL][NB];
  auto fetchw = [&](int c, int k) {
    const int task = tid + c * NT;
    const int g = task % G;
    const bool pair = g != 0 && 2 * g != NsL;
    const float2* wrow = a.memo + ((size_t)lip[c] * nk + k) * (N + 1);
#pragma unroll
    for (int t = 0; t < RL; ++t) pw[c][t] = lok[c] ? wrow[g + t * NsL] : make_float2(0.f, 0.f);
#pragma unroll
    for (int t = 0; t < RL; ++t) pw[c][RL + t] = (lok[c] && pair) ? wrow[NsL - g + t * NsL] : make_float2(0.f, 0.f);
    pw[c][2 * RL] = (lok[c] && g == 0) ? wrow[N] : make_float2(0.f, 0.f);
  };
  if constexpr (PFTCC_ITEMS_PREFETCH_W) {
#pragma unroll
    for (int c = 0; c < TPTL; ++c)
      if ((TL % NT == 0 || tid + c * NT < TL) && a.kcorr > 0) fetchw(c, 0);
  }
  for (int kk = 0; kk < kloop; ++kk) {
    const bool docorr = kk < a.kcorr;
    const bool donorm = kk < a.knorm;
    __syncthreads();   // the previous shell's last pass has finished with buf; phases initialised
    // ---- pass 1: q = ref * |CTF| * shmat (* shell scale); radix-4 DFT ----
#pragma unroll
    for (int c = 0; c < TPT1; ++c) {
      const int task = tid + c * NT;
      if (!(T1 % NT == 0 || task < T1)) continue;
      const int f = task / Q, j = task % Q;
      float2 y[4];
      if (slotok[c]) {
        const int ip = slotip[c];
        float4 ra, rb;
        float2 c01 = make_float2(1.f, 1.f), c23 = c01;
        if constexpr (PFTCC_ITEMS_PREFETCH) {
          ra = pra[c];
          rb = prb[c];
          c01 = pc01[c];
          c23 = pc23[c];
          if (kk + 1 < kloop) fetch1(c, kk + 1);
        } else {
          const float2* r = rrow[c] + (size_t)kk * N;
          ra = *reinterpret_cast<const float4*>(r + 2 * j);
          rb = *reinterpret_cast<const float4*>(r + 2 * j + N / 2);
          if (crow[c]) {
            c01 = *reinterpret_cast<const float2*>(crow[c] + (size_t)kk * N + 2 * j);
            c23 = *reinterpret_cast<const float2*>(crow[c] + (size_t)kk * N + 2 * j + N / 2);
          }
        }
        float2 q0 = make_float2(ra.x * c01.x, ra.y * c01.x), q1 = make_float2(ra.z * c01.y, ra.w * c01.y);
        float2 q2 = make_float2(rb.x * c23.x, rb.y * c23.x), q3 = make_float2(rb.z * c23.y, rb.w * c23.y);
        // shift phases of angles 2j, 2j+1, 2j+N/2, 2j+N/2+1 (owned by this task): use, then advance
        double2* ph = shph + f * N;
        const double2* sp = shst + f * N;
        const int ang[4] = {2 * j, 2 * j + 1, 2 * j + N / 2, 2 * j + N / 2 + 1};
        float2* qs[4] = {&q0, &q1, &q2, &q3};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const double2 p = ph[ang[u]], s = sp[ang[u]];
          *qs[u] = cmul(*qs[u], make_float2((float)p.x, (float)p.y));
          ph[ang[u]] = make_double2(p.x * s.x - p.y * s.y, p.x * s.y + p.y * s.x);
        }
        if (a.refinv && docorr) {
          const int iref = slotref[c];
          const int row = a.optrow ? a.optrow[ip] : iref;
          const float sc = a.ofac[(size_t)row * a.kcorr + kk] *
                           a.refinv[((size_t)(a.iseven[ip] ? 0 : 1) * a.nrefs + iref) * a.kcorr + kk];
          q0 = cscale(q0, sc);
          q1 = cscale(q1, sc);
          q2 = cscale(q2, sc);
          q3 = cscale(q3, sc);
        }
        const float e2 = q0.x * q0.x + q0.y * q0.y + q1.x * q1.x + q1.y * q1.y + q2.x * q2.x + q2.y * q2.y +
                         q3.x * q3.x + q3.y * q3.y;
        if (donorm) sqp[c] += e2;
        if (a.mode == 1 && docorr) offp[c] = fmaf(a.shellw[(size_t)ip * nk + kk], e2, offp[c]);
        y[0] = make_float2(q0.x + q0.y, q1.x + q1.y);
        y[1] = make_float2(q2.x + q2.y, q3.x + q3.y);
        y[2] = make_float2(q0.x - q0.y, q1.x - q1.y);
        y[3] = make_float2(q2.x - q2.y, q3.x - q3.y);
      } else {
        y[0] = y[1] = y[2] = y[3] = make_float2(0.f, 0.f);
      }
      if (docorr) {
        Dft<4>::run(y);
        float4* dst = reinterpret_cast<float4*>(buf + f * N + 4 * j);
        dst[0] = make_float4(y[0].x, y[0].y, y[1].x, y[1].y);
        dst[1] = make_float4(y[2].x, y[2].y, y[3].x, y[3].y);
      }
    }
    if (!docorr) continue;
    __syncthreads();
    lds_passes<N, F, NT, 1, NP - 1>(buf, tw2, tid);
    // ---- last pass + split + conj(V) (W_k / 2) accumulated in registers ----
#pragma unroll
    for (int c = 0; c < TPTL; ++c) {
      const int task = tid + c * NT;
      if (!(TL % NT == 0 || task < TL)) continue;
      const int f = task / G, g = task % G;
      const bool pair = g != 0 && 2 * g != NsL;
      const int j2 = NsL - g;
      const float2* wrow = a.memo + ((size_t)lip[c] * nk + kk) * (N + 1);
      const float wk = (a.mode == 1 && lok[c]) ? a.shellw[(size_t)lip[c] * nk + kk] : 1.0f;
      float2 wcur[NB];
      if constexpr (PFTCC_ITEMS_PREFETCH_W) {
#pragma unroll
        for (int b = 0; b < NB; ++b) wcur[b] = pw[c][b];
        if (kk + 1 < a.kcorr) fetchw(c, kk + 1);
      }
      // the spectrum at bin m, held in slot t (m = g + t NsL), RL + t (m = NsL - g + t NsL) or
      // 2 RL (m = N)
      auto Wb = [&](int slot, int m) {
        const float2 w = PFTCC_ITEMS_PREFETCH_W ? wcur[slot] : (lok[c] ? wrow[m] : make_float2(0.f, 0.f));
        return make_float2(0.5f * w.x, 0.5f * w.y);
      };
      float2 y1[RL], y2[RL];
#pragma unroll
      for (int t = 0; t < RL; ++t) y1[t] = buf[f * N + g + t * NsL];
#pragma unroll
      for (int t = 1; t < RL; ++t) y1[t] = cmul(y1[t], PFTCC_FUSED_TWL ? twL[(t - 1) * NsL + g] : tw2[2 * g * t]);
      Dft<RL>::run(y1);
      if (pair) {
#pragma unroll
        for (int t = 0; t < RL; ++t) y2[t] = buf[f * N + j2 + t * NsL];
#pragma unroll
        for (int t = 1; t < RL; ++t) y2[t] = cmul(y2[t], PFTCC_FUSED_TWL ? twL[(t - 1) * NsL + j2] : tw2[2 * j2 * t]);
        Dft<RL>::run(y2);
      } else {
#pragma unroll
        for (int t = 0; t < RL; ++t) y2[t] = make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int t = 0; t < RL; ++t) {
        const int m = g + t * NsL;
        const float2 b0 = y1[(RL - t) % RL], b1 = y2[RL - 1 - t], b2 = y1[RL - 1 - t];
        const float2 B = vsel(g == 0, b0, vsel(pair, b1, b2));
        const float2 V = split_bin2(y1[t], B, tw2[m]);
        const float2 p = cmulcj(V, Wb(t, m));
        acc[c][t].x = fmaf(wk, p.x, acc[c][t].x);
        acc[c][t].y = fmaf(wk, p.y, acc[c][t].y);
      }
      if (pair) {
#pragma unroll
        for (int t = 0; t < RL; ++t) {
          const int m = j2 + t * NsL;
          const float2 V = split_bin2(y2[t], y1[RL - 1 - t], tw2[m]);
          const float2 p = cmulcj(V, Wb(RL + t, m));
          acc[c][RL + t].x = fmaf(wk, p.x, acc[c][RL + t].x);
          acc[c][RL + t].y = fmaf(wk, p.y, acc[c][RL + t].y);
        }
      }
      if (g == 0) {
        const float2 V = split_bin2(y1[0], y1[0], tw2[N]);
        const float2 p = cmulcj(V, Wb(2 * RL, N));
        acc[c][2 * RL].x = fmaf(wk, p.x, acc[c][2 * RL].x);
        acc[c][2 * RL].y = fmaf(wk, p.y, acc[c][2 * RL].y);
      }
    }
  }
  __syncthreads();
  // ---- normalisation sums (fixed order) ----
#pragma unroll
  for (int c = 0; c < TPT1; ++c) {
    const int task = tid + c * NT;
    if (task < T1) {
      sqs[task] = sqp[c];
      offs[task] = offp[c];
    }
  }
  // ---- Y = pre-processed inverse input, conj'd, from the register spectra ----
#pragma unroll
  for (int c = 0; c < TPTL; ++c) {
    const int task = tid + c * NT;
    if (!(TL % NT == 0 || task < TL)) continue;
    const int f = task / G, g = task % G;
    const bool pair = g != 0 && 2 * g != NsL;
    const int j2 = NsL - g;
#pragma unroll
    for (int t = 0; t < RL; ++t) {
      const int m = g + t * NsL;
      float2 X = acc[c][t];
      const float2 x0 = (t == 0) ? acc[c][2 * RL] : acc[c][(RL - t) % RL];
      const float2 x1 = acc[c][2 * RL - 1 - t], x2 = acc[c][RL - 1 - t];
      float2 Xn = vsel(g == 0, x0, vsel(pair, x1, x2));
      if (m == 0) {
        X.y = 0.f;
        Xn.y = 0.f;
      }
      buf[f * N + m] = unsplit_bin_conj(X, Xn, cconj(tw2[m]));
    }
    if (pair) {
#pragma unroll
      for (int t = 0; t < RL; ++t) {
        const int m = j2 + t * NsL;
        buf[f * N + m] = unsplit_bin_conj(acc[c][RL + t], acc[c][RL - 1 - t], cconj(tw2[m]));
      }
    }
  }
  __syncthreads();
  if (tid < F) {
    float s0 = 0.f, s1 = 0.f;
    for (int u = 0; u < Q; ++u) {
      s0 += sqs[tid * Q + u];
      s1 += offs[tid * Q + u];
    }
    sq[tid] = s0;
    off[tid] = s1;
    const int P = item0 + tid;
    float ps = 1.f;
    if (P < a.np) {
      const int ip = a.plist[P];
      if (a.mode == 1) {
        ps = 0.f;
        for (int kk = 0; kk < a.kcorr; ++kk) ps += a.shellw[(size_t)ip * nk + kk] * a.sqsums_k[(size_t)ip * nk + kk];
      } else {
        ps = a.sqsums[ip];
      }
    }
    pscal[tid] = ps;
  }
  // ---- inverse FFT of every item (all passes in LDS) ----
  lds_passes<N, F, NT, 0, NP>(buf, tw2, tid);
  const float inv4n = 1.0f / (4.0f * (float)N);
  for (int idx = tid; idx < F * N; idx += NT) {
    const int f = idx / N, j = idx % N;
    if (item0 + f >= a.np) continue;
    const float2 y = buf[f * N + j];
    float x0 = y.x, x1 = -y.y;
    if (a.mode == 1) {
      const float o = off[f] + pscal[f];
      x0 = 2.f * x0 * inv4n - o;
      x1 = 2.f * x1 * inv4n - o;
    } else {
      const float sc = inv4n / sqrtf(sq[f] * pscal[f]);
      x0 *= sc;
      x1 *= sc;
    }
    const unsigned int lin0 = 2u * j;
    const unsigned long long k0 = ((unsigned long long)ordered_bits(x0) << 32) | (0xFFFFFFFFu - lin0);
    const unsigned long long k1 = ((unsigned long long)ordered_bits(x1) << 32) | (0xFFFFFFFFu - lin0 - 1u);
    atomicMax(&bestf[f], k0 > k1 ? k0 : k1);
  }
  __syncthreads();
  if (tid < F && item0 + tid < a.np) a.best[item0 + tid] = bestf[tid];
}

}  // namespace pfk
