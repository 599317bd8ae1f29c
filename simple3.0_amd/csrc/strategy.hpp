// strategy.hpp - the reference-loop search strategies, evaluated on the device from the
// correlation vectors of every (particle, reference) pair:
//   greedy 2-D  simple_strategy2D_greedy.f90:33-58   (all classes with members, ">=": last max)
//   snhc 2-D    simple_strategy2D_snhc.f90:33-93     (stochastic order, neighbourhood bound,
//                                                     first improvement chosen by shcloc)
//   greedy 3-D  simple_strategy3D_greedy.f90:40-77   (all projections, then sort_corrs)
//   shc 3-D     simple_strategy3D_shc.f90:36-83      (stochastic order, stop after the first
//                                                     reference whose best beats prev_corr)
//   snhc 3-D    simple_strategy3D_snhc.f90:37-60     (the first szsn references of the order)
//   sort_corrs  simple_strategy3D_utils.f90:100-124  (references ordered by their best
//                                                     correlation, unsearched ones at -huge)
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cstdint>

namespace pfk {

enum { SRCH_GREEDY2D = 0, SRCH_SNHC2D = 1, SRCH_GREEDY3D = 2, SRCH_SHC3D = 3, SRCH_SNHC3D = 4 };

// maxloc over the rotations of each correlation vector (one wavefront per vector): the largest
// value, ties to the lowest rotation as Fortran's maxloc.  cc [nvec][nrots].
__device__ __forceinline__ void wave_maxloc(const float* __restrict__ row, int nrots, float& bo, int& io) {
  const int lane = threadIdx.x & 63;
  float b = -FLT_MAX;
  int bi = INT_MAX;
  for (int r = lane; r < nrots; r += 64) {
    const float x = row[r];
    if (bi == INT_MAX || x > b) {
      b = x;
      bi = r;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const float b2 = __shfl_xor(b, o);
    const int i2 = __shfl_xor(bi, o);
    if (i2 != INT_MAX && (bi == INT_MAX || b2 > b || (b2 == b && i2 < bi))) {
      b = b2;
      bi = i2;
    }
  }
  bo = b;
  io = bi;
}

__global__ void __launch_bounds__(256) k_ref_maxloc(const float* __restrict__ cc, long long nvec, int nrots,
                                                    float* __restrict__ vbest, int* __restrict__ vrot) {
  const long long v = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (v >= nvec) return;
  float b;
  int bi;
  wave_maxloc(cc + v * nrots, nrots, b, bi);
  if ((threadIdx.x & 63) == 0) {
    vbest[v] = b;
    vrot[v] = bi;
  }
}

// A round of the stochastic scans: active particle a (global index act[a]) correlates with the
// w references at positions [i0, i0 + w) of its search order.  Builds the launch's particle and
// per-particle reference lists.
__global__ void k_round_lists(const int* __restrict__ order, const int* __restrict__ act,
                              const int* __restrict__ plocal, int nact, int nrefs, int i0, int w,
                              int* __restrict__ rpl, int* __restrict__ rl) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nact * w) return;
  const int a = (int)(t / w), c = (int)(t - (long long)a * w);
  const int p = act[a];
  rl[t] = order[(size_t)p * nrefs + i0 + c];
  if (c == 0) rpl[a] = plocal[p];
}

// maxloc of a round's vectors [nact][w][nrots] into the particles' per-reference tables.
__global__ void __launch_bounds__(256) k_round_maxloc(const float* __restrict__ cc, int nact, int w, int nrots,
                                                      const int* __restrict__ act, const int* __restrict__ rl,
                                                      int nrefs, float* __restrict__ vbest, int* __restrict__ vrot) {
  const long long v = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (v >= (long long)nact * w) return;
  float b;
  int bi;
  wave_maxloc(cc + v * nrots, nrots, b, bi);
  if ((threadIdx.x & 63) == 0) {
    const size_t o = (size_t)act[v / w] * nrefs + rl[v];
    vbest[o] = b;
    vrot[o] = bi;
  }
}

struct StrategyArgs {
  int kind;
  int nrefs, nrots;
  int bound;                   // snhc 2-D / 3-D: positions of the search order inside the neighbourhood
  int ntop;                    // sorted references written per particle (3-D kinds)
  int sortlen;                 // power of two >= nrefs (3-D kinds): LDS keys of the sort
  const int* order;            // [np][nrefs] 0-based references in search order, or null
  const int8_t* ref_ok;        // [nrefs] class populated / state exists, or null (all)
  const float* prev_corr;      // [np] or null (greedy kinds): -huge
  const float* rnd;            // [np] uniform draws in [0, 1) for shcloc
  const float* cc;             // [np][nrefs][nrots]
  const float* vbest;          // [np][nrefs]  maxloc value
  const int* vrot;             // [np][nrefs]  maxloc rotation, 0-based
  float* best_corr;            // [np]
  int* best_ref;               // [np] 1-based, 0: no reference evaluated
  int* best_rot;               // [np] 1-based
  int* nrefs_eval;             // [np]
  int* found;                  // [np] the stochastic kinds' exit on an improving reference
  float* ref_corrs;            // [np][nrefs] or null: proj_space_corrs, -FLT_MAX where unsearched
  int* ref_rots;               // [np][nrefs] or null: 1-based, 0 where unsearched
  int* sorted;                 // [np][ntop] or null: 1-based references, best first
  // scans resolved by the rounds (k_strategy_round): exit position (INT_MAX: none), and for snhc
  // shcloc's rotation and its correlation; null: k_strategy scans itself
  const int* istar_in;
  const int* srot_in;
  const float* scorr_in;
};

__device__ __forceinline__ unsigned int strat_bits(float v) {
  const unsigned int b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

template <typename T, typename Op>
__device__ __forceinline__ T block_reduce(T v, T* sh, Op op) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  T r = sh[0];
  for (int i = 1; i < nw; ++i) r = op(r, sh[i]);
  return r;
}

// shcloc (simple_rnd.f90:273-295): a uniformly drawn rotation among those with
// corrs - prev_corr >= 0, here the floor(u count)-th of them in rotation order.  Whole work-group;
// -1 if none.
__device__ int shcloc_pick(const float* __restrict__ row, int nrots, float prev, float u, int* scnt, int* srot) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int seg = (nrots + nt - 1) / nt;
  const int j0 = min(tid * seg, nrots), j1 = min(j0 + seg, nrots);
  int cnt = 0;
  for (int j = j0; j < j1; ++j) cnt += row[j] >= prev;
  __syncthreads();
  scnt[tid] = cnt;
  if (tid == 0) *srot = -1;
  __syncthreads();
  if (tid == 0) {
    int s = 0;
    for (int t = 0; t < nt; ++t) {
      const int c = scnt[t];
      scnt[t] = s;
      s += c;
    }
    scnt[0] = s;   // thread 0's exclusive prefix is 0; the total is kept in its slot
  }
  __syncthreads();
  const int total = scnt[0];
  const int excl = tid == 0 ? 0 : scnt[tid];
  int k = (int)floor((double)u * (double)total);
  k = max(0, min(k, total - 1));
  if (total > 0 && k >= excl && k < excl + cnt) {
    int c = excl;
    for (int j = j0; j < j1; ++j)
      if (row[j] >= prev) {
        if (c == k) {
          *srot = j;
          break;
        }
        ++c;
      }
  }
  __syncthreads();
  return *srot;
}

// One work-group per particle.  The scan limit, the first improving position, the strategy's
// choice, the references it searched and (3-D) their order by best correlation.
__global__ void __launch_bounds__(256) k_strategy(StrategyArgs a) {
  extern __shared__ unsigned long long skeys[];
  __shared__ unsigned long long sh64[4];
  __shared__ int sh32[4];
  __shared__ int scnt[256];
  __shared__ int srot;
  const int p = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const int nrefs = a.nrefs;
  const bool stoch = a.kind == SRCH_SNHC2D || a.kind == SRCH_SHC3D || a.kind == SRCH_SNHC3D;
  const bool noexit = a.kind == SRCH_SNHC3D;   // a fixed neighbourhood, every reference searched
  const int* ord = stoch ? a.order + (size_t)p * nrefs : nullptr;
  const float* vb = a.vbest + (size_t)p * nrefs;
  const int* vr = a.vrot + (size_t)p * nrefs;
  const float prev = a.prev_corr ? a.prev_corr[p] : -FLT_MAX;
  auto ref_at = [&](int i) { return ord ? ord[i] : i; };
  auto ok = [&](int r) { return noexit || a.ref_ok == nullptr || a.ref_ok[r] != 0; };

  // positions [0, lim) may be searched; the stochastic kinds stop at the first improving one
  const int lim = (a.kind == SRCH_SNHC2D || noexit) ? min(a.bound, nrefs) : nrefs;
  int istar = INT_MAX;
  if (noexit) {
    istar = INT_MAX;
  } else if (stoch && a.istar_in) {
    istar = a.istar_in[p];
  } else if (stoch) {
    int mine = INT_MAX;
    for (int i = tid; i < lim; i += nt) {
      const int r = ref_at(i);
      if (ok(r)) {
        const float c = vb[r];
        // shcloc finds a rotation when some corrs - prev_corr >= 0; shc counts best > prev_corr
        if (a.kind == SRCH_SNHC2D ? (c >= prev) : (c > prev)) {
          mine = i;
          break;
        }
      }
    }
    istar = block_reduce(mine, sh32, [](int x, int y) { return min(x, y); });
  }
  const bool found = istar != INT_MAX;
  const int last = found ? istar : lim - 1;   // the searched positions are [0, last]

  // proj_space_corrs / inplinds of this particle (each reference written once: order is a
  // permutation)
  if (a.ref_corrs)
    for (int i = tid; i < nrefs; i += nt) {
      const int r = ref_at(i);
      const bool s = i <= last && ok(r);
      a.ref_corrs[(size_t)p * nrefs + r] = s ? vb[r] : -FLT_MAX;
      a.ref_rots[(size_t)p * nrefs + r] = s ? vr[r] + 1 : 0;
    }

  int bref = -1;
  if (a.kind == SRCH_GREEDY2D || (a.kind == SRCH_SNHC2D && !found)) {
    // greedy: "inpl_corr >= corr" over iref = 1..nrefs keeps the last of equal maxima; snhc's
    // global best uses ">" over the search order and keeps the first
    unsigned long long key = 0;
    for (int i = tid; i <= last; i += nt) {
      const int r = ref_at(i);
      if (!ok(r)) continue;
      const unsigned long long lo = a.kind == SRCH_GREEDY2D ? (unsigned)r : 0xFFFFFFFFu - (unsigned)i;
      const unsigned long long k = ((unsigned long long)strat_bits(vb[r]) << 32) | lo;
      if (k > key) key = k;
    }
    key = block_reduce(key, sh64, [](unsigned long long x, unsigned long long y) { return x > y ? x : y; });
    if (key != 0) {
      const unsigned lo = (unsigned)(key & 0xFFFFFFFFu);
      bref = a.kind == SRCH_GREEDY2D ? (int)lo : ref_at((int)(0xFFFFFFFFu - lo));
    }
    if (tid == 0) {
      a.best_ref[p] = bref + 1;
      a.best_rot[p] = bref >= 0 ? vr[bref] + 1 : 0;
      a.best_corr[p] = bref >= 0 ? vb[bref] : prev;
    }
  } else if (a.kind == SRCH_SNHC2D) {
    bref = ref_at(istar);
    int rot;
    float corr;
    if (a.srot_in) {
      rot = a.srot_in[p];
      corr = a.scorr_in[p];
    } else {
      const float* row = a.cc + ((size_t)p * nrefs + bref) * a.nrots;
      rot = shcloc_pick(row, a.nrots, prev, a.rnd[p], scnt, &srot);
      if (rot < 0) rot = vr[bref];
      corr = row[rot];
    }
    if (tid == 0) {
      a.best_ref[p] = bref + 1;
      a.best_rot[p] = rot + 1;
      a.best_corr[p] = corr;
    }
  } else {
    // sort_corrs: the searched references by best correlation, unsearched at -huge; ties to the
    // lower reference index
    const int L = a.sortlen;
    for (int i = tid; i < L; i += nt) skeys[i] = 0ull;
    __syncthreads();
    for (int i = tid; i < nrefs; i += nt) {
      const int r = ref_at(i);
      const bool s = i <= last && ok(r);
      skeys[r] = ((unsigned long long)strat_bits(s ? vb[r] : -FLT_MAX) << 32) | (0xFFFFFFFFu - (unsigned)r);
    }
    __syncthreads();
    for (int k = 2; k <= L; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = tid; i < L; i += nt) {
          const int ixj = i ^ j;
          if (ixj > i) {
            const unsigned long long x = skeys[i], y = skeys[ixj];
            if (((i & k) == 0) ? (x < y) : (x > y)) {
              skeys[i] = y;
              skeys[ixj] = x;
            }
          }
        }
        __syncthreads();
      }
    if (a.sorted)
      for (int t = tid; t < a.ntop; t += nt)
        a.sorted[(size_t)p * a.ntop + t] = (int)(0xFFFFFFFFu - (unsigned)(skeys[t] & 0xFFFFFFFFu)) + 1;
    if (tid == 0) {
      const int r = (int)(0xFFFFFFFFu - (unsigned)(skeys[0] & 0xFFFFFFFFu));
      // the best is a searched reference unless none was searched
      bool any = false;
      for (int i = 0; i <= last && !any; ++i) any = ok(ref_at(i));
      a.best_ref[p] = any ? r + 1 : 0;
      a.best_rot[p] = any ? vr[r] + 1 : 0;
      a.best_corr[p] = any ? vb[r] : prev;
    }
  }

  // evaluated references: greedy all; snhc the position of the exit (nrefs_eval is incremented
  // before the bound test); shc the searched ones with an existing state
  int ne = nrefs;
  if (a.kind == SRCH_SNHC2D) ne = found ? istar + 1 : (a.bound < nrefs ? a.bound + 1 : nrefs);
  if (noexit) ne = lim;
  if (a.kind == SRCH_SHC3D) {
    int c = 0;
    for (int i = tid; i <= last; i += nt) c += ok(ref_at(i));
    ne = block_reduce(c, sh32, [](int x, int y) { return x + y; });
  }
  if (tid == 0) {
    a.nrefs_eval[p] = ne;
    a.found[p] = found ? 1 : 0;
  }
}

// One work-group per active particle of a round: the first improving position in [i0, i1) of
// the scan limit, with shcloc's rotation for snhc (the round's vectors cc [nact][w][nrots] hold
// the exit reference's); done once the exit is found or the scan limit is reached.
__global__ void __launch_bounds__(256) k_strategy_round(StrategyArgs a, const int* __restrict__ act, int i0, int i1,
                                                        const float* __restrict__ cc, int w, int* __restrict__ istar,
                                                        int* __restrict__ srot_out, float* __restrict__ scorr,
                                                        int* __restrict__ done) {
  __shared__ int sh32[4];
  __shared__ int scnt[256];
  __shared__ int srot;
  const int ia = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const int p = act[ia];
  const int nrefs = a.nrefs;
  const int* ord = a.order + (size_t)p * nrefs;
  const float* vb = a.vbest + (size_t)p * nrefs;
  const float prev = a.prev_corr ? a.prev_corr[p] : -FLT_MAX;
  const bool noexit = a.kind == SRCH_SNHC3D;
  const int lim = (a.kind == SRCH_SNHC2D || noexit) ? min(a.bound, nrefs) : nrefs;
  const int iend = min(i1, lim);
  int mine = INT_MAX;
  for (int i = i0 + tid; i < (noexit ? i0 : iend); i += nt) {
    const int r = ord[i];
    if (a.ref_ok == nullptr || a.ref_ok[r] != 0) {
      const float c = vb[r];
      if (a.kind == SRCH_SNHC2D ? (c >= prev) : (c > prev)) {
        mine = i;
        break;
      }
    }
  }
  const int is = block_reduce(mine, sh32, [](int x, int y) { return min(x, y); });
  if (is != INT_MAX && a.kind == SRCH_SNHC2D) {
    const float* row = cc + ((size_t)ia * w + (is - i0)) * a.nrots;
    int rot = shcloc_pick(row, a.nrots, prev, a.rnd[p], scnt, &srot);
    if (rot < 0) rot = a.vrot[(size_t)p * nrefs + ord[is]];
    if (tid == 0) {
      srot_out[p] = rot;
      scorr[p] = row[rot];
    }
  }
  if (tid == 0) {
    istar[p] = is;
    done[p] = (is != INT_MAX || iend >= lim) ? 1 : 0;
  }
}

}  // namespace pfk
