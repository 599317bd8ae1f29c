// pftcc.hip - C ABI (include/pftcc.h) of the gfx950 polar-FT correlation search.
//
// Host-side state mirrors type polarft_corrcalc (simple_polarft_corrcalc.f90:68-198): the polar
// grid, argument-transfer constants and CTF tables are built on the host in the reference's
// precision (FP32 grid, FP64 argtransf) and uploaded once; rings, CTF matrices and the particle
// spectra live in HBM and every calculator runs on the device.  There is no CPU fallback: a
// missing or failing device is reported as PFTCC_E_DEVICE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pftcc.h"
#include "pftcc_kernels.hpp"
#include "corr_fused.hpp"
#include "mfma_search.hpp"
#include "matchfilt.hpp"
#include "polarizer.hpp"
#include "prep.hpp"
#include "shsrch.hpp"
#include "strategy.hpp"

using namespace pfk;

namespace {

// ------------------------------------------------------------------ magic_pftsz (host) ---
// simple_magic_boxes.f90:14-19 (table) and :25-62; simple_math.f90:391-405 (round2even),
// :2051-2084 (locate_2/find_2).  Evaluated in real(sp) like the reference.
const int kBoxSizes[] = {
    32, 36, 40, 48, 52, 56, 64, 66, 70, 72, 80, 84, 88, 100, 104, 108, 112, 120, 128, 130, 132,
    140, 144, 150, 160, 162, 168, 176, 180, 182, 192, 200, 208, 216, 220, 224, 240, 256, 264, 288,
    300, 308, 320, 324, 336, 338, 352, 364, 384, 400, 420, 432, 448, 450, 462, 480, 486, 500, 504,
    512, 520, 528, 546, 560, 576, 588, 600, 640, 648, 650, 660, 672, 686, 700, 702, 704, 720, 726,
    728, 750, 768, 770, 784, 800, 810, 840, 882, 896, 910, 924, 936, 972, 980, 1008, 1014, 1020,
    1024, 1296, 1536, 1728, 1944, 2048, 2304, 2592, 3072, 3200, 3456, 3888, 4096, 4608, 5000, 5184,
    6144, 6250, 6400, 6912, 7776, 8192, 9216, 10240, 12288, 12500};
const int kNBox = sizeof(kBoxSizes) / sizeof(int);
const float kPi32 = 3.14159274101257324f;   // acos(-1.) in real(sp)
const double kDPi = 3.14159265358979323846;

int fnint(float x) { return (int)(x >= 0.f ? std::floor(x + 0.5f) : -std::floor(-x + 0.5f)); }

int round2even(float val) {
  const int r = fnint(val);
  if (r % 2 == 0) return r;
  if (std::fabs((float)(r - 1) - val) <= std::fabs((float)(r + 1) - val)) return r - 1;
  return r + 1;
}

int locate_box(int x) {
  const int n = kNBox;
  int jl = 0, ju = n + 1;
  while (ju - jl > 1) {
    const int jm = (ju + jl) / 2;
    if ((kBoxSizes[n - 1] >= kBoxSizes[0]) == (x >= kBoxSizes[jm - 1])) jl = jm;
    else ju = jm;
  }
  if (x == kBoxSizes[0]) return 1;
  if (x == kBoxSizes[n - 1]) return n - 1;
  return jl;
}

int find_magic_box(int trial) {
  int j = std::max(1, locate_box(trial));
  const int dist1 = kBoxSizes[j - 1] - trial;
  if (j < kNBox) {
    const int dist2 = kBoxSizes[j] - trial;
    if (std::abs(dist1) >= std::abs(dist2)) j = j + 1;
  }
  int best = kBoxSizes[j - 1];
  if (best == kBoxSizes[kNBox - 1] && trial - best > 16) best = round2even((float)trial);
  return best;
}

int magic_pftsz_host(int ring2) {
  const float a = kPi32 * (float)ring2;
  const int old = round2even(a);
  const int p = find_magic_box(fnint(a));
  if ((float)std::abs(p - old) / (float)old > 0.1f) return old;
  return p;
}

}  // namespace

// --------------------------------------------------------------------------------- handle ---
struct pftcc {
  pftcc_cfg cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  int n = 0, nrots = 0, kfrom = 0, kto = 0, kstop = 0, nk = 0, kcorr = 0;
  int nrefs = 0, nptcls = 0, pfrom = 0, pto = 0, ldim = 0;
  bool with_ctf = false, match_filt = false, pssnr = false, clsfrcs = false, euclid = false;
  std::vector<int> pinds;           // pinds[iptcl - pfrom] = local 1-based index, 0 = absent
  std::vector<int8_t> iseven;       // host mirror of the eo flags
  std::vector<float> angtab_rad, angtab_deg, polar_x, polar_y, spafreq, angmat, pxls;
  std::vector<double> argx, argy, onex, oney;
  std::vector<float> sn32, ncn32;   // sin(theta_j), -cos(theta_j) in FP32 (polar = k x these)
  std::vector<float> optlp;         // [nrefs][kcorr]
  std::vector<float> sigma2;        // [nptcls][nk] (local order)
  FftPlan plan{};
  // device state
  float2 *d_refs_e = nullptr, *d_refs_o = nullptr, *d_ptcls = nullptr, *d_memo = nullptr;
  float2 *d_tw2 = nullptr, *d_dptcls = nullptr;
  float *d_ctf = nullptr, *d_sqsums = nullptr, *d_sqsums_k = nullptr;
  float* d_optlp = nullptr;         // [nrefs][kcorr] device copy of ref_optlp (prep_matchfilt)
  bool optlp_dev_valid = false;
  int8_t* d_iseven = nullptr;
  double *d_argx = nullptr, *d_argy = nullptr, *d_onex = nullptr, *d_oney = nullptr;
  float *d_sn32 = nullptr, *d_ncn32 = nullptr;
  double* d_sqp8 = nullptr;         // [nptcls][nk] FP64 shell sums of |P|^2 (built on first use)
  float *d_spafreq = nullptr, *d_angmat = nullptr;
  float *d_refinv_f = nullptr, *d_ofac_f = nullptr, *d_shellw = nullptr;
  double *d_refinv_d = nullptr, *d_ofac_d = nullptr, *d_sigma2w = nullptr, *d_refpw = nullptr;
  int* d_optrow = nullptr;
  // orientation derivatives of the references (set_dref_fcomp, :459-469): [ndslots][3][kcorr][n]
  // per eo class, slots 1..ndslots = max(cfg.nthr, 1) as the reference allocates them (:321-322)
  float2 *d_drefs_e = nullptr, *d_drefs_o = nullptr;
  int ndslots = 0;
  // polarizer tables (init_polarizer): weights [9][nk n] and logical addresses [6][nk n]
  float* d_pz_w = nullptr;
  int* d_pz_hk = nullptr;
  bool pz_ready = false;
  bool memo_valid = false, dptcls_valid = false, refnorm_valid = false, have_sigma2 = false;
  bool sqp8_valid = false;
  bool ctf_ready = false;
  bool use_fused = true;            // PFTCC_GENERIC=1 forces the runtime-size kernel
  // precision form: cached split-BF16 reference operands per eo class
  struct APack {
    bool valid = false;
    std::vector<int> refs;            // 0-based reference ids of the rows
    int rows_pad = 0, KP = 0, kcorr = 0;
    std::vector<float> shifts;        // [ns][2]: rows are (reference, shift), reference-major
    int knorm = 0;                    // shells in sr (kcorr unshifted, nk shifted)
    int fmt = 0;                      // OP_BF16, OP_F16 (split operands hi, lo) or OP_FP8 (a8)
    unsigned int *hi = nullptr, *lo = nullptr;
    double* sr = nullptr;             // [nr][knorm]
    // FP16 and FP8: per-shell mean squares of the reference spectra, shell scales a_k (device
    // copies of a_k and 1/a_k), the operand's largest component and (FP8) the operand itself
    std::vector<double> msV;
    std::vector<float> ak;
    float *d_ak = nullptr, *d_bk = nullptr;
    unsigned int* gmax = nullptr;
    int KP8 = 0;
    unsigned int* a8 = nullptr;
  } apack[2];
  // kernel timing (pftcc_set_timing)
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  std::vector<int> event_ch;
  size_t nev = 0;
  double timed_ms[5] = {0, 0, 0, 0, 0};
  int64_t timed_launches[5] = {0, 0, 0, 0, 0};
  // scratch arena (stream ordered; calls on a handle are serialised by mtx)
  char* d_scratch = nullptr;
  size_t scratch_cap = 0, scratch_off = 0;
  // device copy of pinds (global id - pfrom -> local 1-based slot), kept current by new and
  // reallocate_ptcls, and this GPU's address of the no-allocation search's static scratch
  int* d_pinds = nullptr;
  size_t d_pinds_cap = 0;
  unsigned char* noalloc_scratch = nullptr;
  std::mutex mtx;
  std::string err;
};

namespace {

int fail(pftcc* h, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h) h->err = buf;
  return code;
}

#define HIPCHK(h, call)                                                                        \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return fail((h), PFTCC_E_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),  \
                  __FILE__, __LINE__);                                                         \
  } while (0)

template <typename T>
int dalloc(pftcc* h, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return PFTCC_OK;
  HIPCHK(h, hipMalloc((void**)p, count * sizeof(T)));
  return PFTCC_OK;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// Runs f when the scope ends, on every return path (temporary device buffers).
template <typename F>
struct ScopeExit {
  F f;
  ~ScopeExit() { f(); }
};
template <typename F>
ScopeExit<F> scope_exit(F f) {
  return ScopeExit<F>{f};
}

// Host <-> device copies and fills issued on the handle's stream and completed before returning.
// The stream may be non-blocking (a pool stream of the caller), which the null stream's implicit
// ordering would not cover: a synchronous null-stream copy could overtake pending kernels that
// still read the destination.
hipError_t sync_copy(pftcc* h, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, h->stream);
  return e != hipSuccess ? e : hipStreamSynchronize(h->stream);
}
hipError_t sync_memset(pftcc* h, void* dst, int value, size_t bytes) {
  const hipError_t e = hipMemsetAsync(dst, value, bytes, h->stream);
  return e != hipSuccess ? e : hipStreamSynchronize(h->stream);
}

// Scratch arena: reserve() before a call, take() pieces, all stream ordered.
int scratch_reserve(pftcc* h, size_t bytes) {
  h->scratch_off = 0;
  if (bytes <= h->scratch_cap) return PFTCC_OK;
  if (h->d_scratch) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    (void)hipFree(h->d_scratch);
    h->d_scratch = nullptr;
  }
  size_t cap = std::max(bytes, (size_t)1 << 20);
  HIPCHK(h, hipMalloc((void**)&h->d_scratch, cap));
  h->scratch_cap = cap;
  return PFTCC_OK;
}

template <typename T>
T* take(pftcc* h, size_t count) {
  size_t off = (h->scratch_off + 255) & ~(size_t)255;
  h->scratch_off = off + count * sizeof(T);
  return reinterpret_cast<T*>(h->d_scratch + off);
}

size_t aligned(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

// Dynamic LDS per work-group kept within 64 KiB so several groups share a CU.
constexpr size_t kMaxLds = 64 * 1024;
// Correlation vectors held in HBM per call of the search strategies (a chunk of particles x all
// references x all rotations).
constexpr size_t kStrategyCcBytes = (size_t)1 << 30;
// First round of the stochastic scans, in search-order positions (doubling per round).
constexpr int kStrategyRound0 = 32;

int local_index(pftcc* h, int iptcl, int* ip) {
  if (iptcl < h->pfrom || iptcl > h->pto)
    return fail(h, PFTCC_E_ARG, "iptcl %d outside [%d, %d]", iptcl, h->pfrom, h->pto);
  const int l = h->pinds[iptcl - h->pfrom];
  if (l <= 0) return fail(h, PFTCC_E_ARG, "particle %d is not held (pinds = 0)", iptcl);
  *ip = l - 1;
  return PFTCC_OK;
}

int check_ref(pftcc* h, int iref) {
  if (iref < 1 || iref > h->nrefs) return fail(h, PFTCC_E_ARG, "iref %d outside [1, %d]", iref, h->nrefs);
  return PFTCC_OK;
}

int check_rot(pftcc* h, int irot) {
  if (irot < 1 || irot > h->nrots) return fail(h, PFTCC_E_ARG, "irot %d outside [1, %d]", irot, h->nrots);
  return PFTCC_OK;
}

// Polar grid, argtransf and CTF tables exactly as new() / create_polar_absctfmats build them.
void build_host_grid(pftcc* h) {
  const int n = h->n, R = h->nrots, nk = h->nk;
  const float twopi32 = 2.0f * kPi32;
  const float ang = twopi32 / (float)R;
  h->angtab_rad.resize(R);
  h->angtab_deg.resize(R);
  h->polar_x.resize((size_t)nk * R);
  h->polar_y.resize((size_t)nk * R);
  std::vector<float> sn(R), cn(R);
  for (int i = 0; i < R; ++i) {
    const float a = (float)i * ang;                 // angtab(irot) = real(irot-1)*ang   (:265)
    h->angtab_rad[i] = a;
    sn[i] = std::sin(a);
    cn[i] = std::cos(a);
    for (int kk = 0; kk < nk; ++kk) {
      const float k = (float)(h->kfrom + kk);
      h->polar_x[(size_t)kk * R + i] = sn[i] * k;   // (:268)
      h->polar_y[(size_t)kk * R + i] = -cn[i] * k;  // (:269)
    }
    h->angtab_deg[i] = (a / kPi32) * 180.0f;        // rad2deg (:274)
  }
  const double A = kDPi / (double)(h->ldim / 2);    // (:311)
  h->argx.resize((size_t)nk * n);
  h->argy.resize((size_t)nk * n);
  h->onex.resize(n);
  h->oney.resize(n);
  h->sn32.resize(n);
  h->ncn32.resize(n);
  for (int j = 0; j < n; ++j) {
    h->sn32[j] = sn[j];
    h->ncn32[j] = -cn[j];
    h->onex[j] = (double)sn[j] * A;                 // argtransf_shellone (:313-314)
    h->oney[j] = (double)(-cn[j]) * A;
    for (int kk = 0; kk < nk; ++kk) {
      h->argx[(size_t)kk * n + j] = (double)h->polar_x[(size_t)kk * R + j] * A;   // (:316)
      h->argy[(size_t)kk * n + j] = (double)h->polar_y[(size_t)kk * R + j] * A;   // (:317)
    }
  }
  // spaFreqSq_mat / ang_mat (:927-936)
  const float inv = 1.0f / (float)h->ldim;
  h->spafreq.resize((size_t)nk * n);
  h->angmat.resize((size_t)nk * n);
  for (int kk = 0; kk < nk; ++kk) {
    for (int j = 0; j < n; ++j) {
      const float px = h->polar_x[(size_t)kk * R + j], py = h->polar_y[(size_t)kk * R + j];
      volatile float hinv = px * inv;
      volatile float kinv = py * inv;
      volatile float a2 = hinv * hinv;
      volatile float b2 = kinv * kinv;
      h->spafreq[(size_t)kk * n + j] = a2 + b2;
      h->angmat[(size_t)kk * n + j] = std::atan2(py, px);
    }
  }
  // setup_pxls_p_shell (:879-893)
  h->pxls.assign(nk, 0.0f);
  for (int a = 1; a <= h->kto; ++a)
    for (int b = 1; b <= h->kto; ++b) {
      const int sh = fnint(std::sqrt((float)(a * a + b * b)));
      if (sh >= h->kfrom && sh <= h->kto) h->pxls[sh - h->kfrom] += 1.0f;
    }
}

int alloc_ptcl_arrays(pftcc* h) {
  const size_t ring = (size_t)h->nk * h->n;
  int rc;
  if ((rc = dalloc(h, &h->d_ptcls, (size_t)h->nptcls * ring))) return rc;
  if ((rc = dalloc(h, &h->d_memo, (size_t)h->nptcls * h->nk * (h->n + 1)))) return rc;
  if ((rc = dalloc(h, &h->d_sqsums, (size_t)h->nptcls))) return rc;
  if ((rc = dalloc(h, &h->d_sqsums_k, (size_t)h->nptcls * h->nk))) return rc;
  if ((rc = dalloc(h, &h->d_iseven, (size_t)h->nptcls))) return rc;
  HIPCHK(h, sync_memset(h, h->d_ptcls, 0, (size_t)h->nptcls * ring * sizeof(float2)));
  HIPCHK(h, sync_memset(h, h->d_sqsums, 0, (size_t)h->nptcls * sizeof(float)));
  HIPCHK(h, sync_memset(h, h->d_sqsums_k, 0, (size_t)h->nptcls * h->nk * sizeof(float)));
  h->iseven.assign(h->nptcls, 1);
  HIPCHK(h, sync_copy(h, h->d_iseven, h->iseven.data(), h->nptcls, hipMemcpyHostToDevice));
  return PFTCC_OK;
}

void free_ptcl_arrays(pftcc* h) {
  dfree(h->d_ptcls);
  dfree(h->d_memo);
  dfree(h->d_sqsums);
  dfree(h->d_sqsums_k);
  dfree(h->d_iseven);
  dfree(h->d_ctf);
  dfree(h->d_dptcls);
  dfree(h->d_sqp8);
  dfree(h->d_shellw);
  dfree(h->d_sigma2w);
  dfree(h->d_optrow);
  h->memo_valid = h->dptcls_valid = h->sqp8_valid = h->ctf_ready = h->have_sigma2 = false;
  h->refnorm_valid = false;
}

int upload_iseven(pftcc* h) {
  HIPCHK(h, hipMemcpyAsync(h->d_iseven, h->iseven.data(), h->nptcls, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFTCC_OK;
}

// Device copy of pinds, refreshed whenever the host table changes (new, reallocate_ptcls).
int upload_pinds(pftcc* h) {
  const size_t span = h->pinds.size();
  if (span > h->d_pinds_cap) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    dfree(h->d_pinds);
    h->d_pinds_cap = 0;
    int rc;
    if ((rc = dalloc(h, &h->d_pinds, span))) return rc;
    h->d_pinds_cap = span;
  }
  HIPCHK(h, sync_copy(h, h->d_pinds, h->pinds.data(), span * sizeof(int), hipMemcpyHostToDevice));
  return PFTCC_OK;
}

// ------------------------------------------------------- scratch of the no-allocation search ---
// pftcc_search_dev_noalloc keeps its working memory (index lists, the shift table, the per-particle
// best keys) in this file-scope array instead of allocating: every GPU the library runs on holds
// one copy, at its own address, until the process ends.  Host code never passes the array's name
// to a kernel or a copy (on the host it is only a stand-in for the symbol): pftcc_new resolves
// each device's address once with hipGetSymbolAddress, and the entry point uses that address.
constexpr size_t kNoallocScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];
std::mutex g_scratch_mtx;
std::vector<std::pair<int, unsigned char*>> g_scratch_addr;   // (device, address of its scratch)

// With h->device current: this device's address of scratch, resolved on first use.
int resolve_noalloc_scratch(pftcc* h) {
  std::lock_guard<std::mutex> g(g_scratch_mtx);
  for (const auto& e : g_scratch_addr)
    if (e.first == h->device) {
      h->noalloc_scratch = e.second;
      return PFTCC_OK;
    }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(h, PFTCC_E_DEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s", h->device,
                hipGetErrorString(e));
  g_scratch_addr.emplace_back(h->device, (unsigned char*)p);
  h->noalloc_scratch = (unsigned char*)p;
  return PFTCC_OK;
}

int update_sqsums(pftcc* h, int first, int count) {
  if (count <= 0) return PFTCC_OK;
  hipLaunchKernelGGL(k_sqsums, dim3(count), dim3(256), 0, h->stream, (const float2*)h->d_ptcls, h->nk,
                     h->n, h->kcorr, h->d_sqsums, h->d_sqsums_k, first, count, (const int*)nullptr);
  HIPCHK(h, hipGetLastError());
  return PFTCC_OK;
}

// Shell normalisation factors (match_filt / pssnr): refinv = 1/sqrt(pw) (1 if pw <= 1e-12) and
// ofac = optlp or sqrt(optlp*k), in float and double (shellnorm_and_filter_ref(_8), :756-805).
int ensure_refnorm(pftcc* h) {
  if (!(h->match_filt || h->pssnr) || h->refnorm_valid) return PFTCC_OK;
  const int kc = h->kcorr, nr = h->nrefs;
  const size_t cnt = (size_t)2 * nr * kc;
  if (!h->d_refpw) {
    int rc;
    if ((rc = dalloc(h, &h->d_refpw, cnt))) return rc;
    if ((rc = dalloc(h, &h->d_refinv_d, cnt))) return rc;
    if ((rc = dalloc(h, &h->d_refinv_f, cnt))) return rc;
  }
  hipLaunchKernelGGL(k_ref_power, dim3((unsigned)cnt), dim3(256), 0, h->stream, (const float2*)h->d_refs_e,
                     (const float2*)h->d_refs_o, nr, h->nk, h->n, kc, h->d_refpw);
  HIPCHK(h, hipGetLastError());
  std::vector<double> pw(cnt), inv(cnt);
  std::vector<float> invf(cnt);
  HIPCHK(h, hipMemcpyAsync(pw.data(), h->d_refpw, cnt * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t i = 0; i < cnt; ++i) {
    inv[i] = pw[i] > 1e-12 ? 1.0 / std::sqrt(pw[i]) : 1.0;
    const float pwf = (float)pw[i];
    invf[i] = pwf > 1e-12f ? 1.0f / std::sqrt(pwf) : 1.0f;
  }
  // ofac rows: references (or particles when clsfrcs, rows indexed by the global particle id)
  const int rows = nr;
  std::vector<double> od((size_t)rows * kc);
  std::vector<float> of((size_t)rows * kc);
  for (int r = 0; r < rows; ++r)
    for (int kk = 0; kk < kc; ++kk) {
      const float o = h->optlp.empty() ? 1.0f : h->optlp[(size_t)r * kc + kk];
      const int k = h->kfrom + kk;
      if (h->pssnr) {
        od[(size_t)r * kc + kk] = std::sqrt((double)o * (double)k);
        of[(size_t)r * kc + kk] = std::sqrt(o * (float)k);
      } else {
        od[(size_t)r * kc + kk] = (double)o;
        of[(size_t)r * kc + kk] = o;
      }
    }
  if (!h->d_ofac_d) {
    int rc;
    if ((rc = dalloc(h, &h->d_ofac_d, (size_t)rows * kc))) return rc;
    if ((rc = dalloc(h, &h->d_ofac_f, (size_t)rows * kc))) return rc;
  }
  HIPCHK(h, sync_copy(h, h->d_refinv_d, inv.data(), cnt * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(h, sync_copy(h, h->d_refinv_f, invf.data(), cnt * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(h, sync_copy(h, h->d_ofac_d, od.data(), od.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(h, sync_copy(h, h->d_ofac_f, of.data(), of.size() * sizeof(float), hipMemcpyHostToDevice));
  if (h->clsfrcs) {
    // ref_optlp(k, iptcl) with the global particle id (:975-979)
    std::vector<int> row(h->nptcls, 0);
    for (int g = h->pfrom; g <= h->pto; ++g) {
      const int l = h->pinds[g - h->pfrom];
      if (l > 0) {
        if (g < 1 || g > nr) return fail(h, PFTCC_E_ARG, "clsfrcs needs particle id %d <= nrefs", g);
        row[l - 1] = g - 1;
      }
    }
    if (!h->d_optrow) {
      int rc;
      if ((rc = dalloc(h, &h->d_optrow, (size_t)h->nptcls))) return rc;
    }
    HIPCHK(h, sync_copy(h, h->d_optrow, row.data(), row.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  h->refnorm_valid = true;
  return PFTCC_OK;
}

size_t corr_lds_bytes(int n, int F) {
  return sizeof(float2) * ((size_t)2 * F * n + (size_t)F * (n + 1) + (n + 1) + 2 * n) +
         sizeof(float) * ((size_t)n + (size_t)F * n);
}

// item mode of k_corr_generic (F = 1): two FP64 phase arrays of n after the FP32 buffers
size_t corr_lds_bytes_items(int n) { return corr_lds_bytes(n, 1) + 16 + 2 * sizeof(double2) * (size_t)n; }

int pick_F(int n, int npairs) {
  int F = 8;
  while (F > 1 && corr_lds_bytes(n, F) > kMaxLds) F /= 2;
  while (F > 1 && F / 2 >= npairs) F /= 2;
  return F;
}

int timing_begin(pftcc* h, int ch, hipEvent_t* e1);
int timing_end(pftcc* h, hipEvent_t e1);

// Compile-time ring lengths of the fused kernel: the benchmark shapes (192, 308, 400), common
// box sizes (64 -> 100, 160 -> 240, 320 -> 500, 384 -> 600) and small sizes used in tests.
#ifndef PFTCC_FUSED_F_OVERRIDE
#define PFTCC_FUSED_F_OVERRIDE 0   // tile size for tuning builds (0: fused_F)
#endif
#ifndef PFTCC_FUSED_NT
#define PFTCC_FUSED_NT 256
#endif
#ifndef PFTCC_FUSED_RST
#define PFTCC_FUSED_RST 1           // staged reference rows (k_corr_fused RST) where a tile allows it
#endif
#ifndef PFTCC_FUSED_W_OVERRIDE
#define PFTCC_FUSED_W_OVERRIDE 0    // register budget for tuning builds (0: fused_W_tuned)
#endif
template <int N>
struct FusedLaunch {
  static constexpr int NT = PFTCC_FUSED_NT;
  static constexpr bool OVR = PFTCC_FUSED_F_OVERRIDE > 0 && fused_lds_bytes<N, PFTCC_FUSED_F_OVERRIDE + 0>() <= 65536;
  static constexpr int F = OVR ? PFTCC_FUSED_F_OVERRIDE : fused_F_tuned<N, NT>();
  static constexpr int FR = OVR ? PFTCC_FUSED_F_OVERRIDE : fused_F_rst_tuned<N, NT>();   // staged variant
  static constexpr int W = PFTCC_FUSED_W_OVERRIDE > 0 ? PFTCC_FUSED_W_OVERRIDE : fused_W_tuned<N, NT>();
  static constexpr size_t lds = fused_lds_bytes<N, F>();
  static constexpr size_t lds_r = fused_lds_bytes<N, FR>();
  static_assert(lds <= 65536 && lds_r <= 65536, "fused kernel LDS");
  static void go(unsigned blocks, hipStream_t s, const CorrArgs& a, bool rst) {
    if (rst)
      hipLaunchKernelGGL((k_corr_fused<N, FR, NT, W, true>), dim3(blocks), dim3(NT), lds_r, s, a);
    else
      hipLaunchKernelGGL((k_corr_fused<N, F, NT, W, false>), dim3(blocks), dim3(NT), lds, s, a);
  }
};

#define PFTCC_FUSED_SIZES(X) X(16) X(24) X(32) X(40) X(48) X(64) X(100) X(192) X(240) X(308) X(400) X(500) X(600)

// Tile (pairs per work-group) of the fused kernel for ring length n, 0 if n has no fused
// instantiation.  *rst: the staged-reference variant runs, which needs a tile to span at most two
// references (ns >= its tile), the CC objective and no shell normalisation (rst_ok).
int fused_tile(int n, bool rst_ok, int ns, bool* rst) {
#define X(N)                                                                   \
  if (n == N) {                                                                \
    *rst = PFTCC_FUSED_RST && rst_ok && ns >= FusedLaunch<N>::FR;              \
    return *rst ? FusedLaunch<N>::FR : FusedLaunch<N>::F;                      \
  }
  PFTCC_FUSED_SIZES(X)
#undef X
  *rst = false;
  return 0;
}

// The fused kernel addresses reference and shift rows by 32-bit byte offsets from their bases:
// references x shells x n and shifts x shells x n float2 must stay below 4 GiB.
bool fused_offsets_fit(const pftcc* h, int ns) {
  const size_t row = (size_t)h->nk * h->n * sizeof(float2);
  return (size_t)h->nrefs * row < ((size_t)1 << 32) && (size_t)std::max(ns, 1) * row < ((size_t)1 << 32);
}

void fused_go(int n, unsigned blocks, hipStream_t s, const CorrArgs& a, bool rst) {
#define X(N) if (n == N) { FusedLaunch<N>::go(blocks, s, a, rst); return; }
  PFTCC_FUSED_SIZES(X)
#undef X
}

// Launches the plain FP32 correlation for np particles x nr refs x ns shifts.
// rlist_ld > 0: d_rlist holds one list of nr references per particle slot (row stride rlist_ld).
int launch_corr(pftcc* h, int np, const int* d_plist, int nr, const int* d_rlist, int ns,
                const float2* d_shtab, bool shifted, float* d_cc, unsigned long long* d_best, int rlist_ld = 0) {
  // the staged-reference kernel reads the shift table of every shifted search without testing it
  if (shifted && !d_shtab) return fail(h, PFTCC_E_ARG, "shifted search without a shift table");
  CorrArgs a{};
  a.n = h->n;
  a.nk = h->nk;
  a.kcorr = h->kcorr;
  a.mode = h->euclid ? 1 : 0;
  // gencorrs_cc_2 normalises over all shells (:1531); gencorrs_cc_1 over kfrom..kstop (:984)
  a.knorm = (shifted && !h->euclid) ? h->nk : h->kcorr;
  a.nrefs = h->nrefs;
  a.plan = h->plan;
  a.refs_even = h->d_refs_e;
  a.refs_odd = h->d_refs_o;
  a.memo = h->d_memo;
  a.ctf = h->with_ctf ? h->d_ctf : nullptr;
  a.sqsums = h->d_sqsums;
  a.sqsums_k = h->d_sqsums_k;
  a.iseven = h->d_iseven;
  a.shtab = shifted ? d_shtab : nullptr;
  a.ns = shifted ? ns : 1;
  // gencorrs_euclid_2 does not shell-normalise (:1573-1577)
  const bool norm = (h->match_filt || h->pssnr) && !(h->euclid && shifted);
  a.refinv = norm ? h->d_refinv_f : nullptr;
  a.ofac = norm ? h->d_ofac_f : nullptr;
  a.optrow = (norm && h->clsfrcs) ? h->d_optrow : nullptr;
  a.shellw = h->euclid ? h->d_shellw : nullptr;
  a.plist = d_plist;
  a.np = np;
  a.rlist = d_rlist;
  a.nr = nr;
  a.rlist_ld = rlist_ld;
  a.cc_out = d_cc;
  a.best = d_best;
  a.tw2g = h->d_tw2;
  const int npairs = nr * a.ns;
  bool rst = false;
  const int Ff = h->use_fused && fused_offsets_fit(h, a.ns)
                     ? fused_tile(h->n, a.mode == 0 && !a.refinv && rlist_ld == 0, a.ns, &rst)
                     : 0;
  a.F = Ff > 0 ? Ff : pick_F(h->n, npairs);
  a.ntiles = (npairs + a.F - 1) / a.F;
  const size_t lds = corr_lds_bytes(h->n, a.F);
  const size_t blocks = (size_t)np * a.ntiles;
  if (blocks == 0) return PFTCC_OK;
  if (blocks > 0x7fffffffu) return fail(h, PFTCC_E_ARG, "search too large for one launch (%zu blocks)", blocks);
  hipEvent_t e1 = nullptr;
  int rc;
  if ((rc = timing_begin(h, 0, &e1))) return rc;
  if (Ff > 0) fused_go(h->n, (unsigned)blocks, h->stream, a, rst);
  else hipLaunchKernelGGL(k_corr_generic, dim3((unsigned)blocks), dim3(256), lds, h->stream, a);
  HIPCHK(h, hipGetLastError());
  return timing_end(h, e1);
}

int timing_begin(pftcc* h, int ch, hipEvent_t* e1) {
  *e1 = nullptr;
  if (!h->timing) return PFTCC_OK;
  if (h->nev == h->events.size()) {
    hipEvent_t x, y;
    HIPCHK(h, hipEventCreate(&x));
    HIPCHK(h, hipEventCreate(&y));
    h->events.emplace_back(x, y);
    h->event_ch.push_back(0);
  }
  HIPCHK(h, hipEventRecord(h->events[h->nev].first, h->stream));
  h->event_ch[h->nev] = ch;
  *e1 = h->events[h->nev].second;
  ++h->nev;
  return PFTCC_OK;
}

int timing_end(pftcc* h, hipEvent_t e1) {
  if (e1) HIPCHK(h, hipEventRecord(e1, h->stream));
  return PFTCC_OK;
}

int collect_timing(pftcc* h) {
  for (size_t i = 0; i < h->nev; ++i) {
    HIPCHK(h, hipEventSynchronize(h->events[i].second));
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->events[i].first, h->events[i].second));
    h->timed_ms[h->event_ch[i]] += ms;
    ++h->timed_launches[h->event_ch[i]];
  }
  h->nev = 0;
  return PFTCC_OK;
}

int shift_table(pftcc* h, int ns, const double* d_shifts, float2* d_shtab) {
  const int tot = ns * h->n;
  hipLaunchKernelGGL(k_shift_table, dim3((tot + 255) / 256), dim3(256), 0, h->stream, d_shifts, ns,
                     (const double*)h->d_argx, (const double*)h->d_argy, (const double*)h->d_onex,
                     (const double*)h->d_oney, h->nk, h->n, d_shtab);
  HIPCHK(h, hipGetLastError());
  return PFTCC_OK;
}

int need_memo(pftcc* h) {
  if (!h->memo_valid)
    return fail(h, PFTCC_E_STATE, "particle FFTs are not memoized: call memoize_ffts after setting particles");
  if (h->with_ctf && !h->ctf_ready)
    return fail(h, PFTCC_E_STATE, "CTF matrices missing: call create_polar_absctfmats");
  if (h->euclid && !h->have_sigma2)
    return fail(h, PFTCC_E_STATE, "objfun euclid needs assign_sigma2_noise");
  return ensure_refnorm(h);
}

int need_direct(pftcc* h) {
  if (h->with_ctf && !h->ctf_ready)
    return fail(h, PFTCC_E_STATE, "CTF matrices missing: call create_polar_absctfmats");
  if (h->euclid && !h->have_sigma2)
    return fail(h, PFTCC_E_STATE, "objfun euclid needs assign_sigma2_noise");
  return ensure_refnorm(h);
}

}  // namespace

// Angular derivative rings dP/dpsi of every particle (spectral derivative of the full ring).
__global__ void k_deriv(const float2* __restrict__ memo, int nitems, int nk, int n, FftPlan plan,
                        const float2* __restrict__ tw2g, float2* __restrict__ dptcls, int F) {
  extern __shared__ float2 smem[];
  // 2 transforms per item: even (Ge) and odd (Go) frequency parts
  const int T = 2 * F;
  float2* bufA = smem;
  float2* bufB = bufA + (size_t)T * n;
  float2* tw2 = bufB + (size_t)T * n;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int i = tid; i < 2 * n; i += nt) tw2[i] = tw2g[i];
  const int item0 = blockIdx.x * F;
  const int half = n / 2;
  for (int idx = tid; idx < T * n; idx += nt) {
    const int tr = idx / n, u = idx - tr * n;
    const int f = tr >> 1, odd = tr & 1;
    const int item = item0 + f;
    float2 g = make_float2(0.f, 0.f);
    if (item < nitems) {
      const float2* W = memo + (size_t)item * (n + 1);
      const int mp = u < half ? u : u - n;          // m' in [-n/2, n/2)
      const int m = 2 * mp + odd;                   // frequency on the 2n ring
      if (!(odd == 0 && mp == -half)) {             // m = -n: Nyquist, zero derivative
        const int am = m < 0 ? -m : m;
        float2 Fp = W[am];
        if (m < 0) Fp = cconj(Fp);
        if (odd) Fp = make_float2(-Fp.y, Fp.x);      // odd bins carry a factor i
        // G = i m F_P
        g = make_float2(-(float)m * Fp.y, (float)m * Fp.x);
      }
    }
    bufA[(size_t)tr * n + u] = cconj(g);             // inverse via conj(FFT(conj))
  }
  __syncthreads();
  float2* res = fft_forward_oop(bufA, bufB, T, n, n, plan, tw2, tid, nt);
  const float inv2n = 1.0f / (2.0f * (float)n);
  for (int idx = tid; idx < F * n; idx += nt) {
    const int f = idx / n, t = idx - f * n;
    const int item = item0 + f;
    if (item >= nitems) continue;
    const float2 e = cconj(res[(size_t)(2 * f) * n + t]);
    const float2 o = cconj(res[(size_t)(2 * f + 1) * n + t]);
    const float2 ph = cconj(tw2[t]);                 // exp(+i pi t / n)
    const float2 po = cmul(ph, o);
    dptcls[(size_t)item * n + t] = make_float2((e.x + po.x) * inv2n, (e.y + po.y) * inv2n);
  }
}

namespace {

int ensure_dptcls(pftcc* h) {
  if (h->dptcls_valid) return PFTCC_OK;
  int rc;
  if ((rc = need_memo(h))) return rc;
  if (!h->d_dptcls && (rc = dalloc(h, &h->d_dptcls, (size_t)h->nptcls * h->nk * h->n))) return rc;
  int F = 4;
  while (F > 1 && sizeof(float2) * ((size_t)2 * (2 * F) * h->n + 2 * h->n) > kMaxLds) F /= 2;
  const int nitems = h->nptcls * h->nk;
  const size_t lds = sizeof(float2) * ((size_t)2 * (2 * F) * h->n + 2 * h->n);
  hipLaunchKernelGGL(k_deriv, dim3((nitems + F - 1) / F), dim3(256), lds, h->stream,
                     (const float2*)h->d_memo, nitems, h->nk, h->n, h->plan, (const float2*)h->d_tw2,
                     h->d_dptcls, F);
  HIPCHK(h, hipGetLastError());
  h->dptcls_valid = true;
  return PFTCC_OK;
}

// FP64 per-shell sums of |P|^2 of every particle (the k-weighted objective's particle norm).
int ensure_sqp8(pftcc* h) {
  if (h->sqp8_valid) return PFTCC_OK;
  int rc;
  if (!h->d_sqp8 && (rc = dalloc(h, &h->d_sqp8, (size_t)h->nptcls * h->nk))) return rc;
  const int nitems = h->nptcls * h->nk;
  if (nitems > 0)
    hipLaunchKernelGGL(k_shell_sq8, dim3((nitems + 3) / 4), dim3(256), 0, h->stream, (const float2*)h->d_ptcls, nitems,
                       h->n, h->d_sqp8);
  HIPCHK(h, hipGetLastError());
  h->sqp8_valid = true;
  return PFTCC_OK;
}

// Objective form of the FP64 single-rotation routines: 0 CC k-weighted (match_filt = no), 1 CC
// match_filt, 2 Euclid.  objfun < 0 follows the handle's cc_objfun (gencorr_for_rot_8's dispatch,
// :1632-1643); PFTCC_OBJFUN_CC / _EUCLID select gencorr_cc_* / gencorr_euclid_* directly.
int rot8_mode(const pftcc* h, int objfun) {
  const bool euclid = objfun < 0 ? h->euclid : objfun == PFTCC_OBJFUN_EUCLID;
  return euclid ? 2 : (h->match_filt ? 1 : 0);
}

// Launches k_rot8 on device lists.
int launch_rot8(pftcc* h, int npairs, const int* d_irefs, const int* d_ips, const double* d_shifts,
                const int* d_rhos, double* d_f, double* d_grad, bool angle, int objfun = -1) {
  Rot8Args a{};
  a.n = h->n;
  a.nk = h->nk;
  a.kcorr = h->kcorr;
  a.kfrom = h->kfrom;
  a.nrefs = h->nrefs;
  a.mode = rot8_mode(h, objfun);
  a.refs_even = h->d_refs_e;
  a.refs_odd = h->d_refs_o;
  a.ptcls = h->d_ptcls;
  a.dptcls = angle ? h->d_dptcls : nullptr;
  a.ctf = h->with_ctf ? h->d_ctf : nullptr;
  a.sqsums = h->d_sqsums;
  a.iseven = h->d_iseven;
  const bool norm = (h->match_filt || h->pssnr) && a.mode != 2;
  a.refinv = norm ? h->d_refinv_d : nullptr;
  a.ofac = norm ? h->d_ofac_d : nullptr;
  a.optrow = (norm && h->clsfrcs) ? h->d_optrow : nullptr;
  a.argx = h->d_argx;
  a.argy = h->d_argy;
  a.sn32 = h->d_sn32;
  a.ncn32 = h->d_ncn32;
  a.arg_a = kDPi / (double)(h->ldim / 2);
  a.one_x = h->d_onex;
  a.one_y = h->d_oney;
  a.sigma2w = h->d_sigma2w;
  a.irefs = d_irefs;
  a.ips = d_ips;
  a.shifts = d_shifts;
  a.rhos = d_rhos;
  a.npairs = npairs;
  a.f = d_f;
  a.grad = d_grad;
  if (npairs <= 0) return PFTCC_OK;
  int rc;
  if (a.mode == 0) {
    if ((rc = ensure_sqp8(h))) return rc;
    a.sqp8 = h->d_sqp8;
  }
  hipEvent_t ev;
  if ((rc = timing_begin(h, 4, &ev))) return rc;
  // one thread per ring position j (rounded up to whole wavefronts; k_rot8's bound is 512, larger
  // rings loop over j)
  const int nt = std::min(512, (h->n + 63) / 64 * 64);
  // the shell normalisation (match_filt / pssnr) is a template argument: without it the kernel
  // carries no per-shell filter factor
  const bool nrm = a.refinv != nullptr;
  if (a.mode == 0 && !nrm) hipLaunchKernelGGL((k_rot8<PFTCC_ROT8_AHEAD, 0, false>), dim3(npairs), dim3(nt), 0, h->stream, a);
  else if (a.mode == 0) hipLaunchKernelGGL((k_rot8<PFTCC_ROT8_AHEAD, 0, true>), dim3(npairs), dim3(nt), 0, h->stream, a);
  else if (a.mode == 1 && !nrm) hipLaunchKernelGGL((k_rot8<PFTCC_ROT8_AHEAD, 1, false>), dim3(npairs), dim3(nt), 0, h->stream, a);
  else if (a.mode == 1) hipLaunchKernelGGL((k_rot8<PFTCC_ROT8_AHEAD, 1, true>), dim3(npairs), dim3(nt), 0, h->stream, a);
  else hipLaunchKernelGGL((k_rot8<PFTCC_ROT8_AHEAD, 2, false>), dim3(npairs), dim3(nt), 0, h->stream, a);
  HIPCHK(h, hipGetLastError());
  return timing_end(h, ev);
}

// ------------------------------------------------------------------ precision form ---
int round_up(int x, int m) { return (x + m - 1) / m * m; }

#ifndef PFTCC_ZGEMM_AS
#define PFTCC_ZGEMM_AS 1       // A-stationary contraction kernel (0: one tile per wavefront)
#endif
#ifndef PFTCC_ZGEMM_AS_TR
#define PFTCC_ZGEMM_AS_TR 2
#endif
#ifndef PFTCC_MFMA_CHUNK
#define PFTCC_MFMA_CHUNK 256   // particles per contraction / inverse-stage launch (multiple of 64)
#endif
#ifndef PFTCC_MFMA_ZCAP_GIB
#define PFTCC_MFMA_ZCAP_GIB 4  // cap on the spectra Z of one chunk
#endif
// Operand formats of the precision form.
enum { OP_BF16 = 0, OP_F16 = 1, OP_FP8 = 2 };
int op_format(int prec) {
  return prec == PFTCC_PREC_FP8 ? OP_FP8 : (prec == PFTCC_PREC_FP16 || prec == PFTCC_PREC_SPLIT_FP16) ? OP_F16 : OP_BF16;
}

// A-stationary contraction for KP = 32 NSTEP; false if NSTEP has no instantiation.
template <int SPLIT, bool H>
bool zgemm_as_go(int KP, int rows_pad, int cols_pad, int n, hipStream_t s, const unsigned int* ahi,
                 const unsigned int* alo, const unsigned int* bhi, const unsigned int* blo, float2* Z,
                 const unsigned int* gA, const unsigned int* gB) {
  constexpr int TR = PFTCC_ZGEMM_AS_TR;
  const dim3 grid((unsigned)((size_t)(rows_pad / (16 * TR)) * (n + 1)));
#define X(NS)                                                                                         \
  if (KP == 32 * NS) {                                                                                \
    hipLaunchKernelGGL((k_zgemm_as<SPLIT, TR, NS, H>), grid, dim3(64), 0, s, (const bf16x8*)ahi,       \
                       (const bf16x8*)alo, (const bf16x8*)bhi, (const bf16x8*)blo, rows_pad, cols_pad, n, Z, \
                       gA, gB);                                                                       \
    return true;                                                                                      \
  }
  X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8)
#undef X
  return false;
}

// Compile-time inverse stage for the fused ring lengths; false if n has none.  a.Z16 selects the
// FP16-spectra variant.
bool zifft_ct_go(int n, int cnt, int nr, hipStream_t s, const ZifftArgs& a) {
#define X(N)                                                                                          \
  if (n == N) {                                                                                       \
    const dim3 grid((unsigned)((cnt + ZifftV2<N>::F - 1) / ZifftV2<N>::F * nr));                      \
    if (a.Z16)                                                                                        \
      hipLaunchKernelGGL((k_zifft_v2<N, ZifftV2<N>::F, ZifftV2<N>::NT, true>), grid,                  \
                         dim3(ZifftV2<N>::NT), ZifftV2<N>::lds, s, a);                                \
    else                                                                                              \
      hipLaunchKernelGGL((k_zifft_v2<N, ZifftV2<N>::F, ZifftV2<N>::NT, false>), grid,                 \
                         dim3(ZifftV2<N>::NT), ZifftV2<N>::lds, s, a);                                \
    return true;                                                                                      \
  }
  PFTCC_FUSED_SIZES(X)
#undef X
  return false;
}

bool zifft_ct_has(int n) {
#define X(N) if (n == N) return true;
  PFTCC_FUSED_SIZES(X)
#undef X
  return false;
}

size_t zifft_lds(int F, int n) { return sizeof(float2) * ((size_t)F * (n + 1) + (size_t)F * n + 2 * n); }

void free_apack(pftcc::APack& ap) {
  dfree(ap.hi);
  dfree(ap.lo);
  dfree(ap.sr);
  dfree(ap.d_ak);
  dfree(ap.d_bk);
  dfree(ap.gmax);
  dfree(ap.a8);
  ap.valid = false;
}

// FP8 shell scales a_k = 2^round(log2(msW_k / msV_k) / 2): conj(a_k V) (W / a_k) = conj(V) W.
std::vector<float> fp8_shell_scales(const std::vector<double>& msV, const std::vector<double>& msW) {
  std::vector<float> ak(msV.size(), 1.0f);
  for (size_t k = 0; k < msV.size(); ++k)
    if (msV[k] > 0.0 && msW[k] > 0.0) ak[k] = (float)std::ldexp(1.0, (int)std::lround(0.5 * std::log2(msW[k] / msV[k])));
  return ak;
}

// Per-shell mean |spec|^2 over count listed rows (list null: 0..count-1), deterministic order.
int shell_ms(pftcc* h, const float2* spec, const int* d_list, int count, std::vector<double>& out) {
  const int kc = h->kcorr, nch = 64;
  double* part = nullptr;
  int rc;
  if ((rc = dalloc(h, &part, (size_t)kc * nch))) return rc;
  hipLaunchKernelGGL(k_shell_ms, dim3(kc, nch), dim3(256), 0, h->stream, spec, d_list, count, h->nk, h->n, part);
  std::vector<double> hp((size_t)kc * nch);
  const bool ok = hipGetLastError() == hipSuccess &&
                  hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream) == hipSuccess &&
                  hipStreamSynchronize(h->stream) == hipSuccess;
  dfree(part);
  if (!ok) return fail(h, PFTCC_E_DEVICE, "shell statistics failed");
  out.assign(kc, 0.0);
  for (int k = 0; k < kc; ++k) {
    double s = 0.0;
    for (int c = 0; c < nch; ++c) s += hp[(size_t)k * nch + c];
    out[k] = s / ((double)count * (h->n + 1));
  }
  return PFTCC_OK;
}

// Reference operand of eo class eo over the (0-based) reference list rl; cached.  Split BF16, or
// split FP16 / FP8 e4m3 with shell scales balanced against the particles' mean squares msW.
int ensure_apack(pftcc* h, int eo, const std::vector<int>& rl, const std::vector<float>& shv, int fmt,
                 const std::vector<double>* msW) {
  pftcc::APack& ap = h->apack[eo];
  const bool fp8 = fmt == OP_FP8, scaled = fmt != OP_BF16;
  if (ap.valid && ap.refs == rl && ap.shifts == shv && ap.kcorr == h->kcorr && ap.fmt == fmt &&
      (!scaled || fp8_shell_scales(ap.msV, *msW) == ap.ak))
    return PFTCC_OK;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  free_apack(ap);
  const int nref = (int)rl.size();
  const int ns = (int)shv.size() / 2, nse = ns > 0 ? ns : 1;
  const int nr = nref * nse;                          // operand rows (reference, shift)
  const int knorm = ns > 0 ? h->nk : h->kcorr;
  const int n = h->n, nk = h->nk;
  ap.rows_pad = round_up(nr, 64);
  ap.KP = round_up(2 * h->kcorr, 32);
  ap.KP8 = round_up(2 * h->kcorr, 128);
  const size_t words = fp8 ? (size_t)(n + 1) * ap.rows_pad * (ap.KP8 / 4) : (size_t)(n + 1) * ap.rows_pad * (ap.KP / 2);
  int rc;
  if ((rc = dalloc(h, &ap.sr, (size_t)nref * knorm))) return rc;
  if (scaled && ((rc = dalloc(h, &ap.d_ak, (size_t)h->kcorr)) || (rc = dalloc(h, &ap.d_bk, (size_t)h->kcorr)) ||
                 (rc = dalloc(h, &ap.gmax, (size_t)1))))
    return rc;
  if (fp8) {
    if ((rc = dalloc(h, &ap.a8, words))) return rc;
  } else if ((rc = dalloc(h, &ap.hi, words)) || (rc = dalloc(h, &ap.lo, words))) {
    return rc;
  }
  int* d_list = nullptr;
  float2 *rings = nullptr, *spec = nullptr, *srings = nullptr, *shtab = nullptr;
  double* d_sh = nullptr;
  // the temporaries are released on every return path (hipFree waits for the device)
  const auto temps = scope_exit([&] {
    dfree(d_list);
    dfree(rings);
    dfree(spec);
    dfree(srings);
    dfree(shtab);
    dfree(d_sh);
  });
  if ((rc = dalloc(h, &d_list, (size_t)nref)) || (rc = dalloc(h, &rings, (size_t)nref * nk * n)) ||
      (rc = dalloc(h, &spec, (size_t)nr * nk * (n + 1))))
    return rc;
  HIPCHK(h, hipMemcpyAsync(d_list, rl.data(), nref * sizeof(int), hipMemcpyHostToDevice, h->stream));
  const size_t rowlen = (size_t)nk * n;
  hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((nref * rowlen + 255) / 256)), dim3(256), 0, h->stream,
                     (const float2*)(eo ? h->d_refs_o : h->d_refs_e), (const int*)d_list, nref, rowlen, rings);
  float2* rowrings = rings;
  if (ns > 0) {
    // q = ref x shmat, with the same FP32 shift table as the plain search (gen_shmat, :1123-1130)
    std::vector<double> sh64(shv.begin(), shv.end());
    if ((rc = dalloc(h, &d_sh, sh64.size())) || (rc = dalloc(h, &shtab, (size_t)ns * rowlen)) ||
        (rc = dalloc(h, &srings, (size_t)nr * rowlen)))
      return rc;
    HIPCHK(h, hipMemcpyAsync(d_sh, sh64.data(), sh64.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if ((rc = shift_table(h, ns, d_sh, shtab))) return rc;
    hipLaunchKernelGGL(k_shift_rings, dim3((unsigned)(((size_t)nr * rowlen + 255) / 256)), dim3(256), 0, h->stream,
                       (const float2*)rings, (const float2*)shtab, nref, ns, nk, n, srings);
    rowrings = srings;
  }
  int F = 8;
  while (F > 1 && sizeof(float2) * ((size_t)2 * F * n + 2 * n) > kMaxLds) F /= 2;
  hipLaunchKernelGGL(k_memoize, dim3((nr * nk + F - 1) / F), dim3(256), sizeof(float2) * ((size_t)2 * F * n + 2 * n),
                     h->stream, (const float2*)rowrings, 0, nr, nk, n, h->plan, (const float2*)h->d_tw2, spec, F,
                     (const int*)nullptr);
  if (scaled) {
    if ((rc = shell_ms(h, spec, nullptr, nr, ap.msV))) return rc;
    ap.ak = fp8_shell_scales(ap.msV, *msW);
    std::vector<float> bk(ap.ak.size());
    for (size_t k = 0; k < bk.size(); ++k) bk[k] = 1.0f / ap.ak[k];
    HIPCHK(h, hipMemcpyAsync(ap.d_ak, ap.ak.data(), ap.ak.size() * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(ap.d_bk, bk.data(), bk.size() * 4, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, h->stream, ap.gmax, 1, 0u);
    const size_t tot = (size_t)nr * h->kcorr * (n + 1);
    hipLaunchKernelGGL(k_absmax_scaled, dim3((unsigned)std::min<size_t>(1024, (tot + 255) / 256)), dim3(256), 0, h->stream,
                       (const float2*)spec, (const float*)nullptr, (const float*)ap.d_ak, nr, nk, n, h->kcorr, ap.gmax);
  }
  if (fp8) {
    hipLaunchKernelGGL(k_pack_fp8, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, h->stream, (const float2*)spec,
                       (const float*)nullptr, (const float*)ap.d_ak, (const unsigned int*)ap.gmax, nr, ap.rows_pad, nk,
                       h->kcorr, n, ap.KP8, ap.a8);
  } else if (scaled) {
    hipLaunchKernelGGL(k_pack_split<true>, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, h->stream,
                       (const float2*)spec, (const float*)nullptr, (const float*)ap.d_ak, (const unsigned int*)ap.gmax,
                       nr, ap.rows_pad, nk, h->kcorr, n, ap.KP, ap.hi, ap.lo);
  } else {
    hipLaunchKernelGGL(k_pack_split<false>, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, h->stream,
                       (const float2*)spec, (const float*)nullptr, (const float*)nullptr, (const unsigned int*)nullptr,
                       nr, ap.rows_pad, nk, h->kcorr, n, ap.KP, ap.hi, ap.lo);
  }
  hipLaunchKernelGGL(k_ring_sq64, dim3((nref * knorm + 3) / 4), dim3(64, 4), 0, h->stream, (const float2*)rings, nref,
                     nk, n, knorm, ap.sr);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  dfree(d_list);
  dfree(rings);
  dfree(spec);
  dfree(srings);
  dfree(shtab);
  dfree(d_sh);
  ap.refs = rl;
  ap.shifts = shv;
  ap.knorm = knorm;
  ap.kcorr = h->kcorr;
  ap.fmt = fmt;
  ap.valid = true;
  return PFTCC_OK;
}

#ifndef PFTCC_RESCORE_ORDER
#define PFTCC_RESCORE_ORDER 1   // rescoring work-groups take particles grouped by their best candidate's row
#endif
#ifndef PFTCC_FP8_Z16
#define PFTCC_FP8_Z16 1   // FP8 coarse pass: spectra Z stored as FP16 pairs (half the round trip's bytes)
#endif

// Runs the precision form for particles pl (local ids) that share eo class eo, against the cached
// operand of refs rl.  outpos[i] = position of pl[i] in best/cc_out.
int mfma_group(pftcc* h, int eo, const std::vector<int>& pl, const std::vector<int>& outpos,
               const std::vector<int>& rl, const std::vector<float>& shv, int prec, unsigned long long* d_best,
               float* d_cc, unsigned long long* d_prbest, const std::vector<double>* msW) {
  int rc;
  const int fmt = op_format(prec);
  const bool fp8 = fmt == OP_FP8, f16 = fmt == OP_F16;
  if ((rc = ensure_apack(h, eo, rl, shv, fmt, msW))) return rc;
  const pftcc::APack& ap = h->apack[eo];
  const int nse = std::max(1, (int)shv.size() / 2);
  const int n = h->n, nk = h->nk, KP = fp8 ? ap.KP8 : ap.KP, nr = (int)rl.size() * nse;
  // key id of each operand row: global reference * nse + shift
  std::vector<int> rowkey(nr);
  for (int r = 0; r < nr; ++r) rowkey[r] = rl[r / nse] * nse + r % nse;
  const int total = (int)pl.size();
  // FP8 coarse pass: Z as FP16 pairs, scaled by the power of two zs16 <= 65504 / (KP 448^2), which
  // bounds |Z gA gB| over e4m3 operands (|x| <= 448) so that no value overflows FP16
  const bool z16 = fp8 && PFTCC_FP8_Z16 && zifft_ct_has(n);
  double zs16 = 1.0;
  while (zs16 * KP * 448.0 * 448.0 > 65504.0) zs16 *= 0.5;
  // chunk of particles: keep Z (n+1) x rows_pad x Pc float2 within the cap
  int Pc = PFTCC_MFMA_CHUNK;
  while (Pc > 64 && (size_t)(n + 1) * ap.rows_pad * Pc * 8 > ((size_t)PFTCC_MFMA_ZCAP_GIB << 30)) Pc -= 64;
  Pc = std::min(Pc, round_up(total, 64));
  const size_t words = (size_t)(n + 1) * Pc * (fp8 ? KP / 4 : KP / 2);
  const size_t zsz = (size_t)(n + 1) * ap.rows_pad * Pc;
  if ((rc = scratch_reserve(h, aligned((size_t)total * 4) * 2 + aligned(nr * 4) + aligned((size_t)Pc * nk * (n + 1) * 8) +
                                   aligned((size_t)Pc * nk * 4) + aligned((size_t)Pc * nk * 8) + aligned((size_t)total * 8) +
                                   2 * aligned(words * 4) + aligned(zsz * 8) + aligned(4))))
    return rc;
  int* d_pl = take<int>(h, total);
  int* d_pos = take<int>(h, total);
  int* d_rl = take<int>(h, nr);
  float2* wg = take<float2>(h, (size_t)Pc * nk * (n + 1));
  float* cbar = take<float>(h, (size_t)Pc * nk);
  double* cbar2 = take<double>(h, (size_t)Pc * nk);
  double* spall = take<double>(h, total);
  unsigned int* bhi = take<unsigned int>(h, words);
  unsigned int* blo = take<unsigned int>(h, words);
  float2* Z = take<float2>(h, zsz);
  unsigned int* gmaxB = take<unsigned int>(h, 1);
  HIPCHK(h, hipMemcpyAsync(d_pl, pl.data(), total * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d_pos, outpos.data(), total * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d_rl, rowkey.data(), nr * 4, hipMemcpyHostToDevice, h->stream));
  // FP64 particle sums of squares over the correlated shells, all particles of the group at once
  hipLaunchKernelGGL(k_ring_sq64_list, dim3((total + 3) / 4), dim3(64, 4), 0, h->stream, (const float2*)h->d_ptcls,
                     (const int*)d_pl, total, nk, n, h->kcorr, spall);
  for (int c0 = 0; c0 < total; c0 += Pc) {
    const int cnt = std::min(Pc, total - c0);
    hipEvent_t ev;
    if ((rc = timing_begin(h, 3, &ev))) return rc;
    const size_t rowlen = (size_t)nk * (n + 1);
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((cnt * rowlen + 255) / 256)), dim3(256), 0, h->stream,
                       (const float2*)h->d_memo, (const int*)(d_pl + c0), cnt, rowlen, wg);
    const bool ctf = h->with_ctf;
    if (ctf)
      hipLaunchKernelGGL(k_ring_mean_ctf_list, dim3((cnt * nk + 3) / 4), dim3(64, 4), 0, h->stream,
                         (const float*)h->d_ctf, (const int*)(d_pl + c0), cnt, nk, n, cbar, cbar2);
    if (fp8 || f16) {
      hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, h->stream, gmaxB, 1, 0u);
      const size_t tot = (size_t)cnt * h->kcorr * (n + 1);
      hipLaunchKernelGGL(k_absmax_scaled, dim3((unsigned)std::min<size_t>(1024, (tot + 255) / 256)), dim3(256), 0, h->stream,
                         (const float2*)wg, ctf ? (const float*)cbar : nullptr, (const float*)ap.d_bk, cnt, nk, n,
                         h->kcorr, gmaxB);
    }
    if (fp8) {
      hipLaunchKernelGGL(k_pack_fp8, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, h->stream,
                         (const float2*)wg, ctf ? (const float*)cbar : nullptr, (const float*)ap.d_bk,
                         (const unsigned int*)gmaxB, cnt, Pc, nk, h->kcorr, n, KP, bhi);
    } else if (f16) {
      hipLaunchKernelGGL(k_pack_split<true>, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, h->stream,
                         (const float2*)wg, ctf ? (const float*)cbar : nullptr, (const float*)ap.d_bk,
                         (const unsigned int*)gmaxB, cnt, Pc, nk, h->kcorr, n, KP, bhi, blo);
    } else {
      hipLaunchKernelGGL(k_pack_split<false>, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, h->stream,
                         (const float2*)wg, ctf ? (const float*)cbar : nullptr, (const float*)nullptr,
                         (const unsigned int*)nullptr, cnt, Pc, nk, h->kcorr, n, KP, bhi, blo);
    }
    HIPCHK(h, hipGetLastError());
    if ((rc = timing_end(h, ev)) || (rc = timing_begin(h, 1, &ev))) return rc;
    // FP8: MT consecutive m per wavefront where one k' step covers the shells (KP = 128)
    const int mt8 = KP == 128 ? ZGEMM8_MT : 1;
    const dim3 gg((unsigned)((size_t)(Pc / (16 * ZGEMM8_TJ)) * (ap.rows_pad / 64) * ((n + 1 + mt8 - 1) / mt8)));
    const dim3 gz((unsigned)((size_t)(Pc / (16 * ZGEMM_TP)) * (ap.rows_pad / (16 * ZGEMM_TR)) * (n + 1)));
    if (fp8) {
#define Z8_LAUNCH(Z16, MT)                                                                             \
  hipLaunchKernelGGL((k_zgemm8<Z16, MT>), gg, dim3(64), 0, h->stream, (const i32x8*)ap.a8, (const i32x8*)bhi, \
                     ap.rows_pad, Pc, KP, n, Z, (const unsigned int*)ap.gmax, (const unsigned int*)gmaxB,  \
                     (float)zs16)
      if (z16 && mt8 > 1) Z8_LAUNCH(true, ZGEMM8_MT);
      else if (z16) Z8_LAUNCH(true, 1);
      else if (mt8 > 1) Z8_LAUNCH(false, ZGEMM8_MT);
      else Z8_LAUNCH(false, 1);
#undef Z8_LAUNCH
    } else if (PFTCC_ZGEMM_AS && prec == PFTCC_PREC_BF16 &&
             zgemm_as_go<1, false>(KP, ap.rows_pad, Pc, n, h->stream, ap.hi, ap.lo, bhi, blo, Z, nullptr, nullptr)) {
    } else if (PFTCC_ZGEMM_AS && prec == PFTCC_PREC_SPLIT_BF16 &&
               zgemm_as_go<3, false>(KP, ap.rows_pad, Pc, n, h->stream, ap.hi, ap.lo, bhi, blo, Z, nullptr, nullptr)) {
    } else if (PFTCC_ZGEMM_AS && prec == PFTCC_PREC_FP16 &&
               zgemm_as_go<1, true>(KP, ap.rows_pad, Pc, n, h->stream, ap.hi, ap.lo, bhi, blo, Z, ap.gmax, gmaxB)) {
    } else if (PFTCC_ZGEMM_AS && prec == PFTCC_PREC_SPLIT_FP16 &&
               zgemm_as_go<3, true>(KP, ap.rows_pad, Pc, n, h->stream, ap.hi, ap.lo, bhi, blo, Z, ap.gmax, gmaxB)) {
    } else {
#define ZG_LAUNCH(S, H)                                                                                  \
  hipLaunchKernelGGL((k_zgemm<S, ZGEMM_TR, ZGEMM_TP, H>), gz, dim3(64), 0, h->stream, (const bf16x8*)ap.hi,  \
                     (const bf16x8*)ap.lo, (const bf16x8*)bhi, (const bf16x8*)blo, ap.rows_pad, Pc, KP, n, Z,   \
                     (const unsigned int*)ap.gmax, (const unsigned int*)gmaxB)
      if (prec == PFTCC_PREC_BF16) ZG_LAUNCH(1, false);
      else if (prec == PFTCC_PREC_SPLIT_BF16) ZG_LAUNCH(3, false);
      else if (prec == PFTCC_PREC_FP16) ZG_LAUNCH(1, true);
      else ZG_LAUNCH(3, true);
#undef ZG_LAUNCH
    }
    HIPCHK(h, hipGetLastError());
    if ((rc = timing_end(h, ev)) || (rc = timing_begin(h, 2, &ev))) return rc;
    ZifftArgs za{};
    za.n = n;
    za.nrefs = nr;
    za.np = cnt;
    za.rows_pad = ap.rows_pad;
    za.cols_pad = Pc;
    za.kcorr = h->kcorr;
    za.ns = nse;
    za.knorm = ap.knorm;
    za.Z = Z;
    za.sr = ap.sr;
    za.cbar2 = ctf ? cbar2 : nullptr;
    za.nk = nk;
    za.sp = spall + c0;
    za.rlist = d_rl;
    za.outpos = d_pos + c0;
    za.cc_out = d_cc;
    za.best = d_best;
    za.prbest = d_prbest;
    za.tw2g = h->d_tw2;
    za.plan = h->plan;
    if (z16) {
      za.Z16 = reinterpret_cast<const unsigned int*>(Z);
      za.gmaxA = (const unsigned int*)ap.gmax;
      za.gmaxB = (const unsigned int*)gmaxB;
      za.zs = zs16;
    }
    if (zifft_ct_go(n, cnt, nr, h->stream, za)) {
      HIPCHK(h, hipGetLastError());
      if ((rc = timing_end(h, ev))) return rc;
      continue;
    }
    if (nr > 65535) return fail(h, PFTCC_E_UNSUPPORTED, "pftsz %d: at most 65535 operand rows per call", n);
    int Fz = 8;
    while (Fz > 2 && zifft_lds(Fz, n) > kMaxLds) Fz /= 2;
    const size_t lds = zifft_lds(Fz, n);
    if (Fz == 8)
      hipLaunchKernelGGL(k_zifft<8>, dim3((cnt + 7) / 8, nr), dim3(256), lds, h->stream, za);
    else if (Fz == 4)
      hipLaunchKernelGGL(k_zifft<4>, dim3((cnt + 3) / 4, nr), dim3(256), lds, h->stream, za);
    else
      hipLaunchKernelGGL(k_zifft<2>, dim3((cnt + 1) / 2, nr), dim3(256), lds, h->stream, za);
    HIPCHK(h, hipGetLastError());
    if ((rc = timing_end(h, ev))) return rc;
  }
  return PFTCC_OK;
}

// Common driver: validates, splits particles by eo class and runs both groups.
int mfma_run(pftcc* h, int np, const int32_t* iptcls, int nr, const int32_t* irefs, const std::vector<float>& shv,
             int prec, unsigned long long* d_best, float* d_cc, unsigned long long* d_prbest = nullptr,
             std::vector<int>* plocal = nullptr) {
  int rc;
  if (prec != PFTCC_PREC_BF16 && prec != PFTCC_PREC_SPLIT_BF16 && prec != PFTCC_PREC_FP16 &&
      prec != PFTCC_PREC_SPLIT_FP16 && prec != PFTCC_PREC_FP8)
    return fail(h, PFTCC_E_ARG, "unknown precision %d", prec);
  if (h->euclid || h->match_filt || h->pssnr)
    return fail(h, PFTCC_E_UNSUPPORTED, "the MFMA form implements objfun cc without match filter");
  if (zifft_lds(2, h->n) > kMaxLds) return fail(h, PFTCC_E_UNSUPPORTED, "pftsz %d too large for the MFMA form", h->n);
  const size_t nse = std::max<size_t>(1, shv.size() / 2);
  if ((size_t)h->nrefs * nse * h->nrots >= 0xFFFFFFFFull)
    return fail(h, PFTCC_E_ARG, "too many references x shifts x rotations for the search keys");
  if ((rc = need_memo(h))) return rc;
  std::vector<int> rl(nr);
  for (int i = 0; i < nr; ++i) {
    const int r = irefs ? irefs[i] : i + 1;
    if ((rc = check_ref(h, r))) return rc;
    rl[i] = r - 1;
  }
  std::vector<int> pl[2], pos[2];
  for (int i = 0; i < np; ++i) {
    int ip;
    if ((rc = local_index(h, iptcls[i], &ip))) return rc;
    const int eo = h->iseven[ip] ? 0 : 1;
    pl[eo].push_back(ip);
    pos[eo].push_back(i);
    if (plocal) plocal->push_back(ip);
  }
  // FP16 and FP8: per-shell mean squares of the particles' spectra (for the shell scales)
  std::vector<double> msW;
  if (op_format(prec) != OP_BF16) {
    std::vector<int> all;
    for (int eo = 0; eo < 2; ++eo) all.insert(all.end(), pl[eo].begin(), pl[eo].end());
    int* d_all = nullptr;
    if ((rc = dalloc(h, &d_all, all.size()))) return rc;
    HIPCHK(h, hipMemcpyAsync(d_all, all.data(), all.size() * 4, hipMemcpyHostToDevice, h->stream));
    rc = shell_ms(h, h->d_memo, d_all, (int)all.size(), msW);
    dfree(d_all);
    if (rc) return rc;
  }
  for (int eo = 0; eo < 2; ++eo)
    if (!pl[eo].empty() && (rc = mfma_group(h, eo, pl[eo], pos[eo], rl, shv, prec, d_best, d_cc, d_prbest, &msW)))
      return rc;
  return PFTCC_OK;
}

std::vector<float> shift_vector(int32_t ns, const float* shifts) {
  return ns > 0 && shifts ? std::vector<float>(shifts, shifts + 2 * (size_t)ns) : std::vector<float>();
}

// One (iref, iptcl, shvec, irot) evaluation through k_rot8; returns f and grad[3].
int rot8_single(pftcc* h, int iref, int iptcl, const double shvec[2], int irot, double* f, double* grad,
                bool angle, int objfun = -1) {
  int rc, ip;
  if ((rc = check_ref(h, iref)) || (rc = local_index(h, iptcl, &ip)) || (rc = check_rot(h, irot))) return rc;
  if ((rc = need_direct(h))) return rc;
  if (rot8_mode(h, objfun) == 2 && !h->have_sigma2)
    return fail(h, PFTCC_E_STATE, "the Euclid objective needs assign_sigma2_noise");
  if (angle && (rc = ensure_dptcls(h))) return rc;
  struct {
    int iref, ip, rho, pad;
    double sh[2];
  } in{iref - 1, ip, irot - 1, 0, {shvec ? shvec[0] : 0.0, shvec ? shvec[1] : 0.0}};
  if ((rc = scratch_reserve(h, 4096))) return rc;
  char* d_in = take<char>(h, sizeof in);
  double* d_out = take<double>(h, 4);
  HIPCHK(h, hipMemcpyAsync(d_in, &in, sizeof in, hipMemcpyHostToDevice, h->stream));
  const int* di = reinterpret_cast<const int*>(d_in);
  if ((rc = launch_rot8(h, 1, di, di + 1, reinterpret_cast<const double*>(d_in + 16), di + 2, d_out, d_out + 1,
                        angle, objfun)))
    return rc;
  double out[4];
  HIPCHK(h, hipMemcpyAsync(out, d_out, sizeof out, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (f) *f = out[0];
  if (grad) {
    grad[0] = out[1];
    grad[1] = out[2];
    grad[2] = out[3];
  }
  return PFTCC_OK;
}

}  // namespace

// =============================================================================== C ABI ===
extern "C" {

int32_t pftcc_magic_pftsz(int32_t ring2) { return magic_pftsz_host(ring2); }

const char* pftcc_last_error(const pftcc_t* h) { return h ? h->err.c_str() : "null handle"; }

int pftcc_new(const pftcc_cfg* cfg, int32_t nrefs, int32_t pfrom, int32_t pto, const int8_t* ptcl_mask,
              const int32_t* eoarr, pftcc_t** out) {
  if (!cfg || !out) return PFTCC_E_ARG;
  *out = nullptr;
  pftcc* h = new pftcc();
  h->cfg = *cfg;
  auto bad = [&](int code) {
    *out = h;    // hand the handle back so the caller can read the message, then kill it
    return code;
  };
  if (pto - pfrom + 1 < 1) return bad(fail(h, PFTCC_E_ARG, "nptcls (# of particles) must be > 0; new"));
  if (nrefs < 1) return bad(fail(h, PFTCC_E_ARG, "nrefs (# of reference sections) must be > 0; new"));
  if (cfg->boxmatch <= 0 || cfg->boxmatch % 2 != 0)
    return bad(fail(h, PFTCC_E_ARG, "only even logical dims supported; new (boxmatch %d)", cfg->boxmatch));
  h->ldim = cfg->boxmatch;
  h->n = cfg->pftsz > 0 ? cfg->pftsz : magic_pftsz_host(cfg->ring2);
  if (h->n <= 0 || h->n % 2 != 0) return bad(fail(h, PFTCC_E_ARG, "pftsz %d must be even and > 0", h->n));
  h->plan = make_plan(h->n);
  if (h->plan.np < 0)
    return bad(fail(h, PFTCC_E_UNSUPPORTED, "pftsz %d has a prime factor > 13", h->n));
  h->nrots = 2 * h->n;
  h->kfrom = cfg->kfromto[0];
  h->kto = cfg->kfromto[1];
  h->kstop = cfg->kstop > 0 ? cfg->kstop : h->kto;
  if (h->kfrom < 1 || h->kto < h->kfrom || h->kstop < h->kfrom || h->kstop > h->kto)
    return bad(fail(h, PFTCC_E_ARG, "bad shell range kfromto=[%d,%d] kstop=%d", h->kfrom, h->kto, h->kstop));
  h->nk = h->kto - h->kfrom + 1;
  h->kcorr = h->kstop - h->kfrom + 1;
  h->nrefs = nrefs;
  h->pfrom = pfrom;
  h->pto = pto;
  h->with_ctf = cfg->ctf != 0;
  h->match_filt = cfg->match_filt != 0;
  h->pssnr = cfg->pssnr != 0;
  h->clsfrcs = cfg->clsfrcs != 0;
  h->euclid = cfg->objfun == PFTCC_OBJFUN_EUCLID;
  {
    const char* gen = std::getenv("PFTCC_GENERIC");
    h->use_fused = !(gen && gen[0] == '1');
  }
  // index translation table (:276-290)
  h->pinds.assign(pto - pfrom + 1, 0);
  int cnt = 0;
  for (int i = 0; i < pto - pfrom + 1; ++i) {
    if (!ptcl_mask || ptcl_mask[i]) h->pinds[i] = ++cnt;
  }
  h->nptcls = cnt;
  if (cnt < 1) return bad(fail(h, PFTCC_E_ARG, "ptcl_mask selects no particle; new"));
  if (hipGetDevice(&h->device) != hipSuccess) return bad(fail(h, PFTCC_E_DEVICE, "no HIP device"));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return bad(fail(h, PFTCC_E_DEVICE, "no HIP device"));
  build_host_grid(h);
  int rc;
  if ((rc = resolve_noalloc_scratch(h)) || (rc = upload_pinds(h))) return bad(rc);
  const size_t ring = (size_t)h->nk * h->n;
  if ((rc = dalloc(h, &h->d_refs_e, (size_t)nrefs * ring)) || (rc = dalloc(h, &h->d_refs_o, (size_t)nrefs * ring)))
    return bad(rc);
  if (sync_memset(h, h->d_refs_e, 0, (size_t)nrefs * ring * sizeof(float2)) != hipSuccess ||
      sync_memset(h, h->d_refs_o, 0, (size_t)nrefs * ring * sizeof(float2)) != hipSuccess)
    return bad(fail(h, PFTCC_E_DEVICE, "hipMemset failed"));
  if ((rc = alloc_ptcl_arrays(h))) return bad(rc);
  // eo assignment (:291-307)
  if (eoarr) {
    bool allm1 = true;
    for (int i = 0; i < pto - pfrom + 1; ++i) allm1 = allm1 && eoarr[i] == -1;
    if (!allm1)
      for (int i = 0; i < pto - pfrom + 1; ++i)
        if (h->pinds[i] > 0) h->iseven[h->pinds[i] - 1] = eoarr[i] == 0 ? 1 : 0;
    if (sync_copy(h, h->d_iseven, h->iseven.data(), h->nptcls, hipMemcpyHostToDevice) != hipSuccess)
      return bad(fail(h, PFTCC_E_DEVICE, "copy eo"));
  }
  // constant tables
  std::vector<float2> tw(2 * h->n);
  for (int a = 0; a < 2 * h->n; ++a) {
    const double ang = -kDPi * a / h->n;
    tw[a] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  if ((rc = dalloc(h, &h->d_tw2, tw.size())) || (rc = dalloc(h, &h->d_argx, h->argx.size())) ||
      (rc = dalloc(h, &h->d_argy, h->argy.size())) || (rc = dalloc(h, &h->d_onex, h->onex.size())) ||
      (rc = dalloc(h, &h->d_oney, h->oney.size())) || (rc = dalloc(h, &h->d_spafreq, h->spafreq.size())) ||
      (rc = dalloc(h, &h->d_angmat, h->angmat.size())) || (rc = dalloc(h, &h->d_sn32, h->sn32.size())) ||
      (rc = dalloc(h, &h->d_ncn32, h->ncn32.size())))
    return bad(rc);
  if (sync_copy(h, h->d_tw2, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
      sync_copy(h, h->d_argx, h->argx.data(), h->argx.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      sync_copy(h, h->d_argy, h->argy.data(), h->argy.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      sync_copy(h, h->d_onex, h->onex.data(), h->onex.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      sync_copy(h, h->d_oney, h->oney.data(), h->oney.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      sync_copy(h, h->d_spafreq, h->spafreq.data(), h->spafreq.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      sync_copy(h, h->d_angmat, h->angmat.data(), h->angmat.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      sync_copy(h, h->d_sn32, h->sn32.data(), h->sn32.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      sync_copy(h, h->d_ncn32, h->ncn32.data(), h->ncn32.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
    return bad(fail(h, PFTCC_E_DEVICE, "upload of constant tables failed"));
  if (h->match_filt || h->pssnr) h->optlp.assign((size_t)nrefs * h->kcorr, 1.0f);
  *out = h;
  return PFTCC_OK;
}

int pftcc_set_timing(pftcc_t* h, int32_t enable) {
  if (!h) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc;
  if ((rc = collect_timing(h))) return rc;
  h->timing = enable != 0;
  for (int c = 0; c < 5; ++c) {
    h->timed_ms[c] = 0.0;
    h->timed_launches[c] = 0;
  }
  return PFTCC_OK;
}

int pftcc_get_timing(pftcc_t* h, double* total_ms, int64_t* launches) {
  if (!h || !total_ms || !launches) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc;
  if ((rc = collect_timing(h))) return rc;
  *total_ms = h->timed_ms[0];
  *launches = h->timed_launches[0];
  return PFTCC_OK;
}

int pftcc_get_timing_channel(pftcc_t* h, int32_t ch, double* total_ms, int64_t* launches) {
  if (!h || !total_ms || !launches || ch < 0 || ch > 4) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc;
  if ((rc = collect_timing(h))) return rc;
  *total_ms = h->timed_ms[ch];
  *launches = h->timed_launches[ch];
  return PFTCC_OK;
}

int pftcc_get_corr_tile(const pftcc_t* h, int32_t ns, int32_t* tile, int32_t* staged) {
  if (!h || !tile || !staged || ns < 0) return PFTCC_E_ARG;
  // the same selection as launch_corr for an exhaustive list of references (rlist_ld = 0)
  const bool shifted = ns > 0;
  const bool norm = (h->match_filt || h->pssnr) && !(h->euclid && shifted);
  bool rst = false;
  const int Ff = h->use_fused && fused_offsets_fit(h, ns) ? fused_tile(h->n, !h->euclid && !norm, std::max(ns, 1), &rst) : 0;
  *tile = Ff;
  *staged = rst ? 1 : 0;
  return PFTCC_OK;
}

int pftcc_kill(pftcc_t* h) {
  if (!h) return PFTCC_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto& ev : h->events) {
    (void)hipEventDestroy(ev.first);
    (void)hipEventDestroy(ev.second);
  }
  free_ptcl_arrays(h);
  dfree(h->d_refs_e);
  dfree(h->d_refs_o);
  dfree(h->d_tw2);
  dfree(h->d_argx);
  dfree(h->d_argy);
  dfree(h->d_sn32);
  dfree(h->d_ncn32);
  dfree(h->d_onex);
  dfree(h->d_oney);
  dfree(h->d_spafreq);
  dfree(h->d_angmat);
  dfree(h->d_refinv_f);
  dfree(h->d_ofac_f);
  dfree(h->d_refinv_d);
  dfree(h->d_ofac_d);
  dfree(h->d_refpw);
  dfree(h->d_optlp);
  dfree(h->d_pz_w);
  dfree(h->d_pz_hk);
  dfree(h->d_drefs_e);
  dfree(h->d_drefs_o);
  for (auto& ap : h->apack) free_apack(ap);
  dfree(h->d_pinds);
  if (h->d_scratch) (void)hipFree(h->d_scratch);
  delete h;
  return PFTCC_OK;
}

int pftcc_set_stream(pftcc_t* h, void* s) {
  if (!h) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  h->stream = (hipStream_t)s;
  return PFTCC_OK;
}

int pftcc_sync(pftcc_t* h) {
  if (!h) return PFTCC_E_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFTCC_OK;
}

// ------------------------------------------------------------------------------ setters ---
static int set_refs(pftcc_t* h, int iref_first, int count, const void* src, int iseven, hipMemcpyKind kind) {
  std::lock_guard<std::mutex> g(h->mtx);
  if (!src || count < 1) return fail(h, PFTCC_E_ARG, "no reference data");
  if (iref_first < 1 || iref_first + count - 1 > h->nrefs)
    return fail(h, PFTCC_E_ARG, "refs [%d, %d] outside [1, %d]", iref_first, iref_first + count - 1, h->nrefs);
  const size_t ring = (size_t)h->nk * h->n;
  float2* dst = (iseven ? h->d_refs_e : h->d_refs_o) + (size_t)(iref_first - 1) * ring;
  HIPCHK(h, hipMemcpyAsync(dst, src, (size_t)count * ring * sizeof(float2), kind, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->refnorm_valid = false;
  h->apack[0].valid = h->apack[1].valid = false;
  return PFTCC_OK;
}

int pftcc_set_ref_pft(pftcc_t* h, int32_t iref, const float* pft, int32_t iseven) {
  if (!h) return PFTCC_E_ARG;
  return set_refs(h, iref, 1, pft, iseven, hipMemcpyHostToDevice);
}
int pftcc_set_ref_pfts(pftcc_t* h, int32_t first, int32_t count, const float* pfts, int32_t iseven) {
  if (!h) return PFTCC_E_ARG;
  return set_refs(h, first, count, pfts, iseven, hipMemcpyHostToDevice);
}
int pftcc_set_ref_pfts_dev(pftcc_t* h, int32_t first, int32_t count, const void* d, int32_t iseven) {
  if (!h) return PFTCC_E_ARG;
  return set_refs(h, first, count, d, iseven, hipMemcpyDeviceToDevice);
}

static int set_ptcls(pftcc_t* h, int first, int count, const void* src, hipMemcpyKind kind) {
  std::lock_guard<std::mutex> g(h->mtx);
  if (!src || count < 1) return fail(h, PFTCC_E_ARG, "no particle data");
  const size_t ring = (size_t)h->nk * h->n;
  int ip0, rc;
  if ((rc = local_index(h, first, &ip0))) return rc;
  // consecutive global ids usually map to consecutive local slots; copy runs of them at once
  int i = 0;
  while (i < count) {
    int ip;
    if ((rc = local_index(h, first + i, &ip))) return rc;
    int len = 1;
    while (i + len < count) {
      if (first + i + len > h->pto || h->pinds[first + i + len - h->pfrom] != ip + len + 1) break;
      ++len;
    }
    const char* s = static_cast<const char*>(src) + (size_t)i * ring * sizeof(float2);
    HIPCHK(h, hipMemcpyAsync(h->d_ptcls + (size_t)ip * ring, s, (size_t)len * ring * sizeof(float2), kind,
                             h->stream));
    if ((rc = update_sqsums(h, ip, len))) return rc;
    i += len;
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->memo_valid = false;
  h->dptcls_valid = h->sqp8_valid = false;
  return PFTCC_OK;
}

int pftcc_set_ptcl_pft(pftcc_t* h, int32_t iptcl, const float* pft) {
  if (!h) return PFTCC_E_ARG;
  return set_ptcls(h, iptcl, 1, pft, hipMemcpyHostToDevice);
}
int pftcc_set_ptcl_pfts(pftcc_t* h, int32_t first, int32_t count, const float* pfts) {
  if (!h) return PFTCC_E_ARG;
  return set_ptcls(h, first, count, pfts, hipMemcpyHostToDevice);
}
int pftcc_set_ptcl_pfts_dev(pftcc_t* h, int32_t first, int32_t count, const void* d) {
  if (!h) return PFTCC_E_ARG;
  return set_ptcls(h, first, count, d, hipMemcpyDeviceToDevice);
}

// init_polarizer (simple_polarizer.f90:41-92) for the handle's polar grid and box.
int pftcc_init_polarizer(pftcc_t* h, int32_t griddev) {
  if (!h) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  const int ldim = h->ldim, nk = h->nk, n = h->n, np = nk * n;
  if (h->kto > ldim / 2)
    return fail(h, PFTCC_E_ARG, "kfromto(2) = %d beyond the Nyquist index %d of box %d; init_polarizer", h->kto,
                ldim / 2, ldim);
  int rc;
  if (!h->d_pz_w) {
    if ((rc = dalloc(h, &h->d_pz_w, (size_t)PZ_WLEN * np))) return rc;
    if ((rc = dalloc(h, &h->d_pz_hk, (size_t)2 * PZ_WDIM * np))) return rc;
  }
  const size_t grid = (size_t)nk * h->nrots;
  if ((rc = scratch_reserve(h, 2 * aligned(grid * 4)))) return rc;
  float* d_px = take<float>(h, grid);
  float* d_py = take<float>(h, grid);
  HIPCHK(h, hipMemcpyAsync(d_px, h->polar_x.data(), grid * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d_py, h->polar_y.data(), grid * 4, hipMemcpyHostToDevice, h->stream));
  // kbinterpol(KBWINSZ = 1.5, alpha): W = 3, beta = 7.4 (half-width <= 1.5); alpha does not enter
  // the weights (simple_kbinterpol.f90:92-109, simple_defs.f90:131)
  KbWin kb;
  kb.whalf = 1.5f;
  kb.W = 2.0f * kb.whalf;
  kb.beta = 7.4f;
  kb.twooW = 2.0f / kb.W;
  kb.oneoW = 1.0f / kb.W;
  // loop_lims(3) of the even square image (simple_ftiter.f90:89-114)
  const int hlo = -ldim / 2, hhi = ldim / 2, klo = -ldim / 2, khi = ldim / 2 - 1;
  hipLaunchKernelGGL(k_polar_tables, dim3((np + 255) / 256), dim3(256), 0, h->stream, (const float*)d_px,
                     (const float*)d_py, nk, n, h->nrots, kb, hlo, hhi, klo, khi, griddev ? 1 : 0, h->d_pz_w,
                     h->d_pz_hk);
  HIPCHK(h, hipGetLastError());
  // every address the polarize kernel will read must lie in the half-plane transform
  std::vector<int> hk((size_t)2 * PZ_WDIM * np);
  HIPCHK(h, hipMemcpyAsync(hk.data(), h->d_pz_hk, hk.size() * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t c = 0; c < (size_t)2 * PZ_WDIM; ++c) {
    const int lo = c < PZ_WDIM ? hlo : klo, hi = c < PZ_WDIM ? hhi : khi;
    for (int p = 0; p < np; ++p) {
      const int v = hk[c * np + p];
      if (v < lo || v > hi) return fail(h, PFTCC_E_DEVICE, "polarizer address %d outside [%d, %d]", v, lo, hi);
    }
  }
  h->pz_ready = true;
  return PFTCC_OK;
}

}  // extern "C"

namespace {

// polarize (simple_polarizer.f90:120-151) of count transforms into particle or reference slots.
int polarize_impl(pftcc* h, int count, const int32_t* img_inds, const void* cmats, bool dev, int isptcl,
                  int iseven, const int8_t* mask, int nmask) {
  std::lock_guard<std::mutex> g(h->mtx);
  if (!h->pz_ready) return fail(h, PFTCC_E_STATE, "init_polarizer must come before polarize");
  if (count < 1 || !img_inds || !cmats) return fail(h, PFTCC_E_ARG, "polarize: no images");
  if (mask && nmask < h->kto)
    return fail(h, PFTCC_E_ARG, "polarize: mask has %d entries, shells up to %d need indexing", nmask, h->kto);
  std::vector<int> slot(count);
  std::vector<char> seen(isptcl ? (size_t)h->nptcls : (size_t)h->nrefs, 0);
  for (int b = 0; b < count; ++b) {
    int rc, s;
    if (isptcl) {
      if ((rc = local_index(h, img_inds[b], &s))) return rc;
    } else {
      if ((rc = check_ref(h, img_inds[b]))) return rc;
      s = img_inds[b] - 1;
    }
    if (seen[s]) return fail(h, PFTCC_E_ARG, "polarize: image index %d given twice", img_inds[b]);
    seen[s] = 1;
    slot[b] = s;
  }
  const int ldim = h->ldim, nk = h->nk, n = h->n, np = nk * n;
  const size_t img_bytes = (size_t)ldim * (ldim / 2 + 1) * sizeof(float2);
  const size_t budget = (size_t)256 << 20;    // host transforms are staged through this much HBM
  const int chunk = dev ? count : (int)std::max<size_t>(1, std::min<size_t>(count, budget / img_bytes));
  int rc;
  if ((rc = scratch_reserve(h, aligned((size_t)count * 4) + aligned(nk) + (dev ? 0 : aligned(chunk * img_bytes)))))
    return rc;
  int* d_slot = take<int>(h, count);
  int8_t* d_mask = take<int8_t>(h, nk);
  float2* d_img = dev ? nullptr : take<float2>(h, chunk * img_bytes / sizeof(float2));
  HIPCHK(h, hipMemcpyAsync(d_slot, slot.data(), (size_t)count * 4, hipMemcpyHostToDevice, h->stream));
  if (mask) {
    std::vector<int8_t> mk(nk);
    for (int kk = 0; kk < nk; ++kk) mk[kk] = mask[h->kfrom + kk - 1] ? 1 : 0;   // mask(k), 1-based k
    HIPCHK(h, hipMemcpyAsync(d_mask, mk.data(), nk, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  float2* out = isptcl ? h->d_ptcls : (iseven ? h->d_refs_e : h->d_refs_o);
  for (int b0 = 0; b0 < count; b0 += chunk) {
    const int nb = std::min(chunk, count - b0);
    const float2* src;
    if (dev) {
      src = static_cast<const float2*>(cmats) + (size_t)b0 * (img_bytes / sizeof(float2));
    } else {
      HIPCHK(h, hipMemcpyAsync(d_img, static_cast<const char*>(cmats) + (size_t)b0 * img_bytes, nb * img_bytes,
                               hipMemcpyHostToDevice, h->stream));
      src = d_img;
    }
    hipLaunchKernelGGL(k_polarize, dim3((unsigned)((np + 255) / 256 * ((nb + PZ_IMGS - 1) / PZ_IMGS))), dim3(256), 0,
                       h->stream, src,
                       ldim, nb, (const float*)h->d_pz_w, (const int*)h->d_pz_hk, nk, n, (const int*)d_slot + b0,
                       mask ? (const int8_t*)d_mask : (const int8_t*)nullptr, out);
    HIPCHK(h, hipGetLastError());
  }
  if (isptcl) {
    // set_ptcl_pft refreshes sqsums_ptcls (:438-446)
    hipLaunchKernelGGL(k_sqsums, dim3(count), dim3(256), 0, h->stream, (const float2*)h->d_ptcls, nk, n, h->kcorr,
                       h->d_sqsums, h->d_sqsums_k, 0, count, (const int*)d_slot);
    HIPCHK(h, hipGetLastError());
    h->memo_valid = false;
    h->dptcls_valid = h->sqp8_valid = false;
  } else {
    h->refnorm_valid = false;
    h->apack[0].valid = h->apack[1].valid = false;
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFTCC_OK;
}

}  // namespace

extern "C" {

int pftcc_polarize(pftcc_t* h, int32_t count, const int32_t* img_inds, const float* cmats, int32_t isptcl,
                   int32_t iseven, const int8_t* mask, int32_t nmask) {
  if (!h) return PFTCC_E_ARG;
  return polarize_impl(h, count, img_inds, cmats, false, isptcl, iseven, mask, nmask);
}

int pftcc_polarize_dev(pftcc_t* h, int32_t count, const int32_t* img_inds, const void* d_cmats, int32_t isptcl,
                       int32_t iseven, const int8_t* mask, int32_t nmask) {
  if (!h) return PFTCC_E_ARG;
  return polarize_impl(h, count, img_inds, d_cmats, true, isptcl, iseven, mask, nmask);
}

}  // extern "C"

namespace {

// constructor / init of the ctf object (simple_ctf.f90:54-81, 522-525) in real(sp) arithmetic;
// gate_phaseplate: create_polar_absctfmats adds the phase shift only with a phase plate, while
// apply_serial passes ctfparms%phshift as given.
KCtfConst make_kctf(const pftcc_ctfparams& p, bool gate_phaseplate) {
  KCtfConst c{};
  const float a = 1000.0f * p.kv;
  const float wl = 12.26f / std::sqrt(a + (0.9784f * (a * a)) / 1.0e6f);
  c.wl = wl / p.smpd;
  c.cs = (p.cs * 1.0e7f) / p.smpd;
  c.ampc = std::atan(p.fraca / std::sqrt(1.0f - p.fraca * p.fraca));
  c.dfx = (p.dfx * 1.0e4f) / p.smpd;
  c.dfy = (p.dfy * 1.0e4f) / p.smpd;
  c.angast = (p.angast / 180.0f) * kPi32;
  c.phshift = (!gate_phaseplate || p.l_phaseplate) ? p.phshift : 0.0f;
  c.active = 1;
  return c;
}

// tw2 table of a length-L transform: exp(-i pi a / L), a in [0, 2L), FP64 rounded to FP32.
// Box sizes whose preparation transforms run compile-time passes (others: runtime radices).
#define PFTCC_PREP_SIZES(X) X(64) X(100) X(128) X(160) X(192) X(200) X(240) X(256) X(320) X(384) X(400)

// Calls fn with std::integral_constant<int, L> for a listed size, else with 0.
template <typename Fn>
void prep_dispatch(int L, Fn&& fn) {
#define X(N)                                   \
  if (L == N) {                                \
    fn(std::integral_constant<int, N>{});      \
    return;                                    \
  }
  PFTCC_PREP_SIZES(X)
#undef X
  fn(std::integral_constant<int, 0>{});
}

std::vector<float2> twiddles(int L) {
  std::vector<float2> t(2 * L);
  for (int a = 0; a < 2 * L; ++a) {
    const double ang = -kDPi * a / L;
    t[a] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  return t;
}

// Fortran nint of an FP32 value
int nint32(float x) { return (int)(x >= 0.f ? std::floor(x + 0.5f) : -std::floor(-x + 0.5f)); }

// gen_instrfun_img (simple_gridding.f90:18-55) with the polarizer's window (kbinterpol%instr,
// simple_kbinterpol.f90:92-109, 211-225): w(i) for i = 1..ldim, sinh(a)/a evaluated in FP64.
std::vector<float> instrfun_1d(int ldim) {
  const double whalf = 1.5, W = 2.0 * whalf, beta = 7.4, piW = kDPi * W;
  const double thresh = beta / piW - 1e-20;
  std::vector<float> w(ldim);
  const int center = ldim / 2 + 1;
  for (int i = 1; i <= ldim; ++i) {
    const float x = (float)(i - center) / (float)ldim;
    double r = 1.0;
    if (std::fabs((double)x) < thresh) {
      const double a2 = std::sqrt(beta * beta - (piW * x) * (piW * x));
      r = a2 < 1e-10 ? 1.0 : std::sinh(a2) / a2;
    }
    w[i - 1] = (float)r;
  }
  return w;
}

int prep_impl(pftcc* h, int count, const void* imgs, bool dev, const pftcc_prep_cfg* cfg, const float* shifts,
              const pftcc_ctfparams* ctfp, void* cmats) {
  std::lock_guard<std::mutex> g(h->mtx);
  if (count < 1 || !imgs || !cmats || !cfg) return fail(h, PFTCC_E_ARG, "prep_imgs: no images");
  const int B = cfg->box, M = cfg->boxmatch;
  if (B < 8 || M < 8 || (B & 1) || (M & 1) || M > B)
    return fail(h, PFTCC_E_ARG, "prep_imgs: box %d / boxmatch %d must be even, >= 8, boxmatch <= box", B, M);
  if (M != h->ldim) return fail(h, PFTCC_E_ARG, "prep_imgs: boxmatch %d differs from the handle's %d", M, h->ldim);
  const FftPlan pB = make_plan(B), pM = make_plan(M);
  if (pB.np < 0 || pM.np < 0)
    return fail(h, PFTCC_E_UNSUPPORTED, "prep_imgs: box %d or boxmatch %d has a prime factor above 13", B, M);
  const size_t lds_rows = sizeof(float2) * ((size_t)2 * PREP_RP * B + 2 * B);
  const size_t lds_cols = sizeof(float2) * ((size_t)2 * PREP_CP * B + 2 * B);
  if (lds_rows > kMaxLds || lds_cols > kMaxLds) return fail(h, PFTCC_E_UNSUPPORTED, "prep_imgs: box %d too large", B);
  if (!(cfg->msk > 0.f)) return fail(h, PFTCC_E_ARG, "prep_imgs: mask radius must be positive");
  if (cfg->ctfflag < 0 || cfg->ctfflag > 2) return fail(h, PFTCC_E_ARG, "prep_imgs: unsupported CTF flag %d", cfg->ctfflag);
  const bool flip = cfg->ctfflag == 1;
  if (flip && !ctfp) return fail(h, PFTCC_E_ARG, "prep_imgs: CTF parameters required for the phase flip");
  if (!shifts) return fail(h, PFTCC_E_ARG, "prep_imgs: shifts required");
  // soft-mask radius and window (mask, simple_image.f90:6208-6240; prepimg4align :354-360)
  const float soft = cfg->inner > 0.f ? cfg->msk : (cfg->focusmsk > 0.f ? cfg->focusmsk : cfg->msk);
  const int minlen = std::min(nint32(2.0f * (soft + 6.0f)), M);
  const float shconst = kPi32 / ((float)B / 2.0f);
  const float ratio = (float)(M * M) / (float)(B * B);
  const size_t nhB = B / 2 + 1, nhM = M / 2 + 1;
  const size_t in_img = (size_t)B * B, out_img = (size_t)M * nhM;
  // chunk so that the work buffers stay within ~512 MB
  const size_t per = in_img * 4 * (dev ? 0 : 1) + (size_t)B * nhB * 8 + (size_t)M * M * 4 + out_img * 8 * (dev ? 0 : 1) +
                     8 + 8 + sizeof(KCtfConst);
  const int C = (int)std::max<size_t>(1, std::min<size_t>(count, ((size_t)512 << 20) / per));
  std::vector<float2> twB = twiddles(B), twM = twiddles(M);
  std::vector<float> wI;
  if (cfg->griddev) wI = instrfun_1d(M);
  int rc;
  if ((rc = scratch_reserve(h, 2 * aligned(twB.size() * 8) + aligned(twM.size() * 8) + aligned((size_t)M * 4) +
                                   aligned((size_t)C * 8) * 2 + aligned((size_t)C * sizeof(KCtfConst)) +
                                   aligned((size_t)C * B * nhB * 8) + aligned((size_t)C * M * M * 4) +
                                   (dev ? 0 : aligned((size_t)C * in_img * 4) + aligned((size_t)C * out_img * 8)))))
    return rc;
  float2* d_twB = take<float2>(h, twB.size());
  float2* d_twM = take<float2>(h, twM.size());
  float* d_wI = take<float>(h, M);
  float2* d_stats = take<float2>(h, C);
  float* d_sh = take<float>(h, 2 * (size_t)C);
  KCtfConst* d_ctf = take<KCtfConst>(h, C);
  float2* d_work = take<float2>(h, (size_t)C * B * nhB);
  float* d_img2 = take<float>(h, (size_t)C * M * M);
  float* d_in = dev ? nullptr : take<float>(h, (size_t)C * in_img);
  float2* d_out = dev ? nullptr : take<float2>(h, (size_t)C * out_img);
  HIPCHK(h, hipMemcpyAsync(d_twB, twB.data(), twB.size() * 8, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d_twM, twM.data(), twM.size() * 8, hipMemcpyHostToDevice, h->stream));
  if (cfg->griddev) HIPCHK(h, hipMemcpyAsync(d_wI, wI.data(), (size_t)M * 4, hipMemcpyHostToDevice, h->stream));
  std::vector<KCtfConst> kc;
  for (int b0 = 0; b0 < count; b0 += C) {
    const int nb = std::min(C, count - b0);
    const float* src;
    float2* dst;
    if (dev) {
      src = static_cast<const float*>(imgs) + (size_t)b0 * in_img;
      dst = static_cast<float2*>(cmats) + (size_t)b0 * out_img;
    } else {
      HIPCHK(h, hipMemcpyAsync(d_in, static_cast<const float*>(imgs) + (size_t)b0 * in_img, (size_t)nb * in_img * 4,
                               hipMemcpyHostToDevice, h->stream));
      src = d_in;
      dst = d_out;
    }
    HIPCHK(h, hipMemcpyAsync(d_sh, shifts + 2 * (size_t)b0, (size_t)nb * 8, hipMemcpyHostToDevice, h->stream));
    if (flip) {
      kc.resize(nb);
      for (int i = 0; i < nb; ++i) kc[i] = make_kctf(ctfp[b0 + i], false);
      HIPCHK(h, hipMemcpyAsync(d_ctf, kc.data(), (size_t)nb * sizeof(KCtfConst), hipMemcpyHostToDevice, h->stream));
    }
    hipLaunchKernelGGL(k_prep_stats, dim3(nb), dim3(256), 0, h->stream, src, B, cfg->msk, d_stats);
    prep_dispatch(B, [&](auto lc) {
      hipLaunchKernelGGL(k_prep_rows_fwd<decltype(lc)::value>, dim3((B / 2 + PREP_RP - 1) / PREP_RP, nb), dim3(256),
                         lds_rows, h->stream, src, B, (const float2*)d_stats, (const float2*)d_twB, pB, d_work);
    });
    PrepColArgs ca{};
    ca.data = d_work;
    ca.L = B;
    ca.inverse = 0;
    ca.post = 1;
    ca.tw2g = d_twB;
    ca.plan = pB;
    ca.shifts = d_sh;
    ca.shconst = shconst;
    ca.ctf = flip ? d_ctf : nullptr;
    ca.ratio = 1.0f;
    const dim3 gcB((unsigned)((nhB + PREP_CP - 1) / PREP_CP), nb);
    prep_dispatch(B, [&](auto lc) {
      hipLaunchKernelGGL(k_prep_cols<decltype(lc)::value>, gcB, dim3(256), lds_cols, h->stream, ca);
      ca.inverse = 1;
      ca.post = 0;
      hipLaunchKernelGGL(k_prep_cols<decltype(lc)::value>, gcB, dim3(256), lds_cols, h->stream, ca);
      hipLaunchKernelGGL(k_prep_rows_inv<decltype(lc)::value>, dim3((M / 2 + PREP_RP - 1) / PREP_RP, nb), dim3(256),
                         lds_rows, h->stream, (const float2*)d_work, B, M, (const float2*)d_twB, pB, d_img2);
    });
    hipLaunchKernelGGL(k_prep_mask, dim3(nb), dim3(256), sizeof(float) * 4 * M, h->stream, d_img2, M, soft,
                       cfg->inner, cfg->width, minlen, cfg->griddev ? (const float*)d_wI : nullptr);
    const size_t lds_rows_m = sizeof(float2) * ((size_t)2 * PREP_RP * M + 2 * M);
    const size_t lds_cols_m = sizeof(float2) * ((size_t)2 * PREP_CP * M + 2 * M);
    prep_dispatch(M, [&](auto lc) {
      hipLaunchKernelGGL(k_prep_rows_fwd<decltype(lc)::value>, dim3((M / 2 + PREP_RP - 1) / PREP_RP, nb), dim3(256),
                         lds_rows_m, h->stream, (const float*)d_img2, M, (const float2*)nullptr, (const float2*)d_twM,
                         pM, dst);
    });
    PrepColArgs cm{};
    cm.data = dst;
    cm.L = M;
    cm.inverse = 0;
    cm.post = 2;
    cm.tw2g = d_twM;
    cm.plan = pM;
    cm.ratio = ratio;
    prep_dispatch(M, [&](auto lc) {
      hipLaunchKernelGGL(k_prep_cols<decltype(lc)::value>, dim3((unsigned)((nhM + PREP_CP - 1) / PREP_CP), nb),
                         dim3(256), lds_cols_m, h->stream, cm);
    });
    HIPCHK(h, hipGetLastError());
    if (!dev)
      HIPCHK(h, hipMemcpyAsync(static_cast<float2*>(cmats) + (size_t)b0 * out_img, d_out, (size_t)nb * out_img * 8,
                               hipMemcpyDeviceToHost, h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFTCC_OK;
}

}  // namespace

extern "C" {

int pftcc_prep_imgs(pftcc_t* h, int32_t count, const float* imgs, const pftcc_prep_cfg* cfg, const float* shifts,
                    const pftcc_ctfparams* ctfp, float* cmats) {
  if (!h) return PFTCC_E_ARG;
  return prep_impl(h, count, imgs, false, cfg, shifts, ctfp, cmats);
}

int pftcc_prep_imgs_dev(pftcc_t* h, int32_t count, const void* d_imgs, const pftcc_prep_cfg* cfg,
                        const float* shifts, const pftcc_ctfparams* ctfp, void* d_cmats) {
  if (!h) return PFTCC_E_ARG;
  return prep_impl(h, count, d_imgs, true, cfg, shifts, ctfp, d_cmats);
}

int pftcc_set_ref_fcomp(pftcc_t* h, int32_t iref, int32_t irot, int32_t k, float re, float im, int32_t iseven) {
  if (!h) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc;
  if ((rc = check_ref(h, iref))) return rc;
  if (irot < 1 || irot > h->n || k < h->kfrom || k > h->kto)
    return fail(h, PFTCC_E_ARG, "component (%d, %d) outside the ring block", irot, k);
  const float2 v = make_float2(re, im);
  float2* dst = (iseven ? h->d_refs_e : h->d_refs_o) + ((size_t)(iref - 1) * h->nk + (k - h->kfrom)) * h->n + (irot - 1);
  HIPCHK(h, sync_copy(h, dst, &v, sizeof v, hipMemcpyHostToDevice));
  h->refnorm_valid = false;
  h->apack[0].valid = h->apack[1].valid = false;
  return PFTCC_OK;
}

int pftcc_set_ptcl_fcomp(pftcc_t* h, int32_t iptcl, int32_t irot, int32_t k, float re, float im) {
  if (!h) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc, ip;
  if ((rc = local_index(h, iptcl, &ip))) return rc;
  if (irot < 1 || irot > h->n || k < h->kfrom || k > h->kto)
    return fail(h, PFTCC_E_ARG, "component (%d, %d) outside the ring block", irot, k);
  const float2 v = make_float2(re, im);
  HIPCHK(h, sync_copy(h, h->d_ptcls + ((size_t)ip * h->nk + (k - h->kfrom)) * h->n + (irot - 1), &v, sizeof v,
                      hipMemcpyHostToDevice));
  // set_ptcl_fcomp does not refresh sqsums_ptcls in the reference (:471-476); neither do we.
  h->memo_valid = false;
  h->dptcls_valid = h->sqp8_valid = false;
  return PFTCC_OK;
}

int pftcc_cp_even2odd_ref(pftcc_t* h, int32_t iref) {
  if (!h) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc;
  if ((rc = check_ref(h, iref))) return rc;
  const size_t ring = (size_t)h->nk * h->n;
  HIPCHK(h, sync_copy(h, h->d_refs_o + (size_t)(iref - 1) * ring, h->d_refs_e + (size_t)(iref - 1) * ring,
                      ring * sizeof(float2), hipMemcpyDeviceToDevice));
  h->refnorm_valid = false;
  h->apack[0].valid = h->apack[1].valid = false;
  return PFTCC_OK;
}

int pftcc_cp_even_ref2ptcl(pftcc_t* h, int32_t iref, int32_t iptcl) {
  if (!h) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc, ip;
  if ((rc = check_ref(h, iref)) || (rc = local_index(h, iptcl, &ip))) return rc;
  const size_t ring = (size_t)h->nk * h->n;
  HIPCHK(h, hipMemcpyAsync(h->d_ptcls + (size_t)ip * ring, h->d_refs_e + (size_t)(iref - 1) * ring,
                           ring * sizeof(float2), hipMemcpyDeviceToDevice, h->stream));
  if ((rc = update_sqsums(h, ip, 1))) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->memo_valid = false;
  h->dptcls_valid = h->sqp8_valid = false;
  return PFTCC_OK;
}

int pftcc_swap_ptclsevenodd(pftcc_t* h) {
  if (!h) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  for (auto& e : h->iseven) e = e ? 0 : 1;
  return upload_iseven(h);
}

int pftcc_reallocate_ptcls(pftcc_t* h, int32_t nptcls, const int32_t* pinds) {
  if (!h) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  if (nptcls < 1 || !pinds) return fail(h, PFTCC_E_ARG, "reallocate_ptcls needs nptcls >= 1 and pinds");
  int lo = pinds[0], hi = pinds[0];
  for (int i = 1; i < nptcls; ++i) {
    lo = std::min(lo, (int)pinds[i]);
    hi = std::max(hi, (int)pinds[i]);
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (nptcls != h->nptcls) {
    free_ptcl_arrays(h);
    h->nptcls = nptcls;
    int rc;
    if ((rc = alloc_ptcl_arrays(h))) return rc;
  } else {
    // (:543-546) particles, square sums and eo flags are reset in both branches
    const size_t ring = (size_t)h->nk * h->n;
    HIPCHK(h, sync_memset(h, h->d_ptcls, 0, (size_t)h->nptcls * ring * sizeof(float2)));
    HIPCHK(h, sync_memset(h, h->d_sqsums, 0, (size_t)h->nptcls * sizeof(float)));
    HIPCHK(h, sync_memset(h, h->d_sqsums_k, 0, (size_t)h->nptcls * h->nk * sizeof(float)));
    h->iseven.assign(h->nptcls, 1);
    int rc;
    if ((rc = upload_iseven(h))) return rc;
  }
  h->pfrom = lo;
  h->pto = hi;
  h->pinds.assign(hi - lo + 1, 0);
  for (int i = 0; i < nptcls; ++i) h->pinds[pinds[i] - lo] = i + 1;
  h->memo_valid = h->dptcls_valid = h->sqp8_valid = false;
  h->refnorm_valid = false;
  return upload_pinds(h);
}

int pftcc_set_eo(pftcc_t* h, int32_t iptcl, int32_t is_even) {
  if (!h) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc, ip;
  if ((rc = local_index(h, iptcl, &ip))) return rc;
  h->iseven[ip] = is_even ? 1 : 0;
  return upload_iseven(h);
}

int pftcc_set_eos(pftcc_t* h, const int32_t* eoarr) {
  if (!h || !eoarr) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  bool allm1 = true;
  for (int i = 0; i < h->nptcls; ++i) allm1 = allm1 && eoarr[i] == -1;
  for (int i = 0; i < h->nptcls; ++i) h->iseven[i] = allm1 ? 1 : (eoarr[i] == 0 ? 1 : 0);
  return upload_iseven(h);
}

int pftcc_set_ref_optlp(pftcc_t* h, int32_t iref, const float* optlp) {
  if (!h || !optlp) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc;
  if ((rc = check_ref(h, iref))) return rc;
  if (h->optlp.empty()) h->optlp.assign((size_t)h->nrefs * h->kcorr, 1.0f);
  std::memcpy(&h->optlp[(size_t)(iref - 1) * h->kcorr], optlp, sizeof(float) * h->kcorr);
  h->refnorm_valid = false;
  h->optlp_dev_valid = false;
  return PFTCC_OK;
}

namespace {
// prep_matchfilt for a list of (particle, reference, rotation); iref 0 or match_filt off: the
// particle is only memoized, as in the reference.
int prep_matchfilt_run(pftcc* h, int np, const int32_t* iptcls, const int32_t* irefs, const int32_t* irots) {
  int rc;
  std::vector<int> pl(np), rl(np), rh(np);
  std::vector<char> seen(h->nptcls, 0);
  bool any = false;
  for (int i = 0; i < np; ++i) {
    if ((rc = local_index(h, iptcls[i], &pl[i]))) return rc;
    if (seen[pl[i]]) return fail(h, PFTCC_E_ARG, "particle %d listed twice", iptcls[i]);
    seen[pl[i]] = 1;
    rl[i] = -1;
    rh[i] = 0;
    if (h->match_filt && irefs[i] != 0) {
      if ((rc = check_ref(h, irefs[i])) || (rc = check_rot(h, irots[i]))) return rc;
      rl[i] = irefs[i] - 1;
      rh[i] = irots[i] - 1;
      any = true;
    }
  }
  if (any && h->with_ctf && !h->ctf_ready)
    return fail(h, PFTCC_E_STATE, "CTF matrices missing: call create_polar_absctfmats");
  if (h->kcorr > MF_MAXK) return fail(h, PFTCC_E_UNSUPPORTED, "prep_matchfilt supports at most %d shells", MF_MAXK);
  if (!h->memo_valid) {
    // the rest of the particles need their spectra too (the reference's arrays always hold them)
    int F = 8;
    while (F > 1 && sizeof(float2) * ((size_t)2 * F * h->n + 2 * h->n) > kMaxLds) F /= 2;
    const size_t lds = sizeof(float2) * ((size_t)2 * F * h->n + 2 * h->n);
    hipLaunchKernelGGL(k_memoize, dim3((h->nptcls * h->nk + F - 1) / F), dim3(256), lds, h->stream,
                       (const float2*)h->d_ptcls, 0, h->nptcls, h->nk, h->n, h->plan, (const float2*)h->d_tw2,
                       h->d_memo, F, (const int*)nullptr);
    HIPCHK(h, hipGetLastError());
  }
  if (any && !h->optlp_dev_valid) {
    const size_t cnt = (size_t)h->nrefs * h->kcorr;
    std::vector<float> o = h->optlp.empty() ? std::vector<float>(cnt, 1.0f) : h->optlp;
    if (!h->d_optlp && (rc = dalloc(h, &h->d_optlp, cnt))) return rc;
    HIPCHK(h, hipMemcpyAsync(h->d_optlp, o.data(), cnt * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->optlp_dev_valid = true;
  }
  if ((rc = scratch_reserve(h, 3 * aligned((size_t)np * 4)))) return rc;
  int* d_pl = take<int>(h, np);
  int* d_rl = take<int>(h, np);
  int* d_rh = take<int>(h, np);
  HIPCHK(h, hipMemcpyAsync(d_pl, pl.data(), np * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d_rl, rl.data(), np * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d_rh, rh.data(), np * 4, hipMemcpyHostToDevice, h->stream));
  if (any) {
    MatchFiltArgs a{};
    a.n = h->n;
    a.nk = h->nk;
    a.kcorr = h->kcorr;
    a.kfrom = h->kfrom;
    a.nrefs = h->nrefs;
    a.pssnr = h->pssnr ? 1 : 0;
    a.ptcls = h->d_ptcls;
    a.refs_even = h->d_refs_e;
    a.refs_odd = h->d_refs_o;
    a.ctf = h->with_ctf ? h->d_ctf : nullptr;
    a.optlp = h->d_optlp;
    a.iseven = h->d_iseven;
    a.plist = d_pl;
    a.rlist = d_rl;
    a.rhos = d_rh;
    a.count = np;
    hipLaunchKernelGGL(k_prep_matchfilt, dim3(np), dim3(256), 0, h->stream, a);
    HIPCHK(h, hipGetLastError());
    hipLaunchKernelGGL(k_sqsums, dim3(np), dim3(256), 0, h->stream, (const float2*)h->d_ptcls, h->nk, h->n,
                       h->kcorr, h->d_sqsums, h->d_sqsums_k, 0, np, (const int*)d_pl);
    HIPCHK(h, hipGetLastError());
  }
  int F = 8;
  while (F > 1 && sizeof(float2) * ((size_t)2 * F * h->n + 2 * h->n) > kMaxLds) F /= 2;
  const size_t lds = sizeof(float2) * ((size_t)2 * F * h->n + 2 * h->n);
  hipLaunchKernelGGL(k_memoize, dim3((np * h->nk + F - 1) / F), dim3(256), lds, h->stream, (const float2*)h->d_ptcls, 0,
                     np, h->nk, h->n, h->plan, (const float2*)h->d_tw2, h->d_memo, F, (const int*)d_pl);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->memo_valid = true;
  h->dptcls_valid = h->sqp8_valid = false;
  return PFTCC_OK;
}
}  // namespace

int pftcc_prep_matchfilt(pftcc_t* h, int32_t iptcl, int32_t iref, int32_t irot) {
  if (!h) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  return prep_matchfilt_run(h, 1, &iptcl, &iref, &irot);
}

int pftcc_prep_matchfilt_batch(pftcc_t* h, int32_t np, const int32_t* iptcls, const int32_t* irefs,
                               const int32_t* irots) {
  if (!h || np < 0 || (np > 0 && (!iptcls || !irefs || !irots))) return PFTCC_E_ARG;
  if (np == 0) return PFTCC_OK;
  std::lock_guard<std::mutex> g(h->mtx);
  return prep_matchfilt_run(h, np, iptcls, irefs, irots);
}

int pftcc_assign_sigma2_noise(pftcc_t* h, const float* sigma2) {
  if (!h || !sigma2) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  const int nk = h->nk;
  std::vector<float> shellw((size_t)h->nptcls * nk, 0.0f);
  std::vector<double> s2w((size_t)h->nptcls * nk, 0.0);
  for (int gidx = h->pfrom; gidx <= h->pto; ++gidx) {
    const int l = h->pinds[gidx - h->pfrom];
    if (l <= 0) continue;
    for (int kk = 0; kk < nk; ++kk) {
      const float s = sigma2[(size_t)(gidx - h->pfrom) * nk + kk];
      // gencorrs_euclid: pxls_p_shell(k)/pftsz /(2 sigma2) (:1554-1555); for_rot_8: pxls/pftsz/sigma2
      shellw[(size_t)(l - 1) * nk + kk] = h->pxls[kk] / (float)h->n / (2.0f * s);
      s2w[(size_t)(l - 1) * nk + kk] = (double)h->pxls[kk] / (double)h->n / (double)s;
    }
  }
  int rc;
  if (!h->d_shellw && ((rc = dalloc(h, &h->d_shellw, shellw.size())) || (rc = dalloc(h, &h->d_sigma2w, s2w.size()))))
    return rc;
  HIPCHK(h, sync_copy(h, h->d_shellw, shellw.data(), shellw.size() * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(h, sync_copy(h, h->d_sigma2w, s2w.data(), s2w.size() * sizeof(double), hipMemcpyHostToDevice));
  h->have_sigma2 = true;
  return PFTCC_OK;
}

int pftcc_zero_refs_beyond_kstop(pftcc_t* h) {
  if (!h) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  if (h->kstop == h->kto) return PFTCC_OK;
  const size_t ring = (size_t)h->nk * h->n;
  const size_t off = (size_t)h->kcorr * h->n;
  const size_t len = (size_t)(h->nk - h->kcorr) * h->n;
  for (int r = 0; r < h->nrefs; ++r) {
    HIPCHK(h, hipMemsetAsync(h->d_refs_e + r * ring + off, 0, len * sizeof(float2), h->stream));
    HIPCHK(h, hipMemsetAsync(h->d_refs_o + r * ring + off, 0, len * sizeof(float2), h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->refnorm_valid = false;
  h->apack[0].valid = h->apack[1].valid = false;
  return PFTCC_OK;
}

// ---------------------------------------------------------------------------- memoizers ---
int pftcc_create_polar_absctfmats(pftcc_t* h, const pftcc_ctfparams* params) {
  if (!h || !params) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  std::vector<KCtfConst> cst(h->nptcls);
  for (auto& c : cst) c.active = 0;
  for (int gidx = h->pfrom; gidx <= h->pto; ++gidx) {
    const int l = h->pinds[gidx - h->pfrom];
    if (l <= 0) continue;
    const pftcc_ctfparams& p = params[gidx - h->pfrom];
    KCtfConst c{};
    // constructor (simple_ctf.f90:54-70) and kV2wl (:522-525), real(sp) arithmetic
    const float kv = p.kv;
    const float a = 1000.0f * kv;
    const float wl = 12.26f / std::sqrt(a + (0.9784f * (a * a)) / 1.0e6f);
    c.wl = wl / p.smpd;
    c.cs = (p.cs * 1.0e7f) / p.smpd;
    const float phaseq = std::sqrt(1.0f - p.fraca * p.fraca);
    c.ampc = std::atan(p.fraca / phaseq);
    // init (:73-81)
    c.dfx = (p.dfx * 1.0e4f) / p.smpd;
    c.dfy = (p.dfy * 1.0e4f) / p.smpd;
    c.angast = (p.angast / 180.0f) * kPi32;
    c.phshift = p.l_phaseplate ? p.phshift : 0.0f;
    c.active = 1;
    cst[l - 1] = c;
  }
  int rc;
  const size_t ring = (size_t)h->nk * h->n;
  if (!h->d_ctf && (rc = dalloc(h, &h->d_ctf, (size_t)h->nptcls * ring))) return rc;
  HIPCHK(h, hipMemsetAsync(h->d_ctf, 0, (size_t)h->nptcls * ring * sizeof(float), h->stream));
  if ((rc = scratch_reserve(h, aligned(cst.size() * sizeof(KCtfConst))))) return rc;
  KCtfConst* d_c = take<KCtfConst>(h, cst.size());
  HIPCHK(h, hipMemcpyAsync(d_c, cst.data(), cst.size() * sizeof(KCtfConst), hipMemcpyHostToDevice, h->stream));
  const size_t tot = (size_t)h->nptcls * ring;
  hipLaunchKernelGGL(k_ctfmats, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, h->stream, (const KCtfConst*)d_c,
                     (const float*)h->d_spafreq, (const float*)h->d_angmat, h->nk, h->n, h->nptcls, h->d_ctf);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->ctf_ready = true;
  return PFTCC_OK;
}

int pftcc_memoize_ffts(pftcc_t* h) {
  if (!h) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int F = 8;
  while (F > 1 && sizeof(float2) * ((size_t)2 * F * h->n + 2 * h->n) > kMaxLds) F /= 2;
  const int nitems = h->nptcls * h->nk;
  const size_t lds = sizeof(float2) * ((size_t)2 * F * h->n + 2 * h->n);
  hipLaunchKernelGGL(k_memoize, dim3((nitems + F - 1) / F), dim3(256), lds, h->stream, (const float2*)h->d_ptcls, 0,
                     h->nptcls, h->nk, h->n, h->plan, (const float2*)h->d_tw2, h->d_memo, F, (const int*)nullptr);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->memo_valid = true;
  h->dptcls_valid = h->sqp8_valid = false;
  return PFTCC_OK;
}

// ------------------------------------------------------------------------------ getters ---
int pftcc_get_nrots(const pftcc_t* h, int32_t* v) { if (!h || !v) return PFTCC_E_ARG; *v = h->nrots; return 0; }
int pftcc_get_pftsz(const pftcc_t* h, int32_t* v) { if (!h || !v) return PFTCC_E_ARG; *v = h->n; return 0; }
int pftcc_get_box(const pftcc_t* h, int32_t* v) { if (!h || !v) return PFTCC_E_ARG; *v = h->ldim; return 0; }
int pftcc_get_nrefs(const pftcc_t* h, int32_t* v) { if (!h || !v) return PFTCC_E_ARG; *v = h->nrefs; return 0; }
int pftcc_get_nptcls(const pftcc_t* h, int32_t* v) { if (!h || !v) return PFTCC_E_ARG; *v = h->nptcls; return 0; }
int pftcc_get_pdim(const pftcc_t* h, int32_t pdim[3]) {
  if (!h || !pdim) return PFTCC_E_ARG;
  pdim[0] = h->n;
  pdim[1] = h->kfrom;
  pdim[2] = h->kto;
  return 0;
}
int pftcc_get_rot(const pftcc_t* h, int32_t roind, float* rot) {
  if (!h || !rot) return PFTCC_E_ARG;
  if (roind < 1 || roind > h->nrots)
    return fail(const_cast<pftcc*>(h), PFTCC_E_ARG, "roind is out of range; get_rot (roind %d, nrots %d)", roind,
                h->nrots);
  *rot = h->angtab_deg[roind - 1];
  return 0;
}
int pftcc_get_roind(const pftcc_t* h, float rot, int32_t* ind) {
  if (!h || !ind) return PFTCC_E_ARG;
  int best = 0;
  float bd = 0.f;
  for (int i = 0; i < h->nrots; ++i) {
    float d = std::fabs(h->angtab_deg[i] - rot);
    if (d > 180.f) d = 360.f - d;
    if (i == 0 || d < bd) {
      bd = d;
      best = i;
    }
  }
  *ind = best + 1;
  return 0;
}
int pftcc_get_coord(const pftcc_t* h, int32_t irot, int32_t k, float xy[2]) {
  if (!h || !xy) return PFTCC_E_ARG;
  if (irot < 1 || irot > h->nrots || k < h->kfrom || k > h->kto)
    return fail(const_cast<pftcc*>(h), PFTCC_E_ARG, "get_coord (%d, %d) out of range", irot, k);
  xy[0] = h->polar_x[(size_t)(k - h->kfrom) * h->nrots + irot - 1];
  xy[1] = h->polar_y[(size_t)(k - h->kfrom) * h->nrots + irot - 1];
  return 0;
}
int pftcc_ptcl_iseven(const pftcc_t* h, int32_t iptcl, int32_t* v) {
  if (!h || !v) return PFTCC_E_ARG;
  int ip, rc;
  if ((rc = local_index(const_cast<pftcc*>(h), iptcl, &ip))) return rc;
  *v = h->iseven[ip];
  return 0;
}
int pftcc_get_ref_pft(pftcc_t* h, int32_t iref, int32_t iseven, float* pft) {
  if (!h || !pft) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc;
  if ((rc = check_ref(h, iref))) return rc;
  const size_t ring = (size_t)h->nk * h->n;
  HIPCHK(h, sync_copy(h, pft, (iseven ? h->d_refs_e : h->d_refs_o) + (size_t)(iref - 1) * ring, ring * sizeof(float2),
                      hipMemcpyDeviceToHost));
  return 0;
}
int pftcc_get_ptcl_pft(pftcc_t* h, int32_t iptcl, float* pft) {
  if (!h || !pft) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc, ip;
  if ((rc = local_index(h, iptcl, &ip))) return rc;
  const size_t ring = (size_t)h->nk * h->n;
  HIPCHK(h, sync_copy(h, pft, h->d_ptcls + (size_t)ip * ring, ring * sizeof(float2), hipMemcpyDeviceToHost));
  return 0;
}
int pftcc_get_ctfmat(pftcc_t* h, int32_t iptcl, float* out) {
  if (!h || !out) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc, ip;
  if ((rc = local_index(h, iptcl, &ip))) return rc;
  if (!h->ctf_ready) return fail(h, PFTCC_E_STATE, "CTF matrices not created");
  const size_t ring = (size_t)h->nk * h->n;
  HIPCHK(h, sync_copy(h, out, h->d_ctf + (size_t)ip * ring, ring * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}
int pftcc_get_memo(pftcc_t* h, int32_t iptcl, float* out) {
  if (!h || !out) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc, ip;
  if ((rc = local_index(h, iptcl, &ip))) return rc;
  if (!h->memo_valid) return fail(h, PFTCC_E_STATE, "memo not computed");
  const size_t per = (size_t)h->nk * (h->n + 1);
  HIPCHK(h, sync_copy(h, out, h->d_memo + (size_t)ip * per, per * sizeof(float2), hipMemcpyDeviceToHost));
  return 0;
}
int pftcc_get_sqsum_ptcl(pftcc_t* h, int32_t iptcl, float* out) {
  if (!h || !out) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc, ip;
  if ((rc = local_index(h, iptcl, &ip))) return rc;
  HIPCHK(h, sync_copy(h, out, h->d_sqsums + ip, sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

// ------------------------------------------------------------------------- calculators ---
int pftcc_gencorrs(pftcc_t* h, int32_t iref, int32_t iptcl, const float* shvec, float* cc) {
  if (!h || !cc) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc, ip;
  if ((rc = check_ref(h, iref)) || (rc = local_index(h, iptcl, &ip))) return rc;
  if ((rc = need_memo(h))) return rc;
  const size_t ring = (size_t)h->nk * h->n;
  if ((rc = scratch_reserve(h, aligned(64) + aligned(16) + aligned(ring * 8) + aligned(h->nrots * 4)))) return rc;
  int* d_idx = take<int>(h, 2);
  double* d_sh = take<double>(h, 2);
  float2* d_tab = take<float2>(h, ring);
  float* d_cc = take<float>(h, h->nrots);
  const int idx[2] = {ip, iref - 1};
  HIPCHK(h, hipMemcpyAsync(d_idx, idx, sizeof idx, hipMemcpyHostToDevice, h->stream));
  if (shvec) {
    const double sh[2] = {(double)shvec[0], (double)shvec[1]};
    HIPCHK(h, hipMemcpyAsync(d_sh, sh, sizeof sh, hipMemcpyHostToDevice, h->stream));
    if ((rc = shift_table(h, 1, d_sh, d_tab))) return rc;
  }
  if ((rc = launch_corr(h, 1, d_idx, 1, d_idx + 1, 1, d_tab, shvec != nullptr, d_cc, nullptr))) return rc;
  HIPCHK(h, hipMemcpyAsync(cc, d_cc, h->nrots * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFTCC_OK;
}

int pftcc_gencorr_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2], int32_t irot, double* f) {
  if (!h || !f) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  return rot8_single(h, iref, iptcl, shvec, irot, f, nullptr, false);
}

int pftcc_gencorr_for_rot(pftcc_t* h, int32_t iref, int32_t iptcl, const float shvec[2], int32_t irot, float* f) {
  if (!h || !f) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  const double sh[2] = {shvec ? (double)shvec[0] : 0.0, shvec ? (double)shvec[1] : 0.0};
  double v = 0.0;
  const int rc = rot8_single(h, iref, iptcl, sh, irot, &v, nullptr, false);
  *f = (float)v;
  return rc;
}

int pftcc_gencorr_grad_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2], int32_t irot,
                                 double* f, double grad[2]) {
  if (!h || !f || !grad) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  double gg[3];
  const int rc = rot8_single(h, iref, iptcl, shvec, irot, f, gg, false);
  grad[0] = gg[0];
  grad[1] = gg[1];
  return rc;
}

int pftcc_gencorr_grad_only_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2], int32_t irot,
                                      double grad[2]) {
  if (!h || !grad) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  double gg[3], f;
  const int rc = rot8_single(h, iref, iptcl, shvec, irot, &f, gg, false);
  grad[0] = gg[0];
  grad[1] = gg[1];
  return rc;
}

// The objective-specific forms: the same kernel with the objective fixed instead of read from
// cc_objfun.
int pftcc_gencorr_cc_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2], int32_t irot,
                               double* f) {
  if (!h || !f) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  return rot8_single(h, iref, iptcl, shvec, irot, f, nullptr, false, PFTCC_OBJFUN_CC);
}

int pftcc_gencorr_euclid_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2], int32_t irot,
                                   double* f) {
  if (!h || !f) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  return rot8_single(h, iref, iptcl, shvec, irot, f, nullptr, false, PFTCC_OBJFUN_EUCLID);
}

int pftcc_gencorr_cc_for_rot(pftcc_t* h, int32_t iref, int32_t iptcl, const float shvec[2], int32_t irot, float* f) {
  if (!h || !f) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  const double sh[2] = {shvec ? (double)shvec[0] : 0.0, shvec ? (double)shvec[1] : 0.0};
  double v = 0.0;
  const int rc = rot8_single(h, iref, iptcl, sh, irot, &v, nullptr, false, PFTCC_OBJFUN_CC);
  *f = (float)v;
  return rc;
}

int pftcc_gencorr_euclid_for_rot(pftcc_t* h, int32_t iref, int32_t iptcl, const float shvec[2], int32_t irot,
                                 float* f) {
  if (!h || !f) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  const double sh[2] = {shvec ? (double)shvec[0] : 0.0, shvec ? (double)shvec[1] : 0.0};
  double v = 0.0;
  const int rc = rot8_single(h, iref, iptcl, sh, irot, &v, nullptr, false, PFTCC_OBJFUN_EUCLID);
  *f = (float)v;
  return rc;
}

static int grad_for_rot_obj(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2], int32_t irot, double* f,
                            double grad[2], int objfun) {
  if (!h || !grad) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  double gg[3], fv;
  const int rc = rot8_single(h, iref, iptcl, shvec, irot, &fv, gg, false, objfun);
  if (f) *f = fv;
  grad[0] = gg[0];
  grad[1] = gg[1];
  return rc;
}

int pftcc_gencorr_cc_grad_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2], int32_t irot,
                                    double* f, double grad[2]) {
  if (!f) return PFTCC_E_ARG;
  return grad_for_rot_obj(h, iref, iptcl, shvec, irot, f, grad, PFTCC_OBJFUN_CC);
}

int pftcc_gencorr_cc_grad_only_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                         int32_t irot, double grad[2]) {
  return grad_for_rot_obj(h, iref, iptcl, shvec, irot, nullptr, grad, PFTCC_OBJFUN_CC);
}

int pftcc_gencorr_euclid_grad_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                        int32_t irot, double* f, double grad[2]) {
  if (!f) return PFTCC_E_ARG;
  return grad_for_rot_obj(h, iref, iptcl, shvec, irot, f, grad, PFTCC_OBJFUN_EUCLID);
}

int pftcc_gencorr_euclid_grad_only_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                             int32_t irot, double grad[2]) {
  return grad_for_rot_obj(h, iref, iptcl, shvec, irot, nullptr, grad, PFTCC_OBJFUN_EUCLID);
}

// ---------------------------------------------------- orientation derivatives (cont) ---
static int ensure_drefs(pftcc* h) {
  if (h->d_drefs_e) return PFTCC_OK;
  h->ndslots = std::max(1, h->cfg.nthr);
  const size_t cnt = (size_t)h->ndslots * 3 * h->kcorr * h->n;
  int rc;
  if ((rc = dalloc(h, &h->d_drefs_e, cnt)) || (rc = dalloc(h, &h->d_drefs_o, cnt))) return rc;
  HIPCHK(h, sync_memset(h, h->d_drefs_e, 0, cnt * sizeof(float2)));
  HIPCHK(h, sync_memset(h, h->d_drefs_o, 0, cnt * sizeof(float2)));
  return PFTCC_OK;
}

static int check_dslot(pftcc* h, int iref) {
  int rc;
  if ((rc = check_ref(h, iref)) || (rc = ensure_drefs(h))) return rc;
  if (iref > h->ndslots)
    return fail(h, PFTCC_E_ARG, "iref %d: the derivative slots are 1..nthr = %d (pfts_drefs, :321-322)", iref,
                h->ndslots);
  return PFTCC_OK;
}

int pftcc_set_dref_fcomp(pftcc_t* h, int32_t iref, int32_t irot, int32_t k, const float dcomp[6], int32_t iseven) {
  if (!h || !dcomp) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc;
  if ((rc = check_dslot(h, iref))) return rc;
  if (irot < 1 || irot > h->n || k < h->kfrom || k > h->kstop)
    return fail(h, PFTCC_E_ARG, "component (%d, %d) outside the derivative block (kfrom..kstop)", irot, k);
  float2* base = (iseven ? h->d_drefs_e : h->d_drefs_o) + (size_t)(iref - 1) * 3 * h->kcorr * h->n;
  for (int d = 0; d < 3; ++d) {
    const float2 v = make_float2(dcomp[2 * d], dcomp[2 * d + 1]);
    HIPCHK(h, sync_copy(h, base + ((size_t)d * h->kcorr + (k - h->kfrom)) * h->n + (irot - 1), &v, sizeof v,
                        hipMemcpyHostToDevice));
  }
  return PFTCC_OK;
}

int pftcc_set_dref_pft(pftcc_t* h, int32_t iref, const float* dpft, int32_t iseven) {
  if (!h || !dpft) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc;
  if ((rc = check_dslot(h, iref))) return rc;
  const size_t cnt = (size_t)3 * h->kcorr * h->n;
  float2* base = (iseven ? h->d_drefs_e : h->d_drefs_o) + (size_t)(iref - 1) * cnt;
  HIPCHK(h, sync_copy(h, base, dpft, cnt * sizeof(float2), hipMemcpyHostToDevice));
  return PFTCC_OK;
}

// y, S_q, T1[3], T2[3], gx, gy of k_cont8 for one (iref, iptcl, shift, irot).
static int cont8_single(pftcc* h, int iref, int iptcl, const double shvec[2], int irot, double out[kCont8Sums]) {
  int rc, ip;
  if ((rc = check_dslot(h, iref)) || (rc = local_index(h, iptcl, &ip)) || (rc = check_rot(h, irot))) return rc;
  if (iref > h->nptcls)
    return fail(h, PFTCC_E_ARG, "iref %d: the T1 terms read particle slot iref (:1781, :1836), nptcls = %d", iref,
                h->nptcls);
  if ((rc = need_direct(h))) return rc;
  struct {
    int iref, ip, it1, rho;
    double sh[2];
  } in{iref - 1, ip, iref - 1, irot - 1, {shvec ? shvec[0] : 0.0, shvec ? shvec[1] : 0.0}};
  if ((rc = scratch_reserve(h, 4096))) return rc;
  char* d_in = take<char>(h, sizeof in);
  double* d_out = take<double>(h, kCont8Sums);
  HIPCHK(h, hipMemcpyAsync(d_in, &in, sizeof in, hipMemcpyHostToDevice, h->stream));
  const int* di = reinterpret_cast<const int*>(d_in);
  Cont8Args a{};
  a.n = h->n;
  a.nk = h->nk;
  a.kcorr = h->kcorr;
  a.kfrom = h->kfrom;
  a.nrefs = h->nrefs;
  a.ndslots = h->ndslots;
  a.refs_even = h->d_refs_e;
  a.refs_odd = h->d_refs_o;
  a.drefs_even = h->d_drefs_e;
  a.drefs_odd = h->d_drefs_o;
  a.ptcls = h->d_ptcls;
  a.ctf = h->with_ctf ? h->d_ctf : nullptr;
  a.iseven = h->d_iseven;
  const bool norm = h->match_filt || h->pssnr;
  a.refinv = norm ? h->d_refinv_d : nullptr;
  a.ofac = norm ? h->d_ofac_d : nullptr;
  a.optrow = (norm && h->clsfrcs) ? h->d_optrow : nullptr;
  a.sn32 = h->d_sn32;
  a.ncn32 = h->d_ncn32;
  a.arg_a = kDPi / (double)(h->ldim / 2);
  a.argx = h->d_argx;
  a.argy = h->d_argy;
  a.one_x = h->d_onex;
  a.one_y = h->d_oney;
  a.irefs = di;
  a.ips = di + 1;
  a.it1s = di + 2;
  a.rhos = di + 3;
  a.shifts = reinterpret_cast<const double*>(d_in + 16);
  a.nitems = 1;
  a.out = d_out;
  const int nt = std::min(256, (h->n + 63) / 64 * 64);
  hipLaunchKernelGGL(k_cont8, dim3(1), dim3(nt), 0, h->stream, a);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(out, d_out, kCont8Sums * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFTCC_OK;
}

int pftcc_gencorr_cont_grad_cc_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                         int32_t irot, double* cc, double dcc[3]) {
  if (!h || !cc || !dcc) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  double o[kCont8Sums];
  int rc, ip;
  if ((rc = cont8_single(h, iref, iptcl, shvec, irot, o))) return rc;
  (void)local_index(h, iptcl, &ip);
  float sp;
  HIPCHK(h, sync_copy(h, &sp, h->d_sqsums + ip, sizeof sp, hipMemcpyDeviceToHost));
  // cc = sum(fdf_y) / denom, dcc(j) = (sum(T1(:,j)) - sum(fdf_y) sum(T2(:,j)) / sqsum_ref) / denom (:1777-1786)
  const double denom = sqrt(o[1] * (double)sp);
  *cc = o[0] / denom;
  for (int d = 0; d < 3; ++d) dcc[d] = (o[2 + d] - o[0] * o[5 + d] / o[1]) / denom;
  return PFTCC_OK;
}

int pftcc_gencorr_cont_shift_grad_cc_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                               int32_t irot, double* f, double grad[5]) {
  if (!h || !f || !grad) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  double o[kCont8Sums];
  int rc, ip;
  if ((rc = cont8_single(h, iref, iptcl, shvec, irot, o))) return rc;
  (void)local_index(h, iptcl, &ip);
  float sp;
  HIPCHK(h, sync_copy(h, &sp, h->d_sqsums + ip, sizeof sp, hipMemcpyDeviceToHost));
  // (:1832-1846): three orientation derivatives, then the two shift derivatives
  const double denom = sqrt(o[1] * (double)sp);
  *f = o[0] / denom;
  for (int d = 0; d < 3; ++d) grad[d] = (o[2 + d] - o[0] * o[5 + d] / o[1]) / denom;
  grad[3] = o[8] / denom;
  grad[4] = o[9] / denom;
  return PFTCC_OK;
}

// The reference's Euclid forms of these two are unimplemented stubs (:2050-2069: cc = 0, the
// derivatives left unset); these return 0 everywhere.
int pftcc_gencorr_cont_grad_euclid_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                             int32_t irot, double* cc, double dcc[3]) {
  (void)shvec;
  if (!h || !cc || !dcc) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc, ip;
  if ((rc = check_ref(h, iref)) || (rc = local_index(h, iptcl, &ip)) || (rc = check_rot(h, irot))) return rc;
  *cc = 0.0;
  dcc[0] = dcc[1] = dcc[2] = 0.0;
  return PFTCC_OK;
}

int pftcc_gencorr_cont_shift_grad_euclid_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                                   int32_t irot, double* f, double grad[5]) {
  (void)shvec;
  if (!h || !f || !grad) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc, ip;
  if ((rc = check_ref(h, iref)) || (rc = local_index(h, iptcl, &ip)) || (rc = check_rot(h, irot))) return rc;
  *f = 0.0;
  for (int d = 0; d < 5; ++d) grad[d] = 0.0;
  return PFTCC_OK;
}

// --------------------------------------------------------------- remaining getters ---
int pftcc_get_pfromto(const pftcc_t* h, int32_t pfromto[2]) {
  if (!h || !pfromto) return PFTCC_E_ARG;
  pfromto[0] = h->pfrom;
  pfromto[1] = h->pto;
  return PFTCC_OK;
}

int pftcc_assign_pinds(const pftcc_t* h, int32_t* pinds) {
  if (!h || !pinds) return PFTCC_E_ARG;
  for (size_t i = 0; i < h->pinds.size(); ++i) pinds[i] = h->pinds[i];
  return PFTCC_OK;
}

int pftcc_cp_refs(pftcc_t* h, const pftcc_t* src) {
  if (!h || !src) return PFTCC_E_ARG;
  if (h == src) return PFTCC_OK;
  std::lock_guard<std::mutex> g(h->mtx);
  if (src->nrefs != h->nrefs || src->nk != h->nk || src->n != h->n || src->kfrom != h->kfrom)
    return fail(h, PFTCC_E_ARG, "cp_refs: reference blocks differ in shape (%d x %d x %d vs %d x %d x %d)",
                src->nrefs, src->nk, src->n, h->nrefs, h->nk, h->n);
  if (src->device != h->device) return fail(h, PFTCC_E_UNSUPPORTED, "cp_refs: handles on different GPUs");
  const size_t bytes = (size_t)h->nrefs * h->nk * h->n * sizeof(float2);
  // the source's pending work (its own stream) must be complete before its arrays are read
  HIPCHK(h, hipStreamSynchronize(src->stream));
  HIPCHK(h, sync_copy(h, h->d_refs_e, src->d_refs_e, bytes, hipMemcpyDeviceToDevice));
  HIPCHK(h, sync_copy(h, h->d_refs_o, src->d_refs_o, bytes, hipMemcpyDeviceToDevice));
  h->refnorm_valid = false;
  h->apack[0].valid = h->apack[1].valid = false;
  return PFTCC_OK;
}

int pftcc_calc_ptcl_pspec(pftcc_t* h, int32_t iptcl, float* pspec) {
  if (!h || !pspec) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc, ip;
  if ((rc = local_index(h, iptcl, &ip))) return rc;
  if ((rc = scratch_reserve(h, aligned((size_t)h->nk * 8)))) return rc;
  double* d_s = take<double>(h, h->nk);
  // FP64 shell sums of |p|^2 over the whole ring block, one wavefront per shell
  hipLaunchKernelGGL(k_shell_sq8, dim3((h->nk + 3) / 4), dim3(256), 0, h->stream,
                     (const float2*)(h->d_ptcls + (size_t)ip * h->nk * h->n), h->nk, h->n, d_s);
  HIPCHK(h, hipGetLastError());
  std::vector<double> s(h->nk);
  HIPCHK(h, hipMemcpyAsync(s.data(), d_s, h->nk * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  // pspec(:) = sum(csq_fast(pfts_ptcls(:,:,i)), dim=1) / real(pftsz) (:897-904)
  for (int kk = 0; kk < h->nk; ++kk) pspec[kk] = (float)s[kk] / (float)h->n;
  return PFTCC_OK;
}

int pftcc_gencorrs_batch(pftcc_t* h, int32_t np, const int32_t* iptcls, int32_t nr, const int32_t* irefs, int32_t ns,
                         const float* shifts, float* cc_out) {
  if (!h || !iptcls || !cc_out || np < 1 || nr < 1 || ns < 0) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc;
  std::vector<int> pl(np), rl(nr);
  for (int i = 0; i < np; ++i)
    if ((rc = local_index(h, iptcls[i], &pl[i]))) return rc;
  for (int i = 0; i < nr; ++i) {
    const int r = irefs ? irefs[i] : i + 1;
    if ((rc = check_ref(h, r))) return rc;
    rl[i] = r - 1;
  }
  if ((rc = need_memo(h))) return rc;
  const int nse = std::max(ns, 1);
  const size_t ring = (size_t)h->nk * h->n;
  const size_t ncc = (size_t)np * nr * nse * h->nrots;
  if ((rc = scratch_reserve(h, aligned(np * 4) + aligned(nr * 4) + aligned(nse * 16) + aligned(nse * ring * 8) +
                                   aligned(ncc * 4))))
    return rc;
  int* d_pl = take<int>(h, np);
  int* d_rl = take<int>(h, nr);
  double* d_sh = take<double>(h, 2 * nse);
  float2* d_tab = take<float2>(h, nse * ring);
  float* d_cc = take<float>(h, ncc);
  HIPCHK(h, hipMemcpyAsync(d_pl, pl.data(), np * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d_rl, rl.data(), nr * 4, hipMemcpyHostToDevice, h->stream));
  if (ns > 0) {
    std::vector<double> sh(2 * ns);
    for (int i = 0; i < 2 * ns; ++i) sh[i] = (double)shifts[i];
    HIPCHK(h, hipMemcpyAsync(d_sh, sh.data(), sh.size() * 8, hipMemcpyHostToDevice, h->stream));
    if ((rc = shift_table(h, ns, d_sh, d_tab))) return rc;
  }
  if ((rc = launch_corr(h, np, d_pl, nr, d_rl, nse, d_tab, ns > 0, d_cc, nullptr))) return rc;
  HIPCHK(h, hipMemcpyAsync(cc_out, d_cc, ncc * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFTCC_OK;
}

static int search_impl(pftcc_t* h, int np, const int* d_pl, int nr, const int* d_rl, int ns, const float2* d_tab,
                       unsigned long long* d_best, float* d_cc, int* d_ir, int* d_is, int* d_rot) {
  int rc;
  hipLaunchKernelGGL(k_fill_u64, dim3((np + 255) / 256), dim3(256), 0, h->stream, d_best, np, 0ull);
  HIPCHK(h, hipGetLastError());
  if ((rc = launch_corr(h, np, d_pl, nr, d_rl, std::max(ns, 1), d_tab, ns > 0, nullptr, d_best))) return rc;
  hipLaunchKernelGGL(k_best_decode, dim3((np + 255) / 256), dim3(256), 0, h->stream, (const unsigned long long*)d_best,
                     np, nr, std::max(ns, 1), h->nrots, d_rl, d_cc, d_ir, d_is, d_rot);
  HIPCHK(h, hipGetLastError());
  return PFTCC_OK;
}

int pftcc_search(pftcc_t* h, int32_t np, const int32_t* iptcls, int32_t nr, const int32_t* irefs, int32_t ns,
                 const float* shifts, float* best_cc, int32_t* best_iref, int32_t* best_ishift, int32_t* best_irot) {
  if (!h || !iptcls || np < 1 || nr < 1 || ns < 0 || !best_cc || !best_iref || !best_ishift || !best_irot)
    return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc;
  std::vector<int> pl(np), rl(nr);
  for (int i = 0; i < np; ++i)
    if ((rc = local_index(h, iptcls[i], &pl[i]))) return rc;
  for (int i = 0; i < nr; ++i) {
    const int r = irefs ? irefs[i] : i + 1;
    if ((rc = check_ref(h, r))) return rc;
    rl[i] = r - 1;
  }
  if ((rc = need_memo(h))) return rc;
  const int nse = std::max(ns, 1);
  const size_t ring = (size_t)h->nk * h->n;
  if ((rc = scratch_reserve(h, aligned(np * 4) + aligned(nr * 4) + aligned(nse * 16) + aligned(nse * ring * 8) +
                                   5 * aligned((size_t)np * 8))))
    return rc;
  int* d_pl = take<int>(h, np);
  int* d_rl = take<int>(h, nr);
  double* d_sh = take<double>(h, 2 * nse);
  float2* d_tab = take<float2>(h, nse * ring);
  float* d_cc = take<float>(h, np);
  int* d_ir = take<int>(h, np);
  int* d_is = take<int>(h, np);
  int* d_rot = take<int>(h, np);
  HIPCHK(h, hipMemcpyAsync(d_pl, pl.data(), np * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d_rl, rl.data(), nr * 4, hipMemcpyHostToDevice, h->stream));
  if (ns > 0) {
    std::vector<double> sh(2 * ns);
    for (int i = 0; i < 2 * ns; ++i) sh[i] = (double)shifts[i];
    HIPCHK(h, hipMemcpyAsync(d_sh, sh.data(), sh.size() * 8, hipMemcpyHostToDevice, h->stream));
    if ((rc = shift_table(h, ns, d_sh, d_tab))) return rc;
  }
  if ((rc = search_impl(h, np, d_pl, nr, d_rl, ns, d_tab, take<unsigned long long>(h, np), d_cc, d_ir, d_is, d_rot)))
    return rc;
  HIPCHK(h, hipMemcpyAsync(best_cc, d_cc, np * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(best_iref, d_ir, np * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(best_ishift, d_is, np * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(best_irot, d_rot, np * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFTCC_OK;
}

// Device-list form: d_iptcls are LOCAL 0-based particle slots here would be ambiguous, so the
// global ids are translated on the device side by the caller-visible contract: d_iptcls hold
// global ids; the translation table is uploaded once per call.
__global__ void k_translate(const int* __restrict__ gids, int np, const int* __restrict__ pinds, int pfrom,
                                   int span, int* __restrict__ out, int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int g = gids[i] - pfrom;
  int l = (g >= 0 && g < span) ? pinds[g] : 0;
  if (l <= 0) {
    *bad = 1;
    l = 1;
  }
  out[i] = l - 1;
}

__global__ void k_refs0(const int* __restrict__ irefs, int nr, int nrefs, int* __restrict__ out,
                               int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  int r = irefs ? irefs[i] : i + 1;
  if (r < 1 || r > nrefs) {
    *bad = 1;
    r = 1;
  }
  out[i] = r - 1;
}

__global__ void k_f2d(const float* __restrict__ s, int n, double* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (double)s[i];
}

int pftcc_search_dev(pftcc_t* h, int32_t np, const int32_t* d_iptcls, int32_t nr, const int32_t* d_irefs, int32_t ns,
                     const float* d_shifts, float* d_best_cc, int32_t* d_best_iref, int32_t* d_best_ishift,
                     int32_t* d_best_irot) {
  if (!h || !d_iptcls || np < 1 || nr < 1 || ns < 0) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  int rc;
  if (!d_irefs && nr != h->nrefs) return fail(h, PFTCC_E_ARG, "nr must equal nrefs when irefs is NULL");
  if ((rc = need_memo(h))) return rc;
  const int nse = std::max(ns, 1);
  const size_t ring = (size_t)h->nk * h->n;
  const int span = h->pto - h->pfrom + 1;
  if ((rc = scratch_reserve(h, aligned(np * 4) + aligned(nr * 4) + aligned(nse * 16) + aligned(nse * ring * 8) +
                                   aligned((size_t)span * 4) + aligned(64) + aligned((size_t)np * 8))))
    return rc;
  int* d_pl = take<int>(h, np);
  int* d_rl = take<int>(h, nr);
  double* d_sh = take<double>(h, 2 * nse);
  float2* d_tab = take<float2>(h, nse * ring);
  int* d_pinds = take<int>(h, span);
  int* d_bad = take<int>(h, 16);
  HIPCHK(h, hipMemcpyAsync(d_pinds, h->pinds.data(), (size_t)span * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemsetAsync(d_bad, 0, 4, h->stream));
  hipLaunchKernelGGL(k_translate, dim3((np + 255) / 256), dim3(256), 0, h->stream, d_iptcls, np, (const int*)d_pinds,
                     h->pfrom, span, d_pl, d_bad);
  hipLaunchKernelGGL(k_refs0, dim3((nr + 255) / 256), dim3(256), 0, h->stream, d_irefs, nr, h->nrefs, d_rl, d_bad);
  if (ns > 0) {
    hipLaunchKernelGGL(k_f2d, dim3((2 * ns + 255) / 256), dim3(256), 0, h->stream, d_shifts, 2 * ns, d_sh);
    if ((rc = shift_table(h, ns, d_sh, d_tab))) return rc;
  }
  int bad = 0;
  HIPCHK(h, hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (bad) return fail(h, PFTCC_E_ARG, "search_dev: particle or reference index out of range");
  return search_impl(h, np, d_pl, nr, d_rl, ns, d_tab, take<unsigned long long>(h, np), d_best_cc, d_best_iref,
                     d_best_ishift, d_best_irot);
}

int pftcc_search_dev_noalloc(pftcc_t* h, int32_t np, const int32_t* d_iptcls, int32_t nr, const int32_t* d_irefs,
                             int32_t ns, const float* d_shifts, float* d_best_cc, int32_t* d_best_iref,
                             int32_t* d_best_ishift, int32_t* d_best_irot) {
  if (!h || !d_iptcls || np < 1 || nr < 1 || ns < 0 || !d_best_cc || !d_best_iref || !d_best_ishift || !d_best_irot)
    return PFTCC_E_ARG;
  if (ns > 0 && !d_shifts) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  if (!d_irefs && nr != h->nrefs) return fail(h, PFTCC_E_ARG, "nr must equal nrefs when irefs is NULL");
  if (!h->noalloc_scratch || !h->d_pinds) return fail(h, PFTCC_E_STATE, "search_dev_noalloc: handle not set up");
  // the state checks of need_memo, without building anything: the shell-normalisation tables
  // (match_filt / pssnr) are built by the allocating entry points
  if (!h->memo_valid)
    return fail(h, PFTCC_E_STATE, "particle FFTs are not memoized: call memoize_ffts after setting particles");
  if (h->with_ctf && !h->ctf_ready)
    return fail(h, PFTCC_E_STATE, "CTF matrices missing: call create_polar_absctfmats");
  if (h->euclid && !h->have_sigma2) return fail(h, PFTCC_E_STATE, "objfun euclid needs assign_sigma2_noise");
  if ((h->match_filt || h->pssnr) && !h->refnorm_valid)
    return fail(h, PFTCC_E_STATE,
                "search_dev_noalloc: the shell-normalisation tables are stale; one pftcc_search or "
                "pftcc_search_dev call after the last reference or particle change builds them");
  const int nse = std::max(ns, 1);
  const size_t ring = (size_t)h->nk * h->n;
  const size_t need = aligned((size_t)np * 4) + aligned((size_t)nr * 4) + aligned((size_t)nse * 16) +
                      aligned((size_t)nse * ring * 8) + aligned(64) + aligned((size_t)np * 8);
  if (need > kNoallocScratchBytes)
    return fail(h, PFTCC_E_ARG,
                "search_dev_noalloc: %d particles x %d references x %d shifts need %zu bytes of scratch, "
                "the maximum is %zu (fewer shifts or particles per call)",
                np, nr, nse, need, kNoallocScratchBytes);
  // carve the static scratch as scratch_reserve/take carve the arena
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    unsigned char* p = h->noalloc_scratch + off;
    off += aligned(bytes);
    return p;
  };
  int* d_pl = (int*)carve((size_t)np * 4);
  int* d_rl = (int*)carve((size_t)nr * 4);
  double* d_sh = (double*)carve((size_t)nse * 16);
  float2* d_tab = (float2*)carve((size_t)nse * ring * 8);
  int* d_bad = (int*)carve(64);
  unsigned long long* d_best = (unsigned long long*)carve((size_t)np * 8);
  const int span = h->pto - h->pfrom + 1;
  // the previous call's contents remain: the flag is cleared by a fill kernel (no memset node)
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, h->stream, (unsigned int*)d_bad, 16, 0u);
  hipLaunchKernelGGL(k_translate, dim3((np + 255) / 256), dim3(256), 0, h->stream, d_iptcls, np,
                     (const int*)h->d_pinds, h->pfrom, span, d_pl, d_bad);
  hipLaunchKernelGGL(k_refs0, dim3((nr + 255) / 256), dim3(256), 0, h->stream, d_irefs, nr, h->nrefs, d_rl, d_bad);
  HIPCHK(h, hipGetLastError());
  int rc;
  if (ns > 0) {
    hipLaunchKernelGGL(k_f2d, dim3((2 * ns + 255) / 256), dim3(256), 0, h->stream, d_shifts, 2 * ns, d_sh);
    if ((rc = shift_table(h, ns, d_sh, d_tab))) return rc;
  }
  int bad = 0;
  HIPCHK(h, hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (bad) return fail(h, PFTCC_E_ARG, "search_dev_noalloc: particle or reference index out of range");
  // kernel timing creates its events lazily, so it stays off for this call
  const bool timed = h->timing;
  h->timing = false;
  rc = search_impl(h, np, d_pl, nr, d_rl, ns, d_tab, d_best, d_best_cc, d_best_iref, d_best_ishift, d_best_irot);
  h->timing = timed;
  return rc;
}

int pftcc_strategy_srch(pftcc_t* h, const pftcc_strategy_cfg* cfg, int32_t np, const int32_t* iptcls,
                        const int32_t* srch_order, const int8_t* ref_ok, const float* prev_corr, const float* rnd,
                        float* best_corr, int32_t* best_ref, int32_t* best_rot, int32_t* nrefs_eval,
                        int32_t* found_better, float* ref_corrs, int32_t* ref_rots, int32_t* sorted_refs) {
  if (!h || !cfg || !iptcls || np < 1 || !best_corr || !best_ref || !best_rot || !nrefs_eval || !found_better)
    return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  const int kind = cfg->kind;
  if (kind < PFTCC_SRCH_GREEDY2D || kind > PFTCC_SRCH_SNHC3D) return fail(h, PFTCC_E_ARG, "unknown strategy %d", kind);
  const bool stoch = kind == PFTCC_SRCH_SNHC2D || kind == PFTCC_SRCH_SHC3D || kind == PFTCC_SRCH_SNHC3D;
  const bool is3d = kind == PFTCC_SRCH_GREEDY3D || kind == PFTCC_SRCH_SHC3D || kind == PFTCC_SRCH_SNHC3D;
  const int nrefs = h->nrefs, nrots = h->nrots;
  if (stoch && !srch_order) return fail(h, PFTCC_E_ARG, "the stochastic strategies need srch_order");
  if ((kind == PFTCC_SRCH_SNHC2D || kind == PFTCC_SRCH_SHC3D) && !prev_corr)
    return fail(h, PFTCC_E_ARG, "snhc 2-D and shc 3-D need prev_corr");
  if (kind == PFTCC_SRCH_SNHC2D && !rnd) return fail(h, PFTCC_E_ARG, "snhc needs one uniform draw per particle (rnd)");
  if ((ref_corrs == nullptr) != (ref_rots == nullptr)) return fail(h, PFTCC_E_ARG, "ref_corrs and ref_rots go together");
  const int ntop = is3d && sorted_refs ? cfg->ntop : 0;
  if (is3d && sorted_refs && (ntop < 1 || ntop > nrefs)) return fail(h, PFTCC_E_ARG, "ntop %d outside [1, %d]", cfg->ntop, nrefs);
  int sortlen = 1;
  while (sortlen < nrefs) sortlen <<= 1;
  if (is3d && (size_t)sortlen * 8 > kMaxLds)
    return fail(h, PFTCC_E_UNSUPPORTED, "the 3-D strategies sort at most %zu references per particle", kMaxLds / 8);
  // snhc's neighbourhood: max(2, min(nrefs, nint(real(nrefs) * (1 - stoch_bound)))) in default real
  int bound = nrefs;
  if (kind == PFTCC_SRCH_SNHC2D) {
    if (!(cfg->stoch_bound >= 0.f && cfg->stoch_bound <= 1.f))
      return fail(h, PFTCC_E_ARG, "stoch_bound %g outside [0, 1]", cfg->stoch_bound);
    const float x = (float)nrefs * (1.0f - cfg->stoch_bound);
    bound = std::max(2, std::min(nrefs, (int)lroundf(x)));
  }
  if (kind == PFTCC_SRCH_SNHC3D) {
    if (cfg->szsn < 1 || cfg->szsn > nrefs) return fail(h, PFTCC_E_ARG, "szsn %d outside [1, %d]", cfg->szsn, nrefs);
    bound = cfg->szsn;
  }
  int rc;
  std::vector<int> pl(np);
  for (int i = 0; i < np; ++i)
    if ((rc = local_index(h, iptcls[i], &pl[i]))) return rc;
  std::vector<int> ord;
  if (stoch) {
    // each row a permutation of 1..nrefs (srch_order of prep_strategy2D_batch / prep_strategy3D)
    ord.resize((size_t)np * nrefs);
    std::vector<int> seen(nrefs, -1);
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < nrefs; ++j) {
        const int r = srch_order[(size_t)i * nrefs + j];
        if (r < 1 || r > nrefs || seen[r - 1] == i)
          return fail(h, PFTCC_E_ARG, "srch_order row %d is not a permutation of 1..%d", i + 1, nrefs);
        seen[r - 1] = i;
        ord[(size_t)i * nrefs + j] = r - 1;
      }
  }
  if (kind == PFTCC_SRCH_SNHC2D)
    for (int i = 0; i < np; ++i)
      if (!(rnd[i] >= 0.f && rnd[i] < 1.f)) return fail(h, PFTCC_E_ARG, "rnd[%d] = %g outside [0, 1)", i, rnd[i]);
  if ((rc = need_memo(h))) return rc;
  // Correlation vectors (gencorrs, :1594) of a chunk of particles x references stay in HBM.  The
  // greedy kinds correlate every particle with all references.  The stochastic kinds scan their
  // search order in rounds of kStrategyRound0, 2 kStrategyRound0, ... positions, and only the
  // particles whose scan has not ended (no exit, limit not reached) take part in the next round.
  const int lim = (kind == PFTCC_SRCH_SNHC2D || kind == PFTCC_SRCH_SNHC3D) ? std::min(bound, nrefs) : nrefs;
  const size_t per_ptcl = (size_t)(stoch ? lim : nrefs) * nrots * 4;
  // PFTCC_STRATEGY_CC_BYTES overrides the budget of the correlation vectors held at once (tests
  // use a small one to exercise the chunking)
  size_t budget = kStrategyCcBytes;
  if (const char* e = std::getenv("PFTCC_STRATEGY_CC_BYTES")) budget = std::max<size_t>(1, std::strtoull(e, nullptr, 10));
  const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)np, budget / per_ptcl));
  const size_t ccap = std::max((size_t)chunk * per_ptcl, (size_t)lim * nrots * 4);
  const size_t nref_all = (size_t)np * nrefs;
  const size_t ntab = stoch ? nref_all : (size_t)chunk * nrefs;   // per-reference maxloc tables
  if ((rc = scratch_reserve(h, aligned(np * 4) + (stoch ? aligned(nref_all * 4) : 0) + aligned(nrefs) +
                                   2 * aligned(np * 4) + aligned(ccap) + 2 * aligned(ntab * 4) +
                                   5 * aligned(np * 4) + (ref_corrs ? 2 * aligned(nref_all * 4) : 0) +
                                   aligned((size_t)np * ntop * 4) +
                                   (stoch ? 6 * aligned(np * 4) + aligned((size_t)np * lim * 4) : 0) + aligned(4))))
    return rc;
  int* d_pl = take<int>(h, np);
  int* d_ord = stoch ? take<int>(h, nref_all) : nullptr;
  int8_t* d_ok = ref_ok ? take<int8_t>(h, nrefs) : nullptr;
  float* d_prev = prev_corr ? take<float>(h, np) : nullptr;
  float* d_rnd = rnd ? take<float>(h, np) : nullptr;
  float* d_cc = take<float>(h, ccap / 4);
  float* d_vb = take<float>(h, ntab);
  int* d_vr = take<int>(h, ntab);
  float* d_bc = take<float>(h, np);
  int* d_br = take<int>(h, np);
  int* d_bro = take<int>(h, np);
  int* d_ne = take<int>(h, np);
  int* d_fb = take<int>(h, np);
  float* d_rc = ref_corrs ? take<float>(h, nref_all) : nullptr;
  int* d_rr = ref_corrs ? take<int>(h, nref_all) : nullptr;
  int* d_so = ntop ? take<int>(h, (size_t)np * ntop) : nullptr;
  HIPCHK(h, hipMemcpyAsync(d_pl, pl.data(), (size_t)np * 4, hipMemcpyHostToDevice, h->stream));
  if (d_ord) HIPCHK(h, hipMemcpyAsync(d_ord, ord.data(), nref_all * 4, hipMemcpyHostToDevice, h->stream));
  if (d_ok) HIPCHK(h, hipMemcpyAsync(d_ok, ref_ok, nrefs, hipMemcpyHostToDevice, h->stream));
  if (d_prev) HIPCHK(h, hipMemcpyAsync(d_prev, prev_corr, (size_t)np * 4, hipMemcpyHostToDevice, h->stream));
  if (d_rnd) HIPCHK(h, hipMemcpyAsync(d_rnd, rnd, (size_t)np * 4, hipMemcpyHostToDevice, h->stream));
  StrategyArgs a{};
  a.kind = kind;
  a.nrefs = nrefs;
  a.nrots = nrots;
  a.bound = bound;
  a.ntop = ntop;
  a.sortlen = sortlen;
  a.ref_ok = d_ok;
  if (!stoch) {
    for (int p0 = 0; p0 < np; p0 += chunk) {
      const int c = std::min(chunk, np - p0);
      if ((rc = launch_corr(h, c, d_pl + p0, nrefs, nullptr, 1, nullptr, false, d_cc, nullptr))) return rc;
      const long long nvec = (long long)c * nrefs;
      hipLaunchKernelGGL(k_ref_maxloc, dim3((unsigned)((nvec + 3) / 4)), dim3(256), 0, h->stream, (const float*)d_cc,
                         nvec, nrots, d_vb, d_vr);
      a.order = nullptr;
      a.prev_corr = d_prev ? d_prev + p0 : nullptr;
      a.rnd = nullptr;
      a.cc = d_cc;
      a.vbest = d_vb;
      a.vrot = d_vr;
      a.best_corr = d_bc + p0;
      a.best_ref = d_br + p0;
      a.best_rot = d_bro + p0;
      a.nrefs_eval = d_ne + p0;
      a.found = d_fb + p0;
      a.ref_corrs = d_rc ? d_rc + (size_t)p0 * nrefs : nullptr;
      a.ref_rots = d_rr ? d_rr + (size_t)p0 * nrefs : nullptr;
      a.sorted = d_so ? d_so + (size_t)p0 * ntop : nullptr;
      hipLaunchKernelGGL(k_strategy, dim3((unsigned)c), dim3(256), is3d ? (size_t)sortlen * 8 : 0, h->stream, a);
      HIPCHK(h, hipGetLastError());
    }
  } else {
    int* d_act = take<int>(h, np);
    int* d_rpl = take<int>(h, np);
    int* d_istar = take<int>(h, np);
    int* d_srot = take<int>(h, np);
    float* d_scorr = take<float>(h, np);
    int* d_done = take<int>(h, np);
    int* d_rl = take<int>(h, (size_t)np * lim);
    a.order = d_ord;
    a.prev_corr = d_prev;
    a.rnd = d_rnd;
    a.vbest = d_vb;
    a.vrot = d_vr;
    std::vector<int> act(np), done(np);
    for (int i = 0; i < np; ++i) act[i] = i;
    int i0 = 0, w = kStrategyRound0;
    while (!act.empty() && i0 < lim) {
      const int i1 = std::min(i0 + w, lim), wr = i1 - i0;
      const int nact = (int)act.size();
      HIPCHK(h, hipMemcpyAsync(d_act, act.data(), (size_t)nact * 4, hipMemcpyHostToDevice, h->stream));
      const int sub = (int)std::max<size_t>(1, std::min<size_t>((size_t)nact, std::min(ccap, budget) / ((size_t)wr * nrots * 4)));
      for (int a0 = 0; a0 < nact; a0 += sub) {
        const int c = std::min(sub, nact - a0);
        const long long nv = (long long)c * wr;
        hipLaunchKernelGGL(k_round_lists, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, h->stream,
                           (const int*)d_ord, (const int*)(d_act + a0), (const int*)d_pl, c, nrefs, i0, wr, d_rpl, d_rl);
        if ((rc = launch_corr(h, c, d_rpl, wr, d_rl, 1, nullptr, false, d_cc, nullptr, wr))) return rc;
        hipLaunchKernelGGL(k_round_maxloc, dim3((unsigned)((nv + 3) / 4)), dim3(256), 0, h->stream,
                           (const float*)d_cc, c, wr, nrots, (const int*)(d_act + a0), (const int*)d_rl, nrefs, d_vb,
                           d_vr);
        hipLaunchKernelGGL(k_strategy_round, dim3((unsigned)c), dim3(256), 0, h->stream, a,
                           (const int*)(d_act + a0), i0, i1, (const float*)d_cc, wr, d_istar, d_srot, d_scorr, d_done);
        HIPCHK(h, hipGetLastError());
      }
      HIPCHK(h, hipMemcpyAsync(done.data(), d_done, (size_t)np * 4, hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
      std::vector<int> next;
      for (int p : act)
        if (!done[p]) next.push_back(p);
      act.swap(next);
      i0 = i1;
      w *= 2;
    }
    a.cc = nullptr;
    a.istar_in = d_istar;
    a.srot_in = d_srot;
    a.scorr_in = d_scorr;
    a.best_corr = d_bc;
    a.best_ref = d_br;
    a.best_rot = d_bro;
    a.nrefs_eval = d_ne;
    a.found = d_fb;
    a.ref_corrs = d_rc;
    a.ref_rots = d_rr;
    a.sorted = d_so;
    hipLaunchKernelGGL(k_strategy, dim3((unsigned)np), dim3(256), is3d ? (size_t)sortlen * 8 : 0, h->stream, a);
    HIPCHK(h, hipGetLastError());
  }
  HIPCHK(h, hipMemcpyAsync(best_corr, d_bc, (size_t)np * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(best_ref, d_br, (size_t)np * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(best_rot, d_bro, (size_t)np * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(nrefs_eval, d_ne, (size_t)np * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(found_better, d_fb, (size_t)np * 4, hipMemcpyDeviceToHost, h->stream));
  if (d_rc) {
    HIPCHK(h, hipMemcpyAsync(ref_corrs, d_rc, nref_all * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(ref_rots, d_rr, nref_all * 4, hipMemcpyDeviceToHost, h->stream));
  }
  if (d_so) HIPCHK(h, hipMemcpyAsync(sorted_refs, d_so, (size_t)np * ntop * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFTCC_OK;
}

int pftcc_search_mfma(pftcc_t* h, int32_t np, const int32_t* iptcls, int32_t nr, const int32_t* irefs, int32_t ns,
                      const float* shifts, int32_t prec, int32_t ntop, float* best_cc, int32_t* best_iref,
                      int32_t* best_ishift, int32_t* best_irot) {
  if (!h || !iptcls || np < 1 || nr < 1 || ns < 0 || (ns > 0 && !shifts) || !best_cc || !best_iref || !best_irot)
    return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  if (ntop < 0 || ntop > RESCORE_MAX) return fail(h, PFTCC_E_ARG, "ntop %d outside [0, %d]", ntop, RESCORE_MAX);
  const std::vector<float> shv = shift_vector(ns, shifts);
  const int nse = ns > 0 ? ns : 1;
  const size_t nrows = (size_t)nr * nse;
  unsigned long long *d_best = nullptr, *d_prbest = nullptr, *d_cand = nullptr;
  float* d_cc = nullptr;
  int *d_ir = nullptr, *d_rot = nullptr, *d_ish = nullptr, *d_pl = nullptr;
  int *d_bin = nullptr, *d_hist = nullptr, *d_order = nullptr;
  double* d_sh = nullptr;
  int rc;
  auto release = [&]() {
    (void)hipStreamSynchronize(h->stream);
    dfree(d_best);
    dfree(d_cc);
    dfree(d_ir);
    dfree(d_rot);
    dfree(d_ish);
    dfree(d_prbest);
    dfree(d_cand);
    dfree(d_pl);
    dfree(d_sh);
    dfree(d_bin);
    dfree(d_hist);
    dfree(d_order);
  };
  // rescoring order: one bin per key row (global reference x shift) and one for particles without
  // candidates
  const size_t nbins = (size_t)h->nrefs * nse + 1;
  if ((ntop > 0 && PFTCC_RESCORE_ORDER &&
       ((rc = dalloc(h, &d_bin, (size_t)np)) || (rc = dalloc(h, &d_hist, nbins)) ||
        (rc = dalloc(h, &d_order, (size_t)np)))) ||
      (rc = dalloc(h, &d_best, (size_t)np)) || (rc = dalloc(h, &d_cc, (size_t)np)) || (rc = dalloc(h, &d_ir, (size_t)np)) ||
      (rc = dalloc(h, &d_rot, (size_t)np)) || (rc = dalloc(h, &d_ish, (size_t)np)) ||
      (ntop > 0 && ((rc = dalloc(h, &d_prbest, (size_t)np * nrows)) || (rc = dalloc(h, &d_cand, (size_t)np * ntop)) ||
                    (rc = dalloc(h, &d_pl, (size_t)np)) || (ns > 0 && (rc = dalloc(h, &d_sh, (size_t)2 * ns)))))) {
    release();
    return rc;
  }
  hipLaunchKernelGGL(k_fill_u64, dim3((np + 255) / 256), dim3(256), 0, h->stream, d_best, np, 0ull);
  std::vector<int> plocal;
  rc = mfma_run(h, np, iptcls, nr, irefs, shv, prec, ntop > 0 ? nullptr : d_best, nullptr, d_prbest, &plocal);
  hipEvent_t ev = nullptr;
  if (rc == PFTCC_OK && ntop > 0 && (rc = timing_begin(h, 3, &ev)) == PFTCC_OK) {
    std::vector<double> sh64(shv.begin(), shv.end());
    if (hipMemcpyAsync(d_pl, plocal.data(), np * sizeof(int), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        (ns > 0 && hipMemcpyAsync(d_sh, sh64.data(), sh64.size() * sizeof(double), hipMemcpyHostToDevice,
                                  h->stream) != hipSuccess))
      rc = fail(h, PFTCC_E_DEVICE, "search_mfma: upload failed");
    if (rc == PFTCC_OK) {
      hipLaunchKernelGGL(k_topt, dim3(np), dim3(256), 0, h->stream, (const unsigned long long*)d_prbest, np,
                         (int)nrows, ntop, d_cand);
      if (d_order) {
        const unsigned g = (unsigned)((np + 255) / 256);
        hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nbins + 255) / 256)), dim3(256), 0, h->stream,
                           (unsigned int*)d_hist, (int)nbins, 0u);
        hipLaunchKernelGGL(k_order_hist, dim3(g), dim3(256), 0, h->stream, (const unsigned long long*)d_cand, np, ntop,
                           2 * h->n, (int)nbins, d_bin, d_hist);
        hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(1024), 0, h->stream, d_hist, (int)nbins);
        hipLaunchKernelGGL(k_order_scatter, dim3(g), dim3(256), 0, h->stream, (const int*)d_bin, np, d_hist, d_order);
      }
      RescoreArgs ra{};
      ra.n = h->n;
      ra.nk = h->nk;
      ra.kcorr = h->kcorr;
      ra.T = ntop;
      ra.ns = nse;
      ra.knorm = ns > 0 ? h->nk : h->kcorr;
      ra.shifts = ns > 0 ? d_sh : nullptr;
      ra.argx = h->d_argx;
      ra.argy = h->d_argy;
      ra.refs_even = h->d_refs_e;
      ra.refs_odd = h->d_refs_o;
      ra.ptcls = h->d_ptcls;
      ra.ctf = h->with_ctf ? h->d_ctf : nullptr;
      ra.iseven = h->d_iseven;
      ra.plist = d_pl;
      ra.cand = d_cand;
      ra.np = np;
      ra.best_cc = d_cc;
      ra.best_iref = d_ir;
      ra.best_irot = d_rot;
      ra.best_ishift = d_ish;
      ra.order = d_order;
      // one wavefront per candidate; the staging of a shell needs at least n threads
      const int nw = std::max(std::min(RESCORE_WAVES, ntop), (h->n + 63) / 64);
      const size_t lds = (size_t)2 * h->n * (sizeof(float2) + sizeof(float));
      if (h->n > 64 * RESCORE_WAVES) {
        rc = fail(h, PFTCC_E_UNSUPPORTED, "search_mfma: rescoring supports pftsz <= %d", 64 * RESCORE_WAVES);
      } else if (h->n <= 512) {
        if (ns > 0) hipLaunchKernelGGL((k_rescore64<true, 8>), dim3(np), dim3(64 * nw), lds, h->stream, ra);
        else hipLaunchKernelGGL((k_rescore64<false, 8>), dim3(np), dim3(64 * nw), lds, h->stream, ra);
      } else {
        if (ns > 0) hipLaunchKernelGGL((k_rescore64<true, 16>), dim3(np), dim3(64 * nw), lds, h->stream, ra);
        else hipLaunchKernelGGL((k_rescore64<false, 16>), dim3(np), dim3(64 * nw), lds, h->stream, ra);
      }
      if (hipGetLastError() != hipSuccess) rc = fail(h, PFTCC_E_DEVICE, "search_mfma: rescoring launch failed");
    }
    const int rc2 = timing_end(h, ev);
    if (rc == PFTCC_OK) rc = rc2;
  } else if (rc == PFTCC_OK) {
    hipLaunchKernelGGL(k_best_decode, dim3((np + 255) / 256), dim3(256), 0, h->stream,
                       (const unsigned long long*)d_best, np, h->nrefs, nse, h->nrots, (const int*)nullptr, d_cc, d_ir,
                       d_ish, d_rot);
  }
  if (rc == PFTCC_OK) {
    if (hipMemcpyAsync(best_cc, d_cc, np * 4, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipMemcpyAsync(best_iref, d_ir, np * 4, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipMemcpyAsync(best_irot, d_rot, np * 4, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        (best_ishift && hipMemcpyAsync(best_ishift, d_ish, np * 4, hipMemcpyDeviceToHost, h->stream) != hipSuccess) ||
        hipStreamSynchronize(h->stream) != hipSuccess)
      rc = fail(h, PFTCC_E_DEVICE, "search_mfma: copy of the results failed");
  }
  release();
  return rc;
}

int pftcc_gencorrs_mfma_batch(pftcc_t* h, int32_t np, const int32_t* iptcls, int32_t nr, const int32_t* irefs,
                              int32_t ns, const float* shifts, int32_t prec, float* cc_out) {
  if (!h || !iptcls || np < 1 || nr < 1 || ns < 0 || (ns > 0 && !shifts) || !cc_out) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  const std::vector<float> shv = shift_vector(ns, shifts);
  float* d_cc = nullptr;
  int rc;
  const size_t ncc = (size_t)np * nr * (ns > 0 ? ns : 1) * h->nrots;
  if ((rc = dalloc(h, &d_cc, ncc))) return rc;
  rc = mfma_run(h, np, iptcls, nr, irefs, shv, prec, nullptr, d_cc);
  if (rc == PFTCC_OK && (hipMemcpyAsync(cc_out, d_cc, ncc * 4, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
                         hipStreamSynchronize(h->stream) != hipSuccess))
    rc = fail(h, PFTCC_E_DEVICE, "gencorrs_mfma_batch: copy failed");
  (void)hipStreamSynchronize(h->stream);
  dfree(d_cc);
  return rc;
}

int pftcc_grad_batch(pftcc_t* h, int32_t npairs, const int32_t* irefs, const int32_t* iptcls, const double* shvecs,
                     const int32_t* irots, double* f, double* grad) {
  if (!h || npairs < 0 || (npairs > 0 && (!irefs || !iptcls || !shvecs || !irots || !f))) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  if (npairs == 0) return PFTCC_OK;
  int rc;
  std::vector<int> ir(npairs), ip(npairs), rho(npairs);
  for (int i = 0; i < npairs; ++i) {
    if ((rc = check_ref(h, irefs[i])) || (rc = local_index(h, iptcls[i], &ip[i])) || (rc = check_rot(h, irots[i])))
      return rc;
    ir[i] = irefs[i] - 1;
    rho[i] = irots[i] - 1;
  }
  if ((rc = need_direct(h))) return rc;
  const bool angle = grad != nullptr && !h->euclid;
  if (angle && (rc = ensure_dptcls(h))) return rc;
  if ((rc = scratch_reserve(h, 3 * aligned((size_t)npairs * 4) + aligned((size_t)npairs * 16) +
                                   aligned((size_t)npairs * 8) + aligned((size_t)npairs * 24))))
    return rc;
  int* d_ir = take<int>(h, npairs);
  int* d_ip = take<int>(h, npairs);
  int* d_rho = take<int>(h, npairs);
  double* d_sh = take<double>(h, 2 * (size_t)npairs);
  double* d_f = take<double>(h, npairs);
  double* d_g = take<double>(h, 3 * (size_t)npairs);
  HIPCHK(h, hipMemcpyAsync(d_ir, ir.data(), npairs * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d_ip, ip.data(), npairs * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d_rho, rho.data(), npairs * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d_sh, shvecs, (size_t)npairs * 16, hipMemcpyHostToDevice, h->stream));
  if ((rc = launch_rot8(h, npairs, d_ir, d_ip, d_sh, d_rho, d_f, grad ? d_g : nullptr, angle))) return rc;
  HIPCHK(h, hipMemcpyAsync(f, d_f, (size_t)npairs * 8, hipMemcpyDeviceToHost, h->stream));
  if (grad) HIPCHK(h, hipMemcpyAsync(grad, d_g, (size_t)npairs * 24, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFTCC_OK;
}

int pftcc_grad_batch_dev(pftcc_t* h, int32_t npairs, const int32_t* d_irefs, const int32_t* d_iptcls,
                         const double* d_shvecs, const int32_t* d_irots, double* d_f, double* d_grad) {
  if (!h || npairs < 0) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  if (npairs == 0) return PFTCC_OK;
  int rc;
  if ((rc = need_direct(h))) return rc;
  const bool angle = d_grad != nullptr && !h->euclid;
  if (angle && (rc = ensure_dptcls(h))) return rc;
  const int span = h->pto - h->pfrom + 1;
  if ((rc = scratch_reserve(h, 3 * aligned((size_t)npairs * 4) + aligned((size_t)span * 4) + aligned(64)))) return rc;
  int* d_ir = take<int>(h, npairs);
  int* d_ip = take<int>(h, npairs);
  int* d_rho = take<int>(h, npairs);
  int* d_pinds = take<int>(h, span);
  int* d_bad = take<int>(h, 16);
  HIPCHK(h, hipMemcpyAsync(d_pinds, h->pinds.data(), (size_t)span * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemsetAsync(d_bad, 0, 4, h->stream));
  hipLaunchKernelGGL(k_translate, dim3((npairs + 255) / 256), dim3(256), 0, h->stream, d_iptcls, npairs,
                     (const int*)d_pinds, h->pfrom, span, d_ip, d_bad);
  hipLaunchKernelGGL(k_refs0, dim3((npairs + 255) / 256), dim3(256), 0, h->stream, d_irefs, npairs, h->nrefs, d_ir,
                     d_bad);
  hipLaunchKernelGGL(k_refs0, dim3((npairs + 255) / 256), dim3(256), 0, h->stream, d_irots, npairs, h->nrots, d_rho,
                     d_bad);
  int bad = 0;
  HIPCHK(h, hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (bad) return fail(h, PFTCC_E_ARG, "grad_batch_dev: index out of range");
  return launch_rot8(h, npairs, d_ir, d_ip, d_shvecs, d_rho, d_f, d_grad, angle);
}

// Per-shell terms at one rotation (genfrc / calc_frc / gencorr_sigma_contrib), FP64 direct sums:
// out[kk] = {Re sum q conj(P_rot), sum |q|^2, sum |p|^2, sum |q - P_rot|^2} for every shell.
__global__ void k_shell_terms(const float2* __restrict__ ref, const float2* __restrict__ ptc,
                                     const float* __restrict__ ctf, const float2* __restrict__ shtab, int nk, int n,
                                     int rho, const float* __restrict__ scale /*[nk] or null*/,
                                     double* __restrict__ out) {
  const int kk = blockIdx.x;
  __shared__ double red[4][256];
  double acc[4] = {0, 0, 0, 0};
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    float2 r = ref[(size_t)kk * n + j];
    float c = ctf ? ctf[(size_t)kk * n + j] : 1.0f;
    // FP32 preparation as prep_ref4corr / gencorr_sigma_contrib do
    float2 q = make_float2(r.x * c, r.y * c);
    if (scale) q = make_float2(q.x * scale[kk], q.y * scale[kk]);
    if (shtab) q = cmul(q, shtab[(size_t)kk * n + j]);
    int t = j + rho;
    bool cj = false;
    if (t >= 2 * n) t -= 2 * n;
    if (t >= n) { t -= n; cj = true; }
    const float2 pv = ptc[(size_t)kk * n + t];
    const double pr = pv.x, pim = cj ? -(double)pv.y : (double)pv.y;
    acc[0] += (double)q.x * pr + (double)q.y * pim;
    acc[1] += (double)q.x * q.x + (double)q.y * q.y;
    const float2 pj = ptc[(size_t)kk * n + j];
    acc[2] += (double)pj.x * pj.x + (double)pj.y * pj.y;
    const double dr = q.x - pr, di = q.y - pim;
    acc[3] += dr * dr + di * di;
  }
  for (int v = 0; v < 4; ++v) red[v][threadIdx.x] = acc[v];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int v = 0; v < 4; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int v = 0; v < 4; ++v) out[4 * kk + v] = red[v][0];
}

static int shell_terms(pftcc_t* h, int iref, int iptcl, const float* shvec, int irot, bool shellnorm,
                       std::vector<double>& terms) {
  int rc, ip;
  if ((rc = check_ref(h, iref)) || (rc = local_index(h, iptcl, &ip)) || (rc = check_rot(h, irot))) return rc;
  if ((rc = need_direct(h))) return rc;
  const size_t ring = (size_t)h->nk * h->n;
  if ((rc = scratch_reserve(h, aligned(16) + aligned(ring * 8) + aligned(h->nk * 4) + aligned(h->nk * 32)))) return rc;
  double* d_sh = take<double>(h, 2);
  float2* d_tab = take<float2>(h, ring);
  float* d_scale = take<float>(h, h->nk);
  double* d_out = take<double>(h, 4 * (size_t)h->nk);
  if (shvec) {
    const double sh[2] = {(double)shvec[0], (double)shvec[1]};
    HIPCHK(h, hipMemcpyAsync(d_sh, sh, sizeof sh, hipMemcpyHostToDevice, h->stream));
    if ((rc = shift_table(h, 1, d_sh, d_tab))) return rc;
  }
  const bool even = h->iseven[ip] != 0;
  bool use_scale = shellnorm && (h->match_filt || h->pssnr);
  if (use_scale) {
    // per-shell scale of this (eo, ref): ofac[row] * refinv, computed on the host from the tables
    std::vector<float> inv((size_t)2 * h->nrefs * h->kcorr), of;
    HIPCHK(h, sync_copy(h, inv.data(), h->d_refinv_f, inv.size() * 4, hipMemcpyDeviceToHost));
    std::vector<float> sc(h->nk, 1.0f);
    const int row = h->clsfrcs ? (iptcl - 1) : (iref - 1);
    for (int kk = 0; kk < h->kcorr; ++kk) {
      const float o = h->optlp.empty() ? 1.0f : h->optlp[(size_t)row * h->kcorr + kk];
      const float ofv = h->pssnr ? std::sqrt(o * (float)(h->kfrom + kk)) : o;
      sc[kk] = ofv * inv[((size_t)(even ? 0 : 1) * h->nrefs + (iref - 1)) * h->kcorr + kk];
    }
    HIPCHK(h, hipMemcpyAsync(d_scale, sc.data(), h->nk * 4, hipMemcpyHostToDevice, h->stream));
  }
  const float2* ref = (even ? h->d_refs_e : h->d_refs_o) + (size_t)(iref - 1) * ring;
  hipLaunchKernelGGL(k_shell_terms, dim3(h->nk), dim3(256), 0, h->stream, ref,
                     (const float2*)(h->d_ptcls + (size_t)ip * ring), h->with_ctf ? (const float*)(h->d_ctf + ip * ring) : nullptr,
                     shvec ? (const float2*)d_tab : nullptr, h->nk, h->n, irot - 1,
                     use_scale ? (const float*)d_scale : nullptr, d_out);
  HIPCHK(h, hipGetLastError());
  terms.resize(4 * (size_t)h->nk);
  HIPCHK(h, hipMemcpyAsync(terms.data(), d_out, terms.size() * 8, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFTCC_OK;
}

int pftcc_gencorr_sigma_contrib(pftcc_t* h, int32_t iref, int32_t iptcl, const float shvec[2], int32_t irot,
                                float* sigma_contrib, float* ptcl_sumsq, float* ref_sumsq) {
  if (!h || !sigma_contrib) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  std::vector<double> t;
  // gencorr_sigma_contrib applies CTF and shift, no shell normalisation (:2166-2175)
  int rc = shell_terms(h, iref, iptcl, shvec, irot, false, t);
  if (rc) return rc;
  for (int kk = 0; kk < h->nk; ++kk) {
    sigma_contrib[kk] = (float)(0.5 * t[4 * kk + 3] / (double)h->n);
    if (ptcl_sumsq) ptcl_sumsq[kk] = (float)t[4 * kk + 2];
    if (ref_sumsq) ref_sumsq[kk] = (float)t[4 * kk + 1];
  }
  return PFTCC_OK;
}

int pftcc_specscore(pftcc_t* h, int32_t iref, int32_t iptcl, int32_t irot, const float* shvec, float* score) {
  if (!h || !score) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  std::vector<double> t;
  int rc = shell_terms(h, iref, iptcl, shvec, irot, true, t);
  if (rc) return rc;
  std::vector<float> frc(h->kcorr);
  for (int kk = 0; kk < h->kcorr; ++kk) {
    const float den = (float)std::sqrt(t[4 * kk + 2] * t[4 * kk + 1]);
    frc[kk] = den < 1e-12f ? 0.0f : (float)(t[4 * kk] / std::sqrt(t[4 * kk + 2] * t[4 * kk + 1]));
  }
  // median_nocopy (simple_math.f90:2111-2130): the reference's quickselect selec on the array in
  // place, two successive calls for even sizes (the second sees the first's permutation)
  const int nf = (int)frc.size();
  float med;
  if (nf % 2 == 0) {
    const float v1 = prep_selec(nf / 2, nf, frc.data());
    const float v2 = prep_selec(nf / 2 + 1, nf, frc.data());
    med = (v1 + v2) / 2.0f;
  } else {
    med = prep_selec(nint32((float)nf / 2.0f), nf, frc.data());
  }
  *score = std::max(0.0f, med);
  return PFTCC_OK;
}

int pftcc_calc_roinv_corrmat(pftcc_t* h, float* corrmat) {
  if (!h || !corrmat) return PFTCC_E_ARG;
  int np, nr;
  pftcc_get_nptcls(h, &np);
  pftcc_get_nrefs(h, &nr);
  if (np != nr) return fail(h, PFTCC_E_ARG, "nptcls == nrefs in pftcc required; calc_roinv_corrmat");
  std::vector<int> ids(np);
  {
    std::lock_guard<std::mutex> g(h->mtx);
    int c = 0;
    for (int gidx = h->pfrom; gidx <= h->pto; ++gidx)
      if (h->pinds[gidx - h->pfrom] > 0) ids[c++] = gidx;
  }
  std::vector<float> cc((size_t)np * nr * h->nrots);
  int rc = pftcc_gencorrs_batch(h, np, ids.data(), nr, nullptr, 0, nullptr, cc.data());
  if (rc) return rc;
  // corrmat(iref, iptcl) for iref < iptcl, symmetric, unit diagonal (:2219-2228); Fortran order
  for (int i = 0; i < np * np; ++i) corrmat[i] = 1.0f;
  for (int p = 0; p < np; ++p)
    for (int r = 0; r < p; ++r) {
      const float* v = &cc[((size_t)p * nr + r) * h->nrots];
      float m = v[0];
      for (int t = 1; t < h->nrots; ++t) m = std::max(m, v[t]);
      corrmat[(size_t)p * np + r] = m;
      corrmat[(size_t)r * np + p] = m;
    }
  return PFTCC_OK;
}

}  // extern "C"

// ===================================================================== gradient shift search ===
namespace {

// gencorrs at one FP32 shift per item (the in-plane callback of the shift search): best key per
// item (cc, rotation) from the item mode of k_corr_generic.
#ifndef PFTCC_ITEMS_FUSED
#define PFTCC_ITEMS_FUSED 1       // shift-search callback on k_corr_fused_items for fused ring lengths
#endif

// The callback's correlations with the compile-time kernel for the fused ring lengths; false if
// n has none.
bool items_fused_go(int n, int nitems, hipStream_t s, const CorrArgs& a) {
#define X(N)                                                                                        \
  if (n == N) {                                                                                     \
    constexpr int F = items_F<N>();                                                                 \
    hipLaunchKernelGGL((k_corr_fused_items<N, F, 256>), dim3((unsigned)((nitems + F - 1) / F)), dim3(256), \
                       (items_lds_bytes<N, F>()), s, a);                                            \
    return true;                                                                                    \
  }
  PFTCC_FUSED_SIZES(X)
#undef X
  return false;
}

int launch_corr_items(pftcc* h, int n, const int* d_ip, const int* d_ref, const float* d_sh,
                      unsigned long long* d_best) {
  CorrArgs a{};
  a.n = h->n;
  a.nk = h->nk;
  a.kcorr = h->kcorr;
  a.mode = h->euclid ? 1 : 0;
  a.knorm = h->euclid ? h->kcorr : h->nk;      // the shifted gencorrs (:1531)
  a.nrefs = h->nrefs;
  a.plan = h->plan;
  a.refs_even = h->d_refs_e;
  a.refs_odd = h->d_refs_o;
  a.memo = h->d_memo;
  a.ctf = h->with_ctf ? h->d_ctf : nullptr;
  a.sqsums = h->d_sqsums;
  a.sqsums_k = h->d_sqsums_k;
  a.iseven = h->d_iseven;
  a.shtab = nullptr;
  a.ns = 1;
  const bool norm = (h->match_filt || h->pssnr) && !h->euclid;
  a.refinv = norm ? h->d_refinv_f : nullptr;
  a.ofac = norm ? h->d_ofac_f : nullptr;
  a.optrow = (norm && h->clsfrcs) ? h->d_optrow : nullptr;
  a.shellw = h->euclid ? h->d_shellw : nullptr;
  a.plist = d_ip;
  a.np = n;
  a.rlist = d_ref;
  a.nr = 1;
  a.cc_out = nullptr;
  a.best = d_best;
  a.tw2g = h->d_tw2;
  a.F = 1;
  a.ntiles = 1;
  a.item_shift = d_sh;
  a.argx = h->d_argx;
  a.argy = h->d_argy;
  a.onex = h->d_onex;
  a.oney = h->d_oney;
  const size_t lds = corr_lds_bytes_items(h->n);
  if (lds > kMaxLds) return fail(h, PFTCC_E_UNSUPPORTED, "pftsz %d too large for the shift-search callback", h->n);
  hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, h->stream, d_best, n, 0ull);
  hipEvent_t e1 = nullptr;
  int rc;
  if ((rc = timing_begin(h, 0, &e1))) return rc;
  if (!(PFTCC_ITEMS_FUSED && items_fused_go(h->n, n, h->stream, a))) {
    // one pair per work-group of 256 (64 and 128 were measured 1.6x and 1.2x slower)
    hipLaunchKernelGGL(k_corr_generic, dim3((unsigned)n), dim3(256), lds, h->stream, a);
  }
  HIPCHK(h, hipGetLastError());
  return timing_end(h, e1);
}

}  // namespace

extern "C" {

int pftcc_shsrch_grad_batch(pftcc_t* h, int32_t n, const int32_t* irefs, const int32_t* iptcls, const float* lims,
                            int32_t maxits, int32_t opt_angle, int32_t coarse_init, int32_t* irots, float* cxy) {
  if (!h || n < 0 || (n > 0 && (!irefs || !iptcls || !lims || !irots || !cxy))) return PFTCC_E_ARG;
  std::lock_guard<std::mutex> g(h->mtx);
  if (n == 0) return PFTCC_OK;
  int rc;
  // the rotation callback runs the FP32 search (memoized spectra); the objective alone does not
  if ((rc = opt_angle ? need_memo(h) : need_direct(h))) return rc;
  if (!(lims[0] <= lims[2]) || !(lims[1] <= lims[3]))
    return fail(h, PFTCC_E_ARG, "shift limits [%g, %g] x [%g, %g] are empty", lims[0], lims[2], lims[1], lims[3]);
  if (maxits < 0) return fail(h, PFTCC_E_ARG, "maxits %d < 0", maxits);
  if (opt_angle && corr_lds_bytes_items(h->n) > kMaxLds)
    return fail(h, PFTCC_E_UNSUPPORTED, "pftsz %d too large for the shift-search callback", h->n);
  std::vector<int> ref0(n), ip0(n), rot(n);
  for (int i = 0; i < n; ++i) {
    if ((rc = check_ref(h, irefs[i]))) return rc;
    if ((rc = local_index(h, iptcls[i], &ip0[i]))) return rc;
    ref0[i] = irefs[i] - 1;
    rot[i] = irots[i];
    if (!opt_angle && (rot[i] < 1 || rot[i] > h->nrots))
      return fail(h, PFTCC_E_ARG, "irot %d outside [1, %d]", rot[i], h->nrots);
  }
  // device state: items, two request-list sets, counters
  ShsItem* d_items = nullptr;
  int *d_i = nullptr, *d_cnt = nullptr, *d_rot = nullptr;
  double* d_d = nullptr;
  float* d_f = nullptr;
  unsigned long long* d_key = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(h->stream);
    dfree(d_items);
    dfree(d_i);
    dfree(d_cnt);
    dfree(d_rot);
    dfree(d_d);
    dfree(d_f);
    dfree(d_key);
  };
  const size_t N = (size_t)n;
  // ints per set: fg item, ref, ip, rho; corr item, ref, ip  (7 N); plus 2 x N init ids
  if ((rc = dalloc(h, &d_items, N)) || (rc = dalloc(h, &d_i, 2 * 7 * N + 2 * N)) || (rc = dalloc(h, &d_cnt, 4)) ||
      (rc = dalloc(h, &d_rot, N)) || (rc = dalloc(h, &d_d, 2 * 2 * N + 4 * N)) ||
      (rc = dalloc(h, &d_f, 2 * 2 * N + 3 * N + (size_t)h->nrots)) || (rc = dalloc(h, &d_key, N))) {
    cleanup();
    return rc;
  }
  float* d_out = d_f + 4 * N;            // cxy [n][3]
  float* d_angtab = d_out + 3 * N;
  int* set_i[2] = {d_i, d_i + 7 * N};
  int* d_ref_init = d_i + 14 * N;
  int* d_ip_init = d_ref_init + N;
  double* set_sh[2] = {d_d, d_d + 2 * N};
  double* d_fv = d_d + 4 * N;            // f [n]
  double* d_gr = d_fv + N;               // grad [n][3]
  float* set_csh[2] = {d_f, d_f + 2 * N};
  auto fail_clean = [&](int code) {
    cleanup();
    return code;
  };
  if (hipMemcpyAsync(d_ref_init, ref0.data(), N * 4, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      hipMemcpyAsync(d_ip_init, ip0.data(), N * 4, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      hipMemcpyAsync(d_rot, rot.data(), N * 4, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      hipMemcpyAsync(d_angtab, h->angtab_deg.data(), (size_t)h->nrots * 4, hipMemcpyHostToDevice, h->stream) !=
          hipSuccess)
    return fail_clean(fail(h, PFTCC_E_DEVICE, "shsrch_grad_batch: upload failed"));
  ShsArgs a{};
  a.items = d_items;
  a.nitems = n;
  a.angtab_deg = d_angtab;
  a.max_evals = 5;                        // pftcc_shsrch_grad%max_evals
  a.nrots = h->nrots;
  a.lo0 = lims[0];
  a.lo1 = lims[1];
  a.hi0 = lims[2];
  a.hi1 = lims[3];
  a.maxits = maxits;
  a.opt_angle = opt_angle ? 1 : 0;
  a.coarse_init = coarse_init ? 1 : 0;
  auto bind_out = [&](ShsArgs& s, int set) {
    int* b = set_i[set];
    s.counters_out = d_cnt + 2 * set;
    s.fg_item_out = b;
    s.fg_ref_out = b + N;
    s.fg_ip_out = b + 2 * N;
    s.fg_rho_out = b + 3 * N;
    s.corr_item_out = b + 4 * N;
    s.corr_ref_out = b + 5 * N;
    s.corr_ip_out = b + 6 * N;
    s.fg_sh_out = set_sh[set];
    s.corr_sh_out = set_csh[set];
  };
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, h->stream, (unsigned*)d_cnt, 4, 0u);
  {
    ShsArgs s = a;
    s.init = 1;
    s.irefs = d_ref_init;
    s.ips = d_ip_init;
    s.irots = d_rot;
    bind_out(s, 0);
    hipLaunchKernelGGL(k_shs_step, dim3((n + 63) / 64), dim3(64), 0, h->stream, s);
    if (hipGetLastError() != hipSuccess) return fail_clean(fail(h, PFTCC_E_DEVICE, "k_shs_step launch failed"));
  }
  int cur = 0;
  // every item issues at most max_evals x (maxits + line-search trials + callbacks) requests
  const long max_rounds = 5L * (maxits + 2 * LB_MAXLS + 8) * 4 + 64 + 2 * SH_COARSE_STEPS * SH_COARSE_STEPS;
  long rounds = 0;
  for (;; ++rounds) {
    int cnt[2];
    if (hipMemcpyAsync(cnt, d_cnt + 2 * cur, 8, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
      return fail_clean(fail(h, PFTCC_E_DEVICE, "shsrch_grad_batch: device error"));
    const int nfg = cnt[0], ncorr = cnt[1];
    if (nfg == 0 && ncorr == 0) break;
    if (rounds > max_rounds)
      return fail_clean(fail(h, PFTCC_E_DEVICE, "shsrch_grad_batch: no convergence after %ld rounds", rounds));
    int* b = set_i[cur];
    if (nfg > 0 && (rc = launch_rot8(h, nfg, b + N, b + 2 * N, set_sh[cur], b + 3 * N, d_fv, d_gr, false)))
      return fail_clean(rc);
    if (ncorr > 0 && (rc = launch_corr_items(h, ncorr, b + 6 * N, b + 5 * N, set_csh[cur], d_key)))
      return fail_clean(rc);
    const int nxt = 1 - cur;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, h->stream, (unsigned*)(d_cnt + 2 * nxt), 2, 0u);
    ShsArgs s = a;
    s.init = 0;
    s.nfg_in = nfg;
    s.ncorr_in = ncorr;
    s.fg_item_in = b;
    s.fg_f_in = d_fv;
    s.fg_grad_in = d_gr;
    s.corr_item_in = b + 4 * N;
    s.corr_key_in = d_key;
    bind_out(s, nxt);
    const int tot = nfg + ncorr;
    hipLaunchKernelGGL(k_shs_step, dim3((tot + 63) / 64), dim3(64), 0, h->stream, s);
    if (hipGetLastError() != hipSuccess) return fail_clean(fail(h, PFTCC_E_DEVICE, "k_shs_step launch failed"));
    cur = nxt;
  }
  hipLaunchKernelGGL(k_shs_out, dim3((n + 255) / 256), dim3(256), 0, h->stream, (const ShsItem*)d_items, n, d_out,
                     d_rot);
  if (hipMemcpyAsync(cxy, d_out, N * 12, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipMemcpyAsync(irots, d_rot, N * 4, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    return fail_clean(fail(h, PFTCC_E_DEVICE, "shsrch_grad_batch: download failed"));
  cleanup();
  return PFTCC_OK;
}

}  // extern "C"
