// synth.hip - seeded synthetic polar Fourier transforms for benchmarks and large parity tests.
//
// Phantom: a sum of 3-D Gaussian blobs (centre c_b, width s_b, amplitude a_b, in pixels).  Its
// projection along a unit direction d, seen in the plane spanned by (u, v), is a sum of 2-D
// Gaussians whose Fourier transform is known in closed form:
//   P(f) = sum_b a_b sqrt(2 pi) s_b * 2 pi s_b^2 * exp(-2 pi^2 s_b^2 |f|^2) * exp(-2 pi i f.c_b'),
// with c_b' = (c_b.u, c_b.v) and f in cycles/pixel.  Rings are sampled on the polar grid of the
// correlation search: f = (k sin(theta_j), -k cos(theta_j)) / box, theta_j = j pi / pftsz.
// Particles are a reference rotated in-plane by psi (P(R(-psi) f)), shifted by s
// (exp(-2 pi i f.s)), multiplied by |CTF| (phase-flipped particles, as the search assumes) and
// given white complex Gaussian noise of a fixed variance per sample.
//
// This is input generation only (it replaces the polarizer of real data); the correlation path
// never calls it.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

struct Blob {
  float cx, cy, cz, sigma, amp;
};

__device__ inline uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

__device__ inline float u01(uint32_t a, uint32_t b, uint32_t c) {
  const uint32_t h = hash32(a ^ hash32(b ^ hash32(c * 0x9E3779B9U)));
  return ((h >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

// value of the projected phantom's FT at frequency (fx, fy) in the projection plane (u, v)
__device__ inline float2 proj_ft(const Blob* __restrict__ blobs, int nb, const float* __restrict__ uv, float fx,
                                 float fy) {
  const float f2 = fx * fx + fy * fy;
  float re = 0.f, im = 0.f;
  const float twopi = 6.283185307179586f;
  for (int b = 0; b < nb; ++b) {
    const Blob B = blobs[b];
    const float cu = B.cx * uv[0] + B.cy * uv[1] + B.cz * uv[2];
    const float cv = B.cx * uv[3] + B.cy * uv[4] + B.cz * uv[5];
    const float s2 = B.sigma * B.sigma;
    const float mag = B.amp * 2.5066282746f * B.sigma * twopi * s2 * __expf(-0.5f * twopi * twopi * s2 * f2);
    float sn, cs;
    __sincosf(-twopi * (fx * cu + fy * cv), &sn, &cs);
    re += mag * cs;
    im += mag * sn;
  }
  return make_float2(re, im);
}

// refs[r][kk][j]
__global__ void k_refs(const Blob* __restrict__ blobs, int nb, const float* __restrict__ uvs, int nrefs, int nk, int n,
                       int kfrom, int box, float2* __restrict__ out) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t tot = (size_t)nrefs * nk * n;
  if (idx >= tot) return;
  const int j = (int)(idx % n);
  const int kk = (int)((idx / n) % nk);
  const int r = (int)(idx / ((size_t)n * nk));
  const float th = 3.14159265358979f * j / n;
  const float k = (float)(kfrom + kk);
  const float fx = k * sinf(th) / box, fy = -k * cosf(th) / box;
  out[idx] = proj_ft(blobs, nb, uvs + 6 * r, fx, fy);
}

struct PtclParam {
  int32_t ref;
  float psi;       // in-plane rotation, radians
  float sx, sy;    // shift, pixels
  float wl, cs, ampc, dfx, dfy, angast;   // CTF in pixel units (as the correlation's CTF model)
  float noise_sd;  // per-component noise standard deviation
  uint32_t seed;
};

__global__ void k_ptcls(const Blob* __restrict__ blobs, int nb, const float* __restrict__ uvs,
                        const PtclParam* __restrict__ prm, int np, int nk, int n, int kfrom, int box, int with_ctf,
                        float2* __restrict__ out) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t tot = (size_t)np * nk * n;
  if (idx >= tot) return;
  const int j = (int)(idx % n);
  const int kk = (int)((idx / n) % nk);
  const int p = (int)(idx / ((size_t)n * nk));
  const PtclParam P = prm[p];
  const float th = 3.14159265358979f * j / n;
  const float k = (float)(kfrom + kk);
  const float fx = k * sinf(th) / box, fy = -k * cosf(th) / box;
  // rotated reference: evaluate at R(-psi) f
  float sp, cp;
  sincosf(P.psi, &sp, &cp);
  const float gx = cp * fx + sp * fy, gy = -sp * fx + cp * fy;
  float2 v = proj_ft(blobs, nb, uvs + 6 * P.ref, gx, gy);
  // shift
  float ss, cs;
  sincosf(-6.283185307179586f * (fx * P.sx + fy * P.sy), &ss, &cs);
  v = make_float2(v.x * cs - v.y * ss, v.x * ss + v.y * cs);
  if (with_ctf) {
    const float s2 = fx * fx + fy * fy;
    const float ang = atan2f(fy, fx);
    const float df = 0.5f * (P.dfx + P.dfy + cosf(2.f * (ang - P.angast)) * (P.dfx - P.dfy));
    const float chi = 3.14159265358979f * P.wl * s2 * (df - 0.5f * P.wl * P.wl * s2 * P.cs);
    const float c = fabsf(sinf(chi + P.ampc));
    v = make_float2(v.x * c, v.y * c);
  }
  if (P.noise_sd > 0.f) {
    const uint32_t e = (uint32_t)(((size_t)kk * n + j));
    const float u1 = u01(P.seed, e, 1u), u2 = u01(P.seed, e, 2u);
    const float r = sqrtf(-2.f * __logf(u1)) * P.noise_sd;
    float sn, cn;
    __sincosf(6.283185307179586f * u2, &sn, &cn);
    v.x += r * cn;
    v.y += r * sn;
  }
  out[idx] = v;
}

// Plain copy (one float4 per thread); a kernel of known traffic for calibrating HBM counters.
__global__ void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace

extern "C" {

// dst[0:bytes) = src[0:bytes); bytes a multiple of 16.
int synth_copy(void* dst, const void* src, size_t bytes) {
  const size_t n = bytes / 16;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, (const float4*)src, (float4*)dst, n);
  return hipDeviceSynchronize() == hipSuccess ? 0 : 3;
}

// blobs: [nb][5] (cx, cy, cz, sigma, amp); uvs: [nrefs][6] projection axes u, v.
int synth_refs(int nb, const float* blobs, int nrefs, const float* uvs, int nk, int n, int kfrom, int box,
               void* d_out) {
  Blob* d_b = nullptr;
  float* d_uv = nullptr;
  if (hipMalloc(&d_b, sizeof(Blob) * nb) != hipSuccess) return 3;
  if (hipMalloc(&d_uv, sizeof(float) * 6 * nrefs) != hipSuccess) return 3;
  if (hipMemcpy(d_b, blobs, sizeof(Blob) * nb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_uv, uvs, sizeof(float) * 6 * nrefs, hipMemcpyHostToDevice) != hipSuccess)
    return 3;
  const size_t tot = (size_t)nrefs * nk * n;
  hipLaunchKernelGGL(k_refs, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, d_b, nb, d_uv, nrefs, nk, n, kfrom,
                     box, (float2*)d_out);
  const hipError_t e = hipDeviceSynchronize();
  (void)hipFree(d_b);
  (void)hipFree(d_uv);
  return e == hipSuccess ? 0 : 3;
}

// prm: [np] PtclParam (13 x 4 bytes each)
int synth_ptcls(int nb, const float* blobs, int nrefs, const float* uvs, int np, const void* prm, int nk, int n,
                int kfrom, int box, int with_ctf, void* d_out) {
  Blob* d_b = nullptr;
  float* d_uv = nullptr;
  PtclParam* d_p = nullptr;
  if (hipMalloc(&d_b, sizeof(Blob) * nb) != hipSuccess) return 3;
  if (hipMalloc(&d_uv, sizeof(float) * 6 * nrefs) != hipSuccess) return 3;
  if (hipMalloc(&d_p, sizeof(PtclParam) * np) != hipSuccess) return 3;
  if (hipMemcpy(d_b, blobs, sizeof(Blob) * nb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_uv, uvs, sizeof(float) * 6 * nrefs, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_p, prm, sizeof(PtclParam) * np, hipMemcpyHostToDevice) != hipSuccess)
    return 3;
  const size_t tot = (size_t)np * nk * n;
  hipLaunchKernelGGL(k_ptcls, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, d_b, nb, d_uv, d_p, np, nk, n,
                     kfrom, box, with_ctf, (float2*)d_out);
  const hipError_t e = hipDeviceSynchronize();
  (void)hipFree(d_b);
  (void)hipFree(d_uv);
  (void)hipFree(d_p);
  return e == hipSuccess ? 0 : 3;
}

int synth_param_size() { return (int)sizeof(PtclParam); }

void* synth_malloc(size_t bytes) {
  void* p = nullptr;
  return hipMalloc(&p, bytes) == hipSuccess ? p : nullptr;
}

int synth_free(void* p) { return hipFree(p) == hipSuccess ? 0 : 3; }

int synth_d2h(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : 3;
}

int synth_h2d(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : 3;
}

// HIP events on a stream, for timing from hosts that have no other binding to the runtime
void* synth_event_create() {
  hipEvent_t e = nullptr;
  return hipEventCreate(&e) == hipSuccess ? (void*)e : nullptr;
}
int synth_event_record(void* e, void* stream) { return hipEventRecord((hipEvent_t)e, (hipStream_t)stream) == hipSuccess ? 0 : 3; }
float synth_event_elapsed_ms(void* a, void* b) {
  float ms = -1.f;
  if (hipEventSynchronize((hipEvent_t)b) != hipSuccess) return ms;
  if (hipEventElapsedTime(&ms, (hipEvent_t)a, (hipEvent_t)b) != hipSuccess) return -1.f;
  return ms;
}
int synth_device_sync() { return hipDeviceSynchronize() == hipSuccess ? 0 : 3; }

}  // extern "C"
