// LDS microbenchmark: the cost of ds_add_f64 wavefront instructions on gfx950, for the
// gridder's access pattern and for a conflict-free one.  Measurement tooling (the peak the
// gridder's atomic throughput is compared with), built into its own library, libarlmi_probe.so,
// and not part of the imaging library.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "arlmi_probe.h"

namespace arlmi_probe {
namespace {

thread_local char g_err[256] = "";

#define PROBE_HIP(call)                                                                    \
  do {                                                                                     \
    const hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                                \
      snprintf(g_err, sizeof(g_err), "%s: %s", #call, hipGetErrorString(e_));              \
      return -7;                                                                           \
    }                                                                                      \
  } while (0)

constexpr int kProbeThreads = 512;
constexpr int kProbeTile = 4608;  // 64 rows x 72 doubles, the gridder's accumulator tile

__device__ inline unsigned lcg(unsigned& s) {
  s = s * 1664525u + 1013904223u;
  return s >> 8;
}

// pattern 0: each wavefront adds 1 to an 8 x 8 footprint (row stride 72) at a random origin of
// the tile, real and imaginary planes, as the gridder does; pattern 1: lane l adds to address
// 64 * k + l (64 distinct consecutive doubles: no bank conflict); pattern 2: pattern 0 with
// 64-bit integer adds; pattern 3: pattern 0 with float32 adds (ds_add_f32) on a float tile of the
// same shape; pattern 4: pattern 1 with float32 adds.  All addresses < 2 * 4608.
template <int PATTERN>
__global__ __launch_bounds__(kProbeThreads) void k_probe_lds_add(double* out, int iters, unsigned seed) {
  extern __shared__ double t[];
  for (int i = threadIdx.x; i < 2 * kProbeTile; i += blockDim.x) t[i] = 0.0;
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ta = lane >> 3, tb = lane & 7;
  unsigned s = seed + w * 977u + blockIdx.x * 131u;
  for (int it = 0; it < iters; ++it) {
    const unsigned r = __builtin_amdgcn_readfirstlane(lcg(s));
    int o;
    if (PATTERN != 1 && PATTERN != 4) {
      const int ox = (int)(r % 57u), oy = (int)((r >> 8) % 57u);
      o = (oy + ta) * 72 + ox + tb;
    } else {
      o = (int)(r % 72u) * 64 + lane;
    }
    if (PATTERN == 3 || PATTERN == 4) {  // float32 adds
      float* f = reinterpret_cast<float*>(t);
      __hip_atomic_fetch_add(&f[o], 1.0f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(&f[kProbeTile + o], 2.0f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else if (PATTERN == 2) {  // the gridder pattern with 64-bit integer adds
      unsigned long long* u = reinterpret_cast<unsigned long long*>(t);
      __hip_atomic_fetch_add(&u[o], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(&u[kProbeTile + o], 2ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
      __hip_atomic_fetch_add(&t[o], 1.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(&t[kProbeTile + o], 2.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = t[0] + t[kProbeTile + 100];
}

}  // namespace
}  // namespace arlmi_probe

using namespace arlmi_probe;

extern "C" {

const char* arlmi_probe_last_error(void) { return g_err; }

int arlmi_probe_lds_add(int32_t pattern, int32_t iters, double* cycles_per_instr, float* ms_out) {
  if (pattern < 0 || pattern > 4 || iters <= 0 || !cycles_per_instr) {
    snprintf(g_err, sizeof(g_err), "arlmi_probe_lds_add: bad argument");
    return -1;
  }
  int dev = 0, ncu = 0, clk_khz = 0;
  PROBE_HIP(hipGetDevice(&dev));
  PROBE_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  PROBE_HIP(hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeClockRate, dev));
  const int blocks = 2 * ncu;  // two work groups of 8 wavefronts per CU
  double* out = nullptr;
  PROBE_HIP(hipMalloc(&out, sizeof(double) * blocks));
  hipEvent_t a, b;
  PROBE_HIP(hipEventCreate(&a));
  PROBE_HIP(hipEventCreate(&b));
  const size_t sh = sizeof(double) * 2 * kProbeTile;
  float ms = 0.f;
  for (int rep = 0; rep < 2; ++rep) {  // the second launch is timed
    PROBE_HIP(hipEventRecord(a, nullptr));
    if (pattern == 0)
      hipLaunchKernelGGL(k_probe_lds_add<0>, dim3(blocks), dim3(kProbeThreads), sh, nullptr, out, iters, 1u);
    else if (pattern == 1)
      hipLaunchKernelGGL(k_probe_lds_add<1>, dim3(blocks), dim3(kProbeThreads), sh, nullptr, out, iters, 1u);
    else if (pattern == 2)
      hipLaunchKernelGGL(k_probe_lds_add<2>, dim3(blocks), dim3(kProbeThreads), sh, nullptr, out, iters, 1u);
    else if (pattern == 3)
      hipLaunchKernelGGL(k_probe_lds_add<3>, dim3(blocks), dim3(kProbeThreads), sh, nullptr, out, iters, 1u);
    else
      hipLaunchKernelGGL(k_probe_lds_add<4>, dim3(blocks), dim3(kProbeThreads), sh, nullptr, out, iters, 1u);
    PROBE_HIP(hipGetLastError());
    PROBE_HIP(hipEventRecord(b, nullptr));
    PROBE_HIP(hipEventSynchronize(b));
    PROBE_HIP(hipEventElapsedTime(&ms, a, b));
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(out);
  // wavefront instructions per CU: 2 groups x 8 wavefronts x iters x 2 adds
  const double instr = 2.0 * 8.0 * iters * 2.0;
  const double cycles = (double)ms * 1e-3 * (double)clk_khz * 1e3;
  *cycles_per_instr = cycles / instr;
  if (ms_out) *ms_out = ms;
  return 0;
}

}  // extern "C"
