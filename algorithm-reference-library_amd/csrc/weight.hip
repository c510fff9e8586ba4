// Visibility weighting: gridded weight density and uniform reweighting
// (processing_components/griddata/gridding.py:161-230, imaging/weighting.py:19-40).
//
// Every row's imaging weight is added at its nearest grid pixel (pu, pv) of its image channel
// and grid w plane, then divided by the density at that pixel.  The density grid is float64 and
// is summed with float64 global adds: rows spread over N^2 pixels, so adds rarely meet on one
// address, and this runs once per observation, not per imaging step.
#include "arlmi_internal.h"

#include <map>
#include <mutex>

namespace arlmi {
namespace {

// Nearest pixel, grid w plane and image channel of a row as convolution_mapping computes them
// (gridding.py:40-85), with its asserts: 0 <= pu < N and 0 <= pv < N (no kernel footprint, which
// weight gridding never slices), grid and kernel w planes within the kernel's nz, channel within
// range.  Returns the flat density index, or -1 with err bits set.
__device__ inline int64_t weight_pixel(const DevGeom& g, const double* uvw, const double* freq, int64_t row,
                                       uint32_t* err) {
  const double u = uvw[3 * row], v = uvw[3 * row + 1], w = uvw[3 * row + 2];
  const double pud = rint(world2pix0(u, g.du, g.crpix_uv));
  const double pvd = rint(world2pix0(v, g.dv, g.crpix_uv));
  uint32_t e = 0;
  if (!(pud >= 0.0 && pud < (double)g.n && pvd >= 0.0 && pvd < (double)g.n)) e |= ERRB_UV_OFF_GRID;
  const int pwg = map_wplane(w, g.wstep_grid, g.crpix_wg, g.nz, &e);
  (void)map_wplane(w, g.zstep, g.crpix_z, g.nz, &e);
  int c = 0;
  if (freq) {
    const double pf = rint(((freq[row] - g.freq0) / g.dfreq + 1.0) - 1.0);
    if (!(pf >= 0.0 && pf < (double)g.nchan)) e |= ERRB_FREQ_MAP;
    else c = (int)pf;
  }
  if (pwg >= g.nplanes) e |= ERRB_W_RANGE;
  if (e) {
    *err |= e;
    return -1;
  }
  return (((int64_t)c * g.nplanes + pwg) * g.n + (int64_t)pvd) * g.n + (int64_t)pud;
}

constexpr int kThreads = 256;

// Work group b adds into copy b % ncopy of the density (copies nd doubles apart): work groups
// are dispatched to the eight XCDs in turn, so with eight copies the adds of one XCD meet only
// its own copy.
template <typename R>
__global__ void __launch_bounds__(kThreads)
    k_weight_grid(DevGeom g, const double* __restrict__ uvw, const double* __restrict__ freq, int64_t n,
                  const R* __restrict__ wt, double* __restrict__ density_copies, int ncopy, int64_t nd,
                  double* __restrict__ sumwt, uint32_t* __restrict__ err) {
  double* __restrict__ density = density_copies + (int64_t)(blockIdx.x % (unsigned)ncopy) * nd;
  __shared__ double lsum[256];
  __shared__ uint32_t lerr;
  for (int i = threadIdx.x; i < g.nchan; i += blockDim.x) lsum[i] = 0.0;
  if (threadIdx.x == 0) lerr = 0u;
  __syncthreads();
  uint32_t e = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < n; row += stride) {
    const int64_t k = weight_pixel(g, uvw, freq, row, &e);
    if (k < 0) continue;
    const double w = (double)wt[row];
    unsafeAtomicAdd(&density[k], w);
    const int c = (int)(k / ((int64_t)g.nplanes * g.n * g.n));
    __hip_atomic_fetch_add(&lsum[c], w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  if (e) atomicOr(&lerr, e);
  __syncthreads();
  for (int i = threadIdx.x; i < g.nchan; i += blockDim.x)
    if (lsum[i] != 0.0) unsafeAtomicAdd(&sumwt[i], lsum[i]);
  if (threadIdx.x == 0 && lerr) atomicOr(err, lerr);
}

// apply = 0 only checks the mapping (ARL asserts before it changes any weight); apply = 1 divides.
template <typename R>
__global__ void __launch_bounds__(kThreads)
    k_weight_apply(DevGeom g, const double* __restrict__ uvw, const double* __restrict__ freq, int64_t n,
                   const double* __restrict__ density, R* __restrict__ wt, int apply, uint32_t* __restrict__ err) {
  uint32_t e = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < n; row += stride) {
    const int64_t k = weight_pixel(g, uvw, freq, row, &e);
    if (k < 0 || !apply) continue;
    const double d = density[k];
    if (d != 0.0) wt[row] = (R)((double)wt[row] / d);
  }
  if (e) atomicOr(err, e);
}

__global__ void k_fill_f64(double* __restrict__ p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.0;
}

__global__ void k_fill_u32(uint32_t* p) { *p = 0u; }

// density[i] = sum of the ncopy copies at i, in copy order.
__global__ void __launch_bounds__(kThreads)
    k_sum_copies(const double* __restrict__ copies, int ncopy, int64_t nd, double* __restrict__ density) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nd; i += stride) {
    double t = 0.0;
    for (int c = 0; c < ncopy; ++c) t += copies[(int64_t)c * nd + i];
    density[i] = t;
  }
}

// Per-device density copies (grow only, kept until exit).
constexpr int kCopies = 8;
constexpr size_t kCopiesMaxBytes = (size_t)2 << 30;
std::mutex g_copies_mutex;
std::map<int, std::pair<void*, size_t>> g_copies;

double* density_copies(size_t bytes) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lock(g_copies_mutex);
  auto& c = g_copies[dev];
  if (c.second < bytes) {
    if (c.first) (void)hipFree(c.first);
    c.first = nullptr;
    c.second = 0;
    if (hipMalloc(&c.first, bytes) != hipSuccess) return nullptr;
    c.second = bytes;
  }
  return static_cast<double*>(c.first);
}

// One error word per device, allocated on first use.
std::mutex g_err_mutex;
std::map<int, uint32_t*> g_err;

uint32_t* err_word() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lock(g_err_mutex);
  auto it = g_err.find(dev);
  if (it != g_err.end()) return it->second;
  uint32_t* p = nullptr;
  if (hipMalloc(&p, sizeof(uint32_t)) != hipSuccess) return nullptr;
  g_err[dev] = p;
  return p;
}

unsigned blocks_for(int64_t n) {
  const int64_t b = (n + kThreads - 1) / kThreads;
  return (unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

int finish(uint32_t* d_err, hipStream_t s) {
  uint32_t bits = 0;
  ARLMI_HIP(hipMemcpyAsync(&bits, d_err, sizeof(bits), hipMemcpyDeviceToHost, s));
  ARLMI_HIP(hipStreamSynchronize(s));
  if (bits) {
    const int rc = first_error_code(bits);
    set_error("a row maps outside the grid, w or frequency axis (error bits 0x%x)", bits);
    return rc;
  }
  return ARLMI_OK;
}

}  // namespace
}  // namespace arlmi

using namespace arlmi;

extern "C" {

int arlmi_grid_weights(const arlmi_geom* gin, const double* d_uvw, const double* d_freq, int64_t nrows,
                       const void* d_weight, double* d_density, double* d_sumwt, int precision, void* stream) {
  DevGeom g;
  if (int rc = checked_geom(gin, &g)) return rc;
  if ((precision != 32 && precision != 64) || nrows < 0 || !d_density || !d_sumwt ||
      (nrows > 0 && (!d_uvw || !d_weight))) {
    set_error("arlmi_grid_weights: bad argument");
    return ARLMI_E_ARG;
  }
  hipStream_t s = as_stream(stream);
  uint32_t* d_err = err_word();
  if (!d_err) {
    set_error("arlmi_grid_weights: no device error word");
    return ARLMI_E_HIP;
  }
  const int64_t nd = (int64_t)g.nchan * g.nplanes * g.n * g.n;
  // one density copy per XCD when they fit, else the caller's grid alone
  double* copies = nullptr;
  int ncopy = 1;
  if (nrows > 0 && (size_t)nd * sizeof(double) * kCopies <= kCopiesMaxBytes) {
    copies = density_copies((size_t)nd * sizeof(double) * kCopies);
    if (copies) ncopy = kCopies;
  }
  double* target = copies ? copies : d_density;
  ARLMI_LAUNCH("k_fill_f64", k_fill_f64, dim3(blocks_for(nd * ncopy)), dim3(kThreads), 0, s, target, nd * ncopy);
  ARLMI_LAUNCH("k_fill_f64", k_fill_f64, dim3(1), dim3(kThreads), 0, s, d_sumwt, (int64_t)g.nchan);
  ARLMI_LAUNCH("k_fill_u32", k_fill_u32, dim3(1), dim3(1), 0, s, d_err);
  if (nrows > 0) {
    if (precision == 32)
      ARLMI_LAUNCH("k_weight_grid", k_weight_grid<float>, dim3(blocks_for(nrows)), dim3(kThreads), 0, s, g, d_uvw,
                   d_freq, nrows, (const float*)d_weight, target, ncopy, nd, d_sumwt, d_err);
    else
      ARLMI_LAUNCH("k_weight_grid", k_weight_grid<double>, dim3(blocks_for(nrows)), dim3(kThreads), 0, s, g, d_uvw,
                   d_freq, nrows, (const double*)d_weight, target, ncopy, nd, d_sumwt, d_err);
    if (copies)
      ARLMI_LAUNCH("k_sum_copies", k_sum_copies, dim3(blocks_for(nd)), dim3(kThreads), 0, s, copies, ncopy, nd,
                   d_density);
  }
  ARLMI_LAUNCH_CHECK();
  return finish(d_err, s);
}

int arlmi_reweight(const arlmi_geom* gin, const double* d_uvw, const double* d_freq, int64_t nrows,
                   const double* d_density, void* d_weight, int precision, void* stream) {
  DevGeom g;
  if (int rc = checked_geom(gin, &g)) return rc;
  if ((precision != 32 && precision != 64) || nrows < 0 || (nrows > 0 && (!d_uvw || !d_weight || !d_density))) {
    set_error("arlmi_reweight: bad argument");
    return ARLMI_E_ARG;
  }
  if (nrows == 0) return ARLMI_OK;
  hipStream_t s = as_stream(stream);
  uint32_t* d_err = err_word();
  if (!d_err) {
    set_error("arlmi_reweight: no device error word");
    return ARLMI_E_HIP;
  }
  for (int apply = 0; apply < 2; ++apply) {
    ARLMI_LAUNCH("k_fill_u32", k_fill_u32, dim3(1), dim3(1), 0, s, d_err);
    if (precision == 32)
      ARLMI_LAUNCH("k_weight_apply", k_weight_apply<float>, dim3(blocks_for(nrows)), dim3(kThreads), 0, s, g, d_uvw,
                   d_freq, nrows, d_density, (float*)d_weight, apply, d_err);
    else
      ARLMI_LAUNCH("k_weight_apply", k_weight_apply<double>, dim3(blocks_for(nrows)), dim3(kThreads), 0, s, g, d_uvw,
                   d_freq, nrows, d_density, (double*)d_weight, apply, d_err);
    ARLMI_LAUNCH_CHECK();
    if (int rc = finish(d_err, s)) return rc;
  }
  return ARLMI_OK;
}

}  // extern "C"
