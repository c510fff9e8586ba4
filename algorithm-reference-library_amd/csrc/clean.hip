// Hogbom CLEAN minor cycle (processing_library/arrays/cleaners.py:12-70) on the GPU, float64.
//
// ARL's loop is sequential: find the peak of |res * window|, add gain * peak / max(psf) to the
// component image there, subtract that multiple of the PSF centred on the peak from the
// residual, stop when |res[peak]| < 0.9 * absolutethresh.  Each iteration is two launches.
// k_clean_scan subtracts the previous peak's PSF multiple fused with the search for the next
// peak: every block owns a band of image rows, subtracts the PSF where the band meets the overlap
// region (overlapIndices, cleaners.py:134-154) and, when its band changed, recomputes its largest
// |res * window| (ties to the lowest flat index, as numpy's argmax); bands the PSF did not touch
// keep their maximum from an earlier launch.  k_clean_pick (one block) reduces the band maxima
// in band order, books the component and evaluates ARL's stop test on the peak pixel after its
// subtraction, which the next scan applies.  The launch boundary orders the two, so no
// device-scope fence (an L2 write-back on this GPU) is needed.
//
// Every value is computed with ARL's operations in ARL's order and without contraction
// ((res * gain) / pmax, res - psf * mval), so residual and components equal numpy's bit for bit.
#include "arlmi_internal.h"

#include <map>
#include <mutex>
#include <vector>

namespace arlmi {
namespace {

constexpr int kCleanThreads = 256;
constexpr int kPickThreads = 1024;
constexpr int kCleanBatch = 64;  // launches between two host reads of the end flag

struct CleanCtl {
  int32_t pending;  // a peak whose PSF multiple is still to be subtracted
  int32_t px, py;   // that peak (row, column)
  int32_t stop;     // ARL's stop test held for it: end after its subtraction
  int32_t done;     // the minor cycle has ended
  int32_t iters;    // peaks found (ARL's aiter)
  int32_t pad[2];
  double mval;      // the pending peak's component value
  int64_t scanned;  // pixels of the bands rescanned for maxima (measurement)
  int64_t area;     // pixels the PSF subtractions cover (measurement)
};

struct CleanArgs {
  double* res;
  double* comps;
  const double* psf;
  const double* window;  // [nx][ny] or null
  int nx, ny;            // image rows, columns
  int pnx, pny;          // PSF rows, columns
  double gain, pmax, thr;  // thr = 0.9 * absolutethresh
  int niter;
  int rows;              // image rows per block
  CleanCtl* ctl;
  double* pval;          // per-block largest |res * window|
  int64_t* pidx;         // and its flat index
};

// The residual rows [xa, xb) and columns [ya, yb) the PSF centred on (mx, my) overlaps
// (overlapIndices: res_lower = max(0, peak - psf//2), res_upper = min(n, peak + psf//2)).
struct Overlap {
  int xa, xb, ya, yb;
};
__device__ __forceinline__ Overlap overlap(const CleanArgs& a, int mx, int my) {
  Overlap o;
  o.xa = max(0, mx - a.pnx / 2);
  o.xb = min(a.nx, mx + a.pnx / 2);
  o.ya = max(0, my - a.pny / 2);
  o.yb = min(a.ny, my + a.pny / 2);
  return o;
}

// res - psf * mval, a multiply then a subtraction as numpy evaluates res[...] -= psf[...] * mval.
__device__ __forceinline__ double sub_psf(double r, double p, double m) {
#pragma clang fp contract(off)
  const double t = p * m;
  return r - t;
}

__device__ __forceinline__ void better(double& v, int64_t& i, double v2, int64_t i2) {
  if (v2 > v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}

// Block-wide (value, index) maximum into lane 0 of wave 0.
__device__ inline void block_best(double& v, int64_t& i, double* sv, int64_t* si) {
  for (int off = 32; off > 0; off >>= 1) {
    const double v2 = __shfl_down(v, off);
    const int64_t i2 = __shfl_down(i, off);
    better(v, i, v2, i2);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sv[wave] = v;
    si[wave] = i;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) better(v, i, sv[w], si[w]);
  }
}

// Subtract the pending peak's PSF multiple and rescan the bands it changed (every band on the
// first launch).  Reads the control word only; k_clean_pick, the next launch, writes it.
__global__ void __launch_bounds__(kCleanThreads) k_clean_scan(CleanArgs a) {
  __shared__ double sv[kCleanThreads / 64];
  __shared__ int64_t si[kCleanThreads / 64];
  const CleanCtl c = *a.ctl;
  if (c.done) return;
  const int r0 = blockIdx.x * a.rows, r1 = min(a.nx, r0 + a.rows);
  Overlap o{0, 0, 0, 0};
  bool touched = false;
  if (c.pending) {
    o = overlap(a, c.px, c.py);
    touched = r0 < o.xb && o.xa < r1 && o.ya < o.yb;
  }
  const bool need_peak = !c.stop && c.iters < a.niter;
  const bool rescan = need_peak && (touched || c.iters == 0);
  const int hx = a.pnx / 2, hy = a.pny / 2;
  if (rescan) {
    double bv = -1.0;
    int64_t bi = INT64_MAX;
    for (int x = r0; x < r1; ++x) {
      const bool rowin = touched && x >= o.xa && x < o.xb;
      // PSF row hx + x - px, shifted so that column y reads hy + y - py (in range inside the overlap)
      const double* prow = rowin ? a.psf + (int64_t)(hx + x - c.px) * a.pny + (hy - c.py) : a.psf;
      double* rrow = a.res + (int64_t)x * a.ny;
      const double* wrow = a.window ? a.window + (int64_t)x * a.ny : nullptr;
      for (int y = threadIdx.x; y < a.ny; y += blockDim.x) {
        double r = rrow[y];
        if (rowin && y >= o.ya && y < o.yb) {
          r = sub_psf(r, prow[y], c.mval);
          rrow[y] = r;
        }
        const double v = wrow ? fabs(r * wrow[y]) : fabs(r);
        better(bv, bi, v, (int64_t)x * a.ny + y);
      }
    }
    block_best(bv, bi, sv, si);
    if (threadIdx.x == 0) {
      a.pval[blockIdx.x] = bv;
      a.pidx[blockIdx.x] = bi;
    }
  } else if (touched) {
    for (int x = max(r0, o.xa); x < min(r1, o.xb); ++x) {
      const double* prow = a.psf + (int64_t)(hx + x - c.px) * a.pny + (hy - c.py);
      double* rrow = a.res + (int64_t)x * a.ny;
      for (int y = o.ya + threadIdx.x; y < o.yb; y += blockDim.x) rrow[y] = sub_psf(rrow[y], prow[y], c.mval);
    }
  }
}

// One block: the bands' maxima in band order (ties to the lower flat index) give the next peak;
// book its component, decide ARL's stop test on it, or end the cycle.
__global__ void __launch_bounds__(kPickThreads) k_clean_pick(CleanArgs a, int nbands) {
  __shared__ double sv[kPickThreads / 64];
  __shared__ int64_t si[kPickThreads / 64];
  const CleanCtl c = *a.ctl;
  if (c.done) return;
  CleanCtl nc = c;
  if (c.stop || c.iters >= a.niter) {  // the last subtraction is done
    nc.pending = 0;
    nc.done = 1;
    if (threadIdx.x == 0) *a.ctl = nc;
    return;
  }
  double bv = -1.0;
  int64_t bi = INT64_MAX;
  for (int b = threadIdx.x; b < nbands; b += blockDim.x) better(bv, bi, a.pval[b], a.pidx[b]);
  block_best(bv, bi, sv, si);
  if (threadIdx.x != 0) return;
  if (bi == INT64_MAX) {  // no comparable pixel (all NaN): the cycle ends
    nc.pending = 0;
    nc.done = 1;
    *a.ctl = nc;
    return;
  }
  const int hx = a.pnx / 2, hy = a.pny / 2;
  const int mx = (int)(bi / a.ny), my = (int)(bi % a.ny);
  const double rv = a.res[bi];
  double mval;
  {
#pragma clang fp contract(off)
    mval = rv * a.gain / a.pmax;  // (res * gain) / pmax
  }
  a.comps[bi] = a.comps[bi] + mval;
  const Overlap po = overlap(a, mx, my);
  const bool in = mx >= po.xa && mx < po.xb && my >= po.ya && my < po.yb;
  const double after = in ? sub_psf(rv, a.psf[(int64_t)hx * a.pny + hy], mval) : rv;
  nc.pending = 1;
  nc.px = mx;
  nc.py = my;
  nc.mval = mval;
  nc.iters = c.iters + 1;
  nc.stop = fabs(after) < a.thr ? 1 : 0;
  if (c.iters == 0) nc.scanned += (int64_t)a.nx * a.ny;  // the first scan read every band
  if (po.xa < po.xb && po.ya < po.yb) {
    nc.area += (int64_t)(po.xb - po.xa) * (po.yb - po.ya);
    if (nc.iters < a.niter && !nc.stop) {  // the next scan rereads the bands it touches
      const int b0 = po.xa / a.rows, b1 = (po.xb - 1) / a.rows;
      nc.scanned += (int64_t)(min(a.nx, (b1 + 1) * a.rows) - b0 * a.rows) * a.ny;
    }
  }
  *a.ctl = nc;
}

// max |x| over x (part 0) and max y over y (part 1): per-block partials, then one block.
__global__ void __launch_bounds__(kCleanThreads)
    k_clean_maxes(const double* __restrict__ x, int64_t nxe, const double* __restrict__ y, int64_t nye,
                  double* __restrict__ part) {
  __shared__ double s[2][kCleanThreads];
  double ax = 0.0, my = -INFINITY;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nxe; i += stride) ax = fmax(ax, fabs(x[i]));
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nye; i += stride) my = fmax(my, y[i]);
  s[0][threadIdx.x] = ax;
  s[1][threadIdx.x] = my;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s[0][threadIdx.x] = fmax(s[0][threadIdx.x], s[0][threadIdx.x + h]);
      s[1][threadIdx.x] = fmax(s[1][threadIdx.x], s[1][threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s[0][0];
    part[2 * blockIdx.x + 1] = s[1][0];
  }
}

__global__ void k_clean_init(double* comps, int64_t n, CleanCtl* ctl) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) comps[i] = 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CleanCtl z{};
    *ctl = z;
  }
}

// Per-device scratch: the control word, the band maxima and the partial maxima (grow only).
struct CleanScratch {
  void* mem = nullptr;
  size_t bytes = 0;
};
thread_local int64_t g_clean_scanned = 0, g_clean_area = 0;
std::mutex g_clean_mutex;
std::map<int, CleanScratch> g_clean;

void* clean_scratch(size_t bytes) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lock(g_clean_mutex);
  CleanScratch& s = g_clean[dev];
  if (s.bytes < bytes) {
    if (s.mem) (void)hipFree(s.mem);
    s.mem = nullptr;
    s.bytes = 0;
    if (hipMalloc(&s.mem, bytes) != hipSuccess) return nullptr;
    s.bytes = bytes;
  }
  return s.mem;
}

// Scratch layout of a minor cycle on an nx-row image: the control word, the partial maxima of
// k_clean_maxes and the (value, index) maximum of every band.  About 2048 bands of whole rows:
// enough blocks to fill the GPU, few maxima to reduce, so the layout is at most 49,408 bytes.
constexpr int kMaxBlocks = 1024;  // blocks of the max pass
struct CleanLayout {
  int rows, nb;
  size_t off_part, off_pval, off_pidx, bytes;
};
CleanLayout clean_layout(int nx) {
  CleanLayout l;
  l.rows = std::max(1, (nx + 2047) / 2048);
  l.nb = (nx + l.rows - 1) / l.rows;
  l.off_part = 256;
  l.off_pval = l.off_part + sizeof(double) * 2 * kMaxBlocks;
  l.off_pidx = l.off_pval + sizeof(double) * l.nb;
  l.bytes = l.off_pidx + sizeof(int64_t) * l.nb;
  return l;
}

int check_hogbom_args(const char* who, const double* d_res, const double* d_comps, const double* d_psf, int32_t psf_nx,
                      int32_t psf_ny, int32_t nx, int32_t ny, double gain, int32_t niter) {
  if (!d_res || !d_comps || !d_psf || nx < 1 || ny < 1 || psf_nx < 1 || psf_ny < 1) {
    set_error("%s: bad argument", who);
    return ARLMI_E_ARG;
  }
  if (!(gain > 0.0 && gain < 2.0) || niter < 1) {  // cleaners.py:33-34
    set_error("%s: need 0 < gain < 2 and niter > 0 (gain %g, niter %d)", who, gain, niter);
    return ARLMI_E_ARG;
  }
  return ARLMI_OK;
}

// The minor cycle on `mem` (device memory of at least clean_layout(nx).bytes); allocates nothing.
int hogbom_run(unsigned char* mem, double* d_res, double* d_comps, const double* d_psf, int32_t psf_nx, int32_t psf_ny,
               const double* d_window, int32_t nx, int32_t ny, double gain, double thresh, int32_t niter,
               double fracthresh, int32_t* iterations, hipStream_t s) {
  const int64_t npix = (int64_t)nx * ny, npsf = (int64_t)psf_nx * psf_ny;
  const CleanLayout lay = clean_layout(nx);
  const int rows = lay.rows, nb = lay.nb, nmax = kMaxBlocks;
  CleanCtl* ctl = reinterpret_cast<CleanCtl*>(mem);
  double* part = reinterpret_cast<double*>(mem + lay.off_part);
  ARLMI_LAUNCH("k_clean_init", k_clean_init, dim3(1024), dim3(kCleanThreads), 0, s, d_comps, npix, ctl);
  ARLMI_LAUNCH("k_clean_maxes", k_clean_maxes, dim3(nmax), dim3(kCleanThreads), 0, s, (const double*)d_res, npix,
               d_psf, npsf, part);
  ARLMI_LAUNCH_CHECK();
  double hp[2 * kMaxBlocks];
  if (hipMemcpyAsync(hp, part, sizeof(double) * 2 * nmax, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error("arlmi_hogbom: %s", hipGetErrorString(hipGetLastError()));
    return ARLMI_E_HIP;
  }
  double maxabs = 0.0, pmax = -INFINITY;
  for (int b = 0; b < nmax; ++b) {
    maxabs = std::max(maxabs, hp[2 * b]);
    pmax = std::max(pmax, hp[2 * b + 1]);
  }
  if (!(pmax > 0.0)) {  // cleaners.py:43
    set_error("arlmi_hogbom: the PSF maximum must be positive (%g)", pmax);
    return ARLMI_E_ARG;
  }
  const double absolutethresh = std::max(thresh, fracthresh * maxabs);  // cleaners.py:36
  CleanArgs a;
  a.res = d_res;
  a.comps = d_comps;
  a.psf = d_psf;
  a.window = d_window;
  a.nx = nx;
  a.ny = ny;
  a.pnx = psf_nx;
  a.pny = psf_ny;
  a.gain = gain;
  a.pmax = pmax;
  a.thr = 0.9 * absolutethresh;
  a.niter = niter;
  a.rows = rows;
  a.ctl = ctl;
  a.pval = reinterpret_cast<double*>(mem + lay.off_pval);
  a.pidx = reinterpret_cast<int64_t*>(mem + lay.off_pidx);
  // niter peaks take niter + 1 launches (the last one only subtracts); launches after the end
  // return at once
  CleanCtl hc{};
  for (int64_t launched = 0; launched < (int64_t)niter + 1;) {
    const int64_t batch = std::min<int64_t>(kCleanBatch, (int64_t)niter + 1 - launched);
    for (int64_t k = 0; k < batch; ++k) {
      ARLMI_LAUNCH("k_clean_scan", k_clean_scan, dim3((unsigned)nb), dim3(kCleanThreads), 0, s, a);
      ARLMI_LAUNCH("k_clean_pick", k_clean_pick, dim3(1), dim3(kPickThreads), 0, s, a, nb);
    }
    ARLMI_LAUNCH_CHECK();
    launched += batch;
    if (hipMemcpyAsync(&hc, ctl, sizeof(CleanCtl), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("arlmi_hogbom: %s", hipGetErrorString(hipGetLastError()));
      return ARLMI_E_HIP;
    }
    if (hc.done) break;
  }
  if (!hc.done) {
    set_error("arlmi_hogbom: the minor cycle did not end after %d launches", niter + 1);
    return ARLMI_E_HIP;
  }
  if (iterations) *iterations = hc.iters;
  g_clean_scanned = hc.scanned;
  g_clean_area = hc.area;
  return ARLMI_OK;
}

// Scratch of arlmi_hogbom_noalloc: a file-scope device array, one copy per GPU, held until the
// process ends.  Host code never passes `scratch` itself (a host-side stand-in): it passes the
// device address hipGetSymbolAddress gives, looked up once per device.
constexpr size_t kNoAllocScratchBytes = 64 * 1024;
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mutex;
void* g_noalloc_ptr[kMaxDevices] = {};

}  // namespace
}  // namespace arlmi

static __device__ __align__(256) unsigned char scratch[arlmi::kNoAllocScratchBytes];

namespace arlmi {
namespace {

unsigned char* noalloc_scratch() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("arlmi_hogbom_noalloc: no usable current device (%d)", dev);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("arlmi_hogbom_noalloc: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
      return nullptr;
    }
    g_noalloc_ptr[dev] = p;
  }
  return static_cast<unsigned char*>(g_noalloc_ptr[dev]);
}

}  // namespace
}  // namespace arlmi

using namespace arlmi;

extern "C" {

int arlmi_hogbom(double* d_res, double* d_comps, const double* d_psf, int32_t psf_nx, int32_t psf_ny,
                 const double* d_window, int32_t nx, int32_t ny, double gain, double thresh, int32_t niter,
                 double fracthresh, int32_t* iterations, void* stream) {
  if (int rc = check_hogbom_args("arlmi_hogbom", d_res, d_comps, d_psf, psf_nx, psf_ny, nx, ny, gain, niter)) return rc;
  const size_t bytes = clean_layout(nx).bytes;
  unsigned char* mem = static_cast<unsigned char*>(clean_scratch(bytes));
  if (!mem) {
    set_error("arlmi_hogbom: scratch allocation of %zu bytes failed", bytes);
    return ARLMI_E_HIP;
  }
  return hogbom_run(mem, d_res, d_comps, d_psf, psf_nx, psf_ny, d_window, nx, ny, gain, thresh, niter, fracthresh,
                    iterations, as_stream(stream));
}

int arlmi_hogbom_noalloc(double* d_res, double* d_comps, const double* d_psf, int32_t psf_nx, int32_t psf_ny,
                         const double* d_window, int32_t nx, int32_t ny, double gain, double thresh, int32_t niter,
                         double fracthresh, int32_t* iterations, void* stream) {
  if (int rc = check_hogbom_args("arlmi_hogbom_noalloc", d_res, d_comps, d_psf, psf_nx, psf_ny, nx, ny, gain, niter))
    return rc;
  const size_t bytes = clean_layout(nx).bytes;
  if (bytes > kNoAllocScratchBytes) {
    set_error("arlmi_hogbom_noalloc: %zu bytes of scratch needed, %zu available", bytes, kNoAllocScratchBytes);
    return ARLMI_E_UNSUPPORTED;
  }
  unsigned char* mem = noalloc_scratch();
  if (!mem) return ARLMI_E_HIP;
  return hogbom_run(mem, d_res, d_comps, d_psf, psf_nx, psf_ny, d_window, nx, ny, gain, thresh, niter, fracthresh,
                    iterations, as_stream(stream));
}

int arlmi_hogbom_traffic(int64_t* scanned_pixels, int64_t* subtracted_pixels) {
  if (scanned_pixels) *scanned_pixels = g_clean_scanned;
  if (subtracted_pixels) *subtracted_pixels = g_clean_area;
  return ARLMI_OK;
}

}  // extern "C"
