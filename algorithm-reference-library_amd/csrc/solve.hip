// Scalar gain solution (processing_components/calibration/calibration.py:30-101, solve_gaintable
// for npol = 1): point-source equivalent visibilities per gain-table row
// (visibility/operations.py:241-264 divide_visibility, calibration.py:64-78), then D'Addario's
// iterative substitution (processing_library/calibration/solvers.py:61-140) and the solution
// residual (:302-332).  Float64 throughout, as ARL's numpy arrays.
//
// k_solve_x: one thread per (gain row, antenna i, antenna j, channel) sums its element over the
// integrations of that gain row in time order (numpy.sum over axis 0), divides by the summed
// weight, and writes the Hermitian completion the solver makes (x[i, j] = conj(x[j, i]) for i < j,
// zero diagonal).  k_solve_itsubs: one work group per gain row holds the gains in LDS and runs the
// iterations; every antenna sum runs over the other antenna in order.
#include "arlmi_internal.h"

#include <cmath>

namespace arlmi {
namespace {

constexpr int kSolveThreads = 256;
constexpr int kItsubsThreads = 1024;
constexpr int kMaxPairs = 2048;  // antennas x channels of one gain row held in LDS

struct Z {
  double x, y;
};
__device__ __forceinline__ Z zmul(Z a, Z b) {
#pragma clang fp contract(off)
  return Z{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ Z zconj(Z a) { return Z{a.x, -a.y}; }

// numpy's complex division a / b (Smith's method)
__device__ __forceinline__ Z zdiv(Z a, Z b) {
#pragma clang fp contract(off)
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, den = b.x + b.y * r;
    return Z{(a.x + a.y * r) / den, (a.y - a.x * r) / den};
  }
  const double r = b.x / b.y, den = b.x * r + b.y;
  return Z{(a.x * r + a.y) / den, (a.y * r - a.x) / den};
}

template <typename R>
__global__ void __launch_bounds__(kSolveThreads)
    k_solve_x(const cplx<R>* __restrict__ vis, const cplx<R>* __restrict__ model, const R* __restrict__ weight,
              int ntimes, int nants, int nchan, const int32_t* __restrict__ slot, int ngain,
              double2* __restrict__ xout, double* __restrict__ xwout) {
  const int64_t per_row = (int64_t)nants * nants * nchan;
  const int64_t n = per_row * ngain;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += stride) {
    const int s = (int)(e / per_row);
    const int64_t q = e - (int64_t)s * per_row;
    const int ch = (int)(q % nchan);
    const int ij = (int)(q / nchan);
    const int i = ij / nants, j = ij - (ij / nants) * nants;
    if (i == j) {
      xout[e] = make_double2(0.0, 0.0);
      xwout[e] = 0.0;
      continue;
    }
    // the lower-triangle element [a2, a1] the upper one is completed from
    const int a2 = i > j ? i : j, a1 = i > j ? j : i;
    Z sx{0.0, 0.0};
    double sw = 0.0;
    for (int t = 0; t < ntimes; ++t) {
      if (slot[t] != s) continue;
      const int64_t k = (((int64_t)t * nants + a2) * nants + a1) * nchan + ch;
      const cplx<R> v = vis[k];
      const double w = weight ? (double)weight[k] : 1.0;
      Z xv{(double)v.x, (double)v.y};
      double xw = w;
      if (model) {
#pragma clang fp contract(off)
        const cplx<R> m = model[k];
        const double am = hypot((double)m.x, (double)m.y);
        xw = am * am * w;  // numpy.abs(model) ** 2 * weight
        xv = xw > 0.0 ? zdiv(xv, Z{(double)m.x, (double)m.y}) : Z{0.0, 0.0};
      }
      {
#pragma clang fp contract(off)
        sx.x += xv.x * xw;
        sx.y += xv.y * xw;
      }
      sw += xw;
    }
    Z xr{0.0, 0.0};
    if (fabs(sw) > 0.0) xr = Z{sx.x / sw, sx.y / sw};
    if (i < j) xr = zconj(xr);
    xout[e] = make_double2(xr.x, xr.y);
    xwout[e] = sw;
  }
}

struct SolveArgs {
  const double2* x;   // [ngain][nants][nants][nchan], completed
  const double* xw;   // same shape
  const int32_t* active;  // [ngain] integrations in the row
  double2* gain;      // [ngain][nants][nchan], in: start, out: solution
  double* gwt;        // [ngain][nants][nchan]
  double* residual;   // [ngain][nchan]
  int32_t* iters;     // [ngain] iterations run
  int nants, nchan, niter, refant, phase_only, normalise;
  double tol, damping;
};

__device__ inline double block_max(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double m = red[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, red[w]);
  return m;
}

__device__ inline double block_sum_ordered(double v, double* red) {
  // sum of one value per thread in thread order (deterministic)
  __syncthreads();
  red[threadIdx.x] = v;
  __syncthreads();
  double s = 0.0;
  for (int t = 0; t < (int)blockDim.x; ++t) s += red[t];
  return s;
}

__global__ void __launch_bounds__(kItsubsThreads) k_solve_itsubs(SolveArgs a) {
  __shared__ Z g[kMaxPairs];      // current gains
  __shared__ Z gl[kMaxPairs];     // gains of the last iteration
  __shared__ Z tp[kMaxPairs];     // top of the substitution
  __shared__ double bt[kMaxPairs];  // bot of the substitution
  __shared__ Z rot[64];           // per-channel reference-phase rotation
  __shared__ double red[kItsubsThreads];
  const int s = blockIdx.x;
  if (!a.active[s]) return;
  const int na = a.nants, nc = a.nchan, np = na * nc;
  const double2* X = a.x + (int64_t)s * na * na * nc;
  const double* XW = a.xw + (int64_t)s * na * na * nc;
  double2* G = a.gain + (int64_t)s * np;
  for (int p = threadIdx.x; p < np; p += blockDim.x) g[p] = Z{G[p].x, G[p].y};
  __syncthreads();
  int it = 0;
  for (; it < a.niter; ++it) {
    // gain_substitution_scalar (solvers.py:103-127)
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
      gl[p] = g[p];
      const int ant1 = p / nc, ch = p - (p / nc) * nc;
      Z top{0.0, 0.0};
      double bot = 0.0;
      for (int k = 0; k < na; ++k) {
#pragma clang fp contract(off)
        const int64_t e = ((int64_t)k * na + ant1) * nc + ch;
        const double2 xv = X[e];
        const double w = XW[e];
        const Z gk = g[k * nc + ch];
        const Z xx{xv.x * w, xv.y * w};  // xxwt = x * xwt
        const Z t = zmul(gk, xx);
        top.x += t.x;
        top.y += t.y;
        const Z gcg = zmul(gk, zconj(gk));
        bot += gcg.x * w;  // (gcg * xwt).real
      }
      bt[p] = bot;
      tp[p] = top;
    }
    __syncthreads();
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
      const int ant1 = p / nc;
      bool all = true;  // bot.all() > 0.0: every channel's bot non-zero (NaN counts as non-zero)
      for (int c = 0; c < nc; ++c) all = all && (bt[ant1 * nc + c] != 0.0);
      Z ng{0.0, 0.0};
      if (all) ng = Z{tp[p].x / bt[p], tp[p].y / bt[p]};
      // (another channel of this antenna may read bt[p] as 0 here; it then finds `all` false too,
      // which it does anyway since some channel's bot is 0)
      else bt[p] = 0.0;
      if (a.phase_only) {
        const double ab = hypot(ng.x, ng.y);
        if (ab > 0.0) ng = Z{ng.x / ab, ng.y / ab};
      }
      g[p] = ng;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
      const Z r = g[a.refant * nc + c];
      const double ang = atan2(r.y, r.x);  // numpy.angle
      rot[c] = Z{cos(-ang), sin(-ang)};    // numpy.exp(-1j * angle)
    }
    __syncthreads();
    double change = 0.0;
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
#pragma clang fp contract(off)
      Z ng = zmul(g[p], rot[p % nc]);
      const Z last = gl[p];
      ng = Z{(1.0 - a.damping) * ng.x + a.damping * last.x, (1.0 - a.damping) * ng.y + a.damping * last.y};
      g[p] = ng;
      change = fmax(change, hypot(ng.x - last.x, ng.y - last.y));
    }
    change = block_max(change, red);
    if (change < a.tol) {
      ++it;
      break;
    }
  }
  __syncthreads();
  // solution_residual_scalar (solvers.py:302-332): one value summed over every channel
  double rsum = 0.0, wsum = 0.0;
  for (int64_t e = threadIdx.x; e < (int64_t)na * na * nc; e += blockDim.x) {
#pragma clang fp contract(off)
    const int ch = (int)(e % nc);
    const int ij = (int)(e / nc);
    const int i = ij / na, j = ij - (ij / na) * na;
    const double w = XW[e];
    wsum += w;
    if (i == j) continue;
    const Z sm = zmul(zconj(g[i * nc + ch]), g[j * nc + ch]);  // outer(conj(g), g)
    const Z err{X[e].x - sm.x, X[e].y - sm.y};
    rsum += (err.x * w) * err.x + (err.y * w) * err.y;  // (error * xwt * conj(error)).real
  }
  rsum = block_sum_ordered(rsum, red);
  wsum = block_sum_ordered(wsum, red);
  const double res = wsum > 0.0 ? sqrt(rsum / wsum) : 0.0;
  // normalise_gains and not phase_only: divide by the mean |gain| (calibration.py:79-81)
  double gabs = 1.0;
  if (a.normalise && !a.phase_only) {
    double sabs = 0.0;
    for (int p = threadIdx.x; p < np; p += blockDim.x) sabs += hypot(g[p].x, g[p].y);
    gabs = block_sum_ordered(sabs, red) / (double)np;
  }
  for (int p = threadIdx.x; p < np; p += blockDim.x) {
    Z v = g[p];
    if (a.normalise && !a.phase_only) v = Z{v.x / gabs, v.y / gabs};
    G[p] = make_double2(v.x, v.y);
    a.gwt[(int64_t)s * np + p] = bt[p];
  }
  for (int c = threadIdx.x; c < nc; c += blockDim.x) a.residual[(int64_t)s * nc + c] = res;
  if (threadIdx.x == 0) a.iters[s] = it;
}

}  // namespace
}  // namespace arlmi

using namespace arlmi;

extern "C" {

int arlmi_solve_gains(const void* d_vis, const void* d_model, const void* d_weight, int32_t ntimes, int32_t nants,
                      int32_t nchan, const int32_t* d_slot, int32_t ngain, const int32_t* d_active, void* d_gain,
                      double* d_gain_weight, double* d_residual, int32_t* d_iterations, double* d_work,
                      int phase_only, int32_t niter, double tol, int32_t refant, double damping, int normalise,
                      int precision, void* stream) {
  if (precision != 32 && precision != 64) {
    set_error("arlmi_solve_gains: precision must be 32 or 64");
    return ARLMI_E_ARG;
  }
  if (ntimes < 0 || nants < 1 || nchan < 1 || ngain < 0 || niter < 0 || refant < 0 || refant >= nants ||
      (ntimes > 0 && (!d_vis || !d_slot)) || (ngain > 0 && (!d_active || !d_gain || !d_gain_weight || !d_residual ||
                                                             !d_iterations || !d_work))) {
    set_error("arlmi_solve_gains: bad argument");
    return ARLMI_E_ARG;
  }
  if ((int64_t)nants * nchan > kMaxPairs || nchan > 64) {
    set_error("arlmi_solve_gains: nants * nchan = %d exceeds %d (or nchan > 64)", nants * nchan, kMaxPairs);
    return ARLMI_E_UNSUPPORTED;
  }
  if (ngain == 0) return ARLMI_OK;
  hipStream_t s = as_stream(stream);
  const int64_t nx = (int64_t)ngain * nants * nants * nchan;
  double2* x = reinterpret_cast<double2*>(d_work);
  double* xw = d_work + 2 * nx;
  const unsigned gx = (unsigned)std::min<int64_t>((nx + kSolveThreads - 1) / kSolveThreads, 1 << 16);
  if (precision == 32)
    ARLMI_LAUNCH("k_solve_x", k_solve_x<float>, dim3(gx), dim3(kSolveThreads), 0, s, (const float2*)d_vis,
                 (const float2*)d_model, (const float*)d_weight, ntimes, nants, nchan, d_slot, ngain, x, xw);
  else
    ARLMI_LAUNCH("k_solve_x", k_solve_x<double>, dim3(gx), dim3(kSolveThreads), 0, s, (const double2*)d_vis,
                 (const double2*)d_model, (const double*)d_weight, ntimes, nants, nchan, d_slot, ngain, x, xw);
  SolveArgs a;
  a.x = x;
  a.xw = xw;
  a.active = d_active;
  a.gain = reinterpret_cast<double2*>(d_gain);
  a.gwt = d_gain_weight;
  a.residual = d_residual;
  a.iters = d_iterations;
  a.nants = nants;
  a.nchan = nchan;
  a.niter = niter;
  a.refant = refant;
  a.phase_only = phase_only ? 1 : 0;
  a.normalise = normalise ? 1 : 0;
  a.tol = tol;
  a.damping = damping;
  ARLMI_LAUNCH("k_solve_itsubs", k_solve_itsubs, dim3((unsigned)ngain), dim3(kItsubsThreads), 0, s, a);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

}  // extern "C"
