// Visibility layout and gain convention: BlockVisibility -> row conversion (coalesce.py:331-444),
// scalar gain application on blocks (calibration/operations.py:79-164) and on rows, and the
// adjoint of the row form.
//
// All of these are HBM-bound gathers and elementwise products: one thread per output element,
// consecutive threads on consecutive channels, which are contiguous in both layouts.
#include "arlmi_internal.h"

namespace arlmi {
namespace {

constexpr double kLightSpeed = 299792458.0;  // astropy constants.c (m/s)
constexpr int kVisThreads = 256;
constexpr int kGainLds = 4096;  // gains whose gradient is pre-summed in LDS (2 x 32 KiB of float64)

// Complex products without fused multiply-add, in numpy's operation order.
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
#pragma clang fp contract(off)
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename C>
__device__ __forceinline__ C cconj(C a) {
  a.y = -a.y;
  return a;
}

// 1 / g by Smith's method (numpy's complex division of 1 by g).
template <typename R>
__device__ __forceinline__ cplx<R> crecip(cplx<R> g) {
#pragma clang fp contract(off)
  const R gx = g.x, gy = g.y;
  cplx<R> r;
  if (fabs(gx) >= fabs(gy)) {
    const R q = gy / gx;
    const R d = gx + gy * q;
    r.x = (R)1 / d;
    r.y = -q / d;
  } else {
    const R q = gx / gy;
    const R d = gx * q + gy;
    r.x = q / d;
    r.y = (R)-1 / d;
  }
  return r;
}

// Lower-triangle baseline index bl = a2 (a2 - 1) / 2 + a1 (0 <= a1 < a2) -> (a2, a1).
__device__ __forceinline__ void baseline_pair(int bl, int& a2, int& a1) {
  int a = (int)((1.0 + sqrt(1.0 + 8.0 * (double)bl)) * 0.5);
  if (a * (a - 1) / 2 > bl) --a;
  if ((a + 1) * a / 2 <= bl) ++a;
  a2 = a;
  a1 = bl - a * (a - 1) / 2;
}

template <typename R>
__global__ void __launch_bounds__(kVisThreads) k_block_to_rows(arlmi_blocks b, arlmi_rows r, int64_t nrows) {
  using C = cplx<R>;
  const int nchan = b.nchan, nants = b.nants;
  const int nbl = nants * (nants - 1) / 2;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < nrows; row += stride) {
    const int64_t q = row / nchan;
    const int ch = (int)(row - q * nchan);
    const int64_t t = q / nbl;
    const int bl = (int)(q - t * nbl);
    int a2, a1;
    baseline_pair(bl, a2, a1);
    const int64_t e = (t * nants + a2) * nants + a1;  // element [t, a2, a1]
    const int64_t ec = e * nchan + ch;
    if (r.d_vis) reinterpret_cast<C*>(r.d_vis)[row] = reinterpret_cast<const C*>(b.d_vis)[ec];
    if (r.d_weight && b.d_weight) reinterpret_cast<R*>(r.d_weight)[row] = reinterpret_cast<const R*>(b.d_weight)[ec];
    if (r.d_imaging_weight && b.d_imaging_weight)
      reinterpret_cast<R*>(r.d_imaging_weight)[row] = reinterpret_cast<const R*>(b.d_imaging_weight)[ec];
    const double f = b.d_frequency[ch];
    if (r.d_uvw) {
      const double k = f / kLightSpeed;
      r.d_uvw[3 * row] = b.d_uvw[3 * e] * k;
      r.d_uvw[3 * row + 1] = b.d_uvw[3 * e + 1] * k;
      r.d_uvw[3 * row + 2] = b.d_uvw[3 * e + 2] * k;
    }
    if (r.d_frequency) r.d_frequency[row] = f;
    if (r.d_channel_bandwidth && b.d_channel_bandwidth) r.d_channel_bandwidth[row] = b.d_channel_bandwidth[ch];
    if (r.d_time && b.d_time) r.d_time[row] = b.d_time[t];
    if (r.d_integration_time && b.d_integration_time) r.d_integration_time[row] = b.d_integration_time[t];
    if (r.d_antenna1) r.d_antenna1[row] = a1;
    if (r.d_antenna2) r.d_antenna2[row] = a2;
  }
}

// cindex over every block element: the row of a lower-triangle element, 0 elsewhere.
__global__ void __launch_bounds__(kVisThreads) k_block_cindex(int64_t* cindex, int nants, int nchan, int64_t n) {
  const int nbl = nants * (nants - 1) / 2;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += stride) {
    const int64_t q = e / nchan;
    const int ch = (int)(e - q * nchan);
    const int64_t tij = q;
    const int64_t t = tij / ((int64_t)nants * nants);
    const int ij = (int)(tij - t * nants * nants);
    const int i = ij / nants, j = ij - i * nants;  // [t, i, j]: i = a2, j = a1 when j < i
    int64_t row = 0;
    if (j < i) row = (t * nbl + i * (i - 1) / 2 + j) * nchan + ch;
    cindex[e] = row;
  }
}

template <typename R>
__global__ void __launch_bounds__(kVisThreads)
    k_gains_block(cplx<R>* __restrict__ vis, int nants, int nchan, const cplx<R>* __restrict__ gain,
                  const double* __restrict__ gwt, const int32_t* __restrict__ slot, int nslots, int inverse,
                  int64_t n) {
  using C = cplx<R>;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += stride) {
    const int64_t q = e / nchan;
    const int ch = (int)(e - q * nchan);
    const int64_t t = q / ((int64_t)nants * nants);
    const int ij = (int)(q - t * nants * nants);
    const int i = ij / nants, j = ij - i * nants;
    const int s = slot[t];
    if (s < 0 || s >= nslots) continue;
    const int64_t gi = ((int64_t)s * nants + i) * nchan + ch;
    const int64_t gj = ((int64_t)s * nants + j) * nchan + ch;
    C li = gain[gi], lj = gain[gj];
    if (inverse) {
      // |g| > 0 <=> g != 0 (numpy.abs of a finite complex is 0 only at 0)
      li = (li.x != 0 || li.y != 0) ? crecip<R>(li) : C{(R)1, (R)0};
      lj = (lj.x != 0 || lj.y != 0) ? crecip<R>(lj) : C{(R)1, (R)0};
    }
    const C smueller = cmul(li, cconj(lj));
    C v = cmul(vis[e], smueller);
    if (gwt && gwt[gi] * gwt[gj] == 0.0) v = C{(R)0, (R)0};
    vis[e] = v;
  }
}

template <typename R>
__global__ void __launch_bounds__(kVisThreads)
    k_gains_rows(int64_t n, const int32_t* __restrict__ i1, const int32_t* __restrict__ i2,
                 const cplx<R>* __restrict__ gain, int64_t ngain, const cplx<R>* in, cplx<R>* out) {
  using C = cplx<R>;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += stride) {
    const int32_t a = i1[r], b = i2[r];
    C v = C{(R)0, (R)0};
    if (a >= 0 && a < ngain && b >= 0 && b < ngain) v = cmul(in[r], cmul(gain[b], cconj(gain[a])));
    out[r] = v;
  }
}

// Adjoint of k_gains_rows.  With LDS, each work group pre-sums the gain gradient of its rows in
// float64 LDS accumulators and adds them to the output once per gain it touched; without, every
// row adds straight to the output.
template <typename R, bool LDS>
__global__ void __launch_bounds__(kVisThreads)
    k_gains_rows_adjoint(int64_t n, const int32_t* __restrict__ i1, const int32_t* __restrict__ i2,
                         const cplx<R>* __restrict__ gain, int ngain, const cplx<R>* __restrict__ in,
                         const cplx<R>* __restrict__ gout, cplx<R>* __restrict__ gin, double* __restrict__ ggain) {
  using C = cplx<R>;
  __shared__ double acc[LDS ? 2 * kGainLds : 1];
  if (LDS && ggain) {
    for (int k = threadIdx.x; k < 2 * ngain; k += blockDim.x) acc[k] = 0.0;
    __syncthreads();
  }
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += stride) {
    const int32_t a = i1[r], b = i2[r];
    const bool ok = a >= 0 && a < ngain && b >= 0 && b < ngain;
    const C g = gout[r];
    if (gin) gin[r] = ok ? cmul(g, cconj(cmul(gain[b], cconj(gain[a])))) : C{(R)0, (R)0};
    if (ggain && ok) {
      const C v = in[r];
      const C c2 = cmul(cmul(g, gain[a]), cconj(v));   // -> g[antenna2]
      const C c1 = cmul(cmul(cconj(g), gain[b]), v);   // -> g[antenna1]
      if (LDS) {
        __hip_atomic_fetch_add(&acc[2 * b], (double)c2.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&acc[2 * b + 1], (double)c2.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&acc[2 * a], (double)c1.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&acc[2 * a + 1], (double)c1.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      } else {
        unsafeAtomicAdd(&ggain[2 * b], (double)c2.x);
        unsafeAtomicAdd(&ggain[2 * b + 1], (double)c2.y);
        unsafeAtomicAdd(&ggain[2 * a], (double)c1.x);
        unsafeAtomicAdd(&ggain[2 * a + 1], (double)c1.y);
      }
    }
  }
  if (LDS && ggain) {
    __syncthreads();
    for (int k = threadIdx.x; k < 2 * ngain; k += blockDim.x) {
      const double v = acc[k];
      if (v != 0.0) unsafeAtomicAdd(&ggain[k], v);
    }
  }
}

int prec_check(int precision) {
  if (precision != 32 && precision != 64) {
    set_error("precision must be 32 or 64");
    return ARLMI_E_ARG;
  }
  return ARLMI_OK;
}

// ---------------------------------------------------------------------------
// Phase rotation (visibility/base.py:295-381)
// ---------------------------------------------------------------------------

constexpr double kRotMinusTwoPi = -2.0 * 3.14159265358979323846264338327950288;  // -2 * numpy.pi

struct RotArgs {
  double l, m, n1;  // direction cosines of the new phase centre, n1 = sqrt(1 - l^2 - m^2) - 1
  int inverse;      // multiply by the phasor instead of its conjugate
  int rotate;       // tangent = False: rotate uvw
  double ch0, sh0, cd0, sd0;  // cos / sin of ha0 = -ra0 and dec0 (old phase centre)
  double ch1, sh1, cd1, sd1;  // the same for the new phase centre
};

// exp(-2 pi i (u l + v m + w n1)) as (cos, sin) of (-2 pi) * d: numpy's -2j * pi * dot(uvw, s)
// has real part -0 * d and imaginary part (-2 pi) * d, and exp of (0, phi) is (cos, sin)(phi).
__device__ __forceinline__ double2 rot_phasor(const RotArgs& a, double u, double v, double w) {
#pragma clang fp contract(off)
  const double d = u * a.l + v * a.m + w * a.n1;
  double sn, cs;
  sincos(kRotMinusTwoPi * d, &sn, &cs);
  return make_double2(cs, sn);
}

// vis *= conj(phasor) (or phasor), in complex128 as numpy's in-place product with a complex64
// or complex128 left operand computes it, rounded once to the storage precision.
template <typename R>
__device__ __forceinline__ void rot_apply(cplx<R>* __restrict__ vis, int npol, double2 ph, int inverse) {
  const double2 f = make_double2(ph.x, inverse ? ph.y : -ph.y);
  for (int p = 0; p < npol; ++p) {
    const cplx<R> v = vis[p];
    const double2 r = cmul(make_double2((double)v.x, (double)v.y), f);
    cplx<R> o;
    o.x = (R)r.x;
    o.y = (R)r.y;
    vis[p] = o;
  }
}

// xyz = uvw_to_xyz(uvw, ha0, dec0), uvw = xyz_to_uvw(xyz, ha1, dec1) (coordinate_support.py:59-113),
// operation for operation.
__device__ __forceinline__ void rot_uvw(const RotArgs& a, double& u, double& v, double& w) {
#pragma clang fp contract(off)
  const double v0 = v * a.sd0 - w * a.cd0;
  const double z = v * a.cd0 + w * a.sd0;
  const double x = u * a.ch0 + v0 * a.sh0;
  const double y = -u * a.sh0 + v0 * a.ch0;
  const double u1 = x * a.ch1 - y * a.sh1;
  const double q0 = x * a.sh1 + y * a.ch1;
  w = z * a.sd1 - q0 * a.cd1;
  v = z * a.cd1 + q0 * a.sd1;
  u = u1;
}

// One thread per row: the phasor from the row's uvw before rotation, then the rotated uvw.
template <typename R>
__global__ void __launch_bounds__(kVisThreads)
    k_rotate_rows(double* __restrict__ uvw, cplx<R>* __restrict__ vis, int npol, RotArgs a, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += stride) {
    double u = uvw[3 * r], v = uvw[3 * r + 1], w = uvw[3 * r + 2];
    rot_apply<R>(vis + r * npol, npol, rot_phasor(a, u, v, w), a.inverse);
    if (a.rotate) {
      rot_uvw(a, u, v, w);
      uvw[3 * r] = u;
      uvw[3 * r + 1] = v;
      uvw[3 * r + 2] = w;
    }
  }
}

// One thread per (block element, channel): uvw_m * (f / c) per component, then the phasor.
template <typename R>
__global__ void __launch_bounds__(kVisThreads)
    k_rotate_blocks(const double* __restrict__ uvw, cplx<R>* __restrict__ vis, int nchan, int npol,
                    const double* __restrict__ freq, RotArgs a, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += stride) {
    const int64_t b = e / nchan;
    const int ch = (int)(e - b * nchan);
    const double k = freq[ch] / kLightSpeed;
    const double u = uvw[3 * b] * k, v = uvw[3 * b + 1] * k, w = uvw[3 * b + 2] * k;
    rot_apply<R>(vis + e * npol, npol, rot_phasor(a, u, v, w), a.inverse);
  }
}

// In-place uvw rotation of the block elements (after every phasor has read the old uvw).
__global__ void __launch_bounds__(kVisThreads) k_rotate_uvw(double* __restrict__ uvw, RotArgs a, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < n; b += stride) {
    double u = uvw[3 * b], v = uvw[3 * b + 1], w = uvw[3 * b + 2];
    rot_uvw(a, u, v, w);
    uvw[3 * b] = u;
    uvw[3 * b + 1] = v;
    uvw[3 * b + 2] = w;
  }
}

RotArgs rot_args(const double* lmn, const double* trig, int inverse) {
  RotArgs a{};
  a.l = lmn[0];
  a.m = lmn[1];
  a.n1 = lmn[2];
  a.inverse = inverse ? 1 : 0;
  a.rotate = trig ? 1 : 0;
  if (trig) {
    a.ch0 = trig[0];
    a.sh0 = trig[1];
    a.cd0 = trig[2];
    a.sd0 = trig[3];
    a.ch1 = trig[4];
    a.sh1 = trig[5];
    a.cd1 = trig[6];
    a.sd1 = trig[7];
  }
  return a;
}

unsigned grid_for(int64_t n, int64_t cap) {
  const int64_t b = (n + kVisThreads - 1) / kVisThreads;
  return (unsigned)(b < cap ? b : cap);
}

}  // namespace
}  // namespace arlmi

using namespace arlmi;

extern "C" {

int arlmi_block_to_rows(const arlmi_blocks* blocks, const arlmi_rows* rows, int64_t* d_cindex, int precision,
                        void* stream) {
  if (int rc = prec_check(precision)) return rc;
  if (!blocks || !rows || blocks->ntimes < 0 || blocks->nants < 0 || blocks->nchan <= 0) {
    set_error("arlmi_block_to_rows: bad argument");
    return ARLMI_E_ARG;
  }
  const arlmi_blocks& b = *blocks;
  const int64_t nbl = (int64_t)b.nants * (b.nants - 1) / 2;
  const int64_t nrows = (int64_t)b.ntimes * nbl * b.nchan;
  const int64_t nelem = (int64_t)b.ntimes * b.nants * b.nants * b.nchan;
  if (b.nants > 46341 || nbl * b.nchan > INT32_MAX) {
    set_error("arlmi_block_to_rows: %d antennas x %d channels is beyond this build", b.nants, b.nchan);
    return ARLMI_E_UNSUPPORTED;
  }
  if (nrows > 0 && (!b.d_frequency || (rows->d_vis && !b.d_vis) || (rows->d_uvw && !b.d_uvw))) {
    set_error("arlmi_block_to_rows: a requested column has no input");
    return ARLMI_E_ARG;
  }
  hipStream_t s = as_stream(stream);
  if (nrows > 0) {
    const unsigned g = grid_for(nrows, 1 << 16);
    if (precision == 32)
      ARLMI_LAUNCH("k_block_to_rows", k_block_to_rows<float>, dim3(g), dim3(kVisThreads), 0, s, b, *rows, nrows);
    else
      ARLMI_LAUNCH("k_block_to_rows", k_block_to_rows<double>, dim3(g), dim3(kVisThreads), 0, s, b, *rows, nrows);
    ARLMI_LAUNCH_CHECK();
  }
  if (d_cindex && nelem > 0) {
    ARLMI_LAUNCH("k_block_cindex", k_block_cindex, dim3(grid_for(nelem, 1 << 16)), dim3(kVisThreads), 0, s, d_cindex,
                 b.nants, b.nchan, nelem);
    ARLMI_LAUNCH_CHECK();
  }
  return ARLMI_OK;
}

int arlmi_apply_gaintable(void* d_vis, int32_t ntimes, int32_t nants, int32_t nchan, const void* d_gain,
                          const double* d_gain_weight, const int32_t* d_gain_slot, int32_t ngain_times,
                          int inverse, int precision, void* stream) {
  if (int rc = prec_check(precision)) return rc;
  const int64_t n = (int64_t)ntimes * nants * nants * nchan;
  if (ntimes < 0 || nants < 0 || nchan <= 0 || ngain_times <= 0 || (n > 0 && (!d_vis || !d_gain || !d_gain_slot))) {
    set_error("arlmi_apply_gaintable: bad argument");
    return ARLMI_E_ARG;
  }
  if (nants > 46341) {
    set_error("arlmi_apply_gaintable: %d antennas is beyond this build", nants);
    return ARLMI_E_UNSUPPORTED;
  }
  if (n == 0) return ARLMI_OK;
  hipStream_t s = as_stream(stream);
  const unsigned g = grid_for(n, 1 << 16);
  if (precision == 32)
    ARLMI_LAUNCH("k_gains_block", k_gains_block<float>, dim3(g), dim3(kVisThreads), 0, s, (float2*)d_vis, nants, nchan,
                 (const float2*)d_gain, d_gain_weight, d_gain_slot, ngain_times, inverse, n);
  else
    ARLMI_LAUNCH("k_gains_block", k_gains_block<double>, dim3(g), dim3(kVisThreads), 0, s, (double2*)d_vis, nants,
                 nchan, (const double2*)d_gain, d_gain_weight, d_gain_slot, ngain_times, inverse, n);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

int arlmi_apply_gains_rows(int64_t nrows, const int32_t* d_gidx1, const int32_t* d_gidx2, const void* d_gain,
                           int64_t ngain, const void* d_in, void* d_out, int precision, void* stream) {
  if (int rc = prec_check(precision)) return rc;
  if (nrows < 0 || ngain < 0 || (nrows > 0 && (!d_gidx1 || !d_gidx2 || !d_in || !d_out || (ngain > 0 && !d_gain)))) {
    set_error("arlmi_apply_gains_rows: bad argument");
    return ARLMI_E_ARG;
  }
  if (nrows == 0) return ARLMI_OK;
  hipStream_t s = as_stream(stream);
  const unsigned g = grid_for(nrows, 1 << 16);
  if (precision == 32)
    ARLMI_LAUNCH("k_gains_rows", k_gains_rows<float>, dim3(g), dim3(kVisThreads), 0, s, nrows, d_gidx1, d_gidx2,
                 (const float2*)d_gain, ngain, (const float2*)d_in, (float2*)d_out);
  else
    ARLMI_LAUNCH("k_gains_rows", k_gains_rows<double>, dim3(g), dim3(kVisThreads), 0, s, nrows, d_gidx1, d_gidx2,
                 (const double2*)d_gain, ngain, (const double2*)d_in, (double2*)d_out);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

int arlmi_apply_gains_rows_adjoint(int64_t nrows, const int32_t* d_gidx1, const int32_t* d_gidx2, const void* d_gain,
                                   int64_t ngain, const void* d_in, const void* d_grad_out, void* d_grad_in,
                                   void* d_grad_gain, int precision, void* stream) {
  if (int rc = prec_check(precision)) return rc;
  if (nrows < 0 || ngain < 0 || ngain > INT32_MAX / 2 ||
      (nrows > 0 && (!d_gidx1 || !d_gidx2 || !d_grad_out || (ngain > 0 && !d_gain) || (d_grad_gain && !d_in)))) {
    set_error("arlmi_apply_gains_rows_adjoint: bad argument");
    return ARLMI_E_ARG;
  }
  if (nrows == 0 || (!d_grad_in && !d_grad_gain)) return ARLMI_OK;
  hipStream_t s = as_stream(stream);
  const bool lds = d_grad_gain && ngain <= kGainLds;
  // with LDS pre-sums, few work groups keep the number of flushes (groups x gains) small
  const unsigned g = grid_for(nrows, lds ? 512 : (1 << 16));
  const int ng = (int)ngain;
  double* gg = (double*)d_grad_gain;
  if (precision == 32) {
    auto kern = lds ? k_gains_rows_adjoint<float, true> : k_gains_rows_adjoint<float, false>;
    ARLMI_LAUNCH("k_gains_rows_adjoint", kern, dim3(g), dim3(kVisThreads), 0, s, nrows, d_gidx1, d_gidx2,
                 (const float2*)d_gain, ng, (const float2*)d_in, (const float2*)d_grad_out, (float2*)d_grad_in, gg);
  } else {
    auto kern = lds ? k_gains_rows_adjoint<double, true> : k_gains_rows_adjoint<double, false>;
    ARLMI_LAUNCH("k_gains_rows_adjoint", kern, dim3(g), dim3(kVisThreads), 0, s, nrows, d_gidx1, d_gidx2,
                 (const double2*)d_gain, ng, (const double2*)d_in, (const double2*)d_grad_out, (double2*)d_grad_in,
                 gg);
  }
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

}  // extern "C"

extern "C" {

int arlmi_phaserotate_rows(double* d_uvw, void* d_vis, int64_t nrows, int32_t npol, const double* h_lmn,
                           const double* h_trig, int inverse, int precision, void* stream) {
  if (int rc = prec_check(precision)) return rc;
  if (nrows < 0 || npol < 1 || !h_lmn || (nrows > 0 && (!d_uvw || !d_vis))) {
    set_error("arlmi_phaserotate_rows: bad argument");
    return ARLMI_E_ARG;
  }
  if (nrows == 0) return ARLMI_OK;
  hipStream_t s = as_stream(stream);
  const RotArgs a = rot_args(h_lmn, h_trig, inverse);
  const unsigned g = grid_for(nrows, 1 << 16);
  if (precision == 32)
    ARLMI_LAUNCH("k_rotate_rows", k_rotate_rows<float>, dim3(g), dim3(kVisThreads), 0, s, d_uvw, (float2*)d_vis, npol,
                 a, nrows);
  else
    ARLMI_LAUNCH("k_rotate_rows", k_rotate_rows<double>, dim3(g), dim3(kVisThreads), 0, s, d_uvw, (double2*)d_vis,
                 npol, a, nrows);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

int arlmi_phaserotate_blocks(double* d_uvw, void* d_vis, int64_t nblock, int32_t nchan, int32_t npol,
                             const double* d_frequency, const double* h_lmn, const double* h_trig, int inverse,
                             int precision, void* stream) {
  if (int rc = prec_check(precision)) return rc;
  if (nblock < 0 || nchan < 1 || npol < 1 || !h_lmn || (nblock > 0 && (!d_uvw || !d_vis || !d_frequency))) {
    set_error("arlmi_phaserotate_blocks: bad argument");
    return ARLMI_E_ARG;
  }
  if (nblock == 0) return ARLMI_OK;
  hipStream_t s = as_stream(stream);
  const RotArgs a = rot_args(h_lmn, h_trig, inverse);
  const int64_t n = nblock * nchan;
  const unsigned g = grid_for(n, 1 << 16);
  if (precision == 32)
    ARLMI_LAUNCH("k_rotate_blocks", k_rotate_blocks<float>, dim3(g), dim3(kVisThreads), 0, s, d_uvw, (float2*)d_vis,
                 nchan, npol, d_frequency, a, n);
  else
    ARLMI_LAUNCH("k_rotate_blocks", k_rotate_blocks<double>, dim3(g), dim3(kVisThreads), 0, s, d_uvw,
                 (double2*)d_vis, nchan, npol, d_frequency, a, n);
  ARLMI_LAUNCH_CHECK();
  if (h_trig) {
    ARLMI_LAUNCH("k_rotate_uvw", k_rotate_uvw, dim3(grid_for(nblock, 1 << 16)), dim3(kVisThreads), 0, s, d_uvw, a,
                 nblock);
    ARLMI_LAUNCH_CHECK();
  }
  return ARLMI_OK;
}

}  // extern "C"
