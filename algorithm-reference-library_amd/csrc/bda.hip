// Baseline-dependent averaging of a BlockVisibility into rows (coalesce_visibility /
// average_in_blocks, processing_components/visibility/coalesce.py:36-82, 146-328, with
// average_chunks / average_chunks2, processing_library/util/array_functions.py:49-118), Stokes I.
//
// k_bda_uvdist gives every baseline its max over time of u^2 + v^2 and its summed weight; the host
// turns them into the per-baseline time and frequency factors and row offsets exactly as ARL does
// (a few thousand numbers).  k_bda_rows evaluates ARL's two-stage weighted averages (each
// integration's channel chunk first, then the integrations with those partial weights; a factor
// <= 1 passes values through unchanged) for every column at once, streaming each baseline's
// integrations in order.  k_bda_cindex maps every block element to its row.
#include "arlmi_internal.h"

namespace arlmi {
namespace {

constexpr int kBdaThreads = 256;
constexpr double kC = 299792458.0;  // astropy constants.c (m/s)

template <typename R>
__global__ void __launch_bounds__(kBdaThreads)
    k_bda_uvdist(const double* __restrict__ uvw, const R* __restrict__ wt, int ntimes, int nants, int nchan,
                 double* __restrict__ uvd, double* __restrict__ wsum) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nants * nants) return;
  double m = 0.0, ws = 0.0;
  for (int t = 0; t < ntimes; ++t) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)t * nants * nants + b;
    const double u = uvw[3 * e], v = uvw[3 * e + 1];
    m = fmax(m, u * u + v * v);
    for (int c = 0; c < nchan; ++c) ws += wt ? (double)wt[e * nchan + c] : 1.0;
  }
  uvd[b] = m;
  wsum[b] = ws;
}

// Running weighted sums of one chunk for every column (stage 1 over channels, stage 2 over times).
struct Acc {
  double w, time, freq, u, v, x, it, bw, vr, vi, iw;
};

// Adds w * value of every column, or (first = true, factor <= 1) takes the values themselves.
__device__ __forceinline__ void acc_add(Acc& s, double w, const double* v) {
#pragma clang fp contract(off)
  s.w += w;
  s.time += w * v[0];
  s.freq += w * v[1];
  s.u += w * v[2];
  s.v += w * v[3];
  s.x += w * v[4];
  s.it += w * v[5];
  s.bw += w * v[6];
  s.vr += w * v[7];
  s.vi += w * v[8];
  s.iw += w * v[9];
}
__device__ __forceinline__ void acc_norm(Acc& s) {
  if (s.w > 0.0) {  // putmask(chunks, weights > 0, chunks / weights)
    s.time /= s.w;
    s.freq /= s.w;
    s.u /= s.w;
    s.v /= s.w;
    s.x /= s.w;
    s.it /= s.w;
    s.bw /= s.w;
    s.vr /= s.w;
    s.vi /= s.w;
    s.iw /= s.w;
  }
}

// One thread per (baseline, channel), consecutive threads on consecutive block elements of one
// integration (coalesced).  The thread at the first channel of each frequency chunk streams
// through the integrations: stage 1 sums its chunk's channels in order, stage 2 accumulates the
// chunk's integrations in order with those partial weights and writes the cell when the time
// chunk ends.  The other threads of a chunk only help the reads stay contiguous (they exit).
template <typename R>
__global__ void __launch_bounds__(kBdaThreads)
    k_bda_rows(arlmi_blocks blk, arlmi_rows out, const int32_t* __restrict__ tav, const int32_t* __restrict__ fav,
               const int64_t* __restrict__ off) {
  const int nants = blk.nants, nchan = blk.nchan, ntimes = blk.ntimes, nb = nants * nants;
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)nb * nchan) return;
  const int b = (int)(id / nchan), fl = (int)(id - (int64_t)b * nchan);
  const int ta = tav[b], fa = fav[b];
  const int tstep = ta > 1 ? ta : 1, fstep = fa > 1 ? fa : 1;
  if (fl % fstep) return;
  const int a2 = b / nants, a1 = b - (b / nants) * nants;
  const int nf = (nchan + fstep - 1) / fstep, F = fl / fstep;
  const int f1 = min(nchan, fl + fstep);
  const int64_t cells = off[b + 1] - off[b];
  // the inputs are never written here: restrict lets the next integration's loads issue before
  // this one's stores
  const cplx<R>* __restrict__ vis = reinterpret_cast<const cplx<R>*>(blk.d_vis);
  const R* __restrict__ wt = reinterpret_cast<const R*>(blk.d_weight);
  const R* __restrict__ iwt = reinterpret_cast<const R*>(blk.d_imaging_weight);
  const double* __restrict__ in_uvw = blk.d_uvw;
  const double* __restrict__ in_time = blk.d_time;
  const double* __restrict__ in_itime = blk.d_integration_time;
  const double* __restrict__ in_freq = blk.d_frequency;
  const double* __restrict__ in_bw = blk.d_channel_bandwidth;
  Acc s2{};
#pragma unroll 4
  for (int t = 0; t < ntimes; ++t) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)t * nb + b;
    const double uvw0 = in_uvw[3 * e], uvw1 = in_uvw[3 * e + 1], uvw2 = in_uvw[3 * e + 2];
    const double tt = in_time[t], it = in_itime[t];
    Acc s1{};
    for (int f = fl; f < f1; ++f) {
      const int64_t k = e * nchan + f;
      const double w = wt ? (double)wt[k] : 1.0;
      const double kf = in_freq[f] / kC;  // frequency / constants.c
      const cplx<R> vv = vis[k];
      const double vals[10] = {tt, in_freq[f], uvw0 * kf, uvw1 * kf, uvw2 * kf, it,
                               in_bw[f], (double)vv.x, (double)vv.y, iwt ? (double)iwt[k] : 1.0};
      if (fa <= 1) {  // average_chunks returns the values themselves
        s1 = Acc{w, vals[0], vals[1], vals[2], vals[3], vals[4], vals[5], vals[6], vals[7], vals[8], vals[9]};
      } else {
        acc_add(s1, w, vals);
      }
    }
    if (fa > 1) acc_norm(s1);
    Acc a;
    int T;
    if (ta <= 1) {
      a = s1;
      T = t;
    } else {
      const double v1[10] = {s1.time, s1.freq, s1.u, s1.v, s1.x, s1.it, s1.bw, s1.vr, s1.vi, s1.iw};
      acc_add(s2, s1.w, v1);
      if ((t + 1) % tstep != 0 && t != ntimes - 1) continue;
      a = s2;
      acc_norm(a);
      s2 = Acc{};
      T = t / tstep;
    }
    const int64_t r = off[b] + (int64_t)T * nf + F;
    if (out.d_uvw) {
      out.d_uvw[3 * r] = a.u;
      out.d_uvw[3 * r + 1] = a.v;
      out.d_uvw[3 * r + 2] = a.x;
    }
    if (out.d_vis) {
      cplx<R> o;
      o.x = (R)a.vr;
      o.y = (R)a.vi;
      reinterpret_cast<cplx<R>*>(out.d_vis)[r] = o;
    }
    if (out.d_weight) reinterpret_cast<R*>(out.d_weight)[r] = (R)a.w;
    if (out.d_imaging_weight) reinterpret_cast<R*>(out.d_imaging_weight)[r] = (R)a.iw;
    if (out.d_time) out.d_time[r] = a.time;
    if (out.d_frequency) out.d_frequency[r] = a.freq;
    // ARL's "sum" of these multiplies the averages by the baseline's number of cells
    if (out.d_integration_time) out.d_integration_time[r] = a.it * (double)cells;
    if (out.d_channel_bandwidth) out.d_channel_bandwidth[r] = a.bw * (double)cells;
    if (out.d_antenna1) out.d_antenna1[r] = a1;
    if (out.d_antenna2) out.d_antenna2[r] = a2;
  }
}

// Element (t, a2, a1, ch) -> start + ((t * nchan + ch) mod cells) of its baseline: ARL assigns
// range(cells) to the ntimes x nchan indices through a flat iterator, which cycles the values.
__global__ void __launch_bounds__(kBdaThreads)
    k_bda_cindex(int ntimes, int nants, int nchan, const int64_t* __restrict__ off, int64_t* __restrict__ cindex) {
  const int64_t n = (int64_t)ntimes * nants * nants * nchan;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += stride) {
    const int ch = (int)(e % nchan);
    const int64_t q = e / nchan;
    const int b = (int)(q % ((int64_t)nants * nants));
    const int t = (int)(q / ((int64_t)nants * nants));
    const int64_t cells = off[b + 1] - off[b];
    cindex[e] = off[b] + ((int64_t)t * nchan + ch) % cells;
  }
}

unsigned bda_grid(int64_t n) {
  const int64_t g = (n + kBdaThreads - 1) / kBdaThreads;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 1 << 16));
}

}  // namespace
}  // namespace arlmi

using namespace arlmi;

extern "C" {

int arlmi_bda_uvdist(const double* d_uvw, const void* d_weight, int32_t ntimes, int32_t nants, int32_t nchan,
                     double* d_uvdist2, double* d_wsum, int precision, void* stream) {
  if ((precision != 32 && precision != 64) || ntimes < 1 || nants < 1 || nchan < 1 || !d_uvw || !d_uvdist2 ||
      !d_wsum) {
    set_error("arlmi_bda_uvdist: bad argument");
    return ARLMI_E_ARG;
  }
  hipStream_t s = as_stream(stream);
  const int nb = nants * nants;
  if (precision == 32)
    ARLMI_LAUNCH("k_bda_uvdist", k_bda_uvdist<float>, dim3((nb + kBdaThreads - 1) / kBdaThreads), dim3(kBdaThreads),
                 0, s, d_uvw, (const float*)d_weight, ntimes, nants, nchan, d_uvdist2, d_wsum);
  else
    ARLMI_LAUNCH("k_bda_uvdist", k_bda_uvdist<double>, dim3((nb + kBdaThreads - 1) / kBdaThreads),
                 dim3(kBdaThreads), 0, s, d_uvw, (const double*)d_weight, ntimes, nants, nchan, d_uvdist2, d_wsum);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

int arlmi_bda_average(const arlmi_blocks* blocks, const int32_t* d_tav, const int32_t* d_fav, const int64_t* d_rowoff,
                      int64_t nrows, const arlmi_rows* rows, int64_t* d_cindex, int precision, void* stream) {
  if ((precision != 32 && precision != 64) || !blocks || !rows || !d_tav || !d_fav || !d_rowoff || nrows < 0) {
    set_error("arlmi_bda_average: bad argument");
    return ARLMI_E_ARG;
  }
  const arlmi_blocks& b = *blocks;
  if (b.ntimes < 1 || b.nants < 1 || b.nchan < 1 || !b.d_vis || !b.d_uvw || !b.d_time || !b.d_integration_time ||
      !b.d_frequency || !b.d_channel_bandwidth) {
    set_error("arlmi_bda_average: the blocks need vis, uvw, time, integration time, frequency and bandwidth");
    return ARLMI_E_ARG;
  }
  hipStream_t s = as_stream(stream);
  if (nrows > 0) {
    const int64_t nt = (int64_t)b.nants * b.nants * b.nchan;
    const unsigned g = (unsigned)((nt + kBdaThreads - 1) / kBdaThreads);
    if (precision == 32)
      ARLMI_LAUNCH("k_bda_rows", k_bda_rows<float>, dim3(g), dim3(kBdaThreads), 0, s, b, *rows, d_tav, d_fav,
                   d_rowoff);
    else
      ARLMI_LAUNCH("k_bda_rows", k_bda_rows<double>, dim3(g), dim3(kBdaThreads), 0, s, b, *rows, d_tav, d_fav,
                   d_rowoff);
  }
  if (d_cindex) {
    const int64_t n = (int64_t)b.ntimes * b.nants * b.nants * b.nchan;
    ARLMI_LAUNCH("k_bda_cindex", k_bda_cindex, dim3(bda_grid(n)), dim3(kBdaThreads), 0, s, b.ntimes, b.nants,
                 b.nchan, d_rowoff, d_cindex);
  }
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

}  // extern "C"
