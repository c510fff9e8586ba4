// Gridding plan: per-row convolution mapping, w-slab membership and slab means, and the
// tile-sorted entry lists walked by the gridding and degridding kernels.
//
// Follows processing_components/griddata/gridding.py:24-96 (convolution_mapping),
// processing_components/visibility/iterators.py:91-109 (vis_wslice_iter) and
// processing_components/imaging/wstack_single.py:43,83 (slab mean w).
//
// Layout: the grid of every plane (w slab or grid w plane) and image channel is cut into
// T x T tiles.  A (row, plane) membership produces one entry in every tile its S x S kernel
// footprint touches (1..4 tiles); the gridding kernel accumulates a tile in LDS and writes it
// back once, so no pixel needs a global atomic unless the tile is split into several chunks.
// Within a tile's segment the entries that the degridder uses (footprint-start tile of the
// row's last plane) come first.
#include "arlmi_internal.h"

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <string>
#include <vector>
#include <algorithm>

namespace arlmi {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

const char* last_error() { return g_last_error.c_str(); }

int first_error_code(uint32_t bits) {
  if (bits & ERRB_UV_OFF_GRID) return ARLMI_E_UV_OFF_GRID;
  if (bits & ERRB_SUBSAMPLE) return ARLMI_E_SUBSAMPLE;
  if (bits & ERRB_W_RANGE) return ARLMI_E_W_RANGE;
  if (bits & ERRB_FREQ_MAP) return ARLMI_E_FREQ_MAP;
  return ARLMI_OK;
}

// ---------------------------------------------------------------------------
// per-lane counter updates (inactive lanes do nothing)
// ---------------------------------------------------------------------------

__device__ inline void cnt_add(uint32_t* counters, uint32_t key, bool active) {
  if (active) atomicAdd(&counters[key], 1u);
}

__device__ inline uint32_t cnt_fetch_add(uint32_t* counters, uint32_t key, bool active) {
  return active ? atomicAdd(&counters[key], 1u) : 0u;
}

// ---------------------------------------------------------------------------
// max |w|
// ---------------------------------------------------------------------------

// Per-block max |w| into part[blockIdx.x] (the host takes the max of the partials: no device
// atomics on one address).
__global__ __launch_bounds__(256) void k_wmax(const double* __restrict__ uvw, int64_t n, double* part) {
  __shared__ double red[4];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double m = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
  for (; i + 3 * stride < n; i += 4 * stride) {  // four loads in flight per thread
    const double a = uvw[3 * i + 2], b = uvw[3 * (i + stride) + 2], c = uvw[3 * (i + 2 * stride) + 2],
                 d = uvw[3 * (i + 3 * stride) + 2];
    m = fmax(m, fabs(a));
    m1 = fmax(m1, fabs(b));
    m2 = fmax(m2, fabs(c));
    m3 = fmax(m3, fabs(d));
  }
  for (; i < n; i += stride) m = fmax(m, fabs(uvw[3 * i + 2]));
  m = fmax(fmax(m, m1), fmax(m2, m3));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if (__lane_id() == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// ARL's default slab centres from the k_wmax partials, on the device so that the host need not
// wait for max |w|: boxes = numpy.linspace(-wmax, wmax, P) with numpy's operations (step =
// delta / (P - 1), y_k = k * step + start, the last value = stop; for a zero step y_k =
// (k / (P - 1)) * delta + start; for P = 1, y_0 = 0 * delta + start), and after them the slab
// width boxes[1] - boxes[0], or 2 * wmax for one slab (iterators.py:99-106).  One block.
__global__ __launch_bounds__(256) void k_boxes(const double* __restrict__ part, int nparts, int P, double* boxes) {
  __shared__ double red[4];
  __shared__ double wmax_s;
  double m = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) m = fmax(m, part[i]);
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if (__lane_id() == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) wmax_s = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  __syncthreads();
  const double wmax = wmax_s;
  const double start = -wmax, stop = wmax;
  const double delta = stop - start;
  const int div = P - 1;
  if (div > 0) {
    const double step = delta / (double)div;
    for (int k = threadIdx.x; k < P; k += blockDim.x) {
      double y = (double)k;
      y = step == 0.0 ? (y / (double)div) * delta : y * step;
      y += start;
      boxes[k] = k == P - 1 ? stop : y;
    }
    if (threadIdx.x == 0) {
      const double y0 = (step == 0.0 ? (0.0 / (double)div) * delta : 0.0 * step) + start;
      const double y1 = P == 2 ? stop : (step == 0.0 ? (1.0 / (double)div) * delta : 1.0 * step) + start;
      boxes[P] = y1 - y0;
    }
  } else if (threadIdx.x == 0) {
    boxes[0] = 0.0 * delta + start;
    boxes[1] = 2.0 * wmax;
  }
}

// ---------------------------------------------------------------------------
// w-slab membership: |w - boxes[k]| < 0.5 * wslice (iterators.py:107-109)
// ---------------------------------------------------------------------------

struct Members {
  int n;
  int k[3];
};

// inv_wslice = 1 / wslice only places the candidate window (a rounding difference moves the
// nearest slab by at most one, and the window covers its neighbours); membership itself is the
// exact strict test.
__device__ inline Members slab_members(double w, const double* boxes, int nb, double wslice, double inv_wslice) {
  Members m;
  m.n = 0;
  m.k[0] = m.k[1] = m.k[2] = 0;
  const double halfw = 0.5 * wslice;
  if (!(wslice > 0.0)) return m;  // all |w| zero: |w - box| < 0 never holds
  const double kd = fmin(fmax(rint((w - boxes[0]) * inv_wslice), -2.0), (double)nb + 1.0);
  const int k0 = (int)kd;
  // the three candidate slabs around the nearest one, compacted with selects (no indexed
  // writes, so the result stays in registers)
  const int c0 = k0 - 1, c1 = k0, c2 = k0 + 1;
  const bool v0 = c0 >= 0 && c0 < nb && fabs(w - boxes[c0]) < halfw;
  const bool v1 = c1 >= 0 && c1 < nb && fabs(w - boxes[c1]) < halfw;
  const bool v2 = c2 >= 0 && c2 < nb && fabs(w - boxes[c2]) < halfw;
  m.n = (int)v0 + (int)v1 + (int)v2;
  m.k[0] = v0 ? c0 : (v1 ? c1 : c2);
  m.k[1] = (v0 && v1) ? c1 : c2;
  m.k[2] = c2;
  return m;
}

// Pass A (w-stack): per-slab count and sum of w -> slab means.  Each of at most SLAB_BLOCKS
// blocks writes its partial sums of w - centre; k_slab_means adds them.  The block sums are
// 64-bit integer sums of (w - centre) 2^q (det_round; a member lies within half a slab width of
// its centre, and a block adds at most rows_per_block of them), so they do not depend on the
// order of the LDS adds and the means are the same bits on every build of the plan.
constexpr int SLAB_BLOCKS = 2048;
__global__ void k_slab_sums(const double* __restrict__ uvw, int64_t n, const double* __restrict__ boxes, int nb,
                            double* wsum, uint32_t* wcnt) {
  const double wslice = boxes[nb];  // the slab width follows the centres
  extern __shared__ char smem[];
  long long* ls = reinterpret_cast<long long*>(smem);
  double* lbox = reinterpret_cast<double*>(ls + nb);
  unsigned int* lc = reinterpret_cast<unsigned int*>(lbox + nb);
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    ls[i] = 0;
    lbox[i] = boxes[i];
    lc[i] = 0u;
  }
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const double inv_wslice = 1.0 / wslice;
  const int64_t per_block = ((n + stride - 1) / stride) * (int64_t)blockDim.x;
  double half = 0.5 * wslice;  // members lie within half a slab width of their centre
  if (!(half < INFINITY)) {
    // an unbounded slab (every row a member): bound the offsets of this block's own rows
    __shared__ double red[32];
    double mx = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
      const double w = uvw[3 * i + 2];
      const Members m = slab_members(w, lbox, nb, wslice, inv_wslice);
      for (int j = 0; j < m.n; ++j) mx = fmax(mx, fabs(w - lbox[m.k[j]]));
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    if (__lane_id() == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    half = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) half = fmax(half, red[i]);
  }
  const int q = det_q((double)per_block * half, half);
  const double scale = ldexp(1.0, q);
  for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i0 < n; i0 += 4 * stride) {
    double wv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {  // four loads in flight per thread
      const int64_t i = i0 + u * stride;
      wv[u] = i < n ? uvw[3 * i + 2] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (i0 + u * stride >= n) break;
      const double w = wv[u];
      Members m = slab_members(w, lbox, nb, wslice, inv_wslice);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        if (j < m.n) {
          atomicAdd(reinterpret_cast<unsigned long long*>(&ls[m.k[j]]),
                    (unsigned long long)det_round(w - lbox[m.k[j]], scale));
          atomicAdd(&lc[m.k[j]], 1u);
        }
      }
    }
  }
  __syncthreads();
  const double inv_scale = ldexp(1.0, -q);
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {  // this block's partial sums of w - centre
    wsum[(size_t)blockIdx.x * nb + i] = (double)ls[i] * inv_scale;
    wcnt[(size_t)blockIdx.x * nb + i] = lc[i];
  }
}

// Slab k (blockIdx.x): sums the nblk per-block partials of w - centre in a fixed order
// (deterministic), mean = centre + sum / count (and, when tsum is given, the totals of w,
// count * centre + sum).
__global__ __launch_bounds__(256) void k_slab_means(const double* wsum, const uint32_t* wcnt, int nblk,
                                                    const double* __restrict__ boxes, double* wmean, int nb,
                                                    double* tsum = nullptr, unsigned long long* tcnt = nullptr) {
  __shared__ double rs[4];
  __shared__ unsigned long long rc[4];
  const int k = blockIdx.x;
  double sum = 0.0;
  unsigned long long cnt = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    sum += wsum[(size_t)b * nb + k];
    cnt += wcnt[(size_t)b * nb + k];
  }
  for (int o = 32; o > 0; o >>= 1) {
    sum += __shfl_xor(sum, o);
    cnt += __shfl_xor(cnt, o);
  }
  if (__lane_id() == 0) {
    rs[threadIdx.x >> 6] = sum;
    rc[threadIdx.x >> 6] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = (rs[0] + rs[1]) + (rs[2] + rs[3]);
    const unsigned long long c = rc[0] + rc[1] + rc[2] + rc[3];
    if (wmean) wmean[k] = c ? boxes[k] + t / (double)c : 0.0;
    if (tsum) {
      tsum[k] = (double)c * boxes[k] + t;
      tcnt[k] = c;
    }
  }
}

// ---------------------------------------------------------------------------
// Row enumeration shared by the count and scatter passes.
// ---------------------------------------------------------------------------

struct Enum {
  bool valid;        // row contributes (mapped and in at least one plane)
  bool dropped;
  int nm;            // memberships
  int plane[3];
  int zp[3];         // kernel w plane per membership
  RowMap map;
};

__device__ inline Enum enumerate_row(const DevGeom& g, const double* __restrict__ uvw, const double* __restrict__ freq,
                                     int64_t row, const double* __restrict__ boxes, double wslice, double inv_wslice,
                                     const double* __restrict__ plane_w, uint32_t flags, uint32_t* err_out) {
  Enum e;
  e.valid = false;
  e.dropped = false;
  e.nm = 0;
  e.plane[0] = e.plane[1] = e.plane[2] = 0;
  e.zp[0] = e.zp[1] = e.zp[2] = 0;
  const double u = uvw[3 * row + 0], v = uvw[3 * row + 1], w = uvw[3 * row + 2];
  const double f = freq ? freq[row] : 0.0;
  uint32_t err = 0;
  if (g.mode == ARLMI_MODE_WSTACK) {
    Members m = slab_members(w, boxes, g.nplanes, wslice, inv_wslice);
    if (m.n == 0) return e;  // in no slab: never mapped, never asserted (imaging_serial.py:93-95)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      e.plane[j] = m.k[j];
      e.zp[j] = 0;
      if (j < m.n) {
        const double ws = w - plane_w[m.k[j]];
        // grid WW axis of a 1-plane GridData (nw = 1, wstep = 1e15) is checked against the kernel's nz
        (void)map_wplane(ws, g.wstep_grid, g.crpix_wg, g.nz, &err);
        e.zp[j] = map_wplane(ws, g.zstep, g.crpix_z, g.nz, &err);
      }
    }
    e.nm = m.n;
  } else {
    // 2-D: grid w plane by rounding (gridding.py:58-65), kernel w plane (70-76)
    const int pg = map_wplane(w, g.wstep_grid, g.crpix_wg, g.nz, &err);
    if (pg >= g.nplanes) err |= ERRB_W_RANGE;
    e.plane[0] = pg < g.nplanes ? pg : 0;
    e.zp[0] = map_wplane(w, g.zstep, g.crpix_z, g.nz, &err);
    e.nm = 1;
  }
  e.map = map_row(g, u, v, f, freq != nullptr);
  err |= e.map.err;
  if ((flags & ARLMI_FLAG_DROP_OFF_GRID) && (e.map.err & ERRB_UV_OFF_GRID)) {
    e.dropped = true;
    e.nm = 0;
    return e;
  }
  if (err) {
    *err_out |= err;
    e.nm = 0;
    return e;
  }
  e.valid = true;
  return e;
}

// ---------------------------------------------------------------------------
// Binning without global atomics: a two-level counting sort.
//
// Pass 1 (rows -> planes): each block of K1 rows counts its memberships per plane in LDS
// (plus one bin for rows in no plane) and writes the counts to h1[bin][block].  One exclusive
// scan of h1 gives every (bin, block) its output range; the scatter pass re-enumerates the
// rows and writes one 16-byte record per membership into it (LDS cursors inside a block).
//
// Pass 2 (plane records -> tiles): every plane's record segment is cut into blocks of K2;
// a block counts its entries per bin = 2*tile + kind in LDS, kind 0 = the entry the degridder
// uses, 1 = the rest, and writes its histogram contiguously to h2[plane base + block * 2tpp + bin].
// The per-bin sums over the plane's blocks (k_h2_colsum), scanned once in (plane, tile, kind)
// order, give every bin's first slot in the final tile-sorted entry array (degrid entries first
// in a tile); k_h2_prefix adds the counts of the plane's earlier blocks in the bin to that slot,
// which is the block's cursor s2 (same layout as h2).  Block-major histograms keep the writes of
// k_p2_count and the cursor reads of k_p2_scatter contiguous (bin-major ones, one word per block
// apart, moved ~10x their size through HBM: 326 MB written for 29 MB of counts on the headline
// workload).
// ---------------------------------------------------------------------------

constexpr int K1_ROWS = 4096;  // rows per pass-1 block
constexpr int K2_RECS = 8192;  // records per pass-2 block, at least
// ... and at least K2_HIST times the words of the block's tile histogram (2 tpp), so that with many
// tiles per plane (128^2 at 8192^2) the per-block histograms that are written, summed and read
// back stay small against the records themselves.  Measured on the streamed 8192^2 workload (16
// M-row chunks, 128 slabs, float32): 0 (8192-record blocks) -> 270.5 ms per step, 1 -> 256.4 ms,
// 2 -> 255.3 ms, 4 -> 261.4 ms.
constexpr int K2_HIST = 2;
constexpr int K2_UNROLL = 4;  // record loads in flight per thread

// Membership records: 16 bytes in general, 8 where the grid side is at most 4096, the
// oversampling at most 8, and there is one image channel and one kernel w plane (pu, pv, xoff,
// yoff and the last-membership flag then fit one word beside the row).  The first word is the row
// in both (the degridders read the rows of orphan records with a stride of rec_words).
struct Rec {
  uint32_t row;
  uint32_t uv;   // pu | pv << 16
  uint32_t aux;  // xoff | yoff << 8 | chan << 16 | last-membership << 24
  uint32_t z;    // kernel w plane
};
struct Rec8 {
  uint32_t row;
  uint32_t pk;  // pu | pv << 12 | xoff << 24 | yoff << 27 | last-membership << 30
};
template <typename RT>
struct RecOps;
template <>
struct RecOps<Rec> {
  static __device__ Rec make(uint32_t row, const RowMap& m, bool last, int z) {
    return Rec{row, (uint32_t)m.pu | ((uint32_t)m.pv << 16),
               (uint32_t)m.xoff | ((uint32_t)m.yoff << 8) | ((uint32_t)m.chan << 16) | (last ? (1u << 24) : 0u),
               (uint32_t)z};
  }
  static __device__ Rec orphan(uint32_t row) { return Rec{row, 0u, 0u, 0u}; }
  static __device__ int pu(const Rec& r) { return (int)(r.uv & 0xffffu); }
  static __device__ int pv(const Rec& r) { return (int)(r.uv >> 16); }
  static __device__ int xoff(const Rec& r) { return (int)(r.aux & 0xffu); }
  static __device__ int yoff(const Rec& r) { return (int)((r.aux >> 8) & 0xffu); }
  static __device__ int chan(const Rec& r) { return (int)((r.aux >> 16) & 0xffu); }
  static __device__ bool last(const Rec& r) { return (r.aux >> 24) & 1u; }
  static __device__ int z(const Rec& r) { return (int)r.z; }
};
template <>
struct RecOps<Rec8> {
  static __device__ Rec8 make(uint32_t row, const RowMap& m, bool last, int) {
    return Rec8{row, (uint32_t)m.pu | ((uint32_t)m.pv << 12) | ((uint32_t)m.xoff << 24) | ((uint32_t)m.yoff << 27) |
                         (last ? (1u << 30) : 0u)};
  }
  static __device__ Rec8 orphan(uint32_t row) { return Rec8{row, 0u}; }
  static __device__ int pu(const Rec8& r) { return (int)(r.pk & 0xfffu); }
  static __device__ int pv(const Rec8& r) { return (int)((r.pk >> 12) & 0xfffu); }
  static __device__ int xoff(const Rec8& r) { return (int)((r.pk >> 24) & 7u); }
  static __device__ int yoff(const Rec8& r) { return (int)((r.pk >> 27) & 7u); }
  static __device__ int chan(const Rec8&) { return 0; }
  static __device__ bool last(const Rec8& r) { return (r.pk >> 30) & 1u; }
  static __device__ int z(const Rec8&) { return 0; }
};
// the compact records hold this plan's memberships
inline bool compact_recs(const DevGeom& g) { return g.n <= 4096 && g.os <= 8 && g.nchan == 1 && g.nz == 1; }

__global__ void k_p1_count(DevGeom g, const double* __restrict__ uvw, const double* __restrict__ freq, int64_t n,
                           const double* __restrict__ boxes, const double* __restrict__ plane_w,
                           uint32_t flags, uint32_t* __restrict__ h1, int nb1, uint32_t* misc) {
  // slab centres (and, after them, the slab width) and slab means staged in LDS (looked up per
  // row), then the counters
  extern __shared__ double lsh[];
  const int P = g.nplanes;
  double* lbox = lsh;
  double* lpw = lsh + P + 1;
  uint32_t* lh = reinterpret_cast<uint32_t*>(lsh + 2 * P + 1);
  for (int i = threadIdx.x; i <= P; i += blockDim.x) lbox[i] = boxes[i];
  for (int i = threadIdx.x; i < P; i += blockDim.x) lpw[i] = plane_w[i];
  for (int i = threadIdx.x; i < P + 3; i += blockDim.x) lh[i] = 0u;
  __syncthreads();
  const double wslice = lbox[P];
  const int64_t r0 = (int64_t)blockIdx.x * K1_ROWS;
  const int64_t r1 = r0 + K1_ROWS < n ? r0 + K1_ROWS : n;
  const double inv_wslice = 1.0 / wslice;
  uint32_t err = 0;
  // per-thread tallies of rows in no plane, dropped rows and memberships, summed once per wave
  // at the end (an LDS add of a varying value to one address per row would be expanded into a
  // loop over the active lanes)
  uint32_t n_orph = 0, n_drop = 0, n_mem = 0;
  for (int64_t row = r0 + threadIdx.x; row < r1; row += blockDim.x) {
    Enum e = enumerate_row(g, uvw, freq, row, lbox, wslice, inv_wslice, lpw, flags, &err);
    if (!e.valid) {
      ++n_orph;
      n_drop += e.dropped ? 1u : 0u;
    } else {
#pragma unroll
      for (int j = 0; j < 3; ++j)
        if (j < e.nm) atomicAdd(&lh[e.plane[j]], 1u);
      n_mem += (uint32_t)e.nm;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    n_orph += __shfl_xor(n_orph, o);
    n_drop += __shfl_xor(n_drop, o);
    n_mem += __shfl_xor(n_mem, o);
    err |= __shfl_xor(err, o);
  }
  if (__lane_id() == 0) {
    if (n_orph) atomicAdd(&lh[P], n_orph);
    if (n_drop) atomicAdd(&lh[P + 1], n_drop);
    if (n_mem) atomicAdd(&lh[P + 2], n_mem);
    if (err) atomicOr(&misc[0], err);
  }
  __syncthreads();
  for (int i = threadIdx.x; i <= P; i += blockDim.x) h1[(int64_t)i * nb1 + blockIdx.x] = lh[i];
  if (threadIdx.x == 0) {
    if (lh[P + 1]) atomicAdd(&misc[2], lh[P + 1]);
    if (lh[P + 2]) atomicAdd(&misc[3], lh[P + 2]);
  }
}

template <typename RT>
__global__ void k_p1_scatter(DevGeom g, const double* __restrict__ uvw, const double* __restrict__ freq, int64_t n,
                             const double* __restrict__ boxes, const double* __restrict__ plane_w,
                             uint32_t flags, const uint32_t* __restrict__ s1, int nb1, RT* __restrict__ recs,
                             uint32_t cap) {
  // records at or past cap are not written (the host launches this pass before it knows the
  // record count, into a buffer of cap records, and launches it again into a larger one if the
  // count exceeds cap)
  using O = RecOps<RT>;
  extern __shared__ double lsh[];
  const int P = g.nplanes;
  double* lbox = lsh;
  double* lpw = lsh + P + 1;
  uint32_t* cur = reinterpret_cast<uint32_t*>(lsh + 2 * P + 1);
  for (int i = threadIdx.x; i <= P; i += blockDim.x) lbox[i] = boxes[i];
  for (int i = threadIdx.x; i < P; i += blockDim.x) lpw[i] = plane_w[i];
  for (int i = threadIdx.x; i <= P; i += blockDim.x) cur[i] = s1[(int64_t)i * nb1 + blockIdx.x];
  __syncthreads();
  const double wslice = lbox[P];
  const int64_t r0 = (int64_t)blockIdx.x * K1_ROWS;
  const int64_t r1 = r0 + K1_ROWS < n ? r0 + K1_ROWS : n;
  const double inv_wslice = 1.0 / wslice;
  uint32_t dummy = 0;
  for (int64_t row = r0 + threadIdx.x; row < r1; row += blockDim.x) {
    Enum e = enumerate_row(g, uvw, freq, row, lbox, wslice, inv_wslice, lpw, flags, &dummy);
    if (!e.valid) {
      const uint32_t pos = atomicAdd(&cur[P], 1u);
      if (pos < cap) recs[pos] = O::orphan((uint32_t)row);
      continue;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (j >= e.nm) break;
      const uint32_t pos = atomicAdd(&cur[e.plane[j]], 1u);
      if (pos < cap) recs[pos] = O::make((uint32_t)row, e.map, j == e.nm - 1, e.zp[j]);
    }
  }
}

// gathers s1[i * nb1] for i = 0..P+1 (the plane segment starts and the total) and, after them,
// the 8 words of misc, so that one copy brings both to the host
__global__ void k_p1_bounds(const uint32_t* __restrict__ s1, int nb1, int P, const uint32_t* __restrict__ misc,
                            uint32_t* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= P + 1) out[i] = s1[(int64_t)i * nb1];
  else if (i < P + 10) out[i] = misc[i - (P + 2)];
}

// The four totals the host sizes the work-item lists with, gathered for one copy.
__global__ void k_totals(const uint32_t* __restrict__ s2, int64_t n2, const uint32_t* __restrict__ choff,
                         const uint32_t* __restrict__ choffd, const uint32_t* __restrict__ slotoff, int64_t nt,
                         const uint32_t* __restrict__ sortcnt, uint32_t* out) {
  out[0] = s2[n2];
  out[1] = choff[nt];
  out[2] = choffd[nt];
  out[3] = slotoff[nt];
  for (int k = 0; k < 4; ++k) out[4 + k] = sortcnt ? sortcnt[k] : 0u;
}

// Pass-2 block descriptor: plane, block index in the plane, record range, h2 base, blocks in plane.
struct Blk2 {
  int plane, b, nb;
  uint32_t rs, re;
  uint32_t hb;
};

// Pass-2 layout from the plane segment bounds (one work group): plane pl gets nbp[pl] =
// max(1, ceil(records / rb)) blocks, hbase[pl] = nbin * (blocks of the planes before it) and
// hbase[P] = nbin * all blocks (hbase holds [P+1] bases, then [P] block counts).
__global__ __launch_bounds__(1024) void k_p2_layout(const uint32_t* __restrict__ bounds, int P, uint32_t rb, int nbin,
                                                    uint32_t* __restrict__ hbase) {
  __shared__ uint32_t lsum[1024];
  const int per = (P + 1023) / 1024;  // planes a thread (P <= 4096)
  const int p0 = threadIdx.x * per;
  uint32_t loc = 0;
  for (int k = 0; k < per; ++k) {
    const int pl = p0 + k;
    if (pl < P) loc += max(1u, (bounds[pl + 1] - bounds[pl] + rb - 1) / rb);
  }
  lsum[threadIdx.x] = loc;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan of the per-thread block counts
    const uint32_t v = threadIdx.x >= (unsigned)off ? lsum[threadIdx.x - off] : 0u;
    __syncthreads();
    lsum[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = lsum[threadIdx.x] - loc;
  for (int k = 0; k < per; ++k) {
    const int pl = p0 + k;
    if (pl < P) {
      const uint32_t nb = max(1u, (bounds[pl + 1] - bounds[pl] + rb - 1) / rb);
      hbase[pl] = run * (uint32_t)nbin;
      hbase[P + 1 + pl] = nb;
      run += nb;
    }
  }
  if (threadIdx.x == 1023) hbase[P] = lsum[1023] * (uint32_t)nbin;
}

// This work group's pass-2 block: the plane whose blocks hold blockIdx.x (hbase strictly
// increases, every plane having a block), then its record range.
__device__ inline Blk2 p2_block(const uint32_t* __restrict__ bounds, const uint32_t* __restrict__ hbase, int P,
                                int nbin, uint32_t rb) {
  __shared__ Blk2 sb;
  if (threadIdx.x == 0) {
    const uint32_t key = blockIdx.x * (uint32_t)nbin;
    int lo = 0, hi = P - 1;  // the last plane with hbase <= key
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (hbase[mid] <= key) lo = mid;
      else hi = mid - 1;
    }
    Blk2 k;
    k.plane = lo;
    k.hb = hbase[lo];
    k.b = (int)((key - k.hb) / (uint32_t)nbin);
    k.nb = (int)hbase[P + 1 + lo];
    k.rs = bounds[lo] + (uint32_t)k.b * rb;
    k.re = min(bounds[lo + 1], k.rs + rb);
    sb = k;
  }
  __syncthreads();
  return sb;
}

__device__ inline void rec_tiles(const DevGeom& g, int pu, int pv, int* tx0, int* ty0, int* ntx_t, int* nty_t,
                                 int* xs, int* ys) {
  const int half = g.support / 2;
  *xs = pu - half;
  *ys = pv - half;
  // xs, ys >= 0 (the footprint lies on the grid), so shifts divide exactly
  *tx0 = *xs >> g.tshift;
  *ty0 = *ys >> g.tshift;
  *ntx_t = ((*xs + g.support - 1) >> g.tshift) - *tx0 + 1;
  *nty_t = ((*ys + g.support - 1) >> g.tshift) - *ty0 + 1;
}

template <typename RT>
__global__ void k_p2_count(DevGeom g, const RT* __restrict__ recs, const uint32_t* __restrict__ bounds,
                           const uint32_t* __restrict__ hbase, uint32_t rb, uint32_t* __restrict__ h2) {
  using O = RecOps<RT>;
  extern __shared__ uint32_t lh[];
  const int tpp = g.nchan * g.ntx * g.ntx;
  const Blk2 bk = p2_block(bounds, hbase, g.nplanes, 2 * tpp, rb);
  for (int i = threadIdx.x; i < 2 * tpp; i += blockDim.x) lh[i] = 0u;
  __syncthreads();
  for (uint32_t k0 = bk.rs + threadIdx.x; k0 < bk.re; k0 += K2_UNROLL * blockDim.x) {
    RT rr[K2_UNROLL];
#pragma unroll
    for (int u = 0; u < K2_UNROLL; ++u) {
      const uint32_t k = k0 + u * blockDim.x;
      if (k < bk.re) rr[u] = recs[k];
    }
#pragma unroll
    for (int u = 0; u < K2_UNROLL; ++u) {
      if (k0 + u * blockDim.x >= bk.re) break;
      const RT& r = rr[u];
      int tx0, ty0, ntxt, nty, xs, ys;
      rec_tiles(g, O::pu(r), O::pv(r), &tx0, &ty0, &ntxt, &nty, &xs, &ys);
      const int chan = O::chan(r);
      const bool last = O::last(r);
      for (int dy = 0; dy < nty; ++dy)
        for (int dx = 0; dx < ntxt; ++dx) {
          const int tl = (chan * g.ntx + ty0 + dy) * g.ntx + tx0 + dx;
          const int kind = (dx == 0 && dy == 0 && last) ? 0 : 1;
          atomicAdd(&lh[2 * tl + kind], 1u);
        }
    }
  }
  __syncthreads();
  uint32_t* hrow = h2 + bk.hb + (size_t)bk.b * 2 * tpp;
  for (int i = threadIdx.x; i < 2 * tpp; i += blockDim.x) hrow[i] = lh[i];
}

template <typename RT>
__global__ void k_p2_scatter(DevGeom g, const RT* __restrict__ recs, const uint32_t* __restrict__ bounds,
                             const uint32_t* __restrict__ hbase, uint32_t rb, const uint32_t* __restrict__ s2,
                             uint2* __restrict__ entries, uint16_t* __restrict__ entry_z) {
  using O = RecOps<RT>;
  extern __shared__ uint32_t cur[];
  const int tpp = g.nchan * g.ntx * g.ntx;
  const Blk2 bk = p2_block(bounds, hbase, g.nplanes, 2 * tpp, rb);
  const uint32_t* srow = s2 + bk.hb + (size_t)bk.b * 2 * tpp;
  for (int i = threadIdx.x; i < 2 * tpp; i += blockDim.x) cur[i] = srow[i];
  __syncthreads();
  for (uint32_t k0 = bk.rs + threadIdx.x; k0 < bk.re; k0 += K2_UNROLL * blockDim.x) {
    RT rr[K2_UNROLL];
#pragma unroll
    for (int u = 0; u < K2_UNROLL; ++u) {
      const uint32_t k = k0 + u * blockDim.x;
      if (k < bk.re) rr[u] = recs[k];
    }
#pragma unroll
    for (int u = 0; u < K2_UNROLL; ++u) {
      if (k0 + u * blockDim.x >= bk.re) break;
      const RT& r = rr[u];
      int tx0, ty0, ntxt, nty, xs, ys;
      rec_tiles(g, O::pu(r), O::pv(r), &tx0, &ty0, &ntxt, &nty, &xs, &ys);
      const int chan = O::chan(r);
      const bool last = O::last(r);
      const int xo = O::xoff(r), yo = O::yoff(r);
      for (int dy = 0; dy < nty; ++dy)
        for (int dx = 0; dx < ntxt; ++dx) {
          const int tx = tx0 + dx, ty = ty0 + dy;
          const int tl = (chan * g.ntx + ty) * g.ntx + tx;
          const bool start = dx == 0 && dy == 0;
          const int kind = (start && last) ? 0 : 1;
          const uint32_t pos = atomicAdd(&cur[2 * tl + kind], 1u);
          entries[pos] = make_uint2(r.row, pack_entry(xs - tx * g.tile + g.support, ys - ty * g.tile + g.support,
                                                      xo, yo, start, kind == 0));
          if (entry_z) entry_z[pos] = (uint16_t)O::z(r);
        }
    }
  }
}

// The two kernels below split the pass-2 blocks of a plane (blockIdx.y) into ns consecutive
// slices (a power of two <= 16, chosen on the host from the largest block count of a plane) and
// give each (group of BPT consecutive bins, slice) one thread: a large plane (hundreds of
// blocks at 2048^2) is then no longer one long chain of loads per bin, and the many bins of
// planes with few blocks (128^2 tiles at 8192^2, 4 bins a thread) do not each take a thread of
// their own.  16 loads in flight per thread.
constexpr int H2_THREADS = 1024;

// Sums over the blocks [b0, b1) of this thread's slice of h[b * nbin + u], u < BPT (bins past
// nbin read as 0), H2_BATCH blocks of loads in flight.
template <int BPT, int H2_BATCH = 16 / BPT>
__device__ inline void h2_slice_sum(const uint32_t* __restrict__ h, int nbin, int bin0, uint32_t b0, uint32_t b1,
                                    uint32_t (&sum)[BPT]) {
#pragma unroll
  for (int u = 0; u < BPT; ++u) sum[u] = 0u;
  for (uint32_t b = b0; b < b1; b += H2_BATCH) {
    uint32_t v[H2_BATCH][BPT];
#pragma unroll
    for (int k = 0; k < H2_BATCH; ++k)
#pragma unroll
      for (int u = 0; u < BPT; ++u)
        v[k][u] = (b + k < b1 && bin0 + u < nbin) ? h[(size_t)(b + k) * nbin + bin0 + u] : 0u;
#pragma unroll
    for (int k = 0; k < H2_BATCH; ++k)
#pragma unroll
      for (int u = 0; u < BPT; ++u) sum[u] += v[k][u];
  }
}

// c2[plane * nbin + bin] = sum over the plane's pass-2 blocks of h2[plane base + block * nbin + bin]
// (nbin = 2 tpp), then c2[P * nbin] = 0, so that the exclusive scan of c2 ends with the total.
template <int BPT>
__global__ __launch_bounds__(H2_THREADS) void k_h2_colsum(const uint32_t* __restrict__ h2,
                                                          const uint32_t* __restrict__ hbase,
                                                          const uint32_t* __restrict__ nbp, int P, int nbin, int ns,
                                                          uint32_t* __restrict__ c2) {
  __shared__ uint32_t part[BPT][H2_THREADS];
  const int p = blockIdx.y, bg = H2_THREADS / ns;
  const int bl = threadIdx.x % bg, j = threadIdx.x / bg;
  const int bin0 = (blockIdx.x * bg + bl) * BPT;
  const uint32_t nb = nbp[p], len = (nb + ns - 1) / ns;
  const uint32_t b0 = min(nb, (uint32_t)j * len), b1 = min(nb, b0 + len);
  uint32_t v[BPT];
  h2_slice_sum<BPT>(h2 + hbase[p], nbin, bin0, b0, b1, v);
  if (ns > 1) {
#pragma unroll
    for (int u = 0; u < BPT; ++u) part[u][threadIdx.x] = v[u];
    __syncthreads();
  }
  if (j == 0) {
    for (int k = 1; k < ns; ++k)
#pragma unroll
      for (int u = 0; u < BPT; ++u) v[u] += part[u][k * bg + bl];
#pragma unroll
    for (int u = 0; u < BPT; ++u)
      if (bin0 + u < nbin) c2[(int64_t)p * nbin + bin0 + u] = v[u];
  }
  if (p == 0 && blockIdx.x == 0 && threadIdx.x == 0) c2[(int64_t)P * nbin] = 0u;
}

// Block cursors: s2[plane base + block * nbin + bin] = e2[plane * nbin + bin] + the counts of the
// plane's earlier blocks in the bin (earlier slices from the slice sums, then the slice's own).
template <int BPT>
__global__ __launch_bounds__(H2_THREADS) void k_h2_prefix(const uint32_t* __restrict__ h2,
                                                          const uint32_t* __restrict__ hbase,
                                                          const uint32_t* __restrict__ nbp, int P, int nbin, int ns,
                                                          const uint32_t* __restrict__ e2, uint32_t* __restrict__ s2) {
  __shared__ uint32_t part[BPT][H2_THREADS];
  const int p = blockIdx.y, bg = H2_THREADS / ns;
  const int bl = threadIdx.x % bg, j = threadIdx.x / bg;
  const int bin0 = (blockIdx.x * bg + bl) * BPT;
  const uint32_t nb = nbp[p], len = (nb + ns - 1) / ns;
  const uint32_t b0 = min(nb, (uint32_t)j * len), b1 = min(nb, b0 + len);
  const uint32_t* h = h2 + hbase[p];
  uint32_t* sp = s2 + hbase[p];
  uint32_t run[BPT];
  if (ns > 1) {
    h2_slice_sum<BPT>(h, nbin, bin0, b0, b1, run);
#pragma unroll
    for (int u = 0; u < BPT; ++u) part[u][threadIdx.x] = run[u];
    __syncthreads();
  }
  if (bin0 >= nbin) return;
#pragma unroll
  for (int u = 0; u < BPT; ++u) run[u] = bin0 + u < nbin ? e2[(int64_t)p * nbin + bin0 + u] : 0u;
  for (int k = 0; k < j; ++k)
#pragma unroll
    for (int u = 0; u < BPT; ++u) run[u] += part[u][k * bg + bl];
  constexpr int H2_BATCH = 16 / BPT;
  for (uint32_t b = b0; b < b1; b += H2_BATCH) {
    uint32_t v[H2_BATCH][BPT];
#pragma unroll
    for (int k = 0; k < H2_BATCH; ++k)
#pragma unroll
      for (int u = 0; u < BPT; ++u)
        v[k][u] = (b + k < b1 && bin0 + u < nbin) ? h[(size_t)(b + k) * nbin + bin0 + u] : 0u;
#pragma unroll
    for (int k = 0; k < H2_BATCH; ++k)
#pragma unroll
      for (int u = 0; u < BPT; ++u) {
        if (b + k < b1 && bin0 + u < nbin) sp[(size_t)(b + k) * nbin + bin0 + u] = run[u];
        run[u] += v[k][u];
      }
  }
}

// Per-tile totals from the scanned bin sums e2 (bin 2*key is tile key's degrid entries, 2*key + 1
// the rest, 2*key + 2 the next tile's, or the total): off (first entry), cnt (entries), cntd
// (degrid entries).
__global__ void k_p2_keys(const uint32_t* __restrict__ e2, int64_t nt, uint32_t* cnt, uint32_t* cntd,
                          uint32_t* off) {
  const int64_t key = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (key >= nt) return;
  const uint32_t a = e2[2 * key], d = e2[2 * key + 1], e = e2[2 * key + 2];
  off[key] = a;
  cntd[key] = d - a;
  cnt[key] = e - a;
}

// Entries of every (tile, kind) segment grouped by footprint origin: a counting sort on the
// origin (ly, lx) (W = T + S values a side, NK = W^2 keys) from the entry array into a second
// one, so that the gridder finds the entries of one origin next to each other and adds their
// taps in registers before one LDS add (rows of the short baselines fall on the same pixels:
// 29.7 M rows of the 2048^2 w-stack workload have 10.1 M distinct (slab, pixel) origins, and a
// central tile holds hundreds of entries per origin).  k_sort_copytiny copies segments of fewer
// than SORT_MIN entries as they are, k_sort_lists lists the others: up to SORT_MID entries one work group
// sorts in LDS (k_sort_mid); larger ones are cut into blocks of SORT_BB entries, counted per
// block (k_sort_bcount, block-major histograms), the blocks' counts summed per key
// (k_sort_bksum), turned into every block's cursors (k_sort_bcursor) and scattered
// (k_sort_bscatter).
// (Measured on the 2048^2 w-stack workload: a copy instead of the LDS sort for the middle segments
// 7.04 -> 7.17 ms per step, SORT_MIN 256 or 1024 no different from 64, SORT_BB 4096 slower than
// 8192.)
constexpr int SORT_MIN = 64, SORT_MID = 4096, SORT_BB = 8192, SORT_MAXKEYS = 16384;
// large segments of more than SORT_HUGE blocks get 16 slices of threads over their blocks in the
// key passes, the others one
constexpr int SORT_HUGE = 16;

struct SortSeg {
  uint32_t start, m, blk0, nblk;
};

// Segment lists of the sort without global atomics: segment q = 2 t + kind of tile t, one
// thread a segment.  k_sort_lists_a: each work group's numbers of middle segments, large
// segments and their blocks (wg[3][G]); k_sort_lists_b: the work groups before it summed, its
// threads' offsets by block scans, the lists written.  cnt: [0] middle segments, [1] large
// segments, [2] their blocks.
constexpr int SORT_LIST_THREADS = 1024;
__device__ inline void sort_seg(const uint32_t* __restrict__ off, const uint32_t* __restrict__ tcnt,
                                const uint32_t* __restrict__ tcntd, int64_t q, uint32_t* s, uint32_t* m) {
  const int64_t t = q >> 1;
  const uint32_t nd = tcntd[t];
  *s = (q & 1) ? off[t] + nd : off[t];
  *m = (q & 1) ? tcnt[t] - nd : nd;
}
// exclusive scan over the work group (every thread calls it); *total gets the sum
__device__ inline uint32_t block_excl_u32(uint32_t v, uint32_t* wtot, uint32_t* total) {
  const int lane = __lane_id(), wv = threadIdx.x >> 6;
  uint32_t inc = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(inc, o);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wtot[wv] = inc;
  __syncthreads();
  uint32_t base = 0, tot = 0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
    if (k < wv) base += wtot[k];
    tot += wtot[k];
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}
__device__ inline void sort_class(uint32_t m, uint32_t* nm, uint32_t* nb, uint32_t* nbb, uint32_t* nh) {
  *nm = (m >= (uint32_t)SORT_MIN && m <= (uint32_t)SORT_MID) ? 1u : 0u;
  *nb = m > (uint32_t)SORT_MID ? 1u : 0u;
  *nbb = m > (uint32_t)SORT_MID ? (m + SORT_BB - 1) / SORT_BB : 0u;
  *nh = *nbb > (uint32_t)SORT_HUGE ? 1u : 0u;
}
__global__ __launch_bounds__(SORT_LIST_THREADS) void k_sort_lists_a(const uint32_t* __restrict__ off,
                                                                    const uint32_t* __restrict__ tcnt,
                                                                    const uint32_t* __restrict__ tcntd, int64_t nt,
                                                                    uint32_t* __restrict__ wg) {
  __shared__ uint32_t wtot[SORT_LIST_THREADS / 64];
  const int64_t q = blockIdx.x * (int64_t)SORT_LIST_THREADS + threadIdx.x;
  uint32_t s = 0, m = 0;
  if (q < 2 * nt) sort_seg(off, tcnt, tcntd, q, &s, &m);
  uint32_t a, b, c, h, ta, tb, tc, th;
  sort_class(m, &a, &b, &c, &h);
  block_excl_u32(a, wtot, &ta);
  block_excl_u32(b, wtot, &tb);
  block_excl_u32(c, wtot, &tc);
  block_excl_u32(h, wtot, &th);
  if (threadIdx.x == 0) {
    wg[blockIdx.x] = ta;
    wg[gridDim.x + blockIdx.x] = tb;
    wg[2 * gridDim.x + blockIdx.x] = tc;
    wg[3 * gridDim.x + blockIdx.x] = th;
  }
}
// The four per-work-group counts scanned over the G work groups of the list passes (one work
// group): wg[r][g] becomes their exclusive prefix, cnt[r] their total.
__global__ __launch_bounds__(SORT_LIST_THREADS) void k_sort_lists_scan(uint32_t* __restrict__ wg, int G,
                                                                       uint32_t* __restrict__ cnt) {
  __shared__ uint32_t wtot[SORT_LIST_THREADS / 64];
  const int per = (G + SORT_LIST_THREADS - 1) / SORT_LIST_THREADS;
  const int g0 = threadIdx.x * per;
  for (int r = 0; r < 4; ++r) {
    uint32_t* row = wg + (size_t)r * G;
    uint32_t loc = 0;
    for (int j = 0; j < per; ++j)
      if (g0 + j < G) loc += row[g0 + j];
    uint32_t total;
    uint32_t run = block_excl_u32(loc, wtot, &total);
    for (int j = 0; j < per; ++j)
      if (g0 + j < G) {
        const uint32_t v = row[g0 + j];
        row[g0 + j] = run;
        run += v;
      }
    if (threadIdx.x == 0) cnt[r] = total;
  }
}
__global__ __launch_bounds__(SORT_LIST_THREADS) void k_sort_lists_b(const uint32_t* __restrict__ off,
                                                                    const uint32_t* __restrict__ tcnt,
                                                                    const uint32_t* __restrict__ tcntd, int64_t nt,
                                                                    const uint32_t* __restrict__ wg,
                                                                    uint2* __restrict__ mid, SortSeg* __restrict__ big,
                                                                    uint2* __restrict__ bigblk,
                                                                    uint32_t* __restrict__ segidx,
                                                                    uint32_t* __restrict__ cnt) {
  __shared__ uint32_t wtot[SORT_LIST_THREADS / 64];
  __shared__ uint32_t base[4];
  if (threadIdx.x < 4) base[threadIdx.x] = wg[threadIdx.x * gridDim.x + blockIdx.x];  // scanned
  const int64_t q = blockIdx.x * (int64_t)SORT_LIST_THREADS + threadIdx.x;
  uint32_t s = 0, m = 0;
  if (q < 2 * nt) sort_seg(off, tcnt, tcntd, q, &s, &m);
  uint32_t a, b, c, h, ta, tb, tc, th;
  sort_class(m, &a, &b, &c, &h);
  uint32_t oa = block_excl_u32(a, wtot, &ta);
  uint32_t ob = block_excl_u32(b, wtot, &tb);
  uint32_t oc = block_excl_u32(c, wtot, &tc);
  uint32_t oh = block_excl_u32(h, wtot, &th);
  oa += base[0];
  ob += base[1];
  oc += base[2];
  oh += base[3];
  if (a) mid[oa] = make_uint2(s, m);
  if (b) {
    big[ob] = SortSeg{s, m, oc, c};
    for (uint32_t k = 0; k < c; ++k) bigblk[oc + k] = make_uint2(ob, k);
    // segment index lists for the key passes: the huge ones from the front, the others from the
    // back (cnt[1] - 1 down)
    const uint32_t total_big = cnt[1];
    if (h) segidx[oh] = ob;
    else segidx[total_big - 1 - (ob - oh)] = ob;
  }
}

// Segments of fewer than SORT_MIN entries copied as they are: 16 lanes a segment.
__global__ __launch_bounds__(256) void k_sort_copytiny(const uint2* __restrict__ in, uint2* __restrict__ out,
                                                       const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ tcnt,
                                                       const uint32_t* __restrict__ tcntd, int64_t nt) {
  const int64_t q = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 4;
  const int l = threadIdx.x & 15;
  if (q >= 2 * nt) return;
  uint32_t s, m;
  sort_seg(off, tcnt, tcntd, q, &s, &m);
  if (m >= (uint32_t)SORT_MIN) return;
  for (uint32_t k = l; k < m; k += 16) out[s + k] = in[s + k];
}

__device__ inline int sort_key(uint32_t pk, int W) { return entry_ly(pk) * W + entry_lx(pk); }

// Exclusive scan of hist[0, nkeys) in place by a work group of NTH threads (runs of consecutive
// keys a thread, wavefront scans, wavefront totals).
template <int NTH>
__device__ inline void block_scan_keys(uint32_t* hist, int nkeys, uint32_t* wtot) {
  const int per = (nkeys + NTH - 1) / NTH;
  const int k0 = threadIdx.x * per;
  const int lane = __lane_id(), wv = threadIdx.x >> 6;
  uint32_t loc = 0;
  for (int j = 0; j < per; ++j)
    if (k0 + j < nkeys) loc += hist[k0 + j];
  uint32_t inc = loc;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(inc, o);
    if (lane >= o) inc += v;
  }
  if (lane == 63) wtot[wv] = inc;
  __syncthreads();
  uint32_t base = 0;
  for (int k = 0; k < wv; ++k) base += wtot[k];
  uint32_t run = base + inc - loc;
  for (int j = 0; j < per; ++j)
    if (k0 + j < nkeys) {
      const uint32_t h = hist[k0 + j];
      hist[k0 + j] = run;
      run += h;
    }
  __syncthreads();
}

// One middle segment (at most SORT_MID entries) per work group: entries held in registers,
// histogram and cursors in LDS.
constexpr int SORT_MID_THREADS = 256;
__global__ __launch_bounds__(SORT_MID_THREADS) void k_sort_mid(const uint2* __restrict__ in, uint2* __restrict__ out,
                                                               const uint2* __restrict__ mid, int W) {
  extern __shared__ uint32_t hist[];  // [W * W]
  __shared__ uint32_t wtot[SORT_MID_THREADS / 64];
  constexpr int PER = SORT_MID / SORT_MID_THREADS;
  const uint2 sg = mid[blockIdx.x];
  const uint32_t s = sg.x, m = sg.y;
  const int nkeys = W * W;
  for (int i = threadIdx.x; i < nkeys; i += SORT_MID_THREADS) hist[i] = 0u;
  uint2 e[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t i = threadIdx.x + k * SORT_MID_THREADS;
    e[k] = i < m ? in[s + i] : make_uint2(0u, 0u);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; ++k)
    if (threadIdx.x + k * SORT_MID_THREADS < m) atomicAdd(&hist[sort_key(e[k].y, W)], 1u);
  __syncthreads();
  block_scan_keys<SORT_MID_THREADS>(hist, nkeys, wtot);
#pragma unroll
  for (int k = 0; k < PER; ++k)
    if (threadIdx.x + k * SORT_MID_THREADS < m) out[s + atomicAdd(&hist[sort_key(e[k].y, W)], 1u)] = e[k];
}

// Large segments: block histograms (block-major, NK words a block).
constexpr int SORT_BIG_THREADS = 512;
__global__ __launch_bounds__(SORT_BIG_THREADS) void k_sort_bcount(const uint2* __restrict__ in,
                                                                  const SortSeg* __restrict__ big,
                                                                  const uint2* __restrict__ bigblk, int W,
                                                                  uint32_t* __restrict__ h3) {
  extern __shared__ uint32_t hist[];
  const uint2 bb = bigblk[blockIdx.x];
  const SortSeg sg = big[bb.x];
  const uint32_t s = sg.start + bb.y * SORT_BB, m = min((uint32_t)SORT_BB, sg.m - bb.y * SORT_BB);
  const int nkeys = W * W;
  for (int i = threadIdx.x; i < nkeys; i += SORT_BIG_THREADS) hist[i] = 0u;
  __syncthreads();
  for (uint32_t i0 = threadIdx.x; i0 < m; i0 += 4 * SORT_BIG_THREADS) {
    uint32_t pk[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t i = i0 + u * SORT_BIG_THREADS;
      pk[u] = i < m ? in[s + i].y : 0u;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i0 + u * SORT_BIG_THREADS < m) atomicAdd(&hist[sort_key(pk[u], W)], 1u);
  }
  __syncthreads();
  uint32_t* hrow = h3 + (size_t)blockIdx.x * nkeys;
  for (int i = threadIdx.x; i < nkeys; i += SORT_BIG_THREADS) hrow[i] = hist[i];
}

// Per large segment (segidx[blockIdx.y]) and KCH = 1024 / NS keys (blockIdx.x): the key's count
// over the segment's blocks, ksum[segment * NK + key]; the blocks split into NS slices of threads.
template <int NS>
__global__ __launch_bounds__(1024) void k_sort_bksum(const SortSeg* __restrict__ big,
                                                     const uint32_t* __restrict__ segidx, int W,
                                                     const uint32_t* __restrict__ h3, uint32_t* __restrict__ ksum) {
  constexpr int SORT_KCH = 1024 / NS;
  __shared__ uint32_t part[NS][SORT_KCH];
  const uint32_t gi = segidx[blockIdx.y];
  const SortSeg sg = big[gi];
  const int nkeys = W * W;
  const int kl = threadIdx.x % SORT_KCH, sl = threadIdx.x / SORT_KCH;
  const int k = blockIdx.x * SORT_KCH + kl;
  const uint32_t len = (sg.nblk + NS - 1) / NS, b0 = min(sg.nblk, sl * len), b1 = min(sg.nblk, b0 + len);
  const uint32_t* hk = h3 + (size_t)sg.blk0 * nkeys + k;
  uint32_t t = 0;
  if (k < nkeys)
    for (uint32_t b = b0; b < b1; b += 4) {
      uint32_t v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = b + u < b1 ? hk[(size_t)(b + u) * nkeys] : 0u;
      t += (v[0] + v[1]) + (v[2] + v[3]);
    }
  part[sl][kl] = t;
  __syncthreads();
  if (sl == 0 && k < nkeys) {
    uint32_t u = 0;
#pragma unroll
    for (int q = 0; q < NS; ++q) u += part[q][kl];
    ksum[(size_t)gi * nkeys + k] = u;
  }
}

// Every block's cursors in place of its counts: segment start + the counts of the segment's
// smaller keys + the key's counts in the earlier blocks.  Same grid as k_sort_bksum.
template <int NS>
__global__ __launch_bounds__(1024) void k_sort_bcursor(const SortSeg* __restrict__ big,
                                                       const uint32_t* __restrict__ segidx, int W,
                                                       const uint32_t* __restrict__ ksum, uint32_t* __restrict__ h3) {
  constexpr int SORT_KCH = 1024 / NS;
  __shared__ uint32_t wtot[1024 / 64];
  __shared__ uint32_t keyb[SORT_KCH];
  __shared__ uint32_t part[NS][SORT_KCH];
  const uint32_t gi = segidx[blockIdx.y];
  const SortSeg sg = big[gi];
  const int nkeys = W * W;
  const uint32_t* ks = ksum + (size_t)gi * nkeys;
  const int kbase = blockIdx.x * SORT_KCH;
  const int kl = threadIdx.x % SORT_KCH, sl = threadIdx.x / SORT_KCH;
  const int k = kbase + kl;
  // the counts of the keys below this group's first key, then the group's own keys scanned
  uint32_t below = 0;
  for (int q = threadIdx.x; q < kbase; q += 1024) below += ks[q];
  uint32_t tbelow, tk;
  block_excl_u32(below, wtot, &tbelow);
  const uint32_t own = (sl == 0 && k < nkeys) ? ks[k] : 0u;
  const uint32_t ex = block_excl_u32(own, wtot, &tk);
  if (sl == 0) keyb[kl] = tbelow + ex;
  // this slice's blocks: counts summed (four loads in flight), then the slices' offsets
  const uint32_t len = (sg.nblk + NS - 1) / NS, b0 = min(sg.nblk, sl * len), b1 = min(sg.nblk, b0 + len);
  const uint32_t* hk = h3 + (size_t)sg.blk0 * nkeys + k;
  uint32_t t = 0;
  if (k < nkeys)
    for (uint32_t b = b0; b < b1; b += 4) {
      uint32_t v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = b + u < b1 ? hk[(size_t)(b + u) * nkeys] : 0u;
      t += (v[0] + v[1]) + (v[2] + v[3]);
    }
  part[sl][kl] = t;
  __syncthreads();
  if (k >= nkeys) return;
  uint32_t run = sg.start + keyb[kl];
  for (int q = 0; q < sl; ++q) run += part[q][kl];
  uint32_t* hw = h3 + (size_t)sg.blk0 * nkeys + k;
  for (uint32_t b = b0; b < b1; b += 4) {
    uint32_t v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = b + u < b1 ? hw[(size_t)(b + u) * nkeys] : 0u;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (b + u < b1) {
        hw[(size_t)(b + u) * nkeys] = run;
        run += v[u];
      }
  }
}

// A block's entries sorted by key in LDS first, then written in that order: entries of one key
// go to consecutive places (its global cursor on), so the writes come in runs rather than one
// scattered 8-byte store each.  LDS: [NK] cursor - local base, [NK] local cursors, [SORT_BB] entries.
constexpr int SORT_SCAT_THREADS = 1024;
__global__ __launch_bounds__(SORT_SCAT_THREADS) void k_sort_bscatter(const uint2* __restrict__ in,
                                                                     uint2* __restrict__ out,
                                                                     const SortSeg* __restrict__ big,
                                                                     const uint2* __restrict__ bigblk, int W,
                                                                     const uint32_t* __restrict__ h3) {
  extern __shared__ uint32_t sm[];
  __shared__ uint32_t wtot[SORT_SCAT_THREADS / 64];
  constexpr int PER = SORT_BB / SORT_SCAT_THREADS;
  const int nkeys = W * W;
  uint32_t* delta = sm;
  uint32_t* lcur = sm + nkeys;
  uint2* buf = reinterpret_cast<uint2*>(sm + 2 * nkeys);
  const uint2 bb = bigblk[blockIdx.x];
  const SortSeg sg = big[bb.x];
  const uint32_t s = sg.start + bb.y * SORT_BB, m = min((uint32_t)SORT_BB, sg.m - bb.y * SORT_BB);
  for (int i = threadIdx.x; i < nkeys; i += SORT_SCAT_THREADS) lcur[i] = 0u;
  uint2 e[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t i = threadIdx.x + k * SORT_SCAT_THREADS;
    e[k] = i < m ? in[s + i] : make_uint2(0u, 0u);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; ++k)
    if (threadIdx.x + k * SORT_SCAT_THREADS < m) atomicAdd(&lcur[sort_key(e[k].y, W)], 1u);
  __syncthreads();
  block_scan_keys<SORT_SCAT_THREADS>(lcur, nkeys, wtot);
  const uint32_t* hrow = h3 + (size_t)blockIdx.x * nkeys;
  for (int i = threadIdx.x; i < nkeys; i += SORT_SCAT_THREADS) delta[i] = hrow[i] - lcur[i];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; ++k)
    if (threadIdx.x + k * SORT_SCAT_THREADS < m) buf[atomicAdd(&lcur[sort_key(e[k].y, W)], 1u)] = e[k];
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < m; i += SORT_SCAT_THREADS) {
    const uint2 v = buf[i];
    out[delta[sort_key(v.y, W)] + i] = v;
  }
}

// Chunk counts per tile; tiles with several grid chunks also get partial-tile slots.
__global__ void k_chunks(const uint32_t* cnt, const uint32_t* cntd, int ntiles, int chunk, int sparse, uint32_t* nch,
                         uint32_t* nchd, uint32_t* nslot) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const uint32_t c = cnt[t];
  // empty tiles still write zeros, unless the grids are sparse
  const uint32_t ng = c == 0 ? (sparse ? 0u : 1u) : (c + chunk - 1) / chunk;
  nch[t] = ng;
  nchd[t] = (cntd[t] + chunk - 1) / chunk;
  nslot[t] = ng > 1 ? ng : 0u;
}

// Occupancy masks of sparse grids (pc = plane * nchan + chan, ntx <= 32):
// tilemask[pc * ntx + tx] bit ty: tile (pc, ty, tx) has entries;
// colmask[pc] bit tx: tile column tx of pc has a tile with entries.
// k_tilemask: one thread per (pc, tx); k_colmask: one thread per pc.
__global__ void k_tilemask(const uint32_t* cnt, int npc, int ntx, uint32_t* tilemask) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)npc * ntx) return;
  const int tx = (int)(i % ntx);
  const uint32_t* c = cnt + (i / ntx) * ntx * ntx + tx;
  uint32_t m = 0;
  for (int ty = 0; ty < ntx; ++ty)
    if (c[(size_t)ty * ntx] != 0u) m |= 1u << ty;
  tilemask[i] = m;
}

__global__ void k_colmask(const uint32_t* tilemask, int npc, int ntx, uint32_t* colmask) {
  const int pc = blockIdx.x * blockDim.x + threadIdx.x;
  if (pc >= npc) return;
  uint32_t col = 0;
  for (int tx = 0; tx < ntx; ++tx)
    if (tilemask[(size_t)pc * ntx + tx]) col |= 1u << tx;
  colmask[pc] = col;
}

// Column-band occupancy (arlmi_plan_or_colmask): one thread per (pc, tile column tx) sets bit
// tx >> shift of colmask[pc] when the tile column holds an entry.
__global__ void k_or_colmask(const uint32_t* __restrict__ cnt, int npc, int ntx, int shift, uint32_t* colmask) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)npc * ntx) return;
  const int tx = (int)(i % ntx);
  const int64_t pc = i / ntx;
  const uint32_t* c = cnt + pc * ntx * ntx + tx;
  bool any = false;
  for (int ty = 0; ty < ntx; ++ty) any |= c[(size_t)ty * ntx] != 0u;
  if (any) atomicOr(&colmask[pc], 1u << (tx >> shift));
}

// Work items: grid items {tile, start, count, slot+1 or 0}, degrid items {tile, start, count, 0};
// tiles split over several grid items are listed with their first partial slot.
// Largest t in [0, n) with off[t] <= i (off non-decreasing, off[0] = 0 <= i).
__device__ inline int owner(const uint32_t* __restrict__ off, int n, uint32_t i) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// One thread per work item (a tile with many entries has many items, so a thread per tile
// would serialise them): gridding item i belongs to the tile t with choff[t] <= i < choff[t+1],
// likewise for the degridding items.
__global__ void k_items(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ cntd,
                        const uint32_t* __restrict__ off, const uint32_t* __restrict__ choff,
                        const uint32_t* __restrict__ choffd, const uint32_t* __restrict__ slotoff, int ntiles,
                        int chunk, uint32_t ngi, uint32_t ndi, int4* gitems, int4* ditems, int4* multi,
                        uint32_t* misc) {
  __shared__ uint32_t lcount, lbase;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (threadIdx.x == 0) lcount = 0u;
  __syncthreads();
  int t = 0;
  uint32_t k = 0, ng = 0;
  if (i < ngi) {
    t = owner(choff, ntiles, i);
    k = i - choff[t];
    ng = choff[t + 1] - choff[t];
  }
  // split tiles take consecutive slots of `multi`: an LDS count per block and one device add
  // per block
  const bool first_split = i < ngi && k == 0 && ng > 1;
  const uint32_t li = first_split ? atomicAdd(&lcount, 1u) : 0u;
  __syncthreads();
  if (threadIdx.x == 0) lbase = lcount ? atomicAdd(&misc[4], lcount) : 0u;
  __syncthreads();
  if (i < ngi) {
    const uint32_t c = cnt[t];
    const uint32_t s = k * chunk;
    const uint32_t len = c > s ? min((uint32_t)chunk, c - s) : 0u;
    gitems[i] = make_int4(t, (int)(off[t] + s), (int)len, ng > 1 ? (int)(slotoff[t] + k + 1) : 0);
    if (first_split) multi[lbase + li] = make_int4(t, (int)slotoff[t], (int)ng, 0);
  }
  if (i < ndi) {
    const int td = owner(choffd, ntiles, i);
    const uint32_t kd = i - choffd[td];
    const uint32_t s = kd * chunk;
    ditems[i] = make_int4(td, (int)(off[td] + s), (int)min((uint32_t)chunk, cntd[td] - s), 0);
  }
}

// ---------------------------------------------------------------------------
// exclusive scan of uint32 (three-phase: block scans, scan of block sums, add)
// ---------------------------------------------------------------------------

constexpr int SCAN_T = 1024;                 // threads per scan block
constexpr int SCAN_V = 4;                    // consecutive values per thread
constexpr int SCAN_BLOCK = SCAN_T * SCAN_V;  // values per scan block

// Exclusive scan of one block of SCAN_BLOCK values: per-thread sums, a shuffle scan inside each
// wavefront, then a scan of the wavefront totals; bsum[block] = the block's total.
__global__ __launch_bounds__(SCAN_T) void k_scan_block(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* bsum) {
  __shared__ uint32_t wsum[SCAN_T / 64];
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  constexpr int NW = SCAN_T / 64;
  const int64_t base = blockIdx.x * (int64_t)SCAN_BLOCK + (int64_t)threadIdx.x * SCAN_V;
  uint32_t v[SCAN_V];
  uint32_t t = 0;
#pragma unroll
  for (int j = 0; j < SCAN_V; ++j) {
    v[j] = base + j < n ? in[base + j] : 0u;
    t += v[j];
  }
  uint32_t x = t;  // inclusive scan of the thread sums within the wavefront
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  if (wave == 0) {
    uint32_t ws = lane < NW ? wsum[lane] : 0u;
#pragma unroll
    for (int o = 1; o < NW; o <<= 1) {
      const uint32_t y = __shfl_up(ws, o);
      if (lane >= o) ws += y;
    }
    if (lane < NW) wsum[lane] = ws;
  }
  __syncthreads();
  uint32_t off = (wave ? wsum[wave - 1] : 0u) + x - t;
#pragma unroll
  for (int j = 0; j < SCAN_V; ++j) {
    if (base + j < n) out[base + j] = off;
    off += v[j];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = wsum[NW - 1];
}

// Adds each block's offset; boff[nb] is the grand total, stored at out[n].
__global__ __launch_bounds__(SCAN_T) void k_scan_add(uint32_t* out, int64_t n, const uint32_t* boff, int64_t nb) {
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = boff[nb];
  const uint32_t add = boff[blockIdx.x];
  const int64_t base = blockIdx.x * (int64_t)SCAN_BLOCK + (int64_t)threadIdx.x * SCAN_V;
#pragma unroll
  for (int j = 0; j < SCAN_V; ++j)
    if (base + j < n) out[base + j] += add;
}

size_t scan_tmp_words(int64_t n) {
  size_t words = 0;
  while (n > 1) {
    int64_t nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    words += (size_t)nb + (size_t)nb + 1;
    n = nb;
  }
  return words + 4;
}

int exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp, hipStream_t s) {
  if (n <= 0) {
    ARLMI_HIP(hipMemsetAsync(out, 0, sizeof(uint32_t), s));
    return ARLMI_OK;
  }
  const int64_t nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
  if (nb == 1) {  // one block: its total is out[n]
    ARLMI_LAUNCH("k_scan_block", k_scan_block, dim3(1), dim3(SCAN_T), 0, s, in, out, n, out + n);
    ARLMI_LAUNCH_CHECK();
    return ARLMI_OK;
  }
  uint32_t* bsum = tmp;
  uint32_t* boff = tmp + nb;
  ARLMI_LAUNCH("k_scan_block", k_scan_block, dim3((unsigned)nb), dim3(SCAN_T), 0, s, in, out, n, bsum);
  ARLMI_LAUNCH_CHECK();
  int rc = exclusive_scan_u32(bsum, boff, nb, tmp + 2 * nb + 1, s);
  if (rc) return rc;
  ARLMI_LAUNCH("k_scan_add", k_scan_add, dim3((unsigned)nb), dim3(SCAN_T), 0, s, out, n, boff, nb);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

// convolution_mapping for every row (2-D semantics: grid plane by rounding w), for callers and
// tests that want the indices themselves: out[row] = {pu, pu_off, pv, pv_off, pwg, pwc, pfreq, err}.
__global__ void k_map(DevGeom g, const double* __restrict__ uvw, const double* __restrict__ freq, int64_t n,
                      int32_t* __restrict__ out) {
  const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (row >= n) return;
  const double u = uvw[3 * row], v = uvw[3 * row + 1], w = uvw[3 * row + 2];
  RowMap m = map_row(g, u, v, freq ? freq[row] : 0.0, freq != nullptr);
  uint32_t err = m.err;
  const int pg = map_wplane(w, g.wstep_grid, g.crpix_wg, g.nz, &err);
  const int pz = map_wplane(w, g.zstep, g.crpix_z, g.nz, &err);
  int32_t* o = out + 8 * row;
  const double x = world2pix0(u, g.du, g.crpix_uv), y = world2pix0(v, g.dv, g.crpix_uv);
  o[0] = (int32_t)rint(x);
  o[1] = m.xoff;
  o[2] = (int32_t)rint(y);
  o[3] = m.yoff;
  o[4] = pg;
  o[5] = pz;
  o[6] = m.chan;
  o[7] = (int32_t)err;
}

}  // namespace arlmi

using namespace arlmi;

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

int arlmi::checked_geom(const arlmi_geom* in, DevGeom* g) {
  if (!in) {
    set_error("geometry is NULL");
    return ARLMI_E_ARG;
  }
  const int n = in->npixel;
  if (n < 16 || n > 16384 || (n & (n - 1))) {
    set_error("npixel %d must be a power of two in [16, 16384]", n);
    return ARLMI_E_UNSUPPORTED;
  }
  if (in->support <= 0 || in->support > 64) {
    set_error("support %d must be in [1, 64]", in->support);
    return ARLMI_E_UNSUPPORTED;
  }
  if (in->support & 1) {
    set_error("support %d is odd; ARL's slice-add needs an even support", in->support);
    return ARLMI_E_ODD_SUPPORT;
  }
  if (in->oversampling <= 0 || in->oversampling > 128) {
    set_error("oversampling %d must be in [1, 128]", in->oversampling);
    return ARLMI_E_UNSUPPORTED;
  }
  if (in->nchan <= 0 || in->nz <= 0 || in->nz > 65535 || in->nplanes <= 0) {
    set_error("nchan %d, nz %d, nplanes %d must be positive", in->nchan, in->nz, in->nplanes);
    return ARLMI_E_ARG;
  }
  if (in->nplanes > 4096) {
    set_error("nplanes %d: at most 4096 planes (the per-block slab tables live in LDS)", in->nplanes);
    return ARLMI_E_UNSUPPORTED;
  }
  if (in->mode != ARLMI_MODE_2D && in->mode != ARLMI_MODE_WSTACK) {
    set_error("unknown mode %d", in->mode);
    return ARLMI_E_ARG;
  }
  int tile = in->tile ? in->tile : (n >= 64 ? 64 : n);
  // the pass-2 tile histogram (2 counters per tile of a plane) must fit in LDS
  while (!in->tile && tile < 128 && tile < n && (int64_t)in->nchan * (n / tile) * (n / tile) * 8 > 128 * 1024) tile *= 2;
  if (tile != 16 && tile != 32 && tile != 64 && tile != 128) {
    set_error("tile %d must be 16, 32, 64 or 128", tile);
    return ARLMI_E_UNSUPPORTED;
  }
  if (tile > n) tile = n;
  if ((int64_t)in->nchan * (n / tile) * (n / tile) * 8 > 160 * 1024) {
    set_error("%d channels of %d^2 tiles per plane exceed the tile histogram's LDS", in->nchan, n / tile);
    return ARLMI_E_UNSUPPORTED;
  }
  if (in->nchan > 255) {
    set_error("at most 255 image channels");
    return ARLMI_E_UNSUPPORTED;
  }
  const int64_t ntiles = (int64_t)in->nplanes * in->nchan * (n / tile) * (n / tile);
  if (ntiles > (1ll << 30)) {
    set_error("too many tiles (%lld)", (long long)ntiles);
    return ARLMI_E_UNSUPPORTED;
  }
  g->n = n;
  g->nchan = in->nchan;
  g->support = in->support;
  g->os = in->oversampling;
  g->nz = in->nz;
  g->nplanes = in->nplanes;
  g->mode = in->mode;
  g->tile = tile;
  g->tshift = __builtin_ctz((unsigned)tile);
  g->ntx = n / tile;
  g->du = in->du;
  g->dv = in->dv;
  g->ddu = in->ddu;
  g->ddv = in->ddv;
  g->crpix_uv = in->crpix_uv;
  g->crpix_d = in->crpix_d;
  g->zstep = in->zstep;
  g->crpix_z = in->crpix_z;
  g->wstep_grid = in->wstep_grid;
  g->crpix_wg = in->crpix_wg;
  g->freq0 = in->freq0;
  g->dfreq = in->dfreq;
  g->cellsize = in->cellsize;
  return ARLMI_OK;
}

template <typename T>
static int grow(T** p, int64_t* cap, int64_t need) {
  if (need <= *cap && *p) return ARLMI_OK;
  if (*p) ARLMI_HIP(hipFree(*p));
  *p = nullptr;
  const int64_t c = need + need / 8 + 64;
  ARLMI_HIP(hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (size_t)c));
  *cap = c;
  return ARLMI_OK;
}

static int grow_bytes(void** p, int64_t* cap, int64_t need) {
  if (need <= *cap && *p) return ARLMI_OK;
  if (*p) ARLMI_HIP(hipFree(*p));
  *p = nullptr;
  const int64_t c = need + need / 8 + 256;
  ARLMI_HIP(hipMalloc(p, (size_t)c));
  *cap = c;
  return ARLMI_OK;
}

static void free_plan(arlmi_plan* p) {
  if (!p) return;
  void* ptrs[] = {p->d_boxes, p->d_plane_w, p->d_wsum, p->d_wcnt, p->d_cnt, p->d_tilemask, p->d_colmask, p->d_zeros, p->d_cntd, p->d_off, p->d_nch,
                  p->d_nchd, p->d_nslot, p->d_choff, p->d_choffd, p->d_slotoff, p->d_misc, p->d_bounds,
                  p->d_hbase, p->d_h1, p->d_s1, p->d_h2, p->d_s2, p->d_c2, p->d_scan_tmp, p->d_recs,
                  p->d_entries, p->d_entries_alt, p->d_sort_mid, p->d_sort_big, p->d_sort_bigblk, p->d_sort_h3, p->d_sort_ksum, p->d_sort_wg, p->d_sort_segidx, p->d_sortcnt, p->d_entry_z, p->d_gitems, p->d_ditems, p->d_multi, p->d_partial,
                  p->d_img_part, p->d_det};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (p->h_bounds) (void)hipHostFree(p->h_bounds);
  if (p->ev_bounds) (void)hipEventDestroy(p->ev_bounds);
}

static int plan_alloc_fixed(arlmi_plan* p) {
  const int P = p->g.nplanes;
  const int64_t nt = p->ntiles;
  ARLMI_HIP(hipMalloc(&p->d_boxes, sizeof(double) * (P + 1)));  // centres, then the slab width
  ARLMI_HIP(hipMalloc(&p->d_plane_w, sizeof(double) * P));
  ARLMI_HIP(hipMalloc(&p->d_wsum, sizeof(double) * P * SLAB_BLOCKS));
  ARLMI_HIP(hipMalloc(&p->d_wcnt, sizeof(uint32_t) * P * SLAB_BLOCKS));
  ARLMI_HIP(hipMalloc(&p->d_cnt, sizeof(uint32_t) * nt));
  if (p->flags & ARLMI_FLAG_SPARSE_GRIDS) {
    const int64_t npc = nt / ((int64_t)p->g.ntx * p->g.ntx);
    ARLMI_HIP(hipMalloc(&p->d_tilemask, sizeof(uint32_t) * (size_t)(npc * p->g.ntx)));
    ARLMI_HIP(hipMalloc(&p->d_colmask, sizeof(uint32_t) * (size_t)npc));
  }
  // zeros for the masked-out loads of the transforms (sparse grids, or a caller's band mask)
  ARLMI_HIP(hipMalloc(&p->d_zeros, 1024));
  ARLMI_HIP(hipMemset(p->d_zeros, 0, 1024));
  ARLMI_HIP(hipMalloc(&p->d_cntd, sizeof(uint32_t) * nt));
  ARLMI_HIP(hipMalloc(&p->d_off, sizeof(uint32_t) * nt));
  ARLMI_HIP(hipMalloc(&p->d_nch, sizeof(uint32_t) * nt));
  ARLMI_HIP(hipMalloc(&p->d_nchd, sizeof(uint32_t) * nt));
  ARLMI_HIP(hipMalloc(&p->d_nslot, sizeof(uint32_t) * nt));
  ARLMI_HIP(hipMalloc(&p->d_choff, sizeof(uint32_t) * (nt + 1)));
  ARLMI_HIP(hipMalloc(&p->d_choffd, sizeof(uint32_t) * (nt + 1)));
  ARLMI_HIP(hipMalloc(&p->d_slotoff, sizeof(uint32_t) * (nt + 1)));
  ARLMI_HIP(hipMalloc(&p->d_misc, sizeof(uint32_t) * 8));
  // [P+2] segment bounds, [8] misc copy, [4] totals
  ARLMI_HIP(hipMalloc(&p->d_bounds, sizeof(uint32_t) * (P + 18)));
  ARLMI_HIP(hipMalloc(&p->d_sortcnt, sizeof(uint32_t) * 4));
  ARLMI_HIP(hipHostMalloc(reinterpret_cast<void**>(&p->h_bounds), sizeof(uint32_t) * (P + 18), hipHostMallocDefault));
  ARLMI_HIP(hipEventCreateWithFlags(&p->ev_bounds, hipEventDisableTiming));
  // [P+1] plane bases, then [P] blocks per plane (one copy from the host)
  ARLMI_HIP(hipMalloc(&p->d_hbase, sizeof(uint32_t) * (2 * P + 1)));
  ARLMI_HIP(hipMalloc(&p->d_multi, sizeof(int4) * nt));
  // pass-2 bin sums (2 nt + 1 words) and their scan (2 nt + 2: a one-block scan also writes its
  // total after the last value)
  ARLMI_HIP(hipMalloc(&p->d_c2, sizeof(uint32_t) * (size_t)(4 * nt + 3)));
  if (p->flags & ARLMI_FLAG_DETERMINISTIC)
    ARLMI_HIP(hipMalloc(&p->d_det, sizeof(double) * (size_t)det_words((int64_t)P * p->g.nchan)));
  return ARLMI_OK;
}

static int ensure_scan_tmp(arlmi_plan* p, int64_t n) {
  return grow(&p->d_scan_tmp, &p->cap_scan, (int64_t)scan_tmp_words(n));
}

static int plan_build_steps(arlmi_plan* p, const double* uvw, const double* freq, int64_t nrows,
                            const double* h_boxes, double wslice, hipStream_t s) {
  DevGeom& g = p->g;
  const int P = g.nplanes;
  const int64_t nt = p->ntiles;
  const int tpp = p->tpp;
  if (nrows < 0 || nrows > 0x7fffffffll) {
    set_error("nrows %lld out of range", (long long)nrows);
    return ARLMI_E_ARG;
  }
  if (nrows > 0 && !uvw) {
    set_error("uvw is NULL");
    return ARLMI_E_ARG;
  }
  p->nrows = nrows;
  int rc;
  ARLMI_HIP(hipMemsetAsync(p->d_plane_w, 0, sizeof(double) * P, s));
  std::vector<double> hboxes;  // host staging of the given centres; outlives the copy (the
                               // stream is synchronised before this function returns)
  if (g.mode == ARLMI_MODE_WSTACK) {
    if (h_boxes) {
      hboxes.assign(h_boxes, h_boxes + P);
      hboxes.push_back(wslice);
      ARLMI_HIP(hipMemcpyAsync(p->d_boxes, hboxes.data(), sizeof(double) * (P + 1), hipMemcpyHostToDevice, s));
    } else {
      // ARL's default centres from max |w|, computed on the device (d_wsum holds the partials
      // until the slab sums overwrite it)
      const int wb = (int)std::max<int64_t>(1, std::min<int64_t>((nrows + 1023) / 1024, 1024));
      if (nrows > 0) {
        ARLMI_LAUNCH("k_wmax", k_wmax, dim3((unsigned)wb), dim3(256), 0, s, uvw, nrows, p->d_wsum);
        ARLMI_LAUNCH_CHECK();
      }
      ARLMI_LAUNCH("k_boxes", k_boxes, dim3(1), dim3(256), 0, s, p->d_wsum, nrows > 0 ? wb : 0, P, p->d_boxes);
      ARLMI_LAUNCH_CHECK();
    }
    if (!p->fixed_w.empty()) {  // slab means over all chunks of a stream, given by the caller
      ARLMI_HIP(hipMemcpyAsync(p->d_plane_w, p->fixed_w.data(), sizeof(double) * P, hipMemcpyHostToDevice, s));
    } else {
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((nrows + 1023) / 1024, SLAB_BLOCKS));
    if (nrows > 0) {
      const size_t sh = (2 * sizeof(double) + sizeof(unsigned int)) * P;
      ARLMI_LAUNCH("k_slab_sums", k_slab_sums, dim3((unsigned)blocks), dim3(256), sh, s, uvw, nrows, p->d_boxes, P,
                   p->d_wsum, p->d_wcnt);
      ARLMI_LAUNCH_CHECK();
    } else {
      ARLMI_HIP(hipMemsetAsync(p->d_wsum, 0, sizeof(double) * P, s));
      ARLMI_HIP(hipMemsetAsync(p->d_wcnt, 0, sizeof(uint32_t) * P, s));
    }
    ARLMI_LAUNCH("k_slab_means", k_slab_means, dim3((unsigned)P), dim3(256), 0, s, p->d_wsum, p->d_wcnt, blocks,
                 (const double*)p->d_boxes, p->d_plane_w, P, nullptr, nullptr);
    ARLMI_LAUNCH_CHECK();
    }
  }
  ARLMI_HIP(hipMemsetAsync(p->d_misc, 0, sizeof(uint32_t) * 8, s));

  // pass 1: rows -> plane records
  const int nb1 = (int)std::max<int64_t>(1, (nrows + K1_ROWS - 1) / K1_ROWS);
  const int64_t n1 = (int64_t)(P + 1) * nb1;
  if ((rc = grow(&p->d_h1, &p->cap_h1, n1))) return rc;
  if ((rc = grow(&p->d_s1, &p->cap_s1, n1 + 1))) return rc;
  if ((rc = ensure_scan_tmp(p, n1))) return rc;
  ARLMI_LAUNCH("k_p1_count", k_p1_count, dim3(nb1), dim3(256), sizeof(double) * (2 * P + 1) + sizeof(uint32_t) * (P + 3),
               s, g, uvw, freq, nrows, p->d_boxes, p->d_plane_w, p->flags, p->d_h1, nb1, p->d_misc);
  ARLMI_LAUNCH_CHECK();
  if ((rc = exclusive_scan_u32(p->d_h1, p->d_s1, n1, p->d_scan_tmp, s))) return rc;
  ARLMI_LAUNCH("k_p1_bounds", k_p1_bounds, dim3((P + 10 + 255) / 256), dim3(256), 0, s, p->d_s1, nb1, P, p->d_misc,
               p->d_bounds);
  ARLMI_LAUNCH_CHECK();
  // The segment bounds come back through pinned memory while the scatter pass runs: it is
  // launched before they arrive, into a buffer sized for one record a row and 1/64 more (rows in
  // two slabs are rare), and launched again into a larger one when the records are more.
  ARLMI_HIP(hipMemcpyAsync(p->h_bounds, p->d_bounds, sizeof(uint32_t) * (P + 10), hipMemcpyDeviceToHost, s));
  ARLMI_HIP(hipEventRecord(p->ev_bounds, s));
  const bool compact = compact_recs(g);
  p->rec_words = compact ? 2 : 4;
  const size_t rec_bytes = sizeof(uint32_t) * (size_t)p->rec_words;
  const size_t sh1 = sizeof(double) * (2 * P + 1) + sizeof(uint32_t) * (P + 1);
  auto scatter1 = [&]() -> int {
    const int64_t cap = std::min<int64_t>(p->cap_recs / (int64_t)rec_bytes, 0xffffffffll);
    if (compact)
      ARLMI_LAUNCH("k_p1_scatter", k_p1_scatter<Rec8>, dim3(nb1), dim3(256), sh1, s, g, uvw, freq, nrows, p->d_boxes,
                   p->d_plane_w, p->flags, p->d_s1, nb1, (Rec8*)p->d_recs, (uint32_t)cap);
    else
      ARLMI_LAUNCH("k_p1_scatter", k_p1_scatter<Rec>, dim3(nb1), dim3(256), sh1, s, g, uvw, freq, nrows, p->d_boxes,
                   p->d_plane_w, p->flags, p->d_s1, nb1, (Rec*)p->d_recs, (uint32_t)cap);
    ARLMI_LAUNCH_CHECK();
    return ARLMI_OK;
  };
  if ((rc = grow_bytes(&p->d_recs, &p->cap_recs, (int64_t)rec_bytes * (nrows + nrows / 64 + 1024)))) return rc;
  if ((rc = scatter1())) return rc;
  ARLMI_HIP(hipEventSynchronize(p->ev_bounds));
  const std::vector<uint32_t> bounds(p->h_bounds, p->h_bounds + P + 10);
  const uint32_t* misc = bounds.data() + P + 2;
  if (misc[0]) {
    ARLMI_HIP(hipStreamSynchronize(s));  // the scatter pass has ended before the plan can be freed
    rc = first_error_code(misc[0]);
    const char* what = rc == ARLMI_E_UV_OFF_GRID ? "uv (or kernel footprint) off the grid"
                       : rc == ARLMI_E_SUBSAMPLE ? "kernel sub-sample index equals oversampling"
                       : rc == ARLMI_E_W_RANGE   ? "w plane out of range"
                                                 : "frequency maps outside the image channels";
    set_error("convolution mapping failed: %s", what);
    return rc;
  }
  const int64_t nrec = bounds[P + 1];
  p->ndropped = misc[2];
  p->nmember = misc[3];
  p->orphan_start = bounds[P];
  p->norphans = bounds[P + 1] - bounds[P];
  if (nrec > p->cap_recs / (int64_t)rec_bytes) {
    ARLMI_HIP(hipStreamSynchronize(s));  // the first scatter pass still writes the old buffer
    if ((rc = grow_bytes(&p->d_recs, &p->cap_recs, (int64_t)rec_bytes * nrec))) return rc;
    if ((rc = scatter1())) return rc;
  }

  // pass 2: plane records -> tile entries.  The block layout (k_p2_layout) is made on the device
  // from the bounds it already holds; the host repeats the count for the launch sizes.
  const int64_t rb = std::max<int64_t>(K2_RECS, (int64_t)K2_HIST * 2 * tpp);  // records per block
  std::vector<uint32_t> nbp(P);
  int64_t nblk = 0;
  for (int pl = 0; pl < P; ++pl) {
    nbp[pl] = (uint32_t)std::max<int64_t>(1, ((int64_t)(bounds[pl + 1] - bounds[pl]) + rb - 1) / rb);
    nblk += nbp[pl];
  }
  const int64_t n2 = (int64_t)2 * tpp * nblk;
  if (n2 >= (1ll << 32) || rb >= (1ll << 32)) {
    set_error("tile histogram too large");
    return ARLMI_E_UNSUPPORTED;
  }
  if ((rc = grow(&p->d_h2, &p->cap_h2, n2 + 1))) return rc;
  if ((rc = grow(&p->d_s2, &p->cap_s2, n2 + 2))) return rc;
  ARLMI_LAUNCH("k_p2_layout", k_p2_layout, dim3(1), dim3(1024), 0, s, p->d_bounds, P, (uint32_t)rb, 2 * tpp,
               p->d_hbase);
  ARLMI_LAUNCH_CHECK();
  const size_t sh2 = sizeof(uint32_t) * 2 * (size_t)tpp;
  if (compact)
    ARLMI_LAUNCH("k_p2_count", k_p2_count<Rec8>, dim3((unsigned)nblk), dim3(512), sh2, s, g,
                 (const Rec8*)p->d_recs, p->d_bounds, p->d_hbase, (uint32_t)rb, p->d_h2);
  else
    ARLMI_LAUNCH("k_p2_count", k_p2_count<Rec>, dim3((unsigned)nblk), dim3(512), sh2, s, g,
                 (const Rec*)p->d_recs, p->d_bounds, p->d_hbase, (uint32_t)rb, p->d_h2);
  ARLMI_LAUNCH_CHECK();
  const int nbin = 2 * tpp;
  const int64_t nk = (int64_t)P * nbin;  // == 2 nt
  uint32_t* c2 = p->d_c2;
  uint32_t* e2 = p->d_c2 + (nk + 1);
  if ((rc = ensure_scan_tmp(p, nk + 1))) return rc;
  // slices of a plane's blocks: at least 16 blocks a slice for the largest plane, at most 16 slices
  int ns = 1;
  {
    uint32_t nbmax = 1;
    for (int pl = 0; pl < P; ++pl) nbmax = std::max(nbmax, nbp[pl]);
    while (ns < 16 && (uint32_t)ns * 16 < nbmax) ns *= 2;
  }
  // 4 bins a thread where the bins alone would give more than 2^20 threads
  const bool wide = nk > (1 << 20);
  const int64_t bins_per_group = (int64_t)(H2_THREADS / ns) * (wide ? 4 : 1);
  const dim3 hgrid((unsigned)((nbin + bins_per_group - 1) / bins_per_group), (unsigned)P);
  if (wide)
    ARLMI_LAUNCH("k_h2_colsum", k_h2_colsum<4>, hgrid, dim3(H2_THREADS), 0, s, p->d_h2, p->d_hbase,
                 p->d_hbase + P + 1, P, nbin, ns, c2);
  else
    ARLMI_LAUNCH("k_h2_colsum", k_h2_colsum<1>, hgrid, dim3(H2_THREADS), 0, s, p->d_h2, p->d_hbase,
                 p->d_hbase + P + 1, P, nbin, ns, c2);
  ARLMI_LAUNCH_CHECK();
  if ((rc = exclusive_scan_u32(c2, e2, nk + 1, p->d_scan_tmp, s))) return rc;
  if (wide)
    ARLMI_LAUNCH("k_h2_prefix", k_h2_prefix<4>, hgrid, dim3(H2_THREADS), 0, s, p->d_h2, p->d_hbase,
                 p->d_hbase + P + 1, P, nbin, ns, e2, p->d_s2);
  else
    ARLMI_LAUNCH("k_h2_prefix", k_h2_prefix<1>, hgrid, dim3(H2_THREADS), 0, s, p->d_h2, p->d_hbase,
                 p->d_hbase + P + 1, P, nbin, ns, e2, p->d_s2);
  ARLMI_LAUNCH_CHECK();
  ARLMI_LAUNCH("k_p2_keys", k_p2_keys, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, e2, nt, p->d_cnt,
               p->d_cntd, p->d_off);
  ARLMI_LAUNCH_CHECK();
  // entry buffers sized by an upper bound (each membership's footprint touches at most
  // ((S-1)/T + 2)^2 tiles), so the host need not wait for the exact count here; it is read with
  // the work-item totals below
  const int64_t per_axis = (g.support - 1) / g.tile + 2;
  const int64_t ebound = std::max<int64_t>(1, p->nmember * per_axis * per_axis);
  if ((rc = grow(&p->d_entries, &p->cap_entries, ebound))) return rc;
  if (g.nz > 1 && (rc = grow(&p->d_entry_z, &p->cap_z, ebound))) return rc;
  if (compact)
    ARLMI_LAUNCH("k_p2_scatter", k_p2_scatter<Rec8>, dim3((unsigned)nblk), dim3(512), sh2, s, g,
                 (const Rec8*)p->d_recs, p->d_bounds, p->d_hbase, (uint32_t)rb, p->d_s2, p->d_entries,
                 g.nz > 1 ? p->d_entry_z : nullptr);
  else
    ARLMI_LAUNCH("k_p2_scatter", k_p2_scatter<Rec>, dim3((unsigned)nblk), dim3(512), sh2, s, g,
                 (const Rec*)p->d_recs, p->d_bounds, p->d_hbase, (uint32_t)rb, p->d_s2, p->d_entries,
                 g.nz > 1 ? p->d_entry_z : nullptr);
  ARLMI_LAUNCH_CHECK();
  // entries grouped by footprint origin within their tile segments (one kernel w plane: the
  // kernel w planes of the entries would have to move with them; not for deterministic plans,
  // whose gridder adds every entry and runs slower on grouped entries, consecutive adds to the
  // same addresses): segments classified here, the sorts launched once the host knows the list
  // sizes (with the work-item totals)
  const int sortw = g.tile + g.support;
  p->origin_sorted = g.nz == 1 && !(p->flags & ARLMI_FLAG_DETERMINISTIC) && sortw * sortw <= SORT_MAXKEYS &&
                     sizeof(uint32_t) * 2 * sortw * sortw + sizeof(uint2) * SORT_BB <= 160 * 1024 - 1024;
  if (p->origin_sorted) {
    if ((rc = grow(&p->d_entries_alt, &p->cap_entries_alt, ebound))) return rc;
    if ((rc = grow(&p->d_sort_mid, &p->cap_sort_mid, 2 * nt))) return rc;
    const int64_t nbig = ebound / SORT_MID + 1;
    if ((rc = grow_bytes(&p->d_sort_big, &p->cap_sort_big, (int64_t)sizeof(SortSeg) * nbig))) return rc;
    if ((rc = grow(&p->d_sort_bigblk, &p->cap_sort_bigblk, ebound / SORT_BB + nbig))) return rc;
    const unsigned lg = (unsigned)((2 * nt + SORT_LIST_THREADS - 1) / SORT_LIST_THREADS);
    if ((rc = grow(&p->d_sort_wg, &p->cap_sort_wg, 4 * (int64_t)lg))) return rc;
    if ((rc = grow(&p->d_sort_segidx, &p->cap_sort_segidx, nbig))) return rc;
    ARLMI_LAUNCH("k_sort_lists_a", k_sort_lists_a, dim3(lg), dim3(SORT_LIST_THREADS), 0, s, p->d_off, p->d_cnt,
                 p->d_cntd, nt, p->d_sort_wg);
    ARLMI_LAUNCH_CHECK();
    ARLMI_LAUNCH("k_sort_lists_scan", k_sort_lists_scan, dim3(1), dim3(SORT_LIST_THREADS), 0, s, p->d_sort_wg, (int)lg,
                 p->d_sortcnt);
    ARLMI_LAUNCH_CHECK();
    ARLMI_LAUNCH("k_sort_lists_b", k_sort_lists_b, dim3(lg), dim3(SORT_LIST_THREADS), 0, s, p->d_off, p->d_cnt,
                 p->d_cntd, nt, p->d_sort_wg, p->d_sort_mid, (SortSeg*)p->d_sort_big, p->d_sort_bigblk,
                 p->d_sort_segidx, p->d_sortcnt);
    ARLMI_LAUNCH_CHECK();
  }

  // work items
  ARLMI_LAUNCH("k_chunks", k_chunks, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, p->d_cnt, p->d_cntd,
               (int)nt, p->chunk, (p->flags & (ARLMI_FLAG_SPARSE_GRIDS | ARLMI_FLAG_OCCUPIED_TILES)) ? 1 : 0,
               p->d_nch, p->d_nchd, p->d_nslot);
  ARLMI_LAUNCH_CHECK();
  if (p->flags & ARLMI_FLAG_SPARSE_GRIDS) {
    const int npc = (int)(nt / ((int64_t)g.ntx * g.ntx));
    const int64_t nm = (int64_t)npc * g.ntx;
    ARLMI_LAUNCH("k_tilemask", k_tilemask, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, s, p->d_cnt, npc, g.ntx,
                 p->d_tilemask);
    ARLMI_LAUNCH_CHECK();
    ARLMI_LAUNCH("k_colmask", k_colmask, dim3((unsigned)((npc + 255) / 256)), dim3(256), 0, s, p->d_tilemask, npc,
                 g.ntx, p->d_colmask);
    ARLMI_LAUNCH_CHECK();
  }
  if ((rc = ensure_scan_tmp(p, nt))) return rc;
  if ((rc = exclusive_scan_u32(p->d_nch, p->d_choff, nt, p->d_scan_tmp, s))) return rc;
  if ((rc = exclusive_scan_u32(p->d_nchd, p->d_choffd, nt, p->d_scan_tmp, s))) return rc;
  if ((rc = exclusive_scan_u32(p->d_nslot, p->d_slotoff, nt, p->d_scan_tmp, s))) return rc;
  uint32_t tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  ARLMI_LAUNCH("k_totals", k_totals, dim3(1), dim3(1), 0, s, e2, nk, p->d_choff, p->d_choffd, p->d_slotoff, nt,
               p->origin_sorted ? p->d_sortcnt : nullptr, p->d_bounds + P + 10);
  ARLMI_LAUNCH_CHECK();
  ARLMI_HIP(hipMemcpyAsync(tot, p->d_bounds + P + 10, sizeof(tot), hipMemcpyDeviceToHost, s));
  ARLMI_HIP(hipStreamSynchronize(s));
  const uint32_t total = tot[0];
  const uint32_t* tot3 = tot + 1;
  p->nentries = total;
  if (total > ebound) {
    set_error("internal: %u entries exceed the bound %lld", total, (long long)ebound);
    return ARLMI_E_UNSUPPORTED;
  }
  if (p->origin_sorted) {
    const uint32_t nmid = tot[4], nbig = tot[5], nbigblk = tot[6];
    // the segments under SORT_MIN entries: copied one by one where segments are few against the
    // entries, else with the whole array (the sorts then overwrite their segments): at 8192^2,
    // 4 M segments of 4 entries on average, the per-segment copy took 3 ms a plan
    if (2 * nt <= (int64_t)total / 8) {
      ARLMI_LAUNCH("k_sort_copytiny", k_sort_copytiny, dim3((unsigned)((2 * nt * 16 + 255) / 256)), dim3(256), 0, s,
                   p->d_entries, p->d_entries_alt, p->d_off, p->d_cnt, p->d_cntd, nt);
      ARLMI_LAUNCH_CHECK();
    } else if (total > 0) {
      ARLMI_HIP(hipMemcpyAsync(p->d_entries_alt, p->d_entries, sizeof(uint2) * (size_t)total, hipMemcpyDeviceToDevice,
                               s));
    }
    const int nkeys = sortw * sortw;
    const size_t shk = sizeof(uint32_t) * (size_t)nkeys;
    if (nmid > 0) {
      ARLMI_LAUNCH("k_sort_mid", k_sort_mid, dim3(nmid), dim3(SORT_MID_THREADS), shk, s, p->d_entries,
                   p->d_entries_alt, p->d_sort_mid, sortw);
      ARLMI_LAUNCH_CHECK();
    }
    if (nbig > 0) {
      if ((rc = grow(&p->d_sort_h3, &p->cap_sort_h3, (int64_t)nbigblk * nkeys))) return rc;
      ARLMI_LAUNCH("k_sort_bcount", k_sort_bcount, dim3(nbigblk), dim3(SORT_BIG_THREADS), shk, s, p->d_entries,
                   (const SortSeg*)p->d_sort_big, p->d_sort_bigblk, sortw, p->d_sort_h3);
      ARLMI_LAUNCH_CHECK();
      if ((rc = grow(&p->d_sort_ksum, &p->cap_sort_ksum, (int64_t)nbig * nkeys))) return rc;
      const uint32_t nhuge = tot[7];
      if (nhuge > 0) {  // segidx[0, nhuge): 16 slices
        const dim3 kg((unsigned)((nkeys + 63) / 64), nhuge);
        ARLMI_LAUNCH("k_sort_bksum", k_sort_bksum<16>, kg, dim3(1024), 0, s, (const SortSeg*)p->d_sort_big,
                     p->d_sort_segidx, sortw, p->d_sort_h3, p->d_sort_ksum);
        ARLMI_LAUNCH_CHECK();
        ARLMI_LAUNCH("k_sort_bcursor", k_sort_bcursor<16>, kg, dim3(1024), 0, s, (const SortSeg*)p->d_sort_big,
                     p->d_sort_segidx, sortw, p->d_sort_ksum, p->d_sort_h3);
        ARLMI_LAUNCH_CHECK();
      }
      if (nbig > nhuge) {  // segidx[nhuge, nbig): one slice
        const dim3 kg((unsigned)((nkeys + 1023) / 1024), nbig - nhuge);
        ARLMI_LAUNCH("k_sort_bksum", k_sort_bksum<1>, kg, dim3(1024), 0, s, (const SortSeg*)p->d_sort_big,
                     p->d_sort_segidx + nhuge, sortw, p->d_sort_h3, p->d_sort_ksum);
        ARLMI_LAUNCH_CHECK();
        ARLMI_LAUNCH("k_sort_bcursor", k_sort_bcursor<1>, kg, dim3(1024), 0, s, (const SortSeg*)p->d_sort_big,
                     p->d_sort_segidx + nhuge, sortw, p->d_sort_ksum, p->d_sort_h3);
        ARLMI_LAUNCH_CHECK();
      }
      ARLMI_LAUNCH("k_sort_bscatter", k_sort_bscatter, dim3(nbigblk), dim3(SORT_SCAT_THREADS),
                   2 * shk + sizeof(uint2) * SORT_BB, s, p->d_entries,
                   p->d_entries_alt, (const SortSeg*)p->d_sort_big, p->d_sort_bigblk, sortw, p->d_sort_h3);
      ARLMI_LAUNCH_CHECK();
    }
    std::swap(p->d_entries, p->d_entries_alt);
    std::swap(p->cap_entries, p->cap_entries_alt);
  }
  p->ngitems = tot3[0];
  p->nditems = tot3[1];
  p->nslots = tot3[2];
  if ((rc = grow(&p->d_gitems, &p->cap_gitems, std::max<int64_t>(1, p->ngitems)))) return rc;
  if ((rc = grow(&p->d_ditems, &p->cap_ditems, std::max<int64_t>(1, p->nditems)))) return rc;
  const int64_t nitems = std::max<int64_t>(1, std::max(p->ngitems, p->nditems));
  ARLMI_LAUNCH("k_items", k_items, dim3((unsigned)((nitems + 255) / 256)), dim3(256), 0, s, p->d_cnt, p->d_cntd,
               p->d_off, p->d_choff, p->d_choffd, p->d_slotoff, (int)nt, p->chunk, (uint32_t)p->ngitems,
               (uint32_t)p->nditems, p->d_gitems, p->d_ditems, p->d_multi, p->d_misc);
  ARLMI_LAUNCH_CHECK();
  p->nmulti = -1;  // left on the device (d_misc[4]); the reduction kernel reads it
  return ARLMI_OK;
}

// A build that fails part way may leave kernels and copies of its own in flight (the scatter
// pass, the staged block descriptors): they end before the caller frees or reuses the buffers.
static int plan_build(arlmi_plan* p, const double* uvw, const double* freq, int64_t nrows, const double* h_boxes,
                      double wslice, hipStream_t s) {
  const int rc = plan_build_steps(p, uvw, freq, nrows, h_boxes, wslice, s);
  if (rc) (void)hipStreamSynchronize(s);
  return rc;
}

extern "C" {

int arlmi_version(void) { return 100; }

const char* arlmi_last_error(void) { return arlmi::last_error(); }

int arlmi_wmax(const double* d_uvw, int64_t nrows, void* stream, double* wmax_out) {
  if (!wmax_out || (nrows > 0 && !d_uvw)) {
    set_error("arlmi_wmax: NULL argument");
    return ARLMI_E_ARG;
  }
  hipStream_t s = as_stream(stream);
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((nrows + 1023) / 1024, 1024));
  double* d = nullptr;
  ARLMI_HIP(hipMallocAsync(reinterpret_cast<void**>(&d), sizeof(double) * blocks, s));
  std::vector<double> part(blocks, 0.0);
  if (nrows > 0) {
    ARLMI_LAUNCH("k_wmax", k_wmax, dim3((unsigned)blocks), dim3(256), 0, s, d_uvw, nrows, d);
    ARLMI_LAUNCH_CHECK();
    ARLMI_HIP(hipMemcpyAsync(part.data(), d, sizeof(double) * blocks, hipMemcpyDeviceToHost, s));
  }
  ARLMI_HIP(hipFreeAsync(d, s));
  ARLMI_HIP(hipStreamSynchronize(s));
  double w = 0.0;
  for (double v : part) w = std::max(w, v);
  *wmax_out = w;
  return ARLMI_OK;
}

int arlmi_plan_create(const arlmi_geom* gin, const double* d_uvw, const double* d_freq, int64_t nrows,
                      const double* h_boxes, double wslice, uint32_t flags, void* stream, arlmi_plan** out) {
  if (!out) {
    set_error("plan_out is NULL");
    return ARLMI_E_ARG;
  }
  *out = nullptr;
  arlmi_plan* p = new arlmi_plan();
  int rc = checked_geom(gin, &p->g);
  if (rc) {
    delete p;
    return rc;
  }
  p->flags = flags;
  // the occupancy of a tile column fits one 32-bit word: beyond 32 tiles per side the grids
  // stay dense
  if (p->g.ntx > 32) p->flags &= ~ARLMI_FLAG_SPARSE_GRIDS;
  p->tpp = p->g.nchan * p->g.ntx * p->g.ntx;
  p->ntiles = p->g.nplanes * p->tpp;
  rc = plan_alloc_fixed(p);
  if (!rc) rc = plan_build(p, d_uvw, d_freq, nrows, h_boxes, wslice, as_stream(stream));
  if (rc) {
    free_plan(p);
    delete p;
    return rc;
  }
  *out = p;
  return ARLMI_OK;
}

int arlmi_plan_fix_plane_w(arlmi_plan* p, const double* h_plane_w, int32_t n) {
  if (!p) {
    set_error("plan is NULL");
    return ARLMI_E_ARG;
  }
  if (n == 0 || !h_plane_w) {
    p->fixed_w.clear();
    return ARLMI_OK;
  }
  if (p->g.mode != ARLMI_MODE_WSTACK || n != p->g.nplanes) {
    set_error("arlmi_plan_fix_plane_w: a w-stack plan of %d slabs needs %d means (%d given)", p->g.nplanes,
              p->g.nplanes, n);
    return ARLMI_E_ARG;
  }
  p->fixed_w.assign(h_plane_w, h_plane_w + n);
  return ARLMI_OK;
}

int arlmi_wslab_sums(const double* d_uvw, int64_t nrows, const double* h_boxes, double wslice, int32_t nslabs,
                     double* h_sum, int64_t* h_count, void* stream) {
  if (!h_boxes || !h_sum || !h_count || nslabs < 1 || (nrows > 0 && !d_uvw) || nrows < 0) {
    set_error("arlmi_wslab_sums: bad argument");
    return ARLMI_E_ARG;
  }
  if (nrows == 0) return ARLMI_OK;
  hipStream_t s = as_stream(stream);
  const int P = nslabs;
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((nrows + 1023) / 1024, SLAB_BLOCKS));
  const size_t bytes = sizeof(double) * (P + 1) + (sizeof(double) + sizeof(uint32_t)) * (size_t)blocks * P +
                       (sizeof(double) + sizeof(unsigned long long)) * P + 64;
  unsigned char* mem = nullptr;
  ARLMI_HIP(hipMallocAsync(reinterpret_cast<void**>(&mem), bytes, s));
  double* d_boxes = reinterpret_cast<double*>(mem);
  double* d_part = d_boxes + (P + 1);
  double* d_tsum = d_part + (size_t)blocks * P;
  unsigned long long* d_tcnt = reinterpret_cast<unsigned long long*>(d_tsum + P);
  uint32_t* d_pcnt = reinterpret_cast<uint32_t*>(d_tcnt + P);
  std::vector<double> hb(h_boxes, h_boxes + P);
  hb.push_back(wslice);
  std::vector<double> tsum(P);
  std::vector<unsigned long long> tcnt(P);
  int rc = ARLMI_OK;
  do {
    if (hipMemcpyAsync(d_boxes, hb.data(), sizeof(double) * (P + 1), hipMemcpyHostToDevice, s) != hipSuccess) {
      rc = ARLMI_E_HIP;
      break;
    }
    const size_t sh = (2 * sizeof(double) + sizeof(unsigned int)) * P;
    ARLMI_LAUNCH("k_slab_sums", k_slab_sums, dim3((unsigned)blocks), dim3(256), sh, s, d_uvw, nrows, d_boxes, P, d_part,
                 d_pcnt);
    ARLMI_LAUNCH("k_slab_means", k_slab_means, dim3((unsigned)P), dim3(256), 0, s, d_part, d_pcnt, blocks,
                 (const double*)d_boxes, nullptr, P, d_tsum, d_tcnt);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(tsum.data(), d_tsum, sizeof(double) * P, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(tcnt.data(), d_tcnt, sizeof(unsigned long long) * P, hipMemcpyDeviceToHost, s) != hipSuccess)
      rc = ARLMI_E_HIP;
  } while (0);
  (void)hipFreeAsync(mem, s);
  if (hipStreamSynchronize(s) != hipSuccess) rc = ARLMI_E_HIP;
  if (rc) {
    set_error("arlmi_wslab_sums: %s", hipGetErrorString(hipGetLastError()));
    return rc;
  }
  for (int k = 0; k < P; ++k) {
    h_sum[k] += tsum[k];
    h_count[k] += (int64_t)tcnt[k];
  }
  return ARLMI_OK;
}

int arlmi_plan_update(arlmi_plan* p, const double* d_uvw, const double* d_freq, int64_t nrows,
                      const double* h_boxes, double wslice, void* stream) {
  if (!p) {
    set_error("plan is NULL");
    return ARLMI_E_ARG;
  }
  return plan_build(p, d_uvw, d_freq, nrows, h_boxes, wslice, as_stream(stream));
}

int arlmi_plan_or_colmask(const arlmi_plan* p, uint32_t* d_colmask, int32_t* band_shift_out, void* stream) {
  if (!p || !d_colmask) {
    set_error("arlmi_plan_or_colmask: NULL argument");
    return ARLMI_E_ARG;
  }
  const DevGeom& g = p->g;
  const int bs = band_shift(g);
  if (band_shift_out) *band_shift_out = bs;
  const int npc = g.nplanes * g.nchan;
  const int64_t n = (int64_t)npc * g.ntx;
  if (n == 0) return ARLMI_OK;
  ARLMI_LAUNCH("k_or_colmask", k_or_colmask, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream),
               p->d_cnt, npc, g.ntx, bs - g.tshift, d_colmask);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

int arlmi_plan_use_colmask(arlmi_plan* p, const uint32_t* d_colmask) {
  if (!p) {
    set_error("plan is NULL");
    return ARLMI_E_ARG;
  }
  if (d_colmask && (p->flags & ARLMI_FLAG_SPARSE_GRIDS)) {
    set_error("arlmi_plan_use_colmask: a plan with sparse grids uses its own occupancy");
    return ARLMI_E_ARG;
  }
  p->ext_colmask = d_colmask;
  return ARLMI_OK;
}

int arlmi_plan_set_chunk(arlmi_plan* p, int32_t chunk) {
  if (!p || chunk < 64 || chunk > (1 << 20)) {
    set_error("chunk %d outside [64, 1048576]", chunk);
    return ARLMI_E_ARG;
  }
  p->chunk = chunk;
  return ARLMI_OK;
}

int arlmi_plan_stats_get(const arlmi_plan* p, arlmi_plan_stats* out) {
  if (!p || !out) {
    set_error("NULL argument");
    return ARLMI_E_ARG;
  }
  out->nrows = p->nrows;
  out->nentries = p->nentries;
  out->ngrid_items = p->ngitems;
  out->ndegrid_items = p->nditems;
  out->nmember = p->nmember;
  out->ndropped = p->ndropped;
  out->norphans = p->norphans;
  out->ntiles = p->ntiles;
  out->tile = p->g.tile;
  return ARLMI_OK;
}

int arlmi_plan_plane_w(const arlmi_plan* p, double* h_out, int32_t n) {
  if (!p || !h_out || n < p->g.nplanes) {
    set_error("arlmi_plan_plane_w: bad argument");
    return ARLMI_E_ARG;
  }
  ARLMI_HIP(hipMemcpy(h_out, p->d_plane_w, sizeof(double) * p->g.nplanes, hipMemcpyDeviceToHost));
  return ARLMI_OK;
}

int arlmi_plan_slab_boxes(const arlmi_plan* p, double* h_out, int32_t n) {
  if (!p || !h_out || p->g.mode != ARLMI_MODE_WSTACK || n < p->g.nplanes + 1) {
    set_error("arlmi_plan_slab_boxes: bad argument (w-stack plans only, n >= nplanes + 1)");
    return ARLMI_E_ARG;
  }
  ARLMI_HIP(hipMemcpy(h_out, p->d_boxes, sizeof(double) * (p->g.nplanes + 1), hipMemcpyDeviceToHost));
  return ARLMI_OK;
}

int arlmi_convolution_mapping(const arlmi_geom* gin, const double* d_uvw, const double* d_freq, int64_t nrows,
                              int32_t* d_out, void* stream) {
  DevGeom g;
  int rc = checked_geom(gin, &g);
  if (rc) return rc;
  if (nrows < 0 || (nrows > 0 && (!d_uvw || !d_out))) {
    set_error("arlmi_convolution_mapping: bad argument");
    return ARLMI_E_ARG;
  }
  if (nrows == 0) return ARLMI_OK;
  ARLMI_LAUNCH("k_map", k_map, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, as_stream(stream), g, d_uvw,
                     d_freq, nrows, d_out);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

void arlmi_plan_destroy(arlmi_plan* p) {
  if (!p) return;
  free_plan(p);
  delete p;
}

}  // extern "C"
