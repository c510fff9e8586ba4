// Internal declarations shared by the libarlmi translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <vector>

#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <cmath>

#include "arlmi.h"

namespace arlmi {

void set_error(const char* fmt, ...);
inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

#define ARLMI_HIP(call)                                                                     \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      ::arlmi::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,  \
                         __LINE__);                                                         \
      return ARLMI_E_HIP;                                                                   \
    }                                                                                       \
  } while (0)

#define ARLMI_LAUNCH_CHECK()                                                                \
  do {                                                                                      \
    hipError_t e_ = hipGetLastError();                                                      \
    if (e_ != hipSuccess) {                                                                 \
      ::arlmi::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_),        \
                         __FILE__, __LINE__);                                               \
      return ARLMI_E_HIP;                                                                   \
    }                                                                                       \
  } while (0)

template <typename R> struct cplx_t;
template <> struct cplx_t<float> { using type = float2; };
template <> struct cplx_t<double> { using type = double2; };
template <typename R> using cplx = typename cplx_t<R>::type;

// Error bits raised by the mapping kernels (reduced to the first ARLMI_E_* code on the host).
enum : uint32_t {
  ERRB_UV_OFF_GRID = 1u,
  ERRB_W_RANGE = 2u,
  ERRB_FREQ_MAP = 4u,
  ERRB_SUBSAMPLE = 8u,
};

// Packed entry word: lx | ly << 8 | xoff << 16 | yoff << 23 | START << 30 | DEGRID << 31.
// lx, ly: kernel footprint start relative to the tile origin, plus S (so >= 1).
constexpr uint32_t ENTRY_START = 1u << 30;   // the tile holding the footprint start (counts sumwt)
constexpr uint32_t ENTRY_DEGRID = 1u << 31;  // the (row, plane) degridding writes back

__host__ __device__ inline uint32_t pack_entry(int lx, int ly, int xoff, int yoff, bool start, bool degrid) {
  return (uint32_t)lx | ((uint32_t)ly << 8) | ((uint32_t)xoff << 16) | ((uint32_t)yoff << 23) |
         (start ? ENTRY_START : 0u) | (degrid ? ENTRY_DEGRID : 0u);
}
__host__ __device__ inline int entry_lx(uint32_t p) { return (int)(p & 0xffu); }
__host__ __device__ inline int entry_ly(uint32_t p) { return (int)((p >> 8) & 0xffu); }
__host__ __device__ inline int entry_xoff(uint32_t p) { return (int)((p >> 16) & 0x7fu); }
__host__ __device__ inline int entry_yoff(uint32_t p) { return (int)((p >> 23) & 0x7fu); }

// Result of ARL's convolution_mapping for one row (gridding.py:24-96).
// Work groups of one launch are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  xcd_block maps the hardware block index to a logical one under which
// every run of G consecutive logical blocks (neighbours that read the same cache lines, such
// as adjacent column groups) is dispatched to one XCD in consecutive slots, while successive
// runs go round-robin over the XCDs (so uneven work along the logical order stays balanced).
__device__ inline int xcd_block(int bid, int n, int G) {
  const int span = 8 * G;
  const int full = (n / span) * span;
  if (bid >= full) return bid;
  const int xcd = bid & 7, slot = bid >> 3;
  return ((slot / G) * 8 + xcd) * G + slot % G;
}

struct RowMap {
  int pu, pv;      // nearest grid pixel (numpy.round: half to even)
  int xoff, yoff;  // kernel sub-sample (numpy.floor; a -1 wraps to os-1 like a numpy index)
  int chan;        // image channel
  uint32_t err;    // ERRB_* bits
};

// Geometry in the form the device code reads it.
struct DevGeom {
  int n, nchan, support, os, nz, nplanes, mode, tile, ntx;
  int tshift;  // log2(tile)
  double du, dv, ddu, ddv, crpix_uv, crpix_d, zstep, crpix_z, wstep_grid, crpix_wg, freq0, dfreq;
  double cellsize;
};

// World -> pixel on a linear WCS axis with origin 0, in wcslib's operation order:
// ((world - crval) / cdelt + crpix) - 1 (crval is 0 on the UU/VV/DUU/DVV/WW axes).
__device__ inline double world2pix0(double world, double cdelt, double crpix) {
  return (world / cdelt + crpix) - 1.0;
}

// round((x / step + crpix) - 1) for an integral crpix.  When |x| <= |step| / 4 the quotient lies
// in [-1/4, 1/4] (division is monotone and 1/4 is exact), so the rounded pixel is crpix - 1
// without dividing; otherwise the division is done as wcslib does it.  Identical results.
__device__ inline double round_pix0(double x, double step, double crpix) {
  if (fabs(x) <= 0.25 * fabs(step) && step != 0.0 && crpix == rint(crpix)) return crpix - 1.0;
  return rint(world2pix0(x, step, crpix));
}

__device__ inline RowMap map_row(const DevGeom& g, double u, double v, double f, bool have_freq) {
  RowMap m;
  m.err = 0;
  const double x = world2pix0(u, g.du, g.crpix_uv);
  const double y = world2pix0(v, g.dv, g.crpix_uv);
  const double pud = rint(x);
  const double pvd = rint(y);
  const double half = (double)(g.support / 2);
  // footprint (pu - S/2 .. pu + S/2 - 1) must lie inside [0, N): also implies 0 <= pu < N
  if (!(pud - half >= 0.0 && pud + half <= (double)g.n && pvd - half >= 0.0 && pvd + half <= (double)g.n)) {
    m.err |= ERRB_UV_OFF_GRID;
    m.pu = m.pv = 0;
  } else {
    m.pu = (int)pud;
    m.pv = (int)pvd;
  }
  // pixel -> world (origin 0): cdelt * ((p + 1) - crpix) + crval, crval = 0
  const double wu_grid = g.du * ((pud + 1.0) - g.crpix_uv) + 0.0;
  const double wv_grid = g.dv * ((pvd + 1.0) - g.crpix_uv) + 0.0;
  const double wu_sub = u - wu_grid;
  const double wv_sub = v - wv_grid;
  int xo = (int)floor(world2pix0(wu_sub, g.ddu, g.crpix_d));
  int yo = (int)floor(world2pix0(wv_sub, g.ddv, g.crpix_d));
  if (xo < 0) xo += g.os;  // numpy negative index
  if (yo < 0) yo += g.os;
  if (xo < 0 || xo >= g.os || yo < 0 || yo >= g.os) m.err |= ERRB_SUBSAMPLE;
  m.xoff = xo;
  m.yoff = yo;
  int c = 0;
  if (have_freq) {
    const double pf = round_pix0(f - g.freq0, g.dfreq, 1.0);
    if (!(pf >= 0.0 && pf < (double)g.nchan)) {
      m.err |= ERRB_FREQ_MAP;
    } else {
      c = (int)pf;
    }
  }
  m.chan = c;
  return m;
}

// Kernel w plane of a (shifted) w value: round((w/zstep + crpix_z) - 1) (gridding.py:70-76).
__device__ inline int map_wplane(double w, double step, double crpix, int nmax, uint32_t* err) {
  const double p = round_pix0(w, step, crpix);
  if (!(p >= 0.0 && p < (double)nmax)) {
    *err |= ERRB_W_RANGE;
    return 0;
  }
  return (int)p;
}

int first_error_code(uint32_t bits);

// Validates a caller's geometry and converts it to the device form.
int checked_geom(const arlmi_geom* in, DevGeom* g);

// Exclusive scan of n uint32 values (out has n+1 entries; out[n] = total).  tmp must hold
// scan_tmp_words(n) words.
size_t scan_tmp_words(int64_t n);
int exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp, hipStream_t s);

// Device twiddle table exp(-2 pi i k / n), k < n, cached per (n, precision).
const void* twiddles(int n, int precision);

// Grid rows in the digit-transposed order of the two-level column transform: for a plan with
// ARLMI_FLAG_PRIVATE_GRIDS, arlmi_image_to_grids leaves row v of an n x n plane at row
// perm_row_of(v, n, n1) with n1 = grid_perm_n1(n, precision) (0: natural order).
int grid_perm_n1(int n, int precision);
__host__ __device__ inline int perm_row_of(int v, int n, int n1) {
  if (!n1) return v;
  const int n2 = n / n1;
  return n1 * (v % n2) + v / n2;
}

// Fixed-point accumulation of ARLMI_FLAG_DETERMINISTIC plans: d_det holds [0..7] the call's
// constants (2^qw and 2^-qw of the sums of weights, the single-tap bound X), then [nsw] 64-bit
// integer sums of weights, then DET_BLOCKS x 4 bound partials of the call.
constexpr int DET_BLOCKS = 512;
inline int64_t det_words(int64_t nsw) { return 8 + nsw + 4 * (int64_t)DET_BLOCKS; }

// round(x * scale) as a 64-bit integer, for a power-of-two scale and |x * scale| <= 2^51: the
// fused multiply-add with 1.5 * 2^52 lands in [2^52, 2^53), where the doubles are the integers, and
// its bit pattern minus that of 1.5 * 2^52 is the rounded integer (ties to even).  Integer sums of
// such values are exact in any order.
constexpr long long DET_MAGIC_BITS = 0x4338000000000000ll;  // the bits of 1.5 * 2^52
__device__ __forceinline__ long long det_round(double x, double scale) {
  return __double_as_longlong(__builtin_fma(x, scale, 0x1.8p52)) - DET_MAGIC_BITS;
}

// Exponent e with y < 2^e for y > 0 finite; -1000 for 0.
__device__ inline int det_exp(double y) {
  if (!(y > 0.0)) return -1000;
  int e;
  (void)frexp(y, &e);
  return e;
}

// The scale exponent q for integer sums of terms |x| <= bound_one whose sums stay within
// bound_sum: bound_sum 2^q < 2^60 (64-bit sums cannot overflow) and bound_one 2^q < 2^50 (inside
// det_round's range); 0 when a bound is not finite.
__device__ inline int det_q(double bound_sum, double bound_one) {
  if (!isfinite(bound_sum) || !isfinite(bound_one)) return 0;
  const int a = 60 - det_exp(bound_sum), b = 50 - det_exp(bound_one);
  const int q = a < b ? a : b;
  return q < -1000 ? -1000 : (q > 1000 ? 1000 : q);
}

// log2 of the width of the column bands of an occupancy mask: the tile side, or N / 32 when a plane
// has more than 32 tile columns (a band mask is one 32-bit word per plane-channel).
inline int band_shift(const DevGeom& g) {
  int t = 0, n = g.n / 32;
  while ((1 << t) < g.tile) ++t;
  int b = 0;
  while ((1 << b) < n) ++b;
  return t > b ? t : b;
}

// Opt-in kernel timing (profile.hip).
bool profile_on();
void* profile_begin(const char* name, hipStream_t s);
void profile_end(void* token, hipStream_t s);

// Launch a kernel, bracketed by timing events when profiling is on.
#define ARLMI_LAUNCH(name, kernel, grid, block, shmem, stream, ...)          \
  do {                                                                       \
    void* tok_ = ::arlmi::profile_begin(name, stream);                       \
    hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);     \
    ::arlmi::profile_end(tok_, stream);                                      \
  } while (0)

}  // namespace arlmi

// The plan object (opaque to C callers).
struct arlmi_plan {
  arlmi::DevGeom g;
  uint32_t flags;
  int64_t nrows;
  int ntiles;  // planes * tiles per plane
  int tpp;     // tiles per plane (nchan * ntx * ntx)
  // device buffers (capacities in elements)
  double* d_boxes = nullptr;             // [P]
  double* d_plane_w = nullptr;           // [P] slab mean w
  double* d_wsum = nullptr;              // [blocks][P] partial slab sums of w
  uint32_t* d_wcnt = nullptr;            // [blocks][P] partial slab counts
  uint32_t* d_cnt = nullptr;             // [ntiles] entries per tile
  // occupancy bit masks (sparse grids, at most 32 tiles per side), pc = plane * nchan + chan:
  uint32_t* d_tilemask = nullptr;        // [pc][tx]: bit ty set if tile (pc, ty, tx) has entries
  uint32_t* d_colmask = nullptr;         // [pc]: bit tx set if tile column (pc, tx) has entries
  void* d_zeros = nullptr;               // 1 KiB of zeros (masked-out loads of the FFT passes)
  const uint32_t* ext_colmask = nullptr;  // caller's column-band occupancy for the transforms (or null)
  uint32_t* d_cntd = nullptr;            // [ntiles] degrid entries per tile
  uint32_t* d_off = nullptr;             // [ntiles] first entry of a tile
  uint32_t* d_nch = nullptr;             // [ntiles] grid chunks per tile
  uint32_t* d_nchd = nullptr;            // [ntiles] degrid chunks per tile
  uint32_t* d_nslot = nullptr;           // [ntiles] partial-tile slots per tile
  uint32_t* d_choff = nullptr;           // [ntiles+1]
  uint32_t* d_choffd = nullptr;          // [ntiles+1]
  uint32_t* d_slotoff = nullptr;         // [ntiles+1]
  uint32_t* d_misc = nullptr;            // [8]: err bits, -, dropped, nmember, nmulti
  uint32_t* d_bounds = nullptr;          // [P+2] plane segment starts in the record array, [8] misc
                                         // copy, [8] totals (gathered for single copies)
  uint32_t* h_bounds = nullptr;          // pinned host copy of d_bounds
  hipEvent_t ev_bounds = nullptr;        // recorded after the copy of the segment bounds
  uint32_t* d_hbase = nullptr;           // [P+1] plane bases in h2, then [P] pass-2 blocks per plane
  uint32_t *d_h1 = nullptr, *d_s1 = nullptr;
  uint32_t *d_h2 = nullptr, *d_s2 = nullptr;
  uint32_t* d_c2 = nullptr;              // [2 ntiles + 1] pass-2 bin sums, then [2 ntiles + 2] their scan
  uint32_t* d_scan_tmp = nullptr;
  void* d_recs = nullptr;                // membership records (rec_words 32-bit words each, the row first)
  int rec_words = 4;
  uint2* d_entries = nullptr;            // (row, packed)
  uint2* d_entries_alt = nullptr;        // the other buffer of the origin sort
  bool origin_sorted = false;            // entries grouped by footprint origin in their tile segments
  uint2* d_sort_mid = nullptr;           // origin sort: middle segments (start, entries)
  void* d_sort_big = nullptr;            // large segments (start, entries, first block, blocks)
  uint2* d_sort_bigblk = nullptr;        // their blocks (segment, block)
  uint32_t* d_sort_h3 = nullptr;         // block histograms, then cursors
  uint32_t* d_sort_ksum = nullptr;       // per large segment and key: counts over its blocks
  uint32_t* d_sort_wg = nullptr;         // per work group of the list pass: its list sizes
  uint32_t* d_sort_segidx = nullptr;     // large segments, the huge ones first
  uint32_t* d_sortcnt = nullptr;         // [4] list sizes
  int64_t cap_sort_mid = 0, cap_sort_big = 0, cap_sort_bigblk = 0, cap_sort_h3 = 0, cap_sort_ksum = 0, cap_sort_wg = 0, cap_sort_segidx = 0;
  uint16_t* d_entry_z = nullptr;         // kernel w plane per entry (nz > 1)
  int4* d_gitems = nullptr;              // {tile, start, count, partial slot + 1 or 0}
  int4* d_ditems = nullptr;              // {tile, start, count, 0}
  int4* d_multi = nullptr;               // {tile, first slot, slots, 0} for split tiles
  void* d_partial = nullptr;             // partial tiles of split tiles (grid precision)
  void* d_img_part = nullptr;            // per-plane-group partial images of the invert row pass
  double* d_det = nullptr;               // ARLMI_FLAG_DETERMINISTIC: scales, integer sums of weights,
                                         // bound partials (det_words)
  int64_t cap_img_part_bytes = 0;
  int64_t cap_h1 = 0, cap_s1 = 0, cap_h2 = 0, cap_s2 = 0, cap_scan = 0, cap_recs = 0, cap_entries = 0, cap_entries_alt = 0, cap_z = 0,
          cap_gitems = 0, cap_ditems = 0, cap_partial_bytes = 0;
  // host copies
  int64_t nentries = 0, ngitems = 0, nditems = 0, nmulti = 0, nslots = 0, norphans = 0, ndropped = 0, nmember = 0;
  int64_t orphan_start = 0;  // orphan rows are records [orphan_start, orphan_start + norphans)
  int chunk = 4096;
  // slab means given by the caller (arlmi_plan_fix_plane_w; empty: the rows' own means)
  std::vector<double> fixed_w;
};
