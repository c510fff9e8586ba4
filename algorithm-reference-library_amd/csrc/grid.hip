// Convolutional gridding and degridding over the plan's tile lists.
//
// grid_visibility_to_griddata (processing_components/griddata/gridding.py:99-133):
//   G[chan, :, plane, pv-S/2+a, pu-S/2+b] += conj(CF[chan, :, z, yoff, xoff, a, b]) * vis * weight
// degrid_visibility_from_griddata (gridding.py:233-267):
//   V = sum_ab G[chan, :, plane, pv-S/2+a, pu-S/2+b] * CF[chan, :, z, yoff, xoff, a, b]
//
// One work group owns one tile chunk.  Gridding: every entry whose kernel footprint touches the
// tile is listed in it; one wavefront adds one entry at a time, one kernel tap per lane, into a
// T x T float64 accumulator tile in LDS (taps outside the tile are masked), and the tile is
// stored to HBM in one coalesced pass (split tiles store partial tiles, summed in slot order
// by k_reduce_tiles; no global atomics).  The accumulators are float64 in both precisions:
// on gfx950 an LDS ds_add_f64 costs ~9 cycles per wavefront instruction and ds_add_f32 ~194.
// Degridding: the tile plus an S-wide halo is staged in LDS once and every lane gathers the
// footprint of its own entry from it.
#include "arlmi_internal.h"

#include <algorithm>

namespace arlmi {

template <typename R>
__device__ inline cplx<R> mk(R x, R y) {
  cplx<R> c;
  c.x = x;
  c.y = y;
  return c;
}

struct TilePos {
  int plane, ty, tx;
};
__device__ inline TilePos tile_pos(int tile, int ntx) {
  TilePos t;
  const int per = ntx * ntx;
  t.plane = tile / per;
  const int rem = tile - t.plane * per;
  t.ty = rem / ntx;
  t.tx = rem - t.ty * ntx;
  return t;
}

// Work-group size of the gridders: with the ~75 KiB tile of T=64, two groups of 16 waves
// share a CU; the add loop is latency-bound, and 32 waves per CU measured 18% faster than 16
// and 37% faster than 8.
constexpr int GRID_THREADS = 1024;

__device__ inline void block_add_double(double v, double* dst, double* scratch) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int wid = threadIdx.x >> 6;
  if (__lane_id() == 0) scratch[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += scratch[i];
    if (t != 0.0) unsafeAtomicAdd(dst, t);
  }
}

// ---------------------------------------------------------------------------
// fixed-point accumulation (ARLMI_FLAG_DETERMINISTIC)
//
// A tap's product x becomes the integer round(x 2^q) (det_round); the LDS tiles, partial tiles
// and sums of weights then add 64-bit integers, exact in any order.  q is chosen per tile and
// call: q = det_q(X n_t, X), where X = Kmax * M bounds every single tap and X n_t every grid value
// of tile t and every partial sum of one (n_t entries of the tile), with M the largest
// m = |Re vis w| + |Im vis w| (w for a PSF) over the rows and Kmax = (max |k|)^2 over a separable
// table or max (|Re c| + |Im c|) over a full one.  M and Kmax are maxima and n_t an integer count,
// so q is the same on every build of the plan whatever the order of its entries.  The sums of
// weights take qw from sum |w| and max |w| over the rows.
// ---------------------------------------------------------------------------
// One accumulator add: a float64 LDS add, or (DET) the fixed-point integer add.
template <bool DET>
__device__ __forceinline__ void acc_add(double* p, double v, double scale) {
  if constexpr (DET) {
    __hip_atomic_fetch_add(reinterpret_cast<long long*>(p), det_round(v, scale), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
  } else {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

__device__ inline void block_add_i64(long long v, unsigned long long* dst, long long* scratch) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int wid = threadIdx.x >> 6;
  if (__lane_id() == 0) scratch[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += scratch[i];
    if (t != 0) atomicAdd(dst, (unsigned long long)t);
  }
}

// Sum (op 0) or maximum (op 1) over the block in a fixed order; every thread gets it.
template <int OP>
__device__ inline double block_fold(double v, double* scratch) {
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_xor(v, o);
    v = OP == 0 ? v + u : fmax(v, u);
  }
  __syncthreads();
  if (__lane_id() == 0) scratch[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = scratch[0];
  for (int i = 1; i < (int)(blockDim.x >> 6); ++i) t = OP == 0 ? t + scratch[i] : fmax(t, scratch[i]);
  return t;
}

// Per-block partials of one call (DET_BLOCKS blocks of 256 threads, rows and kernel taps in grid
// stride): sum |w_r|, max |w_r|, max m_r, max tap magnitude, at part[k * DET_BLOCKS + b].
template <typename R>
__global__ __launch_bounds__(256) void k_det_bounds(const cplx<R>* __restrict__ vis, const R* __restrict__ wt,
                                                    int64_t nrows, int dopsf, const R* __restrict__ ktab,
                                                    int64_t nk, int complex_tab, double* __restrict__ part) {
  __shared__ double scratch[4];
  double sw = 0.0, mw = 0.0, mm = 0.0, km = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t r = t0; r < nrows; r += stride) {
    const R w = wt[r];
    const double aw = fabs((double)w);
    sw += aw;
    mw = fmax(mw, aw);
    if (!dopsf) {
      const cplx<R> v = vis[r];
      const R vr = v.x * w, vi = v.y * w;  // as the gridders form them
      mm = fmax(mm, fabs((double)vr) + fabs((double)vi));
    }
  }
  if (dopsf) mm = mw;
  for (int64_t i = t0; i < nk; i += stride) {
    if (complex_tab) {
      const cplx<R> c = reinterpret_cast<const cplx<R>*>(ktab)[i];
      km = fmax(km, fabs((double)c.x) + fabs((double)c.y));
    } else {
      km = fmax(km, fabs((double)ktab[i]));
    }
  }
  const double v[4] = {block_fold<0>(sw, scratch), block_fold<1>(mw, scratch), block_fold<1>(mm, scratch),
                       block_fold<1>(km, scratch)};
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) part[(size_t)k * DET_BLOCKS + blockIdx.x] = v[k];
}

// One block of DET_BLOCKS threads: folds the partials in a fixed order into det[0] = 2^qw,
// det[1] = 2^-qw, det[2] = X = Kmax * M, and zeroes the integer sums of weights det[8 .. 8 + nsw).
__global__ __launch_bounds__(DET_BLOCKS) void k_det_scale(double* __restrict__ det, int64_t nsw, int complex_tab) {
  __shared__ double scratch[DET_BLOCKS / 64];
  const double* part = det + 8 + nsw;
  const int t = threadIdx.x;
  const double sw = block_fold<0>(part[t], scratch), mw = block_fold<1>(part[DET_BLOCKS + t], scratch),
               mm = block_fold<1>(part[2 * DET_BLOCKS + t], scratch),
               km = block_fold<1>(part[3 * DET_BLOCKS + t], scratch);
  if (t == 0) {
    const int qw = det_q(sw, mw);
    det[0] = ldexp(1.0, qw);
    det[1] = ldexp(1.0, -qw);
    det[2] = (complex_tab ? km : km * km) * mm;
  }
  long long* swi = reinterpret_cast<long long*>(det + 8);
  for (int64_t i = t; i < nsw; i += blockDim.x) swi[i] = 0;
}

// The scale exponent of tile t (every item of a tile, and its partial-tile sum, use the same).
__device__ inline int det_tile_q(const double* __restrict__ det, const uint32_t* __restrict__ tcnt, int t) {
  return det_q(det[2] * (double)tcnt[t], det[2]);
}

// sumwt[i] += the integer sum of weights i times 2^-qw.
__global__ void k_det_sumwt(const double* __restrict__ det, double* __restrict__ sumwt, int64_t nsw) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < nsw) sumwt[i] += (double)reinterpret_cast<const long long*>(det + 8)[i] * det[1];
}

// The state of a fixed-point call the gridders read.
struct DetArgs {
  const double* det;     // call constants (det[0..7]) and integer sums of weights
  const uint32_t* tcnt;  // [ntiles] entries per tile
};

// ---------------------------------------------------------------------------
// gridding, separable real kernel (PSWF), compile-time support
//
// One visibility at a time per wavefront, one kernel tap per lane: the 64 lanes of a
// ds_add touch 64 different pixels (no same-address serialisation), and with the planar
// LDS tile's row stride an odd multiple of 8 words the S=8 footprint spreads over all banks.
// A wave loads 64 entries (plus their vis and weight) coalesced, then broadcasts them one
// by one with readlane.
// ---------------------------------------------------------------------------

// Row stride (in elements) of a planar LDS tile of width w: an odd multiple of 8.
__host__ __device__ inline int lds_stride(int w) {
  int s = (w + 7) & ~7;
  if (((s >> 3) & 1) == 0) s += 8;
  return s;
}

__device__ __forceinline__ float bcast(float x, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j));
}
__device__ __forceinline__ double bcast(double x, int j) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), j);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ uint32_t bcast(uint32_t x, int j) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, j);
}

// Writes the T x T accumulator tile (planar float64: real parts from acc[0], imaginary parts
// from acc[im_off], row stride WS): to the grid, or (slot > 0) to partial tile slot-1 of a
// split tile.
// accum: add the tile to what the grid holds (streamed chunks) instead of overwriting it.
template <typename R>
__device__ inline void store_tile(const double* acc, int WS, int T, int N, const TilePos& tp, int slot,
                                  cplx<R>* __restrict__ grids, cplx<R>* __restrict__ partial, int im_off, int accum) {
  cplx<R>* g = grids + (size_t)tp.plane * N * N + (size_t)(tp.ty * T) * N + (size_t)tp.tx * T;
  cplx<R>* d = partial + (size_t)(slot > 0 ? slot - 1 : 0) * T * T;
  for (int i = threadIdx.x; i < T * T; i += blockDim.x) {
    const int y = i / T, x = i - (i / T) * T;
    const int o = y * WS + x;
    if (slot == 0) {
      double re = acc[o], im = acc[im_off + o];
      if (accum) {
        const cplx<R> old = g[(size_t)y * N + x];
        re += (double)old.x;
        im += (double)old.y;
      }
      g[(size_t)y * N + x] = mk<R>((R)re, (R)im);
    } else {
      d[i] = mk<R>((R)acc[o], (R)acc[im_off + o]);
    }
  }
}

// store_tile for fixed-point accumulators: values 2^-q times the integers; partial tiles keep
// the integers (summed exactly by k_reduce_tiles_det).
template <typename R>
__device__ inline void store_tile_det(const long long* acc, int WS, int T, int N, const TilePos& tp, int slot,
                                      cplx<R>* __restrict__ grids, longlong2* __restrict__ partial, int im_off,
                                      int accum, double inv_scale) {
  cplx<R>* g = grids + (size_t)tp.plane * N * N + (size_t)(tp.ty * T) * N + (size_t)tp.tx * T;
  longlong2* d = partial + (size_t)(slot > 0 ? slot - 1 : 0) * T * T;
  for (int i = threadIdx.x; i < T * T; i += blockDim.x) {
    const int y = i / T, x = i - (i / T) * T;
    const int o = y * WS + x;
    if (slot == 0) {
      double re = (double)acc[o] * inv_scale, im = (double)acc[im_off + o] * inv_scale;
      if (accum) {
        const cplx<R> old = g[(size_t)y * N + x];
        re += (double)old.x;
        im += (double)old.y;
      }
      g[(size_t)y * N + x] = mk<R>((R)re, (R)im);
    } else {
      longlong2 v;
      v.x = acc[o];
      v.y = acc[im_off + o];
      d[i] = v;
    }
  }
}

// Sums the partial tiles of every split tile, in slot order, into the grid.
// Grid: (split tiles, T*T/256 pixel blocks); each thread sums one pixel over the tile's slots
// in slot order (four loads in flight).
template <typename R>
__global__ void k_reduce_tiles(const int4* __restrict__ multi, const uint32_t* __restrict__ nmulti,
                               const cplx<R>* __restrict__ partial, cplx<R>* __restrict__ grids, int T, int ntx,
                               int N, int accum) {
  const int i = blockIdx.y * blockDim.x + threadIdx.x;
  if (i >= T * T) return;
  // the number of split tiles is read here, so the host never waits for it
  const uint32_t count = *nmulti;
  for (uint32_t mi = blockIdx.x; mi < count; mi += gridDim.x) {
  const int4 m = multi[mi];
  const TilePos tp = tile_pos(m.x, ntx);
  const size_t tt = (size_t)T * T;
  const cplx<R>* src = partial + (size_t)m.y * tt + i;
  R re = R(0), im = R(0);
  int k = 0;
  for (; k + 4 <= m.z; k += 4) {
    const cplx<R> a = src[(size_t)k * tt], b = src[(size_t)(k + 1) * tt], c = src[(size_t)(k + 2) * tt],
                  d = src[(size_t)(k + 3) * tt];
    re += a.x;
    im += a.y;
    re += b.x;
    im += b.y;
    re += c.x;
    im += c.y;
    re += d.x;
    im += d.y;
  }
  for (; k < m.z; ++k) {
    const cplx<R> a = src[(size_t)k * tt];
    re += a.x;
    im += a.y;
  }
  cplx<R>* dst = grids + (size_t)tp.plane * N * N + (size_t)(tp.ty * T + i / T) * N + (size_t)tp.tx * T + (i % T);
  if (accum) {
    const cplx<R> old = *dst;
    re += old.x;
    im += old.y;
  }
  *dst = mk<R>(re, im);
  }
}

// k_reduce_tiles for fixed-point partial tiles: integer sums, then 2^-q times the total.
template <typename R>
__global__ void k_reduce_tiles_det(const int4* __restrict__ multi, const uint32_t* __restrict__ nmulti,
                                   const longlong2* __restrict__ partial, cplx<R>* __restrict__ grids, int T, int ntx,
                                   int N, int accum, const DetArgs da) {
  const int i = blockIdx.y * blockDim.x + threadIdx.x;
  if (i >= T * T) return;
  const uint32_t count = *nmulti;
  for (uint32_t mi = blockIdx.x; mi < count; mi += gridDim.x) {
    const int4 m = multi[mi];
    const double inv_scale = ldexp(1.0, -det_tile_q(da.det, da.tcnt, m.x));
    const TilePos tp = tile_pos(m.x, ntx);
    const size_t tt = (size_t)T * T;
    const longlong2* src = partial + (size_t)m.y * tt + i;
    long long re = 0, im = 0;
    for (int k = 0; k < m.z; ++k) {
      const longlong2 a = src[(size_t)k * tt];
      re += a.x;
      im += a.y;
    }
    cplx<R>* dst = grids + (size_t)tp.plane * N * N + (size_t)(tp.ty * T + i / T) * N + (size_t)tp.tx * T + (i % T);
    double r = (double)re * inv_scale, q = (double)im * inv_scale;
    if (accum) {
      const cplx<R> old = *dst;
      r += (double)old.x;
      q += (double)old.y;
    }
    *dst = mk<R>((R)r, (R)q);
  }
}

// Per-entry values a wavefront broadcasts while it walks its 64 entries:
//   sa: footprint origin as an accumulator index, cy*WS + cx + SA_BIAS, and bit 31 set when
//       the whole S x S footprint lies inside the tile (no masking needed);
//   ko: offsets of the entry's kernel rows in the LDS kernel table, yoff*S | xoff*S << 16.
constexpr int SA_BIAS = 1 << 20;
constexpr uint32_t SA_INNER = 0x80000000u;

// Kernel values of the taps this lane owns, for the sub-sample offsets in ko.
template <typename R, int TPL>
__device__ __forceinline__ void sep_taps(const R* lk, uint32_t ko, const int* ta, const int* tb, R* ky, R* kx) {
#pragma unroll
  for (int g = 0; g < TPL; ++g) {
    ky[g] = lk[(int)(ko & 0xffffu) + ta[g]];
    kx[g] = lk[(int)(ko >> 16) + tb[g]];
  }
}

// Adds (jr, ji) times this lane's taps of one entry into the planar float64 accumulator tile
// (real parts acc[0, TW + 64), imaginary parts from acc[TW + 64]; the last 64 of each are a
// trash row).  Inner footprints take the unmasked path (one uniform branch); taps
// of an edge footprint outside the tile add 0 to this lane's own slot of the trash row:
// both paths issue the same two LDS adds per tap, so the wait counts the compiler
// derives for the software-pipelined kernel reads stay exact.
template <typename R, int S, int TPL, bool DET>
__device__ __forceinline__ void sep_add(double* acc, int T, int TW, int lane, uint32_t sa, uint32_t pk,
                                        const int* ta, const int* tb, const int* loff, const R* ky, const R* kx,
                                        R jr, R ji, double scale) {
  constexpr int TAPS = S * S;
  const int sb = (int)(sa & ~SA_INNER) - SA_BIAS;
  if (sa & SA_INNER) {
#pragma unroll
    for (int g = 0; g < TPL; ++g) {
      const bool ok = TAPS % 64 == 0 || lane + 64 * g < TAPS;
      const R c = ok ? ky[g] * kx[g] : R(0);
      const int o = ok ? sb + loff[g] : TW + lane;
      acc_add<DET>(acc + o, (double)(c * jr), scale);
      acc_add<DET>(acc + (TW + 64) + o, (double)(c * ji), scale);
    }
  } else {
    const int cx = entry_lx(pk) - S, cy = entry_ly(pk) - S;
#pragma unroll
    for (int g = 0; g < TPL; ++g) {
      const int x = cx + tb[g], y = cy + ta[g];
      const bool ok =
          (TAPS % 64 == 0 || lane + 64 * g < TAPS) && (unsigned)x < (unsigned)T && (unsigned)y < (unsigned)T;
      const R c = ok ? ky[g] * kx[g] : R(0);
      const int o = ok ? sb + loff[g] : TW + lane;
      acc_add<DET>(acc + o, (double)(c * jr), scale);
      acc_add<DET>(acc + (TW + 64) + o, (double)(c * ji), scale);
    }
  }
}

// One work item of the separable gridder (the body of k_grid_sep and of its batched form).
// DET: fixed-point accumulation with the scales of det (ARLMI_FLAG_DETERMINISTIC); partial then
// holds integer tiles (longlong2) and the sums of weights go to det's integer sums.
template <typename R, int S, bool DET>
__device__ __forceinline__ void grid_sep_item(const int4 it, unsigned char* smem, double* red,
                                              const uint2* __restrict__ entries, const cplx<R>* __restrict__ vis,
                                              const R* __restrict__ wt, int dopsf, const R* __restrict__ ktab, int os,
                                              cplx<R>* __restrict__ grids, cplx<R>* __restrict__ partial,
                                              double* __restrict__ sumwt, int T, int ntx, int N, int accum,
                                              const DetArgs da) {
  constexpr int TAPS = S * S;
  constexpr int TPL = (TAPS + 63) / 64;  // taps per lane
  const int WS = lds_stride(T);
  const int TW = T * WS;
  double* acc = reinterpret_cast<double*>(smem);  // [2][TW + 64]: real, imaginary (tile + trash row)
  R* lk = reinterpret_cast<R*>(acc + 2 * (TW + 64));
  if (accum && it.z == 0) return;  // an empty tile adds nothing
  for (int i = threadIdx.x; i < 2 * (TW + 64); i += blockDim.x) acc[i] = 0.0;
  for (int i = threadIdx.x; i < os * S; i += blockDim.x) lk[i] = ktab[i];
  __syncthreads();
  const int lane = __lane_id(), wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwave = blockDim.x >> 6;
  int ta[TPL], tb[TPL], loff[TPL];
#pragma unroll
  for (int g = 0; g < TPL; ++g) {
    const int t = lane + 64 * g;
    ta[g] = t < TAPS ? t / S : 0;
    tb[g] = t < TAPS ? t % S : 0;
    loff[g] = ta[g] * WS + tb[g];
  }
  double scale = 0.0, inv_scale = 0.0, wscale = 0.0;
  if constexpr (DET) {
    const int q = det_tile_q(da.det, da.tcnt, it.x);
    scale = ldexp(1.0, q);
    inv_scale = ldexp(1.0, -q);
    wscale = da.det[0];
  }
  double swt = 0.0;
  long long swi = 0;
  // each wave takes one contiguous share of the item's entries, so every wave of the group
  // has work even for small items
  const int per = __builtin_amdgcn_readfirstlane((it.z + nwave - 1) / nwave);
  const int e0 = min(it.z, wave * per), e1 = min(it.z, e0 + per);
  for (int base = e0; base < e1; base += 64) {
    const int e = base + lane;
    uint32_t pk = 0;  // lanes past the end hold entry 0 of the tile origin: offsets in range
    R vr = R(0), vi = R(0);
    if (e < e1) {
      const uint2 en = entries[it.y + e];
      pk = en.y;
      const R w = wt[en.x];
      if (dopsf) {
        vr = w;
      } else {
        const cplx<R> v = vis[en.x];
        vr = v.x * w;
        vi = v.y * w;
      }
      if (pk & ENTRY_START) {
        if constexpr (DET) swi += det_round((double)w, wscale);
        else swt += (double)w;
      }
    }
    const int cx = entry_lx(pk) - S, cy = entry_ly(pk) - S;
    const bool inner = cx >= 0 && cy >= 0 && cx + S <= T && cy + S <= T;
    const uint32_t sav = (uint32_t)(cy * WS + cx + SA_BIAS) | (inner ? SA_INNER : 0u);
    const uint32_t kov = (uint32_t)(entry_yoff(pk) * S) | ((uint32_t)(entry_xoff(pk) * S) << 16);
    const int cnt = __builtin_amdgcn_readfirstlane(min(64, e1 - base));
    // Software pipeline over two register sets A and B: the kernel values of entry j+1 are read
    // from LDS before the adds of entry j are issued.  LDS operations retire in order, so the
    // wait for a set's reads never includes adds.  The prologue issues A reads, 2*TPL adds of
    // zero to the trash row, then B reads: the same pattern of outstanding operations as the
    // loop's back edge, so the wait counts at the loop head are exact on both paths.
    R kya[TPL], kxa[TPL], kyb[TPL], kxb[TPL];
    sep_taps<R, TPL>(lk, bcast(kov, 0), ta, tb, kya, kxa);
#pragma unroll
    for (int g = 0; g < TPL; ++g) {
      acc_add<DET>(acc + TW + lane, 0.0, scale);
      acc_add<DET>(acc + (TW + 64) + TW + lane, 0.0, scale);
    }
    sep_taps<R, TPL>(lk, bcast(kov, 1), ta, tb, kyb, kxb);
    for (int j = 0;; j += 2) {
      sep_add<R, S, TPL, DET>(acc, T, TW, lane, bcast(sav, j), bcast(pk, j), ta, tb, loff, kya, kxa, bcast(vr, j),
                              bcast(vi, j), scale);
      if (j + 1 >= cnt) break;
      sep_taps<R, TPL>(lk, bcast(kov, (j + 2) & 63), ta, tb, kya, kxa);
      sep_add<R, S, TPL, DET>(acc, T, TW, lane, bcast(sav, j + 1), bcast(pk, j + 1), ta, tb, loff, kyb, kxb,
                              bcast(vr, j + 1), bcast(vi, j + 1), scale);
      if (j + 2 >= cnt) break;
      sep_taps<R, TPL>(lk, bcast(kov, (j + 3) & 63), ta, tb, kyb, kxb);
    }
  }
  __syncthreads();
  const TilePos tp = tile_pos(it.x, ntx);
  if constexpr (DET) {
    store_tile_det<R>(reinterpret_cast<const long long*>(acc), WS, T, N, tp, it.w, grids,
                      reinterpret_cast<longlong2*>(partial), TW + 64, accum, inv_scale);
    block_add_i64(swi, reinterpret_cast<unsigned long long*>(const_cast<double*>(da.det) + 8) + tp.plane,
                  reinterpret_cast<long long*>(red));
  } else {
    store_tile<R>(acc, WS, T, N, tp, it.w, grids, partial, TW + 64, accum);
    block_add_double(swt, &sumwt[tp.plane], red);
  }
}

// The separable gridder over entries grouped by footprint origin within their tile (the plan's
// origin sort; at most 64 taps, one a lane).  A wave walks its 64 entries adding the tap times
// vis * weight into registers and issues the two LDS adds of a run only at its last entry (a
// run: consecutive entries of one origin, cut at every 32nd lane so that no register sum spans
// more than 32 entries; the sums are in the grid precision, as each tap's product is in the
// kernel without runs).  On the 2048^2 w-stack workload 38.2 M entries make about 13 M runs:
// k_grid_sep 2.36 -> 1.88 ms.  Measured and not used: float64 run sums (2.22 ms, the image error
// against float64 unchanged), the lane's taps of all sub-sample offsets held in registers (the
// pick by offset compiles to select chains), the kernel values of the next entry or of groups of
// 4 and 8 entries read ahead (2.09-2.37 ms: the reads then wait behind the run's adds).
template <typename R, int S>
__device__ __forceinline__ void grid_sep_item_merge(const int4 it, unsigned char* smem, double* red,
                                                    const uint2* __restrict__ entries, const cplx<R>* __restrict__ vis,
                                                    const R* __restrict__ wt, int dopsf, const R* __restrict__ ktab,
                                                    int os, cplx<R>* __restrict__ grids, cplx<R>* __restrict__ partial,
                                                    double* __restrict__ sumwt, int T, int ntx, int N, int accum) {
  constexpr int TAPS = S * S;
  static_assert(TAPS <= 64, "one tap a lane");
  const int WS = lds_stride(T);
  const int TW = T * WS;
  double* acc = reinterpret_cast<double*>(smem);  // [2][TW + 64]: real, imaginary (tile + trash row)
  R* lk = reinterpret_cast<R*>(acc + 2 * (TW + 64));
  if (accum && it.z == 0) return;  // an empty tile adds nothing
  for (int i = threadIdx.x; i < 2 * (TW + 64); i += blockDim.x) acc[i] = 0.0;
  for (int i = threadIdx.x; i < os * S; i += blockDim.x) lk[i] = ktab[i];
  __syncthreads();
  const int lane = __lane_id(), wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwave = blockDim.x >> 6;
  const bool tap = TAPS == 64 || lane < TAPS;
  const int ta = tap ? lane / S : 0, tb = tap ? lane % S : 0;
  const int loff = ta * WS + tb;
  double swt = 0.0;
  const int per = __builtin_amdgcn_readfirstlane((it.z + nwave - 1) / nwave);
  const int e0 = min(it.z, wave * per), e1 = min(it.z, e0 + per);
  for (int base = e0; base < e1; base += 64) {
    const int e = base + lane;
    uint32_t pk = 0;
    R vr = R(0), vi = R(0);
    if (e < e1) {
      const uint2 en = entries[it.y + e];
      pk = en.y;
      const R w = wt[en.x];
      if (dopsf) {
        vr = w;
      } else {
        const cplx<R> v = vis[en.x];
        vr = v.x * w;
        vi = v.y * w;
      }
      if (pk & ENTRY_START) swt += (double)w;
    }
    const int cx = entry_lx(pk) - S, cy = entry_ly(pk) - S;
    const bool inner = cx >= 0 && cy >= 0 && cx + S <= T && cy + S <= T;
    const uint32_t sav = (uint32_t)(cy * WS + cx + SA_BIAS) | (inner ? SA_INNER : 0u);
    const int cnt = __builtin_amdgcn_readfirstlane(min(64, e1 - base));
    const uint32_t nsav = (uint32_t)__shfl_down((int)sav, 1);
    const bool last = lane == cnt - 1 || (lane & 31) == 31 || (lane < cnt - 1 && nsav != sav);
    const unsigned long long ends = __ballot(last);
    R sr = R(0), si = R(0);
    for (int j = 0; j < cnt; ++j) {
      const uint32_t pj = bcast(pk, j);
      const R c = tap ? lk[entry_yoff(pj) * S + ta] * lk[entry_xoff(pj) * S + tb] : R(0);
      sr += c * bcast(vr, j);
      si += c * bcast(vi, j);
      if ((ends >> j) & 1ull) {
        const uint32_t sa = bcast(sav, j);
        const int sb = (int)(sa & ~SA_INNER) - SA_BIAS;
        bool ok = tap;
        if (!(sa & SA_INNER)) {  // an edge footprint: taps outside the tile go to the trash row
          const int x = entry_lx(pj) - S + tb, y = entry_ly(pj) - S + ta;
          ok = ok && (unsigned)x < (unsigned)T && (unsigned)y < (unsigned)T;
        }
        const int o = ok ? sb + loff : TW + lane;
        acc_add<false>(acc + o, ok ? (double)sr : 0.0, 0.0);
        acc_add<false>(acc + (TW + 64) + o, ok ? (double)si : 0.0, 0.0);
        sr = R(0);
        si = R(0);
      }
    }
  }
  __syncthreads();
  const TilePos tp = tile_pos(it.x, ntx);
  store_tile<R>(acc, WS, T, N, tp, it.w, grids, partial, TW + 64, accum);
  block_add_double(swt, &sumwt[tp.plane], red);
}

template <typename R, int S>
__global__ __launch_bounds__(GRID_THREADS) void k_grid_sep_merge(const int4* __restrict__ items,
                                                                 const uint2* __restrict__ entries,
                                                                 const cplx<R>* __restrict__ vis,
                                                                 const R* __restrict__ wt, int dopsf,
                                                                 const R* __restrict__ ktab, int os,
                                                                 cplx<R>* __restrict__ grids,
                                                                 cplx<R>* __restrict__ partial,
                                                                 double* __restrict__ sumwt, int T, int ntx, int N,
                                                                 int accum) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[GRID_THREADS / 64];
  grid_sep_item_merge<R, S>(items[blockIdx.x], smem, red, entries, vis, wt, dopsf, ktab, os, grids, partial, sumwt,
                            T, ntx, N, accum);
}

template <typename R, int S, bool DET>
__global__ __launch_bounds__(GRID_THREADS) void k_grid_sep(const int4* __restrict__ items,
                                                           const uint2* __restrict__ entries,
                                                           const cplx<R>* __restrict__ vis, const R* __restrict__ wt,
                                                           int dopsf, const R* __restrict__ ktab, int os,
                                                           cplx<R>* __restrict__ grids, cplx<R>* __restrict__ partial,
                                                           double* __restrict__ sumwt, int T, int ntx, int N,
                                                           int accum, const DetArgs da) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[GRID_THREADS / 64];
  grid_sep_item<R, S, DET>(items[blockIdx.x], smem, red, entries, vis, wt, dopsf, ktab, os, grids, partial, sumwt, T,
                           ntx, N, accum, da);
}

// Per-plan arguments of the batched separable kernels: the plan's work items are blocks
// [base, base + nitems) of the batched launch.
struct BatchDesc {
  const int4* gitems;
  const int4* ditems;
  const uint2* entries;
  const void* vis;
  const void* wt;
  const void* ktab;
  void* grids;
  void* partial;
  double* sumwt;
  void* vis_out;
  const int4* multi;       // split tiles
  const uint32_t* nmulti;  // their number (device)
  const uint32_t* orphans; // records of rows in no plane (orphan_words words each, the row first), or null
  int64_t norphans;
  int orphan_words;
  int os, T, ntx, N, nsw;
  int gbase, ngitems, dbase, nditems;
  int perm_n1;  // row order of the grids the degridder reads (perm_row_of)
  int runs;     // entries grouped by footprint origin: the gridder adds runs (k_grid_sep_merge)
};

// The descriptor whose [base, base + count) block range holds global block b (descriptors in
// block order; a scalar search over at most a few hundred plans).
__device__ __forceinline__ int find_desc(const BatchDesc* __restrict__ d, int nd, int b, bool degrid) {
  int lo = 0, hi = nd - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((degrid ? d[mid].dbase : d[mid].gbase) <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Every plan's separable gridding in one launch: block b runs work item b - gbase of the plan
// that holds it.
template <typename R, int S>
__global__ __launch_bounds__(GRID_THREADS) void k_grid_sep_batch(const BatchDesc* __restrict__ descs, int nd,
                                                                 int dopsf) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[GRID_THREADS / 64];
  const int b = blockIdx.x;
  const BatchDesc& d = descs[find_desc(descs, nd, b, false)];
  if constexpr (S * S <= 64) {
    if (d.runs) {
      grid_sep_item_merge<R, S>(d.gitems[b - d.gbase], smem, red, d.entries, (const cplx<R>*)d.vis, (const R*)d.wt,
                                dopsf, (const R*)d.ktab, d.os, (cplx<R>*)d.grids, (cplx<R>*)d.partial, d.sumwt, d.T,
                                d.ntx, d.N, 0);
      return;
    }
  }
  grid_sep_item<R, S, false>(d.gitems[b - d.gbase], smem, red, d.entries, (const cplx<R>*)d.vis, (const R*)d.wt,
                             dopsf, (const R*)d.ktab, d.os, (cplx<R>*)d.grids, (cplx<R>*)d.partial, d.sumwt, d.T,
                             d.ntx, d.N, 0, DetArgs{});
}

// ---------------------------------------------------------------------------
// gridding, full complex kernel table (AW / box / any CF), runtime support
// ---------------------------------------------------------------------------

// Complex kernel values of this lane's taps (t = lane + 64 q) of one entry, from the table
// slice at kbase (L2-resident: the whole table is read by every tile).  Every lane loads (a lane
// past the last tap loads tap 0 and scales it by 0): with a per-lane condition on the load, the
// compiler lowered the pipelined entry loop with per-lane exit masks and merged the load's lane
// mask into them, so that when the last tap group was only partly filled (S = 14 with 4 taps per
// lane, S = 22 with 8) the lanes past the last tap left the loop after its first entries and those
// entries' later adds were lost; with unconditional loads the loop stays scalar.
template <typename R, int MT>
__device__ __forceinline__ void tab_taps(const cplx<R>* __restrict__ tab, uint32_t kbase, int taps, int lane,
                                         cplx<R>* k) {
#pragma unroll
  for (int q = 0; q < MT; ++q) {
    const int t = lane + 64 * q;
    const R m = t < taps ? R(1) : R(0);
    const cplx<R> v = tab[(size_t)kbase + (t < taps ? t : 0)];
    k[q] = mk<R>(v.x * m, v.y * m);
  }
}

// conj(c) * (jr + i ji) for this lane's taps of one entry, into the interleaved accumulators;
// same masking and trash-row scheme as sep_add.
template <typename R, int MT, bool DET>
__device__ __forceinline__ void tab_add(double* acc, int T, int TW, int S, int taps, int lane, uint32_t sa,
                                        uint32_t pk, const int* ta, const int* tb, const int* loff,
                                        const cplx<R>* k, R jr, R ji, double scale) {
  const int sb = (int)(sa & ~SA_INNER) - SA_BIAS;
  const bool inner = (sa & SA_INNER) != 0;
  const int cx = entry_lx(pk) - S, cy = entry_ly(pk) - S;
#pragma unroll
  for (int q = 0; q < MT; ++q) {
    const int x = cx + tb[q], y = cy + ta[q];
    const bool ok = lane + 64 * q < taps && (inner || ((unsigned)x < (unsigned)T && (unsigned)y < (unsigned)T));
    const int o = ok ? sb + loff[q] : TW + lane;
    const R cr = ok ? k[q].x : R(0), ci = ok ? k[q].y : R(0);
    acc_add<DET>(acc + o, (double)(cr * jr + ci * ji), scale);
    acc_add<DET>(acc + (TW + 64) + o, (double)(cr * ji - ci * jr), scale);
  }
}

// Gridding with a full complex kernel table (AW / box / any CF).  MT = taps per lane held in
// registers (S*S <= 64*MT); the next entry's kernel values are loaded from L2 while the current
// entry's adds run.  MT = 0: any support, taps loaded and added one group of 64 at a time.
template <typename R, int MT, bool DET>
__global__ __launch_bounds__(GRID_THREADS) void k_grid_tab(const int4* __restrict__ items,
                                                           const uint2* __restrict__ entries,
                                                           const uint16_t* __restrict__ entry_z,
                                                           const cplx<R>* __restrict__ vis, const R* __restrict__ wt,
                                                           int dopsf, const cplx<R>* __restrict__ tab, int os, int S,
                                                           cplx<R>* __restrict__ grids, cplx<R>* __restrict__ partial,
                                                           double* __restrict__ sumwt, int T, int ntx, int N,
                                                           int accum, const DetArgs da) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[GRID_THREADS / 64];
  const int WS = lds_stride(T);
  const int TW = T * WS;
  double* acc = reinterpret_cast<double*>(smem);  // [2][TW + 64]: real, imaginary (tile + trash row)
  const int4 it = items[blockIdx.x];
  if (accum && it.z == 0) return;  // an empty tile adds nothing
  for (int i = threadIdx.x; i < 2 * (TW + 64); i += blockDim.x) acc[i] = 0.0;
  __syncthreads();
  // The wave index through readfirstlane, so that the compiler knows the entry loop's bounds to be
  // wave-uniform and keeps the loop scalar (see tab_taps).
  const int lane = __lane_id(), wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwave = blockDim.x >> 6;
  const int taps = S * S;
  constexpr int NQ = MT > 0 ? MT : 1;
  int ta[NQ], tb[NQ], loff[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int t = lane + 64 * q;
    ta[q] = t < taps ? t / S : 0;
    tb[q] = t < taps ? t % S : 0;
    loff[q] = ta[q] * WS + tb[q];
  }
  double scale = 0.0, inv_scale = 0.0, wscale = 0.0;
  if constexpr (DET) {
    const int q = det_tile_q(da.det, da.tcnt, it.x);
    scale = ldexp(1.0, q);
    inv_scale = ldexp(1.0, -q);
    wscale = da.det[0];
  }
  double swt = 0.0;
  long long swi = 0;
  const int per = __builtin_amdgcn_readfirstlane((it.z + nwave - 1) / nwave);
  const int e0 = min(it.z, wave * per), e1 = min(it.z, e0 + per);
  for (int base = e0; base < e1; base += 64) {
    const int e = base + lane;
    uint32_t pk = 0, zz = 0;
    R vr = R(0), vi = R(0);
    if (e < e1) {
      const uint2 en = entries[it.y + e];
      pk = en.y;
      zz = entry_z ? (uint32_t)entry_z[it.y + e] : 0u;
      const R w = wt[en.x];
      if (dopsf) {
        vr = w;
      } else {
        const cplx<R> v = vis[en.x];
        vr = v.x * w;
        vi = v.y * w;
      }
      if (pk & ENTRY_START) {
        if constexpr (DET) swi += det_round((double)w, wscale);
        else swt += (double)w;
      }
    }
    const int cx = entry_lx(pk) - S, cy = entry_ly(pk) - S;
    const bool inner = cx >= 0 && cy >= 0 && cx + S <= T && cy + S <= T;
    const uint32_t sav = (uint32_t)(cy * WS + cx + SA_BIAS) | (inner ? SA_INNER : 0u);
    const uint32_t kbv = (uint32_t)(((int)zz * os + entry_yoff(pk)) * os + entry_xoff(pk)) * (uint32_t)taps;
    const int cnt = __builtin_amdgcn_readfirstlane(min(64, e1 - base));
    if constexpr (MT > 0) {
      cplx<R> ka[MT], kb[MT];
      tab_taps<R, MT>(tab, bcast(kbv, 0), taps, lane, ka);
      tab_taps<R, MT>(tab, bcast(kbv, 1), taps, lane, kb);
      for (int j = 0;; j += 2) {
        tab_add<R, MT, DET>(acc, T, TW, S, taps, lane, bcast(sav, j), bcast(pk, j), ta, tb, loff, ka,
                            bcast(vr, j), bcast(vi, j), scale);
        if (j + 1 >= cnt) break;
        tab_taps<R, MT>(tab, bcast(kbv, (j + 2) & 63), taps, lane, ka);
        tab_add<R, MT, DET>(acc, T, TW, S, taps, lane, bcast(sav, j + 1), bcast(pk, j + 1), ta, tb, loff, kb,
                            bcast(vr, j + 1), bcast(vi, j + 1), scale);
        if (j + 2 >= cnt) break;
        tab_taps<R, MT>(tab, bcast(kbv, (j + 3) & 63), taps, lane, kb);
      }
    } else {
      // Any support: the entry's taps in groups of 64 (group q holds taps 64 q + lane), only the
      // groups whose footprint rows meet the tile, four groups' table reads issued before their
      // adds (the reads come from L2, and one read then one wait per group left the loop
      // latency-bound).  a = t / S by a multiply-high with ceil(2^32 / S), exact for t < 2^32 / S.
      const uint32_t inv_s = 0xffffffffu / (uint32_t)S + 1u;
      for (int j = 0; j < cnt; ++j) {
        const uint32_t p = bcast(pk, j);
        const int sb = (int)(bcast(sav, j) & ~SA_INNER) - SA_BIAS;
        const cplx<R>* k = tab + bcast(kbv, j);
        const R jr = bcast(vr, j), ji = bcast(vi, j);
        const int jx = entry_lx(p) - S, jy = entry_ly(p) - S;
        const int a_lo = max(0, -jy), a_hi = min(S, T - jy);  // footprint rows inside the tile
        const int q_lo = (a_lo * S) >> 6, q_hi = (a_hi * S + 63) >> 6;
        for (int q0 = q_lo; q0 < q_hi; q0 += 4) {
          cplx<R> c[4];
          int o[4];
          bool ok[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int t = lane + 64 * (q0 + u);
            const int a = (int)__umulhi((uint32_t)t, inv_s), b = t - a * S;
            const int x = jx + b, y = jy + a;
            ok[u] = q0 + u < q_hi && t < taps && (unsigned)x < (unsigned)T && (unsigned)y < (unsigned)T;
            o[u] = sb + a * WS + b;
            c[u] = ok[u] ? k[t] : mk<R>(R(0), R(0));
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            if (ok[u]) {
              acc_add<DET>(acc + o[u], (double)(c[u].x * jr + c[u].y * ji), scale);
              acc_add<DET>(acc + (TW + 64) + o[u], (double)(c[u].x * ji - c[u].y * jr), scale);
            }
          }
        }
      }
    }
  }
  __syncthreads();
  const TilePos tp = tile_pos(it.x, ntx);
  if constexpr (DET) {
    store_tile_det<R>(reinterpret_cast<const long long*>(acc), WS, T, N, tp, it.w, grids,
                      reinterpret_cast<longlong2*>(partial), TW + 64, accum, inv_scale);
    block_add_i64(swi, reinterpret_cast<unsigned long long*>(const_cast<double*>(da.det) + 8) + tp.plane,
                  reinterpret_cast<long long*>(red));
  } else {
    store_tile<R>(acc, WS, T, N, tp, it.w, grids, partial, TW + 64, accum);
    block_add_double(swt, &sumwt[tp.plane], red);
  }
}

// ---------------------------------------------------------------------------
// degridding
// ---------------------------------------------------------------------------

// Stage the tile and its halo (Wd x Wd grid values from the tile origin, 0 past the grid edge)
// in LDS.  When every thread has at most MAXL values to move, all of its loads are issued before
// any LDS store, so a block waits for one HBM latency, not one per value.
template <typename R, int MAXL = 16>
__device__ inline void load_halo(const cplx<R>* __restrict__ grids, R* lds, int Wd, int T, int N, const TilePos& tp,
                                 int perm_n1) {
  const cplx<R>* g = grids + (size_t)tp.plane * N * N;
  const int y0 = tp.ty * T, x0 = tp.tx * T;
  const int n = Wd * Wd;
  if (n <= MAXL * (int)blockDim.x) {
    cplx<R> v[MAXL];
#pragma unroll
    for (int j = 0; j < MAXL; ++j) {
      const int i = threadIdx.x + j * (int)blockDim.x;
      v[j] = mk<R>(R(0), R(0));
      if (i < n) {
        const int y = i / Wd, x = i - (i / Wd) * Wd;
        const int gy = y0 + y, gx = x0 + x;
        if (gy < N && gx < N) v[j] = g[(size_t)perm_row_of(gy, N, perm_n1) * N + gx];
      }
    }
#pragma unroll
    for (int j = 0; j < MAXL; ++j) {
      const int i = threadIdx.x + j * (int)blockDim.x;
      if (i < n) {
        lds[2 * i] = v[j].x;
        lds[2 * i + 1] = v[j].y;
      }
    }
    return;
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int y = i / Wd, x = i - (i / Wd) * Wd;
    const int gy = y0 + y, gx = x0 + x;
    cplx<R> v = mk<R>(R(0), R(0));
    if (gy < N && gx < N) v = g[(size_t)perm_row_of(gy, N, perm_n1) * N + gx];
    lds[2 * i] = v.x;
    lds[2 * i + 1] = v.y;
  }
}

// Work-group size of the degridders (one entry per thread; an item holds up to 4096 entries):
// measured best at 512 in single and 1024 in double precision (with origin-grouped entries too:
// 1024 threads in single precision take 0.82 against 0.75 ms on the 2048^2 w-stack workload).
template <typename R>
constexpr int degrid_threads() {
  return sizeof(R) == 4 ? 512 : 1024;
}

// One work item of the separable degridder (the body of k_degrid_sep and of its batched form).
template <typename R, int S>
__device__ __forceinline__ void degrid_sep_item(const int4 it, unsigned char* smem, const uint2* __restrict__ entries,
                                                const cplx<R>* __restrict__ grids, const R* __restrict__ ktab,
                                                int os, cplx<R>* __restrict__ vis_out, int T, int ntx, int N,
                                                int perm_n1) {
  const int Wd = T + S;
  R* lds = reinterpret_cast<R*>(smem);
  R* lk = lds + 2 * Wd * Wd;
  const TilePos tp = tile_pos(it.x, ntx);
  load_halo<R>(grids, lds, Wd, T, N, tp, perm_n1);
  for (int i = threadIdx.x; i < os * S; i += blockDim.x) lk[i] = ktab[i];
  __syncthreads();
  for (int e = threadIdx.x; e < it.z; e += blockDim.x) {
    const uint2 en = entries[it.y + e];
    const uint32_t p = en.y;
    const int fx = entry_lx(p) - S, fy = entry_ly(p) - S, xo = entry_xoff(p), yo = entry_yoff(p);
    R kx[S];
#pragma unroll
    for (int b = 0; b < S; ++b) kx[b] = lk[xo * S + b];
    const R* base = lds + 2 * (fy * Wd + fx);
    R ar = R(0), ai = R(0);
#pragma unroll
    for (int a = 0; a < S; ++a) {
      const R* row = base + 2 * a * Wd;
      R rr = R(0), ri = R(0);
#pragma unroll
      for (int b = 0; b < S; ++b) {
        rr += kx[b] * row[2 * b];
        ri += kx[b] * row[2 * b + 1];
      }
      const R ka = lk[yo * S + a];
      ar += ka * rr;
      ai += ka * ri;
    }
    vis_out[en.x] = mk<R>(ar, ai);
  }
}

template <typename R, int S>
__global__ __launch_bounds__(degrid_threads<R>()) void k_degrid_sep(const int4* __restrict__ items, const uint2* __restrict__ entries,
                                                    const cplx<R>* __restrict__ grids, const R* __restrict__ ktab,
                                                    int os, cplx<R>* __restrict__ vis_out, int T, int ntx, int N,
                                                    int perm_n1) {
  extern __shared__ __align__(16) unsigned char smem[];
  degrid_sep_item<R, S>(items[blockIdx.x], smem, entries, grids, ktab, os, vis_out, T, ntx, N, perm_n1);
}

template <typename R, int S>
__global__ __launch_bounds__(degrid_threads<R>()) void k_degrid_sep_batch(const BatchDesc* __restrict__ descs, int nd) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blockIdx.x;
  const BatchDesc& d = descs[find_desc(descs, nd, b, true)];
  degrid_sep_item<R, S>(d.ditems[b - d.dbase], smem, d.entries, (const cplx<R>*)d.grids, (const R*)d.ktab, d.os,
                        (cplx<R>*)d.vis_out, d.T, d.ntx, d.N, d.perm_n1);
}

// DPP moves (0 where a row is masked off) of a 32-bit or 64-bit value within the wavefront.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over each group of L consecutive lanes (L a power of two, 2..64), all lanes active: quad
// permutations, half-row and row mirrors, then row broadcasts.  Every lane of a group holds the
// sum for L <= 16; for L = 32 and 64 the group's last lane does.
template <int L, typename T>
__device__ __forceinline__ T group_sum(T v) {
  if (L >= 2) v += dpp_move<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  if (L >= 4) v += dpp_move<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  if (L >= 8) v += dpp_move<0x141, 0xf>(v);  // row_half_mirror
  if (L >= 16) v += dpp_move<0x140, 0xf>(v); // row_mirror
  if (L >= 32) v += dpp_move<0x142, 0xa>(v); // row_bcast:15 into rows 1 and 3
  if (L >= 64) v += dpp_move<0x143, 0xc>(v); // row_bcast:31 into rows 2 and 3
  return v;
}

constexpr int DEGRID_TAB_THREADS = 256;

// Degridding with a full complex kernel table (AW / any CF).  A group of L lanes (L the power of
// two >= S) handles one entry, one footprint column per lane: at step b the group reads row b of
// the entry's kernel slice and row fy + b of the LDS tile + halo, S consecutive values each (so
// the table reads coalesce and the LDS reads are consecutive), and at the end it sums its
// columns with DPP; 64 / L entries per wavefront step.
template <typename R, int L>
__global__ __launch_bounds__(DEGRID_TAB_THREADS) void k_degrid_tab(const int4* __restrict__ items,
                                                                  const uint2* __restrict__ entries,
                                                                  const uint16_t* __restrict__ entry_z,
                                                                  const cplx<R>* __restrict__ grids,
                                                                  const cplx<R>* __restrict__ tab, int os, int S,
                                                                  cplx<R>* __restrict__ vis_out, int T, int ntx,
                                                                  int N, int perm_n1) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int E = 64 / L;  // entries per wavefront step
  const int Wd = T + S;
  R* lds = reinterpret_cast<R*>(smem);
  const int4 it = items[blockIdx.x];
  const TilePos tp = tile_pos(it.x, ntx);
  load_halo<R, 32>(grids, lds, Wd, T, N, tp, perm_n1);
  __syncthreads();
  const int lane = __lane_id(), wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwave = blockDim.x >> 6;
  const int slot = lane / L, a = lane % L;
  const bool col_ok = a < S;
  const int ac = col_ok ? a : 0;  // lanes past the footprint read column 0 and contribute 0
  const int per = __builtin_amdgcn_readfirstlane((it.z + nwave - 1) / nwave);
  const int e0 = min(it.z, wave * per), e1 = min(it.z, e0 + per);
  for (int base = e0; base < e1; base += E) {
    const int e = base + slot;
    const bool ok = e < e1;
    uint32_t row = 0, p = 0;
    int z = 0;
    if (ok) {
      const uint2 en = entries[it.y + e];
      row = en.x;
      p = en.y;
      if (entry_z) z = (int)entry_z[it.y + e];
    }
    // a slot past the end works on the tile origin with sub-sample 0 (in range) and stores nothing
    const int fx = ok ? entry_lx(p) - S : 0, fy = ok ? entry_ly(p) - S : 0;
    const int xo = ok ? entry_xoff(p) : 0, yo = ok ? entry_yoff(p) : 0;
    const cplx<R>* k = tab + ((size_t)(z * os + yo) * os + xo) * S * S + ac;
    const R* g = lds + 2 * (fy * Wd + fx + ac);
    R sr = R(0), si = R(0);
#pragma unroll 4
    for (int b = 0; b < S; ++b) {
      const cplx<R> c = k[b * S];
      const R gr = g[2 * b * Wd], gi = g[2 * b * Wd + 1];
      sr += gr * c.x - gi * c.y;
      si += gr * c.y + gi * c.x;
    }
    if (!col_ok) {
      sr = R(0);
      si = R(0);
    }
    sr = group_sum<L>(sr);
    si = group_sum<L>(si);
    if (ok && a == L - 1) vis_out[row] = mk<R>(sr, si);
  }
}

template <typename R>
__global__ void k_zero_rows(const uint32_t* __restrict__ rows, int stride, int64_t n, cplx<R>* __restrict__ vis) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) vis[rows[i * stride]] = mk<R>(R(0), R(0));
}

// Zero n doubles (a kernel rather than a memset node, so that captured graphs replay it).
__global__ void k_zero_f64(double* __restrict__ p, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Batched forms of the small per-plan steps around the batched gridder and degridder.
// Sums of weights of every plan set to 0 (block d: plan d).
__global__ void k_batch_zero_sumwt(const BatchDesc* __restrict__ descs) {
  const BatchDesc& d = descs[blockIdx.x];
  for (int i = threadIdx.x; i < d.nsw; i += blockDim.x) d.sumwt[i] = 0.0;
}

// Split tiles of every plan summed in slot order, as k_reduce_tiles: block (d * G + g, y) takes
// split tiles g, g + G, ... of plan d, pixels y * 256 + threadIdx.x.
template <typename R>
__global__ void k_batch_reduce_tiles(const BatchDesc* __restrict__ descs, int G) {
  const BatchDesc& d = descs[blockIdx.x / G];
  const int gi = blockIdx.x % G;
  const int T = d.T;
  const int i = blockIdx.y * blockDim.x + threadIdx.x;
  if (i >= T * T || !d.multi) return;
  const uint32_t count = *d.nmulti;
  const size_t tt = (size_t)T * T;
  const cplx<R>* partial = (const cplx<R>*)d.partial;
  cplx<R>* grids = (cplx<R>*)d.grids;
  for (uint32_t mi = gi; mi < count; mi += G) {
    const int4 m = d.multi[mi];
    const TilePos tp = tile_pos(m.x, d.ntx);
    const cplx<R>* src = partial + (size_t)m.y * tt + i;
    R re = R(0), im = R(0);
    for (int k = 0; k < m.z; ++k) {
      const cplx<R> a = src[(size_t)k * tt];
      re += a.x;
      im += a.y;
    }
    grids[(size_t)tp.plane * d.N * d.N + (size_t)(tp.ty * T + i / T) * d.N + (size_t)tp.tx * T + (i % T)] =
        mk<R>(re, im);
  }
}

// Rows in no plane of every plan predicted as 0 (block d: plan d).
template <typename R>
__global__ void k_batch_zero_rows(const BatchDesc* __restrict__ descs) {
  const BatchDesc& d = descs[blockIdx.x];
  cplx<R>* vis = (cplx<R>*)d.vis_out;
  for (int64_t i = threadIdx.x; i < d.norphans; i += blockDim.x) vis[d.orphans[(int64_t)d.orphan_words * i]] = mk<R>(R(0), R(0));
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

static size_t lds_limit() { return 160 * 1024; }

// Partial-tile storage for split tiles, grown on demand (one buffer per plan, sized for the
// precision of the call).
template <typename R>
static int partial_buffer(const arlmi_plan* cp, cplx<R>** out) {
  arlmi_plan* p = const_cast<arlmi_plan*>(cp);
  // fixed-point plans keep integer partial tiles (two 64-bit words per value)
  const int64_t bytes = (p->flags & ARLMI_FLAG_DETERMINISTIC) ? (int64_t)sizeof(longlong2) : (int64_t)sizeof(cplx<R>);
  const int64_t need = std::max<int64_t>(1, p->nslots) * p->g.tile * p->g.tile * bytes;
  if (need > p->cap_partial_bytes || !p->d_partial) {
    if (p->d_partial) ARLMI_HIP(hipFree(p->d_partial));
    p->d_partial = nullptr;
    ARLMI_HIP(hipMalloc(&p->d_partial, (size_t)need));
    p->cap_partial_bytes = need;
  }
  *out = reinterpret_cast<cplx<R>*>(p->d_partial);
  return ARLMI_OK;
}

template <typename R, int S>
static int launch_grid_sep(const arlmi_plan* p, const void* vis, const void* wt, int dopsf, const void* ktab,
                           void* grids, cplx<R>* partial, double* sumwt, int accum, const DetArgs* det,
                           hipStream_t s) {
  const DevGeom& g = p->g;
  const size_t sh = sizeof(double) * 2 * ((size_t)g.tile * lds_stride(g.tile) + 64) + sizeof(R) * (size_t)g.os * S;
  if (sh > lds_limit()) {
    set_error("gridding tile needs %zu B of LDS", sh);
    return ARLMI_E_UNSUPPORTED;
  }
  if (p->ngitems == 0) return ARLMI_OK;
  if constexpr (S * S <= 64) {
    if (!det && p->origin_sorted) {
      ARLMI_LAUNCH("k_grid_sep", (k_grid_sep_merge<R, S>), dim3((unsigned)p->ngitems), dim3(GRID_THREADS), sh, s,
                   p->d_gitems, p->d_entries, (const cplx<R>*)vis, (const R*)wt, dopsf, (const R*)ktab, g.os,
                   (cplx<R>*)grids, partial, sumwt, g.tile, g.ntx, g.n, accum);
      ARLMI_LAUNCH_CHECK();
      return ARLMI_OK;
    }
  }
  if (det)
    ARLMI_LAUNCH("k_grid_sep", (k_grid_sep<R, S, true>), dim3((unsigned)p->ngitems), dim3(GRID_THREADS), sh, s,
                 p->d_gitems, p->d_entries, (const cplx<R>*)vis, (const R*)wt, dopsf, (const R*)ktab, g.os,
                 (cplx<R>*)grids, partial, sumwt, g.tile, g.ntx, g.n, accum, *det);
  else
    ARLMI_LAUNCH("k_grid_sep", (k_grid_sep<R, S, false>), dim3((unsigned)p->ngitems), dim3(GRID_THREADS), sh, s,
                 p->d_gitems, p->d_entries, (const cplx<R>*)vis, (const R*)wt, dopsf, (const R*)ktab, g.os,
                 (cplx<R>*)grids, partial, sumwt, g.tile, g.ntx, g.n, accum, DetArgs{});
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

// Row order of the grids a plan's degridder reads (what arlmi_image_to_grids leaves for it).
template <typename R>
static int degrid_perm(const arlmi_plan* p) {
  return (p->flags & ARLMI_FLAG_PRIVATE_GRIDS) ? grid_perm_n1(p->g.n, sizeof(R) == 4 ? 32 : 64) : 0;
}

template <typename R, int S>
static int launch_degrid_sep(const arlmi_plan* p, const void* grids, const void* ktab, void* vis_out, hipStream_t s) {
  const DevGeom& g = p->g;
  const int Wd = g.tile + S;
  const size_t sh = sizeof(R) * (2 * (size_t)Wd * Wd + (size_t)g.os * S);
  if (sh > lds_limit()) {
    set_error("degridding tile needs %zu B of LDS", sh);
    return ARLMI_E_UNSUPPORTED;
  }
  if (p->nditems == 0) return ARLMI_OK;
  ARLMI_LAUNCH("k_degrid_sep", (k_degrid_sep<R, S>), dim3((unsigned)p->nditems), dim3(degrid_threads<R>()), sh, s,
               p->d_ditems,
               p->d_entries, (const cplx<R>*)grids, (const R*)ktab, g.os, (cplx<R>*)vis_out, g.tile, g.ntx, g.n,
               degrid_perm<R>(p));
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

template <typename R>
static int grid_launch(const arlmi_plan* p, const void* vis, const void* wt, int dopsf, const arlmi_kernel* k,
                       void* grids, cplx<R>* partial, double* sumwt, int accum, const DetArgs* det,
                       hipStream_t s) {
  const DevGeom& g = p->g;
  if (k->kind == ARLMI_KERNEL_SEPARABLE) {
    switch (g.support) {
      case 2: return launch_grid_sep<R, 2>(p, vis, wt, dopsf, k->d_table, grids, partial, sumwt, accum, det, s);
      case 4: return launch_grid_sep<R, 4>(p, vis, wt, dopsf, k->d_table, grids, partial, sumwt, accum, det, s);
      case 6: return launch_grid_sep<R, 6>(p, vis, wt, dopsf, k->d_table, grids, partial, sumwt, accum, det, s);
      case 8: return launch_grid_sep<R, 8>(p, vis, wt, dopsf, k->d_table, grids, partial, sumwt, accum, det, s);
      case 10: return launch_grid_sep<R, 10>(p, vis, wt, dopsf, k->d_table, grids, partial, sumwt, accum, det, s);
      case 12: return launch_grid_sep<R, 12>(p, vis, wt, dopsf, k->d_table, grids, partial, sumwt, accum, det, s);
      case 16: return launch_grid_sep<R, 16>(p, vis, wt, dopsf, k->d_table, grids, partial, sumwt, accum, det, s);
      default: break;
    }
    set_error("separable kernel support %d not instantiated; pass the full table", g.support);
    return ARLMI_E_UNSUPPORTED;
  }
  const size_t sh = sizeof(double) * 2 * ((size_t)g.tile * lds_stride(g.tile) + 64);
  if (sh > lds_limit()) {
    set_error("gridding tile (T=%d, S=%d) needs %zu B of LDS; use a smaller tile", g.tile, g.support, sh);
    return ARLMI_E_UNSUPPORTED;
  }
  if (p->ngitems == 0) return ARLMI_OK;
  const int taps = g.support * g.support;
  const uint16_t* ez = g.nz > 1 ? p->d_entry_z : nullptr;
  const cplx<R>* tab = (const cplx<R>*)k->d_table;
#define ARLMI_GRID_TAB(MT)                                                                                        \
  do {                                                                                                            \
    if (det)                                                                                                      \
      ARLMI_LAUNCH("k_grid_tab", (k_grid_tab<R, MT, true>), dim3((unsigned)p->ngitems), dim3(GRID_THREADS), sh, s, \
                   p->d_gitems, p->d_entries, ez, (const cplx<R>*)vis, (const R*)wt, dopsf, tab, g.os, g.support, \
                   (cplx<R>*)grids, partial, sumwt, g.tile, g.ntx, g.n, accum, *det);                             \
    else                                                                                                          \
      ARLMI_LAUNCH("k_grid_tab", (k_grid_tab<R, MT, false>), dim3((unsigned)p->ngitems), dim3(GRID_THREADS), sh,  \
                   s, p->d_gitems, p->d_entries, ez, (const cplx<R>*)vis, (const R*)wt, dopsf, tab, g.os,        \
                   g.support, (cplx<R>*)grids, partial, sumwt, g.tile, g.ntx, g.n, accum, DetArgs{});             \
  } while (0)
  // (8 taps per lane, for supports 18-22, measured slower than the grouped loop: 25.3 against 19.1
  // ms for the S = 22 w-stack + AW workload)
  if (taps <= 64)
    ARLMI_GRID_TAB(1);
  else if (taps <= 256)
    ARLMI_GRID_TAB(4);
  else
    ARLMI_GRID_TAB(0);
#undef ARLMI_GRID_TAB
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

// Constants of a fixed-point call (k_det_bounds over the rows and kernel taps, k_det_scale); the
// integer sums of weights zeroed.
template <typename R>
static int det_prepare(const arlmi_plan* p, const void* vis, const void* wt, int dopsf, const arlmi_kernel* k,
                       int64_t nsw, DetArgs* da, hipStream_t s) {
  const DevGeom& g = p->g;
  const int complex_tab = k->kind == ARLMI_KERNEL_TABLE;
  const int64_t nk = complex_tab ? (int64_t)k->nz * g.os * g.os * g.support * g.support : (int64_t)g.os * g.support;
  double* part = p->d_det + 8 + nsw;
  ARLMI_LAUNCH("k_det_bounds", k_det_bounds<R>, dim3(DET_BLOCKS), dim3(256), 0, s, (const cplx<R>*)vis, (const R*)wt,
               p->nrows, dopsf, (const R*)k->d_table, nk, complex_tab, part);
  ARLMI_LAUNCH_CHECK();
  ARLMI_LAUNCH("k_det_scale", k_det_scale, dim3(1), dim3(DET_BLOCKS), 0, s, p->d_det, nsw, complex_tab);
  ARLMI_LAUNCH_CHECK();
  da->det = p->d_det;
  da->tcnt = p->d_cnt;
  return ARLMI_OK;
}

template <typename R>
static int grid_impl(const arlmi_plan* p, const void* vis, const void* wt, int dopsf, const arlmi_kernel* k,
                     void* grids, double* sumwt, int accum, hipStream_t s) {
  const DevGeom& g = p->g;
  const int64_t nsw = (int64_t)g.nplanes * g.nchan;
  const bool det = (p->flags & ARLMI_FLAG_DETERMINISTIC) != 0;
  if (!accum) {
    ARLMI_LAUNCH("k_zero_f64", k_zero_f64, dim3((unsigned)((nsw + 255) / 256)), dim3(256), 0, s, sumwt, nsw);
    ARLMI_LAUNCH_CHECK();
  }
  int rc = ARLMI_OK;
  DetArgs da{};
  if (det) {
    if (!p->d_det) {
      set_error("deterministic plan without its fixed-point state");
      return ARLMI_E_ARG;
    }
    rc = det_prepare<R>(p, vis, wt, dopsf, k, nsw, &da, s);
    if (rc) return rc;
  }
  cplx<R>* partial = nullptr;
  rc = partial_buffer<R>(p, &partial);
  if (rc) return rc;
  rc = grid_launch<R>(p, vis, wt, dopsf, k, grids, partial, sumwt, accum, det ? &da : nullptr, s);
  if (rc) return rc;
  if (p->nslots > 0) {  // every split tile has at least two slots
    const int64_t gx = std::min<int64_t>(std::max<int64_t>(1, p->nslots / 2), 4096);
    if (det)
      ARLMI_LAUNCH("k_reduce_tiles", k_reduce_tiles_det<R>, dim3((unsigned)gx, (g.tile * g.tile + 255) / 256),
                   dim3(256), 0, s, p->d_multi, p->d_misc + 4, reinterpret_cast<const longlong2*>(partial),
                   (cplx<R>*)grids, g.tile, g.ntx, g.n, accum, da);
    else
      ARLMI_LAUNCH("k_reduce_tiles", k_reduce_tiles<R>, dim3((unsigned)gx, (g.tile * g.tile + 255) / 256),
                   dim3(256), 0, s, p->d_multi, p->d_misc + 4, partial, (cplx<R>*)grids, g.tile, g.ntx, g.n, accum);
    ARLMI_LAUNCH_CHECK();
  }
  if (det) {
    ARLMI_LAUNCH("k_det_sumwt", k_det_sumwt, dim3((unsigned)((nsw + 255) / 256)), dim3(256), 0, s,
                 (const double*)p->d_det, sumwt, nsw);
    ARLMI_LAUNCH_CHECK();
  }
  return ARLMI_OK;
}

template <typename R>
static int degrid_impl(const arlmi_plan* p, const void* grids, const arlmi_kernel* k, void* vis_out, hipStream_t s) {
  const DevGeom& g = p->g;
  if (p->norphans > 0) {
    ARLMI_LAUNCH("k_zero_rows", k_zero_rows<R>, dim3((unsigned)((p->norphans + 255) / 256)), dim3(256), 0, s,
                 reinterpret_cast<const uint32_t*>(p->d_recs) + (int64_t)p->rec_words * p->orphan_start, p->rec_words,
                 p->norphans,
                 (cplx<R>*)vis_out);
    ARLMI_LAUNCH_CHECK();
  }
  if (k->kind == ARLMI_KERNEL_SEPARABLE) {
    switch (g.support) {
      case 2: return launch_degrid_sep<R, 2>(p, grids, k->d_table, vis_out, s);
      case 4: return launch_degrid_sep<R, 4>(p, grids, k->d_table, vis_out, s);
      case 6: return launch_degrid_sep<R, 6>(p, grids, k->d_table, vis_out, s);
      case 8: return launch_degrid_sep<R, 8>(p, grids, k->d_table, vis_out, s);
      case 10: return launch_degrid_sep<R, 10>(p, grids, k->d_table, vis_out, s);
      case 12: return launch_degrid_sep<R, 12>(p, grids, k->d_table, vis_out, s);
      case 16: return launch_degrid_sep<R, 16>(p, grids, k->d_table, vis_out, s);
      default: break;
    }
    set_error("separable kernel support %d not instantiated; pass the full table", g.support);
    return ARLMI_E_UNSUPPORTED;
  }
  const int Wd = g.tile + g.support;
  const size_t sh = sizeof(R) * 2 * (size_t)Wd * Wd;
  if (sh > lds_limit()) {
    set_error("degridding tile needs %zu B of LDS; use a smaller tile", sh);
    return ARLMI_E_UNSUPPORTED;
  }
  if (p->nditems == 0) return ARLMI_OK;
  const uint16_t* ez = g.nz > 1 ? p->d_entry_z : nullptr;
  const cplx<R>* tab = (const cplx<R>*)k->d_table;
#define ARLMI_DEGRID_TAB(L)                                                                                       \
  ARLMI_LAUNCH("k_degrid_tab", (k_degrid_tab<R, L>), dim3((unsigned)p->nditems), dim3(DEGRID_TAB_THREADS), sh, s, \
               p->d_ditems, p->d_entries, ez, (const cplx<R>*)grids, tab, g.os, g.support, (cplx<R>*)vis_out,    \
               g.tile, g.ntx, g.n, degrid_perm<R>(p))
  const int S = g.support;
  if (S <= 2)
    ARLMI_DEGRID_TAB(2);
  else if (S <= 4)
    ARLMI_DEGRID_TAB(4);
  else if (S <= 8)
    ARLMI_DEGRID_TAB(8);
  else if (S <= 16)
    ARLMI_DEGRID_TAB(16);
  else if (S <= 32)
    ARLMI_DEGRID_TAB(32);
  else
    ARLMI_DEGRID_TAB(64);
#undef ARLMI_DEGRID_TAB
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

static int check_kernel(const arlmi_plan* p, const arlmi_kernel* k) {
  if (!p || !k || !k->d_table) {
    set_error("NULL plan or kernel");
    return ARLMI_E_ARG;
  }
  if (k->support != p->g.support || k->oversampling != p->g.os) {
    set_error("kernel (S=%d, os=%d) does not match the plan (S=%d, os=%d)", k->support, k->oversampling,
              p->g.support, p->g.os);
    return ARLMI_E_ARG;
  }
  if (k->kind == ARLMI_KERNEL_TABLE && k->nz != p->g.nz) {
    set_error("kernel nz %d does not match the plan nz %d", k->nz, p->g.nz);
    return ARLMI_E_ARG;
  }
  if (k->kind == ARLMI_KERNEL_SEPARABLE && p->g.nz != 1) {
    set_error("a separable kernel has one w plane; the plan has nz=%d", p->g.nz);
    return ARLMI_E_ARG;
  }
  return ARLMI_OK;
}

// The number of runs k_grid_sep_merge adds (one pair of LDS adds each), counted over the same
// work partition: per work item and wavefront share, per 64 entries, the entries whose origin
// differs from the next one's, the 32nd and 64th and the last.  A measurement aid.
__global__ __launch_bounds__(GRID_THREADS) void k_count_runs(const int4* __restrict__ items,
                                                             const uint2* __restrict__ entries,
                                                             unsigned long long* __restrict__ runs) {
  __shared__ unsigned long long wsum[GRID_THREADS / 64];
  const int4 it = items[blockIdx.x];
  const int lane = __lane_id(), wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int per = (it.z + nwave - 1) / nwave;
  const int e0 = min(it.z, wave * per), e1 = min(it.z, e0 + per);
  unsigned long long n = 0;
  for (int base = e0; base < e1; base += 64) {
    const int e = base + lane;
    const uint32_t org = e < e1 ? (entries[it.y + e].y & 0xffffu) : 0u;  // (lx, ly)
    const int cnt = min(64, e1 - base);
    const uint32_t nxt = (uint32_t)__shfl_down((int)org, 1);
    const bool last = lane == cnt - 1 || (lane & 31) == 31 || (lane < cnt - 1 && nxt != org);
    n += (unsigned long long)__popcll(__ballot(last && lane < cnt));
  }
  if (lane == 0) wsum[wave] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < nwave; ++w) t += wsum[w];
    atomicAdd(runs, t);
  }
}

}  // namespace arlmi

using namespace arlmi;

extern "C" {

int arlmi_plan_grid_runs(const arlmi_plan* p, int32_t support, int64_t* runs_out, void* stream) {
  if (!p || !runs_out) {
    set_error("arlmi_plan_grid_runs: NULL argument");
    return ARLMI_E_ARG;
  }
  // the separable gridder adds runs on origin-grouped plans (at most 64 taps), else every entry
  if (!p->origin_sorted || support * support > 64 || (p->flags & ARLMI_FLAG_DETERMINISTIC) || p->ngitems == 0) {
    *runs_out = p->nentries;
    return ARLMI_OK;
  }
  hipStream_t s = as_stream(stream);
  unsigned long long* d = nullptr;
  ARLMI_HIP(hipMallocAsync(reinterpret_cast<void**>(&d), sizeof(unsigned long long), s));
  ARLMI_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), s));
  ARLMI_LAUNCH("k_count_runs", k_count_runs, dim3((unsigned)p->ngitems), dim3(GRID_THREADS), 0, s, p->d_gitems,
               p->d_entries, d);
  ARLMI_LAUNCH_CHECK();
  unsigned long long h = 0;
  ARLMI_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s));
  ARLMI_HIP(hipFreeAsync(d, s));
  ARLMI_HIP(hipStreamSynchronize(s));
  *runs_out = (int64_t)h;
  return ARLMI_OK;
}

int arlmi_grid(const arlmi_plan* p, const void* d_vis, const void* d_weight, int dopsf, const arlmi_kernel* k,
               void* d_grids, double* d_sumwt, int precision, void* stream) {
  int rc = check_kernel(p, k);
  if (rc) return rc;
  if (!d_grids || !d_sumwt || (p->nentries > 0 && (!d_weight || (!dopsf && !d_vis)))) {
    set_error("arlmi_grid: NULL buffer");
    return ARLMI_E_ARG;
  }
  hipStream_t s = as_stream(stream);
  if (precision == 32) return grid_impl<float>(p, d_vis, d_weight, dopsf, k, d_grids, d_sumwt, 0, s);
  if (precision == 64) return grid_impl<double>(p, d_vis, d_weight, dopsf, k, d_grids, d_sumwt, 0, s);
  set_error("precision must be 32 or 64");
  return ARLMI_E_ARG;
}

}  // extern "C"

namespace arlmi {
namespace {

// One launch group: the plans of a batch that share a kernel support.
struct BatchGroup {
  int support = 0;
  int first = 0, count = 0;  // descriptors [first, first + count)
  int64_t gblocks = 0, dblocks = 0;
  size_t gsh = 0, dsh = 0;   // dynamic LDS of the gridder / degridder launches
};

}  // namespace
}  // namespace arlmi

struct arlmi_batch {
  int precision = 32;
  int n = 0;
  arlmi::BatchDesc* d_descs = nullptr;
  std::vector<arlmi::BatchDesc> h_descs;
  std::vector<arlmi::BatchGroup> groups;
  int reduce_g = 1, max_tile = 0;
  bool any_split = false, any_orphans = false;
};

namespace arlmi {
namespace {

template <typename R>
int batch_build(arlmi_batch* b, const arlmi_batch_item* in, int n) {
  // group by support, keeping each group's descriptors together
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int c) { return in[a].plan->g.support < in[c].plan->g.support; });
  for (int oi = 0; oi < n; ++oi) {
    const arlmi_batch_item& it = in[order[oi]];
    const arlmi_plan* p = it.plan;
    const DevGeom& g = p->g;
    if (!b->groups.empty() && b->groups.back().support == g.support) {
      // same group
    } else {
      BatchGroup grp;
      grp.support = g.support;
      grp.first = oi;
      b->groups.push_back(grp);
    }
    BatchGroup& grp = b->groups.back();
    cplx<R>* partial = nullptr;
    if (int rc = partial_buffer<R>(p, &partial)) return rc;
    BatchDesc d{};
    d.gitems = p->d_gitems;
    d.ditems = p->d_ditems;
    d.entries = p->d_entries;
    d.runs = p->origin_sorted ? 1 : 0;
    d.vis = it.d_vis;
    d.wt = it.d_weight;
    d.ktab = it.kernel.d_table;
    d.grids = it.d_grids;
    d.partial = partial;
    d.sumwt = it.d_sumwt;
    d.vis_out = it.d_vis_out;
    d.multi = p->nslots > 0 ? p->d_multi : nullptr;
    d.nmulti = p->d_misc + 4;
    d.orphans = p->norphans > 0 ? reinterpret_cast<const uint32_t*>(p->d_recs) + (int64_t)p->rec_words * p->orphan_start
                                : nullptr;
    d.orphan_words = p->rec_words;
    d.norphans = p->norphans;
    d.os = g.os;
    d.T = g.tile;
    d.ntx = g.ntx;
    d.N = g.n;
    d.nsw = g.nplanes * g.nchan;
    d.gbase = (int)grp.gblocks;
    d.ngitems = (int)p->ngitems;
    d.dbase = (int)grp.dblocks;
    d.nditems = (int)p->nditems;
    d.perm_n1 = (p->flags & ARLMI_FLAG_PRIVATE_GRIDS) ? grid_perm_n1(g.n, sizeof(R) == 4 ? 32 : 64) : 0;
    grp.gblocks += p->ngitems;
    grp.dblocks += p->nditems;
    if (grp.gblocks > 0x7fffffff || grp.dblocks > 0x7fffffff) {
      set_error("arlmi_batch_create: too many work items");
      return ARLMI_E_UNSUPPORTED;
    }
    const int S = g.support;
    grp.gsh = std::max(grp.gsh, sizeof(double) * 2 * ((size_t)g.tile * lds_stride(g.tile) + 64) + sizeof(R) * (size_t)g.os * S);
    grp.dsh = std::max(grp.dsh, sizeof(R) * (2 * (size_t)(g.tile + S) * (g.tile + S) + (size_t)g.os * S));
    grp.count++;
    b->any_split |= p->nslots > 0;
    b->any_orphans |= p->norphans > 0;
    b->max_tile = std::max(b->max_tile, g.tile);
    b->h_descs.push_back(d);
  }
  for (const BatchGroup& grp : b->groups) {
    if (grp.gsh > lds_limit() || grp.dsh > lds_limit()) {
      set_error("arlmi_batch_create: a tile of support %d exceeds LDS", grp.support);
      return ARLMI_E_UNSUPPORTED;
    }
  }
  ARLMI_HIP(hipMalloc(&b->d_descs, sizeof(BatchDesc) * n));
  ARLMI_HIP(hipMemcpy(b->d_descs, b->h_descs.data(), sizeof(BatchDesc) * n, hipMemcpyHostToDevice));
  b->reduce_g = 64;
  return ARLMI_OK;
}

template <typename R, int S>
int batch_grid_group(const arlmi_batch* b, const BatchGroup& grp, int dopsf, hipStream_t s) {
  if (grp.gblocks == 0) return ARLMI_OK;
  ARLMI_LAUNCH("k_grid_sep", (k_grid_sep_batch<R, S>), dim3((unsigned)grp.gblocks), dim3(GRID_THREADS), grp.gsh, s,
               b->d_descs + grp.first, grp.count, dopsf);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

template <typename R, int S>
int batch_degrid_group(const arlmi_batch* b, const BatchGroup& grp, hipStream_t s) {
  if (grp.dblocks == 0) return ARLMI_OK;
  ARLMI_LAUNCH("k_degrid_sep", (k_degrid_sep_batch<R, S>), dim3((unsigned)grp.dblocks), dim3(degrid_threads<R>()),
               grp.dsh, s, b->d_descs + grp.first, grp.count);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

template <typename R>
int batch_grid(const arlmi_batch* b, int dopsf, hipStream_t s) {
  ARLMI_LAUNCH("k_batch_zero_sumwt", k_batch_zero_sumwt, dim3((unsigned)b->n), dim3(256), 0, s, b->d_descs);
  ARLMI_LAUNCH_CHECK();
  for (const BatchGroup& grp : b->groups) {
    int rc = ARLMI_OK;
    switch (grp.support) {
      case 2: rc = batch_grid_group<R, 2>(b, grp, dopsf, s); break;
      case 4: rc = batch_grid_group<R, 4>(b, grp, dopsf, s); break;
      case 6: rc = batch_grid_group<R, 6>(b, grp, dopsf, s); break;
      case 8: rc = batch_grid_group<R, 8>(b, grp, dopsf, s); break;
      case 10: rc = batch_grid_group<R, 10>(b, grp, dopsf, s); break;
      case 12: rc = batch_grid_group<R, 12>(b, grp, dopsf, s); break;
      case 16: rc = batch_grid_group<R, 16>(b, grp, dopsf, s); break;
      default: set_error("batched gridding: support %d not instantiated", grp.support); return ARLMI_E_UNSUPPORTED;
    }
    if (rc) return rc;
  }
  if (b->any_split) {
    const int T = b->max_tile;
    ARLMI_LAUNCH("k_reduce_tiles", k_batch_reduce_tiles<R>, dim3((unsigned)(b->n * b->reduce_g), (T * T + 255) / 256),
                 dim3(256), 0, s, b->d_descs, b->reduce_g);
    ARLMI_LAUNCH_CHECK();
  }
  return ARLMI_OK;
}

template <typename R>
int batch_degrid(const arlmi_batch* b, hipStream_t s) {
  if (b->any_orphans) {
    ARLMI_LAUNCH("k_zero_rows", k_batch_zero_rows<R>, dim3((unsigned)b->n), dim3(256), 0, s, b->d_descs);
    ARLMI_LAUNCH_CHECK();
  }
  for (const BatchGroup& grp : b->groups) {
    int rc = ARLMI_OK;
    switch (grp.support) {
      case 2: rc = batch_degrid_group<R, 2>(b, grp, s); break;
      case 4: rc = batch_degrid_group<R, 4>(b, grp, s); break;
      case 6: rc = batch_degrid_group<R, 6>(b, grp, s); break;
      case 8: rc = batch_degrid_group<R, 8>(b, grp, s); break;
      case 10: rc = batch_degrid_group<R, 10>(b, grp, s); break;
      case 12: rc = batch_degrid_group<R, 12>(b, grp, s); break;
      case 16: rc = batch_degrid_group<R, 16>(b, grp, s); break;
      default: set_error("batched degridding: support %d not instantiated", grp.support); return ARLMI_E_UNSUPPORTED;
    }
    if (rc) return rc;
  }
  return ARLMI_OK;
}

}  // namespace
}  // namespace arlmi

extern "C" {

int arlmi_batch_create(const arlmi_batch_item* items, int32_t n, int precision, arlmi_batch** out) {
  if (!items || n < 1 || !out || (precision != 32 && precision != 64)) {
    set_error("arlmi_batch_create: bad argument");
    return ARLMI_E_ARG;
  }
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    const arlmi_batch_item& it = items[i];
    if (!it.plan || !it.d_grids || !it.d_sumwt) {
      set_error("arlmi_batch_create: item %d has a NULL plan or buffer", i);
      return ARLMI_E_ARG;
    }
    if (int rc = check_kernel(it.plan, &it.kernel)) return rc;
    if (it.kernel.kind != ARLMI_KERNEL_SEPARABLE) {
      set_error("arlmi_batch_create: item %d: only separable kernels are batched", i);
      return ARLMI_E_UNSUPPORTED;
    }
    if (it.plan->flags & ARLMI_FLAG_OCCUPIED_TILES) {
      set_error("arlmi_batch_create: item %d: plans that leave untouched tiles unwritten are not batched", i);
      return ARLMI_E_UNSUPPORTED;
    }
    if (it.plan->flags & ARLMI_FLAG_DETERMINISTIC) {
      set_error("arlmi_batch_create: item %d: fixed-point (deterministic) plans are not batched", i);
      return ARLMI_E_UNSUPPORTED;
    }
  }
  arlmi_batch* b = new arlmi_batch();
  b->precision = precision;
  b->n = n;
  const int rc = precision == 32 ? batch_build<float>(b, items, n) : batch_build<double>(b, items, n);
  if (rc) {
    if (b->d_descs) (void)hipFree(b->d_descs);
    delete b;
    return rc;
  }
  *out = b;
  return ARLMI_OK;
}

int arlmi_batch_grid(const arlmi_batch* b, int dopsf, void* stream) {
  if (!b) {
    set_error("arlmi_batch_grid: NULL batch");
    return ARLMI_E_ARG;
  }
  hipStream_t s = as_stream(stream);
  return b->precision == 32 ? batch_grid<float>(b, dopsf, s) : batch_grid<double>(b, dopsf, s);
}

int arlmi_batch_degrid(const arlmi_batch* b, void* stream) {
  if (!b) {
    set_error("arlmi_batch_degrid: NULL batch");
    return ARLMI_E_ARG;
  }
  hipStream_t s = as_stream(stream);
  return b->precision == 32 ? batch_degrid<float>(b, s) : batch_degrid<double>(b, s);
}

void arlmi_batch_destroy(arlmi_batch* b) {
  if (!b) return;
  if (b->d_descs) (void)hipFree(b->d_descs);
  delete b;
}

int arlmi_grid_add(const arlmi_plan* p, const void* d_vis, const void* d_weight, int dopsf, const arlmi_kernel* k,
                   void* d_grids, double* d_sumwt, int precision, void* stream) {
  int rc = check_kernel(p, k);
  if (rc) return rc;
  if (!d_grids || !d_sumwt || (p->nentries > 0 && (!d_weight || (!dopsf && !d_vis)))) {
    set_error("arlmi_grid_add: NULL buffer");
    return ARLMI_E_ARG;
  }
  if (p->flags & ARLMI_FLAG_SPARSE_GRIDS) {
    set_error("arlmi_grid_add: the plan has sparse grids, whose untouched tiles are never written; "
              "accumulating chunks needs dense grids");
    return ARLMI_E_ARG;
  }
  hipStream_t s = as_stream(stream);
  if (precision == 32) return grid_impl<float>(p, d_vis, d_weight, dopsf, k, d_grids, d_sumwt, 1, s);
  if (precision == 64) return grid_impl<double>(p, d_vis, d_weight, dopsf, k, d_grids, d_sumwt, 1, s);
  set_error("precision must be 32 or 64");
  return ARLMI_E_ARG;
}

int arlmi_degrid(const arlmi_plan* p, const void* d_grids, const arlmi_kernel* k, void* d_vis_out, int precision,
                 void* stream) {
  int rc = check_kernel(p, k);
  if (rc) return rc;
  if (!d_grids || (!d_vis_out && p->nrows > 0)) {
    set_error("arlmi_degrid: NULL buffer");
    return ARLMI_E_ARG;
  }
  hipStream_t s = as_stream(stream);
  if (precision == 32) return degrid_impl<float>(p, d_grids, k, d_vis_out, s);
  if (precision == 64) return degrid_impl<double>(p, d_grids, k, d_vis_out, s);
  set_error("precision must be 32 or 64");
  return ARLMI_E_ARG;
}

}  // extern "C"
