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
__device__ inline void lds_add(R* p, R v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

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

// Writes the T x T accumulator tile: to the grid, or (slot > 0) to partial tile slot-1 of a
// split tile.
template <typename R>
__device__ inline void store_tile(const double* are, const double* aim, int WS, int T, int N, const TilePos& tp,
                                  int slot, cplx<R>* __restrict__ grids, cplx<R>* __restrict__ partial) {
  if (slot == 0) {
    cplx<R>* g = grids + (size_t)tp.plane * N * N + (size_t)(tp.ty * T) * N + (size_t)tp.tx * T;
    for (int i = threadIdx.x; i < T * T; i += blockDim.x) {
      const int y = i / T, x = i - (i / T) * T;
      const int o = y * WS + x;
      g[(size_t)y * N + x] = mk<R>((R)are[o], (R)aim[o]);
    }
  } else {
    cplx<R>* d = partial + (size_t)(slot - 1) * T * T;
    for (int i = threadIdx.x; i < T * T; i += blockDim.x) {
      const int y = i / T, x = i - (i / T) * T;
      const int o = y * WS + x;
      d[i] = mk<R>((R)are[o], (R)aim[o]);
    }
  }
}

__device__ inline void lds_add_f64(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Sums the partial tiles of every split tile, in slot order, into the grid.
// Grid: (split tiles, T*T/256 pixel blocks); each thread sums one pixel over the tile's slots
// in slot order (four loads in flight).
template <typename R>
__global__ void k_reduce_tiles(const int4* __restrict__ multi, const cplx<R>* __restrict__ partial,
                               cplx<R>* __restrict__ grids, int T, int ntx, int N) {
  const int4 m = multi[blockIdx.x];
  const int i = blockIdx.y * blockDim.x + threadIdx.x;
  if (i >= T * T) return;
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
  grids[(size_t)tp.plane * N * N + (size_t)(tp.ty * T + i / T) * N + (size_t)tp.tx * T + (i % T)] = mk<R>(re, im);
}

template <typename R, int S>
__global__ __launch_bounds__(256) void k_grid_sep(const int4* __restrict__ items, const uint2* __restrict__ entries,
                                                  const cplx<R>* __restrict__ vis, const R* __restrict__ wt,
                                                  int dopsf, const R* __restrict__ ktab, int os,
                                                  cplx<R>* __restrict__ grids, cplx<R>* __restrict__ partial,
                                                  double* __restrict__ sumwt, int T, int ntx, int N) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[16];
  constexpr int TAPS = S * S;
  constexpr int TPL = (TAPS + 63) / 64;  // taps per lane
  const int WS = lds_stride(T);
  double* are = reinterpret_cast<double*>(smem);
  double* aim = are + T * WS;
  R* lk = reinterpret_cast<R*>(aim + T * WS);
  const int4 it = items[blockIdx.x];
  for (int i = threadIdx.x; i < T * WS; i += blockDim.x) {
    are[i] = 0.0;
    aim[i] = 0.0;
  }
  for (int i = threadIdx.x; i < os * S; i += blockDim.x) lk[i] = ktab[i];
  __syncthreads();
  const int lane = __lane_id(), wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  int ta[TPL], tb[TPL];
#pragma unroll
  for (int g = 0; g < TPL; ++g) {
    const int t = lane + 64 * g;
    ta[g] = t < TAPS ? t / S : 0;
    tb[g] = t < TAPS ? t % S : 0;
  }
  double swt = 0.0;
  for (int base = wave * 64; base < it.z; base += nwave * 64) {
    const int e = base + lane;
    uint32_t pk = 0;
    R vr = R(0), vi = R(0);
    if (e < it.z) {
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
    const int cnt = min(64, it.z - base);
    for (int j = 0; j < cnt; ++j) {
      const uint32_t p = bcast(pk, j);
      const R jr = bcast(vr, j), ji = bcast(vi, j);
      const int cx = entry_lx(p) - S, cy = entry_ly(p) - S;  // footprint origin in tile coordinates
      const R* ky = lk + entry_yoff(p) * S;
      const R* kx = lk + entry_xoff(p) * S;
#pragma unroll
      for (int g = 0; g < TPL; ++g) {
        const int x = cx + tb[g], y = cy + ta[g];
        if ((TAPS % 64 == 0 || lane + 64 * g < TAPS) && (unsigned)x < (unsigned)T && (unsigned)y < (unsigned)T) {
          const R c = ky[ta[g]] * kx[tb[g]];
          const int o = y * WS + x;
          lds_add_f64(are + o, (double)(c * jr));
          lds_add_f64(aim + o, (double)(c * ji));
        }
      }
    }
  }
  __syncthreads();
  const TilePos tp = tile_pos(it.x, ntx);
  store_tile<R>(are, aim, WS, T, N, tp, it.w, grids, partial);
  block_add_double(swt, &sumwt[tp.plane], red);
}

// ---------------------------------------------------------------------------
// gridding, full complex kernel table (AW / box / any CF), runtime support
// ---------------------------------------------------------------------------

template <typename R>
__global__ __launch_bounds__(256) void k_grid_tab(const int4* __restrict__ items, const uint2* __restrict__ entries,
                                                  const uint16_t* __restrict__ entry_z,
                                                  const cplx<R>* __restrict__ vis, const R* __restrict__ wt,
                                                  int dopsf, const cplx<R>* __restrict__ tab, int os, int S,
                                                  cplx<R>* __restrict__ grids, cplx<R>* __restrict__ partial,
                                                  double* __restrict__ sumwt, int T, int ntx, int N) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[16];
  const int WS = lds_stride(T);
  double* are = reinterpret_cast<double*>(smem);
  double* aim = are + T * WS;
  const int4 it = items[blockIdx.x];
  for (int i = threadIdx.x; i < T * WS; i += blockDim.x) {
    are[i] = 0.0;
    aim[i] = 0.0;
  }
  __syncthreads();
  const int lane = __lane_id(), wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int taps = S * S;
  double swt = 0.0;
  for (int base = wave * 64; base < it.z; base += nwave * 64) {
    const int e = base + lane;
    uint32_t pk = 0, zz = 0;
    R vr = R(0), vi = R(0);
    if (e < it.z) {
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
      if (pk & ENTRY_START) swt += (double)w;
    }
    const int cnt = min(64, it.z - base);
    for (int j = 0; j < cnt; ++j) {
      const uint32_t p = bcast(pk, j);
      const int z = (int)bcast(zz, j);
      const R jr = bcast(vr, j), ji = bcast(vi, j);
      const int cx = entry_lx(p) - S, cy = entry_ly(p) - S;
      const cplx<R>* k = tab + ((size_t)(z * os + entry_yoff(p)) * os + entry_xoff(p)) * taps;
      for (int t = lane; t < taps; t += 64) {
        const int a = t / S, b = t - a * S;
        const int x = cx + b, y = cy + a;
        if ((unsigned)x < (unsigned)T && (unsigned)y < (unsigned)T) {
          const cplx<R> c = k[t];
          const int o = y * WS + x;
          // conj(c) * v
          lds_add_f64(are + o, (double)(c.x * jr + c.y * ji));
          lds_add_f64(aim + o, (double)(c.x * ji - c.y * jr));
        }
      }
    }
  }
  __syncthreads();
  const TilePos tp = tile_pos(it.x, ntx);
  store_tile<R>(are, aim, WS, T, N, tp, it.w, grids, partial);
  block_add_double(swt, &sumwt[tp.plane], red);
}

// ---------------------------------------------------------------------------
// degridding
// ---------------------------------------------------------------------------

template <typename R>
__device__ inline void load_halo(const cplx<R>* __restrict__ grids, R* lds, int Wd, int T, int N, const TilePos& tp) {
  const cplx<R>* g = grids + (size_t)tp.plane * N * N;
  const int y0 = tp.ty * T, x0 = tp.tx * T;
  for (int i = threadIdx.x; i < Wd * Wd; i += blockDim.x) {
    const int y = i / Wd, x = i - (i / Wd) * Wd;
    const int gy = y0 + y, gx = x0 + x;
    cplx<R> v = mk<R>(R(0), R(0));
    if (gy < N && gx < N) v = g[(size_t)gy * N + gx];
    lds[2 * i] = v.x;
    lds[2 * i + 1] = v.y;
  }
}

template <typename R, int S>
__global__ __launch_bounds__(256) void k_degrid_sep(const int4* __restrict__ items, const uint2* __restrict__ entries,
                                                    const cplx<R>* __restrict__ grids, const R* __restrict__ ktab,
                                                    int os, cplx<R>* __restrict__ vis_out, int T, int ntx, int N) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int Wd = T + S;
  R* lds = reinterpret_cast<R*>(smem);
  R* lk = lds + 2 * Wd * Wd;
  const int4 it = items[blockIdx.x];
  const TilePos tp = tile_pos(it.x, ntx);
  load_halo<R>(grids, lds, Wd, T, N, tp);
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

template <typename R>
__global__ __launch_bounds__(256) void k_degrid_tab(const int4* __restrict__ items, const uint2* __restrict__ entries,
                                                    const uint16_t* __restrict__ entry_z,
                                                    const cplx<R>* __restrict__ grids, const cplx<R>* __restrict__ tab,
                                                    int os, int S, cplx<R>* __restrict__ vis_out, int T, int ntx,
                                                    int N) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int Wd = T + S;
  R* lds = reinterpret_cast<R*>(smem);
  const int4 it = items[blockIdx.x];
  const TilePos tp = tile_pos(it.x, ntx);
  load_halo<R>(grids, lds, Wd, T, N, tp);
  __syncthreads();
  for (int e = threadIdx.x; e < it.z; e += blockDim.x) {
    const uint2 en = entries[it.y + e];
    const uint32_t p = en.y;
    const int fx = entry_lx(p) - S, fy = entry_ly(p) - S, xo = entry_xoff(p), yo = entry_yoff(p);
    const int z = entry_z ? (int)entry_z[it.y + e] : 0;
    const cplx<R>* k = tab + ((size_t)(z * os + yo) * os + xo) * S * S;
    const R* base = lds + 2 * (fy * Wd + fx);
    R ar = R(0), ai = R(0);
    for (int a = 0; a < S; ++a) {
      const R* row = base + 2 * a * Wd;
      for (int b = 0; b < S; ++b) {
        const cplx<R> c = k[a * S + b];
        const R gr = row[2 * b], gi = row[2 * b + 1];
        ar += gr * c.x - gi * c.y;
        ai += gr * c.y + gi * c.x;
      }
    }
    vis_out[en.x] = mk<R>(ar, ai);
  }
}

template <typename R>
__global__ void k_zero_rows(const uint32_t* __restrict__ rows, int stride, int64_t n, cplx<R>* __restrict__ vis) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) vis[rows[i * stride]] = mk<R>(R(0), R(0));
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
  const int64_t need = std::max<int64_t>(1, p->nslots) * p->g.tile * p->g.tile * (int64_t)sizeof(cplx<R>);
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
                           void* grids, cplx<R>* partial, double* sumwt, hipStream_t s) {
  const DevGeom& g = p->g;
  const size_t sh = sizeof(double) * 2 * (size_t)g.tile * lds_stride(g.tile) + sizeof(R) * (size_t)g.os * S;
  if (sh > lds_limit()) {
    set_error("gridding tile needs %zu B of LDS", sh);
    return ARLMI_E_UNSUPPORTED;
  }
  if (p->ngitems == 0) return ARLMI_OK;
  ARLMI_LAUNCH("k_grid_sep", (k_grid_sep<R, S>), dim3((unsigned)p->ngitems), dim3(256), sh, s, p->d_gitems,
               p->d_entries, (const cplx<R>*)vis, (const R*)wt, dopsf, (const R*)ktab, g.os, (cplx<R>*)grids, partial,
               sumwt, g.tile, g.ntx, g.n);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
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
  ARLMI_LAUNCH("k_degrid_sep", (k_degrid_sep<R, S>), dim3((unsigned)p->nditems), dim3(256), sh, s, p->d_ditems,
               p->d_entries, (const cplx<R>*)grids, (const R*)ktab, g.os, (cplx<R>*)vis_out, g.tile, g.ntx, g.n);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

template <typename R>
static int grid_launch(const arlmi_plan* p, const void* vis, const void* wt, int dopsf, const arlmi_kernel* k,
                       void* grids, cplx<R>* partial, double* sumwt, hipStream_t s) {
  const DevGeom& g = p->g;
  if (k->kind == ARLMI_KERNEL_SEPARABLE) {
    switch (g.support) {
      case 2: return launch_grid_sep<R, 2>(p, vis, wt, dopsf, k->d_table, grids, partial, sumwt, s);
      case 4: return launch_grid_sep<R, 4>(p, vis, wt, dopsf, k->d_table, grids, partial, sumwt, s);
      case 6: return launch_grid_sep<R, 6>(p, vis, wt, dopsf, k->d_table, grids, partial, sumwt, s);
      case 8: return launch_grid_sep<R, 8>(p, vis, wt, dopsf, k->d_table, grids, partial, sumwt, s);
      case 10: return launch_grid_sep<R, 10>(p, vis, wt, dopsf, k->d_table, grids, partial, sumwt, s);
      case 12: return launch_grid_sep<R, 12>(p, vis, wt, dopsf, k->d_table, grids, partial, sumwt, s);
      case 16: return launch_grid_sep<R, 16>(p, vis, wt, dopsf, k->d_table, grids, partial, sumwt, s);
      default: break;
    }
    set_error("separable kernel support %d not instantiated; pass the full table", g.support);
    return ARLMI_E_UNSUPPORTED;
  }
  const size_t sh = sizeof(double) * 2 * (size_t)g.tile * lds_stride(g.tile);
  if (sh > lds_limit()) {
    set_error("gridding tile (T=%d, S=%d) needs %zu B of LDS; use a smaller tile", g.tile, g.support, sh);
    return ARLMI_E_UNSUPPORTED;
  }
  if (p->ngitems == 0) return ARLMI_OK;
  ARLMI_LAUNCH("k_grid_tab", k_grid_tab<R>, dim3((unsigned)p->ngitems), dim3(256), sh, s, p->d_gitems, p->d_entries,
               g.nz > 1 ? p->d_entry_z : nullptr, (const cplx<R>*)vis, (const R*)wt, dopsf,
               (const cplx<R>*)k->d_table, g.os, g.support, (cplx<R>*)grids, partial, sumwt, g.tile, g.ntx, g.n);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

template <typename R>
static int grid_impl(const arlmi_plan* p, const void* vis, const void* wt, int dopsf, const arlmi_kernel* k,
                     void* grids, double* sumwt, hipStream_t s) {
  const DevGeom& g = p->g;
  ARLMI_HIP(hipMemsetAsync(sumwt, 0, sizeof(double) * (size_t)g.nplanes * g.nchan, s));
  cplx<R>* partial = nullptr;
  int rc = partial_buffer<R>(p, &partial);
  if (rc) return rc;
  rc = grid_launch<R>(p, vis, wt, dopsf, k, grids, partial, sumwt, s);
  if (rc) return rc;
  if (p->nmulti > 0) {
    ARLMI_LAUNCH("k_reduce_tiles", k_reduce_tiles<R>, dim3((unsigned)p->nmulti, (g.tile * g.tile + 255) / 256),
                 dim3(256), 0, s, p->d_multi, partial, (cplx<R>*)grids, g.tile, g.ntx, g.n);
    ARLMI_LAUNCH_CHECK();
  }
  return ARLMI_OK;
}

template <typename R>
static int degrid_impl(const arlmi_plan* p, const void* grids, const arlmi_kernel* k, void* vis_out, hipStream_t s) {
  const DevGeom& g = p->g;
  if (p->norphans > 0) {
    ARLMI_LAUNCH("k_zero_rows", k_zero_rows<R>, dim3((unsigned)((p->norphans + 255) / 256)), dim3(256), 0, s,
                 reinterpret_cast<const uint32_t*>(p->d_recs) + 4 * p->orphan_start, 4, p->norphans,
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
  ARLMI_LAUNCH("k_degrid_tab", k_degrid_tab<R>, dim3((unsigned)p->nditems), dim3(256), sh, s, p->d_ditems, p->d_entries,
                     g.nz > 1 ? p->d_entry_z : nullptr, (const cplx<R>*)grids, (const cplx<R>*)k->d_table, g.os,
                     g.support, (cplx<R>*)vis_out, g.tile, g.ntx, g.n);
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

}  // namespace arlmi

using namespace arlmi;

extern "C" {

int arlmi_grid(const arlmi_plan* p, const void* d_vis, const void* d_weight, int dopsf, const arlmi_kernel* k,
               void* d_grids, double* d_sumwt, int precision, void* stream) {
  int rc = check_kernel(p, k);
  if (rc) return rc;
  if (!d_weight || !d_grids || !d_sumwt || (!dopsf && !d_vis && p->nentries > 0)) {
    set_error("arlmi_grid: NULL buffer");
    return ARLMI_E_ARG;
  }
  hipStream_t s = as_stream(stream);
  if (precision == 32) return grid_impl<float>(p, d_vis, d_weight, dopsf, k, d_grids, d_sumwt, s);
  if (precision == 64) return grid_impl<double>(p, d_vis, d_weight, dopsf, k, d_grids, d_sumwt, s);
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
