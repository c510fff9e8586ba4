// Shifted 2-D FFTs and the image-plane work fused into them.
//
// ARL's fft / ifft (processing_library/fourier_transforms/fft_support.py:21-89) are
// fftshift(fft2(ifftshift(a))).  For even N this equals (-1)^(k+l) * DFT2((-1)^(x+y) a), so
// the shifts become a sign applied while loading the first pass and while storing the
// second.  Each pass is an LDS-staged Stockham transform (radix 8, remainder radix 4 or 2 in
// the last stage): a row pass works on contiguous rows, a column pass stages NB adjacent
// columns (NB * 8 or 16 bytes per grid row) so that its loads and stores stay coalesced.
//
// Fused around the passes:
//   image -> grids (predict): model * gcf * conj(b_p) for every w plane p, in the row pass
//     (gridding.py:292-302 with wstack_single.py:50-59, b_p = exp(-2 pi i w_p (1 - n)));
//     the columns follow.
//   grids -> image (invert): the columns first (in place), then a row pass that forms
//     Re(b_p * ifft(G_p)) and sums it over planes (gridding.py:270-289, wstack_single.py:93-94,
//     imaging_serial.py:193-196); with several plane groups each group writes a partial image
//     and a finishing pass adds them; the gcf product and the division by the sum of weights
//     (imaging/base.py:84-101) close the last pass.  Both fused passes read and write whole
//     rows, and the planes are split into groups so that a launch has thousands of blocks.
// The screen phase, the gcf product and the weight division are evaluated in float64.
#include "arlmi_internal.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

namespace arlmi {

template <typename R>
__device__ __forceinline__ cplx<R> cmk(R x, R y) {
  cplx<R> c;
  c.x = x;
  c.y = y;
  return c;
}
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) {
  C r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}
template <typename C>
__device__ __forceinline__ C csub(C a, C b) {
  C r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
// multiply by -i (forward) or +i (inverse)
template <typename R, bool INV>
__device__ __forceinline__ cplx<R> mul_mi(cplx<R> a) {
  return INV ? cmk<R>(-a.y, a.x) : cmk<R>(a.y, -a.x);
}

template <typename R, bool INV>
__device__ __forceinline__ void dft2(cplx<R>* v) {
  const cplx<R> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename R, bool INV>
__device__ __forceinline__ void dft4(cplx<R>& a0, cplx<R>& a1, cplx<R>& a2, cplx<R>& a3) {
  const cplx<R> t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_mi<R, INV>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <typename R, bool INV>
__device__ __forceinline__ void dft8(cplx<R>* v) {
  // even / odd split: E = DFT4(v0, v2, v4, v6), O = DFT4(v1, v3, v5, v7)
  cplx<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cplx<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<R, INV>(e0, e1, e2, e3);
  dft4<R, INV>(o0, o1, o2, o3);
  const R h = R(0.70710678118654752440084436210484903928);
  // W8^1 = (1 -+ i)/sqrt2, W8^2 = -+i, W8^3 = (-1 -+ i)/sqrt2
  const cplx<R> w1o1 = INV ? cmk<R>(h * (o1.x - o1.y), h * (o1.x + o1.y)) : cmk<R>(h * (o1.x + o1.y), h * (o1.y - o1.x));
  const cplx<R> w2o2 = mul_mi<R, INV>(o2);
  const cplx<R> w3o3 = INV ? cmk<R>(-h * (o3.x + o3.y), h * (o3.x - o3.y)) : cmk<R>(h * (o3.y - o3.x), -h * (o3.x + o3.y));
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, w1o1);
  v[5] = csub(e1, w1o1);
  v[2] = cadd(e2, w2o2);
  v[6] = csub(e2, w2o2);
  v[3] = cadd(e3, w3o3);
  v[7] = csub(e3, w3o3);
}

template <typename R, int RAD, bool INV>
__device__ __forceinline__ void dft(cplx<R>* v) {
  if constexpr (RAD == 8) dft8<R, INV>(v);
  else if constexpr (RAD == 4) dft4<R, INV>(v[0], v[1], v[2], v[3]);
  else if constexpr (RAD == 2) dft2<R, INV>(v);
}

// LDS layouts of NB transforms of length N:
//   LAY_ROW    [b][n];
//   LAY_ROWPAD [b][n] with one pad element after every 8 (the Stockham writes of the first
//              stages go to stride-8 positions, which would otherwise hit the same banks; the
//              extra index arithmetic costs more than it saves in the VALU-heavy invert row pass);
//   LAY_COL    [n][b].
constexpr int LAY_ROW = 0, LAY_ROWPAD = 1, LAY_COL = 2;
template <int LAY, int N, int NB>
__device__ __forceinline__ int lidx(int b, int n) {
  return LAY == LAY_COL ? n * NB + b : LAY == LAY_ROWPAD ? b * (N + N / 8) + n + (n >> 3) : b * N + n;
}
// elements of LDS a layout of NB transforms occupies
template <int LAY, int N>
__host__ __device__ constexpr size_t lds_elems(int nb) {
  return LAY == LAY_ROWPAD ? (size_t)nb * (N + N / 8) : (size_t)N * nb;
}

// Stockham stages.  Stage NS (1, 8, 64, ...) of radix RAD: butterfly j = t + q*NT of transform b
// reads elements j + r*(N/RAD), applies twiddles exp(-+2 pi i r k / (NS*RAD)) (k = j % NS) and a
// radix-RAD DFT, and writes elements (j / NS)*NS*RAD + k + r*NS.  The first stage (NS = 1, radix
// 8) reads straight from the caller's loader and the last one (NS = NSL, the largest power of 8
// below N, radix N / NSL) hands its outputs, element t + q*NT + r*NSL, to the caller's storer:
// only the stages between go through LDS twice.

// LDS layout of the invert row pass: padded for 16-byte values, whose stride-8 Stockham writes
// would otherwise put every group of 8 lanes on the same banks (float64 at 2048^2: 2.26 -> 2.03 ms);
// unpadded for 8-byte ones, where the index arithmetic costs what the conflicts do.
template <typename R>
__host__ __device__ constexpr int inv_lay() {
  return sizeof(R) == 8 ? LAY_ROWPAD : LAY_ROW;
}

template <int N>
__host__ __device__ constexpr int last_ns() {
  int ns = 1;
  while (ns * 8 < N) ns *= 8;
  return ns;
}

// Twiddle exp(-2 pi i idx / N), idx < N.  From a table holding only the first quarter period
// (exp(-2 pi i j / N), j < N/4), idx = j + m N/4 is the table value turned by m quarter turns
// clockwise, i.e. times (-i)^m: (x, y) -> (y, -x) for odd m, then negated for m >= 2.
// The quarter-period table in double precision, where it frees LDS for the row passes and the
// two-level column passes (float64: 2048^2 predict rows 1.80 -> 1.75 ms, streamed 8192^2 step
// 630 -> 604 ms); the full table in single precision, where the turn's selects cost more than
// the LDS they save (the invert row pass 0.70 -> 0.90 ms with a quarter table).
template <typename R>
__host__ __device__ constexpr bool tw_quarter() {
  return sizeof(R) == 8;
}
template <typename R, int N>
__device__ __forceinline__ cplx<R> tw_get(const cplx<R>* __restrict__ tw, int idx) {
  if constexpr (!tw_quarter<R>()) return tw[idx];
  constexpr int Q = N / 4;
  static_assert((Q & (Q - 1)) == 0 && Q >= 1, "power-of-two transform length");
  const int m = idx / Q;
  const cplx<R> w = tw[idx & (Q - 1)];
  cplx<R> r;
  r.x = (m & 1) ? w.y : w.x;
  r.y = (m & 1) ? -w.x : w.y;
  if (m & 2) {
    r.x = -r.x;
    r.y = -r.y;
  }
  return r;
}

template <typename R, int N, int NT, int RAD, int NS, bool INV>
__device__ __forceinline__ void stage_math(cplx<R> (&v)[(N / RAD) / NT][RAD], int t, const cplx<R>* __restrict__ tw) {
  constexpr int PER = (N / RAD) / NT;
  static_assert(PER * NT * RAD == N, "butterflies must divide among threads");
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int j = t + q * NT;
    const int k = j % NS;
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < RAD; ++r) {
        cplx<R> w = tw_get<R, N>(tw, r * k * (N / (NS * RAD)));
        if (INV) w.y = -w.y;
        v[q][r] = cmul(v[q][r], w);
      }
    }
    dft<R, RAD, INV>(v[q]);
  }
}

template <typename R, int N, int NB, int NT, int RAD, int NS, int LAY>
__device__ __forceinline__ void stage_store_lds(cplx<R>* lds, int b, int t, const cplx<R> (&v)[(N / RAD) / NT][RAD]) {
  constexpr int PER = (N / RAD) / NT;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int j = t + q * NT;
    const int idxD = (j / NS) * NS * RAD + j % NS;
#pragma unroll
    for (int r = 0; r < RAD; ++r) lds[lidx<LAY, N, NB>(b, idxD + r * NS)] = v[q][r];
  }
}

template <typename R, int N, int NB, int NT, int RAD, int LAY>
__device__ __forceinline__ void stage_load_lds(const cplx<R>* lds, int b, int t, cplx<R> (&v)[(N / RAD) / NT][RAD]) {
  constexpr int PER = (N / RAD) / NT, NBF = N / RAD;
#pragma unroll
  for (int q = 0; q < PER; ++q)
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[q][r] = lds[lidx<LAY, N, NB>(b, t + q * NT + r * NBF)];
}

// the radix-8 stages between the first and the last: read, barrier, compute, write, barrier
template <typename R, int N, int NB, int NT, int NS, bool INV, int LAY>
__device__ __forceinline__ void mid_stages(cplx<R>* lds, int b, int t, const cplx<R>* __restrict__ tw) {
  if constexpr (NS < last_ns<N>()) {
    cplx<R> v[(N / 8) / NT][8];
    stage_load_lds<R, N, NB, NT, 8, LAY>(lds, b, t, v);
    __syncthreads();
    stage_math<R, N, NT, 8, NS, INV>(v, t, tw);
    stage_store_lds<R, N, NB, NT, 8, NS, LAY>(lds, b, t, v);
    __syncthreads();
    mid_stages<R, N, NB, NT, NS * 8, INV, LAY>(lds, b, t, tw);
  }
}

// One length-N transform per b over the work group: ld(q, r, n) returns input element
// n = t + q*NT + r*N/8, st(q, r, n, v) receives output element n = t + q*NT + r*NSL.  The pass
// ends right after its last LDS reads: a caller that reuses the LDS must synchronise first.
template <typename R, int N, int NB, int NT, bool INV, int LAY, typename LD, typename ST>
__device__ __forceinline__ void fft_pass(cplx<R>* lds, int b, int t, const cplx<R>* __restrict__ tw, LD&& ld,
                                         ST&& st) {
  constexpr int NSL = last_ns<N>(), RADL = N / NSL;
  static_assert(N >= 64, "at least two stages");
  {
    constexpr int PER = (N / 8) / NT;
    cplx<R> v[PER][8];
#pragma unroll
    for (int q = 0; q < PER; ++q)
#pragma unroll
      for (int r = 0; r < 8; ++r) v[q][r] = ld(q, r, t + q * NT + r * (N / 8));
    stage_math<R, N, NT, 8, 1, INV>(v, t, tw);
    stage_store_lds<R, N, NB, NT, 8, 1, LAY>(lds, b, t, v);
  }
  __syncthreads();
  mid_stages<R, N, NB, NT, 8, INV, LAY>(lds, b, t, tw);
  {
    constexpr int PER = (N / RADL) / NT;
    cplx<R> v[PER][RADL];
    stage_load_lds<R, N, NB, NT, RADL, LAY>(lds, b, t, v);
    stage_math<R, N, NT, RADL, NSL, INV>(v, t, tw);
#pragma unroll
    for (int q = 0; q < PER; ++q)
#pragma unroll
      for (int r = 0; r < RADL; ++r) st(q, r, t + q * NT + r * NSL, v[q][r]);
  }
}

// fft_pass with the first-stage inputs already in v (element t + q*NT + r*N/8 at v[q][r]); after
// the first stage has stored its outputs to LDS, between() runs (it may refill v, e.g. with the
// next transform's inputs, whose loads then overlap the remaining stages).
template <typename R, int N, int NB, int NT, bool INV, int LAY, typename BT, typename ST>
__device__ __forceinline__ void fft_pass_v(cplx<R>* lds, int b, int t, const cplx<R>* __restrict__ tw,
                                           cplx<R> (&v)[(N / 8) / NT][8], BT&& between, ST&& st) {
  constexpr int NSL = last_ns<N>(), RADL = N / NSL;
  static_assert(N >= 64, "at least two stages");
  stage_math<R, N, NT, 8, 1, INV>(v, t, tw);
  stage_store_lds<R, N, NB, NT, 8, 1, LAY>(lds, b, t, v);
  between();
  __syncthreads();
  mid_stages<R, N, NB, NT, 8, INV, LAY>(lds, b, t, tw);
  {
    constexpr int PER = (N / RADL) / NT;
    cplx<R> u[PER][RADL];
    stage_load_lds<R, N, NB, NT, RADL, LAY>(lds, b, t, u);
    stage_math<R, N, NT, RADL, NSL, INV>(u, t, tw);
#pragma unroll
    for (int q = 0; q < PER; ++q)
#pragma unroll
      for (int r = 0; r < RADL; ++r) st(q, r, t + q * NT + r * NSL, u[q][r]);
  }
}

// The twiddle table lives in LDS after the transform data (LDS loads cannot be hoisted
// across the stage barriers, global ones can, which would pin every stage's twiddles in
// registers at once); in double precision only its first quarter period (tw_get turns it).
// Callers size dynamic LDS as data + tw_lds_elems<R, N>() complex values.
template <typename R, int N>
__host__ __device__ constexpr size_t tw_lds_elems() {
  return tw_quarter<R>() ? (size_t)N / 4 : (size_t)N;
}
template <typename R, int N, bool TWG = false>
__device__ __forceinline__ const cplx<R>* stage_twiddles_at(unsigned char* smem, size_t data_bytes,
                                                            const cplx<R>* twg) {
  if constexpr (TWG) return twg;  // the table does not fit beside the data: read it from L2
  cplx<R>* tw = reinterpret_cast<cplx<R>*>(smem + data_bytes);
  for (int i = threadIdx.x; i < (int)tw_lds_elems<R, N>(); i += blockDim.x) tw[i] = twg[i];
  return tw;
}

template <int N>
struct Shape {
  static constexpr int NT = N >= 16 ? N / 16 : 1;       // threads per transform
  static constexpr int ROW_NB = (NT >= 256) ? 1 : 256 / NT;
  // LDS budget: NB transforms plus one twiddle table of N values within 160 KiB
  template <typename R>
  static constexpr int col_nb() {
    constexpr int by_lds = (int)(163840 / (N * sizeof(cplx<R>))) - 1;
    constexpr int by_thr = 1024 / NT;
    constexpr int nb = by_lds < by_thr ? by_lds : by_thr;
    constexpr int c = nb > 16 ? 16 : (nb < 1 ? 1 : nb);
    return c >= 16 ? 16 : c >= 8 ? 8 : c >= 4 ? 4 : c >= 2 ? 2 : 1;  // a power of two dividing N
  }
  // Columns of the split pass (k_cols_split: one radix-2 step in registers, then two transforms of
  // N/2 through LDS): 128-byte row segments where N/2 of them and the twiddles of N/2 fit LDS,
  // else 64-byte ones; 0 = the plain pass.  Used from N = 1024 to 4096, where it measured faster
  // than the plain pass of col_nb() columns (arlmi.fft of dense planes, an input copy included,
  // complex64: 1024^2 2.75 -> 2.58 ms for 256 planes, 2048^2 3.08 -> 2.77 ms for 64, 4096^2 4.21
  // -> 3.29 ms for 16 (64-byte segments; 128-byte ones would not fit); complex128: 1024^2 2.77 ->
  // 2.68 ms, 2048^2 3.16 -> 2.71 ms); 64-byte segments where 128-byte ones fit measured slower
  // (complex64 at 2048^2: 2.87 ms; at 1024^2: 2.79 ms, slower than the plain pass).
  template <typename R>
  static constexpr int split_nb() {
    if (N < 1024 || N > 4096 || col_nb<R>() < 2) return 0;
    constexpr size_t tw = sizeof(cplx<R>) * tw_lds_elems<R, N / 2>();
    constexpr int nb128 = 128 / (int)sizeof(cplx<R>), nb64 = 64 / (int)sizeof(cplx<R>);
    if ((size_t)(N / 2) * 128 + tw <= 163840) return nb128;
    if ((size_t)(N / 2) * 64 + tw <= 163840) return nb64;
    return 0;
  }
  // the invert column pass also holds per-pixel accumulators: half the columns, twice the registers
  template <typename R>
  static constexpr int inv_nb() {
    constexpr int nb = col_nb<R>() / 2;
    return nb < 1 ? 1 : nb;
  }
  // the predict row pass (screens per plane, stores): N/8 threads per row, one radix-8
  // butterfly per thread and stage, fewer elements (and registers) per thread than the
  // plain passes
  static constexpr int FNT = N >= 8 ? N / 8 : 1;
  static constexpr int FNB = FNT >= 512 ? 1 : 512 / FNT;
  // A padded row and the twiddle table together exceed LDS (complex128 at N = 8192): the row
  // passes then read their twiddles from the global table (L2-resident) instead.
  template <typename R>
  static constexpr bool tw_global() {
    return sizeof(cplx<R>) * (lds_elems<LAY_ROWPAD, N>(FNB > ROW_NB ? FNB : ROW_NB) + tw_lds_elems<R, N>()) > 163840;
  }
  template <typename R>
  static constexpr size_t tw_lds() {
    return tw_global<R>() ? 0 : tw_lds_elems<R, N>();
  }
  // Only one column fits LDS beside the twiddles (complex64 at 8192, complex128 from 4096): a
  // column pass would read one 8- or 16-byte value per grid row, so the columns are transformed
  // as rows between two in-place transposes instead (three coalesced passes instead of one
  // scattered one).
  template <typename R>
  static constexpr bool cols_by_transpose() {
    return col_nb<R>() < 2;
  }
  template <typename R>
  static constexpr bool fits() {
    return sizeof(cplx<R>) * (lds_elems<LAY_ROWPAD, N>(ROW_NB) + tw_lds<R>()) <= 163840 &&
           sizeof(cplx<R>) * (lds_elems<LAY_ROWPAD, N>(FNB) + tw_lds<R>()) <= 163840 &&
           (cols_by_transpose<R>() || (size_t)N * sizeof(cplx<R>) * (size_t)(col_nb<R>() + 1) <= 163840);
  }
};

// w-screen helpers ---------------------------------------------------------

// 1 - sqrt(1 - r^2) for pixel (y, x); r2 >= 1 flags the pixel as outside the unit circle.
__device__ __forceinline__ double screen_q(int y, int x, int n, double fov, bool* outside) {
  const double c = (double)(n / 2);
  const double ly = ((double)y - c) / (double)n;
  const double mx = ((double)x - c) / (double)n;
  const double r2 = fov * fov * (ly * ly + mx * mx);
  *outside = r2 >= 1.0;
  return *outside ? 0.0 : 1.0 - sqrt(1.0 - r2);
}

// exp(2 pi i t) for t given in turns, evaluated in float64 up to the range reduction.
// Single precision: t is reduced to [-1/2, 1/2] in float64, then to an octant
// f in [-1/8, 1/8] turns, and sin / cos of 2 pi f come from their Taylor series
// (truncation below 3e-8 on that interval), rotated by the quadrant.
template <typename R>
__device__ __forceinline__ cplx<R> expi_turns(double t);
template <>
__device__ __forceinline__ float2 expi_turns<float>(double t) {
  const double red = t - rint(t);                    // [-1/2, 1/2]
  const float q4 = rintf((float)(red * 4.0));        // quadrant -2..2
  const float f = (float)(red - 0.25 * (double)q4);  // [-1/8, 1/8]
  const float th = 6.28318530717958647692f * f;
  const float t2 = th * th;
  const float sn = th * (1.0f + t2 * (-1.0f / 6.0f + t2 * (1.0f / 120.0f + t2 * (-1.0f / 5040.0f + t2 * (1.0f / 362880.0f)))));
  const float cs = 1.0f + t2 * (-0.5f + t2 * (1.0f / 24.0f + t2 * (-1.0f / 720.0f + t2 * (1.0f / 40320.0f))));
  const int q = ((int)q4) & 3;
  // rotate (cs, sn) by q quarter turns, without branches: odd q swaps the pair, and the cosine is
  // negated for q = 1, 2 and the sine for q = 2, 3 (sign-bit flips)
  const bool swap = (q & 1) != 0;
  const float x = swap ? sn : cs, y = swap ? cs : sn;
  const uint32_t nc = (uint32_t)((q + 1) & 2) << 30, ns = (uint32_t)(q & 2) << 30;
  return make_float2(__uint_as_float(__float_as_uint(x) ^ nc), __uint_as_float(__float_as_uint(y) ^ ns));
}
template <>
__device__ __forceinline__ double2 expi_turns<double>(double t) {
  double s, c;
  sincospi(2.0 * (t - rint(t)), &s, &c);
  return make_double2(c, s);
}

constexpr double kMinusTwoPi = -2.0 * 3.14159265358979323846264338327950288;

// phase -2 pi w (1 - n) of the w screen, in turns
__device__ __forceinline__ double screen_turns(double w, double q) { return -(w * q); }

// The screen phasor exp(-2 pi i w q) of a row pass, with q = 1 - n per pixel (held across the
// planes) and w per plane.  Double precision: the float64 phase and sincospi.  Single precision:
// w and q are each held as two floats (hi + lo, hi the float nearest the double), the product's
// rounding error comes from a fused multiply-add (hi*hi exactly as p + e, plus the cross terms;
// lo*lo, ~2^-48 of the phase, is left out), p is reduced modulo 1 and to an octant exactly in
// float, e is added to the octant remainder (a rounding at |g| <= 1/8 + |e|, 2^-28 turns), and
// sin / cos of the remainder come from the Taylor series of expi_turns: the float64 reduction's
// accuracy without float64 arithmetic.
template <typename R>
struct ScreenQ {
  double q;
};
template <>
struct ScreenQ<float> {
  float hi, lo;
};
template <typename R>
__device__ __forceinline__ ScreenQ<R> screen_split(double q);
template <>
__device__ __forceinline__ ScreenQ<double> screen_split<double>(double q) {
  return ScreenQ<double>{q};
}
template <>
__device__ __forceinline__ ScreenQ<float> screen_split<float>(double q) {
  const float hi = (float)q;
  return ScreenQ<float>{hi, (float)(q - (double)hi)};
}
template <typename R>
__device__ __forceinline__ cplx<R> screen_phasor(const ScreenQ<R>& w, const ScreenQ<R>& q);
template <>
__device__ __forceinline__ double2 screen_phasor<double>(const ScreenQ<double>& w, const ScreenQ<double>& q) {
  return expi_turns<double>(screen_turns(w.q, q.q));
}
template <>
__device__ __forceinline__ float2 screen_phasor<float>(const ScreenQ<float>& w, const ScreenQ<float>& q) {
  // turns t = -(w q) = p + e with w negated on entry
  const float nwh = -w.hi, nwl = -w.lo;
  const float p = nwh * q.hi;
  float e = __builtin_fmaf(nwh, q.hi, -p);
  e = __builtin_fmaf(nwh, q.lo, e);
  e = __builtin_fmaf(nwl, q.hi, e);
  const float f = p - rintf(p);          // exact: [-1/2, 1/2]
  const float q4 = rintf(f * 4.0f);      // quadrant -2..2
  const float g = __builtin_fmaf(-0.25f, q4, f) + e;  // the first two terms cancel exactly
  const float th = 6.28318530717958647692f * g;
  const float t2 = th * th;
  const float sn = th * (1.0f + t2 * (-1.0f / 6.0f + t2 * (1.0f / 120.0f + t2 * (-1.0f / 5040.0f + t2 * (1.0f / 362880.0f)))));
  const float cs = 1.0f + t2 * (-0.5f + t2 * (1.0f / 24.0f + t2 * (-1.0f / 720.0f + t2 * (1.0f / 40320.0f))));
  const int qi = ((int)q4) & 3;
  const bool swap = (qi & 1) != 0;
  const float x = swap ? sn : cs, y = swap ? cs : sn;
  const uint32_t nc = (uint32_t)((qi + 1) & 2) << 30, ns = (uint32_t)(qi & 2) << 30;
  return make_float2(__uint_as_float(__float_as_uint(x) ^ nc), __uint_as_float(__float_as_uint(y) ^ ns));
}

struct FftArgs {
  void* data;            // grids [planes][chan][N][N] (complex)
  int nbatch;            // plain mode: number of planes
  int nplanes, nchan;
  int in_mod, out_mod;   // apply (-1)^(x+y) on load / store
  double scale;
  const void* model;     // [chan][N][N] real (precision R)
  const double* gcf;     // [chan][N][N]
  const double* plane_w; // [nplanes] w-screen centres
  const double* sumwt;   // [nplanes][chan]
  void* img_re;
  void* img_im;
  int normalize, screens, wstack;
  double fov;            // N * cellsize
  int ppg;               // planes per plane group (fused row passes)
  void* part;            // invert: partial images [groups][chan][N][N] (x2 with the imaginary part)
  // sparse grids (at most 32 tiles per side): tiles of side `tile`, ntx per side; for
  // plane-channel pc, colmask[pc] has bit tx set if tile column tx holds entries and
  // tilemask[pc * ntx + tx] bit ty if tile (ty, tx) does; zero_fill: tiles without entries were
  // never written and read as 0; zeros: 64 zero entries
  const uint32_t* colmask;
  const uint32_t* tilemask;
  const void* zeros;
  int zero_fill;
  int tile, ntx;
  int tshift;  // log2(tile)
  // grid rows in digit-transposed order (the two-level column transform's output): row v of a
  // plane is stored at N1 * (v % (N / N1)) + v / (N / N1); 0 = natural order
  int perm_n1;
};


// Masked-out loads read zeros from a small buffer instead of the grid (lane l reads entry
// l & 63, 16 bytes at most per entry), so that every load of a pass is issued unconditionally
// and the loads of a thread stay in flight together.
template <typename R>
__device__ __forceinline__ const cplx<R>* zero_entry(const void* zeros) {
  return reinterpret_cast<const cplx<R>*>(zeros) + (threadIdx.x & 63);
}

// ---------------------------------------------------------------------------
// row passes
// ---------------------------------------------------------------------------

template <typename R, int N, bool INV>
__global__ __launch_bounds__(Shape<N>::NT * Shape<N>::ROW_NB) void k_rows_plain(FftArgs a, const cplx<R>* __restrict__ twg) {
  constexpr int NT = Shape<N>::NT, NB = Shape<N>::ROW_NB;
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem);
  const cplx<R>* tw =
      stage_twiddles_at<R, N, Shape<N>::template tw_global<R>()>(smem, sizeof(cplx<R>) * lds_elems<LAY_ROWPAD, N>(NB), twg);
  const int b = threadIdx.x / NT, t = threadIdx.x % NT;
  const int64_t row = (int64_t)blockIdx.x * NB + b;
  cplx<R>* data = reinterpret_cast<cplx<R>*>(a.data) + row * N;
  const int ym = (int)(row % N);
  fft_pass<R, N, NB, NT, INV, LAY_ROWPAD>(
      lds, b, t, tw,
      [&](int, int, int n) {
        cplx<R> v = data[n];
        if (a.in_mod && ((ym + n) & 1)) v = cmk<R>(-v.x, -v.y);
        return v;
      },
      [&](int, int, int n, cplx<R> v) {
        R sc = (R)a.scale;
        if (a.out_mod && ((ym + n) & 1)) sc = -sc;
        data[n] = cmk<R>(v.x * sc, v.y * sc);
      });
}

// Predict: grids[p][c][y][:] = rowfft((-1)^(x+y) * model[c][y][x] * gcf[c][y][x] * conj(b_p(x,y)))
// Grid: (row blocks of every channel, plane groups).  The model term and the screen's q of
// the elements each thread feeds to the first stage are computed once, before the plane loop.
template <typename R, int N>
__global__ __launch_bounds__(Shape<N>::FNT * Shape<N>::FNB) void k_rows_predict(FftArgs a, const cplx<R>* __restrict__ twg) {
  constexpr int NT = Shape<N>::FNT, NB = Shape<N>::FNB;
  constexpr int PER = (N / 8) / NT;
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem);
  const cplx<R>* tw =
      stage_twiddles_at<R, N, Shape<N>::template tw_global<R>()>(smem, sizeof(cplx<R>) * lds_elems<LAY_ROWPAD, N>(NB), twg);
  cplx<R>* grids = reinterpret_cast<cplx<R>*>(a.data);
  const R* model = reinterpret_cast<const R*>(a.model);
  constexpr int RB = N / NB;  // row blocks per channel
  const int c = blockIdx.x / RB;
  const int b = threadIdx.x / NT, t = threadIdx.x % NT;
  const int y = (blockIdx.x % RB) * NB + b;
  R mv[PER][8];
  ScreenQ<R> qs[PER][8];
#pragma unroll
  for (int q = 0; q < PER; ++q)
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int x = t + q * NT + r * (N / 8);
      const size_t gi = ((size_t)c * N + y) * N + x;
      double m = (double)model[gi] * a.gcf[gi];
      if ((x + y) & 1) m = -m;
      bool outside = false;
      qs[q][r] = screen_split<R>(a.screens ? screen_q(y, x, N, a.fov, &outside) : 0.0);
      mv[q][r] = outside ? R(0) : (R)m;
    }
  const int p0 = blockIdx.y * a.ppg, p1 = min(a.nplanes, p0 + a.ppg);  // this block's plane group
  for (int p = p0; p < p1; ++p) {
    const ScreenQ<R> wsp = screen_split<R>(a.screens ? a.plane_w[p] : 0.0);
    cplx<R>* dst = grids + ((size_t)p * a.nchan + c) * N * N + (size_t)y * N;
    const uint32_t cm = a.colmask ? a.colmask[(size_t)p * a.nchan + c] : ~0u;
    fft_pass<R, N, NB, NT, false, LAY_ROWPAD>(
        lds, b, t, tw,
        [&](int q, int r, int) {
          if (!a.screens) return cmk<R>(mv[q][r], R(0));
          const cplx<R> s = screen_phasor<R>(wsp, qs[q][r]);  // b; conj(b) = (s.x, -s.y)
          return cmk<R>(mv[q][r] * s.x, -mv[q][r] * s.y);
        },
        [&](int, int, int x, cplx<R> v) {
          if ((cm >> (x >> a.tshift)) & 1u) dst[x] = v;
        });
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// column passes
// ---------------------------------------------------------------------------

template <typename R, int N, bool INV>
__global__ __launch_bounds__(Shape<N>::NT * Shape<N>::template col_nb<R>()) void k_cols_plain(FftArgs a, const cplx<R>* __restrict__ twg) {
  constexpr int NT = Shape<N>::NT, NB = Shape<N>::template col_nb<R>();
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem);
  const cplx<R>* tw = stage_twiddles_at<R, N>(smem, sizeof(cplx<R>) * (size_t)N * NB, twg);
  constexpr int CB = N / NB;
  // column groups sharing a 128-byte line run on one XCD
  constexpr int LG = (NB * (int)sizeof(cplx<R>)) >= 128 ? 1 : 128 / (NB * (int)sizeof(cplx<R>));
  const int bid = xcd_block(blockIdx.x, gridDim.x, LG);
  const int64_t plane = bid / CB;
  const int x0 = (bid % CB) * NB;
  uint32_t tm = ~0u;
  if (a.tilemask) {  // sparse grids: skip a column group whose tile column holds no entries
    tm = a.tilemask[(size_t)plane * a.ntx + (x0 >> a.tshift)];  // NB divides the tile side
    if (!tm) return;
    if (!a.zero_fill) tm = ~0u;
  } else if (a.colmask && !((a.colmask[plane] >> (x0 >> a.tshift)) & 1u)) {
    return;  // a column band of zeros (the caller's band mask)
  }
  const int b = threadIdx.x % NB, t = threadIdx.x / NB;
  const int x = x0 + b;
  cplx<R>* col = reinterpret_cast<cplx<R>*>(a.data) + plane * (int64_t)N * N + x;
  fft_pass<R, N, NB, NT, INV, LAY_COL>(
      lds, b, t, tw,
      [&](int, int, int n) {
        // a tile never written (sparse grids) reads as 0
        const cplx<R>* src = ((tm >> (n >> a.tshift)) & 1u) ? col + (size_t)n * N : zero_entry<R>(a.zeros);
        cplx<R> v = *src;
        if (a.in_mod && ((n + x) & 1)) v = cmk<R>(-v.x, -v.y);
        return v;
      },
      [&](int, int, int n, cplx<R> v) {
        R sc = (R)a.scale;
        if (a.out_mod && ((n + x) & 1)) sc = -sc;
        col[(size_t)n * N] = cmk<R>(v.x * sc, v.y * sc);
      });
}

// The column pass with half the LDS: one radix-2 decimation-in-frequency step in registers,
// a_n = x_n + x_(n+N/2) and b_n = (x_n - x_(n+N/2)) W_N^n, then X_2m = DFT_(N/2)(a)_m and
// X_(2m+1) = DFT_(N/2)(b)_m as two LDS-staged transforms of N/2, one after the other, while b
// waits in registers.  NB columns per block, N/32 threads per column (32 values each).
template <typename R, int N, bool INV, int NB>
__global__ __launch_bounds__(Shape<N / 2>::NT * NB) void k_cols_split(FftArgs a, const cplx<R>* __restrict__ twh,
                                                                     const cplx<R>* __restrict__ twn) {
  constexpr int H = N / 2, NT = Shape<H>::NT, K = N / NT, KH = K / 2;
  static_assert(K == 32 && (H / 8) / NT == 2, "32 values per thread, two first-stage butterflies");
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem);
  const cplx<R>* tw = stage_twiddles_at<R, H>(smem, sizeof(cplx<R>) * (size_t)H * NB, twh);
  constexpr int CB = N / NB;
  constexpr int LG = (NB * (int)sizeof(cplx<R>)) >= 128 ? 1 : 128 / (NB * (int)sizeof(cplx<R>));
  const int bid = xcd_block(blockIdx.x, gridDim.x, LG);
  const int64_t plane = bid / CB;
  const int x0 = (bid % CB) * NB;
  uint32_t tm = ~0u;
  if (a.tilemask) {
    tm = a.tilemask[(size_t)plane * a.ntx + (x0 >> a.tshift)];
    if (!tm) return;
    if (!a.zero_fill) tm = ~0u;
  } else if (a.colmask && !((a.colmask[plane] >> (x0 >> a.tshift)) & 1u)) {
    return;
  }
  const int b = threadIdx.x % NB, t = threadIdx.x / NB;
  const int x = x0 + b;
  cplx<R>* col = reinterpret_cast<cplx<R>*>(a.data) + plane * (int64_t)N * N + x;
  // value k of this thread is element t + k*NT; element t + k*NT + N/2 is value k + KH
  cplx<R> lo[KH], hi[KH];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int n = t + k * NT;
    const cplx<R>* src = ((tm >> (n >> a.tshift)) & 1u) ? col + (size_t)n * N : zero_entry<R>(a.zeros);
    cplx<R> v = *src;
    if (a.in_mod && ((n + x) & 1)) v = cmk<R>(-v.x, -v.y);
    if (k < KH) lo[k] = v;
    else hi[k - KH] = v;
  }
  // butterflies: lo becomes a (first-stage layout of the half transform: element t + q*NT + r*H/8
  // is value q + 2r), hi becomes b
  cplx<R> va[2][8], vb[2][8];
#pragma unroll
  for (int k = 0; k < KH; ++k) {
    cplx<R> w = twn[t + k * NT];
    if (INV) w.y = -w.y;
    va[k & 1][k >> 1] = cadd(lo[k], hi[k]);
    vb[k & 1][k >> 1] = cmul(csub(lo[k], hi[k]), w);
  }
  const R sc0 = (R)a.scale;
  auto store = [&](int n, cplx<R> v) {
    R sc = sc0;
    if (a.out_mod && ((n + x) & 1)) sc = -sc;
    col[(size_t)n * N] = cmk<R>(v.x * sc, v.y * sc);
  };
  fft_pass_v<R, H, NB, NT, INV, LAY_COL>(
      lds, b, t, tw, va, [] {}, [&](int, int, int m, cplx<R> v) { store(2 * m, v); });
  __syncthreads();
  fft_pass_v<R, H, NB, NT, INV, LAY_COL>(
      lds, b, t, tw, vb, [] {}, [&](int, int, int m, cplx<R> v) { store(2 * m + 1, v); });
}

// Invert, after the in-place column pass: for plane group g (blockIdx.y) and the NB rows of this
// block, acc = sum_p Re(b_p * (-1)^(x+y) * rowifft(G_p)) over the group's planes.  With one group
// the block writes img = acc * gcf / S_c (imaging_serial.py:193-196 with normalize_sumwt,
// imaging/base.py:84-101); otherwise it writes acc to its partial image for k_invert_finish.
// SCREENS: w-stack planes (real part after the screen); otherwise the planes are summed and
// the imaginary part is kept as well.  Single precision accumulates in float, double in double.
// Each thread accumulates the elements the last stage hands it (x = t + q*NT + r*NSL).
template <typename R, int N, bool SCREENS>
__global__ __launch_bounds__(Shape<N>::NT * Shape<N>::ROW_NB) void k_rows_invert(FftArgs a, const cplx<R>* __restrict__ twg) {
  constexpr int NT = Shape<N>::NT, NB = Shape<N>::ROW_NB;
  constexpr int NSL = last_ns<N>(), RADL = N / NSL, PER = NSL / NT;
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem);
  const cplx<R>* tw =
      stage_twiddles_at<R, N, Shape<N>::template tw_global<R>()>(smem, sizeof(cplx<R>) * lds_elems<inv_lay<R>(), N>(NB), twg);
  const cplx<R>* grids = reinterpret_cast<const cplx<R>*>(a.data);
  constexpr int RB = N / NB;  // row blocks per channel
  const int c = blockIdx.x / RB;
  const int b = threadIdx.x / NT, t = threadIdx.x % NT;
  const int y = (blockIdx.x % RB) * NB + b;
  const int G = gridDim.y;
  const int p0 = blockIdx.y * a.ppg, p1 = min(a.nplanes, p0 + a.ppg);
  R acc[PER][RADL];
  R acci[SCREENS ? 1 : PER][SCREENS ? 1 : RADL];
  ScreenQ<R> qs[SCREENS ? PER : 1][SCREENS ? RADL : 1];
  uint32_t outside = 0;  // bit q * RADL + r: element outside the unit circle (no screen)
  static_assert(PER * RADL <= 32, "outside mask");
#pragma unroll
  for (int q = 0; q < PER; ++q)
#pragma unroll
    for (int r = 0; r < RADL; ++r) {
      acc[q][r] = R(0);
      if constexpr (!SCREENS) {
        acci[q][r] = R(0);
      } else {
        bool o;
        qs[q][r] = screen_split<R>(screen_q(y, t + q * NT + r * NSL, N, a.fov, &o));
        if (o) outside |= 1u << (q * RADL + r);
      }
    }
  // normalize 2 (the list workflows' normalize=False): each w slab's unnormalised image weighted
  // by its sum of weights (imaging_serial.py:193-194)
  const bool weigh = a.wstack && a.normalize == 2;
  // one plane's contribution to the accumulators, element x = t + q*NT + r*NSL of the row transform
  auto accumulate = [&](const ScreenQ<R>& wp, R wsc, int q, int r, int x, cplx<R> v) {
    if ((x + y) & 1) v = cmk<R>(-v.x, -v.y);
    if constexpr (SCREENS) {
      // the screen is 0 outside the unit circle (selected, not branched on: the pixels of a row
      // are in and out together in no useful pattern)
      const cplx<R> s = screen_phasor<R>(wp, qs[q][r]);
      const R val = wsc * (s.x * v.x - s.y * v.y);
      acc[q][r] += ((outside >> (q * RADL + r)) & 1u) ? R(0) : val;
    } else {
      acc[q][r] += v.x;
      acci[q][r] += v.y;
    }
  };
  // normalize_sumwt zeroes the image of a slab with no positive sum of weights
  auto skip = [&](int p) { return a.wstack && a.normalize == 1 && !(a.sumwt[(size_t)p * a.nchan + c] > 0.0); };
  if constexpr (N >= 8192 && sizeof(R) == 4) {
    // One 8-wave work group per CU at this size: the next plane's row is loaded into registers
    // while the current plane's remaining stages run, so that the CU is not left idle during the
    // row loads (with several groups per CU, as at smaller sizes, the extra registers would cost
    // one of them).
    constexpr int PER1 = (N / 8) / NT;
    cplx<R> v1[PER1][8];
    auto load_row = [&](int p) {
      const cplx<R>* row = grids + ((size_t)p * a.nchan + c) * N * N + (size_t)perm_row_of(y, N, a.perm_n1) * N;
      const uint32_t cm = a.colmask ? a.colmask[(size_t)p * a.nchan + c] : ~0u;
#pragma unroll
      for (int q = 0; q < PER1; ++q)
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const int n = t + q * NT + r * (N / 8);
          // sparse grids: a skipped tile column reads as 0
          v1[q][r] = *(((cm >> (n >> a.tshift)) & 1u) ? row + n : zero_entry<R>(a.zeros));
        }
    };
    int p = p0;
    while (p < p1 && skip(p)) ++p;
    if (p < p1) load_row(p);
    while (p < p1) {
      const double sw = a.sumwt[(size_t)p * a.nchan + c];
      const R wsc = weigh ? (R)sw : R(1);
      const ScreenQ<R> wp = screen_split<R>(SCREENS ? a.plane_w[p] : 0.0);
      int pn = p + 1;
      while (pn < p1 && skip(pn)) ++pn;
      fft_pass_v<R, N, NB, NT, true, inv_lay<R>()>(
          lds, b, t, tw, v1,
          [&]() {
            if (pn < p1) load_row(pn);
          },
          [&](int q, int r, int x, cplx<R> v) { accumulate(wp, wsc, q, r, x, v); });
      __syncthreads();
      p = pn;
    }
  } else {
    for (int p = p0; p < p1; ++p) {
      if (skip(p)) continue;
      const double sw = a.sumwt[(size_t)p * a.nchan + c];
      const R wsc = weigh ? (R)sw : R(1);
      const cplx<R>* row = grids + ((size_t)p * a.nchan + c) * N * N + (size_t)perm_row_of(y, N, a.perm_n1) * N;
      const uint32_t cm = a.colmask ? a.colmask[(size_t)p * a.nchan + c] : ~0u;
      const ScreenQ<R> wp = screen_split<R>(SCREENS ? a.plane_w[p] : 0.0);
      fft_pass<R, N, NB, NT, true, inv_lay<R>()>(
          lds, b, t, tw,
          [&](int, int, int n) {
            // sparse grids: a skipped tile column reads as 0
            const cplx<R>* src = ((cm >> (n >> a.tshift)) & 1u) ? row + n : zero_entry<R>(a.zeros);
            return *src;
          },
          [&](int q, int r, int x, cplx<R> v) { accumulate(wp, wsc, q, r, x, v); });
      __syncthreads();
    }
  }
  const size_t rowoff = ((size_t)c * N + y) * N;
  if (G > 1) {
    R* part = reinterpret_cast<R*>(a.part);
    const size_t gsz = (size_t)a.nchan * N * N;
#pragma unroll
    for (int q = 0; q < PER; ++q)
#pragma unroll
      for (int r = 0; r < RADL; ++r) {
        const size_t gi = rowoff + t + q * NT + r * NSL;
        part[(size_t)blockIdx.y * gsz + gi] = acc[q][r];
        if constexpr (!SCREENS) part[(size_t)(G + blockIdx.y) * gsz + gi] = acci[q][r];
      }
    return;
  }
  double tot = 0.0;
  for (int pp = 0; pp < a.nplanes; ++pp) tot += a.sumwt[(size_t)pp * a.nchan + c];
  R* img = reinterpret_cast<R*>(a.img_re);
  R* imi = reinterpret_cast<R*>(a.img_im);
  const bool zero = a.normalize && !(tot > 0.0);
  const bool divide = a.normalize == 1 || weigh;
#pragma unroll
  for (int q = 0; q < PER; ++q)
#pragma unroll
    for (int r = 0; r < RADL; ++r) {
      const size_t gi = rowoff + t + q * NT + r * NSL;
      const double g = a.gcf[gi];
      double re = (double)acc[q][r] * g;
      double im = 0.0;
      if constexpr (!SCREENS) im = (double)acci[q][r] * g;
      if (divide) {
        re /= tot;
        im /= tot;
      }
      if (zero) re = im = 0.0;
      img[gi] = (R)re;
      if (imi) imi[gi] = (R)im;
    }
}

// Adds the G partial images in group order (float64), then applies gcf and the weight
// normalisation as k_rows_invert does for a single group.  Grid: (pixel blocks, channels).
template <typename R>
__global__ __launch_bounds__(256) void k_invert_finish(FftArgs a, int n, int G) {
  __shared__ double stot;
  const int c = blockIdx.y;
  if (threadIdx.x < 64) {
    double sw = 0.0;
    for (int p = threadIdx.x; p < a.nplanes; p += 64) sw += a.sumwt[(size_t)p * a.nchan + c];
    for (int o = 32; o > 0; o >>= 1) sw += __shfl_xor(sw, o);
    if (threadIdx.x == 0) stot = sw;
  }
  __syncthreads();
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * n) return;
  const size_t gsz = (size_t)a.nchan * n * n;
  const size_t gi = (size_t)c * n * n + (size_t)i;
  const R* part = reinterpret_cast<const R*>(a.part);
  double re = 0.0, im = 0.0;
  for (int g = 0; g < G; ++g) {
    re += (double)part[(size_t)g * gsz + gi];
    if (!a.screens) im += (double)part[(size_t)(G + g) * gsz + gi];
  }
  const double tot = stot;
  const double gc = a.gcf[gi];
  re *= gc;
  im *= gc;
  if (a.normalize == 1 || (a.wstack && a.normalize == 2)) {
    re /= tot;
    im /= tot;
  }
  if (a.normalize && !(tot > 0.0)) re = im = 0.0;
  reinterpret_cast<R*>(a.img_re)[gi] = (R)re;
  if (a.img_im) reinterpret_cast<R*>(a.img_im)[gi] = (R)im;
}

// ---------------------------------------------------------------------------
// w screen and direct DFT
// ---------------------------------------------------------------------------

template <typename R>
__global__ void k_w_screen(int n, double fov, double w, cplx<R>* out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * n) return;
  const int y = (int)(i / n), x = (int)(i % n);
  bool outside;
  const double q = screen_q(y, x, n, fov, &outside);
  if (outside) {
    out[i] = cmk<R>(R(0), R(0));
    return;
  }
  double s, c;
  sincos(kMinusTwoPi * w * q, &s, &c);
  out[i] = cmk<R>((R)c, (R)s);
}

// exp(-2 pi i d) from the exact reduction of d modulo 1, e = d - rint(d) in [-1/2, 1/2], and
// r = -2 e in [-1, 1] (exact): cos(pi r) + i sin(pi r).  This is the phasor of d itself (numpy
// evaluates cos(fl(-2 pi d)), whose rounding of 2 pi d adds ~|2 pi d| * 1e-16 rad).  Single
// precision visibilities take the reduced angle in float32 (error ~1e-7, the storage precision);
// double precision keeps float64.
template <typename R>
__device__ __forceinline__ void dft_phasor(double d, double& c, double& s) {
  const double r = -2.0 * (d - rint(d));
  if (sizeof(R) == 4) {
    float sf, cf;
    sincospif((float)r, &sf, &cf);
    s = sf;
    c = cf;
  } else {
    sincospi(r, &s, &c);
  }
}

// One thread per row, every component: component k at (l, m) with n - 1 = sqrt(1 - l^2 - m^2) - 1
// (simulate_point, coordinate_support.py:192-212) and flux f[chan] (imaging/base.py:207-220).
// The components' (l, m, n - 1, flux) are staged in LDS once per block.
template <typename R>
__global__ void __launch_bounds__(256) k_dft(const double* __restrict__ uvw, const int32_t* __restrict__ chan,
                                             int64_t nrows, const double* __restrict__ lmf, int ncomp, int nchan,
                                             cplx<R>* __restrict__ vis) {
  extern __shared__ double comp[];
  const int in_stride = 2 + nchan, stride = 3 + nchan;
  for (int k = threadIdx.x; k < ncomp; k += blockDim.x) {
    const double l = lmf[k * in_stride], m = lmf[k * in_stride + 1];
    double nm1;
    {
#pragma clang fp contract(off)
      nm1 = sqrt(1.0 - l * l - m * m) - 1.0;
    }
    comp[k * stride] = l;
    comp[k * stride + 1] = m;
    comp[k * stride + 2] = nm1;
    for (int c = 0; c < nchan; ++c) comp[k * stride + 3 + c] = lmf[k * in_stride + 2 + c];
  }
  __syncthreads();
  const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (row >= nrows) return;
  const double u = uvw[3 * row], v = uvw[3 * row + 1], w = uvw[3 * row + 2];
  const int ch = chan ? chan[row] : 0;
  double ar = 0.0, ai = 0.0;
  for (int k = 0; k < ncomp; ++k) {
    const double* cp = comp + k * stride;
    const double d = u * cp[0] + v * cp[1] + w * cp[2];
    double s, c;
    dft_phasor<R>(d, c, s);
    const double f = cp[3 + ch];
    ar += f * c;
    ai += f * s;
  }
  cplx<R> o = vis[row];
  vis[row] = cmk<R>(o.x + (R)ar, o.y + (R)ai);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static std::mutex g_tw_mutex;
static std::map<std::pair<int, int>, void*> g_tw;  // (device*100000 + n, precision) -> table

const void* twiddles(int n, int precision) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lock(g_tw_mutex);
  const auto key = std::make_pair(dev * 100000 + n, precision);
  auto it = g_tw.find(key);
  if (it != g_tw.end()) return it->second;
  const size_t esz = precision == 32 ? sizeof(float2) : sizeof(double2);
  std::vector<unsigned char> host(esz * (size_t)n);
  for (int k = 0; k < n; ++k) {
    const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
    const double c = (double)cosl(ang), s = (double)sinl(ang);
    if (precision == 32) {
      float2 v = make_float2((float)c, (float)s);
      memcpy(&host[esz * k], &v, esz);
    } else {
      double2 v = make_double2(c, s);
      memcpy(&host[esz * k], &v, esz);
    }
  }
  void* d = nullptr;
  if (hipMalloc(&d, host.size()) != hipSuccess) return nullptr;
  if (hipMemcpy(d, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  g_tw[key] = d;
  return d;
}

// In-place transpose of every N x N plane of a batch (N a multiple of 32): block (pair, plane)
// exchanges tile (i, j) with tile (j, i) through LDS; a diagonal tile is transposed onto itself.
// Tile rows are 32 contiguous values, so loads and stores stay coalesced.
template <typename R>
__global__ __launch_bounds__(256) void k_transpose(cplx<R>* __restrict__ data, int N) {
  __shared__ cplx<R> ta[32][33], tb[32][33];
  const int64_t p = blockIdx.x;  // pair index over the lower triangle i >= j
  int i = (int)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while ((int64_t)(i + 1) * (i + 2) / 2 <= p) ++i;
  while ((int64_t)i * (i + 1) / 2 > p) --i;
  const int j = (int)(p - (int64_t)i * (i + 1) / 2);
  cplx<R>* base = data + (size_t)blockIdx.y * N * N;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    ta[r][tx] = base[(size_t)(i * 32 + r) * N + j * 32 + tx];
    tb[r][tx] = base[(size_t)(j * 32 + r) * N + i * 32 + tx];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    base[(size_t)(j * 32 + r) * N + i * 32 + tx] = ta[tx][r];
    base[(size_t)(i * 32 + r) * N + j * 32 + tx] = tb[tx][r];
  }
}

// Two-level (four-step) column transform for the sizes where one column fills LDS: length
// N = N1 * N2, NB4 adjacent columns per work group (128-byte row segments).
//   pass A (k_cols4_a), per column group and n1 < N1: the N2 values x[n1 + N1 n2] (rows N1 apart)
//     -> N2-point DFT over n2 -> times W_N^(n1 k2) -> row n1 + N1 k2 (the rows it read);
//   pass B (k_cols4_b), per column group and k2 < N2: the N1 contiguous rows N1 k2 + n1 -> N1-point
//     DFT over n1 -> X[k2 + N2 k1], written to row N1 k2 + k1 (again the rows it read).
// Both passes are in place; the output is in digit-transposed row order (perm_row_of), which the
// consumers read through perm_row_of: the invert row pass, and the degridders for a plan whose grids
// are private.  (-1)^(x+y) input and output modulations and the scale use the natural indices.
constexpr int kCols4N1 = 64;
template <typename R>
__host__ __device__ constexpr int cols4_nb() {
  return sizeof(R) == 4 ? 16 : 8;
}
// transforms of one column group per work group: pass A (length N2 = N / 64) and pass B (64)
constexpr int kCols4MA = 4, kCols4MB = 8;

template <typename R, int N, bool INV>
__global__ __launch_bounds__(Shape<N / kCols4N1>::NT * cols4_nb<R>() * kCols4MA) void k_cols4_a(
    FftArgs a, const cplx<R>* __restrict__ tw2g, const cplx<R>* __restrict__ twN) {
  constexpr int N1 = kCols4N1, N2 = N / N1, NBC = cols4_nb<R>(), M = kCols4MA, NB = NBC * M;
  constexpr int NT = Shape<N2>::NT;
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem);
  const cplx<R>* tw = stage_twiddles_at<R, N2>(smem, sizeof(cplx<R>) * (size_t)N2 * NB, tw2g);
  constexpr int CG = N / NBC;  // column groups per plane
  constexpr int GA = N1 / M;   // n1 groups
  const int g = blockIdx.x % GA;
  const int rest = blockIdx.x / GA;
  const int64_t plane = rest / CG;
  const int x0 = (rest % CG) * NBC;
  const int b = threadIdx.x % NB, t = threadIdx.x / NB;
  const int x = x0 + b % NBC, n1 = g * M + b / NBC;
  if (a.colmask && !((a.colmask[plane] >> (x0 >> a.tshift)) & 1u)) return;  // a column band of zeros
  cplx<R>* col = reinterpret_cast<cplx<R>*>(a.data) + plane * (int64_t)N * N + x;
  fft_pass<R, N2, NB, NT, INV, LAY_COL>(
      lds, b, t, tw,
      [&](int, int, int n2) {
        const int y = n1 + N1 * n2;
        cplx<R> v = col[(size_t)y * N];
        if (a.in_mod && ((y + x) & 1)) v = cmk<R>(-v.x, -v.y);
        return v;
      },
      [&](int, int, int k2, cplx<R> v) {
        cplx<R> w = twN[n1 * k2];  // n1 k2 < N
        if (INV) w.y = -w.y;
        col[(size_t)(n1 + N1 * k2) * N] = cmul(v, w);
      });
}

template <typename R, int N, bool INV>
__global__ __launch_bounds__(Shape<kCols4N1>::NT * cols4_nb<R>() * kCols4MB) void k_cols4_b(
    FftArgs a, const cplx<R>* __restrict__ tw1g) {
  constexpr int N1 = kCols4N1, N2 = N / N1, NBC = cols4_nb<R>(), M = kCols4MB, NB = NBC * M;
  constexpr int NT = Shape<N1>::NT;
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem);
  const cplx<R>* tw = stage_twiddles_at<R, N1>(smem, sizeof(cplx<R>) * (size_t)N1 * NB, tw1g);
  constexpr int CG = N / NBC;
  constexpr int GB = N2 / M;  // k2 groups
  const int g = blockIdx.x % GB;
  const int rest = blockIdx.x / GB;
  const int64_t plane = rest / CG;
  const int x0 = (rest % CG) * NBC;
  const int b = threadIdx.x % NB, t = threadIdx.x / NB;
  const int x = x0 + b % NBC, k2 = g * M + b / NBC;
  if (a.colmask && !((a.colmask[plane] >> (x0 >> a.tshift)) & 1u)) return;  // a column band of zeros
  cplx<R>* col = reinterpret_cast<cplx<R>*>(a.data) + plane * (int64_t)N * N + x + (size_t)(N1 * k2) * N;
  fft_pass<R, N1, NB, NT, INV, LAY_COL>(
      lds, b, t, tw, [&](int, int, int n1) { return col[(size_t)n1 * N]; },
      [&](int, int, int k1, cplx<R> v) {
        const int vrow = k2 + N2 * k1;  // natural index of this output
        R sc = (R)a.scale;
        if (a.out_mod && ((vrow + x) & 1)) sc = -sc;
        col[(size_t)k1 * N] = cmk<R>(v.x * sc, v.y * sc);
      });
}

// Rows per plane of the two-level column transform's digit-transposed order, or 0 where the
// columns are transformed in natural order (one-pass columns) for this size and precision.
template <typename R, int N>
constexpr int cols4_n1() {
  return Shape<N>::template cols_by_transpose<R>() ? kCols4N1 : 0;
}

template <typename R, int N, bool INV>
static int run_cols4(FftArgs c, hipStream_t s) {
  constexpr int N1 = kCols4N1, N2 = N / N1, NB = cols4_nb<R>();
  static_assert(N2 >= 64, "two-level column transform for N >= 4096");
  const cplx<R>* tw1 = reinterpret_cast<const cplx<R>*>(twiddles(N1, sizeof(R) == 4 ? 32 : 64));
  const cplx<R>* tw2 = reinterpret_cast<const cplx<R>*>(twiddles(N2, sizeof(R) == 4 ? 32 : 64));
  const cplx<R>* twN = reinterpret_cast<const cplx<R>*>(twiddles(N, sizeof(R) == 4 ? 32 : 64));
  if (!tw1 || !tw2 || !twN) {
    set_error("twiddle table allocation failed");
    return ARLMI_E_HIP;
  }
  if (c.tilemask) {
    set_error("sparse grids are not supported at N = %d in this precision", N);
    return ARLMI_E_UNSUPPORTED;
  }
  constexpr int MA = kCols4MA, MB = kCols4MB;
  const int64_t groups = (int64_t)c.nbatch * (N / NB);
  const size_t sha = sizeof(cplx<R>) * ((size_t)N2 * NB * MA + tw_lds_elems<R, N2>()),
               shb = sizeof(cplx<R>) * ((size_t)N1 * NB * MB + tw_lds_elems<R, N1>());
  ARLMI_LAUNCH("k_cols4_a", (k_cols4_a<R, N, INV>), dim3((unsigned)(groups * (N1 / MA))),
               dim3(Shape<N2>::NT * NB * MA), sha, s, c, tw2, twN);
  ARLMI_LAUNCH_CHECK();
  ARLMI_LAUNCH("k_cols4_b", (k_cols4_b<R, N, INV>), dim3((unsigned)(groups * (N2 / MB))),
               dim3(Shape<N1>::NT * NB * MB), shb, s, c, tw1);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

// The column transforms of c.nbatch planes in place (modulation and scale as c says): NB adjacent
// columns per block, or, where two columns and the twiddles exceed LDS, transpose + row pass +
// transpose ((-1)^(x+y) is symmetric, so the row pass applies the same modulation).
template <typename R, int N, bool INV>
static int run_cols(FftArgs c, const cplx<R>* tw, hipStream_t s) {
  constexpr int NT = Shape<N>::NT;
  if constexpr (Shape<N>::template cols_by_transpose<R>()) {
    if (c.tilemask) {
      set_error("sparse grids are not supported at N = %d in this precision", N);
      return ARLMI_E_UNSUPPORTED;
    }
    constexpr int RNB = Shape<N>::ROW_NB;
    const size_t rsh = sizeof(cplx<R>) * (lds_elems<LAY_ROWPAD, N>(RNB) + Shape<N>::template tw_lds<R>());
    const int64_t nt = N / 32;
    const dim3 tg((unsigned)(nt * (nt + 1) / 2), (unsigned)c.nbatch);
    ARLMI_LAUNCH("k_transpose", k_transpose<R>, tg, dim3(256), 0, s, (cplx<R>*)c.data, N);
    ARLMI_LAUNCH_CHECK();
    ARLMI_LAUNCH("k_rows_plain", (k_rows_plain<R, N, INV>), dim3((unsigned)((int64_t)c.nbatch * N / RNB)),
                 dim3(NT * RNB), rsh, s, c, tw);
    ARLMI_LAUNCH_CHECK();
    ARLMI_LAUNCH("k_transpose", k_transpose<R>, tg, dim3(256), 0, s, (cplx<R>*)c.data, N);
    ARLMI_LAUNCH_CHECK();
  } else {
    if constexpr (Shape<N>::template split_nb<R>() > 0) {
      constexpr int SNB = Shape<N>::template split_nb<R>();
      const cplx<R>* twh = reinterpret_cast<const cplx<R>*>(twiddles(N / 2, sizeof(R) == 4 ? 32 : 64));
      if (!twh) {
        set_error("twiddle table allocation failed");
        return ARLMI_E_HIP;
      }
      const size_t sh = sizeof(cplx<R>) * ((size_t)(N / 2) * SNB + tw_lds_elems<R, N / 2>());
      ARLMI_LAUNCH("k_cols_split", (k_cols_split<R, N, INV, SNB>), dim3((unsigned)((int64_t)c.nbatch * N / SNB)),
                   dim3(Shape<N / 2>::NT * SNB), sh, s, c, twh, tw);
      ARLMI_LAUNCH_CHECK();
      return ARLMI_OK;
    }
    constexpr int CNB = Shape<N>::template col_nb<R>();
    const size_t csh = sizeof(cplx<R>) * ((size_t)N * CNB + tw_lds_elems<R, N>());
    const unsigned cblocks = (unsigned)((int64_t)c.nbatch * N / CNB);
    ARLMI_LAUNCH("k_cols_plain", (k_cols_plain<R, N, INV>), dim3(cblocks), dim3(NT * CNB), csh, s, c, tw);
    ARLMI_LAUNCH_CHECK();
  }
  return ARLMI_OK;
}

template <typename R, int N>
static int run_fft2d(FftArgs a, int direction, hipStream_t s) {
  const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(twiddles(N, sizeof(R) == 4 ? 32 : 64));
  if (!tw) {
    set_error("twiddle table allocation failed");
    return ARLMI_E_HIP;
  }
  constexpr int NT = Shape<N>::NT, RNB = Shape<N>::ROW_NB;
  const size_t rsh = sizeof(cplx<R>) * (lds_elems<LAY_ROWPAD, N>(RNB) + Shape<N>::template tw_lds<R>());
  const unsigned rblocks = (unsigned)((int64_t)a.nbatch * N / RNB);
  FftArgs r = a, c = a;
  r.in_mod = 1;
  r.out_mod = 0;
  r.scale = 1.0;
  c.in_mod = 0;
  c.out_mod = 1;
  c.scale = direction < 0 ? 1.0 / ((double)N * (double)N) : 1.0;
  if (direction < 0) {
    ARLMI_LAUNCH("k_rows_plain", (k_rows_plain<R, N, true>), dim3(rblocks), dim3(NT * RNB), rsh, s, r, tw);
    ARLMI_LAUNCH_CHECK();
    return run_cols<R, N, true>(c, tw, s);
  }
  ARLMI_LAUNCH("k_rows_plain", (k_rows_plain<R, N, false>), dim3(rblocks), dim3(NT * RNB), rsh, s, r, tw);
  ARLMI_LAUNCH_CHECK();
  return run_cols<R, N, false>(c, tw, s);
}

// Plane groups for the fused row passes: enough blocks to fill the GPU several times over.
static void plane_groups(int rowblocks, int nplanes, int* ppg, int* groups) {
  const int want = std::max(1, std::min(nplanes, std::min(16, (8192 + rowblocks - 1) / rowblocks)));
  *ppg = (nplanes + want - 1) / want;
  *groups = (nplanes + *ppg - 1) / *ppg;
}

template <typename R, int N>
static int run_grids_to_image(FftArgs a, hipStream_t s) {
  const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(twiddles(N, sizeof(R) == 4 ? 32 : 64));
  if (!tw) {
    set_error("twiddle table allocation failed");
    return ARLMI_E_HIP;
  }
  constexpr int NT = Shape<N>::NT, RNB = Shape<N>::ROW_NB;
  const size_t ish = sizeof(cplx<R>) * (lds_elems<inv_lay<R>(), N>(RNB) + Shape<N>::template tw_lds<R>());
  FftArgs c = a;
  c.in_mod = 1;
  c.out_mod = 0;
  c.scale = 1.0;
  c.nbatch = a.nplanes * a.nchan;
  a.perm_n1 = 0;
  if constexpr (Shape<N>::template cols_by_transpose<R>()) {
    // two-level columns; the row pass reads their digit-transposed row order
    if (int rc = run_cols4<R, N, true>(c, s)) return rc;
    a.perm_n1 = kCols4N1;
  } else {
    if (int rc = run_cols<R, N, true>(c, tw, s)) return rc;
  }
  const int rowblocks = a.nchan * (N / RNB);
  int G = 1;
  plane_groups(rowblocks, a.nplanes, &a.ppg, &G);
  if (G > 1 && !a.part) G = 1, a.ppg = a.nplanes;
  if (a.screens)
    ARLMI_LAUNCH("k_rows_invert", (k_rows_invert<R, N, true>), dim3((unsigned)rowblocks, (unsigned)G), dim3(NT * RNB),
                 ish, s, a, tw);
  else
    ARLMI_LAUNCH("k_rows_invert", (k_rows_invert<R, N, false>), dim3((unsigned)rowblocks, (unsigned)G),
                 dim3(NT * RNB), ish, s, a, tw);
  ARLMI_LAUNCH_CHECK();
  if (G > 1) {
    ARLMI_LAUNCH("k_invert_finish", k_invert_finish<R>, dim3((unsigned)(((int64_t)N * N + 255) / 256), (unsigned)a.nchan),
                 dim3(256), 0, s, a, N, G);
    ARLMI_LAUNCH_CHECK();
  }
  return ARLMI_OK;
}

template <typename R, int N>
static int run_image_to_grids(FftArgs a, hipStream_t s) {
  const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(twiddles(N, sizeof(R) == 4 ? 32 : 64));
  if (!tw) {
    set_error("twiddle table allocation failed");
    return ARLMI_E_HIP;
  }
  constexpr int FNT = Shape<N>::FNT, FNB = Shape<N>::FNB;
  const size_t fsh = sizeof(cplx<R>) * (lds_elems<LAY_ROWPAD, N>(FNB) + Shape<N>::template tw_lds<R>());
  const int rowblocks = a.nchan * (N / FNB);
  int G = 1;
  plane_groups(rowblocks, a.nplanes, &a.ppg, &G);
  ARLMI_LAUNCH("k_rows_predict", (k_rows_predict<R, N>), dim3((unsigned)rowblocks, (unsigned)G), dim3(FNT * FNB), fsh,
               s, a, tw);
  ARLMI_LAUNCH_CHECK();
  FftArgs c = a;
  c.in_mod = 0;
  c.out_mod = 1;
  c.scale = 1.0;
  c.nbatch = a.nplanes * a.nchan;
  if constexpr (Shape<N>::template cols_by_transpose<R>()) {
    // grids private to a plan: two-level columns, left in digit-transposed row order for the
    // plan's degridder (a.perm_n1 set by the caller); otherwise natural order through transposes
    if (a.perm_n1) return run_cols4<R, N, false>(c, s);
  }
  return run_cols<R, N, false>(c, tw, s);
}

enum class Op { FFT, TO_IMAGE, TO_GRIDS };

}  // namespace arlmi

// Row order of the grids arlmi_image_to_grids leaves for a plan with private grids: kCols4N1 where
// the two-level column transform is used (size and precision), 0 for natural order.
int arlmi::grid_perm_n1(int n, int precision) {
  const bool p32 = precision == 32;
  switch (n) {
    case 4096: return p32 ? cols4_n1<float, 4096>() : cols4_n1<double, 4096>();
    case 8192: return p32 ? cols4_n1<float, 8192>() : cols4_n1<double, 8192>();
    default: return 0;
  }
}

namespace arlmi {

template <typename R>
static int dispatch_n(Op op, int n, FftArgs a, int direction, hipStream_t s) {
#define ARLMI_CASE(NN)                                                  \
  case NN:                                                              \
    if constexpr (Shape<NN>::template fits<R>()) {                      \
      if (op == Op::FFT) return run_fft2d<R, NN>(a, direction, s);      \
      if (op == Op::TO_IMAGE) return run_grids_to_image<R, NN>(a, s);   \
      return run_image_to_grids<R, NN>(a, s);                           \
    }                                                                   \
    break;
  switch (n) {
    ARLMI_CASE(64)
    ARLMI_CASE(128)
    ARLMI_CASE(256)
    ARLMI_CASE(512)
    ARLMI_CASE(1024)
    ARLMI_CASE(2048)
    ARLMI_CASE(4096)
    ARLMI_CASE(8192)
    default:
      break;
  }
#undef ARLMI_CASE
  set_error("FFT size %d not supported (powers of two 64..8192)", n);
  return ARLMI_E_UNSUPPORTED;
}

static int prec_ok(int precision) {
  if (precision != 32 && precision != 64) {
    set_error("precision must be 32 or 64");
    return ARLMI_E_ARG;
  }
  return ARLMI_OK;
}

}  // namespace arlmi

using namespace arlmi;

extern "C" {

int arlmi_fft2d(void* d_data, int32_t npixel, int64_t nbatch, int direction, int precision, void* stream) {
  if (int rc = prec_ok(precision)) return rc;
  if (!d_data || nbatch < 0 || (direction != 1 && direction != -1)) {
    set_error("arlmi_fft2d: bad argument");
    return ARLMI_E_ARG;
  }
  if (nbatch == 0) return ARLMI_OK;
  FftArgs a = {};
  a.data = d_data;
  a.nbatch = (int)nbatch;
  return precision == 32 ? dispatch_n<float>(Op::FFT, npixel, a, direction, as_stream(stream))
                         : dispatch_n<double>(Op::FFT, npixel, a, direction, as_stream(stream));
}

// FftArgs for a caller's column-band mask (arlmi_plan_use_colmask): column groups of bands whose
// bit is clear are skipped by the column passes, read as zero by the invert row pass and not
// stored by the predict row pass.
static void set_band_mask(const arlmi_plan* p, FftArgs* a) {
  const int bs = band_shift(p->g);
  a->colmask = p->ext_colmask;
  a->tilemask = nullptr;
  a->zeros = p->d_zeros;
  a->zero_fill = 0;
  a->tile = 1 << bs;
  a->tshift = bs;
  a->ntx = p->g.n >> bs;
}

int arlmi_grids_to_image(const arlmi_plan* p, void* d_grids, const double* d_sumwt, const double* d_gcf, int normalize,
                         void* d_img_re, void* d_img_im, int precision, void* stream) {
  if (int rc = prec_ok(precision)) return rc;
  if (!p || !d_grids || !d_sumwt || !d_gcf || !d_img_re) {
    set_error("arlmi_grids_to_image: NULL argument");
    return ARLMI_E_ARG;
  }
  if (normalize < 0 || normalize > 2) {
    set_error("arlmi_grids_to_image: normalize must be 0, 1 or 2");
    return ARLMI_E_ARG;
  }
  if (d_img_im && p->g.mode == ARLMI_MODE_WSTACK) {
    set_error("the imaginary image is only defined without w-stacking");
    return ARLMI_E_ARG;
  }
  FftArgs a = {};
  a.data = d_grids;
  a.nplanes = p->g.nplanes;
  a.nchan = p->g.nchan;
  a.gcf = d_gcf;
  a.plane_w = p->d_plane_w;
  a.sumwt = d_sumwt;
  a.img_re = d_img_re;
  a.img_im = d_img_im;
  a.normalize = normalize;
  a.wstack = p->g.mode == ARLMI_MODE_WSTACK;
  a.screens = a.wstack;
  a.fov = (double)p->g.n * p->g.cellsize;
  if (p->flags & ARLMI_FLAG_SPARSE_GRIDS) {
    a.colmask = p->d_colmask;
    a.tilemask = p->d_tilemask;
    a.zeros = p->d_zeros;
    a.zero_fill = 1;
    a.tile = p->g.tile;
    a.tshift = __builtin_ctz((unsigned)p->g.tile);
    a.ntx = p->g.ntx;
  } else if (p->ext_colmask) {  // the caller's column bands: the others hold zeros
    set_band_mask(p, &a);
  }
  // partial images of the plane groups live with the plan (grown on demand)
  const int n = p->g.n, nt = n >= 16 ? n / 16 : 1, rnb = nt >= 256 ? 1 : 256 / nt;  // Shape<n>::NT, ROW_NB
  int ppg = 1, G = 1;
  plane_groups(a.nchan * (n / rnb), a.nplanes, &ppg, &G);
  if (G > 1) {
    arlmi_plan* mp = const_cast<arlmi_plan*>(p);
    const int64_t need = (int64_t)G * (a.screens ? 1 : 2) * a.nchan * (int64_t)n * n * (precision / 8);
    if (need > mp->cap_img_part_bytes || !mp->d_img_part) {
      if (mp->d_img_part) ARLMI_HIP(hipFree(mp->d_img_part));
      mp->d_img_part = nullptr;
      mp->cap_img_part_bytes = 0;
      ARLMI_HIP(hipMalloc(&mp->d_img_part, (size_t)need));
      mp->cap_img_part_bytes = need;
    }
    a.part = mp->d_img_part;
  }
  return precision == 32 ? dispatch_n<float>(Op::TO_IMAGE, p->g.n, a, 0, as_stream(stream))
                         : dispatch_n<double>(Op::TO_IMAGE, p->g.n, a, 0, as_stream(stream));
}

int arlmi_image_to_grids(const arlmi_plan* p, const void* d_model, const double* d_gcf, void* d_grids, int precision,
                         void* stream) {
  if (int rc = prec_ok(precision)) return rc;
  if (!p || !d_model || !d_gcf || !d_grids) {
    set_error("arlmi_image_to_grids: NULL argument");
    return ARLMI_E_ARG;
  }
  FftArgs a = {};
  a.data = d_grids;
  a.nplanes = p->g.nplanes;
  a.nchan = p->g.nchan;
  a.model = d_model;
  a.gcf = d_gcf;
  a.plane_w = p->d_plane_w;
  a.wstack = p->g.mode == ARLMI_MODE_WSTACK;
  a.screens = a.wstack;
  a.fov = (double)p->g.n * p->g.cellsize;
  if (p->flags & ARLMI_FLAG_SPARSE_GRIDS) {  // only the tile columns the degridder reads
    a.colmask = p->d_colmask;
    a.tilemask = p->d_tilemask;
    a.zeros = p->d_zeros;
    a.zero_fill = 0;
    a.tile = p->g.tile;
    a.tshift = __builtin_ctz((unsigned)p->g.tile);
    a.ntx = p->g.ntx;
  } else if (p->ext_colmask) {  // only the column bands the caller's degridding reads
    set_band_mask(p, &a);
  }
  // grids only this plan's degridder reads may keep the two-level column transform's row order
  if (p->flags & ARLMI_FLAG_PRIVATE_GRIDS) a.perm_n1 = grid_perm_n1(p->g.n, precision);
  return precision == 32 ? dispatch_n<float>(Op::TO_GRIDS, p->g.n, a, 0, as_stream(stream))
                         : dispatch_n<double>(Op::TO_GRIDS, p->g.n, a, 0, as_stream(stream));
}

int arlmi_w_screen(int32_t npixel, double cellsize, double w, void* d_out, int precision, void* stream) {
  if (int rc = prec_ok(precision)) return rc;
  if (!d_out || npixel <= 0 || (npixel & 1)) {
    set_error("arlmi_w_screen: bad argument");
    return ARLMI_E_ARG;
  }
  const int64_t tot = (int64_t)npixel * npixel;
  const unsigned blocks = (unsigned)((tot + 255) / 256);
  const double fov = (double)npixel * cellsize;
  if (precision == 32)
    ARLMI_LAUNCH("k_w_screen", k_w_screen<float>, dim3(blocks), dim3(256), 0, as_stream(stream), npixel, fov, w,
                       (float2*)d_out);
  else
    ARLMI_LAUNCH("k_w_screen", k_w_screen<double>, dim3(blocks), dim3(256), 0, as_stream(stream), npixel, fov, w,
                       (double2*)d_out);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

int arlmi_dft(const double* d_uvw, const int32_t* d_chan, int64_t nrows, const double* d_lmflux, int32_t ncomp,
              int32_t nchan, void* d_vis, int precision, void* stream) {
  if (int rc = prec_ok(precision)) return rc;
  if (nrows < 0 || ncomp < 0 || nchan <= 0 || (nrows > 0 && (!d_uvw || !d_vis)) || (ncomp > 0 && !d_lmflux)) {
    set_error("arlmi_dft: bad argument");
    return ARLMI_E_ARG;
  }
  const size_t sh = sizeof(double) * (size_t)ncomp * (3 + nchan);
  if (sh > 64 * 1024) {
    set_error("too many components for one call (%d)", ncomp);
    return ARLMI_E_UNSUPPORTED;
  }
  if (nrows == 0 || ncomp == 0) return ARLMI_OK;
  const unsigned blocks = (unsigned)((nrows + 255) / 256);
  if (precision == 32)
    ARLMI_LAUNCH("k_dft", k_dft<float>, dim3(blocks), dim3(256), sh, as_stream(stream), d_uvw, d_chan, nrows, d_lmflux,
                       ncomp, nchan, (float2*)d_vis);
  else
    ARLMI_LAUNCH("k_dft", k_dft<double>, dim3(blocks), dim3(256), sh, as_stream(stream), d_uvw, d_chan, nrows, d_lmflux,
                       ncomp, nchan, (double2*)d_vis);
  ARLMI_LAUNCH_CHECK();
  return ARLMI_OK;
}

}  // extern "C"
