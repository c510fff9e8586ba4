/*
 * arlmi.h - C ABI of libarlmi.so, the MI355X (gfx950) implementation of ARL's
 * Fourier-processing hot path: convolutional gridding / degridding with
 * oversampled kernels, w-stacking phase screens, the shifted 2-D FFT and the
 * gridding correction.
 *
 * Conventions
 *   - Every pointer named d_* is DEVICE memory owned by the caller; the library
 *     keeps no reference to it after the call returns.
 *   - Work is enqueued on `stream` (a hipStream_t, NULL = default stream) and is
 *     stream-ordered; calls return once the work is enqueued unless stated.
 *   - Return value: ARLMI_OK (0) or a negative ARLMI_E_* code; the message of the
 *     last failure on the calling thread is returned by arlmi_last_error().
 *   - precision: 32 -> complex64 visibilities / grids, float32 weights and images;
 *                64 -> complex128 / float64 throughout.  uvw, frequencies, w-slab
 *     centres, grid-correction images and sums of weights are always float64.
 *   - Complex numbers are interleaved (re, im) pairs, as numpy/torch store them.
 *   - Graph capture: once a plan has served one call in a precision, arlmi_grid, arlmi_degrid,
 *     arlmi_grids_to_image and arlmi_image_to_grids on it allocate nothing, record no memset
 *     and never synchronise, so a sequence of them can be captured into a HIP graph and
 *     replayed (plan creation and update synchronise and cannot be captured).
 *
 * The ARL interface each entry point replaces is cited per function
 * (file:line in SKA-ScienceDataProcessor/algorithm-reference-library).  The
 * deprecated ARL C FFI (deprecated_code/ffiwrappers/include/arlwrap.h:112-113)
 * exposes the same operation pair as arl_invert_2d / arl_predict_2d; its
 * global-error-flag convention (arlwrap.py:41-47) is what arlmi_last_error mirrors.
 */
#ifndef ARLMI_H
#define ARLMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ARLMI_OK 0
#define ARLMI_E_ARG (-1)          /* invalid argument / shape */
#define ARLMI_E_UV_OFF_GRID (-2)  /* gridding.py:43-46 (and footprint over the edge, :128, :264) */
#define ARLMI_E_W_RANGE (-3)      /* gridding.py:64-65, 75-76 */
#define ARLMI_E_FREQ_MAP (-4)     /* gridding.py:83-84 */
#define ARLMI_E_ODD_SUPPORT (-5)  /* gridding.py:125-128 needs an even support */
#define ARLMI_E_SUBSAMPLE (-6)    /* pu_offset == oversampling (gridding.py:53-54 -> IndexError) */
#define ARLMI_E_HIP (-7)          /* a HIP runtime call failed */
#define ARLMI_E_UNSUPPORTED (-8)  /* size / layout not handled by this build */

#define ARLMI_MODE_2D 0      /* planes chosen by rounding w onto the grid's WW axis (nw planes, summed) */
#define ARLMI_MODE_WSTACK 1  /* planes are w slabs |w - box_k| < wslice/2, w shifted by the slab mean */

#define ARLMI_FLAG_DROP_OFF_GRID 1u /* drop (and count) rows whose kernel footprint leaves the grid */
/* Sparse grids (no counterpart in ARL, whose grids are dense numpy arrays): arlmi_grid leaves the
 * T x T tiles that no row touches unwritten; arlmi_grids_to_image reads only touched tiles (the
 * others count as zero) and skips tile columns with no touched tile; arlmi_image_to_grids writes
 * only the tile columns that hold touched tiles, which are all that arlmi_degrid reads.  Grids
 * of such a plan are meant to pass between these calls, not to be read by the caller.  Plans
 * with more than 32 tiles per side ignore the flag (dense grids). */
#define ARLMI_FLAG_SPARSE_GRIDS 2u
/* Dense grids whose untouched tiles arlmi_grid / arlmi_grid_add leave as they are (no counterpart
 * in ARL): the gridding work items cover only tiles that rows touch.  For chunks added into grids
 * the caller zeroed once (arlmi_grid_add); the FFT-side calls read and write the grids densely. */
#define ARLMI_FLAG_OCCUPIED_TILES 4u
/* The plan's grids pass only between its own calls (no counterpart in ARL): arlmi_image_to_grids
 * may then leave the grid rows in the order of the library's two-level column transform (at 8192
 * in single and 4096 / 8192 in double precision, where one grid column and the twiddles fill the
 * LDS), which arlmi_degrid on the same plan reads.  The grids of such a plan are not in ARL's
 * layout between those two calls. */
#define ARLMI_FLAG_PRIVATE_GRIDS 8u
/* Bitwise-reproducible gridding (no counterpart in ARL, whose numpy loop adds in one fixed order):
 * arlmi_grid / arlmi_grid_add accumulate every grid value and sum of weights as 64-bit integers,
 * so that repeated calls on the same inputs give identical grids and sums of weights whatever the
 * order in which the GPU issues the adds (the default float64 LDS adds vary with it in the last
 * bits).  Each call bounds every grid value of a T x T tile by B_t = Kmax * M * n_t (Kmax the
 * largest tap magnitude of the kernel, M the largest |Re vis w| + |Im vis w| over the rows, n_t the
 * tile's entries) and rounds each tap's product once to a multiple of 2^-q_t with B_t 2^q_t < 2^60:
 * the absolute rounding per tap is at most 2^-60 B_t.  The sums of weights use their own q from
 * sum |w|.  Not batched (arlmi_batch_create returns ARLMI_E_UNSUPPORTED); non-finite visibilities
 * or weights give unspecified grids. */
#define ARLMI_FLAG_DETERMINISTIC 16u

#define ARLMI_KERNEL_SEPARABLE 0 /* real separable table k[os][S]: CF[yo][xo][a][b] = k[yo][a]*k[xo][b] */
#define ARLMI_KERNEL_TABLE 1     /* full complex table CF[nz][os][os][S][S] (ARL layout, chan 0, pol 0) */

/* Geometry of one imaging problem, carrying the WCS numbers ARL stores
 * (griddata/operations.py:91-145, griddata/convolution_functions.py:62-147,
 * processing_library/image/operations.py:33-69).  Fill it with arlmi_geom_init
 * semantics: every double is computed on the host exactly as ARL computes it. */
typedef struct arlmi_geom {
  int32_t npixel;       /* N: image side = grid side (power of two, 64..8192) */
  int32_t nchan;        /* image channels */
  int32_t support;      /* S, even */
  int32_t oversampling; /* os */
  int32_t nz;           /* kernel w planes (1 for PSWF) */
  int32_t nplanes;      /* MODE_WSTACK: number of w slabs; MODE_2D: grid WW planes nw */
  int32_t mode;         /* ARLMI_MODE_* */
  int32_t tile;         /* LDS tile side T (32 or 64); 0 = library default */
  double du, dv;        /* UU, VV cdelt = 1/(N*d2r*cdelt_deg) (wavelengths per pixel) */
  double ddu, ddv;      /* DUU, DVV cdelt = du/os, dv/os */
  double crpix_uv;      /* N//2 + 1 */
  double crpix_d;       /* os//2 + 1 */
  double zstep;         /* kernel WW cdelt */
  double crpix_z;       /* nz//2 + 1 */
  double wstep_grid;    /* grid WW cdelt (MODE_2D) */
  double crpix_wg;      /* nw//2 + 1 (MODE_2D) */
  double freq0, dfreq;  /* image FREQ crval and cdelt (crpix 1) */
  double cellsize;      /* abs(cdelt0_deg)*pi/180, the w-screen pixel size (image/operations.py:389) */
} arlmi_geom;

/* Counts describing a plan (all host values). */
typedef struct arlmi_plan_stats {
  int64_t nrows;        /* rows given */
  int64_t nentries;     /* (row, plane, tile) gridding entries */
  int64_t ngrid_items;  /* gridding work items (tile chunks) */
  int64_t ndegrid_items;
  int64_t nmember;      /* (row, plane) memberships */
  int64_t ndropped;     /* rows dropped by ARLMI_FLAG_DROP_OFF_GRID */
  int64_t norphans;     /* rows in no plane (predicted as 0) */
  int32_t ntiles;       /* planes * nchan * (N/T)^2 */
  int32_t tile;
} arlmi_plan_stats;

/* Convolution kernel description. */
typedef struct arlmi_kernel {
  int32_t kind;         /* ARLMI_KERNEL_* */
  int32_t nz, oversampling, support;
  const void* d_table;  /* SEPARABLE: real [os][S]; TABLE: complex [nz][os][os][S][S]; precision of the call */
} arlmi_kernel;

typedef struct arlmi_plan arlmi_plan; /* opaque; owns its device buffers */

int arlmi_version(void);
const char* arlmi_last_error(void);

/* max |w| over rows: numpy.max(numpy.abs(vis.w)) in vis_wslice_iter
 * (processing_components/visibility/iterators.py:99).  Synchronises `stream`. */
int arlmi_wmax(const double* d_uvw, int64_t nrows, void* stream, double* wmax_out);

/* Build the gridding plan: per-row convolution mapping (gridding.py:24-96), w-slab
 * membership (iterators.py:91-109: slab centres h_boxes[P] and width wslice from the host, or,
 * with h_boxes NULL, ARL's default boxes = numpy.linspace(-max|w|, max|w|, P) and wslice =
 * boxes[1] - boxes[0] (2 max|w| for P = 1) computed on the device with numpy's operations),
 * slab mean w (wstack_single.py:43, :83),
 * and the tile-sorted entry lists the grid/degrid kernels walk.  d_freq may be NULL
 * (all rows in image channel 0).  Synchronises `stream`.  Returns an ARLMI_E_* code if
 * a row violates one of ARL's asserts and ARLMI_FLAG_DROP_OFF_GRID does not cover it. */
int arlmi_plan_create(const arlmi_geom* g, const double* d_uvw, const double* d_freq, int64_t nrows,
                      const double* h_boxes, double wslice, uint32_t flags, void* stream,
                      arlmi_plan** plan_out);
/* Rebuild an existing plan for new uvw (same geometry), reusing its buffers when they fit. */
int arlmi_plan_update(arlmi_plan* plan, const double* d_uvw, const double* d_freq, int64_t nrows,
                      const double* h_boxes, double wslice, void* stream);
/* Work partition (no counterpart in ARL): gridding / degridding work items hold at most `chunk`
 * entries of one tile (default 4096).  Takes effect at the next arlmi_plan_update. */
int arlmi_plan_set_chunk(arlmi_plan* plan, int32_t chunk);
int arlmi_plan_stats_get(const arlmi_plan* plan, arlmi_plan_stats* out);
/* Measurement aid (no counterpart in ARL): the number of (real, imaginary) LDS add pairs the
 * separable gridder of kernel support `support` issues on this plan.  Plans whose entries are
 * grouped by footprint origin (one kernel w plane, at most 64 taps, not deterministic) add each
 * run of one origin once (runs cut at 32 entries and at every 64 entries of a wavefront); other
 * plans add every entry (the entry count is returned).  Synchronises `stream`. */
int arlmi_plan_grid_runs(const arlmi_plan* plan, int32_t support, int64_t* runs_out, void* stream);
/* Streaming (one vis set in chunks of rows, as ARL's list workflows see one Visibility whose rows
 * arrive in pieces; imaging_serial.py:166-197 sums the per-slab images of all its rows).  The slab
 * rule and means are those of the whole set (iterators.py:99-106 takes max|w| over every row,
 * wstack_single.py:43, :83 the mean over every row of a slab), so a first pass gathers them:
 * arlmi_wmax per chunk (the host keeps the largest), then arlmi_wslab_sums per chunk with the slab
 * centres: h_sum[k] += sum of w and h_count[k] += count of the chunk's rows in slab k
 * (|w - h_boxes[k]| < wslice / 2), summed per chunk in a fixed order.  Synchronises `stream`. */
int arlmi_wslab_sums(const double* d_uvw, int64_t nrows, const double* h_boxes, double wslice, int32_t nslabs,
                     double* h_sum, int64_t* h_count, void* stream);
/* Fix the slab means (h_plane_w[nplanes], e.g. sum / count over all chunks) that the following
 * arlmi_plan_update calls of a w-stack plan use for the slab-mean w shift and the screens, instead
 * of the means of the rows each update is given.  n = 0 (or NULL) returns to the rows' own means. */
int arlmi_plan_fix_plane_w(arlmi_plan* plan, const double* h_plane_w, int32_t n);

/* Column-band occupancy for grids that several plans' rows add into (arlmi_grid_add; no
 * counterpart in ARL, whose grids are dense numpy arrays): d_colmask[pc] (one uint32 per plane and
 * image channel, pc = plane * nchan + chan, zeroed by the caller) gets bit b set for every band b of
 * grid columns that holds an entry of this plan; a band is 2^band_shift columns wide, the tile side
 * or N / 32 when a plane has more than 32 tile columns (*band_shift_out, if not NULL).  Stream-
 * ordered. */
int arlmi_plan_or_colmask(const arlmi_plan* plan, uint32_t* d_colmask, int32_t* band_shift_out, void* stream);
/* The transforms of this plan (arlmi_grids_to_image, arlmi_image_to_grids) then treat the column
 * bands whose bit is clear as zero: the column passes skip them, the invert row pass reads them as
 * zero and the predict row pass does not store them (the degridding of the plans whose entries
 * the mask covers reads none of them).  For grids whose other bands hold zeros (invert) or are not
 * read (predict).  NULL returns to dense transforms.  Not for plans with ARLMI_FLAG_SPARSE_GRIDS.
 * The mask is the caller's and must stay valid while the plan uses it. */
int arlmi_plan_use_colmask(arlmi_plan* plan, const uint32_t* d_colmask);

/* Host copy of a w-stack plan's slab centres and, after them, the slab width (nplanes + 1
 * values; iterators.py:99-106). */
int arlmi_plan_slab_boxes(const arlmi_plan* plan, double* h_out, int32_t n);
/* Host copy of the per-plane mean w (numpy.average(vis.w) per slab; 0 in MODE_2D). */
int arlmi_plan_plane_w(const arlmi_plan* plan, double* h_out, int32_t n);
void arlmi_plan_destroy(arlmi_plan* plan);

/* convolution_mapping (gridding.py:24-96) for every row with 2-D semantics: d_out int32
 * [nrows][8] = {pu, pu_offset, pv, pv_offset, pwg, pwc, pfreq, err_bits}; err_bits: 1 uv or
 * footprint off the grid, 2 w plane out of range, 4 frequency out of range, 8 sub-sample index
 * equal to the oversampling.  A negative floor() sub-sample is wrapped like a numpy index. */
int arlmi_convolution_mapping(const arlmi_geom* g, const double* d_uvw, const double* d_freq, int64_t nrows,
                              int32_t* d_out, void* stream);

/* grid_visibility_to_griddata (gridding.py:99-133) for every plane of the plan:
 * d_grids[plane][chan][N][N] (complex) = sum conj(CF) * vis * imaging_weight, and
 * d_sumwt[plane][chan] (float64) = sum imaging_weight.  dopsf replaces vis by 1
 * (imaging/base.py:162-163).  Both outputs are overwritten. */
int arlmi_grid(const arlmi_plan* plan, const void* d_vis, const void* d_weight, int dopsf,
               const arlmi_kernel* kernel, void* d_grids, double* d_sumwt, int precision, void* stream);

/* arlmi_grid for the rows of one chunk of a streamed vis set: adds the chunk's gridded
 * visibilities into d_grids and its sums of weights into d_sumwt instead of overwriting them
 * (the caller zeroes both before the first chunk), so that the grids of all chunks equal those of
 * one call over their union up to the order of the additions.  Needs a plan with dense grids (a
 * plan with ARLMI_FLAG_SPARSE_GRIDS returns ARLMI_E_ARG: each chunk would touch other tiles). */
int arlmi_grid_add(const arlmi_plan* plan, const void* d_vis, const void* d_weight, int dopsf,
                   const arlmi_kernel* kernel, void* d_grids, double* d_sumwt, int precision, void* stream);

/* Batched gridding and degridding of several plans, one launch per kernel support (no
 * counterpart in ARL, which images its vis lists one call at a time, imaging_serial.py:176-197,
 * or as separate Dask tasks, imaging_arlexecute.py:213-221).  Each item names a plan with its
 * separable kernel, gridding inputs (vis, imaging weight), grids and sums of weights (as
 * arlmi_grid / arlmi_degrid take them) and the degridding output.  arlmi_batch_grid equals
 * arlmi_grid (dopsf as there) on every item, arlmi_batch_degrid arlmi_degrid; work items of all
 * plans of one support run in one launch (each block finds its plan), then the split tiles of all
 * plans are summed in one launch.  The batch records the plans' work lists and buffer addresses
 * when it is created: after arlmi_plan_update, or with other buffers, create it again.  Creation
 * synchronises; the calls allocate nothing, record no memset and never synchronise (they can be
 * captured into a HIP graph).  Plans with ARLMI_FLAG_OCCUPIED_TILES or ARLMI_FLAG_DETERMINISTIC
 * and table kernels are not batched (ARLMI_E_UNSUPPORTED). */
typedef struct arlmi_batch arlmi_batch; /* opaque */
typedef struct arlmi_batch_item {
  const arlmi_plan* plan;
  const void* d_vis;      /* [V] complex, gridding input */
  const void* d_weight;   /* [V] real imaging weights */
  arlmi_kernel kernel;    /* separable */
  void* d_grids;          /* [plane][chan][N][N]: gridding output, degridding input */
  double* d_sumwt;        /* [plane][chan] */
  void* d_vis_out;        /* [V] complex, degridding output */
} arlmi_batch_item;
int arlmi_batch_create(const arlmi_batch_item* items, int32_t n, int precision, arlmi_batch** out);
int arlmi_batch_grid(const arlmi_batch* batch, int dopsf, void* stream);
int arlmi_batch_degrid(const arlmi_batch* batch, void* stream);
void arlmi_batch_destroy(arlmi_batch* batch);

/* degrid_visibility_from_griddata (gridding.py:233-267): d_vis_out[row] = sum CF * G
 * (no conjugation), reading the plane each row belongs to.  Rows in no plane get 0
 * (imaging_serial.py:92 zero copy); a row in two w slabs keeps the later slab's value
 * (imaging_serial.py:97). */
int arlmi_degrid(const arlmi_plan* plan, const void* d_grids, const arlmi_kernel* kernel, void* d_vis_out,
                 int precision, void* stream);

/* fft_griddata_to_image (gridding.py:270-289) fused with the w-stack combination
 * (wstack_single.py:93-94), the per-slab sumwt weighting and normalize_sumwt
 * (imaging_serial.py:193-196, imaging/base.py:84-101):
 *   img[c] = gcf[c] / S_c * sum_planes Re(b_p * N^2 ifft(G[p][c]))   (b_p = 1 in MODE_2D)
 *   S_c = sum over all planes of sumwt[p][c]; img = 0 if S_c <= 0; a plane with
 *   sumwt[p][c] <= 0 contributes nothing (normalize_sumwt zeroes its slab image).
 * normalize = 0 skips the division (the plain plane sum: invert_wstack_single / invert_2d with
 * normalize=False, and the adjoint of predict).  normalize = 2 is the list workflows'
 * normalize=False (imaging_serial.py:166-197): every slab's unnormalised image weighted by its
 * sumwt[p][c], summed and divided by S_c (MODE_WSTACK); in MODE_2D the plain sum, 0 if S_c <= 0.
 * d_img_im (may be NULL, MODE_2D only) receives the
 * imaginary part (invert_2d imaginary=True).  d_grids is used as scratch (destroyed).
 * d_gcf: float64 [nchan][N][N]. */
int arlmi_grids_to_image(const arlmi_plan* plan, void* d_grids, const double* d_sumwt, const double* d_gcf,
                         int normalize, void* d_img_re, void* d_img_im, int precision, void* stream);

/* fft_image_to_griddata (gridding.py:292-302) fused with the w-stack model screen
 * (wstack_single.py:50-59): G[p][c] = fft(model[c] * gcf[c] * conj(b_p)); b_p = 1 in MODE_2D. */
int arlmi_image_to_grids(const arlmi_plan* plan, const void* d_model, const double* d_gcf, void* d_grids,
                         int precision, void* stream);

/* ARL fft / ifft (processing_library/fourier_transforms/fft_support.py:21-89) in place on
 * a batch of complex N x N planes: fftshift(fft2(ifftshift(a))) for direction = +1 and
 * fftshift(ifft2(ifftshift(a))) (1/N^2 normalised) for direction = -1. */
int arlmi_fft2d(void* d_data, int32_t npixel, int64_t nbatch, int direction, int precision, void* stream);

/* w-term screen (processing_library/fourier_transforms/convolutional_gridding.py:120-188 via
 * create_w_term_like, image/operations.py:364-408): d_out[N][N] complex =
 * exp(-2 pi i w (1 - sqrt(1 - r^2))), r^2 = fov^2 ((y-cy)^2 + (x-cx)^2)/N^2, fov = N*cellsize,
 * centre cx = cy = N//2; 0 where r^2 >= 1. */
int arlmi_w_screen(int32_t npixel, double cellsize, double w, void* d_out, int precision, void* stream);

/* Direct DFT of point components (predict_skycomponent_visibility, imaging/base.py:207-220;
 * simulate_point, coordinate_support.py:192-212): d_vis[row] += sum_k flux[k][chan(row)] *
 * exp(-2 pi i (u l_k + v m_k + w (sqrt(1 - l_k^2 - m_k^2) - 1))).  d_lmflux: float64 [ncomp][2 + nchan]
 * (l, m, flux per image channel).  d_chan: int32 [nrows] image channel of each row, or NULL (0).
 * The phase d = u l + v m + w (n - 1) is float64 in both precisions; exp(-2 pi i d) is evaluated as
 * cos(pi r) + i sin(pi r) of r = -2 (d - rint(d)) (an exact reduction, so the phasor is that of d
 * itself; numpy's cos(fl(-2 pi d)) carries an extra ~|2 pi d| * 1e-16 rad), with r in float32 for
 * precision 32.  Components are summed per row in float64 and added to d_vis once.  At most
 * 64 KiB / (8 (3 + nchan)) components per call. */
int arlmi_dft(const double* d_uvw, const int32_t* d_chan, int64_t nrows, const double* d_lmflux, int32_t ncomp,
              int32_t nchan, void* d_vis, int precision, void* stream);

/* grid_weight_to_griddata (processing_components/griddata/gridding.py:161-180): every row's
 * imaging weight is added at its nearest grid pixel, d_density[chan][plane][pv][pu] (float64, N x N
 * per channel and grid w plane, zeroed by the call), and to d_sumwt[chan] (float64, zeroed).
 * Rows are mapped as convolution_mapping maps them (gridding.py:40-85; MODE_2D geometry, planes =
 * the grid's w planes), with its asserts: 0 <= pu, pv < N (the kernel footprint is not checked,
 * since weights are not convolved), grid and kernel w planes below nz, channel in range; a row
 * that fails them returns the matching ARLMI_E_* code.  Synchronises `stream`.  The float64 sums
 * are reproducible to rounding (the order of the adds varies). */
int arlmi_grid_weights(const arlmi_geom* g, const double* d_uvw, const double* d_freq, int64_t nrows,
                       const void* d_weight, double* d_density, double* d_sumwt, int precision, void* stream);

/* griddata_reweight (gridding.py:209-230), the uniform weighting of weight_visibility
 * (processing_components/imaging/weighting.py:19-40): d_weight[row] /= d_density at the row's pixel
 * where that density is non-zero (in place, in float64, rounded to the call's precision).  The
 * mapping is checked for every row before any weight changes, as ARL asserts first.
 * Synchronises `stream`. */
int arlmi_reweight(const arlmi_geom* g, const double* d_uvw, const double* d_freq, int64_t nrows,
                   const double* d_density, void* d_weight, int precision, void* stream);

/* Columns of a BlockVisibility (data_models/memory_data_models.py:940-1094), npol = 1.
 * Complex and real columns are in the call's precision. */
typedef struct arlmi_blocks {
  int32_t ntimes, nants, nchan, reserved;
  const void* d_vis;                  /* [ntimes][nants][nants][nchan] complex */
  const double* d_uvw;                /* [ntimes][nants][nants][3] metres */
  const void* d_weight;               /* [ntimes][nants][nants][nchan] real, or NULL */
  const void* d_imaging_weight;       /* [ntimes][nants][nants][nchan] real, or NULL */
  const double* d_time;               /* [ntimes], or NULL */
  const double* d_integration_time;   /* [ntimes], or NULL */
  const double* d_frequency;          /* [nchan] Hz (required) */
  const double* d_channel_bandwidth;  /* [nchan], or NULL */
} arlmi_blocks;

/* Columns of row-form Visibility (memory_data_models.py:745-937), npol = 1.  A NULL pointer
 * skips that column (as does a NULL input column of arlmi_blocks). */
typedef struct arlmi_rows {
  double* d_uvw;                /* [V][3] wavelengths */
  void* d_vis;                  /* [V] complex */
  void* d_weight;               /* [V] real */
  void* d_imaging_weight;       /* [V] real */
  double* d_time;               /* [V] */
  double* d_integration_time;   /* [V] */
  double* d_frequency;          /* [V] */
  double* d_channel_bandwidth;  /* [V] */
  int64_t* d_antenna1;          /* [V] */
  int64_t* d_antenna2;          /* [V] */
} arlmi_rows;

/* convert_blockvisibility_to_visibility / convert_blocks (processing_components/visibility/
 * coalesce.py:85-109, 331-444), no averaging.  Rows are the lower triangle [t, a2, a1] with
 * a1 < a2 of every integration, in numpy.argwhere order (t, a2, a1, chan) (:413-416):
 *   row = (t * nbl + a2 (a2 - 1) / 2 + a1) * nchan + chan,   nbl = nants (nants - 1) / 2,
 * so V = ntimes * nbl * nchan.  antenna1 = a1, antenna2 = a2; uvw = uvw_m[t, a2, a1] * (f / c)
 * with c = 299792458 m/s, evaluated in that order (:435-437), which makes it bit-identical to
 * ARL's float64 result.  d_cindex (int64 [ntimes][nants][nants][nchan], or NULL) receives the row
 * of every lower-triangle element and 0 elsewhere (:418). */
int arlmi_block_to_rows(const arlmi_blocks* blocks, const arlmi_rows* rows, int64_t* d_cindex, int precision,
                        void* stream);

/* Baseline-dependent averaging, coalesce_visibility with time_coal or frequency_coal != 0
 * (processing_components/visibility/coalesce.py:36-82, average_in_blocks :146-328), Stokes I.
 * Step 1: for every baseline b = a2 * nants + a1 (both triangles and the diagonal) the max over
 * time of u*u + v*v (uvw in metres, :198-207) into d_uvdist2[b], and its summed weight over
 * (time, channel) into d_wsum[b] (d_weight NULL = 1).  The caller derives the factors as ARL does
 * (:208-226): uvmax = sqrt(max d_uvdist2); factor = round(coal * uvmax / sqrt(d_uvdist2[b]))
 * (max_coal where that distance is 0), 1 for baselines without weight, clipped to [1, max_coal];
 * a baseline then has ceil(ntimes / tf) x ceil(nchan / ff) cells (factors <= 1 keep every sample). */
int arlmi_bda_uvdist(const double* d_uvw, const void* d_weight, int32_t ntimes, int32_t nants, int32_t nchan,
                     double* d_uvdist2, double* d_wsum, int precision, void* stream);
/* Step 2 (:247-328): d_tav, d_fav int32 [nants^2] the factors, d_rowoff int64 [nants^2 + 1] the
 * exclusive sum of the cells per baseline (nrows = d_rowoff[nants^2]).  Row d_rowoff[b] + T * nf + F
 * of baseline b = (a2, a1) holds the [time chunk T, frequency chunk F] cell, with average_chunks2's
 * two-stage weighted averages (array_functions.py:85-118: each integration's channel chunk first,
 * then the integrations with those partial weights; a factor <= 1 passes values through
 * unchanged): vis (weights: weight), weight = the summed weights, imaging_weight (weights: weight),
 * time, frequency, uvw in wavelengths (uvw_m * (f / c) per sample), integration time and
 * channel bandwidth as averages times the baseline's number of cells (ARL's sum_from_grid),
 * antenna1 = a1, antenna2 = a2.  The blocks need time, integration time, frequency and channel
 * bandwidth; weight and imaging weight may be NULL (1).  d_cindex (int64 [ntimes][nants][nants]
 * [nchan], or NULL) receives d_rowoff[b] + ((t * nchan + ch) mod cells of b) for element
 * (t, a2, a1, ch), as ARL's flat assignment of range(cells) gives it (:278).  The output
 * columns must not overlap the block arrays. */
int arlmi_bda_average(const arlmi_blocks* blocks, const int32_t* d_tav, const int32_t* d_fav, const int64_t* d_rowoff,
                      int64_t nrows, const arlmi_rows* rows, int64_t* d_cindex, int precision, void* stream);

/* apply_gaintable with scalar gains (processing_components/calibration/operations.py:79-164,
 * scalar branch :127-140), in place on a BlockVisibility's vis [ntimes][nants][nants][nchan]:
 *   vis[t][i][j][ch] *= lg[s][i][ch] * conj(lg[s][j][ch]),   s = d_gain_slot[t],
 * with lg = g (inverse = 0) or lg = 1/g where |g| > 0 and 1 elsewhere (inverse = 1, :129-131), and
 * vis set to 0 where gain_weight[s][i][ch] * gain_weight[s][j][ch] == 0 (:137-138).  Every
 * (i, j) pair, both triangles and the diagonal, is processed as in ARL; weights are not
 * changed (ARL's appliedwt is computed but never stored back, :164).  d_gain: complex
 * [ngain_times][nants][nchan]; d_gain_weight: float64, same shape, or NULL (all 1).  d_gain_slot:
 * int32 [ntimes], the gain-table row whose time matches each integration (:112-113); an
 * integration whose slot lies outside [0, ngain_times) is left unchanged. */
int arlmi_apply_gaintable(void* d_vis, int32_t ntimes, int32_t nants, int32_t nchan, const void* d_gain,
                          const double* d_gain_weight, const int32_t* d_gain_slot, int32_t ngain_times,
                          int inverse, int precision, void* stream);

/* The same scalar gain convention on rows (operations.py:135-137 seen through convert_blocks'
 * (a2, a1) rows, so the un-conjugated gain is antenna2's):
 *   d_out[r] = d_in[r] * g[d_gidx2[r]] * conj(g[d_gidx1[r]]),
 * d_gidx1 / d_gidx2: int32 [nrows], index into d_gain (complex [ngain], flattened by the caller,
 * e.g. (slot * nants + antenna) * nchan + chan) of the gain of antenna1 / antenna2.  A row with an
 * index outside [0, ngain) gets 0.  d_out may equal d_in. */
int arlmi_apply_gains_rows(int64_t nrows, const int32_t* d_gidx1, const int32_t* d_gidx2, const void* d_gain,
                           int64_t ngain, const void* d_in, void* d_out, int precision, void* stream);

/* Adjoint of arlmi_apply_gains_rows for a real loss L (no ARL counterpart: ARL has no
 * gradients; this is what gradient-based calibration through the convention above needs).
 * Gradients are complex numbers dL/dRe + i dL/dIm.  With G = d_grad_out and s = g[gidx2] conj(g[gidx1]):
 *   d_grad_in[r]            = G[r] * conj(s_r)                      (precision complex, or NULL)
 *   d_grad_gain[gidx2[r]]  += G[r] * g[gidx1[r]] * conj(d_in[r])
 *   d_grad_gain[gidx1[r]]  += conj(G[r]) * g[gidx2[r]] * d_in[r]    (complex128 [ngain], or NULL)
 * d_grad_gain is accumulated into (the caller zeroes it); rows with an index outside [0, ngain)
 * contribute nothing and get d_grad_in = 0. */
int arlmi_apply_gains_rows_adjoint(int64_t nrows, const int32_t* d_gidx1, const int32_t* d_gidx2, const void* d_gain,
                                   int64_t ngain, const void* d_in, const void* d_grad_out, void* d_grad_in,
                                   void* d_grad_gain, int precision, void* stream);

/* phaserotate_visibility (processing_components/visibility/base.py:295-381), in place on the
 * caller's copy (ARL rotates a copy_visibility; the caller copies first and handles the
 * |n| < 1e-15 early return, :312-313).  h_lmn = {l, m, sqrt(1 - l^2 - m^2) - 1} of the new
 * phase centre relative to the old one (skycoord_to_lmn, coordinate_support.py:150-169), all
 * computed by the caller.  With the uvw held before any rotation, per element
 *   d = u l + v m + w n1,   phasor = exp(-2 pi i d) = cos(-2 pi d) + i sin(-2 pi d)
 * (simulate_point, coordinate_support.py:192-212, evaluated without contraction in that order),
 *   vis *= conj(phasor)  (inverse = 0)  or  vis *= phasor  (inverse = 1),
 * the complex product in float64 (numpy's complex128 order) and stored in the vis precision.
 * h_trig != NULL (tangent = False) then rotates uvw in place (:329-339):
 *   xyz = uvw_to_xyz(uvw, ha0, dec0),  uvw = xyz_to_uvw(xyz, ha1, dec1)
 * (coordinate_support.py:59-113, numpy's operation order) with h_trig = {cos ha0, sin ha0,
 * cos dec0, sin dec0, cos ha1, sin ha1, cos dec1, sin dec1}, ha = -ra of the old (0) and new (1)
 * phase centres (for inverse = 1 ARL uses the same angles, :330-333).
 *
 * Rows (Visibility, :318-340): d_uvw [V][3] wavelengths, d_vis [V][npol]. */
int arlmi_phaserotate_rows(double* d_uvw, void* d_vis, int64_t nrows, int32_t npol, const double* h_lmn,
                           const double* h_trig, int inverse, int precision, void* stream);

/* Blocks (BlockVisibility, :342-379): d_uvw [nblock][3] metres (nblock = ntimes * nants * nants,
 * every (i, j) pair), d_vis [nblock][nchan][npol]; the phasor of channel ch uses
 * uvw_m * (f[ch] / c), c = 299792458 m/s (:344-350), d_frequency [nchan] Hz on the device. */
int arlmi_phaserotate_blocks(double* d_uvw, void* d_vis, int64_t nblock, int32_t nchan, int32_t npol,
                             const double* d_frequency, const double* h_lmn, const double* h_trig, int inverse,
                             int precision, void* stream);

/* hogbom (processing_library/arrays/cleaners.py:12-70), float64, one image plane:
 *   absolutethresh = max(thresh, fracthresh * max|dirty|), pmax = max(psf) (> 0, else ARLMI_E_ARG),
 *   repeat up to niter times: (mx, my) = first index of max |res * window| (|res| for a NULL window,
 *   row-major order, as numpy's argmax), mval = res[mx, my] * gain / pmax, comps[mx, my] += mval,
 *   res[overlap] -= psf[overlap] * mval over overlapIndices (:134-154): rows
 *   [max(0, mx - pnx/2), min(nx, mx + pnx/2)) against PSF rows from pnx/2 + (row - mx), likewise for
 *   columns; stop once |res[mx, my]| < 0.9 * absolutethresh after the subtraction.
 * d_res [nx][ny]: the dirty image on entry, the residual on return; d_comps [nx][ny]: the
 * components (zeroed here); d_psf [psf_nx][psf_ny]; d_window [nx][ny] or NULL.  *iterations (may be
 * NULL) receives ARL's iteration count.  0 < gain < 2 and niter > 0 are required (:33-34).  The
 * operations are ARL's, uncontracted, so the results equal numpy's bit for bit (a NaN is never
 * chosen as the peak, where numpy's argmax would choose it).  Synchronises the stream about every 64
 * iterations to read the end flag.  The control block and band maxima live in a per-device buffer
 * of the library (grown on demand, kept until exit), so calls that overlap on one device race:
 * run them one after another (on one stream, or synchronised). */
int arlmi_hogbom(double* d_res, double* d_comps, const double* d_psf, int32_t psf_nx, int32_t psf_ny,
                 const double* d_window, int32_t nx, int32_t ny, double gain, double thresh, int32_t niter,
                 double fracthresh, int32_t* iterations, void* stream);

/* arlmi_hogbom with no allocation (no counterpart in ARL): for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and synchronisation as arlmi_hogbom, but the control block
 * and band maxima live in a fixed 64 KiB device array of the library (one per GPU, held until the
 * process ends) instead of a buffer grown with hipMalloc; it calls no hipMalloc / hipFree.  Every
 * image size fits (the scratch layout needs at most 49,408 bytes); a larger need would return
 * ARLMI_E_UNSUPPORTED.  All calls on one GPU share that array and the control block is cleared at
 * the start of every call, so calls that overlap on one GPU race: run them one after another. */
int arlmi_hogbom_noalloc(double* d_res, double* d_comps, const double* d_psf, int32_t psf_nx, int32_t psf_ny,
                         const double* d_window, int32_t nx, int32_t ny, double gain, double thresh, int32_t niter,
                         double fracthresh, int32_t* iterations, void* stream);

/* solve_gaintable for scalar gains (processing_components/calibration/calibration.py:30-101,
 * npol = 1, crosspol = False), float64 arithmetic:
 *   for every gain-table row s with at least one integration (d_active[s] != 0; d_slot[t] is the
 *   row of integration t, or -1, as |time - gt.time| < interval / 2 selects it, :61-63):
 *     point-source equivalents (:64-78; divide_visibility, visibility/operations.py:241-264): with
 *       a model, x_t = vis / model and w_t = |model|^2 * weight where that is > 0 (else 0); without,
 *       x_t = vis and w_t = weight; x = sum_t x_t w_t / sum_t w_t (0 where the sum is 0), for the
 *       lower-triangle element [a2, a1] of every (antenna pair, channel);
 *     solve_antenna_gains_itsubs_scalar (processing_library/calibration/solvers.py:61-100):
 *       x[i, j] = conj(x[j, i]) above the diagonal, 0 on it; up to niter iterations of
 *       gain_substitution_scalar (:103-127: new g_a = sum_k g_k x[k, a] w[k, a] / sum_k |g_k|^2 w[k, a],
 *       0 for every channel of an antenna whose denominator is 0 in any channel), |g| = 1 if
 *       phase_only, the phase of antenna refant removed per channel, damping
 *       g = (1 - damping) g + damping g_last, stop when max |g - g_last| < tol;
 *     solution_residual_scalar (:302-332): sqrt(sum over channels and pairs of |x - conj(g_i) g_j|^2 w
 *       / sum of w), one value written for every channel (as ARL's broadcast sum gives);
 *     normalise and not phase_only: g /= mean |g| (:79-81).
 * d_vis, d_model (or NULL), d_weight (or NULL = 1): BlockVisibility [ntimes][nants][nants][nchan] in
 * the given precision; d_gain complex128 [ngain][nants][nchan] (in: start, out: solution);
 * d_gain_weight float64 (the last denominators), d_residual float64 [ngain][nchan],
 * d_iterations int32 [ngain]; d_work: 3 * ngain * nants^2 * nchan doubles of scratch.
 * nants * nchan <= 2048 and nchan <= 64.  Rows without integrations are left unchanged. */
int arlmi_solve_gains(const void* d_vis, const void* d_model, const void* d_weight, int32_t ntimes, int32_t nants,
                      int32_t nchan, const int32_t* d_slot, int32_t ngain, const int32_t* d_active, void* d_gain,
                      double* d_gain_weight, double* d_residual, int32_t* d_iterations, double* d_work,
                      int phase_only, int32_t niter, double tol, int32_t refant, double damping, int normalise,
                      int precision, void* stream);

/* Measurement only (no ARL counterpart): for the last arlmi_hogbom call on the calling thread,
 * the pixels whose bands were scanned for the peak (the whole image once, then the bands of
 * rows each PSF subtraction changed) and the pixels the PSF subtractions covered; the minor
 * cycle's algorithmic HBM bytes are 8 * scanned (+ 8 * scanned with a window) + 16 * subtracted. */
int arlmi_hogbom_traffic(int64_t* scanned_pixels, int64_t* subtracted_pixels);

/* Opt-in kernel timing (no ARL counterpart; measurement only): with profiling on, every library
 * kernel launch is bracketed by two HIP events on its stream.  collect() waits for them and adds
 * the elapsed times per kernel; report() writes "kernel total_ms launches" lines into buf and
 * returns the size needed.  only() restricts the timing to a comma-separated list of kernel
 * names (NULL or "" times every kernel): each timed launch adds two event markers to its
 * stream, a few microseconds of stream time each. */
int arlmi_profile_enable(int on);
int arlmi_profile_only(const char* names);
int arlmi_profile_collect(void);
int arlmi_profile_reset(void);
int arlmi_profile_report(char* buf, int cap);

#ifdef __cplusplus
}
#endif
#endif /* ARLMI_H */
