"""FP64 NumPy restatement of ARL's Fourier-processing hot path.

THIS MODULE IS TEST INFRASTRUCTURE (the oracle). Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker of the HIP product path or as the timed CPU
baseline. The product package ``arlmi`` never imports it and never falls back
to it.

It restates, in complex128/float64 NumPy, the functions of the
SKA-ScienceDataProcessor/algorithm-reference-library (ARL) hot path.  Each
function names the reference file:line it follows.  ARL itself cannot be
imported here (its data models need astropy, which is absent), so the WCS
arithmetic that ARL delegates to astropy/wcslib is written out explicitly, in
the same operation order wcslib uses for linear axes:

    world -> pixel (origin 0):  pix = (world - crval) / cdelt + crpix - 1
    pixel -> world (origin 0):  world = cdelt * ((pix + 1) - crpix) + crval

Pinning (what checks each function; see tests/test_oracle_*.py):
  grdsf, create_pswf_convolutionfunction   ARL known answers
      (tests/processing_components/test_griddata_kernels.py:60-101)
  create_box_convolutionfunction           ARL known answer (:45-58)
  create_awterm_convolutionfunction (w only), apply_bounding_box_convolutionfunction
                                           ARL known answers and symmetries (:104-139, :235-296)
  w_beam, coordinates*                     ARL known answers
      (tests/processing_library/test_convolutional_gridding.py:21-44) + closed form
  fft / ifft                               numpy.fft identity + direct DFT on small arrays
  convolution_mapping, grid / degrid, fft_griddata_to_image, fft_image_to_griddata,
  invert_2d, predict_2d, normalize_sumwt   brute-force direct sums on tiny inputs, the
                                           adjoint identity, box == nearest-pixel gridding,
                                           point-source agreement with the direct DFT
  vis_wslice_iter, *_wstack_single, w-stack list workflows
                                           closed form of the slab partition, adjoint
                                           identity, and ARL's point-source recipe with w != 0
                                           (tests/workflows/test_imaging_serial.py:153-187,
                                           232-243) against the direct DFT: the 101-slab residual
                                           falls to the w = 0 kernel floor, the w-stack +
                                           w-projection hybrid stays under ARL's 1.0 on 100 Jy,
                                           and a conjugated screen or a w shift of the wrong
                                           sign fails (tests/test_oracle_wstack_dft.py); the
                                           normalize semantics of the list workflow by weight
                                           scaling
  simulate_point / predict_skycomponent_visibility
                                           ARL known answers
      (tests/processing_library/test_coordinate_support.py:104-123)
  xyz_to_uvw, xyz_at_latitude, hadec_to_azel
                                           closed forms (rotation invariants, known angles)
  create_blockvisibility, convert_blocks, decoalesce_visibility
                                           the element loop of convert_blocks, the independent
                                           row construction of create_visibility (same rows, same
                                           bits, other order), cindex round trips
  apply_gaintable, apply_gains_rows        ARL's own assertions (test_calibration_operations.py:
                                           92-135: apply changes, apply + inverse = identity to
                                           1e-12, zero gains + inverse = no change) and block/row
                                           agreement through convert_blocks
  apply_gains_rows_adjoint                 central finite differences and the dot-product identity
  grid_weight_to_griddata, griddata_reweight, weight_visibility
                                           a direct per-pixel count on integer-valued weights, and
                                           the uniform-weighting identity (reweighted weights grid
                                           to exactly 1 on every occupied pixel); ARL's own test of
                                           it (test_griddata_gridding.py:283-295) needs LOWBD2 data
                                           and is not reproducible here
  skycoord_to_lmn                          ARL known answers (test_coordinate_support.py:125-141)
  uvw_to_xyz                               inverse of xyz_to_uvw, rotation invariants
  phaserotate_rows, phaserotate_blocks     ARL's identities (test_coordinate_support.py:143-170:
                                           vis x phasor = the rotated prediction, 6 x 6 x 6
                                           centre/source combinations) and ARL's tests
                                           (test_visibility_operations.py:140-245): rotation
                                           equals a prediction made at the new centre with the
                                           hour angle shifted, rotating there and back (and with
                                           inverse) is the identity; blocks agree with rows
                                           through convert_blocks
  pixel_to_skycoord_sin, separation, shift_vis_to_image
                                           closed forms (the reference pixel is crval; on the
                                           equator an x offset gives ra = ra0 + asin(l); known
                                           separations), skycoord_to_lmn of the deprojected
                                           direction returns the pixel's (l, m), and a point
                                           source imaged through the shift peaks at the pixel of
                                           its (l, m) about the image centre
  overlapIndices                          ARL known answer (test_image_deconvolution.py:59-68)
  hogbom, deconvolve_cube_hogbom           closed form of a delta PSF (geometric decay of the
                                           peak, iteration count from the stop rule), flux
                                           conservation dirty - res = comps convolved with the
                                           PSF (an FFT convolution), windows, argmax ties
  solve_gaintable_scalar and its solver    ARL's own assertions (test_calibration_solvers.py:
                                           55-115: residual < 3e-8 and gains away from 1 for
                                           phase-only, amplitude + phase with normalisation,
                                           several integrations per solution, point-source input),
                                           recovery of the simulated gains up to the reference
                                           phase, and a dead antenna's zero gain
  average_chunks(2), bda_factors, average_in_blocks
                                           ARL's assertions (test_visibility_coalesce.py:39-110:
                                           row counts, the minimum frequency kept, decoalesce
                                           sizes), closed forms (factor 1 is the plain conversion;
                                           constant data average to themselves; weighted sums
                                           are conserved), brute-force chunk sums
"""

import numpy

__all__ = [
    "coordinateBounds", "coordinates", "coordinates2", "coordinates2Offset", "grdsf", "w_beam",
    "fft", "ifft", "ImageGeom", "create_image_geom", "create_pswf_convolutionfunction",
    "create_box_convolutionfunction", "create_awterm_convolutionfunction",
    "apply_bounding_box_convolutionfunction", "calculate_bounding_box_convolutionfunction",
    "convolution_mapping", "grid_visibility_to_griddata", "grid_visibility_to_griddata_loop",
    "grid_visibility_to_griddata_fast", "degrid_visibility_from_griddata",
    "degrid_visibility_from_griddata_loop", "fft_griddata_to_image", "fft_image_to_griddata",
    "normalize_sumwt", "invert_2d", "predict_2d", "create_w_term_like", "vis_wslice_iter",
    "invert_wstack_single", "predict_wstack_single", "invert_list_wstack", "predict_list_wstack",
    "simulate_point", "predict_skycomponent_visibility", "xyz_at_latitude", "xyz_to_uvw",
    "hadec_to_azel", "create_visibility", "apply_gains_rows", "create_blockvisibility", "convert_blocks",
    "convert_blocks_loop", "decoalesce_visibility", "apply_gaintable", "apply_gains_rows_adjoint",
    "grid_weight_to_griddata", "griddata_merge_weights", "griddata_reweight", "weight_visibility",
    "uvw_to_xyz", "skycoord_to_lmn", "phaserotate_rows", "phaserotate_blocks", "pixel_to_skycoord_sin",
    "separation", "shift_vis_to_image", "overlapIndices", "hogbom",
    "deconvolve_cube_hogbom", "gain_substitution_scalar", "solution_residual_scalar",
    "solve_antenna_gains_itsubs_scalar", "solve_gaintable_scalar", "average_chunks", "average_chunks2",
    "bda_factors", "average_in_blocks",
]

D2R = numpy.pi / 180.0
C_LIGHT = 299792458.0  # astropy constants.c in m/s


# ---------------------------------------------------------------------------
# processing_library/fourier_transforms/convolutional_gridding.py
# ---------------------------------------------------------------------------

def coordinateBounds(npixel):
    """convolutional_gridding.py:18-34."""
    if npixel % 2 == 0:
        return -0.5, 0.5 * (npixel - 2) / npixel
    return -0.5 * (npixel - 1) / npixel, 0.5 * (npixel - 1) / npixel


def coordinates(npixel):
    """convolutional_gridding.py:37-41: (arange(n) - n//2)/n."""
    return (numpy.arange(npixel) - npixel // 2) / npixel


def coordinates2(npixel):
    """convolutional_gridding.py:44-50."""
    return (numpy.mgrid[0:npixel, 0:npixel] - npixel // 2) / npixel


def coordinates2Offset(npixel, cx, cy, quadrant=False):
    """convolutional_gridding.py:53-70."""
    if cx is None:
        cx = npixel // 2
    if cy is None:
        cy = npixel // 2
    if not quadrant:
        mg = numpy.mgrid[0:npixel, 0:npixel]
    else:
        mg = numpy.mgrid[0:npixel // 2 + 1, 0:npixel // 2 + 1]
    return (mg[0] - cy) / npixel, (mg[1] - cx) / npixel


_GRDSF_P = numpy.array([[8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1],
                        [4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2]])
_GRDSF_Q = numpy.array([[1.0000000e0, 8.212018e-1, 2.078043e-1],
                        [1.0000000e0, 9.599102e-1, 2.918724e-1]])


def grdsf(nu):
    """Schwab's rational PSWF approximation (m=6, alpha=1).

    convolutional_gridding.py:72-117 (coefficients 83-86).  Returns
    (psi(nu), (1 - nu**2) * psi(nu)); both are 0 for |nu| > 1.
    """
    nu = numpy.abs(numpy.asarray(nu, dtype=float))
    nuend = numpy.zeros_like(nu)
    part = numpy.zeros(nu.shape, dtype=int)
    part[(nu >= 0.0) & (nu < 0.75)] = 0
    part[(nu >= 0.75) & (nu <= 1.0)] = 1
    nuend[(nu >= 0.0) & (nu < 0.75)] = 0.75
    nuend[(nu >= 0.75) & (nu <= 1.0)] = 1.0
    delnusq = nu ** 2 - nuend ** 2
    top = _GRDSF_P[part, 0]
    for k in range(1, _GRDSF_P.shape[1]):
        top += _GRDSF_P[part, k] * numpy.power(delnusq, k)
    bot = _GRDSF_Q[part, 0]
    for k in range(1, _GRDSF_Q.shape[1]):
        bot += _GRDSF_Q[part, k] * numpy.power(delnusq, k)
    psi = numpy.zeros_like(nu)
    ok = bot > 0.0
    psi[ok] = top[ok] / bot[ok]
    psi[nu > 1.0] = 0.0
    return psi, (1 - nu ** 2) * psi


def w_beam(npixel, field_of_view, w, cx=None, cy=None, remove_shift=False):
    """W-term phase screen exp(-2 pi i w (1 - sqrt(1 - r^2))), 0 outside the unit circle.

    convolutional_gridding.py:120-188 (the quadrant + reflect-pad form, 172-184).
    """
    if cx is None:
        cx = npixel // 2
    if cy is None:
        cy = npixel // 2
    ly, mx = coordinates2Offset(npixel, cx, cy, quadrant=True)
    r2 = field_of_view ** 2 * (ly ** 2 + mx ** 2)
    with numpy.errstate(invalid="ignore"):
        ph = -2 * numpy.pi * w * (1 - numpy.sqrt(1.0 - r2))
    numpy.putmask(ph, r2 >= 1.0, 0)
    cp = numpy.exp(1j * ph)
    numpy.putmask(cp, r2 >= 1.0, 0 + 0j)
    numpy.putmask(cp, r2 == 0, 1.0 + 0j)
    if remove_shift:
        cp /= cp[-1, -1]
    cp = numpy.pad(cp, ((0, int(cx) + npixel % 2 - 1), (0, int(cy) + npixel % 2 - 1)), 'reflect')
    return cp


# ---------------------------------------------------------------------------
# processing_library/fourier_transforms/fft_support.py (numpy branch)
# ---------------------------------------------------------------------------

def fft(a):
    """Image -> grid: fftshift(fft2(ifftshift(a))) over the last two axes.

    fft_support.py:21-38 (numpy branch used when pyfftw is absent).
    """
    if a.ndim == 4:
        return numpy.fft.fftshift(numpy.fft.fft2(numpy.fft.ifftshift(a, axes=[2, 3])), axes=[2, 3])
    if a.ndim == 5:
        return numpy.fft.fftshift(numpy.fft.fft2(numpy.fft.ifftshift(a, axes=[3, 4])), axes=[3, 4])
    return numpy.fft.fftshift(numpy.fft.fft2(numpy.fft.ifftshift(a)))


def ifft(a):
    """Grid -> image: fftshift(ifft2(ifftshift(a))) over the last two axes (1/N^2 normalised).

    fft_support.py:55-71.
    """
    if a.ndim == 4:
        return numpy.fft.fftshift(numpy.fft.ifft2(numpy.fft.ifftshift(a, axes=[2, 3])), axes=[2, 3])
    if a.ndim == 5:
        return numpy.fft.fftshift(numpy.fft.ifft2(numpy.fft.ifftshift(a, axes=[2, 3, 4])), axes=[2, 3, 4])
    return numpy.fft.fftshift(numpy.fft.ifft2(numpy.fft.ifftshift(a)))


# ---------------------------------------------------------------------------
# Image / grid geometry (the WCS values ARL stores and the arithmetic astropy
# performs on them).  processing_library/image/operations.py:33-69,
# processing_components/imaging/base.py:326-334,
# processing_components/griddata/operations.py:91-145,
# processing_components/griddata/convolution_functions.py:62-147.
# ---------------------------------------------------------------------------

class ImageGeom:
    """The numbers of ARL's image WCS that the hot path reads.

    cdelt[0] = -cell*180/pi, cdelt[1] = +cell*180/pi (degrees), crpix = N//2+1,
    FREQ axis: crval = frequency, cdelt = channel_bandwidth, crpix = 1
    (image/operations.py:61-64; imaging/base.py:328-333).
    """

    def __init__(self, npixel, cellsize, nchan=1, npol=1, frequency=1e8, channel_bandwidth=1e6):
        self.npixel = int(npixel)
        self.nchan = int(nchan)
        self.npol = int(npol)
        self.cellsize = float(cellsize)
        self.cdelt0 = -cellsize * 180.0 / numpy.pi
        self.cdelt1 = cellsize * 180.0 / numpy.pi
        self.crpix0 = float(npixel // 2 + 1)
        self.crpix1 = float(npixel // 2 + 1)
        self.freq0 = float(frequency)
        self.dfreq = float(channel_bandwidth)

    @property
    def shape(self):
        return (self.nchan, self.npol, self.npixel, self.npixel)

    def du(self):
        """UU cdelt of the grid: 1/(N * d2r * cdelt0) (griddata/operations.py:134)."""
        return 1.0 / (self.npixel * D2R * self.cdelt0)

    def dv(self):
        """VV cdelt of the grid (griddata/operations.py:135)."""
        return 1.0 / (self.npixel * D2R * self.cdelt1)

    def wcs_cellsize(self):
        """abs(cdelt0)*pi/180 as create_w_term_like reads it (image/operations.py:389)."""
        return abs(self.cdelt0) * numpy.pi / 180.0


def create_image_geom(npixel, cellsize, nchan=1, frequency=1e8, channel_bandwidth=1e6):
    return ImageGeom(npixel, cellsize, nchan=nchan, frequency=frequency,
                     channel_bandwidth=channel_bandwidth)


class CFGeom:
    """WCS numbers of a ConvolutionFunction (convolution_functions.py:115-131)."""

    def __init__(self, geom, oversampling, support, nz=1, zstep=1e15):
        self.oversampling = int(oversampling)
        self.support = int(support)
        self.nz = int(nz)
        self.zstep = float(zstep)
        self.crpix_z = float(nz // 2 + 1)
        self.du = 1.0 / (geom.npixel * D2R * geom.cdelt0)
        self.dv = 1.0 / (geom.npixel * D2R * geom.cdelt1)
        self.ddu = self.du / oversampling
        self.ddv = self.dv / oversampling
        self.crpix_d = float(oversampling // 2 + 1)


# ---------------------------------------------------------------------------
# processing_components/griddata/kernels.py
# ---------------------------------------------------------------------------

def create_pswf_convolutionfunction(geom, oversampling=8, support=6):
    """PSWF gridding kernel and grid-correction image.

    kernels.py:52-97.  Returns (gcf [nchan,npol,N,N] float64,
    cf [nchan,npol,1,os,os,S,S] complex128, cfgeom).
    """
    kernel = numpy.zeros([oversampling, support])
    for grid in range(support):
        for subsample in range(oversampling):
            nu = ((grid - support // 2) - (subsample - oversampling // 2) / oversampling)
            kernel[subsample, grid] = grdsf([nu / (support // 2)])[1][0]
    kernel /= numpy.sum(numpy.real(kernel[oversampling // 2, :]))
    nchan, npol = geom.nchan, geom.npol
    cf = numpy.zeros([nchan, npol, 1, oversampling, oversampling, support, support]).astype('complex')
    for y in range(oversampling):
        for x in range(oversampling):
            cf[:, :, 0, y, x, :, :] = numpy.outer(kernel[y, :], kernel[x, :])[numpy.newaxis, numpy.newaxis, ...]
    norm = numpy.sum(numpy.real(cf[0, 0, 0, 0, 0, :, :]))
    cf /= norm
    nx = geom.npixel
    nu = numpy.abs(2.0 * coordinates(nx))
    gcf1d = grdsf(nu)[0]
    gcf = numpy.outer(gcf1d, gcf1d)
    gcf[gcf > 0.0] = gcf.max() / gcf[gcf > 0.0]
    gcf_data = numpy.zeros(geom.shape)
    gcf_data[...] = gcf[numpy.newaxis, numpy.newaxis, ...]
    return gcf_data, cf, CFGeom(geom, oversampling, support)


def create_box_convolutionfunction(geom):
    """Box kernel: delta at (...,2,2) of an S=4, os=1 CF; gcf = 1/sinc^2 (kernels.py:19-49)."""
    cf = numpy.zeros([geom.nchan, geom.npol, 1, 1, 1, 4, 4], dtype='complex')
    cf[..., 2, 2] = 1.0 + 0.0j
    nu = numpy.abs(coordinates(geom.npixel))
    gcf1d = numpy.sinc(nu)
    gcf = 1.0 / numpy.outer(gcf1d, gcf1d)
    gcf_data = numpy.zeros(geom.shape)
    gcf_data[...] = gcf[numpy.newaxis, numpy.newaxis, ...]
    return gcf_data, cf, CFGeom(geom, 1, 4)


def create_w_term_like(geom, w, remove_shift=False, cellsize_geom=None, npol=1):
    """Image of the w-term screen centred on crpix (image/operations.py:364-408).

    cellsize = abs(cdelt0)*pi/180 as read back from the WCS (389); the centre is
    crpix-1 because the phasecentre argument is never a SkyCoord class (391-394).
    """
    g = geom if cellsize_geom is None else cellsize_geom
    cellsize = g.wcs_cellsize()
    npixel = geom.npixel
    temp = w_beam(npixel, npixel * cellsize, w=w, cx=geom.crpix0 - 1.0, cy=geom.crpix1 - 1.0,
                  remove_shift=remove_shift)
    out = numpy.zeros([geom.nchan, npol, npixel, npixel], dtype='complex')
    out[:, :, ...] = temp
    return out


def create_awterm_convolutionfunction(geom, nw=1, wstep=1e15, oversampling=8, support=6, use_aaf=True,
                                      maxsupport=512, planes=None):
    """W-only AW (w-projection) kernel, make_pb=None (kernels.py:100-200).

    ``planes`` optionally restricts the (expensive) per-plane FFTs to a subset of
    z indices; the normalisation always uses plane nw//2 (kernels.py:191).
    Returns (gcf, cf, cfgeom).
    """
    nchan, npol, onx = geom.nchan, geom.npol, geom.npixel
    cfg = CFGeom(geom, oversampling, support, nz=nw, zstep=wstep)
    if abs(wstep) > 0.0:
        # wcs_pix2world(range(nw), 0) on the WW axis: cdelt*((z+1) - crpix) + 0
        w_list = [wstep * ((z + 1.0) - cfg.crpix_z) + 0.0 for z in range(nw)]
    else:
        w_list = [0.0]
    nx = max(maxsupport, 2 * oversampling * support)
    ny = nx
    qnx = nx // oversampling
    qny = ny // oversampling
    cf = numpy.zeros([nchan, npol, nw, oversampling, oversampling, support, support], dtype='complex')
    ccell = onx * numpy.abs(D2R * geom.cdelt0) / qnx
    # sub-image geometry: cdelt0 = -ccell/d2r, crpix = qnx//2+1
    sub = ImageGeom(qnx, 1.0, nchan=nchan, npol=npol)
    sub.cdelt0 = -ccell / D2R
    sub.cdelt1 = +ccell / D2R
    sub.crpix0 = float(qnx // 2 + 1)
    sub.crpix1 = float(qny // 2 + 1)
    if use_aaf:
        # create_pswf_convolutionfunction(subim, oversampling=1, support=6) gcf only
        nu = numpy.abs(2.0 * coordinates(qnx))
        g1 = grdsf(nu)[0]
        g2 = numpy.outer(g1, g1)
        g2[g2 > 0.0] = g2.max() / g2[g2 > 0.0]
        norm = 1.0 / g2[numpy.newaxis, numpy.newaxis, ...]
    else:
        norm = 1.0
    z_todo = range(len(w_list)) if planes is None else sorted(set(list(planes) + [nw // 2]))
    ycen, xcen = ny // 2, nx // 2
    for z in z_todo:
        w = w_list[z]
        thisplane = create_w_term_like(sub, w, npol=npol)
        thisplane = thisplane * norm
        padded = numpy.zeros([nchan, npol, ny, nx], dtype='complex')
        ystart = ny // 2 - qny // 2
        xstart = nx // 2 - qnx // 2
        padded[..., ystart:ystart + qny, xstart:xstart + qnx] = thisplane
        padded = ifft(padded)  # fft_image on a RA/DEC image uses ifft (image/operations.py:293)
        for y in range(oversampling):
            ybeg = y + ycen + (support * oversampling) // 2 - oversampling // 2
            yend = y + ycen - (support * oversampling) // 2 - oversampling // 2
            for x in range(oversampling):
                xbeg = x + xcen + (support * oversampling) // 2 - oversampling // 2
                xend = x + xcen - (support * oversampling) // 2 - oversampling // 2
                cf[..., z, y, x, :, :] = padded[..., ybeg:yend:-oversampling, xbeg:xend:-oversampling]
    cf /= numpy.sum(numpy.real(cf[0, 0, nw // 2, oversampling // 2, oversampling // 2, :, :]))
    cf = numpy.conjugate(cf)
    if use_aaf:
        gcf, _, _ = create_pswf_convolutionfunction(geom, oversampling=1, support=6)
    else:
        gcf = numpy.ones(geom.shape)
    return gcf, cf, cfg


def apply_bounding_box_convolutionfunction(cf, fractional_level=1e-4):
    """Clip a CF to |cf| > level*max (convolution_functions.py:159-184). Returns the data only."""
    nx = cf.shape[-1]
    mask = numpy.max(numpy.abs(cf), axis=(0, 1, 2, 3, 4))
    coords = numpy.argwhere(mask > fractional_level * numpy.max(numpy.abs(cf)))
    crpx = int(numpy.round(float(nx // 2) + 1.0))
    crpy = crpx
    x0, y0 = coords.min(axis=0)
    dx = crpx - x0
    dy = crpy - y0
    x0 -= 1
    y0 -= 1
    x1 = crpx + dx - 1
    y1 = crpy + dy - 1
    return cf[..., y0:y1, x0:x1]


def calculate_bounding_box_convolutionfunction(cf, fractional_level=1e-4):
    """Per-z bounding boxes (convolution_functions.py:187-207)."""
    bboxes = []
    threshold = fractional_level * numpy.max(numpy.abs(cf))
    for z in range(cf.shape[2]):
        mask = numpy.max(numpy.abs(cf[:, :, z, ...]), axis=(0, 1, 2, 3))
        coords = numpy.argwhere(mask > threshold)
        x0, y0 = coords.min(axis=0)
        x1, y1 = coords.max(axis=0)
        bboxes.append((z, (y0, y1), (x0, x1)))
    return bboxes


# ---------------------------------------------------------------------------
# processing_components/griddata/gridding.py
# ---------------------------------------------------------------------------

def convolution_mapping(uvw, frequency, geom, cfg, nw_grid=1, wstep_grid=1e15, nchan_cf=None):
    """Per-visibility grid/kernel indices (gridding.py:24-96).

    Raises AssertionError exactly where ARL asserts (43-46, 64-65, 75-76, 83-84).
    Returns pu, pu_off, pv, pv_off, pwg, pwc, pfreq (int arrays).
    """
    n = geom.npixel
    du = 1.0 / (n * D2R * geom.cdelt0)
    dv = 1.0 / (n * D2R * geom.cdelt1)
    crpu = float(n // 2 + 1)
    crpv = float(n // 2 + 1)
    x = (uvw[:, 0] / du + crpu) - 1.0
    y = (uvw[:, 1] / dv + crpv) - 1.0
    pu = numpy.round(x).astype('int')
    pv = numpy.round(y).astype('int')
    assert numpy.min(pu) >= 0, "U axis underflows: %f" % numpy.min(pu)
    assert numpy.max(pu) < n, "U axis overflows: %f" % numpy.max(pu)
    assert numpy.min(pv) >= 0, "V axis underflows: %f" % numpy.min(pv)
    assert numpy.max(pv) < n, "V axis overflows: %f" % numpy.max(pv)
    wu_grid = du * ((pu + 1.0) - crpu) + 0.0
    wv_grid = dv * ((pv + 1.0) - crpv) + 0.0
    wu_sub = uvw[:, 0] - wu_grid
    wv_sub = uvw[:, 1] - wv_grid
    pu_off = numpy.floor((wu_sub / cfg.ddu + cfg.crpix_d) - 1.0).astype('int')
    pv_off = numpy.floor((wv_sub / cfg.ddv + cfg.crpix_d) - 1.0).astype('int')
    crpw = float(nw_grid // 2 + 1)
    pwg = numpy.round((uvw[:, 2] / wstep_grid + crpw) - 1.0).astype('int')
    assert numpy.min(pwg) >= 0, "W axis underflows"
    assert numpy.max(pwg) < cfg.nz, "W axis overflows"
    pwc = numpy.round((uvw[:, 2] / cfg.zstep + cfg.crpix_z) - 1.0).astype('int')
    assert numpy.min(pwc) >= 0, "W axis underflows"
    assert numpy.max(pwc) < cfg.nz, "W axis overflows"
    pfreq = numpy.round(((frequency - geom.freq0) / geom.dfreq + 1.0) - 1.0).astype('int')
    nchan_cf = geom.nchan if nchan_cf is None else nchan_cf
    assert numpy.min(pfreq) >= 0, "Frequency axis underflows"
    assert numpy.max(pfreq) < nchan_cf, "Frequency axis overflows"
    return pu, pu_off, pv, pv_off, pwg, pwc, pfreq


def _check_footprints(pu, pv, n, support):
    """A footprint crossing the grid edge makes ARL's slice-add fail (gridding.py:128, 264)."""
    half = support // 2
    if len(pu) and (numpy.min(pu) - half < 0 or numpy.max(pu) + half > n or
                    numpy.min(pv) - half < 0 or numpy.max(pv) + half > n):
        raise ValueError("could not broadcast: kernel footprint crosses the grid edge")


def grid_visibility_to_griddata_loop(vis, geom, cf, cfg, nw_grid=1, wstep_grid=1e15):
    """ARL's per-visibility gridding loop, unchanged in structure (gridding.py:99-133).

    vis: dict with 'uvw' [V,3] (wavelengths), 'vis' [V,npol], 'imaging_weight' [V,npol],
    'frequency' [V].  Returns (grid [nchan,npol,nw,N,N] c128, sumwt [nchan,npol]).
    """
    nchan, npol, nz, oversampling, _, support, _ = cf.shape
    sumwt = numpy.zeros([geom.nchan, npol])
    pu, pu_off, pv, pv_off, pwg, pwc, pfreq = convolution_mapping(vis['uvw'], vis['frequency'], geom, cfg,
                                                                  nw_grid=nw_grid, wstep_grid=wstep_grid)
    _check_footprints(pu, pv, geom.npixel, support)
    gv, gu = cf.shape[-2:]
    grid = numpy.zeros([geom.nchan, npol, nw_grid, geom.npixel, geom.npixel], dtype='complex')
    cfc = numpy.conjugate(cf)
    du, dv = gu // 2, gv // 2
    vw = vis['vis'] * vis['imaging_weight']
    for v, vwt, chan, uu, uuf, vv, vvf, zzg, zzc in zip(vw, vis['imaging_weight'], pfreq, pu, pu_off, pv,
                                                         pv_off, pwg, pwc):
        grid[chan, :, zzg, (vv - dv):(vv + dv), (uu - du):(uu + du)] += \
            cfc[chan, :, zzc, vvf, uuf, :, :] * v[:, numpy.newaxis, numpy.newaxis]
        sumwt[chan, :] += vwt
    return grid, sumwt


def grid_visibility_to_griddata(vis, geom, cf, cfg, nw_grid=1, chunk=1 << 16, wstep_grid=1e15):
    """Vectorised form of grid_visibility_to_griddata_loop (gridding.py:99-133).

    numpy.add.at adds the contributions to each pixel unbuffered in input order,
    i.e. in visibility order, with the same complex products as the loop, so the
    result is bitwise identical to the loop form.
    """
    nchan, npol, nz, oversampling, _, support, _ = cf.shape
    n = geom.npixel
    pu, pu_off, pv, pv_off, pwg, pwc, pfreq = convolution_mapping(vis['uvw'], vis['frequency'], geom, cfg,
                                                                  nw_grid=nw_grid, wstep_grid=wstep_grid)
    _check_footprints(pu, pv, n, support)
    cfc = numpy.conjugate(cf)
    half = support // 2
    vw = vis['vis'] * vis['imaging_weight']
    size = geom.nchan * npol * nw_grid * n * n
    re = numpy.zeros(size)
    im = numpy.zeros(size)
    a = numpy.arange(support)
    nvis = len(pu)
    for s0 in range(0, nvis, chunk):
        sl = slice(s0, min(nvis, s0 + chunk))
        for pol in range(npol):
            base = (((pfreq[sl] * npol + pol) * nw_grid + pwg[sl]) * n)
            rows = (pv[sl] - half)[:, None] + a[None, :]            # [v, S]
            cols = (pu[sl] - half)[:, None] + a[None, :]            # [v, S]
            idx = ((base[:, None, None] + rows[:, :, None]) * n + cols[:, None, :])
            vals = cfc[pfreq[sl], pol, pwc[sl], pv_off[sl], pu_off[sl], :, :] * vw[sl, pol][:, None, None]
            numpy.add.at(re, idx.ravel(), vals.real.ravel())
            numpy.add.at(im, idx.ravel(), vals.imag.ravel())
    grid = (re + 1j * im).reshape(geom.nchan, npol, nw_grid, n, n)
    sumwt = numpy.zeros([geom.nchan, npol])
    for pol in range(npol):
        sumwt[:, pol] = numpy.bincount(pfreq, weights=vis['imaging_weight'][:, pol], minlength=geom.nchan)
    return grid, sumwt


def grid_visibility_to_griddata_fast(vis, geom, cfg, nw_grid=1):
    """Nearest-pixel gridding with no kernel (gridding.py:136-158)."""
    n = geom.npixel
    npol = vis['vis'].shape[1]
    pu, pu_off, pv, pv_off, pwg, pwc, pfreq = convolution_mapping(vis['uvw'], vis['frequency'], geom, cfg,
                                                                  nw_grid=nw_grid)
    grid = numpy.zeros([geom.nchan, npol, nw_grid, n, n], dtype='complex')
    sumwt = numpy.zeros([geom.nchan, npol])
    for v, vwt, chan, xx, yy, zzg in zip(vis['vis'], vis['imaging_weight'], pfreq, pu, pv, pwg):
        grid[chan, :, zzg, yy, xx] += v * vwt
        sumwt[chan, :] += vwt
    return grid, sumwt


def degrid_visibility_from_griddata_loop(vis, grid, geom, cf, cfg, wstep_grid=1e15):
    """ARL's per-visibility degridding loop (gridding.py:233-267); no conjugation."""
    nchan, npol, nz, oversampling, _, support, _ = cf.shape
    nw_grid = grid.shape[2]
    pu, pu_off, pv, pv_off, pwg, pwc, pfreq = convolution_mapping(vis['uvw'], vis['frequency'], geom, cfg,
                                                                  nw_grid=nw_grid, wstep_grid=wstep_grid)
    _check_footprints(pu, pv, geom.npixel, support)
    gv, gu = cf.shape[-2:]
    du, dv = gu // 2, gv // 2
    out = numpy.zeros_like(vis['vis'], dtype='complex')
    for ivis in range(len(pu)):
        chan, uu, uuf, vv, vvf, zzg, zzc = (pfreq[ivis], pu[ivis], pu_off[ivis], pv[ivis], pv_off[ivis],
                                            pwg[ivis], pwc[ivis])
        out[ivis, :] += numpy.einsum('ijk,ijk->i', grid[chan, :, zzg, (vv - dv):(vv + dv), (uu - du):(uu + du)],
                                     cf[chan, :, zzc, vvf, uuf, :, :])
    return out


def degrid_visibility_from_griddata(vis, grid, geom, cf, cfg, chunk=1 << 15, wstep_grid=1e15):
    """Vectorised form of degrid_visibility_from_griddata_loop (gridding.py:233-267)."""
    nchan, npol, nz, oversampling, _, support, _ = cf.shape
    n = geom.npixel
    nw_grid = grid.shape[2]
    pu, pu_off, pv, pv_off, pwg, pwc, pfreq = convolution_mapping(vis['uvw'], vis['frequency'], geom, cfg,
                                                                  nw_grid=nw_grid, wstep_grid=wstep_grid)
    _check_footprints(pu, pv, n, support)
    half = support // 2
    a = numpy.arange(support)
    out = numpy.zeros(vis['vis'].shape, dtype='complex')
    flat = grid.reshape(-1)
    nvis = len(pu)
    for s0 in range(0, nvis, chunk):
        sl = slice(s0, min(nvis, s0 + chunk))
        for pol in range(npol):
            base = (((pfreq[sl] * npol + pol) * nw_grid + pwg[sl]) * n)
            rows = (pv[sl] - half)[:, None] + a[None, :]
            cols = (pu[sl] - half)[:, None] + a[None, :]
            idx = ((base[:, None, None] + rows[:, :, None]) * n + cols[:, None, :])
            patch = flat[idx]
            k = cf[pfreq[sl], pol, pwc[sl], pv_off[sl], pu_off[sl], :, :]
            out[sl, pol] = numpy.sum(patch * k, axis=(1, 2))
    return out


def fft_griddata_to_image(grid, gcf, imaginary=False):
    """img = ifft(sum_z G) * gcf * nx * ny; real (and imaginary) part (gridding.py:270-289)."""
    projected = numpy.sum(grid, axis=2)
    _, _, ny, nx = projected.shape
    im_data = ifft(projected) * gcf * float(nx) * float(ny)
    if imaginary:
        return im_data.real, im_data.imag
    return im_data.real


def fft_image_to_griddata(im, gcf, nw_grid=1):
    """G[:, :, z] = fft(im * gcf) for every z (gridding.py:292-302)."""
    g = fft(im * gcf)
    return numpy.repeat(g[:, :, numpy.newaxis, ...], nw_grid, axis=2)


# ---------------------------------------------------------------------------
# processing_components/imaging/base.py
# ---------------------------------------------------------------------------

def normalize_sumwt(im, sumwt):
    """img[c,p] /= sumwt[c,p] if > 0 else 0 (imaging/base.py:84-101)."""
    im = numpy.array(im, copy=True)
    nchan, npol, _, _ = im.shape
    for chan in range(nchan):
        for pol in range(npol):
            if sumwt[chan, pol] > 0.0:
                im[chan, pol, :, :] = im[chan, pol, :, :] / sumwt[chan, pol]
            else:
                im[chan, pol, :, :] = 0.0
    return im


def _copy_vis(vis):
    return {k: numpy.array(v, copy=True) for k, v in vis.items()}


def invert_2d(vis, geom, gcf, cf, cfg, dopsf=False, normalize=True, imaginary=False, loop=False, centres=None):
    """Dirty image by convolutional gridding (imaging/base.py:141-189).

    centres: None when the image is centred on the visibility phase centre (shift_vis_to_image,
    base.py:50-81, is then the identity), else (visibility phase centre, image centre) as
    (ra, dec) radians, and the rows are shifted to the image centre after dopsf's unit
    visibilities are set (base.py:160-165).
    """
    svis = _copy_vis(vis)
    if dopsf:
        svis['vis'][...] = 1.0 + 0.0j
    if centres is not None:
        svis['uvw'], svis['vis'], _ = shift_vis_to_image(svis, centres[0], centres[1])
    gridder = grid_visibility_to_griddata_loop if loop else grid_visibility_to_griddata
    grid, sumwt = gridder(svis, geom, cf, cfg)
    if imaginary:
        re, im = fft_griddata_to_image(grid, gcf, imaginary=True)
        if normalize:
            re = normalize_sumwt(re, sumwt)
            im = normalize_sumwt(im, sumwt)
        return re, sumwt, im
    re = fft_griddata_to_image(grid, gcf)
    if normalize:
        re = normalize_sumwt(re, sumwt)
    return re, sumwt


def predict_2d(vis, model, geom, gcf, cf, cfg, loop=False, centres=None):
    """Model visibilities by convolutional degridding (imaging/base.py:104-138). Returns vis array.
    centres as for invert_2d: the degridded rows are shifted back from the image centre
    (shift_vis_to_image with inverse, base.py:136)."""
    grid = fft_image_to_griddata(model, gcf)
    degridder = degrid_visibility_from_griddata_loop if loop else degrid_visibility_from_griddata
    out = degridder(vis, grid, geom, cf, cfg)
    if centres is not None:
        _, out, _ = shift_vis_to_image({'uvw': vis['uvw'], 'vis': out}, centres[0], centres[1], inverse=True)
    return out


# ---------------------------------------------------------------------------
# w-stacking: processing_components/imaging/wstack_single.py,
# processing_components/visibility/iterators.py,
# workflows/serial/imaging/imaging_serial.py
# ---------------------------------------------------------------------------

def vis_wslice_iter(w, vis_slices=1):
    """Boolean row masks |w - box| < step/2 over boxes = linspace(-wmax, wmax, P) (iterators.py:91-109)."""
    wmaxabs = numpy.max(numpy.abs(w))
    boxes = numpy.linspace(-wmaxabs, +wmaxabs, vis_slices)
    if vis_slices > 1:
        wstack = boxes[1] - boxes[0]
    else:
        wstack = 2 * wmaxabs
    for box in boxes:
        yield numpy.abs(w - box) < 0.5 * wstack


def _rows(vis, rows):
    return {k: numpy.array(v[rows], copy=True) for k, v in vis.items()}


def invert_wstack_single(vis, geom, gcf, cf, cfg, dopsf=False, normalize=True, loop=False, centres=None):
    """One w slab: shift w by the slab mean, invert, rotate by the screen (wstack_single.py:67-96).
    An image-centre shift (centres) is evaluated on the w-shifted rows, inside invert_2d."""
    vis = _copy_vis(vis)
    w_average = numpy.average(vis['uvw'][:, 2])
    vis['uvw'][..., 2] -= w_average
    re, sumwt, im = invert_2d(vis, geom, gcf, cf, cfg, dopsf=dopsf, normalize=normalize, imaginary=True,
                              loop=loop, centres=centres)
    b = create_w_term_like(geom, w_average)
    return b.real * re - b.imag * im, sumwt


def predict_wstack_single(vis, model, geom, gcf, cf, cfg, loop=False, centres=None):
    """One w slab: P(Re b * I) - 1j P(Im b * I) with w shifted by the slab mean (wstack_single.py:26-64).
    An image-centre shift (centres) is evaluated on the w-shifted rows, inside predict_2d."""
    vis = _copy_vis(vis)
    w_average = numpy.average(vis['uvw'][:, 2])
    vis['uvw'][..., 2] -= w_average
    b = create_w_term_like(geom, w_average)
    v_re = predict_2d(vis, b.real * model, geom, gcf, cf, cfg, loop=loop, centres=centres)
    v_im = predict_2d(vis, b.imag * model, geom, gcf, cf, cfg, loop=loop, centres=centres)
    return v_re - 1j * v_im


def invert_list_wstack(vis, geom, gcf, cf, cfg, vis_slices, dopsf=False, normalize=True, loop=False, centres=None):
    """Serial w-stack invert of one vis list (imaging_serial.py:181-197).

    Each slab's normalised image is re-weighted by its sumwt, summed, and the
    sum normalised by the total sumwt.
    """
    result = numpy.zeros(geom.shape)
    result_sumwt = numpy.zeros([geom.nchan, geom.npol])
    for rows in vis_wslice_iter(vis['uvw'][:, 2], vis_slices):
        if numpy.sum(rows) == 0:
            continue
        img, sumwt = invert_wstack_single(_rows(vis, rows), geom, gcf, cf, cfg, dopsf=dopsf, normalize=normalize,
                                          loop=loop, centres=centres)
        result += sumwt[:, :, numpy.newaxis, numpy.newaxis] * img
        result_sumwt += sumwt
    return normalize_sumwt(result, result_sumwt), result_sumwt


def predict_list_wstack(vis, model, geom, gcf, cf, cfg, vis_slices, loop=False, centres=None):
    """Serial w-stack predict of one vis list (imaging_serial.py:86-98).

    Rows in no slab stay zero; a row in two slabs keeps the later slab's value.
    """
    out = numpy.zeros(vis['vis'].shape, dtype='complex')
    for rows in vis_wslice_iter(vis['uvw'][:, 2], vis_slices):
        if numpy.sum(rows) == 0:
            continue
        out[rows, ...] = predict_wstack_single(_rows(vis, rows), model, geom, gcf, cf, cfg, loop=loop,
                                               centres=centres)
    return out


# ---------------------------------------------------------------------------
# Direct DFT: processing_library/util/coordinate_support.py:192-212,
# processing_components/imaging/base.py:192-244
# ---------------------------------------------------------------------------

def simulate_point(dist_uvw, l, m):
    """exp(-2 pi i uvw . (l, m, sqrt(1-l^2-m^2)-1)) (coordinate_support.py:192-212)."""
    s = numpy.array([l, m, numpy.sqrt(1 - l ** 2 - m ** 2) - 1.0])
    return numpy.exp(-2j * numpy.pi * numpy.dot(dist_uvw, s))


def predict_skycomponent_visibility(vis, components, chan_map=None):
    """Add the direct-DFT visibilities of point components (imaging/base.py:207-220).

    components: list of (l, m, flux) with flux either scalar or [nchan_img, npol].
    chan_map: image channel of each row; default ARL's get_frequency_map(vis, None) of the
    Visibility branch (processing_library/imaging/imaging_params.py:23-34): the index of each row's
    frequency among numpy.unique(frequency).  A scalar or single-channel flux applies to every row.
    """
    out = numpy.array(vis['vis'], copy=True)
    nvis, npol = out.shape
    if chan_map is None:
        if 'frequency' in vis:
            chan_map = numpy.searchsorted(numpy.unique(vis['frequency']), vis['frequency'])
        else:
            chan_map = numpy.zeros(nvis, dtype=int)
    for (l, m, flux) in components:
        phasor = simulate_point(vis['uvw'], l, m)
        f = numpy.atleast_2d(numpy.asarray(flux, dtype=float))
        if f.shape == (1, 1):
            out += f[0, 0] * phasor[:, numpy.newaxis]
        else:
            out += f[chan_map, :] * phasor[:, numpy.newaxis]
    return out


# ---------------------------------------------------------------------------
# Array geometry: processing_library/util/coordinate_support.py:37-85, 290-313;
# processing_components/visibility/base.py:50-150
# ---------------------------------------------------------------------------

def xyz_at_latitude(local_xyz, lat):
    """Local XYZ -> celestial XYZ (coordinate_support.py:37-56)."""
    x, y, z = numpy.hsplit(local_xyz, 3)
    lat2 = numpy.pi / 2 - lat
    y2 = -z * numpy.sin(lat2) + y * numpy.cos(lat2)
    z2 = z * numpy.cos(lat2) + y * numpy.sin(lat2)
    return numpy.hstack([x, y2, z2])


def xyz_to_uvw(xyz, ha, dec):
    """Rotate earth XYZ to UVW for (ha, dec) (coordinate_support.py:59-85)."""
    x, y, z = numpy.hsplit(xyz, 3)
    u = x * numpy.cos(ha) - y * numpy.sin(ha)
    v0 = x * numpy.sin(ha) + y * numpy.cos(ha)
    w = z * numpy.sin(dec) - v0 * numpy.cos(dec)
    v = z * numpy.cos(dec) + v0 * numpy.sin(dec)
    return numpy.hstack([u, v, w])


def uvw_to_xyz(uvw, ha, dec):
    """Rotate UVW for (ha, dec) back to earth XYZ (coordinate_support.py:88-113)."""
    u, v, w = numpy.hsplit(uvw, 3)
    v0 = v * numpy.sin(dec) - w * numpy.cos(dec)
    z = v * numpy.cos(dec) + w * numpy.sin(dec)
    x = u * numpy.cos(ha) + v0 * numpy.sin(ha)
    y = -u * numpy.sin(ha) + v0 * numpy.cos(ha)
    return numpy.hstack([x, y, z])


def skycoord_to_lmn(ra, dec, ra0, dec0):
    """(l, m, n) of direction (ra, dec) relative to the phase centre (ra0, dec0), radians
    (coordinate_support.py:150-169).  ARL takes astropy's SkyOffsetFrame Cartesian (x towards the
    centre, y east, z north) and returns (y, z, x - 1); written out here as the direction cosines
    of spherical trigonometry:
        x = cos d cos d0 cos(ra - ra0) + sin d sin d0,  y = cos d sin(ra - ra0),
        z = sin d cos d0 - cos d sin d0 cos(ra - ra0)."""
    dra = ra - ra0
    x = numpy.cos(dec) * numpy.cos(dec0) * numpy.cos(dra) + numpy.sin(dec) * numpy.sin(dec0)
    y = numpy.cos(dec) * numpy.sin(dra)
    z = numpy.sin(dec) * numpy.cos(dec0) - numpy.cos(dec) * numpy.sin(dec0) * numpy.cos(dra)
    return y, z, x - 1.0


def phaserotate_rows(vis, phasecentre, newphasecentre, tangent=True, inverse=False):
    """phaserotate_visibility for a row-form Visibility (visibility/base.py:295-340).

    vis: dict with 'uvw' [V, 3] wavelengths and 'vis' [V, npol]; centres are (ra, dec) radians.
    Returns (uvw, vis) of the rotated copy, or the inputs themselves when |n| < 1e-15."""
    l, m, n = skycoord_to_lmn(newphasecentre[0], newphasecentre[1], phasecentre[0], phasecentre[1])
    if numpy.abs(n) < 1e-15:
        return vis['uvw'], vis['vis']
    uvw = numpy.array(vis['uvw'], dtype=float, copy=True)
    out = numpy.array(vis['vis'], copy=True)
    phasor = simulate_point(uvw, l, m)
    if out.ndim > phasor.ndim:
        phasor = phasor[:, numpy.newaxis]
    if inverse:
        out *= phasor
    else:
        out *= numpy.conj(phasor)
    if not tangent:
        xyz = uvw_to_xyz(uvw, ha=-phasecentre[0], dec=phasecentre[1])
        uvw = xyz_to_uvw(xyz, ha=-newphasecentre[0], dec=newphasecentre[1])
    return uvw, out


def phaserotate_blocks(uvw_m, vis, frequency, phasecentre, newphasecentre, tangent=True, inverse=False):
    """phaserotate_visibility for a BlockVisibility (visibility/base.py:342-379): uvw_m
    [nt, na, na, 3] metres, vis [nt, na, na, nchan, npol]; the phasor of channel ch is
    simulate_point(uvw_m * f[ch] / c).  Returns (uvw_m, vis) of the rotated copy."""
    l, m, n = skycoord_to_lmn(newphasecentre[0], newphasecentre[1], phasecentre[0], phasecentre[1])
    if numpy.abs(n) < 1e-15:
        return uvw_m, vis
    k = numpy.asarray(frequency, dtype=float) / C_LIGHT
    uvw = uvw_m[..., numpy.newaxis] * k
    out = numpy.array(vis, copy=True)
    phasor = numpy.ones_like(out, dtype='complex')
    nchan = out.shape[3]
    for chan in range(nchan):
        phasor[:, :, :, chan, :] = simulate_point(uvw[..., chan], l, m)[..., numpy.newaxis]
    if inverse:
        out *= phasor
    else:
        out *= numpy.conj(phasor)
    uvw_out = uvw_m
    if not tangent:
        nt, na, _, _ = uvw_m.shape
        lin = uvw_m.reshape([nt * na * na, 3])
        xyz = uvw_to_xyz(lin, ha=-phasecentre[0], dec=phasecentre[1])
        uvw_out = xyz_to_uvw(xyz, ha=-newphasecentre[0], dec=newphasecentre[1]).reshape([nt, na, na, 3])
    return uvw_out, out


def pixel_to_skycoord_sin(px, py, crpix, cdelt_deg, crval_deg):
    """(ra, dec) radians of the 1-relative pixel (px, py) of a RA---SIN / DEC--SIN image without
    a rotation matrix: astropy's pixel_to_skycoord(px, py, wcs, origin=1), which
    imaging/base.py:72 applies to the FFT centre (N//2 + 1, N//2 + 1).  The intermediate world
    coordinates (l, m) = cdelt * (p - crpix), in radians, are the direction cosines of the slant
    orthographic projection about (ra0, dec0) = crval (FITS SIN with LONPOLE 180), inverted as
        n = sqrt(1 - l^2 - m^2),  dec = asin(m cos dec0 + n sin dec0),
        ra = ra0 + atan2(l, n cos dec0 - m sin dec0)."""
    ra0, dec0 = crval_deg[0] * D2R, crval_deg[1] * D2R
    l = cdelt_deg[0] * D2R * (px - crpix[0])
    m = cdelt_deg[1] * D2R * (py - crpix[1])
    if l == 0.0 and m == 0.0:
        return ra0, dec0
    n = numpy.sqrt(1.0 - l * l - m * m)
    dec = numpy.arcsin(m * numpy.cos(dec0) + n * numpy.sin(dec0))
    ra = ra0 + numpy.arctan2(l, n * numpy.cos(dec0) - m * numpy.sin(dec0))
    return ra, dec


def separation(a, b):
    """Angular distance in radians between directions (ra, dec): astropy's SkyCoord.separation,
    the Vincenty form atan2(sqrt(n1^2 + n2^2), d) of angle_utilities.angular_separation."""
    dl = b[0] - a[0]
    n1 = numpy.cos(b[1]) * numpy.sin(dl)
    n2 = numpy.cos(a[1]) * numpy.sin(b[1]) - numpy.sin(a[1]) * numpy.cos(b[1]) * numpy.cos(dl)
    d = numpy.sin(a[1]) * numpy.sin(b[1]) + numpy.cos(a[1]) * numpy.cos(b[1]) * numpy.cos(dl)
    return numpy.arctan2(numpy.hypot(n1, n2), d)


def shift_vis_to_image(vis, vis_phasecentre, image_centre, tangent=True, inverse=False):
    """imaging/base.py:50-81 for row visibilities: rotate to the image's FFT-centre direction when
    it is more than 1e-15 rad from the visibility phase centre.  Returns (uvw, vis) (the inputs
    themselves when no rotation is made) and the phase centre the result is labelled with."""
    if separation(vis_phasecentre, image_centre) > 1e-15:
        uvw, v = phaserotate_rows(vis, vis_phasecentre, image_centre, tangent=tangent, inverse=inverse)
        return uvw, v, image_centre
    return vis['uvw'], vis['vis'], vis_phasecentre


def hadec_to_azel(ha, dec, latitude):
    """(coordinate_support.py:290-313)."""
    coslat, sinlat = numpy.cos(latitude), numpy.sin(latitude)
    cosdec, sindec = numpy.cos(dec), numpy.sin(dec)
    cosha, sinha = numpy.cos(ha), numpy.sin(ha)
    az = numpy.arctan2(- cosdec * sinha, (coslat * sindec - sinlat * cosdec * cosha))
    el = numpy.arcsin(sinlat * sindec + coslat * cosdec * cosha)
    return az, el


def create_visibility(ants_xyz, times, frequency, channel_bandwidth, dec, zerow=False):
    """Visibility rows ordered (time, a1, a2>a1, chan); uvw = (pos[a2]-pos[a1]) * f/c.

    visibility/base.py:50-150 with elevation_limit=None.  ants_xyz are celestial
    XYZ (already rotated by xyz_at_latitude).  Weights are 1 (base.py:120).
    """
    nch = len(frequency)
    nants = ants_xyz.shape[0]
    rows_uvw, rows_freq, rows_a1, rows_a2, rows_time = [], [], [], [], []
    for ha in times:
        ant_pos = xyz_to_uvw(ants_xyz, ha, dec)
        a1, a2 = numpy.triu_indices(nants, k=1)
        bl = ant_pos[a2, :] - ant_pos[a1, :]
        for_chan = []
        for ch in range(nch):
            k = frequency[ch] / C_LIGHT
            for_chan.append(bl * k)
        uvw = numpy.stack(for_chan, axis=1).reshape(-1, 3)
        rows_uvw.append(uvw)
        rows_freq.append(numpy.tile(frequency, len(a1)))
        rows_a1.append(numpy.repeat(a1, nch))
        rows_a2.append(numpy.repeat(a2, nch))
        rows_time.append(numpy.full(len(a1) * nch, ha * 43200.0 / numpy.pi))
    uvw = numpy.concatenate(rows_uvw)
    if zerow:
        uvw[:, 2] = 0.0
    nrows = uvw.shape[0]
    return {
        'uvw': uvw,
        'frequency': numpy.concatenate(rows_freq),
        'antenna1': numpy.concatenate(rows_a1),
        'antenna2': numpy.concatenate(rows_a2),
        'time': numpy.concatenate(rows_time),
        'vis': numpy.zeros([nrows, 1], dtype='complex'),
        'imaging_weight': numpy.ones([nrows, 1]),
    }


def apply_gains_rows(vis_rows, antenna1, antenna2, gain):
    """Scalar gain convention V_row = g[antenna2] * conj(g[antenna1]) * V.

    calibration/operations.py:127-137 writes applied[t, i, j] = V[t, i, j] * g_i * conj(g_j)
    on the BlockVisibility; rows come from the [t, a2, a1] triangle (coalesce.py:413-416),
    so in row form i = antenna2, j = antenna1.  gain: [nants] or [nrows-aligned] complex.
    """
    return vis_rows * (gain[antenna2] * numpy.conjugate(gain[antenna1]))[:, numpy.newaxis]


# ---------------------------------------------------------------------------
# BlockVisibility layout: visibility/base.py:153-250, visibility/coalesce.py:85-143, 331-444
# ---------------------------------------------------------------------------

def create_blockvisibility(ants_xyz, times, frequency, channel_bandwidth, dec, zerow=False, integration_time=1.0):
    """BlockVisibility columns (visibility/base.py:153-250, no elevation limit): vis
    [ntimes, nants, nants, nchan, 1] zeros, uvw [ntimes, nants, nants, 3] metres with
    uvw[t, a2, a1] = pos[a2] - pos[a1] and uvw[t, a1, a2] = -(that), weights 1."""
    nants = ants_xyz.shape[0]
    nch = len(frequency)
    ntimes = len(times)
    uvw = numpy.zeros([ntimes, nants, nants, 3])
    rtimes = numpy.zeros(ntimes)
    for it, ha in enumerate(times):
        ant_pos = xyz_to_uvw(ants_xyz, ha, dec)
        rtimes[it] = ha * 43200.0 / numpy.pi
        for a1 in range(nants):
            for a2 in range(a1 + 1, nants):
                uvw[it, a2, a1, :] = ant_pos[a2, :] - ant_pos[a1, :]
                uvw[it, a1, a2, :] = ant_pos[a1, :] - ant_pos[a2, :]
    if zerow:
        uvw[..., 2] = 0.0
    shape = [ntimes, nants, nants, nch, 1]
    return {
        'vis': numpy.zeros(shape, dtype='complex'),
        'uvw': uvw,
        'weight': numpy.ones(shape),
        'imaging_weight': numpy.ones(shape),
        'time': rtimes,
        'integration_time': numpy.full(ntimes, integration_time),
        'frequency': numpy.asarray(frequency, dtype=float),
        'channel_bandwidth': numpy.asarray(channel_bandwidth, dtype=float),
    }


def convert_blocks(bvis):
    """Rows of the lower triangle (coalesce.py:331-444), vectorised as ARL does it: boolean
    masks of [t, a2, a1] with a1 < a2, numpy.argwhere order (t, a2, a1, chan),
    uvw = uvw_m * (f / c)."""
    vis, uvw = bvis['vis'], bvis['uvw']
    ntimes, nant, _, nchan, npol = vis.shape
    frequency, bw = bvis['frequency'], bvis['channel_bandwidth']
    cnvis = ntimes * nant * (nant - 1) * nchan // 2
    rowgrid = numpy.arange(ntimes * nant * nant * nchan).reshape(ntimes, nant, nant, nchan)
    cindex = numpy.zeros(rowgrid.size, dtype='int')
    lower = numpy.tril(numpy.ones((nant, nant), dtype=bool), k=-1)
    mask_rowgrid = numpy.broadcast_to(lower[None, :, :, None], rowgrid.shape)
    mask_uvw = numpy.broadcast_to(lower[None, :, :, None], uvw.shape)
    mask_vis = numpy.broadcast_to(lower[None, :, :, None, None], vis.shape)
    rg = numpy.argwhere(mask_rowgrid)
    ca2, ca1 = rg[:, 1], rg[:, 2]
    cindex[rowgrid[rg[:, 0], rg[:, 1], rg[:, 2], rg[:, 3]]] = numpy.arange(cnvis)
    nbl = nant * (nant - 1) // 2
    cfrequency = numpy.tile(frequency, ntimes * nbl)
    cbw = numpy.tile(bw, ntimes * nbl)
    ctime = numpy.repeat(bvis['time'], nchan * nbl)
    cint = numpy.repeat(bvis['integration_time'], nchan * nbl)
    cuvw = numpy.tile(uvw[mask_uvw].reshape(-1, 3), nchan).reshape(-1, 3)
    freq = numpy.repeat(cfrequency, 3).reshape(-1, 3)
    cuvw[..., :] *= freq[:] / C_LIGHT
    return {
        'vis': vis[mask_vis].reshape(-1, npol),
        'uvw': cuvw,
        'weight': bvis['weight'][mask_vis].reshape(-1, npol),
        'imaging_weight': bvis['imaging_weight'][mask_vis].reshape(-1, npol),
        'time': ctime, 'frequency': cfrequency, 'channel_bandwidth': cbw, 'integration_time': cint,
        'antenna1': ca1, 'antenna2': ca2, 'cindex': cindex,
    }


def convert_blocks_loop(bvis):
    """The same rows by an explicit loop over (t, a2, a1 < a2, chan) (the row semantics of
    coalesce.py:386-407 in argwhere order)."""
    vis, uvw = bvis['vis'], bvis['uvw']
    ntimes, nant, _, nchan, npol = vis.shape
    rows = {k: [] for k in ('vis', 'uvw', 'weight', 'imaging_weight', 'time', 'frequency', 'channel_bandwidth',
                            'integration_time', 'antenna1', 'antenna2')}
    cindex = numpy.zeros(ntimes * nant * nant * nchan, dtype='int')
    row = 0
    for t in range(ntimes):
        for a2 in range(nant):
            for a1 in range(a2):
                for ch in range(nchan):
                    f = bvis['frequency'][ch]
                    rows['vis'].append(vis[t, a2, a1, ch])
                    rows['uvw'].append(uvw[t, a2, a1] * (f / C_LIGHT))
                    rows['weight'].append(bvis['weight'][t, a2, a1, ch])
                    rows['imaging_weight'].append(bvis['imaging_weight'][t, a2, a1, ch])
                    rows['time'].append(bvis['time'][t])
                    rows['integration_time'].append(bvis['integration_time'][t])
                    rows['frequency'].append(f)
                    rows['channel_bandwidth'].append(bvis['channel_bandwidth'][ch])
                    rows['antenna1'].append(a1)
                    rows['antenna2'].append(a2)
                    cindex[((t * nant + a2) * nant + a1) * nchan + ch] = row
                    row += 1
    out = {k: numpy.array(v) for k, v in rows.items()}
    out['cindex'] = cindex
    return out


def decoalesce_visibility(rows, bvis):
    """Block element i takes row cindex[i] (coalesce.py:112-143), for vis and both weights."""
    out = {k: numpy.array(v, copy=True) for k, v in bvis.items()}
    shape = bvis['vis'].shape
    for k in ('vis', 'weight', 'imaging_weight'):
        out[k] = rows[k].reshape(-1)[rows['cindex']].reshape(shape)
    return out


def apply_gaintable(bvis_vis, gain, gain_weight, slots, inverse=False):
    """Scalar branch of apply_gaintable (calibration/operations.py:79-164, 127-140) on a
    [ntimes, nants, nants, nchan, 1] vis array, returning a new array.  gain, gain_weight:
    [ngain_times, nants, nchan]; slots[t]: gain row of integration t, or -1 (no gain row matched:
    the integration is left as it is)."""
    applied = numpy.array(bvis_vis, copy=True)
    ntimes, nant, _, nchan, _ = applied.shape
    for t in range(ntimes):
        s = slots[t]
        if s < 0:
            continue
        g = gain[s]
        if inverse:
            lgain = numpy.ones_like(g)
            nz = numpy.abs(g) > 0.0
            lgain[nz] = 1.0 / g[nz]
        else:
            lgain = g
        clgain = numpy.conjugate(lgain)
        for chan in range(nchan):
            smueller = numpy.outer(lgain[:, chan], clgain[:, chan])
            applied[t, :, :, chan, 0] = bvis_vis[t, :, :, chan, 0] * smueller
            antantwt = numpy.outer(gain_weight[s, :, chan], gain_weight[s, :, chan])
            applied[t, :, :, chan, 0][antantwt == 0.0] = 0.0
    return applied


def apply_gains_rows_adjoint(vis_rows, grad_out, idx1, idx2, gain_flat):
    """Gradients of a real loss through out = vis * g[idx2] * conj(g[idx1]), as complex numbers
    dL/dRe + i dL/dIm: (grad_vis, grad_gain)."""
    g1, g2 = gain_flat[idx1], gain_flat[idx2]
    v = vis_rows.reshape(-1)
    G = grad_out.reshape(-1)
    grad_vis = (G * numpy.conjugate(g2 * numpy.conjugate(g1))).reshape(vis_rows.shape)
    grad_gain = numpy.zeros(gain_flat.shape, dtype=complex)
    numpy.add.at(grad_gain, idx2, G * g1 * numpy.conjugate(v))
    numpy.add.at(grad_gain, idx1, numpy.conjugate(G) * g2 * v)
    return grad_vis, grad_gain


# ---------------------------------------------------------------------------
# Weighting: processing_components/griddata/gridding.py:161-230,
# processing_components/imaging/weighting.py:19-40
# ---------------------------------------------------------------------------

def grid_weight_to_griddata(vis, geom, cfg, nw_grid=1):
    """Imaging weights added at the nearest grid pixel, row by row (gridding.py:161-180).
    Returns (grid [nchan, npol, nw, N, N] complex128, sumwt [nchan, npol])."""
    npol = vis['imaging_weight'].shape[1]
    pu, _, pv, _, pwg, _, pfreq = convolution_mapping(vis['uvw'], vis['frequency'], geom, cfg, nw_grid=nw_grid)
    grid = numpy.zeros([geom.nchan, npol, nw_grid, geom.npixel, geom.npixel], dtype='complex')
    sumwt = numpy.zeros([geom.nchan, npol])
    for vwt, chan, xx, yy, zzg in zip(vis['imaging_weight'], pfreq, pu, pv, pwg):
        grid[chan, :, zzg, yy, xx] += vwt
        sumwt[chan, :] += vwt
    return grid, sumwt


def griddata_merge_weights(gd_list):
    """Sum of the (grid, sumwt) pairs (gridding.py:183-206)."""
    grid = numpy.array(gd_list[len(gd_list) // 2][0], copy=True)
    sumwt = numpy.array(gd_list[len(gd_list) // 2][1], copy=True)
    for i, g in enumerate(gd_list):
        if i != len(gd_list) // 2:
            grid += g[0]
            sumwt += g[1]
    return grid, sumwt


def griddata_reweight(vis, grid, geom, cfg, nw_grid=1):
    """imaging_weight /= Re(grid) at the row's pixel where that is non-zero (gridding.py:209-230);
    returns the new imaging_weight array.  (ARL's test `numpy.real(g).all() > 0.0` compares a
    bool with 0, so it is true exactly when every polarisation's density is non-zero.)"""
    pu, _, pv, _, pwg, _, pfreq = convolution_mapping(vis['uvw'], vis['frequency'], geom, cfg, nw_grid=nw_grid)
    wt = numpy.array(vis['imaging_weight'], dtype=float, copy=True)
    for vwt, chan, xx, yy, zzg in zip(wt, pfreq, pu, pv, pwg):
        if numpy.real(grid[chan, :, zzg, yy, xx]).all() > 0.0:
            vwt /= numpy.real(grid[chan, :, zzg, yy, xx])
    return wt


def weight_visibility(vis, geom, cfg):
    """Uniform weighting (imaging/weighting.py:19-40): returns the new imaging_weight."""
    grid, _ = grid_weight_to_griddata(vis, geom, cfg)
    return griddata_reweight(vis, grid, geom, cfg)


# ---------------------------------------------------------------------------
# Hogbom CLEAN: processing_library/arrays/cleaners.py:12-70, 134-154
# ---------------------------------------------------------------------------

def overlapIndices(res, psf, peakx, peaky):
    """Limits in res and in psf where the PSF centred on (peakx, peaky) overlaps the residual
    (cleaners.py:134-154): ((res rows, res cols), (psf rows, psf cols)) as 4-tuples."""
    nx, ny = res.shape[0], res.shape[1]
    psfwidthx, psfwidthy = psf.shape[0] // 2, psf.shape[1] // 2
    psfpeakx, psfpeaky = psf.shape[0] // 2, psf.shape[1] // 2
    res_lower = (max(0, peakx - psfwidthx), max(0, peaky - psfwidthy))
    res_upper = (min(nx, peakx + psfwidthx), min(peaky + psfwidthy, ny))
    psf_lower = (max(0, psfpeakx + (res_lower[0] - peakx)), max(0, psfpeaky + (res_lower[1] - peaky)))
    psf_upper = (min(psf.shape[0], psfpeakx + (res_upper[0] - peakx)),
                 min(psfpeaky + (res_upper[1] - peaky), psf.shape[1]))
    return ((res_lower[0], res_upper[0], res_lower[1], res_upper[1]),
            (psf_lower[0], psf_upper[0], psf_lower[1], psf_upper[1]))


def hogbom(dirty, psf, window, gain, thresh, niter, fracthresh):
    """Hogbom CLEAN of one plane (cleaners.py:12-70), ARL's loop: returns (comps, res, aiter)."""
    assert 0.0 < gain < 2.0
    assert niter > 0
    absolutethresh = max(thresh, fracthresh * numpy.fabs(dirty).max())
    comps = numpy.zeros(dirty.shape)
    res = numpy.array(dirty)
    pmax = psf.max()
    assert pmax > 0.0
    aiter = 0
    for i in range(niter):
        aiter = i + 1
        if window is not None:
            mx, my = numpy.unravel_index((numpy.fabs(res * window)).argmax(), dirty.shape)
        else:
            mx, my = numpy.unravel_index((numpy.fabs(res)).argmax(), dirty.shape)
        mval = res[mx, my] * gain / pmax
        comps[mx, my] += mval
        a1o, a2o = overlapIndices(dirty, psf, mx, my)
        res[a1o[0]:a1o[1], a1o[2]:a1o[3]] -= psf[a2o[0]:a2o[1], a2o[2]:a2o[3]] * mval
        if numpy.abs(res[mx, my]) < 0.9 * absolutethresh:
            break
    return comps, res, aiter


def deconvolve_cube_hogbom(dirty, psf, window_shape=None, window_edge=16, psf_support=None, gain=0.7,
                           threshold=0.0, niter=100, fractional_threshold=0.1):
    """deconvolve_cube with algorithm='hogbom' (processing_components/image/deconvolution.py:
    43-240) on arrays [nchan, npol, ny, nx]: windows, PSF cut, per-plane hogbom; planes whose PSF
    maximum is 0 are skipped and stay 0 in both outputs.  Returns (comp, residual, psf_cut)."""
    shape = dirty.shape
    if window_shape == 'quarter':
        qx, qy = shape[3] // 4, shape[2] // 4
        window = numpy.zeros_like(dirty)
        window[..., (qy + 1):3 * qy, (qx + 1):3 * qx] = 1.0
    elif window_shape == 'no_edge':
        nx, ny = shape[3], shape[2]
        window = numpy.zeros_like(dirty)
        window[..., (window_edge + 1):(ny - window_edge), (window_edge + 1):(nx - window_edge)] = 1.0
    else:
        window = None
    if psf_support is None:
        psf_support = max(shape[2] // 2, shape[3] // 2)
    if psf_support <= psf.shape[2] // 2 and psf_support <= psf.shape[3] // 2:
        c0, c1 = psf.shape[2] // 2, psf.shape[3] // 2
        psf = psf[..., (c0 - psf_support):(c0 + psf_support), (c1 - psf_support):(c1 + psf_support)]
    comp = numpy.zeros(shape)
    residual = numpy.zeros(shape)
    for ch in range(shape[0]):
        for pol in range(shape[1]):
            if psf[ch, pol].max():
                w = None if window is None else window[ch, pol]
                comp[ch, pol], residual[ch, pol], _ = hogbom(dirty[ch, pol], psf[ch, pol], w, gain, threshold,
                                                              niter, fractional_threshold)
    return comp, residual, psf


# ---------------------------------------------------------------------------
# Scalar gain solution: processing_components/calibration/calibration.py:30-101,
# processing_library/calibration/solvers.py:61-140, 302-332,
# processing_components/visibility/operations.py:241-264
# ---------------------------------------------------------------------------

def gain_substitution_scalar(gain, x, xwt):
    """One substitution (solvers.py:103-127); gain [nants, nchan, 1, 1]."""
    nants, nchan, nrec, _ = gain.shape
    newgain = numpy.ones_like(gain, dtype='complex')
    gwt = numpy.zeros_like(gain, dtype='float')
    x = x.reshape([nants, nants, nchan, nrec, nrec])
    xwt = xwt.reshape([nants, nants, nchan, nrec, nrec])
    xxwt = x[:, :, :, 0, 0] * xwt[:, :, :, 0, 0]
    cgain = numpy.conjugate(gain)
    gcg = gain[:, :, 0, 0] * cgain[:, :, 0, 0]
    for ant1 in range(nants):
        top = numpy.sum(gain[:, :, 0, 0] * xxwt[:, ant1, :], axis=0)
        bot = numpy.sum((gcg[:, :] * xwt[:, ant1, :, 0, 0]).real, axis=0)
        if bot.all() > 0.0:
            newgain[ant1, :, 0, 0] = top / bot
            gwt[ant1, :, 0, 0] = bot
        else:
            newgain[ant1, :, 0, 0] = 0.0
            gwt[ant1, :, 0, 0] = 0.0
    return newgain, gwt


def solution_residual_scalar(gain, x, xwt):
    """(solvers.py:302-332): the residual summed over every channel, broadcast to each."""
    nant, nchan, nrec, _ = gain.shape
    x = x.reshape(nant, nant, nchan, nrec, nrec)
    xwt = xwt.reshape(nant, nant, nchan, nrec, nrec)
    residual = numpy.zeros([nchan, nrec, nrec])
    sumwt = numpy.zeros([nchan, nrec, nrec])
    for chan in range(nchan):
        lgain = gain[:, chan, 0, 0]
        clgain = numpy.conjugate(lgain)
        smueller = numpy.outer(clgain, lgain).reshape([nant, nant])
        error = x[:, :, chan, 0, 0] - smueller
        for i in range(nant):
            error[i, i] = 0.0
        residual += numpy.sum(error * xwt[:, :, chan, 0, 0] * numpy.conjugate(error)).real
        sumwt += numpy.sum(xwt[:, :, chan, 0, 0])
    residual[sumwt > 0.0] = numpy.sqrt(residual[sumwt > 0.0] / sumwt[sumwt > 0.0])
    residual[sumwt <= 0.0] = 0.0
    return residual


def solve_antenna_gains_itsubs_scalar(gain, gwt, x, xwt, niter=30, tol=1e-8, phase_only=True, refant=0,
                                      damping=0.5):
    """(solvers.py:61-100): returns (gain, gwt, residual, iterations run)."""
    nants = x.shape[0]
    for ant1 in range(nants):
        x[ant1, ant1, ...] = 0.0
        xwt[ant1, ant1, ...] = 0.0
        for ant2 in range(ant1 + 1, nants):
            x[ant1, ant2, ...] = numpy.conjugate(x[ant2, ant1, ...])
            xwt[ant1, ant2, ...] = xwt[ant2, ant1, ...]
    for it in range(niter):
        gainLast = gain
        gain, gwt = gain_substitution_scalar(gain, x, xwt)
        if phase_only:
            mask = numpy.abs(gain) > 0.0
            gain[mask] = gain[mask] / numpy.abs(gain[mask])
        angles = numpy.angle(gain)
        gain *= numpy.exp(-1j * angles)[refant, ...]
        gain = (1.0 - damping) * gain + damping * gainLast
        change = numpy.max(numpy.abs(gain - gainLast))
        if change < tol:
            return gain, gwt, solution_residual_scalar(gain, x, xwt), it + 1
    return gain, gwt, solution_residual_scalar(gain, x, xwt), niter


def solve_gaintable_scalar(bvis, model=None, gain=None, gtime=None, interval=None, phase_only=True, niter=30,
                           tol=1e-8, normalise_gains=True):
    """solve_gaintable for npol = 1 (calibration.py:30-101) on a BlockVisibility dict ('vis',
    'weight' [nt, na, na, nchan, 1], 'time'), an optional model vis array of the same shape and a
    gain table (gain [ngain, na, nchan, 1, 1], gtime, interval).  Returns (gain, weight,
    residual [ngain, nchan, 1, 1], iterations per row)."""
    vis, wt = bvis['vis'], bvis['weight']
    nt, na, _, nchan, _ = vis.shape
    gain = numpy.array(gain, dtype=complex, copy=True)
    ngain = gain.shape[0]
    gwt_all = numpy.ones(gain.shape)
    res_all = numpy.zeros([ngain, nchan, 1, 1])
    iters = numpy.zeros(ngain, dtype=int)
    for row in range(ngain):
        rows = numpy.abs(bvis['time'] - gtime[row]) < interval[row] / 2.0
        if numpy.sum(rows) == 0:
            continue
        sv, sw = vis[rows], wt[rows]
        if model is not None:
            sm = model[rows]
            xwt_p = numpy.abs(sm) ** 2 * sw
            x_p = numpy.zeros_like(sv)
            mask = xwt_p > 0.0
            x_p[mask] = sv[mask] / sm[mask]
            x = numpy.sum(x_p * xwt_p, axis=0)
            xwt = numpy.sum(xwt_p, axis=0)
        else:
            x = numpy.sum(sv * sw, axis=0)
            xwt = numpy.sum(sw, axis=0)
        mask = numpy.abs(xwt) > 0.0
        x[mask] = x[mask] / xwt[mask]
        x[~mask] = 0.0
        g, gw, r, k = solve_antenna_gains_itsubs_scalar(gain[row], gwt_all[row], x, xwt, niter=niter, tol=tol,
                                                        phase_only=phase_only)
        gain[row], gwt_all[row], res_all[row], iters[row] = g, gw, r, k
        if normalise_gains and not phase_only:
            gain[row] /= numpy.average(numpy.abs(gain[row]))
    return gain, gwt_all, res_all, iters


# ---------------------------------------------------------------------------
# Baseline-dependent averaging: processing_components/visibility/coalesce.py:36-82, 146-328,
# processing_library/util/array_functions.py:49-118
# ---------------------------------------------------------------------------

def average_chunks(arr, wts, chunksize):
    """(array_functions.py:49-82): chunk sums of wts * arr and of wts, the former divided by the
    latter where it is > 0; chunksize <= 1 returns the inputs themselves."""
    if chunksize <= 1:
        return arr, wts
    n = len(arr)
    starts = range(0, n, chunksize)
    chunks = numpy.array([numpy.sum((wts * arr)[i:i + chunksize]) for i in starts])
    weights = numpy.array([numpy.sum(wts[i:i + chunksize]) for i in starts])
    out = chunks.copy()
    pos = weights > 0.0
    out[pos] = chunks[pos] / weights[pos]
    return out, weights


def average_chunks2(arr, wts, chunksize):
    """(array_functions.py:85-118): average_chunks along axis 1 for every row, then along axis 0
    for every column of the result, with the partial weights."""
    wts = wts.reshape(arr.shape)
    l0 = len(average_chunks(arr[:, 0], wts[:, 0], chunksize[0])[0])
    l1 = len(average_chunks(arr[0, :], wts[0, :], chunksize[1])[0])
    tempchunks = numpy.zeros([arr.shape[0], l1], dtype=arr.dtype)
    tempwt = numpy.zeros([arr.shape[0], l1])
    for i in range(arr.shape[0]):
        r = average_chunks(arr[i, :], wts[i, :], chunksize[1])
        tempchunks[i, :], tempwt[i, :] = r[0].flatten(), r[1].flatten()
    chunks = numpy.zeros([l0, l1], dtype=arr.dtype)
    weights = numpy.zeros([l0, l1])
    for i in range(l1):
        r = average_chunks(tempchunks[:, i], tempwt[:, i], chunksize[0])
        chunks[:, i], weights[:, i] = r[0].flatten(), r[1].flatten()
    return chunks, weights


def bda_factors(uvw, wts, time_coal, max_time_coal, frequency_coal, max_frequency_coal):
    """Per-baseline averaging factors of average_in_blocks (coalesce.py:198-233): uvmax over every
    element, uvdist_max over time per (a2, a1), round(coal * uvmax / uvdist_max) clipped to
    [1, max] (max where uvdist_max is 0; baselines without weight also end at 1)."""
    nant = uvw.shape[1]
    uvd = uvw[..., 0] * uvw[..., 0] + uvw[..., 1] * uvw[..., 1]
    uvmax = numpy.sqrt(numpy.max(uvd))
    uvdist_max = numpy.sqrt(numpy.max(uvd, axis=0))
    out = []
    for coal, mx in ((time_coal, max_time_coal), (frequency_coal, max_frequency_coal)):
        f = numpy.ones([nant, nant], dtype=numpy.int64)
        pos = uvdist_max > 0.0
        f[pos] = numpy.round(coal * uvmax / uvdist_max[pos])
        f[~pos] = mx
        f[numpy.sum(wts, axis=(0, 3, 4)) == 0] = 0  # putmask(..., allpwtsgrid_bool == 0, 0)
        f[f < 1] = 1
        f[f > mx] = mx
        out.append(f)
    return out[0], out[1]


def average_in_blocks(vis, uvw, wts, imaging_wts, times, integration_time, frequency, channel_bandwidth,
                      time_coal=1.0, max_time_coal=100, frequency_coal=1.0, max_frequency_coal=100):
    """(coalesce.py:146-328): rows of every (a2, a1) baseline in a2, a1 order (both triangles and the
    diagonal), each the [time chunk, frequency chunk] grid of its averages.  Returns a dict of the
    columns and cindex, whose element (t, a2, a1, ch) takes row start + ((t * nchan + ch) mod rows)
    of its baseline (ARL's flat assignment of range(rows) cycles over the ntimes x nchan indices)."""
    ntimes, nant, _, nchan, npol = vis.shape
    allp = numpy.sum(wts, axis=4)
    ta, fa = bda_factors(uvw, wts, time_coal, max_time_coal, frequency_coal, max_frequency_coal)
    frequency_grid, time_grid = numpy.meshgrid(frequency, times)
    bw_grid, it_grid = numpy.meshgrid(channel_bandwidth, integration_time)
    cols = {k: [] for k in ('vis', 'weight', 'imaging_weight', 'time', 'frequency', 'uvw', 'integration_time',
                            'channel_bandwidth', 'antenna1', 'antenna2')}
    cindex = numpy.zeros(ntimes * nant * nant * nchan, dtype=int)
    start = 0
    for a2 in range(nant):
        for a1 in range(nant):
            w2 = allp[:, a2, a1, :]
            ch = (ta[a2, a1], fa[a2, a1])

            def avg(arr):
                return average_chunks2(arr, w2, ch)[0]

            t_avg = avg(time_grid)
            rows = t_avg.size
            cols['time'].append(t_avg.flatten())
            cols['frequency'].append(avg(frequency_grid).flatten())
            cols['uvw'].append(numpy.stack([avg(numpy.outer(uvw[:, a2, a1, k], frequency / C_LIGHT)).flatten()
                                            for k in range(3)], axis=1))
            cols['integration_time'].append((avg(it_grid) * rows).flatten())
            cols['channel_bandwidth'].append((avg(bw_grid) * rows).flatten())
            pv, pw, pi = [], [], []
            for pol in range(npol):
                r = average_chunks2(vis[:, a2, a1, :, pol], wts[:, a2, a1, :, pol], ch)
                pv.append(r[0].flatten())
                pw.append(r[1].flatten())
                pi.append(average_chunks2(imaging_wts[:, a2, a1, :, pol], wts[:, a2, a1, :, pol], ch)[0].flatten())
            cols['vis'].append(numpy.stack(pv, axis=1))
            cols['weight'].append(numpy.stack(pw, axis=1))
            cols['imaging_weight'].append(numpy.stack(pi, axis=1))
            cols['antenna1'].append(numpy.full(rows, a1))
            cols['antenna2'].append(numpy.full(rows, a2))
            tt, cc = numpy.meshgrid(numpy.arange(ntimes), numpy.arange(nchan), indexing='ij')
            elem = ((tt * nant + a2) * nant + a1) * nchan + cc
            cindex[elem.ravel()] = start + (numpy.arange(ntimes * nchan) % rows)
            start += rows
    out = {k: numpy.concatenate(v) for k, v in cols.items()}
    out['cindex'] = cindex
    return out
