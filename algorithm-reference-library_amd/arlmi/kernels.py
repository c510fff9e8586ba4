"""Gridding kernels and grid-correction images (processing_components/griddata/kernels.py).

The PSWF and box kernels are small tables (os x S numbers, plus an N x N correction
image) built once in float64 on the host, as ARL builds them; the gridding kernels on the
GPU consume them.  The AW (w-projection) kernel is built on the GPU: one w screen per
kernel plane (arlmi_w_screen), the image-plane taper, a batched shifted inverse FFT
(arlmi_fft2d) and a strided extraction of the os x os sub-sampled S x S kernels.
"""

import ctypes

import numpy
import torch

from . import _lib
from .data_models import WCS, ConvolutionFunction, Image, D2R, device, _radec
from .plan import ptr, stream_handle

__all__ = ["grdsf", "create_image", "create_convolutionfunction_from_image", "create_pswf_convolutionfunction",
           "convolutionfunction_from_arrays",
           "create_box_convolutionfunction", "create_awterm_convolutionfunction",
           "apply_bounding_box_convolutionfunction", "calculate_bounding_box_convolutionfunction"]

# Schwab's rational approximation of the m=6, alpha=1 spheroidal function; two ranges of nu
# (processing_library/fourier_transforms/convolutional_gridding.py:83-86).
_P = ((8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1),
      (4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2))
_Q = ((1.0000000e0, 8.212018e-1, 2.078043e-1),
      (1.0000000e0, 9.599102e-1, 2.918724e-1))


def grdsf(nu):
    """psi(nu) and (1 - nu^2) psi(nu) on |nu| <= 1, zero outside (convolutional_gridding.py:72-117)."""
    nu = numpy.abs(numpy.asarray(nu, dtype=numpy.float64))
    hi = (nu >= 0.75) & (nu <= 1.0)
    end = numpy.where(hi, 1.0, numpy.where(nu < 0.75, 0.75, 0.0))
    d = nu ** 2 - end ** 2
    p = numpy.where(hi[..., None], numpy.array(_P[1]), numpy.array(_P[0]))
    q = numpy.where(hi[..., None], numpy.array(_Q[1]), numpy.array(_Q[0]))
    top = p[..., 0].copy()
    for k in range(1, 5):
        top += p[..., k] * numpy.power(d, k)
    bot = q[..., 0].copy()
    for k in range(1, 3):
        bot += q[..., k] * numpy.power(d, k)
    psi = numpy.where(bot > 0.0, top / numpy.where(bot > 0.0, bot, 1.0), 0.0)
    psi = numpy.where(nu > 1.0, 0.0, psi)
    return psi, (1 - nu ** 2) * psi


def _coordinates(n):
    return (numpy.arange(n) - n // 2) / n


def create_image(npixel=512, cellsize=0.000015, frequency=None, channel_bandwidth=None, nchan=None,
                 dtype=torch.float32, dev=None, phasecentre=None):
    """Empty image with ARL's WCS numbers (processing_library/image/operations.py:33-69).

    cdelt = [-cell*180/pi, +cell*180/pi, 1, bandwidth], crpix = [N//2+1, N//2+1, 1, 1],
    crval = [ra, dec (degrees) of ``phasecentre``, 1, frequency].  Without a phase centre the
    image has no sky position (crval 0, 0 and ``wcs.celestial`` False): invert and predict then
    take it to be centred on the visibility phase centre.
    """
    frequency = numpy.atleast_1d(frequency if frequency is not None else numpy.array([1e8]))
    channel_bandwidth = numpy.atleast_1d(channel_bandwidth if channel_bandwidth is not None else numpy.array([1e6]))
    if nchan is None:
        nchan = len(frequency)
    w = WCS(naxis=4)
    w.wcs.cdelt = numpy.array([-cellsize * 180.0 / numpy.pi, cellsize * 180.0 / numpy.pi, 1.0,
                               float(channel_bandwidth[0])])
    w.wcs.crpix = numpy.array([npixel // 2 + 1, npixel // 2 + 1, 1.0, 1.0], dtype=float)
    w.wcs.ctype = ["RA---SIN", "DEC--SIN", "STOKES", "FREQ"]
    w.wcs.crval = numpy.array([0.0, 0.0, 1.0, float(frequency[0])])
    if phasecentre is not None:
        ra, dec = _radec(phasecentre)
        w.wcs.crval[0], w.wcs.crval[1] = numpy.rad2deg(ra), numpy.rad2deg(dec)
        w.celestial = True
    data = torch.zeros((nchan, 1, npixel, npixel), dtype=dtype, device=dev or device())
    return Image(data, w)


def create_convolutionfunction_from_image(im, nz=1, zstep=1e15, ztype="WW", oversampling=8, support=16):
    """Empty CF with ARL's 7-axis WCS (convolution_functions.py:62-147)."""
    n = im.data.shape[3]
    ny = im.data.shape[2]
    cw = WCS(naxis=7)
    cw.wcs.ctype = ["UU", "VV", "DUU", "DVV", ztype, im.wcs.wcs.ctype[2], im.wcs.wcs.ctype[3]]
    cw.wcs.crval = numpy.array([0.0, 0.0, 0.0, 0.0, 0.0, im.wcs.wcs.crval[2], im.wcs.wcs.crval[3]])
    cw.wcs.crpix = numpy.array([support // 2 + 1.0, support // 2 + 1.0, oversampling // 2 + 1.0,
                                oversampling // 2 + 1.0, float(nz // 2 + 1), im.wcs.wcs.crpix[2],
                                im.wcs.wcs.crpix[3]])
    du = 1.0 / (n * D2R * im.wcs.wcs.cdelt[0])
    dv = 1.0 / (ny * D2R * im.wcs.wcs.cdelt[1])
    cw.wcs.cdelt = numpy.array([du, dv, du / oversampling, dv / oversampling, zstep, im.wcs.wcs.cdelt[2],
                                im.wcs.wcs.cdelt[3]])
    nchan, npol = im.data.shape[0], im.data.shape[1]
    data = numpy.zeros([nchan, npol, nz, oversampling, oversampling, support, support], dtype=complex)
    return ConvolutionFunction(data, grid_wcs=cw, projection_wcs=im.wcs.deepcopy())


def _pswf_gcf(n):
    """max / (psi_i psi_j) over the image (kernels.py:87-91)."""
    g1 = grdsf(numpy.abs(2.0 * _coordinates(n)))[0]
    g = numpy.outer(g1, g1)
    pos = g > 0.0
    g[pos] = g.max() / g[pos]
    return g


def _gcf_image(im, g2):
    nchan, npol = im.data.shape[0], im.data.shape[1]
    t = torch.as_tensor(g2, dtype=torch.float64).to(device())
    data = t.reshape(1, 1, *g2.shape).expand(nchan, npol, *g2.shape).contiguous()
    return Image(data, im.wcs.deepcopy(), im.polarisation_frame)


def create_pswf_convolutionfunction(im, oversampling=8, support=6):
    """PSWF kernel and grid correction (kernels.py:52-97).

    k[s, g] = (1 - nu^2) psi(nu) at nu = ((g - S//2) - (s - os//2)/os) / (S//2), normalised so
    that sum_g k[os//2, g] = 1; CF[yo, xo] = outer(k[yo], k[xo]) / sum(CF[0, 0]).  The
    separable factor stored for the GPU is k / sqrt(sum(CF[0, 0])).
    """
    g = numpy.arange(support)
    s = numpy.arange(oversampling)
    nu = ((g[None, :] - support // 2) - (s[:, None] - oversampling // 2) / oversampling) / (support // 2)
    kernel = grdsf(nu)[1]
    kernel = kernel / numpy.sum(numpy.real(kernel[oversampling // 2, :]))
    cf = create_convolutionfunction_from_image(im, oversampling=oversampling, support=support)
    full = kernel[:, None, :, None] * kernel[None, :, None, :]  # [yo, xo, a, b]
    norm = numpy.sum(full[0, 0])
    cf.data[...] = (full / norm)[None, None, None, ...]
    cf.separable = kernel / numpy.sqrt(norm)
    return _gcf_image(im, _pswf_gcf(im.data.shape[3])), cf


def create_box_convolutionfunction(im, oversampling=1, support=1):
    """Box kernel: a delta at (..., 2, 2) of an S=4, os=1 CF; gcf = 1/sinc^2 (kernels.py:19-49)."""
    cf = create_convolutionfunction_from_image(im, oversampling=1, support=4)
    cf.data[..., 2, 2] = 1.0 + 0.0j
    n = im.data.shape[3]
    g1 = numpy.sinc(numpy.abs(_coordinates(n)))
    return _gcf_image(im, 1.0 / numpy.outer(g1, g1)), cf


def create_awterm_convolutionfunction(im, make_pb=None, nw=1, wstep=1e15, oversampling=8, support=6,
                                      use_aaf=True, maxsupport=512):
    """W-projection kernel, built on the GPU (kernels.py:100-200 with make_pb=None).

    For each kernel plane z (w = wstep*((z+1) - (nw//2+1))): the w screen of a qnx = nx/os
    pixel image with cell onx*cell/qnx, times 1/gcf_PSWF(qnx) when use_aaf, zero-padded to
    nx = max(maxsupport, 2*os*S), inverse-transformed with ARL's shifted ifft, then sampled
    at [ybeg:yend:-os, xbeg:xend:-os] for each sub-sample offset; the whole table is divided
    by sum Re CF[nw//2, os//2, os//2] and conjugated.
    """
    if make_pb is not None:
        raise NotImplementedError("A-term (primary beam) kernels need image reprojection; only w-projection "
                                  "kernels are built here")
    L = _lib.lib()
    dev = device()
    onx = im.data.shape[3]
    nchan, npol = im.data.shape[0], im.data.shape[1]
    cf = create_convolutionfunction_from_image(im, nz=nw, zstep=wstep, oversampling=oversampling,
                                               support=support)
    cf.grid_wcs.wcs.crpix[4] = nw // 2 + 1.0
    cf.grid_wcs.wcs.cdelt[4] = wstep
    if abs(wstep) > 0.0:
        w_list = [wstep * ((z + 1.0) - (nw // 2 + 1.0)) + 0.0 for z in range(nw)]
    else:
        w_list = [0.0]
    nx = max(maxsupport, 2 * oversampling * support)
    qnx = nx // oversampling
    if nx & (nx - 1) or nx < 64 or nx > 4096:
        raise NotImplementedError("padded kernel image size %d is not a power of two in [64, 4096]" % nx)
    ccell = onx * numpy.abs(D2R * im.wcs.wcs.cdelt[0]) / qnx
    sub_cell = abs(-ccell / D2R) * numpy.pi / 180.0
    if use_aaf:
        norm = torch.as_tensor(1.0 / _pswf_gcf(qnx), dtype=torch.float64).to(dev)
    else:
        norm = torch.ones((qnx, qnx), dtype=torch.float64, device=dev)
    nzp = len(w_list)
    padded = torch.zeros((nzp, nx, nx), dtype=torch.complex128, device=dev)
    screen = torch.empty((qnx, qnx), dtype=torch.complex128, device=dev)
    y0 = nx // 2 - qnx // 2
    for z, w in enumerate(w_list):
        _lib.check(L.arlmi_w_screen(qnx, sub_cell, float(w), ptr(screen), 64, stream_handle()))
        padded[z, y0:y0 + qnx, y0:y0 + qnx] = screen * norm
    _lib.check(L.arlmi_fft2d(ptr(padded), nx, nzp, -1, 64, stream_handle()))
    cen = nx // 2
    a = torch.arange(support, device=dev)
    off = torch.arange(oversampling, device=dev)
    beg = off + cen + (support * oversampling) // 2 - oversampling // 2      # [os]
    idx = beg[:, None] - oversampling * a[None, :]                             # [os, S]
    table = padded[:, idx[:, None, :, None], idx[None, :, None, :]]            # [nz, os, os, S, S]
    table = table / torch.sum(table[nw // 2 if nzp > 1 else 0, oversampling // 2, oversampling // 2].real)
    table = torch.conj(table).resolve_conj()
    cf.data = table.reshape(1, 1, *table.shape).expand(nchan, npol, *table.shape).contiguous()
    gcf = _gcf_image(im, _pswf_gcf(onx)) if use_aaf else _gcf_image(im, numpy.ones((onx, onx)))
    return gcf, cf


def apply_bounding_box_convolutionfunction(cf, fractional_level=1e-4):
    """Clip a CF to |cf| > level * max (convolution_functions.py:159-184)."""
    data = torch.as_tensor(cf.data)
    nx = data.shape[-1]
    ny = data.shape[-2]
    mag = data.abs()
    mask = mag.amax(dim=(0, 1, 2, 3, 4))
    coords = torch.nonzero(mask > fractional_level * mag.max())
    crpx = int(numpy.round(cf.grid_wcs.wcs.crpix[0]))
    crpy = int(numpy.round(cf.grid_wcs.wcs.crpix[1]))
    x0, y0 = (int(v) for v in coords.min(dim=0).values)
    dx = crpx - x0
    dy = crpy - y0
    x0 -= 1
    y0 -= 1
    x1 = crpx + dx - 1
    y1 = crpy + dy - 1
    new = ConvolutionFunction(data[..., y0:y1, x0:x1].contiguous(), grid_wcs=cf.grid_wcs.deepcopy(),
                              projection_wcs=cf.projection_wcs, polarisation_frame=cf.polarisation_frame)
    nny, nnx = new.data.shape[-2], new.data.shape[-1]
    new.grid_wcs.wcs.crpix[0] += nnx / 2 - nx / 2
    new.grid_wcs.wcs.crpix[1] += nny / 2 - ny / 2
    return new


def calculate_bounding_box_convolutionfunction(cf, fractional_level=1e-4):
    """Per-plane (z, (y0, y1), (x0, x1)) boxes (convolution_functions.py:187-207)."""
    data = torch.as_tensor(cf.data)
    mag = data.abs()
    threshold = fractional_level * mag.max()
    out = []
    for z in range(data.shape[2]):
        mask = mag[:, :, z].amax(dim=(0, 1, 2, 3))
        coords = torch.nonzero(mask > threshold)
        lo = coords.min(dim=0).values
        hi = coords.max(dim=0).values
        out.append((z, (int(lo[1]), int(hi[1])), (int(lo[0]), int(hi[0]))))
    return out


def convolutionfunction_from_arrays(im, gcf_data, cf_data, cf_crpix=None, cf_cdelt=None):
    """(gcf Image, ConvolutionFunction) on the device from another library's arrays, e.g. the
    gcfcf pair ARL's create_pswf_convolutionfunction / create_awterm_convolutionfunction return
    (kernels.py:52-200): gcf_data [nchan, npol, N, N] real, cf_data [nchan, npol, nz, os, os, S, S]
    complex, cf_crpix / cf_cdelt the 7 values of the kernel's grid WCS (ARL's
    cf.grid_wcs.wcs.crpix / .cdelt; by default those create_convolutionfunction_from_image gives a
    kernel of this shape on ``im``, with zstep 1e15).  A kernel whose taps are real and, with one w
    plane, factor exactly as CF[yo, xo, a, b] = k[yo, a] k[xo, b] (every PSWF kernel) is marked
    separable, so the gridders read 2 S numbers a row instead of S^2; any other kernel is used as a
    full table."""
    cf_np = numpy.asarray(cf_data)
    if cf_np.ndim != 7:
        raise ValueError("cf_data must be [nchan, npol, nz, os, os, S, S]")
    nz, os_, _, S = cf_np.shape[2], cf_np.shape[3], cf_np.shape[4], cf_np.shape[5]
    zstep = 1e15 if cf_cdelt is None else float(cf_cdelt[4])
    cf = create_convolutionfunction_from_image(im, nz=nz, zstep=zstep, oversampling=os_, support=S)
    if cf_crpix is not None:
        cf.grid_wcs.wcs.crpix = numpy.asarray(cf_crpix, dtype=numpy.float64).copy()
    if cf_cdelt is not None:
        cf.grid_wcs.wcs.cdelt = numpy.asarray(cf_cdelt, dtype=numpy.float64).copy()
    cf.data = cf_np.astype(numpy.complex128)
    if nz == 1 and not numpy.any(numpy.imag(cf_np)):
        # rank-1 test of M[(yo, a), (xo, b)] = CF[0, 0, 0, yo, xo, a, b]
        m = numpy.real(cf_np[0, 0, 0]).transpose(0, 2, 1, 3).reshape(os_ * S, os_ * S)
        j = int(numpy.argmax(numpy.diag(m)))
        if m[j, j] > 0.0:
            k = m[:, j] / numpy.sqrt(m[j, j])
            if numpy.max(numpy.abs(m - numpy.outer(k, k))) <= 1e-12 * numpy.max(numpy.abs(m)):
                cf.separable = k.reshape(os_, S)
    g = numpy.asarray(gcf_data, dtype=numpy.float64)
    t = torch.as_tensor(g).to(im.data.device)
    nchan, npol = im.data.shape[0], im.data.shape[1]
    t = t.reshape(-1, g.shape[-3] if g.ndim > 3 else 1, g.shape[-2], g.shape[-1])[:nchan, :npol]
    gcf = Image(t.expand(nchan, npol, *g.shape[-2:]).contiguous(), im.wcs.deepcopy(), im.polarisation_frame)
    return gcf, cf
