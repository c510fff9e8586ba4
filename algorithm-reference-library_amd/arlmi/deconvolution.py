"""Hogbom CLEAN of dirty images on the GPU (processing_library/arrays/cleaners.py:12-70) and
deconvolve_cube's hogbom branch (processing_components/image/deconvolution.py:43-240).

The minor cycle runs in libarlmi (csrc/clean.hip) in float64, as ARL's numpy arrays do:
images of any precision are promoted to float64 first, and the component and residual images
come back in float64 (ARL's deconvolve_cube builds them with numpy.zeros, i.e. float64).
"""

import ctypes

import torch

from . import _lib
from .data_models import Image
from .plan import ptr, stream_handle

__all__ = ["hogbom", "deconvolve_cube"]


def _plane(x):
    return torch.as_tensor(x).to(dtype=torch.float64).contiguous()


def hogbom(dirty, psf, window, gain, thresh, niter, fracthresh):
    """(comps, res) for one image plane (cleaners.py:12-70): 2-D dirty image and PSF (device
    tensors), window 2-D or None.  Raises ArlmiError for gain outside (0, 2), niter < 1 or a PSF
    whose maximum is not positive, as ARL's asserts do."""
    res = _plane(dirty).clone()
    if res.dim() != 2:
        raise ValueError("hogbom works on one 2-D plane")
    p = _plane(psf)
    w = None if window is None else _plane(window)
    if w is not None and w.shape != res.shape:
        raise ValueError("window shape %s differs from the image's %s" % (tuple(w.shape), tuple(res.shape)))
    comps = torch.empty_like(res)
    it = ctypes.c_int32(0)
    _lib.check(_lib.lib().arlmi_hogbom(ptr(res), ptr(comps), ptr(p), p.shape[0], p.shape[1],
                                       None if w is None else ptr(w), res.shape[0], res.shape[1], float(gain),
                                       float(thresh), int(niter), float(fracthresh), ctypes.byref(it),
                                       stream_handle()))
    hogbom.last_iterations = it.value
    return comps, res


hogbom.last_iterations = 0


def _window(dirty, window_shape, kwargs):
    """deconvolve_cube's clean windows (deconvolution.py:82-107)."""
    shape = dirty.data.shape
    if window_shape == "quarter":
        qx, qy = shape[3] // 4, shape[2] // 4
        window = torch.zeros(shape, dtype=torch.float64, device=dirty.data.device)
        window[..., (qy + 1):3 * qy, (qx + 1):3 * qx] = 1.0
    elif window_shape == "no_edge":
        edge = kwargs.get("window_edge", 16)
        nx, ny = shape[3], shape[2]
        window = torch.zeros(shape, dtype=torch.float64, device=dirty.data.device)
        window[..., (edge + 1):(ny - edge), (edge + 1):(nx - edge)] = 1.0
    elif window_shape is None:
        window = None
    else:
        raise ValueError("Window shape %s is not recognized" % window_shape)
    mask = kwargs.get("mask", None)
    if isinstance(mask, Image):
        window = mask.data
    return window


def deconvolve_cube(dirty, psf, prefix="", **kwargs):
    """(comp Image, residual Image) by Hogbom clean of each channel and polarisation
    (deconvolution.py:43-240, algorithm='hogbom'; other algorithms are not provided).

    kwargs as ARL: window_shape (None | 'quarter' | 'no_edge', window_edge), mask (Image),
    psf_support (default max(ny // 2, nx // 2): the PSF is cut to its centre +- psf_support when
    that fits), gain (0.7), threshold (0.0), niter (100), fractional_threshold (0.1).  The PSF
    image's data are replaced by the cut, as ARL does."""
    if not isinstance(dirty, Image) or not isinstance(psf, Image):
        raise AssertionError("dirty and psf must be Images")
    algorithm = kwargs.get("algorithm", "msclean")
    if algorithm != "hogbom":
        raise NotImplementedError("deconvolve_cube: only algorithm='hogbom' runs on the GPU (%r given)" % algorithm)
    window = _window(dirty, kwargs.get("window_shape", None), kwargs)
    ds = dirty.data.shape
    psf_support = kwargs.get("psf_support", max(ds[2] // 2, ds[3] // 2))
    ps = psf.data.shape
    if psf_support <= ps[2] // 2 and psf_support <= ps[3] // 2:
        c0, c1 = ps[2] // 2, ps[3] // 2
        psf.data = psf.data[..., (c0 - psf_support):(c0 + psf_support), (c1 - psf_support):(c1 + psf_support)]
    gain = kwargs.get("gain", 0.7)
    assert 0.0 < gain < 2.0, "Loop gain must be between 0 and 2"
    thresh = kwargs.get("threshold", 0.0)
    assert thresh >= 0.0
    niter = kwargs.get("niter", 100)
    assert niter > 0
    fracthresh = kwargs.get("fractional_threshold", 0.1)
    assert 0.0 < fracthresh < 1.0
    comp = torch.zeros(ds, dtype=torch.float64, device=dirty.data.device)
    residual = torch.zeros(ds, dtype=torch.float64, device=dirty.data.device)
    for channel in range(ds[0]):
        for pol in range(ds[1]):
            p = psf.data[channel, pol]
            if float(p.max()) != 0.0:
                w = None if window is None else window[channel, pol]
                comp[channel, pol], residual[channel, pol] = hogbom(dirty.data[channel, pol], p, w, gain, thresh,
                                                                    niter, fracthresh)
    return Image(comp, dirty.wcs, dirty.polarisation_frame), Image(residual, dirty.wcs, dirty.polarisation_frame)
