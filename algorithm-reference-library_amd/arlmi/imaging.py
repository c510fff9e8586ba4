"""invert / predict with ARL's signatures (processing_components/imaging/base.py and
processing_components/imaging/wstack_single.py), running on the GPU through libarlmi.
"""

import ctypes

import numpy
import torch

from . import _lib
from .data_models import BlockVisibility, C_LIGHT, D2R, Image, SkyCoord, Skycomponent, Visibility, device
from .kernels import create_pswf_convolutionfunction, create_image
from .plan import Plan, ptr, stream_handle, grids_buffer
from .gridding import grid_planes, degrid_planes
from .phaserotate import phaserotate_visibility, rotation_lmn, separation

__all__ = ["get_parameter", "normalize_sumwt", "shift_vis_to_image", "image_phasecentre", "invert_2d", "predict_2d",
           "invert_wstack_single", "predict_wstack_single", "predict_skycomponent_visibility",
           "create_image_from_visibility", "Imager"]


def get_parameter(kwargs, key, default=None):
    """data_models/parameters.py:79-96."""
    if kwargs is None:
        return default
    return kwargs.get(key, default)


def normalize_sumwt(im, sumwt):
    """img[c, p] /= sumwt[c, p] where sumwt > 0, else 0 (imaging/base.py:84-101)."""
    sw = torch.as_tensor(numpy.asarray(sumwt, dtype=numpy.float64), device=im.data.device)
    pos = sw > 0.0
    div = torch.where(pos, sw, torch.ones_like(sw))[:, :, None, None]
    data = torch.where(pos[:, :, None, None], im.data.to(torch.float64) / div, torch.zeros((), dtype=torch.float64,
                                                                                           device=im.data.device))
    im.data = data.to(im.data.dtype)
    return im


def image_phasecentre(im):
    """Sky direction of the image's FFT centre, the 1-relative pixel (N//2 + 1, N//2 + 1), as
    imaging/base.py:70 computes it with pixel_to_skycoord (origin 1); None for an image made
    without a sky position.  The RA---SIN / DEC--SIN intermediate coordinates
    (l, m) = cdelt * (pixel - crpix) in radians are direction cosines about crval, inverted with
    n = sqrt(1 - l^2 - m^2): dec = asin(m cos dec0 + n sin dec0),
    ra = ra0 + atan2(l, n cos dec0 - m sin dec0).  At crpix it is crval exactly."""
    if im.wcs is None or not getattr(im.wcs, "celestial", False):
        return None
    w = im.wcs.wcs
    ny, nx = im.data.shape[2], im.data.shape[3]
    l = w.cdelt[0] * D2R * ((nx // 2 + 1) - w.crpix[0])
    m = w.cdelt[1] * D2R * ((ny // 2 + 1) - w.crpix[1])
    if l == 0.0 and m == 0.0:
        return im.phasecentre
    ra0, dec0 = w.crval[0] * D2R, w.crval[1] * D2R
    n = numpy.sqrt(1.0 - l * l - m * m)
    dec = numpy.arcsin(m * numpy.cos(dec0) + n * numpy.sin(dec0))
    ra = ra0 + numpy.arctan2(l, n * numpy.cos(dec0) - m * numpy.sin(dec0))
    return SkyCoord(ra, dec)


def _shift_lmn(vis, im):
    """The rotation shift_vis_to_image makes for (vis, im) as [l, m, n - 1] float64, or None: no
    sky position on either side, centres within 1e-15 rad (imaging/base.py:71), or |n| < 1e-15."""
    centre = image_phasecentre(im)
    if centre is None or vis.phasecentre is None:
        return None
    if separation(vis.phasecentre, centre) <= 1e-15:
        return None
    return rotation_lmn(centre, vis.phasecentre)


def _rotate_rows(uvw, data, lmn, inverse, precision):
    """Multiply row visibilities ``data`` [V, npol] in place by the conjugate phasor of the
    direction lmn (by the phasor itself for ``inverse``) evaluated on ``uvw`` [V, 3], which is only
    read: the tangent-plane rotation of phaserotate_visibility (visibility/base.py:318-340)."""
    npol = data.shape[1] if data.dim() > 1 else 1
    _lib.check(_lib.lib().arlmi_phaserotate_rows(ptr(uvw), ptr(data), data.shape[0], npol,
                                                 lmn.ctypes.data_as(ctypes.c_void_p), None, int(bool(inverse)),
                                                 precision, stream_handle()))


def _labelled(out, vis, im):
    """Predicted visibilities carry the image's phase centre when a shift was made, as
    shift_vis_to_image(inverse=True) leaves them (imaging/base.py:71-80)."""
    centre = image_phasecentre(im)
    if centre is not None and vis.phasecentre is not None and separation(vis.phasecentre, centre) > 1e-15:
        out.phasecentre = im.phasecentre
    return out


def shift_vis_to_image(vis, im, tangent=True, inverse=False):
    """Shift visibilities to the image's FFT-centre direction (imaging/base.py:50-81).

    When that direction is more than 1e-15 rad from the visibility phase centre, the result is
    phaserotate_visibility(vis, centre, tangent, inverse) with its phase centre set to the
    image's (for ``inverse`` too, as ARL labels it); otherwise ``vis`` itself.  A visibility or
    image without a sky position is returned unchanged."""
    if not isinstance(vis, (Visibility, BlockVisibility)):
        raise ValueError("vis is not a Visibility or BlockVisibility: %r" % (vis,))
    centre = image_phasecentre(im)
    if centre is None or vis.phasecentre is None:
        return vis
    if separation(vis.phasecentre, centre) > 1e-15:
        vis = phaserotate_visibility(vis, centre, tangent=tangent, inverse=inverse)
        vis.phasecentre = im.phasecentre
    return vis


def _gcfcf(im, gcfcf, kwargs):
    if gcfcf is None:
        return create_pswf_convolutionfunction(im, support=get_parameter(kwargs, "support", 6),
                                               oversampling=get_parameter(kwargs, "oversampling", 128))
    return gcfcf


def _real_dtype(precision):
    return torch.float32 if precision == 32 else torch.float64


class Imager:
    """invert and predict for one visibility set, image geometry and kernel.

    The plan (convolution mapping, w-plane membership, tile lists) is built once on the GPU
    and reused by invert and predict; ``update(vis)`` rebuilds it for new uvw.
      mode "2d":     planes are the grid w planes (one by default), as invert_2d / predict_2d;
      mode "wstack": vis_slices w slabs with ARL's slab rule, screens and slab-mean w shift,
                     as the w-stack list workflows (imaging_serial.py:41-216).
    deterministic=True grids in 64-bit fixed point: repeated inverts of the same rows give
    bitwise-identical images and sums of weights (predict is deterministic in either mode).
    """

    def __init__(self, vis, im, gcfcf=None, mode="2d", vis_slices=1, slab_boxes=None, tile=0,
                 drop_off_grid=False, sparse=True, chunk=0, deterministic=False, **kwargs):
        self.gcf, self.cf = _gcfcf(im, gcfcf, kwargs)
        self.im = im
        self.precision = vis.precision
        # the imager's grids never leave it, so they can be sparse (tiles no row touches are
        # neither written nor read)
        if mode == "2d":
            self.plan = Plan(vis, im, self.cf, _lib.MODE_2D, 1, tile=tile, drop_off_grid=drop_off_grid,
                             sparse=sparse, chunk=chunk, private=True, deterministic=deterministic)
        elif mode == "wstack":
            self.plan = Plan(vis, im, self.cf, _lib.MODE_WSTACK, vis_slices, tile=tile, drop_off_grid=drop_off_grid,
                             slab_boxes=slab_boxes, sparse=sparse, chunk=chunk, private=True,
                             deterministic=deterministic)
        else:
            raise ValueError("mode must be '2d' or 'wstack'")
        self.mode = mode
        self.dev = device()
        self.grids = grids_buffer(self.plan, self.precision, self.dev)
        self.sumwt_planes = torch.empty((self.plan.nplanes, self.plan.nchan), dtype=torch.float64, device=self.dev)
        self.gcf_dev = self.gcf.data.to(torch.float64).contiguous()

    def update(self, vis, slab_boxes=None):
        self.plan.update(vis, slab_boxes)
        self._shift_cache = None

    def invert(self, vis, dopsf=False, normalize=True, imaginary=False, out=None, weigh_slabs=False):
        """Dirty image (or PSF): grid, shifted inverse FFT, screens, gcf and sum-of-weights.

        normalize=False returns the plain sum over planes (invert_2d / invert_wstack_single with
        normalize=False); with weigh_slabs it returns what the list workflows return for
        normalize=False instead: each slab's unnormalised image weighted by the slab's sum of
        weights, the sum divided by the total (imaging_serial.py:166-197)."""
        if vis.precision != self.precision:
            raise TypeError("visibility precision changed")
        lmn = _shift_lmn(vis, self.im)
        if lmn is not None:  # shift_vis_to_image on a copy (imaging/base.py:160-165)
            rot = torch.ones_like(vis.vis) if dopsf else vis.vis.clone()
            _rotate_rows(self._shift_uvw(vis), rot, lmn, False, self.precision)
            vis, dopsf = vis.with_vis(rot), False
        grid_planes(self.plan, vis, self.cf, self.grids, self.sumwt_planes, dopsf=dopsf)
        return self.grids_to_image(normalize=normalize, imaginary=imaginary, out=out, weigh_slabs=weigh_slabs)

    def grids_to_image(self, normalize=True, imaginary=False, out=None, weigh_slabs=False):
        """The image of the gridded planes (the second half of invert): shifted inverse FFTs,
        screens, plane sum, gcf and the sum-of-weights normalisation.  The grids are used as
        scratch."""
        n, nchan = self.plan.npixel, self.plan.nchan
        rdt = _real_dtype(self.precision)
        re = out if out is not None else torch.empty((nchan, 1, n, n), dtype=rdt, device=self.dev)
        imo = torch.empty_like(re) if imaginary else None
        _lib.check(_lib.lib().arlmi_grids_to_image(self.plan.handle, ptr(self.grids), ptr(self.sumwt_planes),
                                                   ptr(self.gcf_dev), 1 if normalize else (2 if weigh_slabs else 0),
                                                   ptr(re), ptr(imo),
                                                   self.precision, stream_handle()))
        return re, imo

    def image_to_grids(self, model_data):
        """The planes predict degrids (the first half of predict): gcf * conj(screen) * model,
        shifted FFTs.  Returns the grids written."""
        rdt = _real_dtype(self.precision)
        model = model_data.to(device=self.dev, dtype=rdt).contiguous()
        grids = getattr(self, "_pgrids", None)
        grids = self.grids if grids is None else grids
        _lib.check(_lib.lib().arlmi_image_to_grids(self.plan.handle, ptr(model), ptr(self.gcf_dev), ptr(grids),
                                                   self.precision, stream_handle()))
        return grids

    def adjoint(self, vis, vis_data, out=None):
        """Adjoint of predict for a real image: Re(sum over planes of screen * gcf * F^H grid(v)),
        which is invert with unit weights and no normalisation (F^H = N^2 ifft2 with the shifts,
        grid with the conjugated kernel).  vis_data: [V, 1] complex rows of ``vis``."""
        ones = self._unit_weights(vis)
        v = vis.with_vis(vis_data.to(vis.vis.dtype).contiguous())
        v.data["imaging_weight"] = ones
        re, _ = self.invert(v, dopsf=False, normalize=False, out=out)
        return re

    def _unit_weights(self, vis):
        w = getattr(self, "_ones", None)
        if w is None or w.shape[0] != vis.nvis:
            w = torch.ones((vis.nvis, 1), dtype=_real_dtype(self.precision), device=self.dev)
            self._ones = w
        return w

    def sumwt(self):
        """[nchan, 1] float64 (host) sum of weights of the last invert."""
        return self.sumwt_planes.sum(dim=0).cpu().numpy()[:, None]

    def separate_predict_grids(self, on=True):
        """Give predict grids of its own, so that a predict and an invert on the same plan may run
        at the same time on two streams (they share nothing else the calls write).  Costs one more
        grid buffer."""
        self._pgrids = grids_buffer(self.plan, self.precision, self.dev) if on else None

    def predict(self, vis, model_data, out=None):
        """Model visibilities: gcf * conj(screen) * model, shifted FFT, degridding."""
        grids = self.image_to_grids(model_data)
        vis_out = out if out is not None else torch.empty_like(vis.vis)
        degrid_planes(self.plan, grids, self.cf, vis_out, self.precision)
        lmn = _shift_lmn(vis, self.im)
        if lmn is not None:  # back from the image centre (imaging/base.py:135-136)
            _rotate_rows(self._shift_uvw(vis), vis_out, lmn, True, self.precision)
        return vis_out

    def _shift_uvw(self, vis):
        """The uvw the image-centre shift is evaluated on: the rows' own in 2-D mode; in w-stack
        mode w less the mean w of the row's slab (rows placed by |w - box| < width/2), since
        invert_wstack_single / predict_wstack_single subtract it before invert_2d / predict_2d
        shift (wstack_single.py:43-47, 83-86)."""
        uvw = vis.data["uvw"]
        if self.mode != "wstack":
            return uvw
        # kept until the plan is rebuilt: the slab centres and means are read back to the host,
        # which a call under stream capture (after an eager call) must not do
        cache = getattr(self, "_shift_cache", None)
        if cache is not None and cache[0] is uvw:
            return cache[1]
        boxes, width = self.plan.slab_boxes()
        wbar = self.plan.plane_w()
        w = uvw[:, 2]
        off = torch.zeros_like(w)
        for box, mean in zip(boxes.tolist(), wbar.tolist()):
            off = torch.where(torch.abs(w - box) < 0.5 * width, torch.full_like(w, mean), off)
        out = uvw.clone()
        out[:, 2] = w - off
        self._shift_cache = (uvw, out)
        return out


def invert_2d(vis, im, dopsf=False, normalize=True, gcfcf=None, **kwargs):
    """Dirty image by convolutional gridding (imaging/base.py:141-189).

    Returns (Image, sumwt[nchan, npol]), or (re, sumwt, im) with imaginary=True.
    """
    imager = Imager(vis, im, gcfcf=gcfcf, mode="2d", **kwargs)
    imaginary = get_parameter(kwargs, "imaginary", False)
    re, imo = imager.invert(vis, dopsf=dopsf, normalize=normalize, imaginary=imaginary)
    sumwt = imager.sumwt()
    out_re = Image(re, im.wcs.deepcopy(), im.polarisation_frame)
    if imaginary:
        return out_re, sumwt, Image(imo, im.wcs.deepcopy(), im.polarisation_frame)
    return out_re, sumwt


def predict_2d(vis, model, gcfcf=None, **kwargs):
    """Model visibilities by convolutional degridding (imaging/base.py:104-138); returns a new
    Visibility."""
    if model is None:
        return vis
    imager = Imager(vis, model, gcfcf=gcfcf, mode="2d", **kwargs)
    return _labelled(vis.with_vis(imager.predict(vis, model.data)), vis, model)


def invert_wstack_single(vis, im, dopsf, normalize=True, remove=True, gcfcf=None, **kwargs):
    """One w slab: all rows, w shifted by their mean, screen applied (wstack_single.py:67-96)."""
    if not remove:
        raise NotImplementedError("remove=False is not supported")
    imager = Imager(vis, im, gcfcf=gcfcf, mode="wstack", vis_slices=1,
                    slab_boxes=(numpy.zeros(1), float("inf")), **kwargs)
    re, _ = imager.invert(vis, dopsf=dopsf, normalize=normalize)
    return Image(re, im.wcs.deepcopy(), im.polarisation_frame), imager.sumwt()


def predict_wstack_single(vis, model, remove=True, gcfcf=None, **kwargs):
    """One w slab predict: P(conj(b) * I) with w shifted by the mean (wstack_single.py:26-64)."""
    if not remove:
        raise NotImplementedError("remove=False is not supported")
    imager = Imager(vis, model, gcfcf=gcfcf, mode="wstack", vis_slices=1,
                    slab_boxes=(numpy.zeros(1), float("inf")), **kwargs)
    return _labelled(vis.with_vis(imager.predict(vis, model.data)), vis, model)


def predict_skycomponent_visibility(vis, sc, chan_map=None):
    """Add the direct-DFT visibilities of point components (imaging/base.py:192-220), on the GPU.

    sc: Skycomponent or list of them, directions given as (l, m); flux [nchan_img, npol].
    chan_map: optional int32 device tensor, the image channel of each row.  By default, for fluxes
    of several channels, ARL's map (get_frequency_map(vis, None),
    processing_library/imaging/imaging_params.py:23-34): the index of each row's frequency among
    the sorted distinct frequencies.  A single-channel flux applies to every row.  A map outside the
    flux's channels raises IndexError, as numpy's comp.flux[map] would.
    """
    if sc is None:
        return vis
    if isinstance(sc, Skycomponent):
        sc = [sc]
    if not sc:
        return vis
    nchan = max(c.flux.shape[0] for c in sc)
    lmf = numpy.zeros((len(sc), 2 + nchan))
    for i, c in enumerate(sc):
        lmf[i, 0], lmf[i, 1] = c.l, c.m
        f = c.flux[:, 0]
        lmf[i, 2:] = f if len(f) == nchan else f[0]
    d_lmf = torch.as_tensor(lmf).to(device())
    out = vis.vis.clone()
    uvw = vis.data["uvw"]
    if chan_map is not None:
        chan = chan_map.to(device=device(), dtype=torch.int32).contiguous()
    elif nchan > 1:
        _, inv = torch.unique(vis.data["frequency"], sorted=True, return_inverse=True)
        chan = inv.to(torch.int32).contiguous()
    else:
        chan = None
    if chan is not None and chan.numel():
        lo, hi = int(chan.min().item()), int(chan.max().item())
        if lo < 0 or hi >= nchan:
            raise IndexError("frequency map reaches image channels %d..%d; the component fluxes have %d"
                             % (lo, hi, nchan))
    per_call = max(1, (64 * 1024) // (8 * (3 + nchan)))  # components staged in one block's LDS
    for k0 in range(0, len(sc), per_call):
        part = d_lmf[k0:k0 + per_call]
        _lib.check(_lib.lib().arlmi_dft(ptr(uvw), ptr(chan), uvw.shape[0], ptr(part), part.shape[0], nchan,
                                        ptr(out), vis.precision, stream_handle()))
    vis.data["vis"] = out
    return vis


def create_image_from_visibility(vis, **kwargs):
    """Empty image for a visibility set (imaging/base.py:247-339), on the visibilities' device.

    Reference frequency: the first of ``frequency`` (default the visibilities' frequencies, per
    row for a Visibility, per channel for a BlockVisibility); nchan defaults to the number of
    distinct frequencies, and an image of several channels needs several in the data (else
    ValueError, as ARL's unknown spectral mode).  Cell: half the critical cell 1/(2 max|u|) by
    default, capped at the critical cell unless override_cellsize=False, with u in wavelengths
    (metres times max(frequency)/c for a BlockVisibility).  Centred on the ``phasecentre``
    keyword, by default the visibility phase centre (base.py:269, 333)."""
    if not isinstance(vis, (Visibility, BlockVisibility)):
        raise AssertionError("vis is not a Visibility or a BlockVisibility: %r" % (vis,))
    uvw = vis.data["uvw"]
    if isinstance(vis, BlockVisibility):
        vfreq = numpy.asarray(vis.frequency, dtype=numpy.float64)
        vbw = numpy.atleast_1d(numpy.asarray(vis.channel_bandwidth, dtype=numpy.float64))
    else:
        vfreq = vis.data["frequency"].cpu().numpy()
        vbw = vis.data["channel_bandwidth"][0:1].cpu().numpy() if "channel_bandwidth" in vis.data else \
            numpy.array([1e6])
    vnchan = len(numpy.unique(vfreq))
    frequency = numpy.atleast_1d(numpy.asarray(get_parameter(kwargs, "frequency", vfreq), dtype=numpy.float64))
    nchan = get_parameter(kwargs, "nchan", vnchan)
    channel_bandwidth = get_parameter(kwargs, "channel_bandwidth", 0.99999999999 * float(vbw[0]))
    # spectral modes (base.py:280-300): one image channel per data channel, or MFS / regridded
    # channels, which need a non-zero channel width
    if not (nchan == vnchan and vnchan > 1):
        if not ((nchan == 1 and vnchan >= 1) or (nchan > 1 and vnchan > 1)):
            raise ValueError("create_image_from_visibility: unknown spectral mode ")
        assert numpy.abs(channel_bandwidth) > 0.0, "Channel width must be non-zero for mfs mode"
    npixel = get_parameter(kwargs, "npixel", 512)
    uvmax = float(torch.max(torch.abs(uvw[..., 0:1])).item())
    if isinstance(vis, BlockVisibility):
        uvmax *= numpy.max(frequency) / C_LIGHT
    critical = 1.0 / (uvmax * 2.0)
    cellsize = get_parameter(kwargs, "cellsize", 0.5 * critical)
    if get_parameter(kwargs, "override_cellsize", True) and cellsize > critical:
        cellsize = critical
    dtype = torch.float32 if vis.precision == 32 else torch.float64
    phasecentre = get_parameter(kwargs, "phasecentre", vis.phasecentre)
    return create_image(npixel=npixel, cellsize=cellsize, frequency=frequency[:1],
                        channel_bandwidth=numpy.atleast_1d(channel_bandwidth), nchan=nchan, dtype=dtype,
                        phasecentre=phasecentre, dev=uvw.device)
