"""Invert and predict of one visibility set whose rows are processed in chunks (a set too large for
one plan, e.g. ~600 M rows at 8192^2), with the result of processing all rows at once.

ARL's w-stack list workflow treats the set as one Visibility: the slab centres come from max |w|
over every row (processing_components/visibility/iterators.py:99-106), each slab's w shift is the
mean w of all its rows (processing_components/imaging/wstack_single.py:43, :83), and each slab's
image sums all its rows (workflows/serial/imaging/imaging_serial.py:166-197).  Here a first pass
over the chunks gathers max |w| and the per-slab sums and counts of w (arlmi_wmax,
arlmi_wslab_sums); then one plan, rebuilt per chunk with those centres and means fixed
(arlmi_plan_fix_plane_w), grids every chunk into resident grids (arlmi_grid_add) and the planes are
transformed, screened and summed once.  Predict transforms the model once and degrids each chunk
from the resident grids.  Grids are dense (each chunk touches other tiles).
"""

import ctypes

import numpy
import torch

from . import _lib
from .data_models import device
from .imaging import _gcfcf, _real_dtype
from .plan import Plan, ptr, stream_handle, grids_buffer, wslice_boxes, kernel_struct

__all__ = ["StreamingImager"]


def _chunks(source):
    return source() if callable(source) else iter(source)


class StreamingImager:
    """One image geometry and kernel over a visibility set given as chunks.

    chunks: a list of Visibility objects, or a callable returning an iterator over them (called
    once per pass, so chunks may be produced on the fly).  mode "wstack" (vis_slices slabs, ARL's
    slab rule over the whole set) or "2d".  The image is centred on the visibilities' phase
    centre (no image-centre shift).  deterministic=True grids every chunk in 64-bit fixed point, so
    that repeated passes give bitwise-identical images.
    """

    def __init__(self, im, gcfcf=None, mode="wstack", vis_slices=1, precision=32, tile=0, chunk=0, keep_plans=False,
                 deterministic=False, band_masks=True, **kwargs):
        if mode not in ("2d", "wstack"):
            raise ValueError("mode must be '2d' or 'wstack'")
        self.gcf, self.cf = _gcfcf(im, gcfcf, kwargs)
        self.im = im
        self.mode = mode
        self.vis_slices = vis_slices if mode == "wstack" else 1
        self.precision = precision
        self.tile = tile
        self.chunk = chunk
        self.dev = device()
        self.plan = None
        self._planned = None
        # keep_plans: one plan per chunk, kept between calls (invert rebuilds them for the chunks'
        # uvw, predict reuses them), at the cost of every chunk's plan buffers
        self.keep_plans = keep_plans
        # deterministic: every chunk gridded in 64-bit fixed point (bitwise-reproducible passes)
        self.deterministic = deterministic
        # band_masks: the transforms skip the column bands no chunk's rows reach (they hold zeros);
        # the union of the chunks' bands is gathered while they are gridded (arlmi_plan_or_colmask)
        self.band_masks = band_masks
        self._bands = None        # device [nplanes * nchan] uint32 union of the last invert's chunks
        self._bands_valid = False  # the kept plans still are those the union was gathered from
        self.plans = {}
        self.grids = None
        self.boxes = None
        self.plane_w = None
        self.sumwt_planes = None
        self.gcf_dev = self.gcf.data.to(torch.float64).contiguous()

    # -- pass 1 ---------------------------------------------------------------------------
    def scan(self, chunks):
        """max |w| over all rows -> slab centres and width; per-slab sums and counts of w ->
        slab means (0 for an empty slab).  Returns (boxes, width, means, counts)."""
        L = _lib.lib()
        if self.mode != "wstack":
            self.boxes = None
            return None
        wmax = 0.0
        for v in _chunks(chunks):
            w = ctypes.c_double()
            uvw = v.data["uvw"]
            _lib.check(L.arlmi_wmax(ptr(uvw), uvw.shape[0], stream_handle(), ctypes.byref(w)))
            wmax = max(wmax, w.value)
        boxes, width = wslice_boxes(wmax, self.vis_slices)
        P = self.vis_slices
        sums = numpy.zeros(P)
        counts = numpy.zeros(P, dtype=numpy.int64)
        bptr = boxes.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        for v in _chunks(chunks):
            uvw = v.data["uvw"]
            _lib.check(L.arlmi_wslab_sums(ptr(uvw), uvw.shape[0], bptr, width, P,
                                          sums.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), stream_handle()))
        means = numpy.where(counts > 0, sums / numpy.maximum(counts, 1), 0.0)
        self.boxes = (boxes, width)
        self.plane_w = numpy.ascontiguousarray(means, dtype=numpy.float64)
        return boxes, width, self.plane_w, counts

    def _new_plan(self, vis):
        mode = _lib.MODE_WSTACK if self.mode == "wstack" else _lib.MODE_2D
        # dense grids; gridding work items only for the tiles a chunk touches (the grids are zeroed
        # once per pass, and every chunk adds into them)
        plan = Plan(vis, self.im, self.cf, mode, self.vis_slices, tile=self.tile, slab_boxes=self.boxes,
                    sparse=False, chunk=self.chunk, occupied_only=True, private=True,
                    deterministic=self.deterministic)
        if self.mode == "wstack":
            _lib.check(_lib.lib().arlmi_plan_fix_plane_w(
                plan.handle, self.plane_w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(self.plane_w)))
            plan.update(vis, self.boxes)
        return plan

    def _plan_for(self, vis, index=0, rebuild=True):
        """The plan of one chunk with the whole set's slab centres and means."""
        if self.keep_plans:
            plan = self.plans.get(index)
            if plan is None:
                plan = self.plans[index] = self._new_plan(vis)
            elif rebuild:
                plan.update(vis, self.boxes)
            self.plan = plan
            return plan
        if self.plan is not None and self._planned is vis:
            return self.plan
        self._planned = vis
        if self.plan is None:
            self.plan = self._new_plan(vis)
        else:
            self.plan.update(vis, self.boxes)
        return self.plan

    def _ensure_grids(self, vis):
        if self.grids is None:
            self.grids = grids_buffer(self.plan, self.precision, self.dev)
            self.sumwt_planes = torch.empty((self.plan.nplanes, self.plan.nchan), dtype=torch.float64,
                                            device=self.dev)

    def release(self):
        """Free the resident grids and the plans' buffers."""
        self.grids = None
        self._planned = None
        self._bands = None
        self._bands_valid = False
        for p in self.plans.values():
            p.close()
        self.plans = {}
        if self.plan is not None:
            self.plan.close()
            self.plan = None

    # -- pass 2 ---------------------------------------------------------------------------
    def invert(self, chunks, dopsf=False, normalize=True, rebuild=True):
        """Dirty image (or PSF) of all rows: [nchan, 1, N, N] in the precision, and sumwt
        [nchan, 1] (host).  normalize=False gives the list workflows' normalize=False form."""
        if self.mode == "wstack" and self.boxes is None:
            self.scan(chunks)
        L = _lib.lib()
        first = True
        self._bands_valid = False
        for i, v in enumerate(_chunks(chunks)):
            if v.precision != self.precision:
                raise TypeError("chunk precision differs from the imager's")
            plan = self._plan_for(v, i, rebuild)
            if first:
                self._ensure_grids(v)
                self.grids.zero_()
                self.sumwt_planes.zero_()
                if self.band_masks:
                    if self._bands is None:
                        self._bands = torch.zeros(self.plan.nplanes * self.plan.nchan, dtype=torch.int32,
                                                  device=self.dev)
                    self._bands.zero_()
                first = False
            if self.band_masks:
                _lib.check(L.arlmi_plan_or_colmask(plan.handle, ptr(self._bands), None, stream_handle()))
            k, keep = kernel_struct(self.cf, self.precision)
            _lib.check(L.arlmi_grid_add(plan.handle, ptr(v.data["vis"]), ptr(v.data["imaging_weight"]),
                                        int(bool(dopsf)), ctypes.byref(k), ptr(self.grids), ptr(self.sumwt_planes),
                                        self.precision, stream_handle()))
            del keep
        if first:
            raise ValueError("no chunks")
        n, nchan = self.plan.npixel, self.plan.nchan
        re = torch.empty((nchan, 1, n, n), dtype=_real_dtype(self.precision), device=self.dev)
        self._use_bands(self.plan, self.band_masks)
        try:
            _lib.check(L.arlmi_grids_to_image(self.plan.handle, ptr(self.grids), ptr(self.sumwt_planes),
                                              ptr(self.gcf_dev), 1 if normalize else 2, ptr(re), None,
                                              self.precision, stream_handle()))
        finally:
            self._use_bands(self.plan, False)
        # the union stays valid for predicts on the same kept plans
        self._bands_valid = self.band_masks and self.keep_plans
        return re, self.sumwt_planes.sum(dim=0).cpu().numpy()[:, None]

    def _use_bands(self, plan, on):
        if self._bands is not None or not on:
            _lib.check(_lib.lib().arlmi_plan_use_colmask(plan.handle, ptr(self._bands) if on else None))

    def predict(self, model_data, chunks, sink=None, out=None, rebuild=None):
        """Model visibilities of every chunk: gcf * conj(screen) * model, shifted FFTs once, then
        degridding per chunk.  Returns the list of [V, 1] complex tensors (written into the tensors
        of ``out`` when given, one per chunk), or, with ``sink``, calls sink(index, vis_chunk,
        predicted) per chunk and returns None."""
        if self.mode == "wstack" and self.boxes is None:
            self.scan(chunks)
        L = _lib.lib()
        rdt = _real_dtype(self.precision)
        model = model_data.to(device=self.dev, dtype=rdt).contiguous()
        if rebuild is None:  # kept plans were built for these chunks by the last invert
            rebuild = not self.keep_plans
        # the last invert's column bands cover these chunks when its kept plans are reused
        bands = self._bands_valid and not rebuild
        if rebuild:
            self._bands_valid = False
        result = [] if sink is None else None
        first = True
        for i, v in enumerate(_chunks(chunks)):
            plan = self._plan_for(v, i, rebuild)
            if first:
                self._ensure_grids(v)
                self._use_bands(plan, bands)
                try:
                    _lib.check(L.arlmi_image_to_grids(plan.handle, ptr(model), ptr(self.gcf_dev), ptr(self.grids),
                                                      self.precision, stream_handle()))
                finally:
                    self._use_bands(plan, False)
                first = False
            pred = out[i] if out is not None else torch.empty_like(v.vis)
            k, keep = kernel_struct(self.cf, self.precision)
            _lib.check(L.arlmi_degrid(plan.handle, ptr(self.grids), ctypes.byref(k), ptr(pred), self.precision,
                                      stream_handle()))
            del keep
            if sink is None:
                result.append(pred)
            else:
                sink(i, v, pred)
        return result
