"""Gridding plans: the device-side index structures of one (visibility set, image, kernel).

A plan holds, for every row, ARL's convolution mapping (gridding.py:24-96) and its
w-plane membership, binned into T x T grid tiles so that the gridding kernels accumulate
in LDS and write each tile once.  Plans are built on the GPU from device uvw; the host
only supplies the WCS numbers (computed exactly as ARL computes them) and, for
w-stacking, ARL's slab centres numpy.linspace(-wmax, wmax, P) (iterators.py:91-109).
"""

import ctypes
import math

import numpy
import torch

from . import _lib
from .data_models import D2R

__all__ = ["Plan", "make_geom", "stream_handle", "ptr"]


def stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    """Device pointer of a contiguous tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("tensor must live on the GPU")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def make_geom(im, cf, mode, nplanes, nw_grid=1, wstep_grid=1e15, tile=0):
    """arlmi_geom from an Image and a ConvolutionFunction.

    UU/VV cdelt = 1/(N * d2r * cdelt_deg) (griddata/operations.py:134-135,
    convolution_functions.py:125-128), DUU/DVV = cdelt/os, crpix = N//2+1, os//2+1, nz//2+1,
    FREQ crval/cdelt of the image (operations.py:116-138).
    """
    w = im.wcs.wcs
    n = im.data.shape[3]
    ny = im.data.shape[2]
    if n != ny:
        raise ValueError("square images only")
    g = _lib.Geom()
    g.npixel = n
    g.nchan = im.data.shape[0]
    g.support = cf.support
    g.oversampling = cf.oversampling
    g.nz = cf.nz
    g.nplanes = nplanes
    g.mode = mode
    g.tile = tile
    g.du = 1.0 / (n * D2R * w.cdelt[0])
    g.dv = 1.0 / (ny * D2R * w.cdelt[1])
    g.ddu = g.du / cf.oversampling
    g.ddv = g.dv / cf.oversampling
    g.crpix_uv = float(n // 2 + 1)
    g.crpix_d = float(cf.oversampling // 2 + 1)
    cw = cf.grid_wcs.wcs if cf.grid_wcs is not None else None
    g.zstep = float(cw.cdelt[4]) if cw is not None else 1e15
    g.crpix_z = float(cw.crpix[4]) if cw is not None else float(cf.nz // 2 + 1)
    g.wstep_grid = float(wstep_grid)
    g.crpix_wg = float(nw_grid // 2 + 1)
    g.freq0 = float(w.crval[3])
    g.dfreq = float(w.cdelt[3])
    g.cellsize = abs(w.cdelt[0]) * numpy.pi / 180.0
    return g


def _all_in_channel_zero(freq, freq0, dfreq):
    """Whether ARL's frequency map (gridding.py:80-85) sends every row to image channel 0:
    round((f - f0) / df + 1 - 1) == 0 with numpy's rounding, evaluated on the device in the plan's
    own float64 operation order.  Cached on the tensor until it is modified in place."""
    key = (freq._version, freq.data_ptr(), freq.shape[0], float(freq0), float(dfreq))
    hit = getattr(freq, "_arlmi_chan0", None)
    if hit is not None and hit[0] == key:
        return hit[1]
    if freq.shape[0] == 0:
        ok = True
    else:
        p = torch.round(((freq - freq0) / dfreq + 1.0) - 1.0)
        ok = bool(((p >= 0.0) & (p < 1.0)).all())
    freq._arlmi_chan0 = (key, ok)
    return ok


def wslice_boxes(wmaxabs, vis_slices):
    """Slab centres and width of vis_wslice_iter (iterators.py:99-106)."""
    boxes = numpy.linspace(-wmaxabs, +wmaxabs, vis_slices)
    if vis_slices > 1:
        wstack = boxes[1] - boxes[0]
    else:
        wstack = 2 * wmaxabs
    return numpy.ascontiguousarray(boxes, dtype=numpy.float64), float(wstack)


def table_tile(npixel, support, precision):
    """Tile side for a full kernel table: the library default (0, i.e. 64) unless the table
    degridder's tile plus an S-wide halo (2 (T + S)^2 real values) would exceed 160 KiB of LDS, as
    for support 64 in double precision; then the largest tile that fits."""
    rb = 4 if precision == 32 else 8
    t = 64 if npixel >= 64 else npixel
    if 2 * rb * (t + support) ** 2 <= 160 * 1024:
        return 0
    while t > 16 and 2 * rb * (t + support) ** 2 > 160 * 1024:
        t //= 2
    return t


class Plan:
    """Owns an arlmi_plan handle.

    mode: _lib.MODE_2D (nplanes grid w planes, rows placed by rounding w) or
    _lib.MODE_WSTACK (nplanes w slabs, rows placed by |w - box| < width/2; slab_boxes=None
    derives ARL's linspace from max |w|; slab_boxes=(boxes, width) gives them).
    sparse: grids hold only the tiles rows touch (ARLMI_FLAG_SPARSE_GRIDS); private: grids pass only
    between this plan's calls, so predict may leave them in the library's own row order
    (ARLMI_FLAG_PRIVATE_GRIDS); both for grids that only pass between grid, grids_to_image,
    image_to_grids and degrid.  deterministic: grids and sums of weights accumulated in 64-bit fixed
    point, bitwise identical between repeated calls (ARLMI_FLAG_DETERMINISTIC).
    """

    def __init__(self, vis, im, cf, mode=_lib.MODE_2D, nplanes=1, nw_grid=1, wstep_grid=1e15, tile=0,
                 drop_off_grid=False, slab_boxes=None, use_freq=None, sparse=False, chunk=0, occupied_only=False,
                 private=False, deterministic=False):
        self.L = _lib.lib()
        self.mode = mode
        self.nplanes = nplanes
        if not tile and cf is not None and getattr(cf, "separable", None) is None and "vis" in vis.data:
            tile = table_tile(im.data.shape[-1], cf.support, 32 if vis.data["vis"].dtype == torch.complex64 else 64)
        self.geom = make_geom(im, cf, mode, nplanes, nw_grid=nw_grid, wstep_grid=wstep_grid, tile=tile)
        self.flags = ((_lib.FLAG_DROP_OFF_GRID if drop_off_grid else 0) | (_lib.FLAG_SPARSE_GRIDS if sparse else 0) |
                      (_lib.FLAG_OCCUPIED_TILES if occupied_only else 0) | (_lib.FLAG_PRIVATE_GRIDS if private else 0) |
                      (_lib.FLAG_DETERMINISTIC if deterministic else 0))
        self.handle = ctypes.c_void_p()
        self.npixel = self.geom.npixel
        self.nchan = self.geom.nchan
        self.use_freq = use_freq
        self._build(vis, slab_boxes, create=True)
        if chunk:
            _lib.check(self.L.arlmi_plan_set_chunk(self.handle, int(chunk)))
            self._build(vis, slab_boxes, create=False)

    def _freq_ptr(self, vis):
        use = self.use_freq if self.use_freq is not None else (self.nchan > 1 or self.geom.dfreq != 0.0)
        if not use or "frequency" not in vis.data:
            return None
        f = vis.data["frequency"]
        if self.nchan == 1 and _all_in_channel_zero(f, self.geom.freq0, self.geom.dfreq):
            return None  # every row maps to image channel 0: the plan need not read the column
        return ptr(f)

    def _build(self, vis, slab_boxes, create):
        uvw = vis.data["uvw"]
        if uvw.dtype != torch.float64:
            raise TypeError("uvw must be float64")
        nrows = uvw.shape[0]
        boxes_arr, width = None, 0.0
        if self.mode == _lib.MODE_WSTACK and slab_boxes is not None:
            boxes, width = slab_boxes
            boxes = numpy.ascontiguousarray(boxes, dtype=numpy.float64)
            self.boxes, self.width = boxes, width
            boxes_arr = boxes.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        # with no centres given, the library computes ARL's default ones (numpy.linspace over
        # +-max|w|) on the device
        fptr = self._freq_ptr(vis)
        if create:
            _lib.check(self.L.arlmi_plan_create(ctypes.byref(self.geom), ptr(uvw), fptr, nrows, boxes_arr, width,
                                                self.flags, stream_handle(), ctypes.byref(self.handle)))
        else:
            _lib.check(self.L.arlmi_plan_update(self.handle, ptr(uvw), fptr, nrows, boxes_arr, width,
                                                stream_handle()))
        self.nrows = nrows

    def update(self, vis, slab_boxes=None):
        """Rebuild for new uvw with the same geometry (buffers reused)."""
        self._build(vis, slab_boxes, create=False)

    def stats(self):
        s = _lib.PlanStats()
        _lib.check(self.L.arlmi_plan_stats_get(self.handle, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _lib.PlanStats._fields_}

    def grid_runs(self, support):
        """LDS add pairs the separable gridder of this support issues on the plan (runs of one
        footprint origin on origin-grouped plans, else one per entry)."""
        n = ctypes.c_int64()
        _lib.check(self.L.arlmi_plan_grid_runs(self.handle, int(support), ctypes.byref(n), stream_handle()))
        return int(n.value)

    def plane_w(self):
        out = numpy.zeros(self.nplanes)
        _lib.check(self.L.arlmi_plan_plane_w(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                             self.nplanes))
        return out

    def slab_boxes(self):
        """(centres [P], width) of a w-stack plan, as the library holds them."""
        out = numpy.zeros(self.nplanes + 1)
        _lib.check(self.L.arlmi_plan_slab_boxes(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                self.nplanes + 1))
        return out[:-1], float(out[-1])

    def close(self):
        if self.handle:
            self.L.arlmi_plan_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kernel_struct(cf, precision):
    """arlmi_kernel for a ConvolutionFunction (separable table when it has one)."""
    k = _lib.Kernel()
    k.nz = cf.nz
    k.oversampling = cf.oversampling
    k.support = cf.support
    sep = cf.separable_table(precision) if cf.nz == 1 else None
    if sep is not None:
        k.kind = _lib.KERNEL_SEPARABLE
        k.d_table = sep.data_ptr()
        keep = sep
    else:
        tab = cf.table(precision)
        k.kind = _lib.KERNEL_TABLE
        k.d_table = tab.data_ptr()
        keep = tab
    return k, keep


def grids_buffer(plan, precision, dev):
    cdt = torch.complex64 if precision == 32 else torch.complex128
    return torch.empty((plan.nplanes, plan.nchan, plan.npixel, plan.npixel), dtype=cdt, device=dev)


def is_finite_positive(x):
    return math.isfinite(x) and x > 0
