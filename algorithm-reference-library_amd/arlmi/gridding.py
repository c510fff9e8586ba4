"""Gridding and degridding with ARL's signatures (processing_components/griddata/gridding.py).

Every function runs on the GPU through libarlmi; arrays are device tensors.
"""

import ctypes

import numpy
import torch

from . import _lib
from .data_models import GridData, Image, WCS, D2R, device
from .plan import Plan, ptr, stream_handle, kernel_struct, make_geom

__all__ = ["convolution_mapping", "grid_visibility_to_griddata", "grid_visibility_to_griddata_fast",
           "degrid_visibility_from_griddata", "fft_griddata_to_image", "fft_image_to_griddata",
           "create_griddata_from_image", "grid_planes", "degrid_planes"]


def create_griddata_from_image(im, nw=1, wstep=1e15, dtype=None):
    """GridData [nchan, npol, nw, N, N] with ARL's grid WCS (griddata/operations.py:91-145)."""
    n, ny = im.data.shape[3], im.data.shape[2]
    gw = WCS(naxis=5)
    gw.wcs.crval = numpy.array([0.0, 0.0, 0.0, im.wcs.wcs.crval[2], im.wcs.wcs.crval[3]])
    gw.wcs.crpix = numpy.array([n // 2 + 1, ny // 2 + 1, nw // 2 + 1, im.wcs.wcs.crpix[2], im.wcs.wcs.crpix[3]],
                               dtype=float)
    gw.wcs.ctype = ["UU", "VV", "WW", im.wcs.wcs.ctype[2], im.wcs.wcs.ctype[3]]
    gw.wcs.cdelt = numpy.array([1.0 / (n * D2R * im.wcs.wcs.cdelt[0]), 1.0 / (ny * D2R * im.wcs.wcs.cdelt[1]),
                                wstep, im.wcs.wcs.cdelt[2], im.wcs.wcs.cdelt[3]])
    if dtype is None:
        dtype = torch.complex64 if im.data.dtype == torch.float32 else torch.complex128
    data = torch.zeros((im.data.shape[0], im.data.shape[1], nw, ny, n), dtype=dtype, device=device())
    return GridData(data, grid_wcs=gw, projection_wcs=im.wcs.deepcopy(), polarisation_frame=im.polarisation_frame)


def _grid_nw(griddata):
    return griddata.data.shape[2], float(griddata.grid_wcs.wcs.cdelt[2])


def _image_of(griddata):
    """Image template carrying the projection WCS of a GridData."""
    shp = griddata.data.shape
    return Image(torch.empty((shp[0], shp[1], shp[3], shp[4]), device=device()), griddata.projection_wcs)


def convolution_mapping(vis, griddata, cf, channel_tolerance=1e-8):
    """Per-row grid and kernel indices, computed on the GPU (gridding.py:24-96).

    Returns (pu_grid, pu_offset, pv_grid, pv_offset, pwg_grid, None, pwc_grid, None, pfreq_grid)
    as numpy int arrays (ARL's float fractions are not produced) and raises AssertionError
    where ARL asserts.
    """
    nw, wstep = _grid_nw(griddata)
    im = _image_of(griddata)
    g = make_geom(im, cf, _lib.MODE_2D, nw, nw_grid=nw, wstep_grid=wstep)
    uvw = vis.data["uvw"]
    n = uvw.shape[0]
    out = torch.empty((n, 8), dtype=torch.int32, device=uvw.device)
    freq = ptr(vis.data["frequency"]) if "frequency" in vis.data else None
    _lib.check(_lib.lib().arlmi_convolution_mapping(ctypes.byref(g), ptr(uvw), freq, n, ptr(out), stream_handle()))
    o = out.cpu().numpy()
    err = numpy.bitwise_or.reduce(o[:, 7]) if n else 0
    assert not (err & 1), "U/V axis or kernel footprint off the grid"
    assert not (err & 2), "W axis out of range"
    assert not (err & 4), "Frequency axis out of range"
    if err & 8:
        raise IndexError("kernel sub-sample index equals the oversampling")
    return o[:, 0], o[:, 1], o[:, 2], o[:, 3], o[:, 4], None, o[:, 5], None, o[:, 6]


def grid_planes(plan, vis, cf, grids, sumwt, dopsf=False):
    """arlmi_grid over every plane of a plan (device outputs, overwritten)."""
    precision = vis.precision
    k, keep = kernel_struct(cf, precision)
    v = vis.data["vis"]
    wt = vis.data["imaging_weight"]
    if v.shape[1] != 1:
        raise NotImplementedError("Stokes I (npol = 1) only")
    _lib.check(_lib.lib().arlmi_grid(plan.handle, ptr(v), ptr(wt), int(bool(dopsf)), ctypes.byref(k), ptr(grids),
                                     ptr(sumwt), precision, stream_handle()))
    del keep


def degrid_planes(plan, grids, cf, vis_out, precision):
    k, keep = kernel_struct(cf, precision)
    _lib.check(_lib.lib().arlmi_degrid(plan.handle, ptr(grids), ctypes.byref(k), ptr(vis_out), precision,
                                       stream_handle()))
    del keep


def grid_visibility_to_griddata(vis, griddata, cf, plan=None):
    """Grid visibilities onto a GridData (gridding.py:99-133).

    Fills griddata.data in place and returns (griddata, sumwt[nchan, npol]).  Rows land on
    grid w plane round(nw//2 + w/wstep) as in ARL; the kernel is conjugated.
    """
    nw, wstep = _grid_nw(griddata)
    im = _image_of(griddata)
    if plan is None:
        plan = Plan(vis, im, cf, _lib.MODE_2D, nw, nw_grid=nw, wstep_grid=wstep)
    nchan = griddata.data.shape[0]
    grids = torch.empty((nw, nchan) + tuple(griddata.data.shape[3:]), dtype=vis.vis.dtype, device=device())
    sumwt = torch.empty((nw, nchan), dtype=torch.float64, device=device())
    grid_planes(plan, vis, cf, grids, sumwt)
    griddata.data = grids.permute(1, 0, 2, 3).unsqueeze(1).contiguous()
    return griddata, sumwt.sum(dim=0).cpu().numpy()[:, None]


def grid_visibility_to_griddata_fast(vis, griddata, cf, gcf=None):
    """Nearest-pixel gridding (gridding.py:136-158): the same as gridding with a one-tap box kernel."""
    from .kernels import create_box_convolutionfunction
    _, box = create_box_convolutionfunction(_image_of(griddata))
    # box: S=4 with the single unit tap at (2, 2) -> exactly G[pv, pu] += v * w
    return grid_visibility_to_griddata(vis, griddata, box)


def degrid_visibility_from_griddata(vis, griddata, cf, plan=None, **kwargs):
    """Degrid a GridData onto a new Visibility (gridding.py:233-267); the kernel is not conjugated."""
    nw, wstep = _grid_nw(griddata)
    im = _image_of(griddata)
    if plan is None:
        plan = Plan(vis, im, cf, _lib.MODE_2D, nw, nw_grid=nw, wstep_grid=wstep)
    grids = griddata.data[:, 0].permute(1, 0, 2, 3).contiguous()
    out = torch.empty_like(vis.vis)
    degrid_planes(plan, grids, cf, out, vis.precision)
    return vis.with_vis(out)


def _geometry_plan(im, cf_like_nz1, nplanes):
    """A plan with no rows: carries N, channels and planes for the FFT-side calls."""
    from .kernels import create_convolutionfunction_from_image
    cf = create_convolutionfunction_from_image(im, oversampling=1, support=2)
    empty = type("V", (), {})()
    empty.data = {"uvw": torch.zeros((0, 3), dtype=torch.float64, device=device())}
    return Plan(empty, im, cf, _lib.MODE_2D, nplanes)


def fft_griddata_to_image(griddata, gcf, imaginary=False):
    """img = Re(ifft(sum_z G) * gcf * N^2) (gridding.py:270-289); not normalised."""
    nw = griddata.data.shape[2]
    im = _image_of(griddata)
    plan = _geometry_plan(im, None, nw)
    grids = griddata.data[:, 0].permute(1, 0, 2, 3).contiguous()
    precision = 32 if grids.dtype == torch.complex64 else 64
    rdt = torch.float32 if precision == 32 else torch.float64
    nchan, n = grids.shape[1], grids.shape[3]
    sumwt = torch.ones((nw, nchan), dtype=torch.float64, device=device())
    re = torch.empty((nchan, 1, n, n), dtype=rdt, device=device())
    imo = torch.empty_like(re) if imaginary else None
    g = gcf.data.to(torch.float64).contiguous()
    _lib.check(_lib.lib().arlmi_grids_to_image(plan.handle, ptr(grids), ptr(sumwt), ptr(g), 0, ptr(re), ptr(imo),
                                               precision, stream_handle()))
    out_re = Image(re, griddata.projection_wcs.deepcopy(), griddata.polarisation_frame)
    if imaginary:
        return out_re, Image(imo, griddata.projection_wcs.deepcopy(), griddata.polarisation_frame)
    return out_re


def fft_image_to_griddata(im, griddata, gcf):
    """G[:, :, z] = fft(im * gcf) for every z (gridding.py:292-302); fills griddata in place."""
    nw = griddata.data.shape[2]
    plan = _geometry_plan(im, None, nw)
    precision = 32 if griddata.data.dtype == torch.complex64 else 64
    rdt = torch.float32 if precision == 32 else torch.float64
    model = im.data.to(rdt).contiguous()
    grids = torch.empty((nw, im.data.shape[0]) + tuple(griddata.data.shape[3:]), dtype=griddata.data.dtype,
                        device=device())
    g = gcf.data.to(torch.float64).contiguous()
    _lib.check(_lib.lib().arlmi_image_to_grids(plan.handle, ptr(model), ptr(g), ptr(grids), precision,
                                               stream_handle()))
    griddata.data = grids.permute(1, 0, 2, 3).unsqueeze(1).contiguous()
    return griddata
