"""Visibility weighting on the GPU: gridded weight density and uniform reweighting
(processing_components/griddata/gridding.py:161-230, processing_components/imaging/weighting.py:19-40,
workflows/serial/imaging/imaging_serial.py:426-473)."""

import ctypes

import numpy
import torch

from . import _lib
from .data_models import GridData, device
from .gridding import create_griddata_from_image, _image_of, _grid_nw
from .plan import ptr, stream_handle, make_geom

__all__ = ["grid_weight_to_griddata", "griddata_merge_weights", "griddata_reweight", "weight_visibility",
           "weight_list_serial_workflow"]


def _geom(griddata, cf):
    nw, wstep = _grid_nw(griddata)
    return make_geom(_image_of(griddata), cf, _lib.MODE_2D, nw, nw_grid=nw, wstep_grid=wstep)


def _freq(vis):
    return ptr(vis.data["frequency"]) if "frequency" in vis.data else None


def grid_weight_to_griddata(vis, griddata, cf):
    """Imaging weights summed at each row's nearest grid pixel (gridding.py:161-180).

    Fills griddata.data ([nchan, 1, nw, N, N] complex128, the density in the real part, as ARL's
    complex grid holds it) and returns (griddata, sumwt[nchan, 1])."""
    if vis.npol != 1:
        raise NotImplementedError("Stokes I (npol = 1) only")
    g = _geom(griddata, cf)
    nchan, nw, n = griddata.data.shape[0], griddata.data.shape[2], griddata.data.shape[-1]
    dev = vis.data["uvw"].device
    density = torch.empty((nchan, nw, n, n), dtype=torch.float64, device=dev)
    sumwt = torch.empty(nchan, dtype=torch.float64, device=dev)
    wt = vis.data["imaging_weight"].contiguous()
    _lib.check(_lib.lib().arlmi_grid_weights(ctypes.byref(g), ptr(vis.data["uvw"]), _freq(vis), vis.nvis, ptr(wt),
                                             ptr(density), ptr(sumwt), vis.precision, stream_handle()))
    griddata.data = density.to(torch.complex128).unsqueeze(1)
    return griddata, sumwt.cpu().numpy()[:, None]


def griddata_merge_weights(gd_list, algorithm="uniform"):
    """Sum of weight grids and sums of weights (gridding.py:183-206); the result carries the
    centre grid's WCS with the summed bandwidth and mean frequency.  The inputs are not changed."""
    centre = len(gd_list) // 2
    c = gd_list[centre][0]
    gd = GridData(c.data.clone(), grid_wcs=c.grid_wcs.deepcopy(), projection_wcs=c.projection_wcs.deepcopy(),
                  polarisation_frame=c.polarisation_frame)
    sumwt = numpy.array(gd_list[centre][1], dtype=float, copy=True)
    frequency = 0.0
    bandwidth = 0.0
    for i, g in enumerate(gd_list):
        if i != centre:
            gd.data += g[0].data
            sumwt += g[1]
        frequency += g[0].grid_wcs.wcs.crval[4]
        bandwidth += g[0].grid_wcs.wcs.cdelt[4]
    gd.grid_wcs.wcs.cdelt[4] = bandwidth
    gd.grid_wcs.wcs.crval[4] = frequency / len(gd_list)
    return gd, sumwt


def griddata_reweight(vis, griddata, cf):
    """Divide every row's imaging weight by the gridded density at its pixel where that is
    non-zero (gridding.py:209-230), in place; returns vis."""
    g = _geom(griddata, cf)
    density = griddata.data[:, 0].real.to(torch.float64).contiguous()
    wt = vis.data["imaging_weight"]
    if not wt.is_contiguous():
        raise ValueError("imaging_weight must be contiguous")
    _lib.check(_lib.lib().arlmi_reweight(ctypes.byref(g), ptr(vis.data["uvw"]), _freq(vis), vis.nvis, ptr(density),
                                         ptr(wt), vis.precision, stream_handle()))
    return vis


def weight_visibility(vis, model, gcfcf=None, weighting="uniform", **kwargs):
    """Uniform weighting of one visibility set (imaging/weighting.py:19-40)."""
    if weighting != "uniform":
        raise NotImplementedError("weighting %r" % weighting)
    if gcfcf is None:
        from .kernels import create_pswf_convolutionfunction
        gcfcf = create_pswf_convolutionfunction(model)
    griddata = create_griddata_from_image(model)
    griddata, _ = grid_weight_to_griddata(vis, griddata, gcfcf[1])
    return griddata_reweight(vis, griddata, gcfcf[1])


def weight_list_serial_workflow(vis_list, model_imagelist, gcfcf=None, weighting="uniform", **kwargs):
    """Uniform weighting over a list: the weight grids of every set are summed, then every set is
    reweighted by the sum (imaging_serial.py:426-473)."""
    if weighting != "uniform":
        raise NotImplementedError("weighting %r" % weighting)
    centre = len(model_imagelist) // 2
    if gcfcf is None:
        from .kernels import create_pswf_convolutionfunction
        gcfcf = [create_pswf_convolutionfunction(model_imagelist[centre])]
    cf = gcfcf[0][1]
    grids = []
    for vis, model in zip(vis_list, model_imagelist):
        if vis is None or model is None:
            continue
        grids.append(grid_weight_to_griddata(vis, create_griddata_from_image(model), cf))
    merged, _ = griddata_merge_weights(grids)
    out = []
    for vis, model in zip(vis_list, model_imagelist):
        if vis is None:
            out.append(None)
            continue
        agd = create_griddata_from_image(model)
        agd.data = merged.data
        out.append(griddata_reweight(vis, agd, cf))
    return out
