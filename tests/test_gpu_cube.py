"""Spectral cubes and grid WW planes on the GPU against the oracle.

Cubes: image channels mapped by pfreq = round((f - f0)/df) (gridding.py:80-85) with per-channel
grids, sums of weights, gcf and normalisation, in 2-D and w-stack invert and predict.  WW planes:
a GridData with nw planes wstep apart (create_griddata_from_image(nw, wstep),
griddata/operations.py:91-145) gridded with a w-projection kernel, each row on plane
round(w/wstep) + nw//2 (gridding.py:58-65), transformed with the plane sum of
fft_griddata_to_image (:270-289), and degridded from per-plane grids.
Tolerances as in test_gpu_parity: 1e-10 (float64), 1e-5 (float32; images over the inner half).
"""

import numpy
import pytest

from oracle import arl_oracle as O
from helpers import random_rows, to_device_vis, image_errors, rel_err

pytestmark = pytest.mark.gpu

TOL = {64: 1e-10, 32: 1e-5}
FREQS = (1.0e8, 1.1e8, 1.2e8)


def _cube_image(n, cell, precision, data=None):
    import torch
    import arlmi
    dt = torch.float32 if precision == 32 else torch.float64
    im = arlmi.create_image(npixel=n, cellsize=cell, frequency=list(FREQS), channel_bandwidth=[1.0e7] * 3, dtype=dt)
    if data is not None:
        im.data = torch.as_tensor(data).to(device=im.data.device, dtype=dt).reshape(im.data.shape)
    return im


def _assert_image(a, b, precision):
    l2, inner, full = image_errors(a, b)
    assert inner < TOL[precision] and (precision == 32 or l2 < TOL[64]), (l2, inner, full)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("context,P", [("2d", 1), ("wstack", 7)])
def test_spectral_cube_invert_predict(gpu, precision, context, P):
    import arlmi
    rng = numpy.random.default_rng(51)
    n, cell, S = 256, 1e-3, 8
    d = random_rows(rng, 12000, n, cell, S, wmax=150.0 if context == "wstack" else 0.0)
    d["frequency"] = rng.choice(FREQS, 12000)
    d["imaging_weight"][d["frequency"] == FREQS[2]] *= 3.0  # channels with different sums of weights
    vis = to_device_vis(d, precision)
    im = _cube_image(n, cell, precision)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=S)
    g = O.ImageGeom(n, cell, nchan=3, frequency=FREQS[0], channel_bandwidth=1.0e7)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    (img, sw), = arlmi.invert_list_serial_workflow([vis], [im], context=context, vis_slices=P, gcfcf=[(gcf, cf)])
    if context == "wstack":
        ref, rsw = O.invert_list_wstack(d, g, ogcf, ocf, cfg, P)
    else:
        ref, rsw = O.invert_2d(d, g, ogcf, ocf, cfg)
    assert img.data.shape[0] == 3
    _assert_image(img.numpy(), ref, precision)
    assert numpy.max(numpy.abs(sw - rsw)) < 1e-9 * numpy.max(rsw)
    assert len(set(numpy.round(rsw[:, 0]))) == 3
    model = numpy.zeros((3, 1, n, n))
    for c in range(3):
        model[c, 0, rng.integers(60, 196, 15), rng.integers(60, 196, 15)] = rng.uniform(0.5, 2.0, 15)
    mim = _cube_image(n, cell, precision, data=model)
    pv, = arlmi.predict_list_serial_workflow([vis], [mim], context=context, vis_slices=P, gcfcf=[(gcf, cf)])
    if context == "wstack":
        rpv = O.predict_list_wstack(d, model, g, ogcf, ocf, cfg, P)
    else:
        rpv = O.predict_2d(d, model, g, ogcf, ocf, cfg)
    assert rel_err(pv.vis.cpu().numpy(), rpv) < TOL[precision]


@pytest.mark.parametrize("precision", [64, 32])
def test_grid_ww_planes(gpu, precision):
    import arlmi
    rng = numpy.random.default_rng(52)
    n, cell, nw, wstep, S = 128, 2e-3, 7, 20.0, 12
    d = random_rows(rng, 8000, n, cell, S, wmax=65.0)
    vis = to_device_vis(d, precision)
    import torch
    dt = torch.float32 if precision == 32 else torch.float64
    im = arlmi.create_image(npixel=n, cellsize=cell, frequency=[1e8], channel_bandwidth=[1e6], dtype=dt)
    gcf, cf = arlmi.create_awterm_convolutionfunction(im, nw=nw, wstep=wstep, oversampling=8, support=S)
    g = O.ImageGeom(n, cell)
    ogcf, ocf, cfg = O.create_awterm_convolutionfunction(g, nw=nw, wstep=wstep, oversampling=8, support=S)
    cdt = torch.complex64 if precision == 32 else torch.complex128
    gd = arlmi.create_griddata_from_image(im, nw=nw, wstep=wstep, dtype=cdt)
    gd, sw = arlmi.grid_visibility_to_griddata(vis, gd, cf)
    ref, rsw = O.grid_visibility_to_griddata(d, g, ocf, cfg, nw_grid=nw, wstep_grid=wstep)
    got = gd.data.cpu().numpy()
    assert got.shape == ref.shape
    assert rel_err(got, ref) < TOL[precision]
    for z in range(nw):  # every plane holds its own rows
        assert rel_err(got[:, :, z], ref[:, :, z]) < 10 * TOL[precision] or numpy.max(numpy.abs(ref[:, :, z])) == 0
    assert abs(sw[0, 0] - rsw[0, 0]) < 1e-9 * rsw[0, 0]
    img = arlmi.fft_griddata_to_image(gd, gcf)
    rimg = O.fft_griddata_to_image(ref, ogcf)
    _assert_image(img.data.cpu().numpy(), rimg, precision)
    # degridding from per-plane grids (each plane different)
    gd2 = arlmi.create_griddata_from_image(im, nw=nw, wstep=wstep, dtype=cdt)
    gd2.data = torch.as_tensor(ref).to(device=gd2.data.device, dtype=cdt)
    pv = arlmi.degrid_visibility_from_griddata(vis, gd2, cf)
    rpv = O.degrid_visibility_from_griddata(d, ref, g, ocf, cfg, wstep_grid=wstep)
    assert rel_err(pv.vis.cpu().numpy(), rpv) < TOL[precision]
