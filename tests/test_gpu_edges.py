"""Edge cases of the GPU path against the FP64 oracle: no rows, one row, all-zero w (no slab
holds a row: ARL's slab width is then 0), rows exactly on a slab boundary (in no slab), zero
weights, and rows in the outermost kernel-footprint positions of the grid."""

import numpy
import pytest

from oracle import arl_oracle as O
from helpers import rel_err, random_rows, to_device_vis, device_image, image_errors

pytestmark = pytest.mark.gpu


def _setup(precision, n=128, cell=2e-3, S=8):
    import arlmi
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=S)
    g = O.ImageGeom(n, cell)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    return im, (gcf, cf), g, (ogcf, ocf, cfg)


def _empty(n):
    return {"uvw": numpy.zeros((0, 3)), "vis": numpy.zeros((0, 1), dtype=complex),
            "imaging_weight": numpy.zeros((0, 1)), "frequency": numpy.zeros(0)}


@pytest.mark.parametrize("precision", [64, 32])
def test_no_rows(gpu, precision):
    import arlmi
    im, gc, g, og = _setup(precision)
    vis = to_device_vis(_empty(0), precision)
    img, sw = arlmi.invert_2d(vis, im, gcfcf=gc)
    assert numpy.all(img.numpy() == 0) and sw[0, 0] == 0
    (img, sw), = arlmi.invert_list_serial_workflow([vis], [im], context="wstack", vis_slices=4, gcfcf=[gc])
    assert numpy.all(img.numpy() == 0) and sw[0, 0] == 0
    pv = arlmi.predict_2d(vis, im, gcfcf=gc)
    assert pv.vis.shape[0] == 0


@pytest.mark.parametrize("precision", [64, 32])
def test_one_row(gpu, precision):
    import arlmi
    im, gc, g, (ogcf, ocf, cfg) = _setup(precision)
    d = random_rows(numpy.random.default_rng(1), 1, 128, 2e-3, 8)
    vis = to_device_vis(d, precision)
    img, sw = arlmi.invert_2d(vis, im, gcfcf=gc)
    rimg, rsw = O.invert_2d(d, g, ogcf, ocf, cfg)
    l2, inner, full = image_errors(img.numpy(), rimg)
    assert inner < (1e-10 if precision == 64 else 1e-5), (l2, inner, full)
    assert abs(sw[0, 0] - rsw[0, 0]) < 1e-6 * rsw[0, 0]


def test_all_w_zero_wstack(gpu):
    """max |w| = 0: boxes = linspace(0, 0, P), width 0, |w - box| < 0 never holds, so every row
    is in no slab: the image is 0 with sum of weights 0, and the prediction is 0 (ARL's list
    workflows leave rows in no slab at zero)."""
    import arlmi
    im, gc, g, (ogcf, ocf, cfg) = _setup(64)
    d = random_rows(numpy.random.default_rng(2), 500, 128, 2e-3, 8)
    vis = to_device_vis(d, 64)
    (img, sw), = arlmi.invert_list_serial_workflow([vis], [im], context="wstack", vis_slices=5, gcfcf=[gc])
    rimg, rsw = O.invert_list_wstack(d, g, ogcf, ocf, cfg, 5)
    assert numpy.all(img.numpy() == 0) and sw[0, 0] == 0 and rsw[0, 0] == 0 and numpy.all(rimg == 0)
    model = numpy.zeros((1, 1, 128, 128))
    model[0, 0, 60, 70] = 1.0
    pv, = arlmi.predict_list_serial_workflow([vis], [device_image(128, 2e-3, 64, data=model)], context="wstack",
                                             vis_slices=5, gcfcf=[gc])
    assert numpy.all(pv.vis.cpu().numpy() == 0)
    assert numpy.all(O.predict_list_wstack(d, model, g, ogcf, ocf, cfg, 5) == 0)


def test_rows_on_slab_boundaries(gpu):
    """Rows exactly half-way between two slab centres satisfy neither strict test and are in no
    slab (iterators.py:107-109): excluded from the image and predicted as 0, as in the oracle."""
    import arlmi
    im, gc, g, (ogcf, ocf, cfg) = _setup(64)
    rng = numpy.random.default_rng(3)
    d = random_rows(rng, 2000, 128, 2e-3, 8, wmax=100.0)
    P = 5
    d["uvw"][0, 2] = 100.0      # max |w| is exactly 100: boxes -100, -50, 0, 50, 100, width 50
    d["uvw"][1:40, 2] = 25.0    # exactly between boxes 0 and 50
    d["uvw"][40:80, 2] = -75.0  # exactly between -100 and -50
    vis = to_device_vis(d, 64)
    (img, sw), = arlmi.invert_list_serial_workflow([vis], [im], context="wstack", vis_slices=P, gcfcf=[gc])
    rimg, rsw = O.invert_list_wstack(d, g, ogcf, ocf, cfg, P)
    assert image_errors(img.numpy(), rimg)[2] < 1e-8
    assert abs(sw[0, 0] - rsw[0, 0]) < 1e-9 * rsw[0, 0]
    model = numpy.zeros((1, 1, 128, 128))
    model[0, 0, 50, 80] = 1.0
    pv, = arlmi.predict_list_serial_workflow([vis], [device_image(128, 2e-3, 64, data=model)], context="wstack",
                                             vis_slices=P, gcfcf=[gc])
    rpv = O.predict_list_wstack(d, model, g, ogcf, ocf, cfg, P)
    got = pv.vis.cpu().numpy()
    assert numpy.all(got[1:80] == 0) and numpy.all(rpv[1:80] == 0)
    assert rel_err(got, rpv) < 1e-10


@pytest.mark.parametrize("precision", [64, 32])
def test_zero_weights_and_grid_edges(gpu, precision):
    """Half the rows with zero weight, and rows whose kernel footprint touches the first and last
    grid rows and columns (pu - S/2 = 0 and pu + S/2 = N)."""
    import arlmi
    n, cell, S = 128, 2e-3, 8
    im, gc, g, (ogcf, ocf, cfg) = _setup(precision, n, cell, S)
    rng = numpy.random.default_rng(4)
    d = random_rows(rng, 3000, n, cell, S)
    d["imaging_weight"][::2] = 0.0
    du = g.du()
    # x = N//2 + u/du: footprint start x - S/2 = 0 -> x = S/2; end x + S/2 = N -> x = N - S/2
    d["uvw"][:4, 0] = (numpy.array([S // 2, n - S // 2, S // 2 + 0.3, n - S // 2 - 0.4]) - n // 2) * du
    d["uvw"][:4, 1] = (numpy.array([n - S // 2, S // 2, S // 2, n - S // 2 - 0.2]) - n // 2) * g.dv()
    vis = to_device_vis(d, precision)
    img, sw = arlmi.invert_2d(vis, im, gcfcf=gc)
    rimg, rsw = O.invert_2d(d, g, ogcf, ocf, cfg)
    l2, inner, full = image_errors(img.numpy(), rimg)
    assert inner < (1e-10 if precision == 64 else 1e-5) and full < (1e-8 if precision == 64 else 5e-3)
    assert abs(sw[0, 0] - rsw[0, 0]) < 1e-6 * rsw[0, 0]
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, 30, 90] = 1.0
    pv = arlmi.predict_2d(vis, device_image(n, cell, precision, data=model), gcfcf=gc)
    assert rel_err(pv.vis.cpu().numpy(), O.predict_2d(d, model, g, ogcf, ocf, cfg)) < (1e-10 if precision == 64
                                                                                      else 1e-5)
