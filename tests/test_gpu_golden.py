"""The HIP path against the committed golden fixtures (tests/golden/golden_small.npz: inputs and
FP64 oracle outputs of small seeded problems).  Tolerances as in test_gpu_parity: float64 1e-10,
float32 1e-5 (images over the inner field, 5e-3 over the full field); the BlockVisibility
conversion bit for bit."""

import os

import numpy
import pytest

from helpers import rel_err, image_errors, to_device_vis, device_image

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = {32: 1e-5, 64: 1e-10}


@pytest.fixture(scope="module")
def golden():
    return dict(numpy.load(os.path.join(HERE, "golden_small.npz"), allow_pickle=False))


def _rows(golden, prefix):
    return {k: golden[prefix + k] for k in ("uvw", "vis", "imaging_weight", "frequency")}


def _check_image(a, b, precision):
    l2, inner, full = image_errors(a, b)
    assert inner < TOL[precision] and full < (5e-3 if precision == 32 else 1e-8), (l2, inner, full)


@pytest.mark.parametrize("precision", [64, 32])
def test_golden_2d_and_wstack(gpu, golden, precision):
    import arlmi
    n, cell, S, os_ = golden["2d_geom"]
    n, S, os_, cell = int(n), int(S), int(os_), float(cell)
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=os_, support=S)
    vis = to_device_vis(_rows(golden, "2d_"), precision)
    img, sw = arlmi.invert_2d(vis, im, gcfcf=(gcf, cf))
    _check_image(img.numpy(), golden["2d_image"], precision)
    assert abs(sw[0, 0] - golden["2d_sumwt"][0, 0]) < 1e-6 * golden["2d_sumwt"][0, 0]
    model = device_image(n, cell, precision, data=golden["2d_model"])
    pv = arlmi.predict_2d(vis, model, gcfcf=(gcf, cf))
    assert rel_err(pv.vis.cpu().numpy(), golden["2d_predict"]) < TOL[precision]
    P = int(golden["ws_slices"][0])
    vis = to_device_vis(_rows(golden, "ws_"), precision)
    (img, sw), = arlmi.invert_list_serial_workflow([vis], [im], context="wstack", vis_slices=P, gcfcf=[(gcf, cf)])
    _check_image(img.numpy(), golden["ws_image"], precision)
    pv, = arlmi.predict_list_serial_workflow([vis], [model], context="wstack", vis_slices=P, gcfcf=[(gcf, cf)])
    assert rel_err(pv.vis.cpu().numpy(), golden["ws_predict"]) < TOL[precision]
    if precision == 64:
        w = arlmi.weight_visibility(vis, im)
        assert rel_err(w.imaging_weight.cpu().numpy(), golden["wt_new"]) < 1e-14


def test_golden_block_conversion(gpu, golden):
    import arlmi
    b = {k[len("blk_in_"):]: v for k, v in golden.items() if k.startswith("blk_in_")}
    bv = arlmi.BlockVisibility.from_arrays(b["uvw"], vis=b["vis"], weight=b["weight"],
                                           imaging_weight=b["imaging_weight"], time=b["time"],
                                           integration_time=b["integration_time"], frequency=b["frequency"],
                                           channel_bandwidth=b["channel_bandwidth"], precision=64)
    vis = arlmi.convert_blockvisibility_to_visibility(bv)
    for k in ("uvw", "vis", "weight", "imaging_weight", "time", "frequency", "antenna1", "antenna2"):
        assert numpy.array_equal(vis.data[k].cpu().numpy().reshape(golden["blk_out_" + k].shape),
                                 golden["blk_out_" + k]), k
    assert numpy.array_equal(vis.cindex.cpu().numpy(), golden["blk_out_cindex"])
