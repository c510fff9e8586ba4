"""The committed golden fixtures: ARL's own known answers (tests/golden/arl_known_answers.json)
and the FP64 oracle's outputs on small seeded problems (tests/golden/golden_small.npz, written
by tests/golden/make_golden.py).  The oracle must keep reproducing both; the GPU path is
compared with the npz in tests/test_gpu_golden.py."""

import json
import os

import numpy
import pytest

from oracle import arl_oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def known():
    with open(os.path.join(HERE, "arl_known_answers.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def golden():
    return dict(numpy.load(os.path.join(HERE, "golden_small.npz"), allow_pickle=False))


@pytest.mark.parametrize("key", ["pswf_os8_s6", "pswf_os1_s6"])
def test_known_pswf(known, key):
    k = known[key]
    g = O.ImageGeom(k["npixel"], k["cellsize"])
    _, cf, _ = O.create_pswf_convolutionfunction(g, oversampling=k["oversampling"], support=k["support"])
    pk = numpy.unravel_index(numpy.argmax(numpy.abs(cf)), cf.shape)
    assert list(pk) == k["peak_index"]
    assert abs(cf[pk] - k["peak_value"]) < k["tol"]


def test_known_box(known):
    k = known["box"]
    _, cf, _ = O.create_box_convolutionfunction(O.ImageGeom(k["npixel"], k["cellsize"]))
    pk = numpy.unravel_index(numpy.argmax(numpy.abs(cf)), cf.shape)
    assert list(pk) == k["peak_index"] and abs(cf[pk] - k["peak_value"]) < k["tol"]


def test_known_wterm(known):
    k = known["wterm_nw201"]
    g = O.ImageGeom(k["npixel"], k["cellsize"])
    _, cf, _ = O.create_awterm_convolutionfunction(g, nw=k["nw"], wstep=k["wstep"], oversampling=k["oversampling"],
                                                   support=k["support"], use_aaf=True)
    pk = numpy.unravel_index(numpy.argmax(numpy.abs(cf)), cf.shape)
    assert list(pk) == k["peak_index"] and abs(cf[pk] - k["peak_value"]) < k["tol"]
    clipped = O.apply_bounding_box_convolutionfunction(cf, fractional_level=k["clipped_level"])
    pk2 = numpy.unravel_index(numpy.argmax(numpy.abs(clipped)), clipped.shape)
    assert list(pk2) == k["clipped_peak_index"]


def test_known_w_beam(known):
    k = known["w_beam_centre"]
    assert abs(O.w_beam(k["npixel"], k["fov"], k["w"])[0, 0] - k["value_00"]) < k["tol"]


def _rows(golden, prefix):
    return {k: golden[prefix + k] for k in ("uvw", "vis", "imaging_weight", "frequency")}


def test_oracle_reproduces_golden(golden):
    """The oracle's outputs on the committed inputs are the committed outputs, bit for bit."""
    n, cell, S, os_ = golden["2d_geom"]
    n, S, os_ = int(n), int(S), int(os_)
    g = O.ImageGeom(n, float(cell))
    gcf, cf, cfg = O.create_pswf_convolutionfunction(g, os_, S)
    d = _rows(golden, "2d_")
    img, sw = O.invert_2d(d, g, gcf, cf, cfg)
    assert numpy.array_equal(img, golden["2d_image"]) and numpy.array_equal(sw, golden["2d_sumwt"])
    assert numpy.array_equal(O.predict_2d(d, golden["2d_model"], g, gcf, cf, cfg), golden["2d_predict"])
    d = _rows(golden, "ws_")
    P = int(golden["ws_slices"][0])
    img, sw = O.invert_list_wstack(d, g, gcf, cf, cfg, P)
    assert numpy.array_equal(img, golden["ws_image"])
    assert numpy.array_equal(O.predict_list_wstack(d, golden["2d_model"], g, gcf, cf, cfg, P), golden["ws_predict"])
    assert numpy.array_equal(O.weight_visibility(d, g, cfg), golden["wt_new"])
    b = {k[len("blk_in_"):]: v for k, v in golden.items() if k.startswith("blk_in_")}
    rows = O.convert_blocks(b)
    for k, v in rows.items():
        assert numpy.array_equal(v, golden["blk_out_" + k]), k
