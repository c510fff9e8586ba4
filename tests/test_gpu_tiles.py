"""The tile sides and work-item sizes the autotuner can choose (arlmi.batch.Autotuner: T = 32 or 64,
1024 to 16384 entries per tile chunk; 128 to split many tiles) give the oracle's images and visibilities, for the separable
gridder and degridder (w-stack with screens, sparse and dense grids) and for the table kernels
(AW, with the entry loop's partly filled last tap group); and the smallest separable support,
S = 2 (4 taps on 64 lanes)."""

import numpy
import pytest

from oracle import arl_oracle as O
from helpers import rel_err, random_rows, to_device_vis, oracle_geom, device_image, image_errors

pytestmark = pytest.mark.gpu

TOL = {32: 1e-5, 64: 1e-10}
FULL = {32: 5e-3, 64: 1e-8}


def _check_image(a, b, precision):
    l2, inner, full = image_errors(a, b)
    assert inner < TOL[precision] and full < FULL[precision], (l2, inner, full)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("tile,chunk,sparse", [(32, 0, True), (32, 128, False), (64, 128, True), (64, 1024, True),
                                               (64, 16384, True), (32, 16384, True)])
def test_wstack_tiles_and_chunks(gpu, precision, tile, chunk, sparse):
    import arlmi
    rng = numpy.random.default_rng(41 + tile + chunk)
    n, cell, S, P = 256, 1e-3, 8, 7
    d = random_rows(rng, 30000, n, cell, S, wmax=250.0)
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=S)
    vis = to_device_vis(d, precision)
    imager = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode="wstack", vis_slices=P, tile=tile, chunk=chunk,
                          sparse=sparse)
    st = imager.plan.stats()
    if chunk == 128:
        assert st["ngrid_items"] > st["ntiles"], st  # split tiles: partial tiles summed
    img, _ = imager.invert(vis)
    g = oracle_geom(n, cell)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    rimg, rsw = O.invert_list_wstack(d, g, ogcf, ocf, cfg, P)
    _check_image(img.cpu().numpy(), rimg, precision)
    assert abs(imager.sumwt()[0, 0] - rsw[0, 0]) < 1e-9 * rsw[0, 0]
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, rng.integers(40, 216, 25), rng.integers(40, 216, 25)] = rng.uniform(0.5, 2.0, 25)
    mim = device_image(n, cell, precision, data=model)
    pv = imager.predict(vis, mim.data)
    rpv = O.predict_list_wstack(d, model, g, ogcf, ocf, cfg, P)
    assert rel_err(pv.cpu().numpy(), rpv) < TOL[precision]


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("S,tile", [(14, 32), (6, 32), (22, 32)])
def test_table_kernel_small_tiles(gpu, precision, S, tile):
    import arlmi
    rng = numpy.random.default_rng(50 + S)
    n, cell = 256, 5e-4
    d = random_rows(rng, 8000, n, cell, S, wmax=80.0)
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_awterm_convolutionfunction(im, nw=9, wstep=20.0, oversampling=4, support=S)
    g = oracle_geom(n, cell)
    ogcf, ocf, cfg = O.create_awterm_convolutionfunction(g, nw=9, wstep=20.0, oversampling=4, support=S)
    vis = to_device_vis(d, precision)
    imager = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode="2d", tile=tile, chunk=1024)
    img, _ = imager.invert(vis)
    rimg, _ = O.invert_2d(d, g, ogcf, ocf, cfg)
    _check_image(img.cpu().numpy(), rimg, precision)
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, 128, 100] = 1.0
    model[0, 0, 90, 150] = 0.5
    mim = device_image(n, cell, precision, data=model)
    pv = imager.predict(vis, mim.data)
    rpv = O.predict_2d(d, model, g, ogcf, ocf, cfg)
    assert rel_err(pv.cpu().numpy(), rpv) < TOL[precision]


@pytest.mark.parametrize("precision", [64, 32])
def test_support_two(gpu, precision):
    """S = 2: the separable gridder's 4 taps on 64 lanes (60 lanes add to the trash row)."""
    import arlmi
    rng = numpy.random.default_rng(61)
    n, cell, S = 128, 5e-4, 2
    d = random_rows(rng, 20000, n, cell, S)
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=S)
    vis = to_device_vis(d, precision)
    gd = arlmi.create_griddata_from_image(im)
    gd, sumwt = arlmi.grid_visibility_to_griddata(vis, gd, cf)
    g = oracle_geom(n, cell)
    _, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    ref, rsum = O.grid_visibility_to_griddata(d, g, ocf, cfg)
    assert rel_err(gd.data.cpu().numpy(), ref) < TOL[precision] / 10
    out = arlmi.degrid_visibility_from_griddata(vis, gd, cf)
    rv = O.degrid_visibility_from_griddata(d, ref, g, ocf, cfg)
    assert rel_err(out.vis.cpu().numpy(), rv) < TOL[precision]


def _clustered_rows(rng, n, cell, S, wmax):
    """Rows of all three sizes of tile segment the plan's origin sort tells apart: 30,000 rows on
    40 pixels of one tile (more than 4096 entries, hundreds a footprint origin), 6,000 near the
    centre, and a few hundred spread over the grid (segments under 64 entries)."""
    du = 1.0 / (n * cell)
    d = random_rows(rng, 36600, n, cell, S, wmax=wmax)
    px = rng.integers(-40, -20, (40, 2))  # pixel offsets from the centre, inside one 64-wide tile
    pick = rng.integers(0, 40, 30000)
    d["uvw"][:30000, 0] = (px[pick, 0] + rng.uniform(-0.45, 0.45, 30000)) * du
    d["uvw"][:30000, 1] = (px[pick, 1] + rng.uniform(-0.45, 0.45, 30000)) * du
    d["uvw"][30000:36000, :2] = rng.uniform(-30 * du, 30 * du, (6000, 2))
    return d


@pytest.mark.parametrize("precision", [64, 32])
def test_origin_grouped_gridding(gpu, precision):
    """Entries grouped by footprint origin (every segment class of the sort) and added in runs:
    images and visibilities equal the oracle's, the gridder adds far fewer runs than entries, and
    a deterministic plan (which adds every entry) gives the same images."""
    import arlmi
    rng = numpy.random.default_rng(77)
    n, cell, S, P = 256, 1e-3, 8, 3
    d = _clustered_rows(rng, n, cell, S, 40.0)
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=S)
    vis = to_device_vis(d, precision)
    imager = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode="wstack", vis_slices=P)
    st = imager.plan.stats()
    runs = imager.plan.grid_runs(S)
    assert 0 < runs < 0.5 * st["nentries"], (runs, st["nentries"])
    img, _ = imager.invert(vis)
    g = oracle_geom(n, cell)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    rimg, rsw = O.invert_list_wstack(d, g, ogcf, ocf, cfg, P)
    _check_image(img.cpu().numpy(), rimg, precision)
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, rng.integers(40, 216, 25), rng.integers(40, 216, 25)] = rng.uniform(0.5, 2.0, 25)
    mim = device_image(n, cell, precision, data=model)
    pv = imager.predict(vis, mim.data)
    rpv = O.predict_list_wstack(d, model, g, ogcf, ocf, cfg, P)
    assert rel_err(pv.cpu().numpy(), rpv) < TOL[precision]
    det = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode="wstack", vis_slices=P, deterministic=True)
    assert det.plan.grid_runs(S) == det.plan.stats()["nentries"]
    dimg, _ = det.invert(vis)
    _check_image(dimg.cpu().numpy(), rimg, precision)
