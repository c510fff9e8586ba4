"""Fixed-point gridding (ARLMI_FLAG_DETERMINISTIC): repeated grids, images and sums of weights are
bitwise identical, also from a rebuilt plan, and they equal the oracle and the float64-add path to
the usual tolerances.  Covered: the separable gridder with split tiles (integer partial tiles), the
table gridder in its three tap layouts (64 and 256 taps per lane group, and the generic loop), the
PSF, and chunks added into resident grids."""

import ctypes

import numpy
import pytest

from oracle import arl_oracle as O
from helpers import random_rows, to_device_vis, oracle_geom, device_image, image_errors

pytestmark = pytest.mark.gpu

TOL = {32: 1e-5, 64: 1e-10}
FULL = {32: 5e-3, 64: 1e-8}
SW_TOL = {32: 1e-9, 64: 1e-12}  # float32 weights are rounded copies of the oracle's float64 ones
# Fixed point against float64 adds.  float64: the taps are rounded to 2^-60 of their tile's bound
# (grids agree to ~7e-15 L2), and the grid correction amplifies that uniform rounding towards the
# image corners: inner field ~1e-14, L2 ~1e-11, full-field max ~5e-11 measured at 256^2.  float32:
# the grid values of the two modes differ by the last bit of a float, amplified the same way.
VS_ADD = {64: dict(l2=1e-10, inner=1e-12, full=1e-9), 32: dict(l2=1e-3, inner=1e-6, full=5e-3)}


def _close_to_add_path(a, b, precision):
    l2, inner, full = image_errors(a, b)
    t = VS_ADD[precision]
    assert l2 < t["l2"] and inner < t["inner"] and full < t["full"], (l2, inner, full)


def _pswf(n, cell, S, precision):
    import arlmi
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=S)
    return im, gcf, cf


def _same(a, b):
    import torch
    return torch.equal(a, b)


@pytest.mark.parametrize("precision", [64, 32])
def test_wstack_bitwise_repeatable(gpu, precision):
    import arlmi
    rng = numpy.random.default_rng(71)
    n, cell, S, P = 256, 1e-3, 8, 9
    d = random_rows(rng, 40000, n, cell, S, wmax=300.0)
    im, gcf, cf = _pswf(n, cell, S, precision)
    vis = to_device_vis(d, precision)
    kw = dict(gcfcf=(gcf, cf), mode="wstack", vis_slices=P, sparse=False, chunk=128)
    a = arlmi.Imager(vis, im, deterministic=True, **kw)
    st = a.plan.stats()
    assert st["ngrid_items"] > st["ntiles"], st  # split tiles: integer partial tiles are summed
    img0 = a.invert(vis)[0].clone()
    sw0 = a.sumwt()
    grids0 = a.grids.clone()  # the transformed planes the invert leaves in its scratch grids
    for _ in range(2):
        img, _ = a.invert(vis)
        assert _same(img, img0) and numpy.array_equal(a.sumwt(), sw0)
        assert _same(a.grids, grids0)
    # a second plan over the same rows (its own entry order and work items)
    b = arlmi.Imager(vis, im, deterministic=True, **kw)
    img_b, _ = b.invert(vis)
    assert _same(img_b, img0) and numpy.array_equal(b.sumwt(), sw0) and _same(b.grids, grids0)
    psf0 = a.invert(vis, dopsf=True)[0].clone()
    psf1, _ = b.invert(vis, dopsf=True)
    assert _same(psf0, psf1)
    # the oracle, and the float64-add path
    g = oracle_geom(n, cell)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    rimg, rsw = O.invert_list_wstack(d, g, ogcf, ocf, cfg, P)
    l2, inner, full = image_errors(img0.cpu().numpy(), rimg)
    assert inner < TOL[precision] and full < FULL[precision], (l2, inner, full)
    assert abs(sw0[0, 0] - rsw[0, 0]) < SW_TOL[precision] * rsw[0, 0]
    c = arlmi.Imager(vis, im, **kw)
    img_c, _ = c.invert(vis)
    _close_to_add_path(img0.cpu().numpy(), img_c.cpu().numpy(), precision)
    # predict is unchanged by the flag
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, rng.integers(40, 216, 20), rng.integers(40, 216, 20)] = 1.0
    mim = device_image(n, cell, precision, data=model)
    assert _same(a.predict(vis, mim.data), c.predict(vis, mim.data))


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("S", [8, 16, 22])
def test_table_kernel_bitwise_repeatable(gpu, precision, S):
    """AW kernel table: 64 taps (one per lane), 256 taps (four per lane) and the generic loop."""
    import arlmi
    rng = numpy.random.default_rng(72 + S)
    n, cell = 256, 5e-4
    d = random_rows(rng, 12000, n, cell, S, wmax=80.0)
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_awterm_convolutionfunction(im, nw=9, wstep=20.0, oversampling=4, support=S)
    vis = to_device_vis(d, precision)
    a = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode="2d", deterministic=True, chunk=256)
    img0 = a.invert(vis)[0].clone()
    sw0 = a.sumwt()
    img1, _ = a.invert(vis)
    sw1 = a.sumwt()
    b = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode="2d", deterministic=True, chunk=256)
    img2, _ = b.invert(vis)
    sw2 = b.sumwt()
    assert _same(img0, img1) and _same(img0, img2)
    assert numpy.array_equal(sw0, sw1) and numpy.array_equal(sw0, sw2)
    g = oracle_geom(n, cell)
    ogcf, ocf, cfg = O.create_awterm_convolutionfunction(g, nw=9, wstep=20.0, oversampling=4, support=S)
    rre, rsw = O.invert_2d(d, g, ogcf, ocf, cfg)
    l2, inner, full = image_errors(img0.cpu().numpy(), rre)
    assert inner < TOL[precision] and full < FULL[precision], (l2, inner, full)
    assert abs(sw0[0, 0] - rsw[0, 0]) < SW_TOL[precision] * rsw[0, 0]


@pytest.mark.parametrize("precision", [64, 32])
def test_stream_bitwise_repeatable(gpu, precision):
    """Chunks added into resident grids: two passes give the same bits; and the image of the
    float64-add streaming pass to rounding."""
    import arlmi
    rng = numpy.random.default_rng(73)
    n, cell, S, P = 512, 5e-4, 8, 8
    d = random_rows(rng, 30000, n, cell, S, wmax=200.0)
    cuts = [0, 9000, 21000, 30000]
    chunks = [to_device_vis({k: v[cuts[i]:cuts[i + 1]] for k, v in d.items()}, precision) for i in range(3)]
    im, gcf, cf = _pswf(n, cell, S, precision)
    st = arlmi.StreamingImager(im, gcfcf=(gcf, cf), mode="wstack", vis_slices=P, precision=precision,
                               deterministic=True)
    img0, sw0 = st.invert(chunks)
    img0 = img0.clone()
    img1, sw1 = st.invert(chunks)
    assert _same(img0, img1) and numpy.array_equal(sw0, sw1)
    ref = arlmi.StreamingImager(im, gcfcf=(gcf, cf), mode="wstack", vis_slices=P, precision=precision)
    rimg, rsw = ref.invert(chunks)
    _close_to_add_path(img0.cpu().numpy(), rimg.cpu().numpy(), precision)
    assert abs(sw0[0, 0] - rsw[0, 0]) < SW_TOL[precision] * rsw[0, 0]
    st.release()
    ref.release()


def test_fixed_point_plans_are_not_batched(gpu):
    import arlmi
    from arlmi.plan import kernel_struct
    rng = numpy.random.default_rng(74)
    d = random_rows(rng, 2000, 256, 1e-3, 8)
    vis = to_device_vis(d, 32)
    im, gcf, cf = _pswf(256, 1e-3, 8, 32)
    imager = arlmi.Imager(vis, im, gcfcf=(gcf, cf), deterministic=True)
    pv = vis.vis.clone()
    items = (arlmi._lib.BatchItem * 1)()
    k, keep = kernel_struct(cf, 32)
    items[0].plan = imager.plan.handle.value
    items[0].d_vis = vis.vis.data_ptr()
    items[0].d_weight = vis.data["imaging_weight"].data_ptr()
    items[0].kernel = k
    items[0].d_grids = imager.grids.data_ptr()
    items[0].d_sumwt = imager.sumwt_planes.data_ptr()
    items[0].d_vis_out = pv.data_ptr()
    handle = ctypes.c_void_p()
    rc = arlmi._lib.lib().arlmi_batch_create(items, 1, 32, ctypes.byref(handle))
    assert rc == -8 and not handle.value, rc
    assert b"deterministic" in arlmi._lib.lib().arlmi_last_error()
