"""invert / predict onto an image centred away from the visibility phase centre
(shift_vis_to_image, imaging/base.py:50-81, called at :136 and :165) against the oracle.

The GPU rotation runs in float64 and stores in the visibility precision; tolerances are those of
the unshifted invert / predict parity tests (1e-10 in float64, 1e-5 in float32 over the inner
image).  Also: shift_vis_to_image is phaserotate_visibility to the image centre, bit for bit, and
labels the result; centres within 1e-15 rad, or a side without a sky position, change nothing;
the adjoint identity of predict and its adjoint holds through the shift.
"""

import numpy
import pytest

from helpers import rel_err, random_rows, to_device_vis, image_errors
from oracle import arl_oracle as O

pytestmark = pytest.mark.gpu

TOL = {32: 1e-5, 64: 1e-10}
FULL = {32: 5e-3, 64: 1e-8}
D2R = numpy.pi / 180.0
C0 = (40.0 * D2R, -30.0 * D2R)


def _centre(n, cell, dx, dy, c0=C0):
    crpix = (float(n // 2 + 1),) * 2
    return O.pixel_to_skycoord_sin(crpix[0] + dx, crpix[1] + dy, crpix, (-cell / D2R, cell / D2R),
                                   (c0[0] / D2R, c0[1] / D2R))


def _image(n, cell, precision, centre, data=None):
    import torch
    import arlmi
    dt = torch.float32 if precision == 32 else torch.float64
    pc = None if centre is None else arlmi.SkyCoord(*centre)
    im = arlmi.create_image(npixel=n, cellsize=cell, frequency=[1e8], channel_bandwidth=[1e6], dtype=dt,
                            phasecentre=pc)
    if data is not None:
        im.data = torch.as_tensor(data).to(device=im.data.device, dtype=dt).reshape(im.data.shape)
    return im


def _assert_image(a, b, precision):
    l2, inner, full = image_errors(a, b)
    assert inner < TOL[precision] and full < FULL[precision], (l2, inner, full)


@pytest.mark.parametrize("precision", [64, 32])
def test_invert_predict_shifted_parity(gpu, precision):
    import arlmi
    rng = numpy.random.default_rng(11)
    n, cell, S = 256, 5e-4, 8
    d = random_rows(rng, 20000, n, cell, S)
    if precision == 32:
        d["vis"] = d["vis"].astype(numpy.complex64).astype(complex)
        d["imaging_weight"] = d["imaging_weight"].astype(numpy.float32).astype(float)
    c1 = _centre(n, cell, -9, 5)
    vis = to_device_vis(d, precision)
    vis.phasecentre = arlmi.SkyCoord(*C0)
    im = _image(n, cell, precision, c1)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=S)
    g = O.ImageGeom(n, cell)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    dirty, sw = arlmi.invert_2d(vis, im, gcfcf=(gcf, cf))
    rdirty, rsw = O.invert_2d(d, g, ogcf, ocf, cfg, centres=(C0, c1))
    _assert_image(dirty.numpy(), rdirty, precision)
    assert rel_err(sw, rsw) < 1e-12
    # the PSF: unit visibilities rotated like data (ARL sets vis = 1 before the shift)
    psf, _ = arlmi.invert_2d(vis, im, dopsf=True, gcfcf=(gcf, cf))
    rpsf, _ = O.invert_2d(d, g, ogcf, ocf, cfg, dopsf=True, centres=(C0, c1))
    _assert_image(psf.numpy(), rpsf, precision)
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, rng.integers(60, 196, 20), rng.integers(60, 196, 20)] = rng.uniform(0.5, 2.0, 20)
    pv = arlmi.predict_2d(vis, _image(n, cell, precision, c1, model), gcfcf=(gcf, cf))
    rpv = O.predict_2d(d, model, g, ogcf, ocf, cfg, centres=(C0, c1))
    assert rel_err(pv.vis.cpu().numpy(), rpv) < TOL[precision]
    assert (pv.phasecentre.ra, pv.phasecentre.dec) == (im.phasecentre.ra, im.phasecentre.dec)
    assert (vis.phasecentre.ra, vis.phasecentre.dec) == C0


@pytest.mark.parametrize("P", [1, 5])
def test_wstack_shifted_parity(gpu, P):
    """The w-stack imager shifts each slab's rows on their slab-mean-shifted w, as
    invert_wstack_single / predict_wstack_single call invert_2d / predict_2d: the list
    workflows with P slabs (and w != 0) against the oracle's."""
    import arlmi
    rng = numpy.random.default_rng(12)
    n, cell, S = 256, 5e-4, 8
    d = random_rows(rng, 8000, n, cell, S, wmax=100.0)
    c1 = _centre(n, cell, 6, -11)
    vis = to_device_vis(d, 64)
    vis.phasecentre = arlmi.SkyCoord(*C0)
    im = _image(n, cell, 64, c1)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=S)
    g = O.ImageGeom(n, cell)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    (dirty, sw), = arlmi.invert_list_serial_workflow([vis], [im], context="wstack", vis_slices=P, gcfcf=[(gcf, cf)])
    rdirty, rsw = O.invert_list_wstack(d, g, ogcf, ocf, cfg, P, centres=(C0, c1))
    _assert_image(dirty.numpy(), rdirty, 64)
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, rng.integers(60, 196, 10), rng.integers(60, 196, 10)] = 1.0
    pv, = arlmi.predict_list_serial_workflow([vis], [_image(n, cell, 64, c1, model)], context="wstack",
                                             vis_slices=P, gcfcf=[(gcf, cf)])
    rpv = O.predict_list_wstack(d, model, g, ogcf, ocf, cfg, P, centres=(C0, c1))
    assert rel_err(pv.vis.cpu().numpy(), rpv) < TOL[64]
    if P == 1:   # the single-slab functions take every row (ARL's one-slab list keeps w < 0 only)
        one, _ = arlmi.invert_wstack_single(vis, im, False, gcfcf=(gcf, cf))
        rone, _ = O.invert_wstack_single(d, g, ogcf, ocf, cfg, centres=(C0, c1))
        _assert_image(one.numpy(), rone, 64)
        pone = arlmi.predict_wstack_single(vis, _image(n, cell, 64, c1, model), gcfcf=(gcf, cf))
        rpone = O.predict_wstack_single(d, model, g, ogcf, ocf, cfg, centres=(C0, c1))
        assert rel_err(pone.vis.cpu().numpy(), rpone) < TOL[64]


def test_shift_is_phaserotate_and_no_ops(gpu):
    import torch
    import arlmi
    rng = numpy.random.default_rng(13)
    n, cell = 128, 1e-3
    d = random_rows(rng, 3000, n, cell, 6)
    vis = to_device_vis(d, 64)
    vis.phasecentre = arlmi.SkyCoord(*C0)
    c1 = _centre(n, cell, 4, 4)
    im = _image(n, cell, 64, c1)
    got = arlmi.shift_vis_to_image(vis, im)
    want = arlmi.phaserotate_visibility(vis, arlmi.image_phasecentre(im))
    assert torch.equal(got.vis, want.vis) and torch.equal(got.data["uvw"], vis.data["uvw"])
    assert got is not vis and (got.phasecentre.ra, got.phasecentre.dec) == (im.phasecentre.ra, im.phasecentre.dec)
    _, sv, _ = O.shift_vis_to_image(d, C0, c1)
    assert rel_err(got.vis.cpu().numpy(), sv) < 1e-11   # host l, m: two routes, 2e-15 apart
    # the same centre, an image without a sky position, a visibility without one: unchanged
    assert arlmi.shift_vis_to_image(vis, _image(n, cell, 64, C0)) is vis
    assert arlmi.shift_vis_to_image(vis, _image(n, cell, 64, None)) is vis
    vis.phasecentre = None
    assert arlmi.shift_vis_to_image(vis, im) is vis
    # images at the phase centre and without a sky position image identically
    vis.phasecentre = arlmi.SkyCoord(*C0)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=6)
    a, _ = arlmi.invert_2d(vis, _image(n, cell, 64, C0), gcfcf=(gcf, cf))
    b, _ = arlmi.invert_2d(vis, _image(n, cell, 64, None), gcfcf=(gcf, cf))
    _assert_image(a.numpy(), b.numpy(), 64)   # the LDS atomics add in varying order


def test_point_source_peak_and_adjoint(gpu):
    """A unit source 7 pixels right and 4 down of a shifted image centre peaks there on the GPU;
    Re<predict(x), y> = <x, adjoint(y)> through the shift."""
    import torch
    import arlmi
    n, cell, S = 256, 1e-3, 8
    c1 = _centre(n, cell, -5, 3)
    crpix = (float(n // 2 + 1),) * 2
    src = O.pixel_to_skycoord_sin(crpix[0] + 7, crpix[1] - 4, crpix, (-cell / D2R, cell / D2R),
                                  (c1[0] / D2R, c1[1] / D2R))
    rng = numpy.random.default_rng(14)
    d = random_rows(rng, 30000, n, cell, S)
    l, m, _ = O.skycoord_to_lmn(src[0], src[1], C0[0], C0[1])
    d["vis"] = O.simulate_point(d["uvw"], l, m)[:, None]
    vis = to_device_vis(d, 64)
    vis.phasecentre = arlmi.SkyCoord(*C0)
    im = _image(n, cell, 64, c1)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=S)
    dirty, _ = arlmi.invert_2d(vis, im, gcfcf=(gcf, cf))
    img = dirty.numpy()[0, 0]
    assert numpy.unravel_index(numpy.argmax(img), img.shape) == (n // 2 - 4, n // 2 + 7)
    imager = arlmi.Imager(vis, im, gcfcf=(gcf, cf))
    x = torch.as_tensor(rng.normal(size=(1, 1, n, n))).to("cuda")
    y = torch.as_tensor(rng.normal(size=(vis.nvis, 1)) + 1j * rng.normal(size=(vis.nvis, 1))).to("cuda")
    ax = imager.predict(vis, x)
    aty = imager.adjoint(vis, y)
    lhs = float(torch.sum(ax.real * y.real + ax.imag * y.imag))
    rhs = float(torch.sum(x * aty))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_batch_graph_with_shift(gpu):
    """A shifted w-stack observation in a captured batch: three replays, with eager work between
    them, each equal to the eager batch (image to float32 accumulation order, visibilities to
    rounding); the eager batch's image equals the single imager's."""
    import torch
    import arlmi
    rng = numpy.random.default_rng(15)
    n, cell, S = 256, 5e-4, 8
    d = random_rows(rng, 6000, n, cell, S, wmax=50.0)
    vis = to_device_vis(d, 32)
    vis.phasecentre = arlmi.SkyCoord(*C0)
    im = _image(n, cell, 32, _centre(n, cell, 3, -7))
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=S)
    model = torch.zeros((1, 1, n, n), dtype=torch.float32, device="cuda")
    model[0, 0, 100, 140] = 1.0
    task = {"vis": vis, "im": im, "gcfcf": (gcf, cf), "mode": "wstack", "vis_slices": 4, "model": model}
    eager = arlmi.BatchImager([task], graph=False)
    eager.run()
    torch.cuda.synchronize()
    ref_img, ref_pv = eager.images[0].cpu().numpy(), eager.predicted[0].cpu().numpy()
    (single, _), = arlmi.invert_list_serial_workflow([vis], [im], context="wstack", vis_slices=4, gcfcf=[(gcf, cf)])
    _assert_image(ref_img, single.numpy(), 32)
    batch = arlmi.BatchImager([task], graph=True)
    for _ in range(3):
        batch.images[0].zero_()
        batch.predicted[0].zero_()
        torch.cuda.synchronize()
        batch.run()
        torch.cuda.synchronize()
        _assert_image(batch.images[0].cpu().numpy(), ref_img, 32)
        assert rel_err(batch.predicted[0].cpu().numpy(), ref_pv) < 1e-6
