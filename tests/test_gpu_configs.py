"""The workloads at their full sizes.

Configuration 1 (plain 2-D imaging, 120,960 rows) is small enough for the FP64 oracle, so it is
compared directly.  The larger ones are checked through properties that hold at any size:
float32 against float64 (the precision target, 1e-5 relative image error over the inner field),
the adjoint identity <predict(x), y> = <x, invert(y)> of predict and its hand-written adjoint,
linearity, exact sums of unit weights, and finite differences of the gradient workload.
"""

import numpy
import pytest

from oracle import arl_oracle as O
from helpers import rel_err, image_errors

pytestmark = pytest.mark.gpu


def host_rows(vis):
    d = {k: vis.data[k].cpu().numpy() for k in ("uvw", "vis", "imaging_weight", "frequency")}
    d["vis"] = d["vis"].astype(numpy.complex128)
    d["imaging_weight"] = d["imaging_weight"].astype(numpy.float64)
    return d


def dot_test(imager, vis, n, rng, dt, inner=False):
    """Relative mismatch of Re<predict(x), y> and <x, adjoint(y)> (random x, y).  inner: x only
    over the central half of each axis, where the grid correction is below ~10 (at the corners it
    reaches ~6e4 and multiplies float32 rounding by as much)."""
    import torch
    xs = rng.normal(size=(1, 1, n, n))
    if inner:
        q = n // 4
        mask = numpy.zeros((n, n))
        mask[q:n - q, q:n - q] = 1.0
        xs = xs * mask
    x = torch.as_tensor(xs).to("cuda", dt)
    y = torch.as_tensor(rng.normal(size=(vis.nvis, 1)) + 1j * rng.normal(size=(vis.nvis, 1))).to(
        "cuda", vis.vis.dtype)
    ax = imager.predict(vis, x)
    aty = imager.adjoint(vis, y)
    lhs = float(torch.sum(ax.real.double() * y.real.double() + ax.imag.double() * y.imag.double()))
    rhs = float(torch.sum(x.double() * aty.double()))
    return abs(lhs - rhs) / abs(lhs)


def test_config1_full_size_against_oracle(gpu):
    """64 antennas, 120,960 rows, 512^2, PSWF S=8, w = 0, complex128: invert (image, PSF) and
    predict against the FP64 oracle, and the direct DFT of the 10 sources."""
    import arlmi
    from arlmi.simulation import config1
    vis, im, comps = config1(precision=64)
    assert vis.nvis == 120960
    d = host_rows(vis)
    g = O.ImageGeom(512, 2.5e-4, frequency=1e8, channel_bandwidth=1e6)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, 8)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=8)
    dirty, sw = arlmi.invert_2d(vis, im, gcfcf=(gcf, cf))
    rdirty, rsw = O.invert_2d(d, g, ogcf, ocf, cfg)
    l2, inner, full = image_errors(dirty.numpy(), rdirty)
    assert l2 < 1e-10 and full < 1e-8, (l2, inner, full)
    assert sw[0, 0] == rsw[0, 0] == 120960.0
    psf, _ = arlmi.invert_2d(vis, im, dopsf=True, gcfcf=(gcf, cf))
    rpsf, _ = O.invert_2d(d, g, ogcf, ocf, cfg, dopsf=True)
    assert image_errors(psf.numpy(), rpsf)[2] < 1e-8
    model = numpy.zeros((1, 1, 512, 512))
    for c in comps:
        model[0, 0, int(round(256 + c.m / 2.5e-4)), int(round(256 - c.l / 2.5e-4))] += c.flux[0, 0]
    from helpers import device_image
    pv = arlmi.predict_2d(vis, device_image(512, 2.5e-4, 64, data=model), gcfcf=(gcf, cf))
    assert rel_err(pv.vis.cpu().numpy(), O.predict_2d(d, model, g, ogcf, ocf, cfg)) < 1e-10
    # the simulated visibilities are the direct DFT of the components plus noise
    dft = numpy.zeros((vis.nvis, 1), dtype=complex)
    for c in comps:
        dft[:, 0] += c.flux[0, 0] * O.simulate_point(d["uvw"], c.l, c.m)
    assert numpy.std(d["vis"] - dft) < 2e-3


def test_config2_full_size_precision_and_adjoint(gpu):
    """29.7 M rows, 2048^2, 64 w slabs: the float32 form within 1e-5 (inner field) of the float64
    form, exact unit-weight sums, and the adjoint identity of the float64 imager."""
    import torch
    import arlmi
    from arlmi import simulation
    out = {}
    for precision in (64, 32):
        vis, im, comps = simulation.config2(precision=precision)
        gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=8)
        imager = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode="wstack", vis_slices=64)
        img, _ = imager.invert(vis)
        out[precision] = (img.double().cpu().numpy(), imager.sumwt()[0, 0], imager.plan.stats()["nmember"])
        if precision == 64:
            rng = numpy.random.default_rng(9)
            assert dot_test(imager, vis, 2048, rng, torch.float64) < 1e-10
        del imager, vis
        torch.cuda.empty_cache()
    l2, inner, full = image_errors(out[32][0], out[64][0])
    assert inner < 1e-5, (l2, inner, full)
    # unit weights: the sum of weights counts every (row, slab) membership once
    assert out[64][1] == out[32][1] == float(out[64][2])


def test_config2_aw32_full_size(gpu):
    """The same 29.7 M rows and 64 slabs with the w-projection part of the configuration at its
    largest support: an AW kernel of S=32 (11 kernel w planes a quarter of a slab width apart, os=8)
    for each row's residual w within its slab.  The float32 form within 1e-5 (inner field) of the
    float64 form, and the adjoint identity of the float64 imager (table gridder and degridder)."""
    import torch
    import arlmi
    from arlmi import simulation
    out = {}
    for precision in (64, 32):
        vis, im, comps = simulation.config2(precision=precision, support=32)
        wmax = float(vis.data["uvw"][:, 2].abs().max())
        width = 2.0 * wmax / 63
        gcf, cf = arlmi.create_awterm_convolutionfunction(im, nw=11, wstep=width / 4.0, oversampling=8, support=32)
        imager = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode="wstack", vis_slices=64)
        img, _ = imager.invert(vis)
        out[precision] = img.double().cpu().numpy()
        if precision == 64:
            rng = numpy.random.default_rng(10)
            assert dot_test(imager, vis, 2048, rng, torch.float64) < 1e-10
        del imager, vis, cf
        torch.cuda.empty_cache()
    l2, inner, full = image_errors(out[32], out[64])
    assert inner < 1e-5, (l2, inner, full)


def test_config4_full_size_gradients(gpu):
    """1.95 M rows from a 128-antenna BlockVisibility, 1024^2: image and gain gradients of the
    float64 form against central finite differences (step 1e-3) at 200 coordinates (100 image
    pixels, the real and imaginary parts of 50 gains)."""
    import torch
    import arlmi
    from arlmi.simulation import config4
    c = config4(precision=64)
    vis, im, gt = c["vis"], c["im"], c["gaintable"]
    assert vis.nvis == 1950720
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=8)
    imager = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode="wstack", vis_slices=16)
    i1, i2 = arlmi.gain_row_indices(vis, gt)
    truth = torch.as_tensor(c["image"]).to("cuda")
    with torch.no_grad():
        vis.data["vis"] = arlmi.predict_with_gains(truth, gt.gain, imager, vis, i1, i2).contiguous()
    img = (0.8 * truth).requires_grad_(True)
    g = (gt.gain * 1.05).requires_grad_(True)
    arlmi.visibility_loss(img, g, imager, vis, i1, i2).backward()

    def L(x, gg):
        with torch.no_grad():
            return float(arlmi.visibility_loss(x, gg, imager, vis, i1, i2))

    rng = numpy.random.default_rng(12)
    h = 1e-3
    si, sg = float(img.grad.abs().max()), float(g.grad.abs().max())
    x0, g0 = img.detach(), g.detach()
    for _ in range(100):
        y, x = rng.integers(300, 724, 2)
        xp, xm = x0.clone(), x0.clone()
        xp[0, 0, y, x] += h
        xm[0, 0, y, x] -= h
        fd = (L(xp, g0) - L(xm, g0)) / (2 * h)
        assert abs(fd - float(img.grad[0, 0, y, x])) <= 1e-4 * si
    flat, gflat = g0.reshape(-1), g.grad.reshape(-1)
    for k in rng.choice(flat.numel(), 50, replace=False):
        for unit in (1.0, 1j):
            gp, gm = flat.clone(), flat.clone()
            gp[k] += h * unit
            gm[k] -= h * unit
            fd = (L(x0, gp.reshape(g0.shape)) - L(x0, gm.reshape(g0.shape))) / (2 * h)
            an = float(gflat[k].real if unit == 1.0 else gflat[k].imag)
            assert abs(fd - an) <= 1e-4 * sg


def test_config5_chunk_8192(gpu):
    """One chunk of the streaming workload at its full image size: 8192^2, 128 w slabs, PSWF S=8,
    float32 (two integrations, 2.47 M rows before selection): adjoint identity and linearity."""
    import torch
    import arlmi
    from arlmi.simulation import config5
    vis, im = config5(precision=32, integrations=2, first=200)
    assert vis.nvis > 1_000_000
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=8)
    imager = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode="wstack", vis_slices=128)
    rng = numpy.random.default_rng(13)
    err = dot_test(imager, vis, 8192, rng, torch.float32, inner=True)
    print("config5 chunk: %d rows, adjoint mismatch %.2e" % (vis.nvis, err))
    assert err < 1e-5
    n = 8192
    x = torch.zeros((1, 1, n, n), dtype=torch.float32, device="cuda")
    y = torch.zeros_like(x)
    x[0, 0, rng.integers(n // 4, 3 * n // 4, 50), rng.integers(n // 4, 3 * n // 4, 50)] = 1.0
    y[0, 0, rng.integers(n // 4, 3 * n // 4, 50), rng.integers(n // 4, 3 * n // 4, 50)] = 0.5
    px = imager.predict(vis, x).clone()
    py = imager.predict(vis, y).clone()
    pxy = imager.predict(vis, 2.0 * x + 3.0 * y)
    assert rel_err((pxy).cpu().numpy(), (2.0 * px + 3.0 * py).cpu().numpy()) < 1e-5
