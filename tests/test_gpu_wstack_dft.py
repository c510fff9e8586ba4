"""The HIP w-stack path against the direct DFT of a point source with w != 0 (ARL's recipe,
tests/workflows/test_imaging_serial.py:153-187, 232-243), against the oracle on the same rows,
and the list workflows' normalize semantics (imaging_serial.py:166-197).

Thresholds: ARL's max |residual image| < 1.0 on 100 Jy for 101 slabs (PSWF S=6, os=8) and for
the w-stack + w-projection hybrid (11 slabs, AW kernels); the HIP residual within 2 % of the
oracle's residual (float64: its residual image agrees with the oracle's to 1e-8 of the flux;
float32: 1e-5).
"""

import numpy
import pytest

from oracle import arl_oracle as O
from helpers import device_image, rel_err, random_rows, to_device_vis, oracle_geom, image_errors
from test_oracle_wstack_dft import (N, CELL, FLUX, PX, PY, low_like_rows, with_component, model_image,
                                    residual_image, wstack)

pytestmark = pytest.mark.gpu

RES_TOL = {64: 1e-8, 32: 1e-5}


def _resid_diff(a, b, precision):
    """max |a - b| over the full field (float64) or the inner half of each axis (float32, whose
    rounding ARL's grid correction multiplies by up to 6e4 at the corners)."""
    if precision == 32:
        q = N // 4
        a, b = a[..., q:N - q, q:N - q], b[..., q:N - q, q:N - q]
    return float(numpy.max(numpy.abs(a - b)))


def _gpu_residual(d, precision, gcfcf, P):
    """Predict (w-stack list workflow) subtracted from the DFT rows, then inverted; also the dirty
    image of the DFT rows."""
    import arlmi
    vis = to_device_vis(d, precision)
    mim = device_image(N, CELL, precision, data=model_image())
    pv, = arlmi.predict_list_serial_workflow([vis], [mim], context="wstack", vis_slices=P, gcfcf=[gcfcf])
    res = dict(d)
    res["vis"] = d["vis"] - pv.vis.cpu().numpy()
    rvis = to_device_vis(res, precision)
    im = device_image(N, CELL, precision)
    (dirty, _), = arlmi.invert_list_serial_workflow([rvis], [im], context="wstack", vis_slices=P, gcfcf=[gcfcf])
    (img, _), = arlmi.invert_list_serial_workflow([vis], [im], context="wstack", vis_slices=P, gcfcf=[gcfcf])
    return dirty.numpy(), img.numpy(), pv.vis.cpu().numpy()


@pytest.fixture(scope="module")
def obs():
    return with_component(low_like_rows(13))


@pytest.mark.parametrize("precision", [64, 32])
def test_wstack_point_source_dft(gpu, obs, precision):
    import arlmi
    im = device_image(N, CELL, precision)
    gcfcf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=6)
    dirty, img, pv = _gpu_residual(obs, precision, gcfcf, 101)
    g = O.ImageGeom(N, CELL)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, 6)
    rdirty, _ = residual_image(obs, *wstack(g, ogcf, ocf, cfg, 101))
    assert numpy.max(numpy.abs(dirty)) < 1.0                                   # ARL's threshold
    assert abs(numpy.max(numpy.abs(dirty)) - numpy.max(numpy.abs(rdirty))) < 0.02 * numpy.max(numpy.abs(rdirty))
    assert _resid_diff(dirty, rdirty, precision) < RES_TOL[precision] * FLUX
    assert numpy.unravel_index(numpy.argmax(img), img.shape) == (0, 0, PY, PX)
    assert abs(img[0, 0, PY, PX] - FLUX) < 0.01 * FLUX
    rpv = O.predict_list_wstack(obs, model_image(), g, ogcf, ocf, cfg, 101)
    assert rel_err(pv, rpv) < (1e-10 if precision == 64 else 1e-5)


@pytest.mark.parametrize("precision", [64, 32])
def test_wstack_wprojection_point_source_dft(gpu, obs, precision):
    """The hybrid: 11 slabs with AW kernels for the residual w inside each (the slab-mean w shift
    is what places a row on its kernel plane)."""
    import arlmi
    im = device_image(N, CELL, precision)
    gcfcf = arlmi.create_awterm_convolutionfunction(im, nw=81, wstep=16.0, oversampling=8, support=16,
                                                    use_aaf=True)
    dirty, img, pv = _gpu_residual(obs, precision, gcfcf, 11)
    g = O.ImageGeom(N, CELL)
    ogcf, ocf, cfg = O.create_awterm_convolutionfunction(g, nw=81, wstep=16.0, oversampling=8, support=16,
                                                         use_aaf=True)
    rdirty, _ = residual_image(obs, *wstack(g, ogcf, ocf, cfg, 11))
    assert numpy.max(numpy.abs(dirty)) < 1.0
    assert _resid_diff(dirty, rdirty, precision) < RES_TOL[precision] * FLUX
    rpv = O.predict_list_wstack(obs, model_image(), g, ogcf, ocf, cfg, 11)
    assert rel_err(pv, rpv) < (1e-10 if precision == 64 else 1e-5)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("context,P", [("wstack", 5), ("2d", 1)])
def test_list_invert_normalize_false(gpu, precision, context, P):
    """normalize=False: sum_p sw_p raw_p / sum_p sw_p (w-stack), the plain image in 2-D, as the
    oracle (imaging_serial.py:166-197) gives it."""
    import arlmi
    rng = numpy.random.default_rng(22)
    n, cell, S = 128, 2e-3, 8
    d = random_rows(rng, 6000, n, cell, S, wmax=60.0 if context == "wstack" else 0.0)
    vis = to_device_vis(d, precision)
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=S)
    g = oracle_geom(n, cell)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    for normalize in (False, True):
        (img, sw), = arlmi.invert_list_serial_workflow([vis], [im], context=context, vis_slices=P,
                                                       normalize=normalize, gcfcf=[(gcf, cf)])
        if context == "wstack":
            ref, rsw = O.invert_list_wstack(d, g, ogcf, ocf, cfg, P, normalize=normalize)
        else:
            ref, rsw = O.invert_2d(d, g, ogcf, ocf, cfg, normalize=normalize)
        l2, inner, full = image_errors(img.numpy(), ref)
        assert inner < (1e-10 if precision == 64 else 1e-5) and l2 < (1e-10 if precision == 64 else 1e-3), \
            (normalize, l2, inner, full)
        assert abs(sw[0, 0] - rsw[0, 0]) < 1e-9 * rsw[0, 0]


def _wprojection_residual(d, precision, gcfcf):
    """2-D (w-projection) predict subtracted from the DFT rows, then inverted
    (test_imaging_serial.py:153-170 with context '2d' and an AW gcfcf)."""
    import arlmi
    vis = to_device_vis(d, precision)
    mim = device_image(N, CELL, precision, data=model_image())
    pv = arlmi.predict_2d(vis, mim, gcfcf=gcfcf)
    res = dict(d)
    res["vis"] = d["vis"] - pv.vis.cpu().numpy()
    im = device_image(N, CELL, precision)
    dirty, _ = arlmi.invert_2d(to_device_vis(res, precision), im, gcfcf=gcfcf)
    return dirty.numpy(), pv.vis.cpu().numpy()


@pytest.mark.parametrize("precision", [64, 32])
def test_wprojection_arl_sizes(gpu, obs, precision):
    """ARL's own w-projection settings (test_imaging_serial.py:121-131, 222-230): nw=61 kernel
    planes 16 wavelengths apart, support 64, oversampling 8, and the same kernel clipped at 1e-3
    by apply_bounding_box_convolutionfunction (convolution_functions.py:159-184); ARL's
    threshold 1.0 on 100 Jy for both.  A 2,000-row subset against the oracle with the same
    kernels."""
    import arlmi
    im = device_image(N, CELL, precision)
    gcf, cf = arlmi.create_awterm_convolutionfunction(im, nw=61, wstep=16.0, oversampling=8, support=64,
                                                      use_aaf=True)
    clipped = arlmi.apply_bounding_box_convolutionfunction(cf, fractional_level=1e-3)
    assert clipped.support < 64
    g = O.ImageGeom(N, CELL)
    ogcf, ocf, cfg = O.create_awterm_convolutionfunction(g, nw=61, wstep=16.0, oversampling=8, support=64,
                                                         use_aaf=True)
    ocl = O.apply_bounding_box_convolutionfunction(ocf, 1e-3)
    sub = {k: v[:2000] for k, v in obs.items()}
    model = numpy.zeros((1, 1, N, N))
    model[0, 0, PY, PX] = FLUX
    for cfx, ocfx in ((cf, ocf), (clipped, ocl)):
        dirty, _ = _wprojection_residual(obs, precision, (gcf, cfx))
        assert numpy.max(numpy.abs(dirty)) < 1.0, (cfx.support, numpy.max(numpy.abs(dirty)))
        S = ocfx.shape[-1]
        ocfg = O.CFGeom(g, 8, S, nz=61, zstep=16.0)
        vis = to_device_vis(sub, precision)
        mim = device_image(N, CELL, precision, data=model)
        pv = arlmi.predict_2d(vis, mim, gcfcf=(gcf, cfx)).vis.cpu().numpy()
        rpv = O.predict_2d(sub, model, g, ogcf, ocfx, ocfg)
        assert rel_err(pv, rpv) < (1e-10 if precision == 64 else 1e-5), S
        img, _ = arlmi.invert_2d(vis, im, gcfcf=(gcf, cfx))
        rimg, _ = O.invert_2d(sub, g, ogcf, ocfx, ocfg)
        l2, inner, full = image_errors(img.numpy(), rimg)
        assert inner < (1e-10 if precision == 64 else 1e-5), (S, l2, inner, full)
