"""The drop-in route of the ARL-side binding: ARL's own (gcf, cf) arrays (here the oracle's, which
restate ARL's create_pswf_convolutionfunction / create_awterm_convolutionfunction) converted with
convolutionfunction_from_arrays and passed to invert_2d / predict_2d give the oracle's images and
visibilities for that kernel (a caller's gcfcf is used, not a default one)."""

import numpy
import pytest

from oracle import arl_oracle as O
from helpers import random_rows, to_device_vis, device_image, image_errors, rel_err

pytestmark = pytest.mark.gpu

TOL = {64: 1e-10, 32: 1e-5}


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("kind", ["pswf", "aw"])
def test_foreign_gcfcf_invert_predict(gpu, precision, kind):
    import arlmi
    rng = numpy.random.default_rng(31)
    n, cell = 256, 1e-3
    g = O.ImageGeom(n, cell)
    if kind == "pswf":
        ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, 6)   # ARL's default S=6, os=8
        wmax = 0.0
    else:
        ogcf, ocf, cfg = O.create_awterm_convolutionfunction(g, nw=9, wstep=20.0, oversampling=8, support=12,
                                                             use_aaf=True)
        wmax = 75.0
    S = ocf.shape[-1]
    d = random_rows(rng, 8000, n, cell, S, wmax=wmax)
    vis = to_device_vis(d, precision)
    im = device_image(n, cell, precision)
    crpix = [S // 2 + 1.0, S // 2 + 1.0, 5.0, 5.0, float(cfg.nz // 2 + 1), 1.0, 1.0]
    cdelt = [cfg.du, cfg.dv, cfg.ddu, cfg.ddv, cfg.zstep, 1.0, 1e6]
    gcf, cf = arlmi.convolutionfunction_from_arrays(im, ogcf, ocf, crpix, cdelt)
    assert (cf.separable is not None) == (kind == "pswf")
    img, sw = arlmi.invert_2d(vis, im, gcfcf=(gcf, cf))
    ref, rsw = O.invert_2d(d, g, ogcf, ocf, cfg)
    l2, inner, full = image_errors(img.data.cpu().numpy(), ref)
    assert inner < TOL[precision] and (precision == 32 or l2 < TOL[64]), (l2, inner, full)
    assert abs(sw[0, 0] - rsw[0, 0]) < 1e-9 * rsw[0, 0]
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, rng.integers(60, 196, 20), rng.integers(60, 196, 20)] = rng.uniform(0.5, 2.0, 20)
    mim = device_image(n, cell, precision, data=model)
    pv = arlmi.predict_2d(vis, mim, gcfcf=(gcf, cf))
    rpv = O.predict_2d(d, model, g, ogcf, ocf, cfg)
    assert rel_err(pv.vis.cpu().numpy(), rpv) < TOL[precision]
