"""Invert and predict at the sizes whose column transforms are two-level (four-step, the grid rows
of a private plan left in digit-transposed order for its degridder): complex128 at 4096^2 and
complex64 at 8192^2, against the oracle (numpy.fft with ARL's shifts), in 2-D and w-stack mode."""

import numpy
import pytest

from oracle import arl_oracle as O
from helpers import random_rows, to_device_vis, device_image, image_errors, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,precision,mode,P", [(4096, 64, "2d", 1), (4096, 64, "wstack", 3),
                                                (8192, 32, "2d", 1)])
def test_two_level_columns(gpu, n, precision, mode, P):
    import arlmi
    rng = numpy.random.default_rng(n + P)
    cell, S = 1.0 / (n * 300.0), 8  # 300 wavelengths per pixel
    d = random_rows(rng, 6000, n, cell, S, wmax=150.0 if mode == "wstack" else 0.0)
    vis = to_device_vis(d, precision)
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=S)
    imager = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode=mode, vis_slices=P)
    img = imager.invert(vis)[0].cpu().numpy()
    g = O.ImageGeom(n, cell)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    if mode == "wstack":
        ref, _ = O.invert_list_wstack(d, g, ogcf, ocf, cfg, P)
    else:
        ref, _ = O.invert_2d(d, g, ogcf, ocf, cfg)
    l2, inner, full = image_errors(img, ref)
    assert inner < (1e-10 if precision == 64 else 1e-5), (l2, inner, full)
    if precision == 64:
        assert l2 < 1e-10, l2
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, rng.integers(n // 4, 3 * n // 4, 12), rng.integers(n // 4, 3 * n // 4, 12)] = 1.0
    mim = device_image(n, cell, precision, data=model)
    pv = imager.predict(vis, mim.data).cpu().numpy()
    if mode == "wstack":
        rpv = O.predict_list_wstack(d, model, g, ogcf, ocf, cfg, P)
    else:
        rpv = O.predict_2d(d, model, g, ogcf, ocf, cfg)
    assert rel_err(pv, rpv) < (1e-10 if precision == 64 else 1e-5)
    # the public GridData route keeps ARL's natural row order (transposed columns)
    if mode == "2d":
        gd = arlmi.create_griddata_from_image(mim)
        gd = arlmi.fft_image_to_griddata(mim, gd, gcf)
        rg = O.fft_image_to_griddata(model, ogcf)
        assert rel_err(gd.data.cpu().numpy(), rg) < (1e-12 if precision == 64 else 1e-5)
