"""GPU parity of the weight density and uniform reweighting against the FP64 oracle
(gridding.py:161-230, weighting.py:19-40, imaging_serial.py:426-473)."""

import numpy
import pytest

from oracle import arl_oracle as O
from helpers import rel_err, to_device_vis, device_image

pytestmark = pytest.mark.gpu


def _rows(seed, n, cell, nvis=20000, integer=False):
    rng = numpy.random.default_rng(seed)
    du = 1.0 / (n * cell)
    uvw = numpy.zeros((nvis, 3))
    # clustered so that pixels hold many rows; a few rows on the outermost pixels (allowed: no
    # kernel footprint is checked for weights)
    uvw[:, 0] = rng.normal(0, n / 8, nvis) * du
    uvw[:, 1] = rng.normal(0, n / 8, nvis) * du
    uvw[:, :2] = numpy.clip(uvw[:, :2], -(n // 2 - 1) * du, (n // 2 - 1) * du)
    uvw[:4, 0] = numpy.array([-(n // 2 - 1), n // 2 - 1.4, 0.0, 0.0]) * du
    uvw[:4, 1] = numpy.array([0.0, 0.0, n // 2 - 1.2, -(n // 2 - 1)]) * du
    wt = rng.integers(1, 6, (nvis, 1)).astype(float) if integer else rng.uniform(0.2, 3.0, (nvis, 1))
    return {"uvw": uvw, "vis": numpy.zeros((nvis, 1), dtype=complex), "imaging_weight": wt,
            "frequency": numpy.full(nvis, 1e8)}


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("integer", [True, False])
def test_grid_weights_and_reweight_parity(gpu, precision, integer):
    import arlmi
    n, cell = 256, 1e-3
    d = _rows(5, n, cell, integer=integer)
    if precision == 32:
        d["imaging_weight"] = d["imaging_weight"].astype(numpy.float32).astype(float)
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=8)
    g = O.ImageGeom(n, cell)
    _, _, cfg = O.create_pswf_convolutionfunction(g, 8, 8)
    vis = to_device_vis(d, precision)
    gd, sumwt = arlmi.grid_weight_to_griddata(vis, arlmi.create_griddata_from_image(im), cf)
    rgrid, rsumwt = O.grid_weight_to_griddata(d, g, cfg)
    got = gd.data.cpu().numpy()
    assert got.shape == rgrid.shape
    if integer:
        assert numpy.array_equal(got, rgrid)             # integer sums are exact in any order
        assert sumwt[0, 0] == rsumwt[0, 0]
    else:
        assert rel_err(got, rgrid) < 1e-13
        assert abs(sumwt[0, 0] - rsumwt[0, 0]) < 1e-12 * rsumwt[0, 0]
    vis = arlmi.griddata_reweight(vis, gd, cf)
    rw = O.griddata_reweight(d, rgrid, g, cfg)
    tol = 1e-14 if precision == 64 else 1e-7
    assert rel_err(vis.imaging_weight.cpu().numpy(), rw) < tol


def test_weight_visibility_and_list_workflow(gpu):
    import arlmi
    n, cell = 256, 1e-3
    d1, d2 = _rows(6, n, cell), _rows(7, n, cell, nvis=15000)
    im = device_image(n, cell, 64)
    g = O.ImageGeom(n, cell)
    _, _, cfg = O.create_pswf_convolutionfunction(g, 8, 8)
    v = arlmi.weight_visibility(to_device_vis(d1, 64), im)
    assert rel_err(v.imaging_weight.cpu().numpy(), O.weight_visibility(d1, g, cfg)) < 1e-14
    out = arlmi.weight_list_serial_workflow([to_device_vis(d1, 64), to_device_vis(d2, 64)], [im, im])
    merged, _ = O.griddata_merge_weights([O.grid_weight_to_griddata(d1, g, cfg), O.grid_weight_to_griddata(d2, g, cfg)])
    for o, d in zip(out, (d1, d2)):
        assert rel_err(o.imaging_weight.cpu().numpy(), O.griddata_reweight(d, merged, g, cfg)) < 1e-14
    # the uniform-weighting identity through the GPU: reweighted weights grid to 1
    gd, _ = arlmi.grid_weight_to_griddata(v, arlmi.create_griddata_from_image(im), arlmi.create_pswf_convolutionfunction(im)[1])
    dens = gd.data.real.cpu().numpy()
    occ = dens != 0
    assert numpy.allclose(dens[occ], 1.0, rtol=0, atol=1e-12)


def test_weights_off_grid_raise_and_leave_weights(gpu):
    import arlmi
    n, cell = 256, 1e-3
    d = _rows(8, n, cell, nvis=1000)
    d["uvw"][10, 0] = (n // 2 + 3) / (n * cell)  # pu >= N
    im = device_image(n, cell, 64)
    _, cf = arlmi.create_pswf_convolutionfunction(im)
    vis = to_device_vis(d, 64)
    with pytest.raises(AssertionError):
        arlmi.grid_weight_to_griddata(vis, arlmi.create_griddata_from_image(im), cf)
    gd = arlmi.create_griddata_from_image(im)
    gd.data = gd.data.to(dtype=__import__("torch").complex128) + 1.0
    before = vis.imaging_weight.cpu().numpy().copy()
    with pytest.raises(AssertionError):
        arlmi.griddata_reweight(vis, gd, cf)
    assert numpy.array_equal(vis.imaging_weight.cpu().numpy(), before)


def test_full_size_uniform_identity(gpu):
    """The w-stack workload's 29.7 M rows (unit float32 weights) on its 2048^2 grid: the density
    counts every row exactly; after uniform weighting the new weights grid to 1 on every occupied
    pixel (within the float32 rounding of 1/count) and to 0 elsewhere."""
    import torch
    import arlmi
    from arlmi.simulation import config2
    vis, im, _ = config2(precision=32, sky=False)
    vis.data["imaging_weight"].fill_(1.0)
    _, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=8)
    gd0, sw0 = arlmi.grid_weight_to_griddata(vis, arlmi.create_griddata_from_image(im), cf)
    count = gd0.data.real
    occ = count != 0
    assert float(count.sum()) == float(vis.nvis) and float(numpy.sum(sw0)) == float(vis.nvis)
    arlmi.weight_visibility(vis, im, gcfcf=(None, cf))
    gd, _ = arlmi.grid_weight_to_griddata(vis, arlmi.create_griddata_from_image(im), cf)
    dens = gd.data.real
    assert torch.equal(dens != 0, occ)
    assert float(torch.max(torch.abs(dens[occ] - 1.0))) < 1e-6
    print("full-size weighting: %d rows on %d occupied pixels" % (vis.nvis, int(occ.sum())))
