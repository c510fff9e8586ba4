"""GPU parity of BlockVisibility -> rows, the scalar gain convention, its adjoint and the
differentiable predict + gains against the FP64 oracle.

Tolerances: the row conversion is a gather and a product in ARL's operation order, so every
column is compared bit for bit.  Gain products: 1e-14 (float64) and 1e-6 (float32) relative to
the peak.  Gradients: central finite differences of the float64 path within 1e-4, and the
float32 path against the float64 oracle adjoint within 1e-4 relative to the peak.
"""

import numpy
import pytest

from oracle import arl_oracle as O
from helpers import rel_err, image_errors
from test_oracle_vis import small_block, _array, _gains

pytestmark = pytest.mark.gpu


def device_block(b, precision):
    import arlmi
    return arlmi.BlockVisibility.from_arrays(b['uvw'], vis=b['vis'], weight=b['weight'],
                                             imaging_weight=b['imaging_weight'], time=b['time'],
                                             integration_time=b['integration_time'], frequency=b['frequency'],
                                             channel_bandwidth=b['channel_bandwidth'], precision=precision)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("nants,ntimes,nchan", [(7, 3, 3), (2, 1, 1), (40, 2, 5), (1, 2, 2)])
def test_block_to_rows_bitexact(gpu, precision, nants, ntimes, nchan):
    import arlmi
    b, *_ = small_block(seed=nants, nants=nants, ntimes=ntimes, nchan=nchan)
    ref = O.convert_blocks(b)
    vis = arlmi.convert_blockvisibility_to_visibility(device_block(b, precision))
    cdt = numpy.complex64 if precision == 32 else numpy.complex128
    rdt = numpy.float32 if precision == 32 else numpy.float64
    assert vis.nvis == len(ref['antenna1'])
    got = {k: v.cpu().numpy() for k, v in vis.data.items()}
    assert numpy.array_equal(got['uvw'], ref['uvw'].reshape(-1, 3))
    assert numpy.array_equal(got['vis'], ref['vis'].astype(cdt))
    assert numpy.array_equal(got['weight'], ref['weight'].astype(rdt))
    assert numpy.array_equal(got['imaging_weight'], ref['imaging_weight'].astype(rdt))
    for k in ('time', 'frequency', 'channel_bandwidth', 'integration_time', 'antenna1', 'antenna2'):
        assert numpy.array_equal(got[k], ref[k]), k
    assert numpy.array_equal(vis.cindex.cpu().numpy(), ref['cindex'])
    if vis.nvis:
        back = arlmi.decoalesce_visibility(vis)
        rb = O.decoalesce_visibility(ref, b)
        assert numpy.array_equal(back.data['vis'].cpu().numpy(), rb['vis'].astype(cdt))


@pytest.mark.parametrize("precision", [64, 32])
def test_apply_gaintable_parity(gpu, precision):
    """GPU apply_gaintable against the oracle, plus ARL's own assertions
    (test_calibration_operations.py:92-135)."""
    import torch
    import arlmi
    b, *_ = small_block(seed=11, nants=9, ntimes=4, nchan=3)
    b['vis'] *= 100.0
    rng = numpy.random.default_rng(11)
    tol = 1e-14 if precision == 64 else 1e-6
    cdt = torch.complex64 if precision == 32 else torch.complex128
    bv = device_block(b, precision)
    gt = arlmi.create_gaintable_from_blockvisibility(bv)
    assert gt.ntimes == 4 and gt.gain.shape == (4, 9, 3, 1, 1)
    g = _gains(rng, 4, 9, 3, amp=0.1, phase=0.1)
    gt.gain = torch.as_tensor(g.reshape(gt.gain.shape)).to(device=gt.gain.device, dtype=cdt)
    wt = numpy.ones(g.shape)
    wt[1, 4, :] = 0.0
    gt.weight = torch.as_tensor(wt.reshape(gt.gain.shape)).to(gt.gain.device)
    slots = numpy.arange(4)
    original = bv.data['vis'].cpu().numpy().astype(numpy.complex128)
    arlmi.apply_gaintable(bv, gt)
    got = bv.data['vis'].cpu().numpy()
    gref = g.astype(numpy.complex64).astype(complex) if precision == 32 else g
    ref = O.apply_gaintable(original, gref, wt, slots)
    assert rel_err(got, ref) < tol
    assert numpy.max(numpy.abs(got - original)) > 10.0
    assert numpy.all(got[1, 4] == 0) and numpy.all(got[1, :, 4] == 0)
    # inverse undoes it where the gain weight is non-zero
    arlmi.apply_gaintable(bv, gt, inverse=True)
    back = bv.data['vis'].cpu().numpy()
    keep = numpy.ones(back.shape, dtype=bool)
    keep[1, 4] = False
    keep[1, :, 4] = False
    assert numpy.max(numpy.abs(back[keep] - original[keep])) < (1e-12 if precision == 64 else 1e-3)
    # zero gains with inverse leave the visibilities unchanged
    bz = device_block(b, precision)
    gz = arlmi.create_gaintable_from_blockvisibility(bz)
    gz.gain = gz.gain * 0.0
    before = bz.data['vis'].cpu().numpy()
    arlmi.apply_gaintable(bz, gz, inverse=True)
    assert numpy.max(numpy.abs(bz.data['vis'].cpu().numpy()[:, 0, 1] - before[:, 0, 1])) < 1e-12


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("ngain", [12, 5000])
def test_gains_rows_and_adjoint_parity(gpu, precision, ngain):
    """Row gains and their adjoint (LDS pre-summed path for <= 4096 gains, direct path above)."""
    import torch
    import arlmi
    rng = numpy.random.default_rng(ngain)
    n = 30000
    v = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    G = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    i1 = rng.integers(0, ngain, n).astype(numpy.int32)
    i2 = rng.integers(0, ngain, n).astype(numpy.int32)
    i1[:5] = -1  # rows with no gain
    g = _gains(rng, 1, ngain, 1, amp=0.2, phase=0.5).reshape(-1)
    cdt = torch.complex64 if precision == 32 else torch.complex128
    tol = 1e-13 if precision == 64 else 2e-6
    dv = torch.as_tensor(v).to(device=gpu, dtype=cdt)
    dG = torch.as_tensor(G).to(device=gpu, dtype=cdt)
    dg = torch.as_tensor(g).to(device=gpu, dtype=cdt)
    d1 = torch.as_tensor(i1).to(gpu)
    d2 = torch.as_tensor(i2).to(gpu)
    out = arlmi.apply_gains_rows(dv, d1, d2, dg).cpu().numpy()
    ok = (i1 >= 0)[:, None]
    ref = numpy.where(ok, O.apply_gains_rows(v, numpy.maximum(i1, 0), i2, g), 0)
    assert rel_err(out, ref) < tol
    assert numpy.all(out[:5] == 0)
    gv, gg = arlmi.apply_gains_rows_adjoint(dv, dG, d1, d2, dg)
    sel = i1 >= 0
    rgv, rgg = O.apply_gains_rows_adjoint(v[sel], G[sel], i1[sel], i2[sel], g)
    gv = gv.cpu().numpy()
    assert rel_err(gv[sel], rgv) < tol
    assert numpy.all(gv[~sel] == 0)
    assert gg.dtype == torch.complex128
    assert rel_err(gg.cpu().numpy(), rgg) < (1e-12 if precision == 64 else 2e-6)


def _imaging_problem(precision, seed=21, nants=24, ntimes=6, nchan=2, n=256, cell=2e-4, P=6):
    """A BlockVisibility converted to rows on the GPU, its w-stack imager and gain indices."""
    import torch
    import arlmi
    rng = numpy.random.default_rng(seed)
    xyz = _array(rng, nants, radius=1500.0)
    ha = numpy.linspace(-0.4, 0.4, ntimes)
    freq = numpy.linspace(1.2e8, 1.4e8, nchan)
    b = O.create_blockvisibility(xyz, ha, freq, numpy.full(nchan, 1e6), -0.6)
    shape = b['vis'].shape
    b['vis'] = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    bv = device_block(b, precision)
    vis = arlmi.convert_blockvisibility_to_visibility(bv)
    dt = torch.float32 if precision == 32 else torch.float64
    im = arlmi.create_image(npixel=n, cellsize=cell, frequency=[freq[0]], channel_bandwidth=[1e9], dtype=dt)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=8)
    imager = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode="wstack", vis_slices=P)
    gt = arlmi.create_gaintable_from_blockvisibility(bv)
    g = _gains(rng, gt.ntimes, nants, nchan, amp=0.1, phase=numpy.pi / 8)
    gain = torch.as_tensor(g.reshape(gt.gain.shape)).to(device=gpu_device(), dtype=vis.vis.dtype)
    i1, i2 = arlmi.gain_row_indices(vis, gt)
    image = numpy.zeros((1, 1, n, n))
    image[0, 0, rng.integers(64, 192, 25), rng.integers(64, 192, 25)] = rng.uniform(0.5, 2.0, 25)
    return b, vis, im, imager, gain, i1, i2, image, (gcf, cf)


def gpu_device():
    import torch
    return torch.device("cuda", 0)


def test_predict_adjoint_dot_product(gpu):
    """Re<predict(x), y> = <x, adjoint(y)> for the w-stack imager (float64)."""
    import torch
    b, vis, im, imager, gain, i1, i2, image, _ = _imaging_problem(64)
    rng = numpy.random.default_rng(3)
    n = image.shape[-1]
    x = torch.as_tensor(rng.normal(size=(1, 1, n, n))).to(gpu)
    y = torch.as_tensor(rng.normal(size=(vis.nvis, 1)) + 1j * rng.normal(size=(vis.nvis, 1))).to(gpu)
    ax = imager.predict(vis, x)
    aty = imager.adjoint(vis, y)
    lhs = float(torch.sum(ax.real * y.real + ax.imag * y.imag))
    rhs = float(torch.sum(x * aty))
    assert abs(lhs - rhs) < 1e-10 * max(abs(lhs), 1.0), (lhs, rhs)


def test_autograd_finite_differences(gpu):
    """Gradients of ||V - g_p conj(g_q) predict(I)||^2 through the hand-written adjoints against
    central finite differences (float64 path) at random image pixels and gain entries."""
    import torch
    import arlmi
    b, vis, im, imager, gain, i1, i2, image, _ = _imaging_problem(64)
    img = torch.as_tensor(image).to(gpu).requires_grad_(True)
    gain = gain.clone().requires_grad_(True)
    loss = arlmi.visibility_loss(img, gain, imager, vis, i1, i2)
    loss.backward()
    gi = img.grad.detach().clone()
    gg = gain.grad.detach().clone()
    rng = numpy.random.default_rng(7)
    n = image.shape[-1]

    def L(x, g):
        with torch.no_grad():
            return float(arlmi.visibility_loss(x, g, imager, vis, i1, i2))

    scale_i = float(gi.abs().max())
    for _ in range(12):
        y, x = rng.integers(n // 4, 3 * n // 4, 2)
        h = 1e-3
        xp = img.detach().clone()
        xm = img.detach().clone()
        xp[0, 0, y, x] += h
        xm[0, 0, y, x] -= h
        fd = (L(xp, gain.detach()) - L(xm, gain.detach())) / (2 * h)
        an = float(gi[0, 0, y, x])
        assert abs(fd - an) <= 1e-4 * scale_i, (y, x, fd, an)
    flat = gain.detach().reshape(-1)
    gflat = gg.reshape(-1)
    scale_g = float(gflat.abs().max())
    for k in rng.integers(0, flat.numel(), 10):
        for unit in (1.0, 1j):
            h = 1e-4
            gp = flat.clone()
            gm = flat.clone()
            gp[k] += h * unit
            gm[k] -= h * unit
            fd = (L(img.detach(), gp.reshape(gain.shape)) - L(img.detach(), gm.reshape(gain.shape))) / (2 * h)
            an = float(gflat[k].real if unit == 1.0 else gflat[k].imag)
            assert abs(fd - an) <= 1e-4 * scale_g, (k, unit, fd, an)


@pytest.mark.parametrize("precision", [32, 64])
def test_autograd_matches_oracle_adjoint(gpu, precision):
    """Image and gain gradients of the GPU path against the FP64 oracle's model and adjoint
    (predict_list_wstack, then per-slab invert with unit weights and no normalisation)."""
    import torch
    import arlmi
    b, vis, im, imager, gain, i1, i2, image, _ = _imaging_problem(precision)
    n, P = image.shape[-1], 6
    dt = torch.float32 if precision == 32 else torch.float64
    img = torch.as_tensor(image).to(device=gpu, dtype=dt).requires_grad_(True)
    gain = gain.clone().requires_grad_(True)
    loss = arlmi.visibility_loss(img, gain, imager, vis, i1, i2)
    loss.backward()
    # oracle
    rows = {k: v.cpu().numpy().astype(numpy.float64 if k != 'vis' else numpy.complex128)
            for k, v in vis.data.items() if k in ('uvw', 'vis', 'frequency')}
    rows['imaging_weight'] = numpy.ones((vis.nvis, 1))
    g = O.create_image_geom(n, 2e-4, frequency=vis.data['frequency'][0].item(), channel_bandwidth=1e9)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, 8)
    M = O.predict_list_wstack(rows, image, g, ogcf, ocf, cfg, P)
    gf = gain.detach().cpu().numpy().astype(numpy.complex128).reshape(-1)
    a1 = i1.cpu().numpy()
    a2 = i2.cpu().numpy()
    out = O.apply_gains_rows(M, a1, a2, gf)
    G = 2.0 * (out - rows['vis'])
    gM, rgg = O.apply_gains_rows_adjoint(M, G, a1, a2, gf)
    radj = numpy.zeros((1, 1, n, n))
    adj = dict(rows)
    adj['vis'] = gM
    for sel in O.vis_wslice_iter(rows['uvw'][:, 2], P):
        if numpy.sum(sel) == 0:
            continue
        img_k, _ = O.invert_wstack_single(O._rows(adj, sel), g, ogcf, ocf, cfg, normalize=False)
        radj += img_k
    # float32: the inner field and the L2 norm (the grid correction amplifies float32 rounding at
    # the image corners by up to ~6e4, as for every float32 image)
    l2, inner, full = image_errors(img.grad.detach().cpu().numpy(), radj)
    if precision == 32:
        assert l2 < 1e-4 and inner < 1e-4, (l2, inner, full)
    else:
        assert full < 1e-9, (l2, inner, full)
    assert rel_err(gain.grad.detach().cpu().numpy().reshape(-1), rgg) < (1e-4 if precision == 32 else 1e-9)
