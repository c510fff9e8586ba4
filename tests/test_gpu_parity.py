"""GPU parity of the HIP path against the FP64 NumPy oracle on identical seeded inputs.

Tolerances: integer index work is compared bit for bit.  Images: the float64 GPU path within
1e-10 (L2 and inner-field max-abs relative) and 1e-8 full-field max-abs; the float32 path
within 1e-5 inner-field max-abs relative (the precision target; inner field = the central half
of each axis) and 5e-3 full-field max-abs relative, because ARL's grid correction max/psi^2
multiplies the image corners by up to ~6e4 and with them float32 rounding.  Grids and visibilities: 1e-10 (float64) and 1e-5
(float32) relative to the peak.
"""

import ctypes

import numpy
import pytest

from oracle import arl_oracle as O
from helpers import rel_err, random_rows, to_device_vis, oracle_geom, device_image, image_errors, vis_errors

pytestmark = pytest.mark.gpu

TOL = {32: 1e-5, 64: 1e-10}
# full field: float32 transform rounding times ARL's grid correction, which reaches ~1e3-1e4 at
# the corners of these small images; measured 6.6e-4 to 1.1e-3 over the float32 cases here
FULL = {32: 2.5e-3, 64: 1e-8}


def assert_image(a, b, precision):
    l2, inner, full = image_errors(a, b)
    assert inner < TOL[precision] and full < FULL[precision], (l2, inner, full)
    if precision == 64:
        assert l2 < TOL[64], l2


def _pswf(npixel, cellsize, os_, S, precision):
    import arlmi
    im = device_image(npixel, cellsize, precision)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=os_, support=S)
    return im, gcf, cf


@pytest.mark.parametrize("os_,S", [(8, 8), (8, 6), (1, 6), (4, 4), (3, 6)])
def test_convolution_mapping_bitexact(gpu, os_, S):
    """Grid pixel, sub-sample, w plane and channel indices equal the oracle's exactly."""
    import arlmi
    rng = numpy.random.default_rng(11)
    n, cell = 256, 5e-4
    d = random_rows(rng, 20000, n, cell, S)
    # add rows exactly on pixel centres and half-way between (rounding / floor edge cases)
    du = 1.0 / (n * cell)
    k = numpy.arange(-100, 100)
    extra = numpy.zeros((len(k) * 2, 3))
    extra[:len(k), 0] = k * du
    extra[len(k):, 0] = (k + 0.5) * du
    extra[:, 1] = -extra[:, 0] * 0.5
    d["uvw"] = numpy.concatenate([d["uvw"], extra])
    d["vis"] = numpy.ones((len(d["uvw"]), 1), complex)
    d["imaging_weight"] = numpy.ones((len(d["uvw"]), 1))
    d["frequency"] = numpy.full(len(d["uvw"]), 1e8)
    import torch
    from arlmi.plan import make_geom, ptr, stream_handle
    im, gcf, cf = _pswf(n, cell, os_, S, 64)
    vis = to_device_vis(d, 64)
    geo = make_geom(im, cf, 0, 1)
    out = torch.empty((len(d["uvw"]), 8), dtype=torch.int32, device="cuda")
    L = arlmi._lib.lib()
    arlmi._lib.check(L.arlmi_convolution_mapping(ctypes.byref(geo), ptr(vis.uvw), ptr(vis.frequency),
                                                 len(d["uvw"]), ptr(out), stream_handle()))
    got = out.cpu().numpy()
    g = oracle_geom(n, cell)
    _, _, cfg = O.create_pswf_convolutionfunction(g, os_, S)
    pu, pu_off, pv, pv_off, pwg, pwc, pfreq = O.convolution_mapping(d["uvw"], d["frequency"], g, cfg)
    # a negative floor() indexes the kernel like numpy (wraps); floor() == os is ARL's IndexError
    bad = (pu_off == os_) | (pv_off == os_)
    assert numpy.array_equal((got[:, 7] & 8) != 0, bad)
    pu_off = numpy.where(pu_off < 0, pu_off + os_, pu_off)
    pv_off = numpy.where(pv_off < 0, pv_off + os_, pv_off)
    ok = ~bad
    for name, col, ref in [("pu", 0, pu), ("pu_off", 1, pu_off), ("pv", 2, pv), ("pv_off", 3, pv_off),
                           ("pwg", 4, pwg), ("pwc", 5, pwc), ("pfreq", 6, pfreq)]:
        assert numpy.array_equal(got[ok, col], ref[ok]), name
    assert numpy.all(got[ok, 7] == 0)
    if bad.any():
        with pytest.raises(IndexError):
            arlmi.convolution_mapping(vis, arlmi.create_griddata_from_image(im), cf)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("S", [4, 6, 8, 10, 12, 16])
def test_grid_parity(gpu, precision, S):
    """grid_visibility_to_griddata: GPU grid vs oracle grid (relative to peak), sumwt."""
    import arlmi
    rng = numpy.random.default_rng(3)
    n, cell = 256, 5e-4
    d = random_rows(rng, 30000, n, cell, S)
    im, gcf, cf = _pswf(n, cell, 8, S, precision)
    vis = to_device_vis(d, precision)
    gd = arlmi.create_griddata_from_image(im)
    gd, sumwt = arlmi.grid_visibility_to_griddata(vis, gd, cf)
    g = oracle_geom(n, cell)
    _, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    ref, rsum = O.grid_visibility_to_griddata(d, g, ocf, cfg)
    got = gd.data.cpu().numpy()
    assert rel_err(got, ref) < TOL[precision] / 10
    assert abs(sumwt[0, 0] - rsum[0, 0]) < 1e-9 * rsum[0, 0]


@pytest.mark.parametrize("precision", [64, 32])
def test_degrid_parity(gpu, precision):
    import torch
    import arlmi
    rng = numpy.random.default_rng(4)
    n, cell, S = 256, 5e-4, 8
    d = random_rows(rng, 30000, n, cell, S)
    im, gcf, cf = _pswf(n, cell, 8, S, precision)
    vis = to_device_vis(d, precision)
    grid = rng.normal(size=(1, 1, 1, n, n)) + 1j * rng.normal(size=(1, 1, 1, n, n))
    gd = arlmi.create_griddata_from_image(im)
    gd.data = torch.as_tensor(grid).to(device=gd.data.device, dtype=gd.data.dtype)
    out = arlmi.degrid_visibility_from_griddata(vis, gd, cf)
    g = oracle_geom(n, cell)
    _, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    ref = O.degrid_visibility_from_griddata(d, grid, g, ocf, cfg)
    assert rel_err(out.vis.cpu().numpy(), ref) < TOL[precision]


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("n", [64, 256, 512, 1024, 2048, 4096])
def test_fft_parity(gpu, precision, n):
    """ARL's shifted fft/ifft against numpy.fft on the same planes."""
    import torch
    import arlmi
    rng = numpy.random.default_rng(n)
    a = rng.normal(size=(2, 1, n, n)) + 1j * rng.normal(size=(2, 1, n, n))
    dt = torch.complex64 if precision == 32 else torch.complex128
    t = torch.as_tensor(a).to(device="cuda", dtype=dt)
    f = arlmi.fft(t).cpu().numpy()
    fi = arlmi.ifft(t).cpu().numpy()
    tol = 2e-6 if precision == 32 else 1e-13
    assert rel_err(f, O.fft(a)) < tol
    assert rel_err(fi, O.ifft(a)) < tol


@pytest.mark.parametrize("precision", [64, 32])
def test_invert_predict_2d_parity(gpu, precision):
    """invert_2d (normalised, imaginary, dopsf) and predict_2d against the oracle."""
    import torch
    import arlmi
    rng = numpy.random.default_rng(5)
    n, cell, S = 256, 5e-4, 8
    d = random_rows(rng, 20000, n, cell, S)
    im, gcf, cf = _pswf(n, cell, 8, S, precision)
    vis = to_device_vis(d, precision)
    g = oracle_geom(n, cell)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    re, sw, imag = arlmi.invert_2d(vis, im, gcfcf=(gcf, cf), imaginary=True)
    rre, rsw, rim = O.invert_2d(d, g, ogcf, ocf, cfg, imaginary=True)
    assert_image(re.numpy(), rre, precision)
    assert_image(imag.numpy(), rim, precision)
    psf, _ = arlmi.invert_2d(vis, im, dopsf=True, gcfcf=(gcf, cf))
    rpsf, _ = O.invert_2d(d, g, ogcf, ocf, cfg, dopsf=True)
    assert_image(psf.numpy(), rpsf, precision)
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, rng.integers(60, 196, 20), rng.integers(60, 196, 20)] = rng.uniform(0.5, 2.0, 20)
    mim = device_image(n, cell, precision, data=model)
    pv = arlmi.predict_2d(vis, mim, gcfcf=(gcf, cf))
    rpv = O.predict_2d(d, model, g, ogcf, ocf, cfg)
    assert rel_err(pv.vis.cpu().numpy(), rpv) < TOL[precision]


@pytest.mark.parametrize("precision", [64, 32])
def test_wstack_parity(gpu, precision):
    """w-stack list workflows (slab partition, slab-mean w shift, screens, per-slab weighting)."""
    import arlmi
    rng = numpy.random.default_rng(6)
    n, cell, S, P = 256, 1e-3, 8, 9
    d = random_rows(rng, 20000, n, cell, S, wmax=300.0)
    im, gcf, cf = _pswf(n, cell, 8, S, precision)
    vis = to_device_vis(d, precision)
    g = oracle_geom(n, cell)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    (img, sw), = arlmi.invert_list_serial_workflow([vis], [im], context="wstack", vis_slices=P, gcfcf=[(gcf, cf)])
    rimg, rsw = O.invert_list_wstack(d, g, ogcf, ocf, cfg, P)
    assert_image(img.numpy(), rimg, precision)
    assert abs(sw[0, 0] - rsw[0, 0]) < 1e-9 * rsw[0, 0]
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, rng.integers(40, 216, 30), rng.integers(40, 216, 30)] = rng.uniform(0.5, 2.0, 30)
    mim = device_image(n, cell, precision, data=model)
    pv, = arlmi.predict_list_serial_workflow([vis], [mim], context="wstack", vis_slices=P, gcfcf=[(gcf, cf)])
    rpv = O.predict_list_wstack(d, model, g, ogcf, ocf, cfg, P)
    assert rel_err(pv.vis.cpu().numpy(), rpv) < TOL[precision]


@pytest.mark.parametrize("precision", [64, 32])
def test_sparse_grids_match_dense(gpu, precision):
    """Sparse grids (tiles no row touches neither written nor read, unused tile columns skipped
    by the FFT passes) give the images and visibilities of dense grids, and match the oracle.
    (Not bit for bit: the order of the float64 LDS adds of different wavefronts varies from run
    to run, which the image corners' grid correction amplifies.)  The rows cover only part of the uv plane so that whole tiles and tile columns are
    empty, and w is tied to u so that the outer slabs are sparser still."""
    import arlmi
    rng = numpy.random.default_rng(12)
    n, cell, S, P = 256, 1e-3, 8, 7
    du = 1.0 / (n * cell)
    lim = (n // 2 - S // 2 - 1) * du
    nv = 12000
    u = rng.uniform(0.1 * lim, lim, nv)
    v = rng.uniform(-0.4 * lim, 0.9 * lim, nv)
    w = 0.5 * u + rng.uniform(-20.0, 20.0, nv)
    d = {"uvw": numpy.stack([u, v, w], 1), "vis": rng.normal(size=(nv, 1)) + 1j * rng.normal(size=(nv, 1)),
         "imaging_weight": rng.uniform(0.5, 2.0, (nv, 1)), "frequency": numpy.full(nv, 1e8)}
    im, gcf, cf = _pswf(n, cell, 8, S, precision)
    vis = to_device_vis(d, precision)
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, rng.integers(40, 216, 20), rng.integers(40, 216, 20)] = rng.uniform(0.5, 2.0, 20)
    mim = device_image(n, cell, precision, data=model)
    out = {}
    for sparse in (True, False):
        imager = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode="wstack", vis_slices=P, sparse=sparse)
        img = imager.invert(vis)[0].cpu().numpy()
        pv = imager.predict(vis, mim.data).cpu().numpy()
        out[sparse] = (img, pv)
    assert_image(out[True][0], out[False][0], precision)
    assert rel_err(out[True][1], out[False][1]) < TOL[precision] / 10
    g = oracle_geom(n, cell)
    ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    rimg, _ = O.invert_list_wstack(d, g, ogcf, ocf, cfg, P)
    assert_image(out[True][0].reshape(rimg.shape), rimg, precision)
    rpv = O.predict_list_wstack(d, model, g, ogcf, ocf, cfg, P)
    assert rel_err(out[True][1], rpv) < TOL[precision]


def test_box_kernel_equals_nearest_pixel(gpu):
    """Box-kernel gridding == nearest-pixel gridding (test_griddata_gridding.py:160-180 invariant)."""
    import arlmi
    rng = numpy.random.default_rng(7)
    n, cell = 128, 5e-4
    d = random_rows(rng, 5000, n, cell, 4)
    im = device_image(n, cell, 64)
    gcf, box = arlmi.create_box_convolutionfunction(im)
    vis = to_device_vis(d, 64)
    gd = arlmi.create_griddata_from_image(im)
    gd, sw = arlmi.grid_visibility_to_griddata(vis, gd, box)
    g = oracle_geom(n, cell)
    _, _, cfg = O.create_box_convolutionfunction(g)
    ref, rsw = O.grid_visibility_to_griddata_fast(d, g, cfg)
    assert rel_err(gd.data.cpu().numpy(), ref) < 1e-13


def test_dft_parity(gpu):
    import torch
    import arlmi
    rng = numpy.random.default_rng(8)
    d = random_rows(rng, 5000, 256, 5e-4, 8, wmax=200.0)
    comps = [(rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1), rng.uniform(0.1, 1.0)) for _ in range(7)]
    for precision in (64, 32):
        vis = to_device_vis(d, precision)
        vis.data["vis"].zero_()
        arlmi.predict_skycomponent_visibility(vis, [arlmi.Skycomponent(l, m, [[f]]) for l, m, f in comps])
        dz = dict(d)
        dz["vis"] = numpy.zeros_like(d["vis"])
        ref = O.predict_skycomponent_visibility(dz, comps)
        assert rel_err(vis.vis.cpu().numpy(), ref) < (1e-12 if precision == 64 else 1e-6)


def test_dft_multichannel_frequency_map(gpu):
    """Fluxes of several channels follow ARL's map (the index of each row's frequency among the
    distinct frequencies, imaging_params.py:23-34); a map past the flux's channels raises."""
    import arlmi
    rng = numpy.random.default_rng(18)
    d = random_rows(rng, 4000, 256, 5e-4, 8, wmax=100.0)
    d["frequency"] = rng.choice([1.2e8, 1.0e8], 4000)
    comps = [(rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1), rng.uniform(0.1, 1.0, (2, 1))) for _ in range(5)]
    dz = dict(d)
    dz["vis"] = numpy.zeros_like(d["vis"])
    ref = O.predict_skycomponent_visibility(dz, comps)
    # the oracle's map: rows at 1.0e8 take flux channel 0, rows at 1.2e8 channel 1
    lo = d["frequency"] == 1.0e8
    ph = sum(f[0, 0] * O.simulate_point(d["uvw"][lo], l, m) for l, m, f in comps)
    assert numpy.max(numpy.abs(ref[lo, 0] - ph)) < 1e-12
    for precision in (64, 32):
        vis = to_device_vis(d, precision)
        vis.data["vis"].zero_()
        arlmi.predict_skycomponent_visibility(vis, [arlmi.Skycomponent(l, m, f) for l, m, f in comps])
        assert rel_err(vis.vis.cpu().numpy(), ref) < (1e-12 if precision == 64 else 1e-6)
    d3 = dict(d)
    d3["frequency"] = rng.choice([1.2e8, 1.0e8, 1.4e8], 4000)
    vis = to_device_vis(d3, 64)
    with pytest.raises(IndexError):
        arlmi.predict_skycomponent_visibility(vis, [arlmi.Skycomponent(l, m, f) for l, m, f in comps])


def test_awterm_kernel_parity(gpu):
    """GPU-built w-projection kernel against the oracle's (kernels.py:100-200)."""
    import arlmi
    n, cell = 256, 5e-4
    im = device_image(n, cell, 64)
    gcf, cf = arlmi.create_awterm_convolutionfunction(im, nw=9, wstep=20.0, oversampling=4, support=16)
    g = oracle_geom(n, cell)
    ogcf, ocf, _ = O.create_awterm_convolutionfunction(g, nw=9, wstep=20.0, oversampling=4, support=16)
    assert rel_err(cf.data.cpu().numpy(), ocf) < 1e-12
    assert rel_err(gcf.data.cpu().numpy(), ogcf) < 1e-14


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("S", [6, 8, 10, 14, 16, 22, 32])
def test_wprojection_grid_degrid_parity(gpu, precision, S):
    """Gridding/degridding with a complex multi-plane (AW) kernel table; the supports cover
    every taps-per-lane variant of the table gridder (up to 64 and up to 256 taps in registers, the
    grouped loop beyond), with full and partly filled last tap groups (S = 14: 196 taps, the last
    of four groups holding 4; a compiler lowering of the entry loop once lost the adds of the lanes
    past the last tap there)."""
    import arlmi
    rng = numpy.random.default_rng(9)
    n, cell = 256, 5e-4
    d = random_rows(rng, 8000, n, cell, S, wmax=80.0)
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_awterm_convolutionfunction(im, nw=9, wstep=20.0, oversampling=4, support=S)
    g = oracle_geom(n, cell)
    ogcf, ocf, cfg = O.create_awterm_convolutionfunction(g, nw=9, wstep=20.0, oversampling=4, support=S)
    vis = to_device_vis(d, precision)
    re, sw = arlmi.invert_2d(vis, im, gcfcf=(gcf, cf))
    rre, rsw = O.invert_2d(d, g, ogcf, ocf, cfg)
    assert_image(re.numpy(), rre, precision)
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, 128, 100] = 1.0
    model[0, 0, 90, 150] = 0.5
    mim = device_image(n, cell, precision, data=model)
    pv = arlmi.predict_2d(vis, mim, gcfcf=(gcf, cf))
    rpv = O.predict_2d(d, model, g, ogcf, ocf, cfg)
    assert rel_err(pv.vis.cpu().numpy(), rpv) < TOL[precision]


def test_off_grid_raises(gpu):
    """A row off the grid raises like ARL's assert (gridding.py:43-46); dropping it is opt-in."""
    import arlmi
    rng = numpy.random.default_rng(10)
    n, cell = 128, 5e-4
    d = random_rows(rng, 100, n, cell, 8)
    d["uvw"][5, 0] = 1.0 / (2 * cell) * 1.5
    im, gcf, cf = _pswf(n, cell, 8, 8, 64)
    vis = to_device_vis(d, 64)
    with pytest.raises(AssertionError):
        arlmi.invert_2d(vis, im, gcfcf=(gcf, cf))
    imager = arlmi.Imager(vis, im, gcfcf=(gcf, cf), drop_off_grid=True)
    assert imager.plan.stats()["ndropped"] == 1


@pytest.mark.parametrize("P", [1, 2, 3, 9, 64, 128])
def test_device_slab_boxes_match_numpy(gpu, P):
    """With no centres given, the plan's slab centres and width equal numpy.linspace(-max|w|,
    max|w|, P) and ARL's width (iterators.py:99-106) bit for bit, including all-zero w."""
    import torch
    import arlmi
    from arlmi.plan import Plan, wslice_boxes
    rng = numpy.random.default_rng(P)
    n, cell = 256, 1e-3
    im, gcf, cf = _pswf(n, cell, 8, 8, 64)
    for wmax in (987.654321, 3.0e-3, 0.0):
        d = random_rows(rng, 5000, n, cell, 8, wmax=wmax)
        vis = to_device_vis(d, 64)
        plan = Plan(vis, im, cf, arlmi._lib.MODE_WSTACK, P)
        boxes, width = plan.slab_boxes()
        rb, rw = wslice_boxes(float(numpy.max(numpy.abs(d["uvw"][:, 2]))), P)
        assert numpy.array_equal(boxes, rb), (P, wmax)
        assert width == rw, (P, wmax)
        plan.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_wstack_with_wprojection_parity(gpu, precision):
    """The hybrid of configuration 2: w slabs (screens at the slab mean) with an AW kernel for the
    residual w within each slab (the kernel's w plane from w - mean, gridding.py:70-76 on the
    shifted rows of wstack_single.py:83-90), against the oracle's list workflow."""
    import arlmi
    rng = numpy.random.default_rng(21)
    n, cell, S, P = 256, 5e-4, 16, 5
    d = random_rows(rng, 12000, n, cell, S, wmax=150.0)
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_awterm_convolutionfunction(im, nw=9, wstep=20.0, oversampling=4, support=S)
    g = oracle_geom(n, cell)
    ogcf, ocf, cfg = O.create_awterm_convolutionfunction(g, nw=9, wstep=20.0, oversampling=4, support=S)
    vis = to_device_vis(d, precision)
    (img, sw), = arlmi.invert_list_serial_workflow([vis], [im], context="wstack", vis_slices=P, gcfcf=[(gcf, cf)])
    rimg, rsw = O.invert_list_wstack(d, g, ogcf, ocf, cfg, P)
    assert_image(img.numpy(), rimg, precision)
    assert abs(sw[0, 0] - rsw[0, 0]) < 1e-9 * rsw[0, 0]
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, 120, 100] = 1.0
    model[0, 0, 150, 160] = 0.5
    mim = device_image(n, cell, precision, data=model)
    pv, = arlmi.predict_list_serial_workflow([vis], [mim], context="wstack", vis_slices=P, gcfcf=[(gcf, cf)])
    rpv = O.predict_list_wstack(d, model, g, ogcf, ocf, cfg, P)
    assert rel_err(pv.vis.cpu().numpy(), rpv) < TOL[precision]


def test_wstack_w_range_assert(gpu):
    """ARL's w-range assert (gridding.py:64-65) on the slab-mean-shifted w of a w-stack plan with a
    one-plane kernel: a grid WW step of 1 wavelength rejects rows more than half a wavelength from
    their slab mean.  Raised when the plan computes the slab means itself and when the caller
    fixes them (arlmi_plan_fix_plane_w)."""
    import ctypes
    import numpy
    import arlmi
    from arlmi.plan import Plan
    rng = numpy.random.default_rng(19)
    n, cell, S = 256, 1e-3, 8
    d = random_rows(rng, 3000, n, cell, S, wmax=100.0)
    vis = to_device_vis(d, 64)
    im, gcf, cf = _pswf(n, cell, 8, S, 64)
    with pytest.raises(arlmi._lib.ArlmiError) as e:
        Plan(vis, im, cf, arlmi._lib.MODE_WSTACK, 4, wstep_grid=1.0)
    assert e.value.code == -3
    ref = Plan(vis, im, cf, arlmi._lib.MODE_WSTACK, 4)
    means, boxes = ref.plane_w(), ref.slab_boxes()
    flat = dict(d)
    flat["uvw"] = d["uvw"].copy()
    flat["uvw"][:, 2] = boxes[0][0]  # every row on the first slab centre: no assert
    plan = Plan(to_device_vis(flat, 64), im, cf, arlmi._lib.MODE_WSTACK, 4, wstep_grid=1.0, slab_boxes=boxes)
    arlmi._lib.check(arlmi._lib.lib().arlmi_plan_fix_plane_w(
        plan.handle, means.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 4))
    with pytest.raises(arlmi._lib.ArlmiError) as e:
        plan.update(vis, boxes)
    assert e.value.code == -3
