"""Pins the FP64 oracle: ARL's own known answers, closed forms, and invariants chosen so that a
plausible mistake (dropped conjugate, wrong sign, transposed axis, off-by-one pixel) fails."""

import numpy
import pytest

from oracle import arl_oracle as O
from helpers import rel_err, random_rows

GEOM512 = O.ImageGeom(512, 0.0005)  # test_griddata_kernels.py setUp: 512^2, cell 5e-4


# --- ARL known answers (tests/processing_components/test_griddata_kernels.py) ---------------

def test_pswf_os8_known_answer():
    _, cf, _ = O.create_pswf_convolutionfunction(GEOM512, oversampling=8, support=6)
    pk = numpy.unravel_index(numpy.argmax(numpy.abs(cf)), cf.shape)
    assert pk == (0, 0, 0, 4, 4, 3, 3)                        # :78
    assert abs(cf[pk] - 0.1871722034219655) < 1e-7            # :77


def test_pswf_os1_known_answer():
    _, cf, _ = O.create_pswf_convolutionfunction(GEOM512, oversampling=1, support=6)
    pk = numpy.unravel_index(numpy.argmax(numpy.abs(cf)), cf.shape)
    assert pk == (0, 0, 0, 0, 0, 3, 3)                        # :99
    assert abs(cf[pk] - 0.18712109669890536) < 1e-7          # :98


def test_box_known_answer():
    _, cf, _ = O.create_box_convolutionfunction(GEOM512)
    pk = numpy.unravel_index(numpy.argmax(numpy.abs(cf)), cf.shape)
    assert pk == (0, 0, 0, 0, 0, 2, 2)                        # :57
    assert abs(cf[pk] - 1.0) < 1e-15                         # :56


@pytest.fixture(scope="module")
def wkernel201():
    return O.create_awterm_convolutionfunction(GEOM512, nw=201, wstep=8.0, oversampling=8, support=60,
                                               use_aaf=True)


def test_wterm_known_answer(wkernel201):
    _, cf, _ = wkernel201
    pk = numpy.unravel_index(numpy.argmax(numpy.abs(cf)), cf.shape)
    assert pk == (0, 0, 100, 4, 4, 30, 30)                    # :112
    assert abs(cf[pk] - (0.18704481681878257 - 0j)) < 1e-7   # :113-114
    clipped = O.apply_bounding_box_convolutionfunction(cf, 1e-3)
    pk = numpy.unravel_index(numpy.argmax(numpy.abs(clipped)), clipped.shape)
    assert pk == (0, 0, 100, 4, 4, 27, 27)                    # :130
    assert abs(clipped[pk] - (0.18704481681878257 - 0j)) < 1e-7


def test_wterm_symmetry_and_os_agreement():
    """test_griddata_kernels.py:235-296: u/v mirror symmetry, cf(w) = conj(cf(-w)), and the os=8
    kernel sampled at sub-sample (4, 4) agrees with the os=1 kernel of the same w to 5e-6."""
    _, cf, _ = O.create_awterm_convolutionfunction(GEOM512, nw=110, wstep=8, oversampling=8, support=60,
                                                   use_aaf=True, planes=[0, 25, 45, 55, 85])
    assert abs(cf[0, 0, 55, 4, 4, 30, 30] - 0.18704481681878257) < 1e-12
    assert abs(cf[0, 0, 55, 4, 4, 25, 25] - cf[0, 0, 55, 4, 4, 35, 35]) < 1e-15
    assert abs(cf[0, 0, 55, 4, 4, 25, 25] - cf[0, 0, 55, 4, 4, 25, 35]) < 1e-15
    assert abs(cf[0, 0, 45, 4, 4, 25, 25] - cf[0, 0, 45, 4, 4, 35, 25]) < 1e-15
    assert abs(cf[0, 0, 25, 4, 4, 25, 25] - numpy.conjugate(cf[0, 0, 85, 4, 4, 25, 25])) < 1e-15
    assert abs(cf[0, 0, 25, 4, 4, 25, 25] - numpy.conjugate(cf[0, 0, 85, 4, 4, 35, 35])) < 1e-15
    _, cfn, _ = O.create_awterm_convolutionfunction(GEOM512, nw=111, wstep=8, oversampling=1, support=60,
                                                    use_aaf=True, planes=[0])
    # both plane-0 kernels are at w = 8 * (1 - 56) = -440
    assert abs(numpy.real(cf[0, 0, 0, 4, 4, 30, 30]) - numpy.real(cfn[0, 0, 0, 0, 0, 30, 30])) < 5e-6


# --- processing_library known answers ------------------------------------------------------

def test_coordinates_known_answers():
    """test_convolutional_gridding.py:21-34."""
    for n in [4, 5, 6, 7, 8, 9, 1000, 1001, 1002, 1003]:
        low, high = O.coordinateBounds(n)
        c = O.coordinates(n)
        cx, cy = O.coordinates2(n)
        assert abs(c.min() - low) < 1e-12 and abs(c.max() - high) < 1e-12
        assert abs(cx.min() - low) < 1e-12 and abs(cy.max() - high) < 1e-12
        assert c[n // 2] == 0 and (cx[n // 2, :] == 0).all() and (cy[:, n // 2] == 0).all()


def test_w_beam_closed_form():
    """w_beam(5, 0.1, 0)[0, 0] == 1 (test_convolutional_gridding.py:41) and the full formula."""
    assert abs(O.w_beam(5, 0.1, 0)[0, 0] - 1.0) < 1e-15
    n, fov, w = 64, 0.3, 123.4
    b = O.w_beam(n, fov, w)
    y, x = numpy.mgrid[0:n, 0:n]
    r2 = fov ** 2 * (((y - n // 2) / n) ** 2 + ((x - n // 2) / n) ** 2)
    ref = numpy.exp(-2j * numpy.pi * w * (1 - numpy.sqrt(1 - r2)))
    assert numpy.max(numpy.abs(b - ref)) < 1e-12
    assert b[n // 2, n // 2] == 1.0
    big = O.w_beam(16, 3.0, 5.0)  # corners outside the unit circle are zero
    assert big[0, 0] == 0


def test_simulate_point_known_answers():
    """test_coordinate_support.py:104-123."""
    uvw = numpy.concatenate(numpy.concatenate(numpy.transpose(numpy.mgrid[-3:4, -3:4, 0:1])))
    bls = numpy.array([uvw[j] - uvw[i] for i in range(len(uvw)) for j in range(i + 1, len(uvw))], dtype=float)
    assert numpy.allclose(O.simulate_point(bls, 0, 0), 1.0)
    bl_even = 1 - 2 * (numpy.sum(bls, axis=1) % 2)
    for l, m in [(0.5, 0.5), (-0.5, 0.5), (0.5, -0.5), (-0.5, -0.5)]:
        assert numpy.allclose(O.simulate_point(bls, l, m), bl_even)


def test_xyz_to_uvw_closed_form():
    """Rotation: |uvw| = |xyz|, ha=0 dec=90deg maps (x, y, z) -> (x, y, z)."""
    rng = numpy.random.default_rng(0)
    xyz = rng.normal(size=(10, 3))
    uvw = O.xyz_to_uvw(xyz, 0.7, -0.4)
    assert numpy.allclose(numpy.linalg.norm(uvw, axis=1), numpy.linalg.norm(xyz, axis=1))
    assert numpy.allclose(O.xyz_to_uvw(xyz, 0.0, numpy.pi / 2), xyz)
    az, el = O.hadec_to_azel(0.0, 0.0, 0.0)
    assert abs(el - numpy.pi / 2) < 1e-12


# --- FFT -------------------------------------------------------------------------------------

def test_fft_shift_convention_against_direct_dft():
    """ARL's fft: fftshift(fft2(ifftshift)), the DFT centred on pixel N//2."""
    rng = numpy.random.default_rng(1)
    n = 6
    a = rng.normal(size=(1, 1, n, n)) + 1j * rng.normal(size=(1, 1, n, n))
    k = numpy.arange(n) - n // 2
    e = numpy.exp(-2j * numpy.pi * numpy.outer(k, k) / n)
    ref = e @ a[0, 0] @ e.T
    assert numpy.max(numpy.abs(O.fft(a)[0, 0] - ref)) < 1e-12
    assert numpy.max(numpy.abs(O.ifft(O.fft(a)) - a)) < 1e-12


# --- gridding: brute force, loops, adjoint, box == nearest -----------------------------------

def _small_case(rng, n=32, nvis=40, S=6, os_=4, cell=0.01):
    g = O.ImageGeom(n, cell)
    gcf, cf, cfg = O.create_pswf_convolutionfunction(g, os_, S)
    d = random_rows(rng, nvis, n, cell, S)
    return g, gcf, cf, cfg, d


def test_grid_brute_force():
    """Each visibility adds conj(CF[yoff, xoff]) * V * w at pixels pu-S/2.., pv-S/2.. (rows = v)."""
    rng = numpy.random.default_rng(2)
    g, gcf, cf, cfg, d = _small_case(rng)
    grid, sumwt = O.grid_visibility_to_griddata(d, g, cf, cfg)
    ref = numpy.zeros((g.npixel, g.npixel), complex)
    du = 1.0 / (g.npixel * O.D2R * g.cdelt0)
    dv = 1.0 / (g.npixel * O.D2R * g.cdelt1)
    S, os_ = 6, 4
    for (u, v, _), vis, w in zip(d["uvw"], d["vis"][:, 0], d["imaging_weight"][:, 0]):
        x = g.npixel // 2 + u / du
        y = g.npixel // 2 + v / dv
        pu, pv = int(numpy.round(x)), int(numpy.round(y))
        xo = int(numpy.floor(os_ // 2 + os_ * (x - pu)))
        yo = int(numpy.floor(os_ // 2 + os_ * (y - pv)))
        ref[pv - S // 2:pv + S // 2, pu - S // 2:pu + S // 2] += numpy.conjugate(cf[0, 0, 0, yo, xo]) * vis * w
    assert rel_err(grid[0, 0, 0], ref) < 1e-12
    assert abs(sumwt[0, 0] - d["imaging_weight"].sum()) < 1e-12


def test_vectorised_grid_equals_arl_loop_bitwise():
    rng = numpy.random.default_rng(3)
    g, gcf, cf, cfg, d = _small_case(rng, n=64, nvis=500)
    a, sa = O.grid_visibility_to_griddata(d, g, cf, cfg, chunk=97)
    b, sb = O.grid_visibility_to_griddata_loop(d, g, cf, cfg)
    assert numpy.array_equal(a, b) and numpy.array_equal(sa, sb)


def test_vectorised_degrid_equals_arl_loop():
    rng = numpy.random.default_rng(4)
    g, gcf, cf, cfg, d = _small_case(rng, n=64, nvis=500)
    grid = rng.normal(size=(1, 1, 1, 64, 64)) + 1j * rng.normal(size=(1, 1, 1, 64, 64))
    a = O.degrid_visibility_from_griddata(d, grid, g, cf, cfg, chunk=77)
    b = O.degrid_visibility_from_griddata_loop(d, grid, g, cf, cfg)
    assert rel_err(a, b) < 1e-14


def test_grid_degrid_adjoint():
    """<degrid(G), V> = <G, grid(V)>: gridding uses conj(CF), degridding CF (weights 1)."""
    rng = numpy.random.default_rng(5)
    g, gcf, cf, cfg, d = _small_case(rng, n=64, nvis=300)
    d["imaging_weight"][:] = 1.0
    G = rng.normal(size=(1, 1, 1, 64, 64)) + 1j * rng.normal(size=(1, 1, 1, 64, 64))
    lhs = numpy.vdot(d["vis"], O.degrid_visibility_from_griddata(d, G, g, cf, cfg))
    grid, _ = O.grid_visibility_to_griddata(d, g, cf, cfg)
    rhs = numpy.vdot(grid, G)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_invert_predict_adjoint():
    """Re <predict_2d(I), V> = <I, invert_2d(V, normalize=False)> (real image, unit weights)."""
    rng = numpy.random.default_rng(6)
    g, gcf, cf, cfg, d = _small_case(rng, n=64, nvis=300)
    d["imaging_weight"][:] = 1.0
    I = rng.normal(size=(1, 1, 64, 64))
    pv = O.predict_2d(d, I, g, gcf, cf, cfg)
    img, _ = O.invert_2d(d, g, gcf, cf, cfg, normalize=False)
    lhs = numpy.real(numpy.vdot(d["vis"], pv))
    rhs = numpy.sum(I * img)
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_wstack_adjoint():
    rng = numpy.random.default_rng(7)
    n, cell, S = 64, 0.01, 6
    g = O.ImageGeom(n, cell)
    gcf, cf, cfg = O.create_pswf_convolutionfunction(g, 4, S)
    d = random_rows(rng, 200, n, cell, S, wmax=40.0)
    d["imaging_weight"][:] = 1.0
    I = rng.normal(size=(1, 1, n, n))
    pv = O.predict_list_wstack(d, I, g, gcf, cf, cfg, 5)
    img, sw = O.invert_list_wstack(d, g, gcf, cf, cfg, 5)
    lhs = numpy.real(numpy.vdot(d["vis"], pv))
    rhs = numpy.sum(I * img) * sw[0, 0]
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_box_gridding_equals_nearest_pixel():
    """test_griddata_gridding.py:160-180: box kernel == _fast (nearest pixel) gridding."""
    rng = numpy.random.default_rng(8)
    g = O.ImageGeom(64, 0.01)
    gcf, cf, cfg = O.create_box_convolutionfunction(g)
    d = random_rows(rng, 300, 64, 0.01, 4)
    a, sa = O.grid_visibility_to_griddata(d, g, cf, cfg)
    b, sb = O.grid_visibility_to_griddata_fast(d, g, cfg)
    assert numpy.array_equal(a, b)


def test_invert_point_source_matches_dft():
    """A point source's dirty image peaks at its pixel, and predict_2d matches the direct DFT.

    Support 6: ARL's grid correction psi(|2 x/N|) is the Fourier pair of its kernel only for the
    m=6 spheroidal over 6 cells, so that is where predict reproduces the DFT (max error 3e-2 of
    the unit flux here at os=64, set by the kernel; a flipped sign or transposed axis gives O(1)).
    With support 8 ARL's correction is mismatched and the predicted amplitude falls off with
    distance from the centre (0.84 at 15 pixels on this grid); the oracle keeps that behaviour.
    """
    rng = numpy.random.default_rng(9)
    n, cell, S = 128, 0.002, 6
    g = O.ImageGeom(n, cell)
    gcf, cf, cfg = O.create_pswf_convolutionfunction(g, 64, S)
    d = random_rows(rng, 2000, n, cell, S, margin=n // 4)
    x, y = 70, 50
    l, m = -(x - n // 2) * cell, (y - n // 2) * cell
    d["vis"] = O.predict_skycomponent_visibility({"uvw": d["uvw"], "vis": numpy.zeros((2000, 1), complex)},
                                                 [(l, m, 1.0)])
    img, _ = O.invert_2d(d, g, gcf, cf, cfg)
    assert numpy.unravel_index(numpy.argmax(img), img.shape) == (0, 0, y, x)
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, y, x] = 1.0
    pv = O.predict_2d(d, model, g, gcf, cf, cfg)
    assert rel_err(pv, d["vis"]) < 3e-2
    assert abs(numpy.mean(numpy.abs(pv / d["vis"])) - 1.0) < 2e-3


def test_wslice_iter_partition():
    """Slab rule: |w - box| < step/2 over linspace(-wmax, wmax, P); a tie with the midpoint drops."""
    w = numpy.array([-10.0, -5.0, -2.5, 0.0, 2.5, 7.4, 10.0])
    masks = list(O.vis_wslice_iter(w, 5))  # boxes -10, -5, 0, 5, 10; step 5
    assert [int(m.sum()) for m in masks] == [1, 1, 1, 1, 1]
    assert not any(m[2] for m in masks)  # -2.5 is exactly half way: in no slab
    assert not any(m[4] for m in masks)  # so is 2.5
    assert masks[3][5] and masks[4][6]


def test_gain_convention():
    """V_row = g[antenna2] conj(g[antenna1]) V (calibration/operations.py:127-137 via coalesce.py:413)."""
    g = numpy.array([1.0 + 1j, 2.0, 0.5j])
    v = numpy.ones((1, 1), complex)
    out = O.apply_gains_rows(v, numpy.array([0]), numpy.array([2]), g)
    assert abs(out[0, 0] - 0.5j * numpy.conjugate(1.0 + 1j)) < 1e-15


def test_grid_w_planes_partition_rows():
    """A grid with nw WW planes (create_griddata_from_image(nw, wstep), gridding.py:58-65) puts each
    row on plane round(w/wstep) + nw//2 and nowhere else: every plane equals the single-plane grid
    of exactly those rows, and the planes sum to the single-plane grid of all rows."""
    rng = numpy.random.default_rng(41)
    n, cell, nw, wstep = 64, 0.01, 5, 20.0
    g = O.ImageGeom(n, cell)
    gcf, cf, cfg = O.create_awterm_convolutionfunction(g, nw=nw, wstep=wstep, oversampling=4, support=8)
    d = random_rows(rng, 600, n, cell, 8, wmax=45.0)
    grid, sw = O.grid_visibility_to_griddata(d, g, cf, cfg, nw_grid=nw, wstep_grid=wstep)
    one, sw1 = O.grid_visibility_to_griddata(d, g, cf, cfg)
    assert numpy.max(numpy.abs(grid.sum(axis=2) - one[:, :, 0])) < 1e-12 * numpy.max(numpy.abs(one))
    plane = numpy.round(d["uvw"][:, 2] / wstep).astype(int) + nw // 2
    assert set(numpy.unique(plane)) == set(range(nw))
    for z in range(nw):
        sel = plane == z
        part, _ = O.grid_visibility_to_griddata(O._rows(d, sel), g, cf, cfg)
        assert numpy.max(numpy.abs(grid[:, :, z] - part[:, :, 0])) < 1e-12 * numpy.max(numpy.abs(one))
    assert abs(sw[0, 0] - numpy.sum(d["imaging_weight"])) < 1e-9


def test_spectral_cube_channels_partition_rows():
    """Image channel c of a cube (pfreq = round((f - f0)/df), gridding.py:80-85) grids exactly the
    rows whose frequency maps to c, and its sum of weights is theirs."""
    rng = numpy.random.default_rng(42)
    n, cell = 64, 0.01
    g3 = O.ImageGeom(n, cell, nchan=3, frequency=1.0e8, channel_bandwidth=1.0e7)
    g1 = O.ImageGeom(n, cell)
    gcf, cf, cfg = O.create_pswf_convolutionfunction(g3, 4, 6)
    d = random_rows(rng, 500, n, cell, 6)
    d["frequency"] = rng.choice([1.0e8, 1.1e8, 1.2e8], 500)
    grid, sw = O.grid_visibility_to_griddata(d, g3, cf, cfg)
    _, cf1, cfg1 = O.create_pswf_convolutionfunction(g1, 4, 6)
    for c, f in enumerate((1.0e8, 1.1e8, 1.2e8)):
        sel = d["frequency"] == f
        part, psw = O.grid_visibility_to_griddata(dict(O._rows(d, sel), frequency=numpy.full(sel.sum(), 1e8)),
                                                  g1, cf1, cfg1)
        assert numpy.array_equal(grid[c], part[0])
        assert abs(sw[c, 0] - psw[0, 0]) < 1e-12
    with pytest.raises(AssertionError):
        bad = dict(d, frequency=numpy.full(500, 1.3e8))
        O.grid_visibility_to_griddata(bad, g3, cf, cfg)
