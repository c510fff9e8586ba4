"""Phase rotation on the GPU (arlmi_phaserotate_rows / _blocks) against the oracle.

The uvw rotation (tangent=False) uses the same float64 trigonometric factors and ARL's operation
order without contraction, so it is compared bit for bit.  The visibilities are compared with a
tolerance: the library computes (l, m) by astropy's route (frame rotations of a Cartesian vector)
and the oracle by spherical trigonometry, which differ by ~1e-16, i.e. ~1e-12 rad of phase at
these baselines; float32 storage adds its own rounding (6e-8).
"""

import numpy
import pytest

from oracle import arl_oracle as O

pytestmark = pytest.mark.gpu

D = numpy.pi / 180.0
PC = (180.0 * D, -35.0 * D)
NEW = (181.3 * D, -33.8 * D)
CASES = [(True, False), (False, False), (False, True), (True, True)]


def _tol(precision):
    return 1e-11 if precision == 64 else 2e-7


def _rel(a, b):
    return float(numpy.max(numpy.abs(a - b)) / numpy.max(numpy.abs(b)))


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("tangent,inverse", CASES)
def test_rows_against_oracle(gpu, precision, tangent, inverse):
    import arlmi
    rng = numpy.random.default_rng(10)
    n = 20000
    uvw = rng.uniform(-3000, 3000, (n, 3))
    uvw[:, 2] *= 0.1
    vis = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    if precision == 32:
        vis = vis.astype(numpy.complex64).astype(numpy.complex128)
    v = arlmi.Visibility.from_arrays(uvw, vis=vis, precision=precision)
    v.phasecentre = arlmi.SkyCoord(*PC)
    out = arlmi.phaserotate_visibility(v, arlmi.SkyCoord(*NEW), tangent=tangent, inverse=inverse)
    ruvw, rvis = O.phaserotate_rows({'uvw': uvw, 'vis': vis}, PC, NEW, tangent=tangent, inverse=inverse)
    got_uvw = out.data["uvw"].cpu().numpy()
    assert numpy.array_equal(got_uvw, ruvw)
    assert _rel(out.vis.cpu().numpy().astype(numpy.complex128), rvis) < _tol(precision)
    # the input is not modified; the phase centre moves only when the uvw are rotated
    assert numpy.array_equal(v.data["uvw"].cpu().numpy(), uvw)
    assert numpy.array_equal(v.vis.cpu().numpy().astype(numpy.complex128), vis)
    assert (out.phasecentre.ra, out.phasecentre.dec) == ((NEW if not tangent else PC))


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("tangent,inverse", CASES)
def test_blocks_against_oracle(gpu, precision, tangent, inverse):
    import arlmi
    rng = numpy.random.default_rng(11)
    xyz = O.xyz_at_latitude(numpy.stack([rng.uniform(-4e3, 4e3, 12), rng.uniform(-4e3, 4e3, 12),
                                         rng.uniform(-5, 5, 12)], 1), -26.7 * D)
    freq = numpy.array([1.0e8, 1.25e8, 1.5e8, 1.8e8])
    b = O.create_blockvisibility(xyz, numpy.linspace(-0.4, 0.4, 6), freq, numpy.full(4, 1e6), PC[1])
    b['vis'] = rng.normal(size=b['vis'].shape) + 1j * rng.normal(size=b['vis'].shape)
    if precision == 32:
        b['vis'] = b['vis'].astype(numpy.complex64).astype(numpy.complex128)
    bv = arlmi.BlockVisibility.from_arrays(b['uvw'], vis=b['vis'], frequency=freq, precision=precision)
    bv.phasecentre = arlmi.SkyCoord(*PC)
    out = arlmi.phaserotate_visibility(bv, arlmi.SkyCoord(*NEW), tangent=tangent, inverse=inverse)
    ruvw, rvis = O.phaserotate_blocks(b['uvw'], b['vis'], freq, PC, NEW, tangent=tangent, inverse=inverse)
    assert numpy.array_equal(out.data["uvw"].cpu().numpy(), ruvw)
    assert _rel(out.vis.cpu().numpy().astype(numpy.complex128), rvis) < _tol(precision)


def test_same_centre_is_a_no_op(gpu):
    import arlmi
    v = arlmi.Visibility.from_arrays(numpy.ones((4, 3)), vis=numpy.ones((4, 1)), precision=64)
    v.phasecentre = arlmi.SkyCoord(*PC)
    assert arlmi.phaserotate_visibility(v, arlmi.SkyCoord(*PC), tangent=False) is v


def test_rows_agree_with_blocks_through_convert(gpu):
    """Rotating a BlockVisibility and converting it to rows equals converting first and rotating
    the rows (uvw scaled by f / c before or after the rotation: equal to rounding)."""
    import arlmi
    rng = numpy.random.default_rng(12)
    xyz = O.xyz_at_latitude(numpy.stack([rng.uniform(-2e3, 2e3, 9), rng.uniform(-2e3, 2e3, 9),
                                         rng.uniform(-5, 5, 9)], 1), -26.7 * D)
    freq = numpy.array([1.0e8, 1.5e8])
    b = O.create_blockvisibility(xyz, numpy.linspace(-0.3, 0.3, 5), freq, numpy.full(2, 1e6), PC[1])
    b['vis'] = rng.normal(size=b['vis'].shape) + 1j * rng.normal(size=b['vis'].shape)
    bv = arlmi.BlockVisibility.from_arrays(b['uvw'], vis=b['vis'], frequency=freq, precision=64)
    bv.phasecentre = arlmi.SkyCoord(*PC)
    rows_then = arlmi.phaserotate_visibility(arlmi.convert_blockvisibility_to_visibility(bv), arlmi.SkyCoord(*NEW),
                                             tangent=False)
    then_rows = arlmi.convert_blockvisibility_to_visibility(
        arlmi.phaserotate_visibility(bv, arlmi.SkyCoord(*NEW), tangent=False))
    assert _rel(then_rows.vis.cpu().numpy(), rows_then.vis.cpu().numpy()) < 1e-11
    assert _rel(then_rows.data["uvw"].cpu().numpy(), rows_then.data["uvw"].cpu().numpy()) < 1e-13


def test_full_size_round_trip(gpu):
    """The 29.7 M rows of the w-stacked 2048^2 workload in float32: rotating 2 degrees away and
    back (tangent=False) restores the uvw to float64 rounding and the visibilities to float32
    rounding; the rotation of one row in 100,000 matches the oracle."""
    import torch
    import arlmi
    from arlmi import simulation
    vis, im, _ = simulation.config2(precision=32, sky=False)
    g = torch.Generator(device="cuda").manual_seed(5)
    vis.data["vis"] = torch.complex(torch.randn(vis.nvis, 1, device="cuda", generator=g),
                                    torch.randn(vis.nvis, 1, device="cuda", generator=g)).contiguous()
    vis.phasecentre = arlmi.SkyCoord(0.0, -30.0 * D)
    there = arlmi.SkyCoord(2.0 * D, -31.5 * D)
    r1 = arlmi.phaserotate_visibility(vis, there, tangent=False)
    r2 = arlmi.phaserotate_visibility(r1, vis.phasecentre, tangent=False)
    u0, u2 = vis.data["uvw"], r2.data["uvw"]
    assert float((u2 - u0).abs().max()) <= 1e-12 * float(u0.abs().max())
    dv = (r2.vis - vis.vis).abs()
    assert float(dv.max()) < 2e-6 * float(vis.vis.abs().max())
    idx = numpy.arange(0, vis.nvis, 100000)
    sub = {'uvw': vis.data["uvw"][idx].cpu().numpy(), 'vis': vis.vis[idx].cpu().numpy().astype(numpy.complex128)}
    ruvw, rvis = O.phaserotate_rows(sub, (0.0, -30.0 * D), (2.0 * D, -31.5 * D), tangent=False)
    assert numpy.array_equal(r1.data["uvw"][idx].cpu().numpy(), ruvw)
    assert _rel(r1.vis[idx].cpu().numpy().astype(numpy.complex128), rvis) < 2e-7
