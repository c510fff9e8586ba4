"""The oracle's phase rotation pinned by ARL's own tests and identities (CPU only).

- skycoord_to_lmn: the known answers of test_coordinate_support.py:125-141.
- test_coordinate_support.py:143-170: for every (phase centre, new centre, source) among six
  positions, the source's visibility at the old centre times the phasor of the old centre seen
  from the new one equals the source's visibility predicted on the rotated uvw.
- test_visibility_operations.py:140-245: rotation with tangent=False equals a prediction made at
  the new centre with the hour angles shifted by -(ra_new - ra_old); rotating there and back, with
  or without inverse, is the identity.
- The block form agrees with the row form through convert_blocks.
"""

import numpy
import pytest
from numpy.testing import assert_allclose

from oracle import arl_oracle as O

D = numpy.pi / 180.0


def test_skycoord_to_lmn_known_answers():
    center, north, south = (0.0, 0.0), (0.0, 90 * D), (0.0, -90 * D)
    east, west = (90 * D, 0.0), (-90 * D, 0.0)

    def lmn(a, b):
        return O.skycoord_to_lmn(a[0], a[1], b[0], b[1])

    assert_allclose(lmn(center, center), (0, 0, 0))
    assert_allclose(lmn(north, center), (0, 1, -1))
    assert_allclose(lmn(south, center), (0, -1, -1))
    assert_allclose(lmn(south, north), (0, 0, -2), atol=1e-14)
    assert_allclose(lmn(east, center), (1, 0, -1))
    assert_allclose(lmn(west, center), (-1, 0, -1))
    assert_allclose(lmn(center, west), (1, 0, -1))
    assert_allclose(lmn(north, west), (0, 1, -1), atol=1e-14)
    assert_allclose(lmn(south, west), (0, -1, -1), atol=1e-14)
    assert_allclose(lmn(north, east), (0, 1, -1), atol=1e-14)
    assert_allclose(lmn(south, east), (0, -1, -1), atol=1e-14)


def test_uvw_to_xyz_inverts_xyz_to_uvw():
    rng = numpy.random.default_rng(0)
    xyz = rng.normal(size=(50, 3)) * 1e3
    for ha, dec in [(0.3, -0.5), (-1.2, 0.9), (2.0, -1.5)]:
        uvw = O.xyz_to_uvw(xyz, ha, dec)
        assert_allclose(O.uvw_to_xyz(uvw, ha, dec), xyz, atol=1e-9)
        assert_allclose(numpy.linalg.norm(uvw, axis=1), numpy.linalg.norm(xyz, axis=1), rtol=1e-14)


POSITIONS = [(17 * D, 35 * D), (17 * D, 30 * D), (12 * D, 30 * D), (11 * D, 35 * D), (51 * D, 35 * D),
             (15 * D, 70 * D)]


def test_phase_rotate_identity():
    """test_coordinate_support.py:143-170 with the oracle's functions."""
    uvw = numpy.array([(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)])
    for pc in POSITIONS:
        for npc in POSITIONS:
            xyz = O.uvw_to_xyz(uvw, -pc[0], pc[1])
            uvw_rotated = O.xyz_to_uvw(xyz, -npc[0], npc[1])
            l_p, m_p, _ = O.skycoord_to_lmn(pc[0], pc[1], npc[0], npc[1])
            phasor = O.simulate_point(uvw_rotated, l_p, m_p)
            for src in POSITIONS:
                l, m, _ = O.skycoord_to_lmn(src[0], src[1], pc[0], pc[1])
                vis = O.simulate_point(uvw, l, m)
                l_r, m_r, _ = O.skycoord_to_lmn(src[0], src[1], npc[0], npc[1])
                vis_rotated = O.simulate_point(uvw_rotated, l_r, m_r)
                assert_allclose(vis * phasor, vis_rotated, atol=1e-10)


def _array(rng, nants=24, radius=1500.0):
    enu = numpy.stack([rng.uniform(-radius, radius, nants), rng.uniform(-radius, radius, nants),
                       rng.uniform(-5, 5, nants)], 1)
    return O.xyz_at_latitude(enu, -26.7 * D)


def _predict(uvw, comps):
    out = numpy.zeros((uvw.shape[0], 1), dtype=complex)
    for (l, m, f) in comps:
        out[:, 0] += f * O.simulate_point(uvw, l, m)
    return out


PC = (180.0 * D, -35.0 * D)
NEW = (181.0 * D, -35.0 * D)


def test_rotation_equals_prediction_at_shifted_hour_angle():
    """test_visibility_operations.py:158-175: a component at the new centre, predicted at the old
    centre and rotated (tangent=False), equals the prediction made at the new centre with the hour
    angle shifted by -(ra_new - ra_old)."""
    rng = numpy.random.default_rng(1)
    xyz = _array(rng)
    times = (numpy.pi / 43200.0) * numpy.arange(0.0, 300.0, 30.0)
    freq = numpy.linspace(1.0e8, 1.1e8, 3)
    vis = O.create_visibility(xyz, times, freq, numpy.full(3, 1e7), PC[1])
    l, m, _ = O.skycoord_to_lmn(NEW[0], NEW[1], PC[0], PC[1])
    vis['vis'] = _predict(vis['uvw'], [(l, m, 100.0)])
    uvw_r, vis_r = O.phaserotate_rows(vis, PC, NEW, tangent=False)
    ha_diff = -(NEW[0] - PC[0])
    pred = O.create_visibility(xyz, times + ha_diff, freq, numpy.full(3, 1e7), NEW[1])
    pred_vis = _predict(pred['uvw'], [(0.0, 0.0, 100.0)])
    assert_allclose(vis_r, pred_vis, rtol=3e-6)
    assert_allclose(uvw_r, pred['uvw'], rtol=3e-6, atol=1e-9)
    # and a component elsewhere: rotated = predicted at its offsets from the new centre
    src = (180.6 * D, -34.7 * D)
    ls, ms, _ = O.skycoord_to_lmn(src[0], src[1], PC[0], PC[1])
    vis['vis'] = _predict(vis['uvw'], [(ls, ms, 3.0)])
    _, vis_r = O.phaserotate_rows(vis, PC, NEW, tangent=False)
    ln, mn, _ = O.skycoord_to_lmn(src[0], src[1], NEW[0], NEW[1])
    assert_allclose(vis_r, _predict(pred['uvw'], [(ln, mn, 3.0)]), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("inverse", [False, True])
def test_rotation_round_trip(inverse):
    """test_visibility_operations.py:140-156, 196-210 (rtol 1e-7 there; here 1e-10: phases reach
    ~6e3 rad, whose float64 rounding is ~1e-12)."""
    rng = numpy.random.default_rng(2)
    xyz = _array(rng)
    times = (numpy.pi / 43200.0) * numpy.arange(0.0, 300.0, 30.0)
    freq = numpy.linspace(1.0e8, 1.1e8, 3)
    vis = O.create_visibility(xyz, times, freq, numpy.full(3, 1e7), PC[1])
    vis['vis'] = rng.normal(size=vis['vis'].shape) + 1j * rng.normal(size=vis['vis'].shape)
    there = (250.0 * D, -60.0 * D)
    for target in [there, (182 * D, -35 * D), (180 * D, -70 * D)]:
        uvw1, vis1 = O.phaserotate_rows(vis, PC, target, tangent=False, inverse=inverse)
        uvw2, vis2 = O.phaserotate_rows({'uvw': uvw1, 'vis': vis1}, target, PC, tangent=False, inverse=inverse)
        assert_allclose(uvw2, vis['uvw'], rtol=1e-12, atol=1e-9)
        assert_allclose(vis2, vis['vis'], rtol=1e-10, atol=1e-10)


def test_same_centre_returns_inputs():
    vis = {'uvw': numpy.ones((3, 3)), 'vis': numpy.ones((3, 1), dtype=complex)}
    uvw, v = O.phaserotate_rows(vis, PC, PC, tangent=False)
    assert uvw is vis['uvw'] and v is vis['vis']


@pytest.mark.parametrize("tangent,inverse", [(True, False), (False, False), (False, True)])
def test_blocks_agree_with_rows(tangent, inverse):
    rng = numpy.random.default_rng(3)
    xyz = _array(rng, nants=10)
    times = numpy.linspace(-0.2, 0.2, 4)
    freq = numpy.array([1.0e8, 1.2e8, 1.5e8])
    b = O.create_blockvisibility(xyz, times, freq, numpy.full(3, 1e6), PC[1])
    b['vis'] = rng.normal(size=b['vis'].shape) + 1j * rng.normal(size=b['vis'].shape)
    uvw_b, vis_b = O.phaserotate_blocks(b['uvw'], b['vis'], freq, PC, NEW, tangent=tangent, inverse=inverse)
    rows = O.convert_blocks(b)
    uvw_r, vis_r = O.phaserotate_rows(rows, PC, NEW, tangent=tangent, inverse=inverse)
    b2 = dict(b)
    b2['uvw'], b2['vis'] = uvw_b, vis_b
    rows2 = O.convert_blocks(b2)
    assert_allclose(rows2['vis'], vis_r, rtol=1e-12, atol=1e-12)
    assert_allclose(rows2['uvw'], uvw_r, rtol=1e-12, atol=1e-9)


def test_product_lmn_matches_oracle():
    """The library's host-side direction cosines (astropy's route: Cartesian vector, frame
    rotations) against the oracle's spherical-trigonometry form."""
    import arlmi
    rng = numpy.random.default_rng(4)
    for _ in range(200):
        ra, ra0 = rng.uniform(0, 2 * numpy.pi, 2)
        dec, dec0 = rng.uniform(-numpy.pi / 2, numpy.pi / 2, 2)
        got = arlmi.skycoord_to_lmn(arlmi.SkyCoord(ra, dec), arlmi.SkyCoord(ra0, dec0))
        assert_allclose(got, O.skycoord_to_lmn(ra, dec, ra0, dec0), rtol=0, atol=2e-15)
