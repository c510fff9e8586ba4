"""The oracle's shift of visibilities to an image centre (shift_vis_to_image, imaging/base.py:50-81)
pinned without retyping it (CPU only).

- pixel_to_skycoord_sin: the reference pixel is crval exactly; on the equator an x offset lands at
  dec 0 and ra0 + asin(l); skycoord_to_lmn (pinned by ARL's known answers) of a deprojected
  direction returns that pixel's intermediate coordinates cdelt * (p - crpix), for any centre.
- separation: 0 for one direction, pi/2 from a pole to the equator, |d ra| along the equator and
  |d dec| along a meridian.
- shift_vis_to_image: centres closer than 1e-15 rad give the inputs back; a point source simulated
  at the phase centre and imaged through the shift onto an image centred elsewhere peaks at the
  pixel of its (l, m) about the image centre (the inverse rotation, a sign slip, does not), and the
  inverse shift undoes the shift.
"""

import numpy
import pytest
from numpy.testing import assert_allclose

from helpers import random_rows
from oracle import arl_oracle as O

D2R = numpy.pi / 180.0


def _wcs(n, cell):
    crpix = (float(n // 2 + 1), float(n // 2 + 1))
    cdelt = (-cell / D2R, cell / D2R)
    return crpix, cdelt


def test_reference_pixel_and_equator():
    crpix, cdelt = _wcs(64, 1e-3)
    assert O.pixel_to_skycoord_sin(crpix[0], crpix[1], crpix, cdelt, (37.0, -51.0)) == (37.0 * D2R, -51.0 * D2R)
    for k in (-20, -3, 1, 17):
        ra, dec = O.pixel_to_skycoord_sin(crpix[0] + k, crpix[1], crpix, cdelt, (10.0, 0.0))
        l = -1e-3 * k   # cdelt0 is negative: ra falls to the right
        assert abs(dec) < 1e-15
        assert_allclose(ra, 10.0 * D2R + numpy.arcsin(l), rtol=0, atol=1e-15)


@pytest.mark.parametrize("centre", [(0.0, 0.0), (15.0, -45.0), (200.0, 60.0), (359.0, -89.0)])
def test_deprojection_inverts_lmn(centre):
    crpix, cdelt = _wcs(256, 2e-3)
    rng = numpy.random.default_rng(1)
    for px, py in rng.uniform(1, 256, (20, 2)):
        ra, dec = O.pixel_to_skycoord_sin(px, py, crpix, cdelt, centre)
        l, m, n = O.skycoord_to_lmn(ra, dec, centre[0] * D2R, centre[1] * D2R)
        assert_allclose((l, m), (cdelt[0] * D2R * (px - crpix[0]), cdelt[1] * D2R * (py - crpix[1])),
                        rtol=0, atol=2e-15)
        assert_allclose(n, numpy.sqrt(1 - l * l - m * m) - 1.0, rtol=0, atol=1e-15)


def test_separation_known_angles():
    assert O.separation((0.3, -0.2), (0.3, -0.2)) == 0.0
    assert_allclose(O.separation((1.0, numpy.pi / 2), (2.5, 0.0)), numpy.pi / 2, rtol=1e-15)
    assert_allclose(O.separation((0.1, 0.0), (0.45, 0.0)), 0.35, rtol=1e-14)
    assert_allclose(O.separation((2.0, -0.7), (2.0, -0.2)), 0.5, rtol=1e-14)
    assert_allclose(O.separation((0.0, 0.0), (1e-9, 0.0)), 1e-9, rtol=1e-12)


def _point_source_image(inverse):
    """Dirty image (oracle invert_2d, 128^2, 2e-3 rad cells) of a unit source 7 pixels right and 4
    down of the image centre C1, which lies 5 pixels left and 3 up of the phase centre C0."""
    n, cell, S = 128, 2e-3, 6
    crpix, cdelt = _wcs(n, cell)
    c0 = (40.0, -30.0)
    c1 = numpy.array(O.pixel_to_skycoord_sin(crpix[0] - 5, crpix[1] + 3, crpix, cdelt, c0)) / D2R
    src = O.pixel_to_skycoord_sin(crpix[0] + 7, crpix[1] - 4, crpix, cdelt, c1)
    rng = numpy.random.default_rng(2)
    d = random_rows(rng, 6000, n, cell, S)
    l, m, _ = O.skycoord_to_lmn(src[0], src[1], c0[0] * D2R, c0[1] * D2R)
    d["vis"] = O.simulate_point(d["uvw"], l, m)[:, None]
    uvw, v, label = O.shift_vis_to_image(d, (c0[0] * D2R, c0[1] * D2R), (c1[0] * D2R, c1[1] * D2R),
                                         inverse=inverse)
    g = O.ImageGeom(n, cell)
    gcf, cf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    dirty, _ = O.invert_2d(dict(d, uvw=uvw, vis=v), g, gcf, cf, cfg)
    return dirty[0, 0], n, label, c1


def test_point_source_lands_on_its_pixel():
    img, n, label, c1 = _point_source_image(False)
    assert numpy.unravel_index(numpy.argmax(img), img.shape) == (n // 2 - 4, n // 2 + 7)
    assert img.max() > 0.95
    assert_allclose(label, c1 * D2R)
    wrong, _, _, _ = _point_source_image(True)
    assert numpy.unravel_index(numpy.argmax(wrong), wrong.shape) != (n // 2 - 4, n // 2 + 7)


def test_no_shift_and_round_trip():
    rng = numpy.random.default_rng(3)
    d = random_rows(rng, 500, 64, 1e-3, 6)
    c0 = (0.5, -0.4)
    uvw, v, label = O.shift_vis_to_image(d, c0, c0)
    assert uvw is d["uvw"] and v is d["vis"] and label == c0
    c1 = (0.52, -0.41)
    _, v1, _ = O.shift_vis_to_image(d, c0, c1)
    assert numpy.max(numpy.abs(v1 - d["vis"])) > 0.1
    _, v2, _ = O.shift_vis_to_image(dict(d, vis=v1), c0, c1, inverse=True)
    assert_allclose(v2, d["vis"], rtol=0, atol=1e-12)
