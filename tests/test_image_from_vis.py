"""create_image_from_visibility (imaging/base.py:247-339) on host tensors: the WCS numbers ARL
writes (cdelt, crpix, crval from the phase centre and the first frequency), the cell rule (half
the critical cell 1/(2 max|u|), capped at it unless override_cellsize=False), u in wavelengths for
a BlockVisibility (metres times max frequency / c), and ARL's
spectral-mode checks."""

import numpy
import pytest
import torch

from oracle import arl_oracle as O

CPU = torch.device("cpu")
D2R = numpy.pi / 180.0


def _blocks(nchan=3):
    rng = numpy.random.default_rng(4)
    xyz = O.xyz_at_latitude(numpy.stack([rng.uniform(-900, 900, 6), rng.uniform(-900, 900, 6),
                                         rng.uniform(-1, 1, 6)], 1), -0.45)
    freq = 1.2e8 + 1e6 * numpy.arange(nchan)
    b = O.create_blockvisibility(xyz, numpy.linspace(-0.3, 0.3, 5), freq, numpy.full(nchan, 1e6), -0.6)
    return b, freq


def _bvis(b, freq, pc=None):
    import arlmi
    bv = arlmi.BlockVisibility.from_arrays(b['uvw'], vis=b['vis'], time=b['time'], frequency=freq,
                                           channel_bandwidth=numpy.full(len(freq), 1e6), precision=64, dev=CPU)
    bv.phasecentre = pc
    return bv


def test_wcs_and_cell_rule():
    import arlmi
    b, freq = _blocks()
    pc = arlmi.SkyCoord.from_degrees(15.0, -45.0)
    bv = _bvis(b, freq, pc)
    im = arlmi.create_image_from_visibility(bv, npixel=128)
    umax = numpy.max(numpy.abs(b['uvw'][..., 0])) * numpy.max(freq) / O.C_LIGHT
    crit = 1.0 / (2.0 * umax)
    w = im.wcs.wcs
    assert im.data.shape == (3, 1, 128, 128) and im.data.device == CPU
    numpy.testing.assert_allclose(w.cdelt[:2], [-0.5 * crit / D2R, 0.5 * crit / D2R], rtol=1e-14)
    assert list(w.crpix[:2]) == [65.0, 65.0]
    assert (w.crval[0], w.crval[1], w.crval[3]) == (numpy.rad2deg(pc.ra), numpy.rad2deg(pc.dec), freq[0])
    assert w.cdelt[3] == 0.99999999999 * 1e6
    assert im.phasecentre.ra == pc.ra and im.phasecentre.dec == pc.dec
    # a cell above the critical one is capped, unless override_cellsize=False
    big = arlmi.create_image_from_visibility(bv, npixel=64, cellsize=2 * crit)
    assert abs(big.cellsize - crit) < 1e-15 * crit
    kept = arlmi.create_image_from_visibility(bv, npixel=64, cellsize=2 * crit, override_cellsize=False)
    assert abs(kept.cellsize - 2 * crit) < 1e-15 * crit
    # MFS: one channel; no phase centre: no sky position
    mfs = arlmi.create_image_from_visibility(_bvis(b, freq), npixel=64, nchan=1)
    assert mfs.data.shape[0] == 1 and mfs.phasecentre is None and not mfs.wcs.celestial


def test_rows_first_frequency_and_spectral_modes():
    import arlmi
    rng = numpy.random.default_rng(5)
    uvw = rng.uniform(-300, 300, (40, 3))
    freq = numpy.repeat([1.1e8, 1.0e8], 20)  # the first row's frequency is the reference, not the smallest
    vis = arlmi.Visibility.from_arrays(uvw, vis=numpy.ones((40, 1)), frequency=freq,
                                       channel_bandwidth=numpy.full(40, 2e6), precision=32, dev=CPU)
    im = arlmi.create_image_from_visibility(vis, npixel=32)
    assert im.data.shape == (2, 1, 32, 32) and im.data.dtype == torch.float32
    assert im.wcs.wcs.crval[3] == 1.1e8 and im.wcs.wcs.cdelt[3] == 0.99999999999 * 2e6
    assert abs(im.cellsize - 0.5 / (2.0 * numpy.max(numpy.abs(uvw[:, 0])))) < 1e-18
    one = arlmi.Visibility.from_arrays(uvw, vis=numpy.ones((40, 1)), frequency=numpy.full(40, 1e8),
                                       channel_bandwidth=numpy.full(40, 2e6), dev=CPU)
    with pytest.raises(ValueError):
        arlmi.create_image_from_visibility(one, npixel=32, nchan=3)
    with pytest.raises(AssertionError):
        arlmi.create_image_from_visibility(one, npixel=32, channel_bandwidth=0.0)
    with pytest.raises(AssertionError):
        arlmi.create_image_from_visibility("not visibilities")
