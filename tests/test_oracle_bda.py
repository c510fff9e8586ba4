"""The oracle's baseline-dependent averaging (average_in_blocks) pinned without retyping it.

- ARL's assertions (test_visibility_coalesce.py:39-110): the smallest frequency survives the
  averaging and decoalescence gives one value per block element.
- Factors by hand: antennas on a line 100 m and 300 m from the first give baselines of 100, 200
  and 300 m, so round(300/100) = 3, round(300/200) = 2 (numpy rounds half to even) and 1; the
  diagonal (length 0) takes the maximum.
- Factors of 1 keep every sample: rows equal the block elements (both triangles and the diagonal,
  in a2, a1 order) bit for bit, uvw in wavelengths.
- Constant visibilities average to themselves; the summed weights and the weighted visibility sum
  of every baseline are conserved.
"""

import numpy
import pytest
from numpy.testing import assert_allclose

from oracle import arl_oracle as O


def _blocks(rng, nants=8, ntimes=12, nchan=3, radius=300.0):
    xyz = O.xyz_at_latitude(numpy.stack([rng.uniform(-radius, radius, nants), rng.uniform(-radius, radius, nants),
                                         rng.uniform(-1, 1, nants)], 1), -0.45)
    times = (numpy.pi / 43200.0) * numpy.arange(0.0, ntimes * 3.76, 3.76)
    df = 27343.75
    freq = 1e8 + df * (numpy.arange(nchan) - nchan // 2)
    b = O.create_blockvisibility(xyz, times, freq, numpy.full(nchan, df), -0.6, integration_time=3.76)
    b['vis'] = rng.normal(size=b['vis'].shape) + 1j * rng.normal(size=b['vis'].shape)
    b['weight'] = rng.uniform(0.5, 2.0, b['weight'].shape)
    b['imaging_weight'] = rng.uniform(0.1, 1.0, b['weight'].shape)
    return b


def _run(b, **kw):
    return O.average_in_blocks(b['vis'], b['uvw'], b['weight'], b['imaging_weight'], b['time'],
                               b['integration_time'], b['frequency'], b['channel_bandwidth'], **kw)


def test_arl_assertions():
    rng = numpy.random.default_rng(1)
    b = _blocks(rng, ntimes=30)
    b['weight'][...] = 1.0  # as ARL's test (weight=1.0): a weighted mean of one value can round off it
    for kw in [dict(time_coal=1.0, frequency_coal=1.0), dict(time_coal=0.0, max_time_coal=1, frequency_coal=1.0),
               dict(time_coal=1.0, frequency_coal=0.0, max_frequency_coal=1)]:
        out = _run(b, **kw)
        assert numpy.min(out['frequency']) == numpy.min(b['frequency'])
        dvis = out['vis'][out['cindex']]
        assert dvis.shape[0] == b['vis'].size


def test_factors_by_hand():
    uvw = numpy.zeros((1, 3, 3, 3))
    x = numpy.array([0.0, 100.0, 300.0])
    uvw[0, :, :, 0] = x[:, None] - x[None, :]
    wts = numpy.ones((1, 3, 3, 1, 1))
    ta, fa = O.bda_factors(uvw, wts, 1.0, 7, 2.0, 5)
    assert ta[1, 0] == 3 and ta[2, 1] == 2 and ta[2, 0] == 1 and ta[0, 1] == 3
    assert ta[0, 0] == ta[1, 1] == 7
    # round(2*300/100) = 6, capped at 5; round(2*300/200) = 3; round(2*300/300) = 2
    assert fa[1, 0] == 5 and fa[2, 1] == 3 and fa[2, 0] == 2
    wts[0, 1, 0] = 0.0
    ta, _ = O.bda_factors(uvw, wts, 1.0, 7, 1.0, 7)
    assert ta[1, 0] == 1   # no weight: 1


def test_factor_one_keeps_samples():
    rng = numpy.random.default_rng(2)
    b = _blocks(rng, nants=5, ntimes=4, nchan=2)
    # tiny coalescence factors round to 0 (then 1); the diagonal takes the maximum, here also 1
    out = _run(b, time_coal=1e-9, frequency_coal=1e-9, max_time_coal=1, max_frequency_coal=1)
    nt, na, _, nc, _ = b['vis'].shape
    assert out['vis'].shape[0] == nt * na * na * nc
    order = b['vis'][:, :, :, :, 0].transpose(1, 2, 0, 3).reshape(-1)   # (a2, a1, t, ch)
    assert numpy.array_equal(out['vis'][:, 0], order)
    assert numpy.array_equal(out['weight'][:, 0], b['weight'][..., 0].transpose(1, 2, 0, 3).reshape(-1))
    k = b['frequency'] / O.C_LIGHT
    u = (b['uvw'][:, :, :, 0][..., None] * k).transpose(1, 2, 0, 3).reshape(-1)
    assert numpy.array_equal(out['uvw'][:, 0], u)
    assert numpy.array_equal(out['integration_time'], numpy.full(out['vis'].shape[0], 3.76 * nt * nc))
    assert numpy.array_equal(out['antenna2'], numpy.repeat(numpy.arange(na), na * nt * nc))


@pytest.mark.parametrize("tc,fc", [(1.0, 1.0), (2.5, 0.0), (0.0, 3.0)])
def test_constant_data_and_conservation(tc, fc):
    rng = numpy.random.default_rng(3)
    b = _blocks(rng, ntimes=17, nchan=5)
    kw = dict(time_coal=tc, frequency_coal=fc, max_time_coal=6, max_frequency_coal=4)
    if tc == 0.0:
        kw['max_time_coal'] = 1
    if fc == 0.0:
        kw['max_frequency_coal'] = 1
    out = _run(b, **kw)
    assert_allclose(numpy.sum(out['weight']), numpy.sum(b['weight']), rtol=1e-12)
    # weighted visibility sum per baseline
    nt, na, _, nc, _ = b['vis'].shape
    for a2, a1 in [(3, 1), (0, 5), (2, 2)]:
        rows = (out['antenna2'] == a2) & (out['antenna1'] == a1)
        got = numpy.sum(out['vis'][rows, 0] * out['weight'][rows, 0])
        want = numpy.sum(b['vis'][:, a2, a1, :, 0] * b['weight'][:, a2, a1, :, 0])
        assert_allclose(got, want, rtol=1e-12)
    c = dict(b)
    c['vis'] = numpy.full(b['vis'].shape, 2.0 - 3.0j)
    assert_allclose(_run(c, **kw)['vis'], 2.0 - 3.0j, rtol=1e-14)
