"""Baseline-dependent averaging on the GPU (arlmi_bda_uvdist / arlmi_bda_average) against the
oracle: factors, row counts, antennas and cindex exactly; averaged columns to float64 rounding
(the two-stage sums run in time and channel order; numpy's chunk sums go through a BLAS
product), complex64 visibilities and weights to float32 rounding."""

import numpy
import pytest

from oracle import arl_oracle as O

pytestmark = pytest.mark.gpu


def _blocks(rng, nants=9, ntimes=20, nchan=4, radius=400.0):
    xyz = O.xyz_at_latitude(numpy.stack([rng.uniform(-radius, radius, nants), rng.uniform(-radius, radius, nants),
                                         rng.uniform(-1, 1, nants)], 1), -0.45)
    times = (numpy.pi / 43200.0) * numpy.arange(0.0, ntimes * 3.76, 3.76)
    freq = 1e8 + 27343.75 * numpy.arange(nchan)
    b = O.create_blockvisibility(xyz, times, freq, numpy.full(nchan, 27343.75), -0.6, integration_time=3.76)
    b['vis'] = rng.normal(size=b['vis'].shape) + 1j * rng.normal(size=b['vis'].shape)
    b['weight'] = rng.uniform(0.5, 2.0, b['weight'].shape)
    b['weight'][:, 2, 5] = 0.0   # a baseline without weight
    b['imaging_weight'] = rng.uniform(0.1, 1.0, b['weight'].shape)
    return b


def _device(b, precision):
    import torch
    import arlmi
    bv = arlmi.BlockVisibility.from_arrays(b['uvw'], vis=b['vis'], weight=b['weight'],
                                           imaging_weight=b['imaging_weight'], time=b['time'],
                                           integration_time=b['integration_time'], frequency=b['frequency'],
                                           channel_bandwidth=b['channel_bandwidth'], precision=precision)
    return bv


def _rel(a, b):
    d = numpy.max(numpy.abs(b))
    return float(numpy.max(numpy.abs(a - b)) / (d if d > 0 else 1.0))


CASES = [dict(time_coal=1.0, frequency_coal=1.0), dict(time_coal=2.0, max_time_coal=7, frequency_coal=0.0,
                                                       max_frequency_coal=1),
         dict(time_coal=0.0, max_time_coal=1, frequency_coal=1.5, max_frequency_coal=3),
         dict(time_coal=0.7, frequency_coal=0.4)]


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_average_in_blocks_against_oracle(gpu, precision, case):
    import arlmi
    kw = CASES[case]
    rng = numpy.random.default_rng(10 + case)
    b = _blocks(rng)
    if precision == 32:
        for k in ('vis', 'weight', 'imaging_weight'):
            b[k] = b[k].astype(numpy.complex64 if k == 'vis' else numpy.float32).astype(b[k].dtype)
    full = dict(time_coal=1.0, max_time_coal=100, frequency_coal=1.0, max_frequency_coal=100)
    full.update(kw)
    want = O.average_in_blocks(b['vis'], b['uvw'], b['weight'], b['imaging_weight'], b['time'],
                               b['integration_time'], b['frequency'], b['channel_bandwidth'], **full)
    got = arlmi.average_in_blocks(_device(b, precision), **full)
    ta, fa = O.bda_factors(b['uvw'], b['weight'], full['time_coal'], full['max_time_coal'], full['frequency_coal'],
                           full['max_frequency_coal'])
    assert numpy.array_equal(got['time_average'], ta) and numpy.array_equal(got['frequency_average'], fa)
    assert got['vis'].shape[0] == want['vis'].shape[0]
    for k in ('antenna1', 'antenna2', 'cindex'):
        assert numpy.array_equal(got[k].cpu().numpy(), want[k]), k
    tol = 1e-12 if precision == 64 else 2e-6
    for k in ('time', 'frequency', 'uvw', 'integration_time', 'channel_bandwidth'):
        assert _rel(got[k].cpu().numpy(), want[k]) < 1e-12, k
    for k in ('vis', 'weight', 'imaging_weight'):
        assert _rel(got[k].cpu().numpy().astype(want[k].dtype), want[k]) < tol, k


def test_coalesce_and_decoalesce(gpu):
    """coalesce_visibility with the ARL test's settings and decoalesce_visibility: every block
    element takes the value of its cindex row; zero factors give the plain conversion."""
    import arlmi
    rng = numpy.random.default_rng(30)
    b = _blocks(rng, ntimes=30, nchan=3)
    b['weight'][...] = 1.0
    bv = _device(b, 64)
    cvis = arlmi.coalesce_visibility(bv, time_coal=1.0, frequency_coal=1.0)
    assert float(cvis.data["frequency"].min()) == numpy.min(b['frequency'])
    want = O.average_in_blocks(b['vis'], b['uvw'], b['weight'], b['imaging_weight'], b['time'],
                               b['integration_time'], b['frequency'], b['channel_bandwidth'])
    dvis = arlmi.decoalesce_visibility(cvis)
    expect = want['vis'][want['cindex'], 0].reshape(b['vis'].shape)
    assert _rel(dvis.data["vis"].cpu().numpy(), expect) < 1e-12
    plain = arlmi.coalesce_visibility(bv)
    assert plain.nvis == 30 * (9 * 8 // 2) * 3


def test_full_size_128_antennas(gpu):
    """The differentiation workload's BlockVisibility (128 antennas, 60 integrations, 4 channels)
    averaged with time_coal = frequency_coal = 1: the factors exactly, the total weight conserved,
    and six baselines' rows against the oracle's average_chunks2 with those factors (the oracle's
    loop over all 16,384 baselines would take minutes)."""
    import torch
    import arlmi
    from arlmi.simulation import config4
    c = config4(precision=64)
    bv = c["bvis"]
    rng = numpy.random.default_rng(31)
    shape = tuple(bv.data["vis"].shape)
    v = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    w = rng.uniform(0.5, 2.0, shape)
    bv.data["vis"] = torch.as_tensor(v).to("cuda").contiguous()
    bv.data["weight"] = torch.as_tensor(w).to("cuda").contiguous()
    got = arlmi.average_in_blocks(bv, 1.0, 100, 1.0, 100)
    uvw = bv.data["uvw"].cpu().numpy()
    ta, fa = O.bda_factors(uvw, w, 1.0, 100, 1.0, 100)
    assert numpy.array_equal(got['time_average'], ta) and numpy.array_equal(got['frequency_average'], fa)
    gw = got['weight'].cpu().numpy()
    assert abs(gw.sum() - w.sum()) < 1e-10 * w.sum()
    a1s, a2s = got['antenna1'].cpu().numpy(), got['antenna2'].cpu().numpy()
    gv = got['vis'].cpu().numpy()[:, 0]
    for a2, a1 in [(5, 3), (127, 0), (64, 64), (10, 90), (100, 99), (1, 0)]:
        rows = (a2s == a2) & (a1s == a1)
        ref, refw = O.average_chunks2(v[:, a2, a1, :, 0], w[:, a2, a1, :, 0], (ta[a2, a1], fa[a2, a1]))
        assert _rel(gv[rows], ref.reshape(-1)) < 1e-12
        assert _rel(gw[rows, 0], refw.reshape(-1)) < 1e-12
