"""Scalar gain solution on the GPU (arlmi_solve_gains) against the oracle and ARL's assertions.

Both sides iterate in float64 on the same (precision-rounded) visibilities; they stop when the
largest gain change falls below tol, and rounding can move that stop by an iteration, so gains
are compared to 1e-7 (the size of a change below tol = 1e-8 after damping), weights to 1e-6,
residuals to each other and, for float64 data, to ARL's bound 3e-8 (complex64 storage of
noise-free data alone leaves a residual of ~3e-8).
"""

import numpy
import pytest
from numpy.testing import assert_allclose

from oracle import arl_oracle as O

pytestmark = pytest.mark.gpu


def _case(rng, nants=12, nchan=2, ntimes=1, per_row=1, phase_sd=10.0, amp_sd=0.0, precision=64):
    xyz = O.xyz_at_latitude(numpy.stack([rng.uniform(-2e3, 2e3, nants), rng.uniform(-2e3, 2e3, nants),
                                         rng.uniform(-5, 5, nants)], 1), -0.5)
    freq = numpy.linspace(1e8, 1.2e8, nchan)
    b = O.create_blockvisibility(xyz, numpy.linspace(-0.2, 0.2, ntimes), freq, numpy.full(nchan, 1e6), -0.7)
    model = (rng.normal(size=b['vis'].shape) + 1j * rng.normal(size=b['vis'].shape)) * 10.0
    model = 0.5 * (model + numpy.conj(numpy.swapaxes(model, 1, 2)))
    ngain = ntimes // per_row
    g = numpy.exp(1j * rng.normal(0.0, phase_sd, (ngain, nants, nchan))) * \
        (1.0 + rng.normal(0.0, amp_sd, (ngain, nants, nchan)))
    slots = numpy.arange(ntimes) // per_row
    vis = O.apply_gaintable(model, g, numpy.ones(g.shape), slots)
    if precision == 32:  # the oracle sees the values the GPU holds
        vis = vis.astype(numpy.complex64).astype(numpy.complex128)
        model = model.astype(numpy.complex64).astype(numpy.complex128)
    times = b['time']
    gtime = numpy.array([times[slots == s].mean() for s in range(ngain)])
    interval = numpy.array([(times[slots == s].max() - times[slots == s].min()) + 1.0 for s in range(ngain)])
    return b, vis, model, g, gtime, interval, freq


def _gpu(b, vis, model, gtime, interval, freq, precision, weight=None, **kw):
    import torch
    import arlmi
    bv = arlmi.BlockVisibility.from_arrays(b['uvw'], vis=vis, weight=weight, time=b['time'], frequency=freq,
                                           precision=precision)
    mv = None
    if model is not None:
        mv = arlmi.BlockVisibility.from_arrays(b['uvw'], vis=model, time=b['time'], frequency=freq,
                                               precision=precision)
    na, nc = b['uvw'].shape[1], len(freq)
    gt = arlmi.GainTable(torch.ones((len(gtime), na, nc, 1, 1), dtype=torch.complex128, device="cuda"), gtime,
                         interval, weight=torch.ones((len(gtime), na, nc, 1, 1), dtype=torch.float64, device="cuda"),
                         frequency=freq)
    sol = arlmi.solve_gaintable(bv, mv, gt=gt, **kw)
    return (sol.gain.cpu().numpy(), sol.weight.cpu().numpy(), sol.residual.cpu().numpy(),
            sol.iterations.cpu().numpy())


def _compare(got, want, precision=64):
    assert_allclose(got[0], want[0], atol=1e-7)
    assert_allclose(got[1], want[1], rtol=1e-6, atol=1e-9)
    assert_allclose(got[2], want[2], rtol=1e-4, atol=1e-12)
    if precision == 64:  # ARL's bound; complex64 storage of the data alone leaves ~3e-8
        assert numpy.max(got[2]) < 3e-8 and numpy.max(want[2]) < 3e-8


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("nchan,ntimes,per_row", [(1, 1, 1), (3, 4, 2), (2, 6, 3)])
def test_phase_only(gpu, precision, nchan, ntimes, per_row):
    rng = numpy.random.default_rng(nchan * 10 + ntimes)
    b, vis, model, g, gtime, interval, freq = _case(rng, nchan=nchan, ntimes=ntimes, per_row=per_row,
                                                    precision=precision)
    g0 = numpy.ones((len(gtime), 12, nchan, 1, 1), dtype=complex)
    want = O.solve_gaintable_scalar({'vis': vis, 'weight': numpy.ones(vis.shape), 'time': b['time']}, model, g0,
                                    gtime, interval, phase_only=True, niter=200)
    got = _gpu(b, vis, model, gtime, interval, freq, precision, phase_only=True, niter=200)
    _compare(got, want, precision)
    assert numpy.max(numpy.abs(got[0] - 1.0)) > 0.1


@pytest.mark.parametrize("precision", [64, 32])
def test_amplitude_phase_normalised(gpu, precision):
    rng = numpy.random.default_rng(3)
    b, vis, model, g, gtime, interval, freq = _case(rng, nchan=2, phase_sd=1.0, amp_sd=0.1, precision=precision)
    g0 = numpy.ones((1, 12, 2, 1, 1), dtype=complex)
    want = O.solve_gaintable_scalar({'vis': vis, 'weight': numpy.ones(vis.shape), 'time': b['time']}, model, g0,
                                    gtime, interval, phase_only=False, niter=200, normalise_gains=True)
    got = _gpu(b, vis, model, gtime, interval, freq, precision, phase_only=False, niter=200, normalise_gains=True)
    _compare(got, want, precision)


def test_point_source_and_dead_antenna(gpu):
    rng = numpy.random.default_rng(4)
    b, vis, model, g, gtime, interval, freq = _case(rng, nchan=2)
    pvis = O.apply_gaintable(numpy.ones_like(model), g, numpy.ones(g.shape), numpy.zeros(1, dtype=int))
    w = numpy.ones(vis.shape)
    w[:, 4] = 0.0
    w[:, :, 4] = 0.0
    g0 = numpy.ones((1, 12, 2, 1, 1), dtype=complex)
    want = O.solve_gaintable_scalar({'vis': pvis, 'weight': w, 'time': b['time']}, None, g0, gtime, interval,
                                    phase_only=True, niter=200)
    got = _gpu(b, pvis, None, gtime, interval, freq, 64, weight=w, phase_only=True, niter=200)
    _compare(got, want)
    assert numpy.all(got[1][0, 4] == 0) and numpy.all(numpy.abs(got[0][0, 4]) < 1e-7)


def test_niter_caps_and_unmatched_rows(gpu):
    """niter = 3 stops early (iteration counts equal); a gain row with no integration is left as
    it was (gain 1, weight 1, residual 0)."""
    rng = numpy.random.default_rng(6)
    b, vis, model, g, gtime, interval, freq = _case(rng, nchan=1, ntimes=2, per_row=1)
    gtime = numpy.append(gtime, gtime[-1] + 1000.0)
    interval = numpy.append(interval, 1.0)
    g0 = numpy.ones((3, 12, 1, 1, 1), dtype=complex)
    want = O.solve_gaintable_scalar({'vis': vis, 'weight': numpy.ones(vis.shape), 'time': b['time']}, model, g0,
                                    gtime, interval, phase_only=True, niter=3)
    got = _gpu(b, vis, model, gtime, interval, freq, 64, phase_only=True, niter=3)
    assert_allclose(got[0], want[0], atol=1e-12)
    assert numpy.array_equal(got[3][:2], want[3][:2]) and got[3][2] == 0
    assert numpy.all(got[0][2] == 1.0) and numpy.all(got[1][2] == 1.0) and numpy.all(got[2][2] == 0.0)


def test_full_size_128_antennas(gpu):
    """The differentiation workload's BlockVisibility (128 antennas, 4 channels, 60 integrations,
    one gain row per integration): random Hermitian model, its simulated gains applied on the GPU,
    solved on the GPU; residual below 3e-8, the gains recovered (conj(g) with antenna 0's phase
    removed) and every row against the oracle."""
    import torch
    import arlmi
    from arlmi.simulation import config4
    c = config4(precision=64)
    bv, gt = c["bvis"], c["gaintable"]
    rng = numpy.random.default_rng(21)
    shape = tuple(bv.data["vis"].shape)
    m = (rng.normal(size=shape) + 1j * rng.normal(size=shape)) * 5.0
    m = 0.5 * (m + numpy.conj(numpy.swapaxes(m, 1, 2)))
    model = bv.copy()
    model.data["vis"] = torch.as_tensor(m).to("cuda").contiguous()
    obs = model.copy()
    arlmi.apply_gaintable(obs, gt)
    start = arlmi.create_gaintable_from_blockvisibility(bv)
    start.gain = start.gain.to(torch.complex128)
    sol = arlmi.solve_gaintable(obs, model, gt=start, phase_only=False, niter=200, normalise_gains=False)
    assert float(sol.residual.max()) < 3e-8
    gtrue = gt.gain.cpu().numpy()[..., 0, 0]
    e = numpy.conj(gtrue)
    e = e * numpy.exp(-1j * numpy.angle(e[:, 0:1, :]))
    assert_allclose(sol.gain.cpu().numpy()[..., 0, 0], e, atol=1e-6)
    want = O.solve_gaintable_scalar({'vis': obs.data["vis"].cpu().numpy(), 'weight': numpy.ones(shape),
                                     'time': bv.time.cpu().numpy()}, m, numpy.ones((gt.ntimes, 128, 4, 1, 1),
                                                                                   dtype=complex),
                                    start.time, start.interval, phase_only=False, niter=200,
                                    normalise_gains=False)
    assert_allclose(sol.gain.cpu().numpy(), want[0], atol=1e-7)
