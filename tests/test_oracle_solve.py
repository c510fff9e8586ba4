"""The oracle's scalar gain solution pinned by ARL's own assertions and closed forms (CPU only).

test_calibration_solvers.py:55-115 simulate gains, apply them to a model BlockVisibility, solve
against the model and assert that the solution residual is below 3e-8 and that the gains are not
all 1.  Here the same, plus: the solution reproduces the simulated gains up to the reference
phase (x[a2, a1] = g_a2 conj(g_a1) from apply_gaintable's convention, and the solver's own
convention x[i, j] = conj(g_i) g_j, so the solution is conj(g) rotated to a real reference),
several integrations per solution interval, the point-source form (model = None), amplitude and
phase with normalisation, and an antenna with no weight (its weight is 0 and its gain decays
to 0 under the damping).
"""

import numpy
import pytest
from numpy.testing import assert_allclose

from oracle import arl_oracle as O


def _setup(rng, nants=12, nchan=2, ntimes=1, phase_sd=10.0, amp_sd=0.0, per_row=1):
    xyz = O.xyz_at_latitude(numpy.stack([rng.uniform(-2e3, 2e3, nants), rng.uniform(-2e3, 2e3, nants),
                                         rng.uniform(-5, 5, nants)], 1), -0.5)
    freq = numpy.linspace(1e8, 1.2e8, nchan)
    b = O.create_blockvisibility(xyz, numpy.linspace(-0.2, 0.2, ntimes), freq, numpy.full(nchan, 1e6), -0.7)
    model = (rng.normal(size=b['vis'].shape) + 1j * rng.normal(size=b['vis'].shape)) * 10.0
    # Hermitian like a real correlator: [t, i, j] = conj([t, j, i]), real diagonal
    model = 0.5 * (model + numpy.conj(numpy.swapaxes(model, 1, 2)))
    ngain = ntimes // per_row
    g = numpy.exp(1j * rng.normal(0.0, phase_sd, (ngain, nants, nchan))) * \
        (1.0 + rng.normal(0.0, amp_sd, (ngain, nants, nchan)))
    slots = numpy.arange(ntimes) // per_row
    vis = O.apply_gaintable(model, g, numpy.ones(g.shape), slots)
    times = b['time']
    gtime = numpy.array([times[slots == s].mean() for s in range(ngain)])
    interval = numpy.array([(times[slots == s].max() - times[slots == s].min()) + 1.0 for s in range(ngain)])
    bv = {'vis': vis, 'weight': numpy.ones(vis.shape, dtype=float), 'time': times}
    return bv, model, g, gtime, interval


def _expected(g, refant=0):
    """conj(g), rotated so that the reference antenna's phase is 0 (per row and channel)."""
    c = numpy.conj(g)
    return c * numpy.exp(-1j * numpy.angle(c[:, refant:refant + 1, :]))


@pytest.mark.parametrize("nchan,ntimes,per_row", [(1, 1, 1), (3, 4, 2), (2, 6, 3)])
def test_phase_only(nchan, ntimes, per_row):
    rng = numpy.random.default_rng(nchan * 10 + ntimes)
    bv, model, g, gtime, interval = _setup(rng, nchan=nchan, ntimes=ntimes, per_row=per_row)
    g0 = numpy.ones((len(gtime), 12, nchan, 1, 1), dtype=complex)
    gain, gwt, res, it = O.solve_gaintable_scalar(bv, model, g0, gtime, interval, phase_only=True, niter=200)
    assert numpy.max(res) < 3e-8
    assert numpy.max(numpy.abs(gain - 1.0)) > 0.1
    assert_allclose(gain[..., 0, 0], _expected(g), atol=1e-7)
    assert numpy.all(it <= 200)


def test_amplitude_and_phase_normalised():
    rng = numpy.random.default_rng(3)
    bv, model, g, gtime, interval = _setup(rng, nchan=2, phase_sd=1.0, amp_sd=0.1)
    g = g * 2.0
    bv['vis'] = O.apply_gaintable(model, g, numpy.ones(g.shape), numpy.zeros(1, dtype=int))
    g0 = numpy.ones((1, 12, 2, 1, 1), dtype=complex)
    gain, gwt, res, it = O.solve_gaintable_scalar(bv, model, g0, gtime, interval, phase_only=False, niter=200,
                                                  normalise_gains=True)
    assert numpy.max(res) < 3e-8
    assert numpy.max(numpy.abs(gain - 1.0)) > 0.1
    e = _expected(g)
    e = e / numpy.average(numpy.abs(e))
    assert_allclose(gain[..., 0, 0], e, atol=1e-7)


def test_point_source_form():
    """model None: vis already divided (ARL's test_solve_gaintable_scalar_pointsource)."""
    rng = numpy.random.default_rng(4)
    bv, model, g, gtime, interval = _setup(rng, nchan=2)
    bv['vis'] = O.apply_gaintable(numpy.ones_like(model), g, numpy.ones(g.shape), numpy.zeros(1, dtype=int))
    g0 = numpy.ones((1, 12, 2, 1, 1), dtype=complex)
    gain, _, res, _ = O.solve_gaintable_scalar(bv, None, g0, gtime, interval, phase_only=True, niter=200)
    assert numpy.max(res) < 3e-8
    assert_allclose(gain[..., 0, 0], _expected(g), atol=1e-7)


def test_dead_antenna():
    rng = numpy.random.default_rng(5)
    bv, model, g, gtime, interval = _setup(rng, nchan=2)
    bv['weight'][:, 4, :, :, :] = 0.0
    bv['weight'][:, :, 4, :, :] = 0.0
    g0 = numpy.ones((1, 12, 2, 1, 1), dtype=complex)
    gain, gwt, res, _ = O.solve_gaintable_scalar(bv, model, g0, gtime, interval, phase_only=True, niter=200)
    # its substitution gives 0 each iteration, which damping halves towards: 0.5^k of the start
    assert numpy.all(numpy.abs(gain[0, 4]) < 1e-7) and numpy.all(gwt[0, 4] == 0)
    others = [a for a in range(12) if a != 4]
    assert_allclose(gain[0, others, :, 0, 0], _expected(g)[0, others], atol=1e-7)
