"""Pins the oracle's BlockVisibility conversion and gain convention: the element loop, an
independent row construction, ARL's own calibration assertions and finite differences."""

import numpy
import pytest

from oracle import arl_oracle as O


def _array(rng, nants, radius=2000.0):
    r = radius * numpy.sqrt(rng.uniform(0, 1, nants))
    th = rng.uniform(0, 2 * numpy.pi, nants)
    enu = numpy.stack([r * numpy.cos(th), r * numpy.sin(th), rng.normal(0, 5, nants)], axis=1)
    return O.xyz_at_latitude(enu, -26.7 * numpy.pi / 180.0)


def small_block(seed=0, nants=7, ntimes=3, nchan=3):
    rng = numpy.random.default_rng(seed)
    xyz = _array(rng, nants)
    ha = numpy.linspace(-0.3, 0.3, ntimes)
    freq = numpy.linspace(1.0e8, 1.5e8, nchan)
    b = O.create_blockvisibility(xyz, ha, freq, numpy.full(nchan, 1e6), -0.5)
    shape = b['vis'].shape
    b['vis'] = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    b['weight'] = rng.uniform(0.5, 2.0, shape)
    b['imaging_weight'] = rng.uniform(0.5, 2.0, shape)
    return b, xyz, ha, freq


def test_convert_blocks_equals_element_loop():
    b, *_ = small_block()
    v = O.convert_blocks(b)
    l = O.convert_blocks_loop(b)
    for k in l:
        assert numpy.array_equal(v[k].reshape(l[k].shape), l[k]), k


def test_convert_blocks_equals_direct_row_construction():
    """create_visibility (visibility/base.py:50-150) builds the same rows from the antenna
    positions directly (a1-major order); as a set they must equal the converted block rows, bit
    for bit, with antenna1 < antenna2 and uvw = pos[antenna2] - pos[antenna1]."""
    b, xyz, ha, freq = small_block(seed=1)
    v = O.convert_blocks(b)
    d = O.create_visibility(xyz, ha, freq, numpy.full(len(freq), 1e6), -0.5)
    key_v = numpy.lexsort((v['frequency'], v['antenna1'], v['antenna2'], v['time']))
    key_d = numpy.lexsort((d['frequency'], d['antenna1'], d['antenna2'], d['time']))
    assert numpy.array_equal(v['uvw'][key_v], d['uvw'][key_d])
    assert numpy.array_equal(v['antenna1'][key_v], d['antenna1'][key_d])
    assert numpy.array_equal(v['antenna2'][key_v], d['antenna2'][key_d])
    assert numpy.all(v['antenna1'] < v['antenna2'])


def test_convert_row_order_and_cindex_round_trip():
    b, *_ = small_block(seed=2, nants=5, ntimes=2, nchan=2)
    v = O.convert_blocks(b)
    nants, nchan = 5, 2
    # row r = (t * nbl + a2 (a2 - 1) / 2 + a1) * nchan + chan
    nbl = nants * (nants - 1) // 2
    r = numpy.arange(len(v['antenna1']))
    bl = (r // nchan) % nbl
    assert numpy.array_equal(bl, v['antenna2'] * (v['antenna2'] - 1) // 2 + v['antenna1'])
    assert list(zip(v['antenna2'][:nchan * 4:nchan], v['antenna1'][:nchan * 4:nchan])) == [(1, 0), (2, 0), (2, 1),
                                                                                            (3, 0)]
    d = O.decoalesce_visibility(v, b)
    lower = numpy.tril(numpy.ones((nants, nants), dtype=bool), k=-1)
    m = numpy.broadcast_to(lower[None, :, :, None, None], b['vis'].shape)
    assert numpy.array_equal(d['vis'][m], b['vis'][m])
    # outside the lower triangle cindex is 0, so decoalescing fills in row 0
    assert numpy.all(d['vis'][~m] == v['vis'][0, 0])


def _gains(rng, ntg, nants, nchan, amp=0.1, phase=0.1):
    return ((1.0 + amp * rng.normal(size=(ntg, nants, nchan))) *
            numpy.exp(1j * phase * rng.normal(size=(ntg, nants, nchan))))


def test_apply_gaintable_arl_assertions():
    """test_calibration_operations.py:92-135 on the scalar path."""
    b, *_ = small_block(seed=3)
    rng = numpy.random.default_rng(3)
    ntimes, nants, _, nchan, _ = b['vis'].shape
    b['vis'] *= 100.0
    g = _gains(rng, ntimes, nants, nchan, amp=0.01, phase=0.1)
    wt = numpy.ones(g.shape)
    slots = numpy.arange(ntimes)
    applied = O.apply_gaintable(b['vis'], g, wt, slots)
    assert numpy.max(numpy.abs(applied - b['vis'])) > 10.0                     # :101
    g2 = _gains(rng, ntimes, nants, nchan, amp=0.1, phase=0.1)
    back = O.apply_gaintable(O.apply_gaintable(b['vis'], g2, wt, slots), g2, wt, slots, inverse=True)
    assert numpy.max(numpy.abs(back - b['vis'])) < 1e-12                       # :125
    zero = O.apply_gaintable(b['vis'], 0.0 * g2, wt, slots, inverse=True)
    assert numpy.max(numpy.abs(zero[:, 0, 1] - b['vis'][:, 0, 1])) < 1e-12     # :135
    # a zero gain weight zeroes every product that involves the antenna
    wt0 = wt.copy()
    wt0[:, 2, :] = 0.0
    z = O.apply_gaintable(b['vis'], g, wt0, slots)
    assert numpy.all(z[:, 2, :] == 0) and numpy.all(z[:, :, 2] == 0)
    assert numpy.all(z[:, 1, 0] != 0)


def test_gain_convention_block_equals_rows():
    """Applying gains to the block then converting equals converting then applying
    V g[antenna2] conj(g[antenna1]) to the rows."""
    b, *_ = small_block(seed=4)
    rng = numpy.random.default_rng(4)
    ntimes, nants, _, nchan, _ = b['vis'].shape
    g = _gains(rng, ntimes, nants, nchan)
    applied = O.apply_gaintable(b['vis'], g, numpy.ones(g.shape), numpy.arange(ntimes))
    bb = dict(b)
    bb['vis'] = applied
    rows_after = O.convert_blocks(bb)['vis']
    v = O.convert_blocks(b)
    t = numpy.searchsorted(b['time'], v['time'])
    ch = numpy.searchsorted(b['frequency'], v['frequency'])
    flat = g.reshape(-1)
    i1 = (t * nants + v['antenna1']) * nchan + ch
    i2 = (t * nants + v['antenna2']) * nchan + ch
    rows_before = O.apply_gains_rows(v['vis'], i1, i2, flat)
    assert numpy.max(numpy.abs(rows_before - rows_after)) < 1e-13
    # the un-conjugated gain is antenna2's: swapping the roles breaks the agreement
    swapped = O.apply_gains_rows(v['vis'], i2, i1, flat)
    assert numpy.max(numpy.abs(swapped - rows_after)) > 1e-3


def test_gain_adjoint_finite_differences():
    rng = numpy.random.default_rng(5)
    n, ng = 200, 12
    v = rng.normal(size=n) + 1j * rng.normal(size=n)
    obs = rng.normal(size=n) + 1j * rng.normal(size=n)
    i1 = rng.integers(0, ng, n)
    i2 = rng.integers(0, ng, n)
    g = _gains(rng, 1, ng, 1, amp=0.2, phase=0.5).reshape(-1)

    def loss(vv, gg):
        d = obs - vv * gg[i2] * numpy.conjugate(gg[i1])
        return numpy.sum(numpy.abs(d) ** 2)

    out = v * g[i2] * numpy.conjugate(g[i1])
    G = 2.0 * (out - obs)  # dL/dRe(out) + i dL/dIm(out) of |obs - out|^2
    gv, gg = O.apply_gains_rows_adjoint(v, G, i1, i2, g)
    h = 1e-6
    for k in (0, 5, 11):
        for unit in (1.0, 1j):
            gp, gm = g.copy(), g.copy()
            gp[k] += h * unit
            gm[k] -= h * unit
            fd = (loss(v, gp) - loss(v, gm)) / (2 * h)
            an = gg[k].real if unit == 1.0 else gg[k].imag
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (k, unit, fd, an)
    for r in (0, 77, 199):
        for unit in (1.0, 1j):
            vp, vm = v.copy(), v.copy()
            vp[r] += h * unit
            vm[r] -= h * unit
            fd = (loss(vp, g) - loss(vm, g)) / (2 * h)
            an = gv[r].real if unit == 1.0 else gv[r].imag
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
    # dot-product identity of the vis part: Re<A x, y> = Re<x, A^H y>
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    y = rng.normal(size=n) + 1j * rng.normal(size=n)
    ax = x * g[i2] * numpy.conjugate(g[i1])
    aty, _ = O.apply_gains_rows_adjoint(v, y, i1, i2, g)
    assert abs(numpy.vdot(ax, y).real - numpy.vdot(x, aty).real) < 1e-10 * n


# --- weighting (gridding.py:161-230) --------------------------------------------------------

def _weight_rows(seed, n=64, nvis=3000, cell=0.01):
    rng = numpy.random.default_rng(seed)
    du = 1.0 / (n * cell)
    lim = (n // 2 - 1) * du
    uvw = numpy.zeros((nvis, 3))
    # few distinct pixels so that many rows share one
    uvw[:, 0] = rng.integers(-8, 8, nvis) * du + rng.uniform(-0.3, 0.3, nvis) * du
    uvw[:, 1] = rng.integers(-8, 8, nvis) * du + rng.uniform(-0.3, 0.3, nvis) * du
    return {'uvw': uvw, 'frequency': numpy.full(nvis, 1e8),
            'imaging_weight': rng.integers(1, 5, (nvis, 1)).astype(float),
            'vis': numpy.zeros((nvis, 1), dtype=complex)}


def test_grid_weight_counts_per_pixel():
    """Integer weights: the density equals a direct count per nearest pixel, sumwt the total."""
    d = _weight_rows(1)
    g = O.ImageGeom(64, 0.01)
    _, _, cfg = O.create_pswf_convolutionfunction(g, 8, 8)
    grid, sumwt = O.grid_weight_to_griddata(d, g, cfg)
    x = numpy.round(64 // 2 + d["uvw"][:, 0] / g.du()).astype(int)
    y = numpy.round(64 // 2 + d["uvw"][:, 1] / g.dv()).astype(int)
    direct = numpy.zeros((64, 64))
    numpy.add.at(direct, (y, x), d['imaging_weight'][:, 0])
    assert numpy.array_equal(grid[0, 0, 0].real, direct)
    assert numpy.all(grid.imag == 0)
    assert sumwt[0, 0] == d['imaging_weight'].sum()


def test_uniform_weighting_identity():
    """After uniform reweighting the weights of every occupied pixel sum to 1, and weights are
    only ever divided by their own pixel's density."""
    d = _weight_rows(2)
    g = O.ImageGeom(64, 0.01)
    _, _, cfg = O.create_pswf_convolutionfunction(g, 8, 8)
    new = O.weight_visibility(d, g, cfg)
    d2 = dict(d)
    d2['imaging_weight'] = new
    grid2, _ = O.grid_weight_to_griddata(d2, g, cfg)
    occ = grid2[0, 0, 0].real != 0
    assert numpy.allclose(grid2[0, 0, 0].real[occ], 1.0, rtol=0, atol=1e-14)
    assert occ.sum() <= 256
    # zero weights leave a pixel empty and are not divided (no division by zero)
    d3 = dict(d)
    w = d['imaging_weight'].copy()
    w[:] = 0.0
    d3['imaging_weight'] = w
    grid3, s3 = O.grid_weight_to_griddata(d3, g, cfg)
    assert numpy.all(O.griddata_reweight(d3, grid3, g, cfg) == 0.0) and s3[0, 0] == 0.0
    # merging two lists doubles the density
    gm, sm = O.griddata_merge_weights([O.grid_weight_to_griddata(d, g, cfg)] * 2)
    grid, sumwt = O.grid_weight_to_griddata(d, g, cfg)
    assert numpy.array_equal(gm, 2 * grid) and numpy.array_equal(sm, 2 * sumwt)
