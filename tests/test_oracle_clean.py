"""The oracle's Hogbom CLEAN pinned without retyping it (CPU only).

- overlapIndices: ARL's known answer (test_image_deconvolution.py:59-68) and its equal-area check.
- A delta PSF: the peak decays geometrically, res = A (1 - g)^k, and the loop stops at the first k
  with (1 - g)^k < 0.9 * fracthresh (a closed-form iteration count).
- Flux conservation: dirty - res equals the components convolved with the PSF, computed by an FFT
  convolution (scipy), for even and odd PSF sizes and PSFs smaller and larger than the image.
- Windows restrict the peaks; equal peaks go to the first index in row-major order.
"""

import numpy
import pytest
from numpy.testing import assert_allclose
from scipy.signal import fftconvolve

from oracle import arl_oracle as O


def test_overlap_known_answer():
    res = numpy.zeros([512, 512])
    psf = numpy.zeros([100, 100])
    s1, s2 = O.overlapIndices(res, psf, 499, 249)
    assert s1 == (449, 512, 199, 299)
    assert s2 == (0, 63, 0, 100)
    a1, a2 = numpy.zeros_like(res), numpy.zeros_like(psf)
    a1[s1[0]:s1[1], s1[2]:s1[3]] = 1
    a2[s2[0]:s2[1], s2[2]:s2[3]] = 1
    assert numpy.sum(a1) == numpy.sum(a2)


@pytest.mark.parametrize("gain,frac,expect", [(0.1, 0.1, 23), (0.5, 0.01, 7), (0.7, 0.2, 2)])
def test_delta_psf_closed_form(gain, frac, expect):
    """(1-g)^k < 0.9 frac first at k = expect: 0.9^23 = 0.0886 < 0.09 <= 0.9^22; 0.5^7 = 0.0078 <
    0.009 <= 0.5^6; 0.3^2 = 0.09 < 0.18 <= 0.3."""
    n, A, pmax = 64, 10.0, 2.0
    dirty = numpy.zeros((n, n))
    dirty[20, 37] = A
    psf = numpy.zeros((16, 16))
    psf[8, 8] = pmax
    comps, res, it = O.hogbom(dirty, psf, None, gain, 0.0, 1000, frac)
    assert it == expect
    assert_allclose(res[20, 37], A * (1 - gain) ** expect, rtol=1e-12)
    assert_allclose(comps[20, 37], A * (1 - (1 - gain) ** expect) / pmax, rtol=1e-12)
    assert numpy.count_nonzero(comps) == 1 and numpy.count_nonzero(res) == 1
    # an absolute threshold above the fractional one takes over
    _, _, it2 = O.hogbom(dirty, psf, None, gain, 0.5 * A, 1000, frac)
    k = 1
    while not (A * (1 - gain) ** k < 0.9 * 0.5 * A):
        k += 1
    assert it2 == k
    # niter caps the loop
    _, _, it3 = O.hogbom(dirty, psf, None, gain, 0.0, 3, frac)
    assert it3 == min(3, expect)


def _conserved(dirty, psf, comps, res):
    """dirty - res == sum over component pixels m of comps[m] * psf[h + x - m] for x - m in [-h, h)
    (h = psf // 2 per axis): the 'full' convolution with the PSF's first 2h rows and columns,
    cropped at offset h."""
    hx, hy = psf.shape[0] // 2, psf.shape[1] // 2
    k = psf[:2 * hx, :2 * hy]
    full = fftconvolve(comps, k, mode="full")
    model = full[hx:hx + dirty.shape[0], hy:hy + dirty.shape[1]]
    return numpy.max(numpy.abs((dirty - res) - model)) / numpy.max(numpy.abs(dirty))


def _sky(rng, n, psf, nsrc=12):
    comps = numpy.zeros((n, n))
    comps[rng.integers(n // 8, 7 * n // 8, nsrc), rng.integers(n // 8, 7 * n // 8, nsrc)] = rng.uniform(0.5, 3, nsrc)
    hx, hy = psf.shape[0] // 2, psf.shape[1] // 2
    return fftconvolve(comps, psf[:2 * hx, :2 * hy], mode="full")[hx:hx + n, hy:hy + n]


def _psf(n, sigma=2.5, odd=False):
    m = n + (1 if odd else 0)
    y, x = numpy.mgrid[0:m, 0:m] - m // 2
    p = numpy.exp(-(x ** 2 + y ** 2) / (2 * sigma ** 2)) * (1 + 0.15 * numpy.cos(0.7 * x) * numpy.cos(0.5 * y))
    return p


@pytest.mark.parametrize("psf_n,odd", [(64, False), (33, True), (128, False), (200, False)])
def test_flux_conservation(psf_n, odd):
    rng = numpy.random.default_rng(psf_n)
    n = 96
    psf = _psf(psf_n - (1 if odd else 0), odd=odd)
    dirty = _sky(rng, n, psf) + rng.normal(0, 1e-3, (n, n))
    comps, res, it = O.hogbom(dirty, psf, None, 0.2, 0.0, 400, 0.05)
    assert it > 10
    assert _conserved(dirty, psf, comps, res) < 1e-12
    assert numpy.max(numpy.abs(res)) < numpy.max(numpy.abs(dirty))


def test_window_restricts_peaks_and_ties_go_first():
    rng = numpy.random.default_rng(5)
    n = 64
    psf = _psf(32)
    dirty = _sky(rng, n, psf, nsrc=8)
    window = numpy.zeros((n, n))
    window[20:40, 10:50] = 1.0
    comps, res, it = O.hogbom(dirty, psf, window, 0.1, 0.0, 200, 0.01)
    assert numpy.all(comps[window == 0] == 0) and numpy.any(comps != 0)
    assert _conserved(dirty, psf, comps, res) < 1e-12
    # two equal peaks: the first in row-major order is taken first
    d = numpy.zeros((n, n))
    d[30, 40] = d[10, 50] = d[30, 5] = 4.0
    p = numpy.zeros((8, 8))
    p[4, 4] = 1.0
    c, _, _ = O.hogbom(d, p, None, 1.0, 0.0, 1, 0.5)
    assert c[10, 50] == 4.0 and numpy.count_nonzero(c) == 1


def test_deconvolve_cube_windows_and_skips():
    rng = numpy.random.default_rng(6)
    n = 64
    psf1 = _psf(n)
    dirty = numpy.stack([_sky(rng, n, psf1), _sky(rng, n, psf1)])[:, None]
    psf = numpy.stack([psf1, numpy.zeros_like(psf1)])[:, None]
    comp, residual, pcut = O.deconvolve_cube_hogbom(dirty, psf, window_shape='quarter', psf_support=16, gain=0.1,
                                                    niter=500, fractional_threshold=0.01)
    assert pcut.shape[-2:] == (32, 32)
    q = n // 4
    inside = numpy.zeros((n, n), bool)
    inside[(q + 1):3 * q, (q + 1):3 * q] = True
    assert numpy.all(comp[0, 0][~inside] == 0) and numpy.any(comp[0, 0] != 0)
    assert numpy.all(comp[1] == 0) and numpy.all(residual[1] == 0)  # zero PSF: plane skipped
