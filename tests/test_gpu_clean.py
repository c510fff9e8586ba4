"""Hogbom CLEAN on the GPU (arlmi_hogbom) against the oracle, bit for bit: both run ARL's float64
operations in ARL's order (numpy does not fuse them, and the kernel is compiled without
contraction), the peak search breaks ties to the first index as numpy's argmax, and the
iteration count is ARL's."""

import numpy
import pytest
from scipy.signal import fftconvolve

from oracle import arl_oracle as O

pytestmark = pytest.mark.gpu


def _psf(m, sigma=2.5):
    y, x = numpy.mgrid[0:m, 0:m] - m // 2
    return numpy.exp(-(x ** 2 + y ** 2) / (2 * sigma ** 2)) * (1 + 0.15 * numpy.cos(0.7 * x) * numpy.cos(0.5 * y))


def _sky(rng, n, psf, nsrc=20):
    comps = numpy.zeros((n, n))
    comps[rng.integers(n // 8, 7 * n // 8, nsrc), rng.integers(n // 8, 7 * n // 8, nsrc)] = rng.uniform(0.5, 3, nsrc)
    hx, hy = psf.shape[0] // 2, psf.shape[1] // 2
    return fftconvolve(comps, psf[:2 * hx, :2 * hy], mode="full")[hx:hx + n, hy:hy + n] + rng.normal(0, 1e-3, (n, n))


def _run(dirty, psf, window, gain, thresh, niter, frac):
    import torch
    import arlmi
    d = torch.as_tensor(dirty).cuda()
    p = torch.as_tensor(psf).cuda()
    w = None if window is None else torch.as_tensor(window).cuda()
    comps, res = arlmi.hogbom(d, p, w, gain, thresh, niter, frac)
    return comps.cpu().numpy(), res.cpu().numpy(), arlmi.hogbom.last_iterations


def _same(got, want):
    assert got[2] == want[2], (got[2], want[2])
    assert numpy.array_equal(got[0], want[0]), numpy.max(numpy.abs(got[0] - want[0]))
    assert numpy.array_equal(got[1], want[1]), numpy.max(numpy.abs(got[1] - want[1]))


@pytest.mark.parametrize("n,psf_n,window", [(128, 128, None), (256, 64, None), (200, 31, "box"), (96, 250, None),
                                            (256, 256, "box")])
def test_hogbom_bit_exact(gpu, n, psf_n, window):
    rng = numpy.random.default_rng(n + psf_n)
    psf = _psf(psf_n)
    dirty = _sky(rng, n, psf)
    win = None
    if window == "box":
        win = numpy.zeros((n, n))
        win[n // 5:4 * n // 5, n // 6:5 * n // 6] = 1.0
    want = O.hogbom(dirty, psf, win, 0.1, 0.0, 600, 0.01)
    _same(_run(dirty, psf, win, 0.1, 0.0, 600, 0.01), want)


def test_hogbom_stops_and_caps(gpu):
    rng = numpy.random.default_rng(3)
    psf = _psf(64)
    dirty = _sky(rng, 128, psf, nsrc=4)
    for gain, thresh, niter, frac in [(0.7, 0.0, 1000, 0.2), (0.1, 0.0, 5, 0.001), (1.5, 0.3, 1000, 0.01),
                                      (0.1, 0.0, 1, 0.5)]:
        want = O.hogbom(dirty, psf, None, gain, thresh, niter, frac)
        _same(_run(dirty, psf, None, gain, thresh, niter, frac), want)


def test_delta_psf_and_ties(gpu):
    d = numpy.zeros((64, 64))
    d[30, 40] = d[10, 50] = d[30, 5] = 4.0
    p = numpy.zeros((8, 8))
    p[4, 4] = 2.0
    _same(_run(d, p, None, 0.3, 0.0, 50, 0.05), O.hogbom(d, p, None, 0.3, 0.0, 50, 0.05))
    z = numpy.zeros((32, 32))  # an empty image: peak 0 at index 0 every time, until niter
    _same(_run(z, p, None, 0.3, 0.0, 7, 0.05), O.hogbom(z, p, None, 0.3, 0.0, 7, 0.05))


def test_hogbom_rejects_arl_asserts(gpu):
    import arlmi
    d = numpy.ones((16, 16))
    with pytest.raises(arlmi._lib.ArlmiError):
        _run(d, numpy.zeros((4, 4)), None, 0.1, 0.0, 10, 0.1)  # pmax = 0
    with pytest.raises(arlmi._lib.ArlmiError):
        _run(d, numpy.ones((4, 4)), None, 2.0, 0.0, 10, 0.1)


def test_deconvolve_cube_against_oracle(gpu):
    import torch
    import arlmi
    rng = numpy.random.default_rng(8)
    n = 128
    psf1 = _psf(n)
    dirty = numpy.stack([_sky(rng, n, psf1), _sky(rng, n, psf1), _sky(rng, n, psf1)])[:, None]
    psf = numpy.stack([psf1, numpy.zeros_like(psf1), 0.5 * psf1])[:, None]
    for kw in [dict(window_shape="quarter", psf_support=32), dict(window_shape="no_edge", window_edge=10),
               dict()]:
        args = dict(gain=0.1, niter=400, fractional_threshold=0.01, **kw)
        im = arlmi.Image(torch.as_tensor(dirty).cuda().float(), None, None)
        pim = arlmi.Image(torch.as_tensor(psf).cuda().float(), None, None)
        # float32 images are promoted to float64 first: give the oracle the same values
        rc, rr, _ = O.deconvolve_cube_hogbom(dirty.astype(numpy.float32).astype(numpy.float64),
                                             psf.astype(numpy.float32).astype(numpy.float64), **args)
        comp, res = arlmi.deconvolve_cube(im, pim, algorithm="hogbom", **args)
        assert comp.data.dtype == torch.float64
        assert numpy.array_equal(comp.data.cpu().numpy(), rc)
        assert numpy.array_equal(res.data.cpu().numpy(), rr)


def test_real_psf_512(gpu):
    """Dirty image and PSF of the 120,960-row 2-D imaging workload (512^2): 800 iterations."""
    import arlmi
    from arlmi.simulation import config1
    vis, im, _ = config1(precision=64)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=8)
    dirty, _ = arlmi.invert_2d(vis, im, gcfcf=(gcf, cf))
    psf, _ = arlmi.invert_2d(vis, im, dopsf=True, gcfcf=(gcf, cf))
    d = dirty.numpy()[0, 0]
    p = psf.numpy()[0, 0]
    want = O.hogbom(d, p, None, 0.1, 0.0, 800, 0.001)
    _same(_run(d, p, None, 0.1, 0.0, 800, 0.001), want)


def test_full_size_2048(gpu):
    """The w-stacked 2048^2 workload's float32 dirty image and PSF (promoted to float64), PSF cut
    to +-1024 (ARL's default psf_support): 60 iterations against the oracle."""
    import torch
    import arlmi
    from arlmi import simulation
    vis, im, _ = simulation.config2(precision=32)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=8)
    imager = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode="wstack", vis_slices=64)
    dirty, _ = imager.invert(vis)
    d = dirty.double().cpu().numpy()[0, 0]
    psf, _ = imager.invert(vis, dopsf=True)
    p = psf.double().cpu().numpy()[0, 0]
    del imager, vis
    torch.cuda.empty_cache()
    want = O.hogbom(d, p, None, 0.1, 0.0, 60, 0.001)
    _same(_run(d, p, None, 0.1, 0.0, 60, 0.001), want)
