"""arlmi_hogbom_noalloc (the minor cycle with its scratch in a fixed device array of the library,
for callers that cannot allocate) against arlmi_hogbom, through the C ABI, bit for bit: two calls
in a row on one device, the second on a smaller image, so the second runs on scratch the first
left behind; the smaller case also against the oracle."""

import ctypes

import numpy
import pytest

from oracle import arl_oracle as O
from test_gpu_clean import _psf, _sky

pytestmark = pytest.mark.gpu


def _call(name, dirty, psf, window, gain, thresh, niter, frac):
    import torch
    import arlmi
    from arlmi.plan import ptr, stream_handle
    res = torch.as_tensor(dirty).cuda().contiguous()
    comps = torch.empty_like(res)
    p = torch.as_tensor(psf).cuda().contiguous()
    w = None if window is None else torch.as_tensor(window).cuda().contiguous()
    it = ctypes.c_int32(-1)
    fn = getattr(arlmi._lib.lib(), name)
    arlmi._lib.check(fn(ptr(res), ptr(comps), ptr(p), p.shape[0], p.shape[1], ptr(w), res.shape[0], res.shape[1],
                        gain, thresh, niter, frac, ctypes.byref(it), stream_handle()))
    torch.cuda.synchronize()
    return comps.cpu().numpy(), res.cpu().numpy(), it.value


def test_hogbom_noalloc_matches_hogbom(gpu):
    rng = numpy.random.default_rng(77)
    cases = []
    for n, psf_n, win in ((512, 128, True), (200, 64, False)):
        psf = _psf(psf_n)
        dirty = _sky(rng, n, psf, nsrc=30)
        window = None
        if win:
            window = numpy.zeros((n, n))
            window[n // 6:5 * n // 6, n // 5:4 * n // 5] = 1.0
        cases.append((dirty, psf, window))
    for dirty, psf, window in cases:
        args = (dirty, psf, window, 0.1, 0.0, 400, 0.01)
        got = _call("arlmi_hogbom_noalloc", *args)
        want = _call("arlmi_hogbom", *args)
        assert got[2] == want[2] and got[2] > 0
        assert numpy.array_equal(got[0], want[0]) and numpy.array_equal(got[1], want[1])
    dirty, psf, window = cases[1]
    ref = O.hogbom(dirty, psf, window, 0.1, 0.0, 400, 0.01)
    assert numpy.array_equal(got[0], ref[0]) and numpy.array_equal(got[1], ref[1]) and got[2] == ref[2]


def test_hogbom_noalloc_rejects_bad_arguments(gpu):
    import arlmi
    rng = numpy.random.default_rng(5)
    psf = _psf(32)
    dirty = _sky(rng, 64, psf, nsrc=3)
    with pytest.raises(arlmi._lib.ArlmiError):
        _call("arlmi_hogbom_noalloc", dirty, psf, None, 2.5, 0.0, 10, 0.01)
