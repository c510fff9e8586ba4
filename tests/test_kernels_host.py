"""Host-side kernel functions of the package (no GPU): ARL's bounding-box clipping of w-projection
kernels against ARL's known answer and the oracle, and the conversion of another library's
(gcf, cf) arrays into the package's data models."""

import numpy
import pytest
import torch

from oracle import arl_oracle as O

CPU = torch.device("cpu")


def _image(n=512, cell=5e-4):
    import arlmi
    return arlmi.create_image(npixel=n, cellsize=cell, frequency=[1e8], channel_bandwidth=[1e6], dev=CPU)


@pytest.fixture(scope="module")
def wkernel():
    """test_griddata_kernels.py:122-139 (nw=201, wstep=8, os=8, S=60 on 512^2 at 5e-4): three of
    the 201 planes, enough for the clipping: the box is set by the widest planes (the end ones,
    |w| = 800) and the peak by the centre plane."""
    g = O.ImageGeom(512, 5e-4)
    _, cf, _ = O.create_awterm_convolutionfunction(g, nw=201, wstep=8.0, oversampling=8, support=60, use_aaf=True,
                                                   planes=[0, 100, 200])
    return cf


def test_apply_bounding_box_known_answer(wkernel):
    import arlmi
    c = arlmi.create_convolutionfunction_from_image(_image(), nz=201, zstep=8.0, oversampling=8, support=60)
    c.data = wkernel
    clipped = arlmi.apply_bounding_box_convolutionfunction(c, fractional_level=1e-3)
    d = numpy.asarray(clipped.data)
    pk = numpy.unravel_index(numpy.argmax(numpy.abs(d)), d.shape)
    assert pk == (0, 0, 100, 4, 4, 27, 27)                                     # :130
    assert abs(d[pk] - (0.18704481681878257 - 0j)) < 1e-7                      # :131-132
    assert numpy.array_equal(d, O.apply_bounding_box_convolutionfunction(wkernel, 1e-3))
    # crpix moves with the clipped size (convolution_functions.py:182-183): 31 + 54/2 - 60/2
    assert list(clipped.grid_wcs.wcs.crpix[:2]) == [28.0, 28.0]
    assert list(c.grid_wcs.wcs.crpix[:2]) == [31.0, 31.0]                      # the input is not changed


def test_calculate_bounding_box_matches_oracle(wkernel):
    import arlmi
    c = arlmi.create_convolutionfunction_from_image(_image(), nz=201, zstep=8.0, oversampling=8, support=60)
    planes = [0, 100, 200]
    part = wkernel[:, :, planes]  # the computed planes (an all-zero plane has no box in ARL either)
    c.data = part
    for level in (1e-4, 1e-3, 1e-2):
        got = arlmi.calculate_bounding_box_convolutionfunction(c, level)
        want = O.calculate_bounding_box_convolutionfunction(part, level)
        assert [(z, tuple(map(int, a)), tuple(map(int, b))) for z, a, b in want] == got
        # the end planes (|w| = 800) are at least as wide as the centre plane
        assert got[0][1][1] - got[0][1][0] >= got[1][1][1] - got[1][1][0]


def test_convolutionfunction_from_arrays_detects_separable():
    """ARL's PSWF table (here the oracle's) factors exactly: it is marked separable, its factor
    rebuilds every tap, and it equals the package's own PSWF factor up to sign."""
    import arlmi
    im = _image(256, 1e-3)
    g = O.ImageGeom(256, 1e-3)
    ogcf, ocf, _ = O.create_pswf_convolutionfunction(g, 8, 6)
    gcf, cf = arlmi.convolutionfunction_from_arrays(im, ogcf, ocf)
    assert cf.separable is not None
    k = cf.separable
    rebuilt = k[:, None, :, None] * k[None, :, None, :]
    assert numpy.max(numpy.abs(rebuilt - ocf[0, 0, 0].real)) < 1e-15
    assert numpy.array_equal(gcf.data.numpy()[0, 0], ogcf[0, 0])
    assert cf.oversampling == 8 and cf.support == 6 and cf.nz == 1


def test_convolutionfunction_from_arrays_keeps_tables():
    """A complex (w-projection) kernel, or a real one that does not factor, stays a full table with
    the kernel WCS given."""
    import arlmi
    im = _image(256, 1e-3)
    g = O.ImageGeom(256, 1e-3)
    ogcf, ocf, _ = O.create_awterm_convolutionfunction(g, nw=3, wstep=10.0, oversampling=4, support=8, use_aaf=True)
    crpix = [5.0, 5.0, 3.0, 3.0, 2.0, 1.0, 1.0]
    cdelt = [1.0, 1.0, 0.25, 0.25, 10.0, 1.0, 1e6]
    gcf, cf = arlmi.convolutionfunction_from_arrays(im, ogcf, ocf, crpix, cdelt)
    assert cf.separable is None and cf.nz == 3
    assert cf.grid_wcs.wcs.cdelt[4] == 10.0 and cf.grid_wcs.wcs.crpix[4] == 2.0
    assert numpy.array_equal(cf.data, ocf)
    rng = numpy.random.default_rng(1)
    _, pcf, _ = O.create_pswf_convolutionfunction(g, 4, 6)
    noisy = pcf + 1e-6 * rng.normal(size=pcf.shape)
    assert arlmi.convolutionfunction_from_arrays(im, ogcf, noisy)[1].separable is None


def test_kernel_channels_must_agree():
    """The gridders read one kernel table for every image channel; ARL indexes the table by the
    row's image channel (gridding.py:123-130), so a table whose channels differ is refused rather
    than read as channel 0, and one whose channels agree is accepted."""
    import arlmi
    rng = numpy.random.default_rng(5)
    im = arlmi.create_image(npixel=64, cellsize=1e-3, frequency=[1e8, 1.1e8], channel_bandwidth=[1e7, 1e7], dev=CPU)
    one = rng.normal(size=(1, 1, 1, 4, 4, 6, 6)) + 1j * rng.normal(size=(1, 1, 1, 4, 4, 6, 6))
    same = numpy.repeat(one, 2, axis=0)
    gcf = numpy.ones((2, 1, 64, 64))
    _, cf = arlmi.convolutionfunction_from_arrays(im, gcf, same)
    cf.check_uniform()
    diff = same.copy()
    diff[1, 0, 0, 2, 1, 3, 3] += 1e-9
    _, cf2 = arlmi.convolutionfunction_from_arrays(im, gcf, diff)
    with pytest.raises(NotImplementedError):
        cf2.check_uniform()
