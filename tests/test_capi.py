"""The C-ABI library loads and exports every entry point include/arlmi.h declares, and the
host-side geometry equals the oracle's WCS arithmetic (no GPU calls)."""

import ctypes
import os
import re

import numpy
import pytest

from oracle import arl_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header="arlmi.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(arlmi_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from arlmi import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(_lib.EXPORTS) <= set(names)
    _lib.lib()  # every signature binds
    assert lib.arlmi_version() >= 100
    assert not hasattr(lib, "arlmi_probe_lds_add")  # measurement tooling lives in its own library


def test_probe_library_exports_every_declared_symbol():
    from arlmi import _lib
    lib = ctypes.CDLL(_lib.PROBE_LIB_PATH)
    names = _declared("arlmi_probe.h")
    assert names == ["arlmi_probe_last_error", "arlmi_probe_lds_add"]
    for name in names:
        assert hasattr(lib, name), name
    _lib.probe_lib()


def test_error_reporting_without_gpu():
    """Argument errors are reported through the return code and arlmi_last_error."""
    from arlmi import _lib
    L = _lib.lib()
    g = _lib.Geom()
    g.npixel = 100  # not a power of two
    h = ctypes.c_void_p()
    rc = L.arlmi_plan_create(ctypes.byref(g), None, None, 0, None, 0.0, 0, None, ctypes.byref(h))
    assert rc == -8
    assert b"power of two" in L.arlmi_last_error()
    g.npixel = 256
    g.support = 7
    g.oversampling = 8
    rc = L.arlmi_plan_create(ctypes.byref(g), None, None, 0, None, 0.0, 0, None, ctypes.byref(h))
    assert rc == -5
    with pytest.raises(_lib.ArlmiError):
        _lib.check(rc)


def test_geometry_matches_oracle_wcs():
    """UU/VV/DUU/DVV cdelt and crpix handed to the library equal ARL's WCS arithmetic bit for bit."""
    import torch
    import arlmi
    from arlmi.plan import make_geom
    n, cell = 2048, 1.5e-4
    im = arlmi.create_image(npixel=n, cellsize=cell, frequency=[1e8], channel_bandwidth=[1e9],
                            dev=torch.device("cpu"))
    gcf_im = arlmi.create_image(npixel=8, cellsize=cell, dev=torch.device("cpu"))
    cf = arlmi.create_convolutionfunction_from_image(im, oversampling=8, support=8)
    g = make_geom(im, cf, 1, 64)
    og = O.ImageGeom(n, cell, frequency=1e8, channel_bandwidth=1e9)
    cfg = O.CFGeom(og, 8, 8)
    assert g.du == og.du() and g.dv == og.dv()
    assert g.ddu == cfg.ddu and g.ddv == cfg.ddv
    assert g.crpix_uv == og.crpix0 and g.crpix_d == cfg.crpix_d
    assert g.cellsize == og.wcs_cellsize()
    assert g.freq0 == og.freq0 and g.dfreq == og.dfreq


def test_pswf_tables_match_oracle():
    """The separable table reproduces the oracle's CF and gcf (host float64, no GPU)."""
    import torch
    import arlmi
    for os_, S in [(8, 6), (8, 8), (1, 6), (128, 6)]:
        g = O.ImageGeom(256, 5e-4)
        ogcf, ocf, _ = O.create_pswf_convolutionfunction(g, os_, S)
        k = arlmi.kernels
        n = numpy.arange(S)
        s = numpy.arange(os_)
        nu = ((n[None, :] - S // 2) - (s[:, None] - os_ // 2) / os_) / (S // 2)
        kern = k.grdsf(nu)[1]
        kern = kern / numpy.sum(kern[os_ // 2])
        full = kern[:, None, :, None] * kern[None, :, None, :]
        full = full / numpy.sum(full[0, 0])
        assert numpy.max(numpy.abs(full - ocf[0, 0, 0])) < 1e-15
        sep = kern / numpy.sqrt(numpy.sum(kern[0]) ** 2)
        assert numpy.max(numpy.abs(sep[:, None, :, None] * sep[None, :, None, :] - ocf[0, 0, 0])) < 1e-15
        assert numpy.max(numpy.abs(k._pswf_gcf(256) - ogcf[0, 0])) == 0.0


def test_wslice_boxes_match_oracle():
    from arlmi.plan import wslice_boxes
    for wmax, p in [(1234.5, 64), (10.0, 1), (3.0, 2)]:
        boxes, width = wslice_boxes(wmax, p)
        ob = numpy.linspace(-wmax, wmax, p)
        assert numpy.array_equal(boxes, ob)
        assert width == (ob[1] - ob[0] if p > 1 else 2 * wmax)
