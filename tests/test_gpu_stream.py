"""Streaming one visibility set in chunks (StreamingImager: global slab centres and means from a
first pass, chunks gridded into resident grids with arlmi_grid_add, planes transformed once) gives
the images and visibilities of one plan over all rows, and the oracle's over all rows; the FFTs at
8192^2 (the streamed workload's size, where complex128 columns go through in-place transposes)
against numpy.fft with ARL's shifts."""

import numpy
import pytest

from oracle import arl_oracle as O
from helpers import random_rows, to_device_vis, device_image, image_errors, rel_err

pytestmark = pytest.mark.gpu

TOL = {64: 1e-10, 32: 1e-5}


def _split(d, k):
    n = len(d["uvw"])
    cuts = numpy.linspace(0, n, k + 1).astype(int)
    return [{key: val[cuts[i]:cuts[i + 1]] for key, val in d.items()} for i in range(k)]


@pytest.mark.parametrize("precision", [64, 32])
def test_stream_equals_one_plan(gpu, precision):
    import torch
    import arlmi
    rng = numpy.random.default_rng(61)
    n, cell, S, P = 1024, 2.5e-4, 8, 16
    d = random_rows(rng, 60000, n, cell, S, wmax=400.0)
    # chunks with different w ranges, so per-chunk slab rules and means would differ
    order = numpy.argsort(d["uvw"][:, 2] + rng.normal(0, 150.0, len(d["uvw"])))
    d = {k: v[order] for k, v in d.items()}
    vis = to_device_vis(d, precision)
    chunks = [to_device_vis(c, precision) for c in _split(d, 4)]
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=S)
    whole = arlmi.Imager(vis, im, gcfcf=(gcf, cf), mode="wstack", vis_slices=P, sparse=False)
    ref_img = whole.invert(vis)[0].cpu().numpy()
    ref_sw = whole.sumwt()
    st = arlmi.StreamingImager(im, gcfcf=(gcf, cf), mode="wstack", vis_slices=P, precision=precision,
                               keep_plans=(precision == 32))  # both plan-holding modes
    boxes, width, means, counts = st.scan(chunks)
    wb, ww = whole.plan.slab_boxes()
    assert numpy.array_equal(boxes, wb) and width == ww
    assert numpy.max(numpy.abs(means - whole.plan.plane_w())) < 1e-12 * numpy.max(numpy.abs(means))
    img, sw = st.invert(chunks)
    l2, inner, full = image_errors(img.cpu().numpy(), ref_img)
    # float32: each chunk rounds its tile sums to complex64 before adding them to the grid
    assert l2 < (1e-12 if precision == 64 else 1e-4) and inner < TOL[precision], (l2, inner, full)
    assert abs(sw[0, 0] - ref_sw[0, 0]) < 1e-10 * ref_sw[0, 0]
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, rng.integers(200, 824, 40), rng.integers(200, 824, 40)] = rng.uniform(0.5, 2.0, 40)
    mim = device_image(n, cell, precision, data=model)
    ref_pv = whole.predict(vis, mim.data).cpu().numpy()
    pv = numpy.concatenate([p.cpu().numpy() for p in st.predict(mim.data, chunks)])
    assert rel_err(pv, ref_pv) < TOL[precision] / 10
    # and the oracle over all rows (same slab rule: the oracle recomputes it from every row)
    if precision == 64:
        g = O.ImageGeom(n, cell)
        ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
        oimg, osw = O.invert_list_wstack(d, g, ogcf, ocf, cfg, P)
        l2, inner, full = image_errors(img.cpu().numpy(), oimg)
        assert l2 < 1e-10 and inner < 1e-10, (l2, inner, full)
        opv = O.predict_list_wstack(d, model, g, ogcf, ocf, cfg, P)
        assert rel_err(pv, opv) < 1e-10
    # normalize=False: the list workflows' sumwt-weighted form
    img_f, _ = st.invert(chunks, normalize=False)
    (wimg, _), = arlmi.invert_list_serial_workflow([vis], [im], context="wstack", vis_slices=P, normalize=False,
                                                   gcfcf=[(gcf, cf)])
    l2, inner, _ = image_errors(img_f.cpu().numpy(), wimg.numpy())
    assert inner < TOL[precision], inner
    st.release()


def test_grid_add_needs_dense_grids(gpu):
    import ctypes
    import arlmi
    from arlmi.plan import Plan, kernel_struct, ptr, stream_handle, grids_buffer
    import torch
    rng = numpy.random.default_rng(62)
    d = random_rows(rng, 2000, 256, 1e-3, 8)
    vis = to_device_vis(d, 32)
    im = device_image(256, 1e-3, 32)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=8)
    plan = Plan(vis, im, cf, arlmi._lib.MODE_2D, 1, sparse=True)
    grids = grids_buffer(plan, 32, gpu)
    sw = torch.zeros((1, 1), dtype=torch.float64, device=gpu)
    k, keep = kernel_struct(cf, 32)
    with pytest.raises(arlmi._lib.ArlmiError):
        arlmi._lib.check(arlmi._lib.lib().arlmi_grid_add(plan.handle, ptr(vis.vis), ptr(vis.data["imaging_weight"]),
                                                         0, ctypes.byref(k), ptr(grids), ptr(sw), 32,
                                                         stream_handle()))


@pytest.mark.parametrize("direction", [1, -1])
def test_fft_8192(gpu, direction):
    """One 8192^2 plane through arlmi.fft / ifft in both precisions against numpy.fft with ARL's
    shifts (relative L2, and max error against the plane's peak)."""
    import torch
    import arlmi
    rng = numpy.random.default_rng(63 + direction)
    n = 8192
    a = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))).astype(numpy.complex64)
    f = numpy.fft.fft2 if direction > 0 else numpy.fft.ifft2
    ref = numpy.fft.fftshift(f(numpy.fft.ifftshift(a.astype(numpy.complex128))))
    for precision in (32, 64):
        cdt = torch.complex64 if precision == 32 else torch.complex128
        t = torch.as_tensor(a).to(device=gpu, dtype=cdt).reshape(1, 1, n, n).contiguous()
        out = (arlmi.fft(t) if direction > 0 else arlmi.ifft(t)).cpu().numpy()[0, 0]
        del t
        l2 = numpy.linalg.norm(out - ref) / numpy.linalg.norm(ref)
        assert l2 < (1e-6 if precision == 32 else 1e-14), (precision, l2)
        assert rel_err(out, ref) < (1e-5 if precision == 32 else 1e-13), precision


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("n", [1024, 4096])
def test_stream_band_masks_exact(gpu, precision, n):
    """The transforms of a streamed pass skip the column bands no chunk reaches (their grid columns
    are zeros): images and predicted visibilities are bit for bit those of dense transforms, at a
    size with one band per tile column (1024^2, T = 64) and at one with 128-column bands over 64
    tile columns (4096^2; in float64 the columns there take the two-level transform).  Gridding in
    fixed point, so that the two passes' grids are the same bits."""
    import torch
    import arlmi
    rng = numpy.random.default_rng(67 + n)
    cell, S, P = 2.5e-4 * 1024 / n, 8, 6
    d = random_rows(rng, 40000, n, cell, S, wmax=300.0)
    # most rows in the inner third of the uv plane, a few in one outer corner: outer bands empty
    d["uvw"][:, 0] *= numpy.where(numpy.arange(40000) % 50 == 0, 1.0, 1.0 / 3.0)
    d["uvw"][:, 1] *= numpy.where(numpy.arange(40000) % 50 == 0, 1.0, 1.0 / 3.0)
    sel = numpy.arange(40000) % 50 == 0
    d["uvw"][sel, 0] = numpy.abs(d["uvw"][sel, 0])
    chunks = [to_device_vis(c, precision) for c in _split(d, 3)]
    im = device_image(n, cell, precision)
    gcf, cf = arlmi.create_pswf_convolutionfunction(im, oversampling=8, support=S)
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, rng.integers(n // 4, 3 * n // 4, 30), rng.integers(n // 4, 3 * n // 4, 30)] = 1.0
    mim = device_image(n, cell, precision, data=model)
    res = []
    for masks in (True, False):
        st = arlmi.StreamingImager(im, gcfcf=(gcf, cf), mode="wstack", vis_slices=P, precision=precision,
                                   keep_plans=True, band_masks=masks, deterministic=True)
        st.scan(chunks)
        img, sw = st.invert(chunks)
        pv = [p.clone() for p in st.predict(mim.data, chunks)]
        if masks:
            bands = st._bands.cpu().numpy().view(numpy.uint32)
            nb = min(32, n // 64)
            full = (1 << nb) - 1
            assert numpy.any(bands != 0) and numpy.any(bands != full), bands  # some bands skipped
        res.append((img.clone(), sw, pv))
        st.release()
    (a, swa, pa), (b, swb, pb) = res
    assert torch.equal(a, b) and numpy.array_equal(swa, swb)
    assert all(torch.equal(x, y) for x, y in zip(pa, pb))
