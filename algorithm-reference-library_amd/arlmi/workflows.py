"""Imaging context registry and the serial list workflows
(workflows/shared/imaging/imaging_shared.py:21-145, workflows/serial/imaging/imaging_serial.py:41-216).

The w-stack workflow is where ARL loops over slabs (one copy, grid, FFT and screen per slab
per vis list); here each vis list is one plan with vis_slices planes, gridded, transformed
and screen-accumulated in batched GPU launches, with the same slab rule, slab-mean w shift,
per-slab weighting and normalisation.
"""

import numpy
import torch

from .imaging import (Imager, invert_2d, predict_2d, invert_wstack_single, predict_wstack_single,
                      normalize_sumwt)
from .data_models import Image

__all__ = ["imaging_contexts", "imaging_context", "invert_list_serial_workflow", "predict_list_serial_workflow",
           "sum_invert_results", "vis_null_iter", "vis_wslice_iter"]


def vis_null_iter(vis, vis_slices=1):
    """All rows (visibility/iterators.py vis_null_iter)."""
    yield torch.ones(vis.nvis, dtype=torch.bool, device=vis.uvw.device)


def vis_wslice_iter(vis, vis_slices=1):
    """Row masks |w - box| < step/2, boxes = linspace(-wmax, wmax, P) (iterators.py:91-109)."""
    w = vis.w
    wmaxabs = float(torch.max(torch.abs(w)).item())
    boxes = numpy.linspace(-wmaxabs, +wmaxabs, vis_slices)
    wstack = boxes[1] - boxes[0] if vis_slices > 1 else 2 * wmaxabs
    for box in boxes:
        yield torch.abs(w - box) < 0.5 * wstack


def imaging_contexts():
    """Context -> {predict, invert, vis_iterator} (imaging_shared.py:21-88), for the contexts of
    the Fourier hot path.  'wprojection' is '2d' with an AW kernel passed as gcfcf."""
    return {
        "2d": {"predict": predict_2d, "invert": invert_2d, "vis_iterator": vis_null_iter},
        "wprojection": {"predict": predict_2d, "invert": invert_2d, "vis_iterator": vis_null_iter},
        "wstack": {"predict": predict_wstack_single, "invert": invert_wstack_single,
                   "vis_iterator": vis_wslice_iter},
    }


def imaging_context(context="2d"):
    contexts = imaging_contexts()
    assert context in contexts, context
    return contexts[context]


def sum_invert_results(image_list, normalize=True):
    """Sumwt-weighted sum of (image, sumwt) pairs (imaging_shared.py:125-145)."""
    if len(image_list) == 1:
        return image_list[0]
    first = image_list[0][0]
    acc = torch.zeros_like(first.data, dtype=torch.float64)
    sumwt = numpy.zeros_like(image_list[0][1])
    for img, sw in image_list:
        acc += torch.as_tensor(sw, device=acc.device)[..., None, None] * img.data.to(torch.float64)
        sumwt = sumwt + sw
    im = Image(acc.to(first.data.dtype), first.wcs.deepcopy(), first.polarisation_frame)
    if normalize:
        im = normalize_sumwt(im, sumwt)
    return im, sumwt


def _mode(context):
    if context in ("2d", "wprojection"):
        return "2d"
    if context == "wstack":
        return "wstack"
    raise ValueError("context %r is outside the Fourier hot path (2d, wprojection, wstack)" % context)


def invert_list_serial_workflow(vis_list, template_model_imagelist, dopsf=False, normalize=True, facets=1,
                                vis_slices=1, context="2d", gcfcf=None, **kwargs):
    """List of (image, sumwt), one per vis list (imaging_serial.py:125-216, facets = 1)."""
    if facets != 1:
        raise NotImplementedError("facets are outside the Fourier hot path")
    if not isinstance(template_model_imagelist, (list, tuple)):
        template_model_imagelist = [template_model_imagelist]
    mode = _mode(context)
    results = []
    for ivis, vis in enumerate(vis_list):
        model = template_model_imagelist[ivis]
        g = None if gcfcf is None else (gcfcf[ivis] if len(gcfcf) > 1 else gcfcf[0])
        if g is None:
            from .kernels import create_pswf_convolutionfunction
            g = create_pswf_convolutionfunction(template_model_imagelist[0])
        imager = Imager(vis, model, gcfcf=g, mode=mode, vis_slices=vis_slices if mode == "wstack" else 1, **kwargs)
        # ARL passes normalize to each slab's invert, weights each slab's image by its sumwt and
        # normalises the sum (imaging_serial.py:166-197)
        re, _ = imager.invert(vis, dopsf=dopsf, normalize=normalize, weigh_slabs=True)
        results.append((Image(re, model.wcs.deepcopy(), model.polarisation_frame), imager.sumwt()))
    return results


def predict_list_serial_workflow(vis_list, model_imagelist, context, vis_slices=1, facets=1, gcfcf=None, **kwargs):
    """List of predicted Visibility, one per vis list (imaging_serial.py:41-122, facets = 1)."""
    if facets != 1:
        raise NotImplementedError("facets are outside the Fourier hot path")
    mode = _mode(context)
    if gcfcf is None:
        from .kernels import create_pswf_convolutionfunction
        gcfcf = [create_pswf_convolutionfunction(m) for m in model_imagelist]
    out = []
    for ivis, vis in enumerate(vis_list):
        g = gcfcf[ivis] if len(gcfcf) > 1 else gcfcf[0]
        model = model_imagelist[ivis]
        imager = Imager(vis, model, gcfcf=g, mode=mode, vis_slices=vis_slices if mode == "wstack" else 1, **kwargs)
        out.append(vis.with_vis(imager.predict(vis, model.data)))
    return out
