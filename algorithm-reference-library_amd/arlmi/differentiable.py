"""Predict and gain application as torch.autograd Functions with hand-written adjoints.

The forward model of gradient-based imaging and calibration,
    V_model[r] = g[antenna2] conj(g[antenna1]) * predict(I)[r],
(gain convention of calibration/operations.py:127-140 in row form; predict as
imaging_serial.py:41-98 or imaging/base.py:104-138), is linear in the real image I.  Its
backward is the exact adjoint, not a traced graph: the image gradient is the invert of the
gain-adjusted residual gradient with unit weights and no normalisation (grid with the
conjugated kernel, F^H, conjugate screens of predict, gcf), and the gain gradient is
summed per antenna by a HIP kernel.  Gradients follow torch's convention for complex
tensors (dL/dRe + i dL/dIm for a real loss L).
"""

import torch

from .calibration import apply_gains_rows, apply_gains_rows_adjoint

__all__ = ["PredictFunction", "GainsFunction", "predict_with_gains", "visibility_loss"]


class PredictFunction(torch.autograd.Function):
    """rows = imager.predict(vis, image); image [nchan, 1, N, N] real (float32 / float64)."""

    @staticmethod
    def forward(ctx, image, imager, vis):
        ctx.imager = imager
        ctx.vis = vis
        return imager.predict(vis, image.detach())

    @staticmethod
    def backward(ctx, grad_rows):
        if not ctx.needs_input_grad[0]:
            return None, None, None
        return ctx.imager.adjoint(ctx.vis, grad_rows), None, None


class GainsFunction(torch.autograd.Function):
    """out[r] = model[r] g[gidx2[r]] conj(g[gidx1[r]]); gain of any shape, indexed flat."""

    @staticmethod
    def forward(ctx, model_rows, gain, gidx1, gidx2):
        ctx.save_for_backward(model_rows, gain)
        ctx.gidx = (gidx1, gidx2)
        return apply_gains_rows(model_rows.detach(), gidx1, gidx2, gain.detach())

    @staticmethod
    def backward(ctx, grad_out):
        model_rows, gain = ctx.saved_tensors
        i1, i2 = ctx.gidx
        gv, gg = apply_gains_rows_adjoint(model_rows, grad_out, i1, i2, gain, need_vis=ctx.needs_input_grad[0],
                                          need_gain=ctx.needs_input_grad[1])
        if gg is not None:
            gg = gg.to(gain.dtype)
        return gv, gg, None, None


def predict_with_gains(image, gain, imager, vis, gidx1, gidx2):
    """Differentiable V_model = g[antenna2] conj(g[antenna1]) predict(image) on the rows of vis."""
    rows = PredictFunction.apply(image, imager, vis)
    return GainsFunction.apply(rows, gain, gidx1, gidx2)


def visibility_loss(image, gain, imager, vis, gidx1, gidx2, weight=None):
    """sum_r w_r |V_r - g[antenna2] conj(g[antenna1]) predict(image)_r|^2 (w = 1 by default)."""
    model = predict_with_gains(image, gain, imager, vis, gidx1, gidx2)
    d = vis.vis - model
    a = d.real ** 2 + d.imag ** 2
    if weight is not None:
        a = a * weight
    return a.sum()
