"""Scalar gain convention on the GPU (processing_components/calibration/operations.py).

V_obs = g_i conj(g_j) V_model for the [t, i, j] element of a BlockVisibility
(GainTable docstring, memory_data_models.py:143-146; apply_gaintable, operations.py:79-164).
In row form (coalesce.py:413-416, rows from the [t, antenna2, antenna1] triangle) the
un-conjugated gain is antenna2's: V_row = g[antenna2] conj(g[antenna1]) V.
"""

import numpy
import torch

from . import _lib
from .data_models import BlockVisibility, GainTable
from .plan import ptr, stream_handle

__all__ = ["create_gaintable_from_blockvisibility", "apply_gaintable", "gain_slots", "gain_row_indices",
           "apply_gains_rows", "apply_gains_rows_adjoint", "solve_gaintable"]


def _np(x):
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else numpy.asarray(x)


def create_gaintable_from_blockvisibility(bvis, timeslice=None):
    """Unit gains consistent with a BlockVisibility (operations.py:25-77), npol = 1: one gain row
    per solution interval of ``timeslice`` seconds (default: the smallest integration time)."""
    if not isinstance(bvis, BlockVisibility):
        raise TypeError("vis is not a BlockVisibility: %r" % (bvis,))
    time = _np(bvis.time).astype(numpy.float64)
    if timeslice is None or timeslice == "auto":
        timeslice = float(numpy.min(_np(bvis.data["integration_time"])))
    utimes = timeslice * numpy.unique(numpy.round((time - time[0]) / timeslice))
    ntimes = len(utimes)
    ufreq = numpy.unique(numpy.asarray(bvis.frequency, dtype=numpy.float64))
    shape = (ntimes, bvis.nants, len(ufreq), 1, 1)
    cdt = bvis.data["vis"].dtype
    dev = bvis.data["vis"].device
    gain = torch.ones(shape, dtype=cdt, device=dev)
    weight = torch.ones(shape, dtype=torch.float64, device=dev)
    return GainTable(gain, utimes + time[0], timeslice * numpy.ones(ntimes), weight=weight, frequency=ufreq)


def gain_slots(times, gt):
    """Gain-table row of each integration time: the row with |gt.time - t| < gt.interval / 2
    (operations.py:111-113), or -1 where none matches (ARL then leaves those visibilities as they
    are).  Each integration is its own time chunk, which is what vis_timeslice_iter's default
    slicing gives for a BlockVisibility with one integration per distinct time."""
    t = _np(times).astype(numpy.float64)
    match = numpy.abs(gt.time[None, :] - t[:, None]) < gt.interval[None, :] / 2.0
    nmatch = match.sum(axis=1)
    if numpy.any(nmatch > 1):
        raise NotImplementedError("an integration matches more than one gain-table row")
    return numpy.where(nmatch == 1, numpy.argmax(match, axis=1), -1).astype(numpy.int32)


def apply_gaintable(bvis, gt, inverse=False):
    """Apply scalar gains to a BlockVisibility in place (operations.py:79-164, scalar branch
    127-140) and return it.  Weights are not changed, as in ARL."""
    if not isinstance(bvis, BlockVisibility):
        raise TypeError("vis is not a BlockVisibility: %r" % (bvis,))
    if not isinstance(gt, GainTable):
        raise TypeError("gt is not a GainTable: %r" % (gt,))
    if gt.gain.shape[-2:] != (1, 1) or bvis.data["vis"].shape[4] != 1:
        raise NotImplementedError("scalar gains on Stokes I (npol = 1) only")
    if gt.nants != bvis.nants or gt.nchan != bvis.nchan:
        raise ValueError("gain table does not match the visibility (antennas / channels)")
    vis = bvis.data["vis"]
    if not vis.is_contiguous():
        raise ValueError("vis must be contiguous")
    dev = vis.device
    gain = gt.gain.to(device=dev, dtype=vis.dtype).contiguous()
    gwt = None if gt.weight is None else gt.weight.to(device=dev, dtype=torch.float64).contiguous()
    slots = torch.as_tensor(gain_slots(bvis.time, gt)).to(dev)
    _lib.check(_lib.lib().arlmi_apply_gaintable(ptr(vis), bvis.ntimes, bvis.nants, bvis.nchan, ptr(gain), ptr(gwt),
                                                ptr(slots), gt.ntimes, int(bool(inverse)), bvis.precision,
                                                stream_handle()))
    return bvis


def gain_row_indices(vis, gt):
    """int32 device indices (gidx1, gidx2) into gt.gain flattened as [slot, antenna, chan] of the
    gains of antenna1 and antenna2 of every row; rows with no matching gain row get -1 (their
    gain product is taken as 0 by the row kernels)."""
    times = _np(vis.data["time"]).astype(numpy.float64)
    ut, inv = numpy.unique(times, return_inverse=True)
    slot = gain_slots(ut, gt)[inv]
    freqs = _np(vis.data["frequency"]).astype(numpy.float64)
    chan = numpy.searchsorted(gt.frequency, freqs)
    chan = numpy.clip(chan, 0, gt.nchan - 1)
    if not numpy.array_equal(gt.frequency[chan], freqs):
        raise ValueError("row frequency not in the gain table")
    a1 = _np(vis.data["antenna1"]).astype(numpy.int64)
    a2 = _np(vis.data["antenna2"]).astype(numpy.int64)
    nants, nchan = gt.nants, gt.nchan
    i1 = numpy.where(slot >= 0, (slot * nants + a1) * nchan + chan, -1).astype(numpy.int32)
    i2 = numpy.where(slot >= 0, (slot * nants + a2) * nchan + chan, -1).astype(numpy.int32)
    dev = vis.data["uvw"].device
    return torch.as_tensor(i1).to(dev), torch.as_tensor(i2).to(dev)


def _flat_gain(gain, cdt, dev):
    return gain.reshape(-1).to(device=dev, dtype=cdt).contiguous()


def apply_gains_rows(vis_rows, gidx1, gidx2, gain):
    """out[r] = vis[r] g[gidx2[r]] conj(g[gidx1[r]]) (a new tensor, same shape as vis_rows)."""
    precision = 32 if vis_rows.dtype == torch.complex64 else 64
    v = vis_rows.contiguous()
    g = _flat_gain(gain, v.dtype, v.device)
    out = torch.empty_like(v)
    _lib.check(_lib.lib().arlmi_apply_gains_rows(v.numel(), ptr(gidx1), ptr(gidx2), ptr(g), g.numel(), ptr(v),
                                                 ptr(out), precision, stream_handle()))
    return out


def apply_gains_rows_adjoint(vis_rows, grad_out, gidx1, gidx2, gain, need_vis=True, need_gain=True):
    """Gradients of a real loss through apply_gains_rows (torch's convention dL/dRe + i dL/dIm):
    returns (grad_vis or None, grad_gain complex128 of gain's shape or None)."""
    precision = 32 if vis_rows.dtype == torch.complex64 else 64
    v = vis_rows.contiguous()
    go = grad_out.to(v.dtype).contiguous()
    g = _flat_gain(gain, v.dtype, v.device)
    gv = torch.empty_like(v) if need_vis else None
    gg = torch.zeros(g.numel(), dtype=torch.complex128, device=v.device) if need_gain else None
    _lib.check(_lib.lib().arlmi_apply_gains_rows_adjoint(v.numel(), ptr(gidx1), ptr(gidx2), ptr(g), g.numel(), ptr(v),
                                                         ptr(go), ptr(gv), ptr(gg), precision, stream_handle()))
    return gv, (gg.reshape(gain.shape) if gg is not None else None)


def solve_gaintable(vis, modelvis=None, gt=None, phase_only=True, niter=30, tol=1e-8, crosspol=False,
                    normalise_gains=True, refant=0, damping=0.5, **kwargs):
    """Scalar gain solution on the GPU (calibration.py:30-101 with solvers.py:61-140, 302-332).

    vis, modelvis: BlockVisibility (npol = 1); gt: a GainTable to start from, else
    create_gaintable_from_blockvisibility(vis, **kwargs).  Returns a GainTable with complex128
    gains, float64 weights (the last substitution's denominators) and ``residual`` [ntimes, nchan,
    1, 1]; ``iterations`` holds the iterations each row ran.  refant and damping are the solver's
    defaults, which ARL's solve_gaintable does not forward."""
    if not isinstance(vis, BlockVisibility):
        raise AssertionError("vis is not a BlockVisibility: %r" % (vis,))
    if modelvis is not None:
        if not isinstance(modelvis, BlockVisibility):
            raise AssertionError("modelvis is not a BlockVisibility: %r" % (modelvis,))
        if not float(modelvis.data["vis"].abs().max()) > 0.0:
            raise AssertionError("Model visibility is zero")
    if crosspol or vis.data["vis"].shape[4] != 1:
        raise NotImplementedError("scalar (npol = 1) gain solutions only")
    if gt is None:
        gt = create_gaintable_from_blockvisibility(vis, **kwargs)
    dev = vis.data["vis"].device
    ntimes, nants, nchan = vis.ntimes, vis.nants, vis.nchan
    if gt.nants != nants or gt.nchan != nchan:
        raise ValueError("gain table does not match the visibility (antennas / channels)")
    slots = gain_slots(vis.time, gt)
    active = numpy.bincount(slots[slots >= 0], minlength=gt.ntimes)[:gt.ntimes].astype(numpy.int32)
    v = vis.data["vis"].contiguous()
    m = None if modelvis is None else modelvis.data["vis"].to(v.dtype).contiguous()
    w = vis.data.get("weight")
    w = None if w is None else w.to(dtype=v.real.dtype).contiguous()
    gain = gt.gain.to(device=dev, dtype=torch.complex128).reshape(gt.ntimes, nants, nchan).contiguous().clone()
    gwt = torch.zeros((gt.ntimes, nants, nchan), dtype=torch.float64, device=dev)
    if gt.weight is not None:
        gwt.copy_(gt.weight.to(device=dev, dtype=torch.float64).reshape(gt.ntimes, nants, nchan))
    residual = torch.zeros((gt.ntimes, nchan), dtype=torch.float64, device=dev)
    old_res = getattr(gt, "residual", None)
    if old_res is not None:
        residual.copy_(torch.as_tensor(old_res).to(device=dev, dtype=torch.float64).reshape(gt.ntimes, nchan))
    iters = torch.zeros(gt.ntimes, dtype=torch.int32, device=dev)
    work = torch.empty(3 * gt.ntimes * nants * nants * nchan, dtype=torch.float64, device=dev)
    d_slots = torch.as_tensor(slots).to(dev)
    d_active = torch.as_tensor(active).to(dev)
    _lib.check(_lib.lib().arlmi_solve_gains(ptr(v), ptr(m), ptr(w), ntimes, nants, nchan, ptr(d_slots), gt.ntimes,
                                            ptr(d_active), ptr(gain), ptr(gwt), ptr(residual), ptr(iters), ptr(work),
                                            int(bool(phase_only)), int(niter), float(tol), int(refant),
                                            float(damping), int(bool(normalise_gains)), vis.precision,
                                            stream_handle()))
    out = GainTable(gain.reshape(gt.ntimes, nants, nchan, 1, 1), gt.time, gt.interval,
                    weight=gwt.reshape(gt.ntimes, nants, nchan, 1, 1), frequency=gt.frequency)
    out.residual = residual.reshape(gt.ntimes, nchan, 1, 1)
    out.iterations = iters
    return out
