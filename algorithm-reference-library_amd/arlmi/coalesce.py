"""BlockVisibility -> Visibility rows on the GPU (processing_components/visibility/coalesce.py).

Only the no-averaging conversion (convert_blockvisibility_to_visibility, coalesce.py:85-109,
and its convert_blocks, :331-444) is provided; baseline-dependent averaging is not.
"""

import ctypes

import numpy
import torch

from . import _lib
from .data_models import BlockVisibility, Visibility, device
from .plan import ptr, stream_handle

__all__ = ["convert_blockvisibility_to_visibility", "convert_blocks", "decoalesce_visibility"]


def convert_blocks(bvis):
    """Row columns of a BlockVisibility (coalesce.py:331-444) as a dict of device tensors:
    uvw [V, 3] (wavelengths), vis, weight, imaging_weight [V, 1], time, integration_time,
    frequency, channel_bandwidth [V], antenna1, antenna2 [V] (int64) and cindex
    [ntimes * nants * nants * nchan] (int64).  Rows are ordered (time, antenna2, antenna1, chan)
    with antenna1 < antenna2."""
    if not isinstance(bvis, BlockVisibility):
        raise TypeError("vis is not a BlockVisibility: %r" % (bvis,))
    vis = bvis.data["vis"]
    if vis.shape[4] != 1:
        raise NotImplementedError("Stokes I (npol = 1) only")
    dev = vis.device
    precision = bvis.precision
    rdt = torch.float32 if precision == 32 else torch.float64
    ntimes, nants, nchan = bvis.ntimes, bvis.nants, bvis.nchan
    nrows = ntimes * (nants * (nants - 1) // 2) * nchan
    freq = torch.as_tensor(numpy.asarray(bvis.frequency, dtype=numpy.float64)).to(dev)
    bw = torch.as_tensor(numpy.asarray(bvis.channel_bandwidth, dtype=numpy.float64)).to(dev)
    uvw_b = bvis.data["uvw"].to(torch.float64).contiguous()
    cols = {k: bvis.data[k].contiguous() for k in ("vis", "weight", "imaging_weight")}
    cols["weight"] = cols["weight"].to(rdt)
    cols["imaging_weight"] = cols["imaging_weight"].to(rdt)
    times = bvis.data["time"].to(torch.float64).contiguous()
    itimes = bvis.data["integration_time"].to(torch.float64).contiguous()
    out = {
        "uvw": torch.empty((nrows, 3), dtype=torch.float64, device=dev),
        "vis": torch.empty((nrows, 1), dtype=vis.dtype, device=dev),
        "weight": torch.empty((nrows, 1), dtype=rdt, device=dev),
        "imaging_weight": torch.empty((nrows, 1), dtype=rdt, device=dev),
        "time": torch.empty(nrows, dtype=torch.float64, device=dev),
        "integration_time": torch.empty(nrows, dtype=torch.float64, device=dev),
        "frequency": torch.empty(nrows, dtype=torch.float64, device=dev),
        "channel_bandwidth": torch.empty(nrows, dtype=torch.float64, device=dev),
        "antenna1": torch.empty(nrows, dtype=torch.int64, device=dev),
        "antenna2": torch.empty(nrows, dtype=torch.int64, device=dev),
    }
    cindex = torch.empty(ntimes * nants * nants * nchan, dtype=torch.int64, device=dev)
    b = _lib.Blocks(ntimes=ntimes, nants=nants, nchan=nchan, reserved=0, d_vis=ptr(cols["vis"]), d_uvw=ptr(uvw_b),
                    d_weight=ptr(cols["weight"]), d_imaging_weight=ptr(cols["imaging_weight"]), d_time=ptr(times),
                    d_integration_time=ptr(itimes), d_frequency=ptr(freq), d_channel_bandwidth=ptr(bw))
    r = _lib.Rows(d_uvw=ptr(out["uvw"]), d_vis=ptr(out["vis"]), d_weight=ptr(out["weight"]),
                  d_imaging_weight=ptr(out["imaging_weight"]), d_time=ptr(out["time"]),
                  d_integration_time=ptr(out["integration_time"]), d_frequency=ptr(out["frequency"]),
                  d_channel_bandwidth=ptr(out["channel_bandwidth"]), d_antenna1=ptr(out["antenna1"]),
                  d_antenna2=ptr(out["antenna2"]))
    _lib.check(_lib.lib().arlmi_block_to_rows(ctypes.byref(b), ctypes.byref(r), ptr(cindex), precision,
                                              stream_handle()))
    out["cindex"] = cindex
    return out


def convert_blockvisibility_to_visibility(bvis):
    """Visibility rows of a BlockVisibility, no averaging (coalesce.py:85-109); the result keeps
    ``cindex`` (row of every lower-triangle block element) and ``blockvis``."""
    cols = convert_blocks(bvis)
    cindex = cols.pop("cindex")
    vis = Visibility(cols, bvis.polarisation_frame, bvis.phasecentre)
    vis.cindex = cindex
    vis.blockvis = bvis
    return vis


def decoalesce_visibility(vis):
    """Rows back into the BlockVisibility layout (coalesce.py:112-143): every block element i
    takes vis, weight and imaging_weight of row cindex[i]; elements outside the lower triangle
    have cindex 0 and so take row 0, as in ARL.  Returns a new BlockVisibility (a torch gather
    on the device)."""
    bvis = vis.blockvis
    shape = bvis.data["vis"].shape
    if int(vis.cindex.max()) >= vis.nvis:
        raise AssertionError("Incorrect template used in decoalescing")
    data = dict(bvis.data)
    for k in ("vis", "weight", "imaging_weight"):
        data[k] = vis.data[k].reshape(-1)[vis.cindex].reshape(shape).contiguous()
    return BlockVisibility(data, frequency=bvis.frequency, channel_bandwidth=bvis.channel_bandwidth,
                           polarisation_frame=bvis.polarisation_frame, phasecentre=bvis.phasecentre)
