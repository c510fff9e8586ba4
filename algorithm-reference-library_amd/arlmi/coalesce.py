"""BlockVisibility -> Visibility rows on the GPU (processing_components/visibility/coalesce.py):
the plain conversion (convert_blockvisibility_to_visibility, coalesce.py:85-109, and its
convert_blocks, :331-444), baseline-dependent averaging (coalesce_visibility / average_in_blocks,
:36-82, 146-328) and decoalescence (:112-143).
"""

import ctypes

import numpy
import torch

from . import _lib
from .data_models import BlockVisibility, Visibility, device
from .plan import ptr, stream_handle

__all__ = ["convert_blockvisibility_to_visibility", "convert_blocks", "decoalesce_visibility",
           "coalesce_visibility", "average_in_blocks"]


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


def _bda_factors(uvd2, wsum, coal, max_coal):
    """coalesce.py:208-226 on the host, numpy operation for operation."""
    uvmax = numpy.sqrt(numpy.max(uvd2))
    dist = numpy.sqrt(uvd2)
    f = numpy.ones(uvd2.shape, dtype=numpy.int64)
    pos = dist > 0.0
    f[pos] = numpy.round(coal * uvmax / dist[pos])
    f[~pos] = max_coal
    numpy.putmask(f, wsum == 0, 0)
    numpy.putmask(f, f < 1, 1)
    numpy.putmask(f, f > max_coal, max_coal)
    return f


def average_in_blocks(bvis, time_coal=1.0, max_time_coal=100, frequency_coal=1.0, max_frequency_coal=100):
    """average_in_blocks (coalesce.py:146-328) of a Stokes-I BlockVisibility on the GPU: a dict of
    the averaged row columns (device tensors) and cindex, as convert_blocks returns them, plus the
    per-baseline factors ``time_average`` and ``frequency_average`` ([nants, nants], numpy)."""
    if not isinstance(bvis, BlockVisibility):
        raise TypeError("vis is not a BlockVisibility: %r" % (bvis,))
    vis = bvis.data["vis"]
    if vis.shape[4] != 1:
        raise NotImplementedError("Stokes I (npol = 1) only")
    dev = vis.device
    precision = bvis.precision
    rdt = torch.float32 if precision == 32 else torch.float64
    ntimes, nants, nchan = bvis.ntimes, bvis.nants, bvis.nchan
    L = _lib.lib()
    uvw_b = bvis.data["uvw"].to(torch.float64).contiguous()
    wts = bvis.data["weight"].to(rdt).contiguous()
    uvd2 = torch.empty(nants * nants, dtype=torch.float64, device=dev)
    wsum = torch.empty_like(uvd2)
    _lib.check(L.arlmi_bda_uvdist(ptr(uvw_b), ptr(wts), ntimes, nants, nchan, ptr(uvd2), ptr(wsum), precision,
                                  stream_handle()))
    uvd2_h = uvd2.cpu().numpy().reshape(nants, nants)
    wsum_h = wsum.cpu().numpy().reshape(nants, nants)
    ta = _bda_factors(uvd2_h, wsum_h, time_coal, max_time_coal)
    fa = _bda_factors(uvd2_h, wsum_h, frequency_coal, max_frequency_coal)
    nt = numpy.where(ta > 1, (ntimes + ta - 1) // ta, ntimes)
    nf = numpy.where(fa > 1, (nchan + fa - 1) // fa, nchan)
    cells = (nt * nf).reshape(-1).astype(numpy.int64)
    rowoff = numpy.concatenate([[0], numpy.cumsum(cells)]).astype(numpy.int64)
    nrows = int(rowoff[-1])
    freq = torch.as_tensor(numpy.asarray(bvis.frequency, dtype=numpy.float64)).to(dev)
    bw = torch.as_tensor(numpy.asarray(bvis.channel_bandwidth, dtype=numpy.float64)).to(dev)
    cols = {k: bvis.data[k].contiguous() for k in ("vis", "imaging_weight")}
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
    d_ta = torch.as_tensor(ta.reshape(-1).astype(numpy.int32)).to(dev)
    d_fa = torch.as_tensor(fa.reshape(-1).astype(numpy.int32)).to(dev)
    d_off = torch.as_tensor(rowoff).to(dev)
    b = _lib.Blocks(ntimes=ntimes, nants=nants, nchan=nchan, reserved=0, d_vis=ptr(cols["vis"]), d_uvw=ptr(uvw_b),
                    d_weight=ptr(wts), d_imaging_weight=ptr(cols["imaging_weight"]), d_time=ptr(times),
                    d_integration_time=ptr(itimes), d_frequency=ptr(freq), d_channel_bandwidth=ptr(bw))
    r = _lib.Rows(d_uvw=ptr(out["uvw"]), d_vis=ptr(out["vis"]), d_weight=ptr(out["weight"]),
                  d_imaging_weight=ptr(out["imaging_weight"]), d_time=ptr(out["time"]),
                  d_integration_time=ptr(out["integration_time"]), d_frequency=ptr(out["frequency"]),
                  d_channel_bandwidth=ptr(out["channel_bandwidth"]), d_antenna1=ptr(out["antenna1"]),
                  d_antenna2=ptr(out["antenna2"]))
    _lib.check(L.arlmi_bda_average(ctypes.byref(b), ptr(d_ta), ptr(d_fa), ptr(d_off), nrows, ctypes.byref(r),
                                   ptr(cindex), precision, stream_handle()))
    out["cindex"] = cindex
    out["time_average"] = ta
    out["frequency_average"] = fa
    return out


def coalesce_visibility(bvis, **kwargs):
    """Visibility rows of a BlockVisibility with baseline-dependent averaging (coalesce.py:36-82):
    time_coal and frequency_coal (default 0) scale the per-baseline factors, max_time_coal and
    max_frequency_coal (default 100) cap them; both 0 is the plain conversion.  The result keeps
    ``cindex`` and ``blockvis`` for decoalesce_visibility."""
    if not isinstance(bvis, BlockVisibility):
        raise AssertionError("vis is not a BlockVisibility: %r" % (bvis,))
    time_coal = kwargs.get("time_coal", 0.0)
    max_time_coal = kwargs.get("max_time_coal", 100)
    frequency_coal = kwargs.get("frequency_coal", 0.0)
    max_frequency_coal = kwargs.get("max_frequency_coal", 100)
    if time_coal == 0.0 and frequency_coal == 0.0:
        return convert_blockvisibility_to_visibility(bvis)
    cols = average_in_blocks(bvis, time_coal, max_time_coal, frequency_coal, max_frequency_coal)
    cindex = cols.pop("cindex")
    cols.pop("time_average")
    cols.pop("frequency_average")
    vis = Visibility(cols, bvis.polarisation_frame, bvis.phasecentre)
    vis.cindex = cindex
    vis.blockvis = bvis
    return vis
