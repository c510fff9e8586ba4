"""Shared helpers for the tests: seeded inputs and conversions between the arlmi device
data models and the oracle's plain arrays."""

import numpy

from oracle import arl_oracle as O


def rel_err(a, b):
    """max |a - b| / max |b| (both numpy)."""
    a = numpy.asarray(a)
    b = numpy.asarray(b)
    den = numpy.max(numpy.abs(b))
    return float(numpy.max(numpy.abs(a - b)) / (den if den > 0 else 1.0))


def image_errors(a, b):
    """(L2 relative, max-abs relative over the inner half of each axis, max-abs relative over the
    whole image) of image a against reference b.  ARL's grid correction max/psi(nu)^2 grows to
    ~6e4 at the image corners, so float32 rounding there is amplified by that much; the inner
    half (correction below ~10) and the L2 norm measure the transform itself."""
    a = numpy.asarray(a, dtype=float)
    b = numpy.asarray(b, dtype=float)
    n = b.shape[-1]
    q = n // 4
    inner = (slice(None),) * (b.ndim - 2) + (slice(q, n - q), slice(q, n - q))
    l2 = float(numpy.linalg.norm(a - b) / numpy.linalg.norm(b))
    mi = float(numpy.max(numpy.abs(a[inner] - b[inner])) / numpy.max(numpy.abs(b[inner])))
    return l2, mi, rel_err(a, b)


def vis_errors(a, b):
    """(L2 relative, max-abs relative to the peak) of visibilities a against reference b."""
    a = numpy.asarray(a)
    b = numpy.asarray(b)
    return float(numpy.linalg.norm(a - b) / numpy.linalg.norm(b)), rel_err(a, b)


def random_rows(rng, nvis, npixel, cellsize, support, wmax=0.0, margin=1, nfreq=1):
    """uvw rows whose kernel footprint stays on the grid, plus complex vis and weights."""
    du = 1.0 / (npixel * cellsize)
    lim = (npixel // 2 - support // 2 - margin) * du
    uvw = numpy.zeros((nvis, 3))
    uvw[:, 0] = rng.uniform(-lim, lim, nvis)
    uvw[:, 1] = rng.uniform(-lim, lim, nvis)
    uvw[:, 2] = rng.uniform(-wmax, wmax, nvis) if wmax > 0 else 0.0
    vis = rng.normal(size=(nvis, 1)) + 1j * rng.normal(size=(nvis, 1))
    wt = rng.uniform(0.5, 2.0, (nvis, 1))
    freq = numpy.full(nvis, 1e8)
    return {"uvw": uvw, "vis": vis, "imaging_weight": wt, "frequency": freq}


def to_device_vis(d, precision):
    import arlmi
    return arlmi.Visibility.from_arrays(d["uvw"], vis=d["vis"], weight=d["imaging_weight"],
                                        imaging_weight=d["imaging_weight"], frequency=d["frequency"],
                                        precision=precision)


def oracle_geom(npixel, cellsize, nchan=1, frequency=1e8, bw=1e6):
    return O.ImageGeom(npixel, cellsize, nchan=nchan, frequency=frequency, channel_bandwidth=bw)


def device_image(npixel, cellsize, precision, frequency=1e8, bw=1e6, data=None):
    import torch
    import arlmi
    dt = torch.float32 if precision == 32 else torch.float64
    im = arlmi.create_image(npixel=npixel, cellsize=cellsize, frequency=[frequency], channel_bandwidth=[bw], dtype=dt)
    if data is not None:
        im.data = torch.as_tensor(numpy.asarray(data)).to(device=im.data.device, dtype=dt).reshape(im.data.shape)
    return im
