"""Seeded synthetic observations for the benchmark workloads.

Array geometry follows ARL's conventions (processing_library/util/coordinate_support.py:37-85:
local ENU -> celestial XYZ at the site latitude -> (u, v, w) per hour angle) and its row order
(processing_components/visibility/base.py:103-129: time, antenna1, antenna2 > antenna1, channel;
uvw = (pos[a2] - pos[a1]) * f / c in wavelengths).  Layouts, sky models and noise are seeded
with numpy.random.default_rng.  Row columns are generated on the GPU (float64) so that the
tens-of-millions-row workloads are built in seconds.
"""

import math

import numpy
import torch

from .data_models import Visibility, Skycomponent, C_LIGHT, device
from .kernels import create_image

__all__ = ["site_xyz", "uvw_rows", "make_observation", "config1", "config2", "config3", "config4", "config5",
           "power_law_sources", "uniform_disk", "log_radial", "gaussian_blobs", "block_uvw", "LATITUDE"]

LATITUDE = -26.82 * numpy.pi / 180.0  # site latitude used by every workload


def uniform_disk(rng, nants, radius):
    r = radius * numpy.sqrt(rng.uniform(0.0, 1.0, nants))
    phi = rng.uniform(0.0, 2 * numpy.pi, nants)
    return numpy.stack([r * numpy.cos(phi), r * numpy.sin(phi), numpy.zeros(nants)], axis=1)


def log_radial(rng, nants, rmax, rmin=20.0):
    r = 10.0 ** rng.uniform(numpy.log10(rmin), numpy.log10(rmax), nants)
    phi = rng.uniform(0.0, 2 * numpy.pi, nants)
    return numpy.stack([r * numpy.cos(phi), r * numpy.sin(phi), numpy.zeros(nants)], axis=1)


def site_xyz(enu, lat=LATITUDE):
    """Local (east, north, up) -> celestial XYZ (xyz_at_latitude, coordinate_support.py:37-56)."""
    x, y, z = enu[:, 0], enu[:, 1], enu[:, 2]
    lat2 = numpy.pi / 2 - lat
    y2 = -z * numpy.sin(lat2) + y * numpy.cos(lat2)
    z2 = z * numpy.cos(lat2) + y * numpy.sin(lat2)
    return numpy.stack([x, y2, z2], axis=1)


def uvw_rows(xyz, hour_angles, frequencies, dec, dev=None, zerow=False):
    """uvw [ntimes*nbl*nchan, 3] (wavelengths, float64, device) in ARL row order.

    xyz_to_uvw (coordinate_support.py:59-85) per hour angle, baselines a1 < a2 ordered
    (a1, a2), uvw = (pos[a2] - pos[a1]) * f / c (visibility/base.py:116-127)."""
    dev = dev or device()
    X = torch.as_tensor(xyz, dtype=torch.float64, device=dev)
    ha = torch.as_tensor(numpy.asarray(hour_angles, dtype=numpy.float64), device=dev)
    f = torch.as_tensor(numpy.asarray(frequencies, dtype=numpy.float64), device=dev)
    nants = X.shape[0]
    a1, a2 = numpy.triu_indices(nants, k=1)
    a1 = torch.as_tensor(a1, device=dev)
    a2 = torch.as_tensor(a2, device=dev)
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    ch, sh = torch.cos(ha)[:, None], torch.sin(ha)[:, None]
    u = x[None, :] * ch - y[None, :] * sh
    v0 = x[None, :] * sh + y[None, :] * ch
    w = z[None, :] * math.sin(dec) - v0 * math.cos(dec)
    v = z[None, :] * math.cos(dec) + v0 * math.sin(dec)
    pos = torch.stack([u, v, w], dim=-1)               # [nt, nants, 3]
    bl = pos[:, a2, :] - pos[:, a1, :]                  # [nt, nbl, 3]
    k = (f / C_LIGHT)                                   # [nchan]
    uvw = bl[:, :, None, :] * k[None, None, :, None]    # [nt, nbl, nchan, 3]
    uvw = uvw.reshape(-1, 3).contiguous()
    if zerow:
        uvw[:, 2] = 0.0
    nt, nbl, nch = ha.shape[0], a1.shape[0], f.shape[0]
    freq = f.repeat(nt * nbl)
    ant1 = a1.repeat_interleave(nch).repeat(nt)
    ant2 = a2.repeat_interleave(nch).repeat(nt)
    time = (ha * 43200.0 / math.pi).repeat_interleave(nbl * nch)
    return uvw, freq, ant1, ant2, time


def power_law_sources(rng, n, smin, smax, index, lmmax):
    """Fluxes with dN/dS ~ S^index on [smin, smax] and (l, m) uniform in +-lmmax."""
    a = index + 1.0
    u = rng.uniform(0.0, 1.0, n)
    flux = (smin ** a + u * (smax ** a - smin ** a)) ** (1.0 / a)
    l = rng.uniform(-lmmax, lmmax, n)
    m = rng.uniform(-lmmax, lmmax, n)
    return [Skycomponent(li, mi, [[fi]]) for li, mi, fi in zip(l, m, flux)]


def _select_uv(uvw, npixel, cellsize, support):
    """Rows whose kernel footprint stays on the grid: |u|, |v| <= (N/2 - S/2 - 1) * du."""
    du = 1.0 / (npixel * cellsize)
    lim = (npixel // 2 - support // 2 - 1) * du
    return (uvw[:, 0].abs() <= lim) & (uvw[:, 1].abs() <= lim)


def make_observation(enu, hour_angles, frequencies, dec, precision=32, npixel=None, cellsize=None, support=8,
                     zerow=False, select=True, dev=None):
    """Visibility (zero vis, unit weights) with optional on-grid row selection."""
    dev = dev or device()
    xyz = site_xyz(enu)
    uvw, freq, a1, a2, time = uvw_rows(xyz, hour_angles, frequencies, dec, dev=dev, zerow=zerow)
    total = uvw.shape[0]
    if select and npixel is not None:
        keep = _select_uv(uvw, npixel, cellsize, support)
        uvw, freq, a1, a2, time = uvw[keep].contiguous(), freq[keep], a1[keep], a2[keep], time[keep]
    cdt = torch.complex64 if precision == 32 else torch.complex128
    rdt = torch.float32 if precision == 32 else torch.float64
    n = uvw.shape[0]
    bw = float(frequencies[1] - frequencies[0]) if len(frequencies) > 1 else 1e6
    data = {
        "uvw": uvw,
        "vis": torch.zeros((n, 1), dtype=cdt, device=dev),
        "weight": torch.ones((n, 1), dtype=rdt, device=dev),
        "imaging_weight": torch.ones((n, 1), dtype=rdt, device=dev),
        "frequency": freq.contiguous(),
        "time": time.contiguous(),
        "channel_bandwidth": torch.full((n,), bw, dtype=torch.float64, device=dev),
        "integration_time": torch.ones((n,), dtype=torch.float64, device=dev),
        "antenna1": a1.contiguous(),
        "antenna2": a2.contiguous(),
    }
    vis = Visibility(data)
    vis.total_rows = total
    return vis


def config1(precision=64, seed=1, dev=None):
    """Plain 2-D imaging: 64 antennas uniform in a 4 km (diameter) disk, 100 MHz, 60 integrations
    over HA -1h..+1h at dec -45 deg, 10 sources U(0.1, 1) Jy within +-0.05 rad, noise 1e-3 Jy,
    512^2 image, cell 2.5e-4 rad, PSWF S=8 os=8, w set to zero."""
    from .imaging import predict_skycomponent_visibility
    rng = numpy.random.default_rng(seed)
    enu = uniform_disk(rng, 64, 2000.0)
    ha = numpy.linspace(-numpy.pi / 12, numpy.pi / 12, 60)
    freq = numpy.array([1e8])
    vis = make_observation(enu, ha, freq, -45.0 * numpy.pi / 180.0, precision=precision, zerow=True, select=False,
                           dev=dev)
    comps = [Skycomponent(rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05), [[rng.uniform(0.1, 1.0)]])
             for _ in range(10)]
    predict_skycomponent_visibility(vis, comps)
    noise = rng.normal(0.0, 1e-3, (vis.nvis, 2))
    vis.data["vis"] += torch.as_tensor(noise[:, 0] + 1j * noise[:, 1]).to(vis.vis.device, vis.vis.dtype)[:, None]
    dtype = torch.float32 if precision == 32 else torch.float64
    im = create_image(npixel=512, cellsize=2.5e-4, frequency=[1e8], channel_bandwidth=[1e6], dtype=dtype)
    return vis, im, comps


def config2(precision=32, seed=2, npixel=2048, cellsize=1.5e-4, support=8, nsources=200, dev=None, sky=True):
    """W-stacked wide field: 256 antennas log-radial to 10 km, 8 channels 100-150 MHz, 120
    integrations over HA -3h..+3h at dec -30 deg; 200 power-law (index -1.5, 1 mJy - 1 Jy) sources
    within +-0.15 rad; 2048^2 image at 1.5e-4 rad.  Rows whose kernel footprint would leave the
    grid are not generated (reported as vis.total_rows vs vis.nvis).  The image is MFS: its FREQ
    cdelt (1 GHz) maps every channel to image channel 0."""
    from .imaging import predict_skycomponent_visibility
    rng = numpy.random.default_rng(seed)
    enu = log_radial(rng, 256, 10000.0)
    ha = numpy.linspace(-numpy.pi / 4, numpy.pi / 4, 120)
    freq = numpy.linspace(100e6, 150e6, 8)
    vis = make_observation(enu, ha, freq, -30.0 * numpy.pi / 180.0, precision=precision, npixel=npixel,
                           cellsize=cellsize, support=support, select=True, dev=dev)
    comps = power_law_sources(rng, nsources, 1e-3, 1.0, -1.5, 0.15)
    if sky:
        predict_skycomponent_visibility(vis, comps)
    dtype = torch.float32 if precision == 32 else torch.float64
    im = create_image(npixel=npixel, cellsize=cellsize, frequency=[100e6], channel_bandwidth=[1e9], dtype=dtype)
    return vis, im, comps


def config3(nobs=64, seed=3, precision=32, dev=None, max_rows=4_000_000, min_rows=10_000):
    """Batching: nobs independent observations, each with a seeded antenna count in 32-512,
    channel count in 1-16 and integration count in 10-240, redrawn until its row count lies in
    [min_rows, max_rows]; image side from {256, 512, 1024, 2048}, PSWF support from {6, 8, 10},
    8-64 w slabs; arrays log-radial to 10 km, channels over 100-150 MHz, HA -1h..+1h at dec
    -30 deg, cell half the critical cell (imaging/base.py:303-318), MFS image.  Visibilities are
    seeded noise (the workload shape, not the sky, is what is measured).  Returns a list of
    dicts: vis, im, support, vis_slices, rows_generated."""
    from .imaging import create_image_from_visibility
    rng = numpy.random.default_rng(seed)
    dev = dev or device()
    out = []
    while len(out) < nobs:
        nants = int(rng.integers(32, 513))
        nchan = int(rng.integers(1, 17))
        nint = int(rng.integers(10, 241))
        rows = nint * nants * (nants - 1) // 2 * nchan
        if not (min_rows <= rows <= max_rows):
            continue
        npixel = int(rng.choice([256, 512, 1024, 2048]))
        support = int(rng.choice([6, 8, 10]))
        slices = int(rng.integers(8, 65))
        enu = log_radial(rng, nants, 10000.0)
        ha = numpy.linspace(-numpy.pi / 12, numpy.pi / 12, nint)
        freq = numpy.linspace(100e6, 150e6, nchan) if nchan > 1 else numpy.array([100e6])
        vis = make_observation(enu, ha, freq, -30.0 * numpy.pi / 180.0, precision=precision, select=False, dev=dev)
        tmpl = create_image_from_visibility(vis, npixel=npixel)
        cell = float(tmpl.cellsize)
        keep = _select_uv(vis.data["uvw"], npixel, cell, support)
        if not bool(keep.all()):
            vis = Visibility({k: v[keep].contiguous() for k, v in vis.data.items()})
        n = vis.nvis
        noise = torch.randn((n, 2), dtype=torch.float64, device=dev,
                            generator=torch.Generator(device=dev).manual_seed(int(rng.integers(1 << 30))))
        vis.data["vis"] = torch.complex(noise[:, 0], noise[:, 1]).to(vis.vis.dtype)[:, None].contiguous()
        dtype = torch.float32 if precision == 32 else torch.float64
        im = create_image(npixel=npixel, cellsize=cell, frequency=[freq[0]], channel_bandwidth=[1e9], dtype=dtype)
        out.append({"vis": vis, "im": im, "support": support, "vis_slices": slices, "rows_generated": rows})
    return out


def gaussian_blobs(rng, npixel, nblobs, fwhm_min, fwhm_max, flux_min=0.1, flux_max=1.0, margin=None):
    """[1, 1, N, N] float64 sum of circular Gaussians with FWHM U(fwhm_min, fwhm_max) pixels."""
    margin = margin if margin is not None else npixel // 4
    img = numpy.zeros((npixel, npixel))
    y, x = numpy.mgrid[0:npixel, 0:npixel]
    for _ in range(nblobs):
        cy, cx = rng.uniform(margin, npixel - margin, 2)
        sig = rng.uniform(fwhm_min, fwhm_max) / (2.0 * math.sqrt(2.0 * math.log(2.0)))
        r = max(1, int(5 * sig))
        iy, ix = int(cy), int(cx)
        sl = (slice(iy - r, iy + r + 1), slice(ix - r, ix + r + 1))
        img[sl] += rng.uniform(flux_min, flux_max) * numpy.exp(-((y[sl] - cy) ** 2 + (x[sl] - cx) ** 2) / (2 * sig ** 2))
    return img[None, None]


def block_uvw(xyz, hour_angles, dec, dev=None):
    """BlockVisibility uvw [ntimes, nants, nants, 3] metres, uvw[t, i, j] = pos[i] - pos[j]
    (visibility/base.py:224-233), float64 on the device."""
    dev = dev or device()
    X = torch.as_tensor(xyz, dtype=torch.float64, device=dev)
    ha = torch.as_tensor(numpy.asarray(hour_angles, dtype=numpy.float64), device=dev)
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    ch, sh = torch.cos(ha)[:, None], torch.sin(ha)[:, None]
    u = x[None, :] * ch - y[None, :] * sh
    v0 = x[None, :] * sh + y[None, :] * ch
    w = z[None, :] * math.sin(dec) - v0 * math.cos(dec)
    v = z[None, :] * math.cos(dec) + v0 * math.sin(dec)
    pos = torch.stack([u, v, w], dim=-1)  # [nt, nants, 3]
    return (pos[:, :, None, :] - pos[:, None, :, :]).contiguous()


def config4(seed=4, precision=32, npixel=1024, dev=None):
    """Differentiation: 128 antennas uniform in a 6 km (diameter) disk, 4 channels 120-140 MHz, 60
    integrations over HA -1h..+1h at dec -30 deg, as a BlockVisibility converted to rows on the
    GPU (8,128 baselines x 60 x 4 = 1,950,720 rows); a 1024^2 image of 50 Gaussian blobs (FWHM 2-6
    pixels) at half the critical cell; gains amplitude 1 +- 0.1 (uniform), phase U(-pi/4, pi/4) per
    antenna per channel (one gain-table row per integration, all equal).  Returns a dict: bvis,
    vis, im, image (numpy [1, 1, N, N]), gaintable."""
    from .coalesce import convert_blockvisibility_to_visibility
    from .calibration import create_gaintable_from_blockvisibility
    from .data_models import BlockVisibility
    from .imaging import create_image_from_visibility
    rng = numpy.random.default_rng(seed)
    dev = dev or device()
    nants, nchan, nint = 128, 4, 60
    enu = uniform_disk(rng, nants, 3000.0)
    ha = numpy.linspace(-numpy.pi / 12, numpy.pi / 12, nint)
    freq = numpy.linspace(120e6, 140e6, nchan)
    dec = -30.0 * numpy.pi / 180.0
    uvw = block_uvw(site_xyz(enu), ha, dec, dev=dev)
    cdt = torch.complex64 if precision == 32 else torch.complex128
    rdt = torch.float32 if precision == 32 else torch.float64
    shape = (nint, nants, nants, nchan, 1)
    data = {
        "uvw": uvw,
        "vis": torch.zeros(shape, dtype=cdt, device=dev),
        "weight": torch.ones(shape, dtype=rdt, device=dev),
        "imaging_weight": torch.ones(shape, dtype=rdt, device=dev),
        "time": torch.as_tensor(ha * 43200.0 / numpy.pi).to(dev),
        "integration_time": torch.full((nint,), 10.0, dtype=torch.float64, device=dev),
    }
    bvis = BlockVisibility(data, frequency=freq, channel_bandwidth=numpy.full(nchan, 5e6))
    vis = convert_blockvisibility_to_visibility(bvis)
    tmpl = create_image_from_visibility(vis, npixel=npixel)
    dtype = torch.float32 if precision == 32 else torch.float64
    im = create_image(npixel=npixel, cellsize=float(tmpl.cellsize), frequency=[freq[0]], channel_bandwidth=[1e9],
                      dtype=dtype)
    image = gaussian_blobs(rng, npixel, 50, 2.0, 6.0)
    gt = create_gaintable_from_blockvisibility(bvis)
    amp = rng.uniform(0.9, 1.1, (nants, nchan))
    phase = rng.uniform(-numpy.pi / 4, numpy.pi / 4, (nants, nchan))
    g = (amp * numpy.exp(1j * phase))[None, :, :, None, None].repeat(gt.ntimes, axis=0)
    gt.gain = torch.as_tensor(g).to(device=dev, dtype=cdt)
    return {"bvis": bvis, "vis": vis, "im": im, "image": image, "gaintable": gt}


def core_spiral(rng, ncore=133, narm=64, core_radius=1000.0, rmax=80000.0, arms=3):
    """Dense core (uniform disk) plus log-spiral arms out to rmax (metres, local ENU)."""
    core = uniform_disk(rng, ncore, core_radius)
    k = numpy.arange(narm)
    arm = k % arms
    r = core_radius * (rmax / core_radius) ** ((k // arms + 1) / math.ceil(narm / arms))
    phi = 2 * numpy.pi * arm / arms + 0.8 * numpy.log(r / core_radius) + rng.normal(0, 0.02, narm)
    spiral = numpy.stack([r * numpy.cos(phi), r * numpy.sin(phi), numpy.zeros(narm)], axis=1)
    return numpy.concatenate([core, spiral])


def config5(precision=32, seed=5, integrations=13, first=0, npixel=8192, cellsize=3e-6, support=8, dev=None):
    """Streaming wide band: 197 dishes (133 in a 1 km core, 64 on three log-spiral arms to 80 km),
    64 channels 0.95-1.76 GHz, 480 integrations over HA -4h..+4h at dec -30 deg; one chunk of
    `integrations` consecutive integrations starting at `first` (13 integrations = 16.06 M rows
    before the on-grid selection), 8192^2 image at 3e-6 rad.  Rows whose kernel footprint would
    leave the grid are not generated (vis.total_rows vs vis.nvis).  Visibilities are zero (the
    caller fills them)."""
    rng = numpy.random.default_rng(seed)
    enu = core_spiral(rng)
    ha_all = numpy.linspace(-numpy.pi / 3, numpy.pi / 3, 480)
    ha = ha_all[first:first + integrations]
    freq = numpy.linspace(0.95e9, 1.76e9, 64)
    vis = make_observation(enu, ha, freq, -30.0 * numpy.pi / 180.0, precision=precision, npixel=npixel,
                           cellsize=cellsize, support=support, select=True, dev=dev)
    dtype = torch.float32 if precision == 32 else torch.float64
    im = create_image(npixel=npixel, cellsize=cellsize, frequency=[freq[0]], channel_bandwidth=[1e10], dtype=dtype)
    return vis, im
