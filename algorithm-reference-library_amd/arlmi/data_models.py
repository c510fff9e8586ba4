"""ARL's data models for the Fourier hot path, holding their arrays as device tensors.

Mirrors data_models/memory_data_models.py (Image 349-434, GridData 437-524,
ConvolutionFunction 527-614, Visibility 745-937, BlockVisibility 940-1094) and the
parts of their astropy WCS that the hot path reads.  The WCS is a small linear
stand-in (``WCS``) with the same ``wcs.crpix / cdelt / crval / ctype`` arrays and the
same values ARL stores; no celestial transformation is implemented (image centre and
visibility phase centre always coincide here).
"""

import copy

import numpy
import torch

__all__ = ["WCS", "PolarisationFrame", "Image", "GridData", "ConvolutionFunction", "Visibility",
           "BlockVisibility", "GainTable", "Skycomponent", "device", "C_LIGHT"]

C_LIGHT = 299792458.0
D2R = numpy.pi / 180.0


def device():
    """The GPU the arrays live on (torch's current device)."""
    return torch.device("cuda", torch.cuda.current_device())


class _WcsPrm:
    def __init__(self, naxis):
        self.naxis = naxis
        self.crpix = numpy.zeros(naxis)
        self.cdelt = numpy.ones(naxis)
        self.crval = numpy.zeros(naxis)
        self.ctype = [""] * naxis


class WCS:
    """Linear WCS with the attributes the hot path reads (astropy.wcs.WCS stand-in)."""

    def __init__(self, naxis=4):
        self.naxis = naxis
        self.wcs = _WcsPrm(naxis)
        # True once crval of the RA---SIN / DEC--SIN axes holds a sky position
        self.celestial = False

    def deepcopy(self):
        return copy.deepcopy(self)

    def world2pix0(self, axis, world):
        """Origin-0 pixel of a linear axis: ((world - crval)/cdelt + crpix) - 1."""
        w = self.wcs
        return ((numpy.asarray(world) - w.crval[axis]) / w.cdelt[axis] + w.crpix[axis]) - 1.0

    def pix2world0(self, axis, pix):
        w = self.wcs
        return w.cdelt[axis] * ((numpy.asarray(pix) + 1.0) - w.crpix[axis]) + w.crval[axis]


class PolarisationFrame:
    """Only Stokes I (npol = 1) is handled by the hot path (data_models/polarisation.py:66-113)."""

    def __init__(self, name="stokesI"):
        if name != "stokesI":
            raise NotImplementedError("only stokesI is supported")
        self.type = name
        self.npol = 1

    def __eq__(self, other):
        return isinstance(other, PolarisationFrame) and other.type == self.type


class Image:
    """Image: data [nchan, npol, ny, nx] plus WCS (memory_data_models.py:349-434)."""

    def __init__(self, data=None, wcs=None, polarisation_frame=None):
        self.data = data
        self.wcs = wcs
        self.polarisation_frame = polarisation_frame or PolarisationFrame()

    @property
    def shape(self):
        return tuple(self.data.shape)

    @property
    def nchan(self):
        return self.data.shape[0]

    @property
    def npol(self):
        return self.data.shape[1]

    @property
    def npixel(self):
        return self.data.shape[3]

    @property
    def phasecentre(self):
        """SkyCoord of crval's (ra, dec) (memory_data_models.py:422-424), or None for an image
        made without a sky position."""
        if self.wcs is None or not getattr(self.wcs, "celestial", False):
            return None
        return SkyCoord.from_degrees(self.wcs.wcs.crval[0], self.wcs.wcs.crval[1])

    @property
    def cellsize(self):
        """abs(cdelt0) * pi / 180 as ARL reads it back (image/operations.py:389)."""
        return abs(self.wcs.wcs.cdelt[0]) * numpy.pi / 180.0

    def numpy(self):
        d = self.data
        return d.detach().cpu().numpy() if isinstance(d, torch.Tensor) else numpy.asarray(d)


class GridData:
    """uv grid [nchan, npol, nw, ny, nx] complex (memory_data_models.py:437-524)."""

    def __init__(self, data=None, grid_wcs=None, projection_wcs=None, polarisation_frame=None):
        self.data = data
        self.grid_wcs = grid_wcs
        self.projection_wcs = projection_wcs
        self.polarisation_frame = polarisation_frame or PolarisationFrame()

    @property
    def shape(self):
        return tuple(self.data.shape)


class ConvolutionFunction:
    """Oversampled kernel [nchan, npol, nz, os, os, S, S] (memory_data_models.py:527-614).

    ``separable`` optionally holds the real 1-D table k[os][S] with
    data[..., yo, xo, a, b] == k[yo, a] * k[xo, b] (the PSWF case); the gridding kernels
    then read 2*S numbers per visibility instead of S*S.
    """

    def __init__(self, data=None, grid_wcs=None, projection_wcs=None, polarisation_frame=None, separable=None):
        self.data = data
        self.grid_wcs = grid_wcs
        self.projection_wcs = projection_wcs
        self.polarisation_frame = polarisation_frame or PolarisationFrame()
        self.separable = separable
        self._cache = {}

    @property
    def shape(self):
        return tuple(self.data.shape)

    @property
    def nz(self):
        return self.data.shape[2]

    @property
    def oversampling(self):
        return self.data.shape[3]

    @property
    def support(self):
        return self.data.shape[5]

    def check_uniform(self):
        """The kernels read one table for every image channel and polarisation; ARL's gridders index
        the table by the row's image channel (gridding.py:123-130, 259-265), so a table whose
        channels or polarisations differ is refused instead of being read as channel 0.  ARL's own
        constructors (kernels.py:19-200) give every channel the same kernel."""
        if "uniform" in self._cache:
            return
        d = torch.as_tensor(self.data)
        if d.shape[0] * d.shape[1] > 1 and not torch.equal(d, d[:1, :1].expand_as(d)):
            raise NotImplementedError("a convolution function whose image channels or polarisations hold "
                                      "different kernels is not supported")
        self._cache["uniform"] = True

    def table(self, precision):
        """Device table in the kernels' layout for the given precision (cached)."""
        key = ("tab", precision)
        if key not in self._cache:
            self.check_uniform()
            cdt = torch.complex64 if precision == 32 else torch.complex128
            self._cache[key] = torch.as_tensor(self.data)[0, 0].to(device=device(), dtype=cdt).contiguous()
        return self._cache[key]

    def separable_table(self, precision):
        key = ("sep", precision)
        if self.separable is None:
            return None
        if key not in self._cache:
            self.check_uniform()
            rdt = torch.float32 if precision == 32 else torch.float64
            self._cache[key] = torch.as_tensor(self.separable, dtype=torch.float64).to(
                device=device(), dtype=rdt).contiguous()
        return self._cache[key]


class Visibility:
    """Row-form visibilities (memory_data_models.py:745-937).

    One row per (time, baseline, channel); uvw in wavelengths.  Columns are held in
    ``data`` as device tensors: uvw f64 [V,3], vis complex [V,npol], weight and
    imaging_weight real [V,npol], frequency f64 [V], time, channel_bandwidth,
    integration_time f64 [V], antenna1/antenna2 int64 [V].  complex64 vis selects the
    single precision path, complex128 the double precision one.
    """

    COLUMNS = ("uvw", "vis", "weight", "imaging_weight", "frequency", "time", "channel_bandwidth",
               "integration_time", "antenna1", "antenna2")

    def __init__(self, data=None, polarisation_frame=None, phasecentre=None):
        self.data = data or {}
        self.polarisation_frame = polarisation_frame or PolarisationFrame()
        self.phasecentre = phasecentre

    @classmethod
    def from_arrays(cls, uvw, vis=None, weight=None, imaging_weight=None, frequency=None, time=None,
                    antenna1=None, antenna2=None, channel_bandwidth=None, integration_time=None,
                    precision=32, dev=None):
        dev = dev or device()
        cdt = torch.complex64 if precision == 32 else torch.complex128
        rdt = torch.float32 if precision == 32 else torch.float64
        uvw = torch.as_tensor(numpy.asarray(uvw, dtype=numpy.float64)).to(dev)
        n = uvw.shape[0]

        def col(x, shape, dt, fill):
            if x is None:
                return torch.full(shape, fill, dtype=dt, device=dev)
            t = torch.as_tensor(numpy.asarray(x)).to(device=dev, dtype=dt)
            return t.reshape(shape).contiguous()

        data = {
            "uvw": uvw.contiguous(),
            "vis": col(vis, (n, 1), cdt, 0.0),
            "weight": col(weight, (n, 1), rdt, 1.0),
            "imaging_weight": col(imaging_weight if imaging_weight is not None else weight, (n, 1), rdt, 1.0),
            "frequency": col(frequency, (n,), torch.float64, 1e8),
            "time": col(time, (n,), torch.float64, 0.0),
            "channel_bandwidth": col(channel_bandwidth, (n,), torch.float64, 1e6),
            "integration_time": col(integration_time, (n,), torch.float64, 1.0),
            "antenna1": col(antenna1, (n,), torch.int64, 0),
            "antenna2": col(antenna2, (n,), torch.int64, 0),
        }
        return cls(data)

    @property
    def nvis(self):
        return self.data["uvw"].shape[0]

    @property
    def npol(self):
        return self.data["vis"].shape[1]

    @property
    def uvw(self):
        return self.data["uvw"]

    @property
    def u(self):
        return self.data["uvw"][:, 0]

    @property
    def v(self):
        return self.data["uvw"][:, 1]

    @property
    def w(self):
        return self.data["uvw"][:, 2]

    @property
    def vis(self):
        return self.data["vis"]

    @property
    def weight(self):
        return self.data["weight"]

    @property
    def imaging_weight(self):
        return self.data["imaging_weight"]

    @property
    def frequency(self):
        return self.data["frequency"]

    @property
    def precision(self):
        return 32 if self.data["vis"].dtype == torch.complex64 else 64

    def copy(self, zero=False):
        """copy_visibility (visibility/base.py:33-47)."""
        new = Visibility({k: v.clone() for k, v in self.data.items()}, self.polarisation_frame, self.phasecentre)
        if zero:
            new.data["vis"].zero_()
        return new

    def with_vis(self, vis):
        """A shallow copy that shares every column but ``vis``."""
        d = dict(self.data)
        d["vis"] = vis
        return Visibility(d, self.polarisation_frame, self.phasecentre)


class BlockVisibility:
    """Correlator-layout visibilities (memory_data_models.py:940-1094).

    vis [ntimes, nants, nants, nchan, npol] complex, uvw [ntimes, nants, nants, 3] in
    metres, frequency [nchan], channel_bandwidth [nchan], time [ntimes].
    """

    COLUMNS = ("vis", "uvw", "weight", "imaging_weight", "time", "integration_time")

    def __init__(self, data=None, frequency=None, channel_bandwidth=None, polarisation_frame=None,
                 phasecentre=None):
        self.data = data or {}
        self.frequency = frequency
        self.channel_bandwidth = channel_bandwidth
        self.polarisation_frame = polarisation_frame or PolarisationFrame()
        self.phasecentre = phasecentre

    @classmethod
    def from_arrays(cls, uvw, vis=None, weight=None, imaging_weight=None, time=None, integration_time=None,
                    frequency=None, channel_bandwidth=None, precision=32, dev=None):
        """Device columns from arrays: uvw [ntimes, nants, nants, 3] metres; vis, weight and
        imaging_weight [ntimes, nants, nants, nchan, 1] (vis 0, weights 1 when omitted)."""
        dev = dev or device()
        cdt = torch.complex64 if precision == 32 else torch.complex128
        rdt = torch.float32 if precision == 32 else torch.float64
        uvw = torch.as_tensor(numpy.asarray(uvw, dtype=numpy.float64)).to(dev).contiguous()
        ntimes, nants = uvw.shape[0], uvw.shape[1]
        frequency = numpy.atleast_1d(numpy.asarray(frequency, dtype=numpy.float64))
        nchan = len(frequency)
        if channel_bandwidth is None:
            channel_bandwidth = numpy.ones(nchan)
        shape = (ntimes, nants, nants, nchan, 1)

        def col(x, dt, fill):
            if x is None:
                return torch.full(shape, fill, dtype=dt, device=dev)
            return torch.as_tensor(numpy.asarray(x)).to(device=dev, dtype=dt).reshape(shape).contiguous()

        t = numpy.zeros(ntimes) if time is None else numpy.asarray(time, dtype=numpy.float64)
        it = numpy.ones(ntimes) if integration_time is None else numpy.asarray(integration_time, dtype=numpy.float64)
        data = {
            "uvw": uvw,
            "vis": col(vis, cdt, 0.0),
            "weight": col(weight, rdt, 1.0),
            "imaging_weight": col(imaging_weight, rdt, 1.0),
            "time": torch.as_tensor(t).to(dev),
            "integration_time": torch.as_tensor(it).to(dev),
        }
        return cls(data, frequency=frequency, channel_bandwidth=numpy.asarray(channel_bandwidth, dtype=numpy.float64))

    @property
    def nants(self):
        return self.data["uvw"].shape[1]

    @property
    def nchan(self):
        return self.data["vis"].shape[3]

    @property
    def ntimes(self):
        return self.data["uvw"].shape[0]

    @property
    def time(self):
        return self.data["time"]

    @property
    def precision(self):
        return 32 if self.data["vis"].dtype == torch.complex64 else 64

    @property
    def vis(self):
        return self.data["vis"]

    @property
    def uvw(self):
        return self.data["uvw"]

    def copy(self, zero=False):
        """copy_visibility (visibility/base.py:33-47) for the block layout."""
        new = BlockVisibility({k: v.clone() for k, v in self.data.items()}, frequency=self.frequency,
                              channel_bandwidth=self.channel_bandwidth, polarisation_frame=self.polarisation_frame,
                              phasecentre=self.phasecentre)
        if zero:
            new.data["vis"].zero_()
        return new


class SkyCoord:
    """An ICRS direction (ra, dec) in radians: the part of astropy's SkyCoord that phase centres
    use here (astropy is not a dependency).  ``ra`` and ``dec`` are floats."""

    def __init__(self, ra, dec):
        self.ra = float(ra)
        self.dec = float(dec)

    @classmethod
    def from_degrees(cls, ra, dec):
        return cls(numpy.deg2rad(ra), numpy.deg2rad(dec))

    def __repr__(self):
        return "SkyCoord(ra=%r, dec=%r)" % (self.ra, self.dec)


def _radec(pos):
    """(ra, dec) in radians of a SkyCoord here, an (ra, dec) pair in radians, or an object with
    astropy's .ra.rad / .dec.rad."""
    if isinstance(pos, SkyCoord):
        return pos.ra, pos.dec
    if hasattr(pos, "ra") and hasattr(pos.ra, "rad"):
        return float(pos.ra.rad), float(pos.dec.rad)
    ra, dec = pos
    return float(ra), float(dec)


class GainTable:
    """Scalar antenna gains (memory_data_models.py:132-234): gain [ntimes, nants, nchan, 1, 1]
    complex, weight the same shape (float64), time and interval [ntimes] (float64), frequency
    [nchan].  The convention is V_obs = g_i conj(g_j) V_model for the [t, i, j] element."""

    def __init__(self, gain, time, interval, weight=None, frequency=None):
        self.gain = gain
        self.time = numpy.asarray(time, dtype=numpy.float64)
        self.interval = numpy.asarray(interval, dtype=numpy.float64)
        self.weight = weight
        self.frequency = frequency

    @property
    def ntimes(self):
        return self.gain.shape[0]

    @property
    def nants(self):
        return self.gain.shape[1]

    @property
    def nchan(self):
        return self.gain.shape[2]


class Skycomponent:
    """Point component at direction cosines (l, m) with flux [nchan, npol]
    (memory_data_models.py:617-689; the direction is given in (l, m) directly)."""

    def __init__(self, l, m, flux):
        self.l = float(l)
        self.m = float(m)
        self.flux = numpy.atleast_2d(numpy.asarray(flux, dtype=float))
