"""Phase rotation of visibilities to a new phase centre on the GPU
(processing_components/visibility/base.py:295-381).

The direction cosines of the new centre and the trigonometric factors of the uvw rotation are a
handful of scalars computed here in float64 with numpy, as ARL computes them; every row (or
block element and channel) is then rotated by one HIP kernel (csrc/vis.hip), which evaluates
simulate_point's phasor (coordinate_support.py:192-212) and the uvw_to_xyz / xyz_to_uvw pair
(coordinate_support.py:59-113) in ARL's operation order.
"""

import ctypes

import numpy
import torch

from . import _lib
from .data_models import BlockVisibility, SkyCoord, Visibility, _radec
from .plan import ptr, stream_handle

__all__ = ["skycoord_to_lmn", "separation", "phaserotate_visibility", "SkyCoord"]


def _rotation(angle, axis):
    """The coordinate-frame rotation matrix of ``angle`` about ``axis`` (astropy's
    matrix_utilities.rotation_matrix convention)."""
    c, s = numpy.cos(angle), numpy.sin(angle)
    if axis == "z":
        return numpy.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]])
    if axis == "y":
        return numpy.array([[c, 0.0, -s], [0.0, 1.0, 0.0], [s, 0.0, c]])
    raise ValueError(axis)


def skycoord_to_lmn(pos, phasecentre):
    """(l, m, n) of ``pos`` relative to ``phasecentre`` (coordinate_support.py:150-169): the
    Cartesian coordinates (x, y, z) of ``pos`` in the sky-offset frame whose origin is the phase
    centre (x towards it, y east, z north), returned as (y, z, x - 1).  The frame is reached as
    astropy's SkyOffsetFrame reaches it: rotate by the centre's ra about z, then by -dec about y."""
    ra, dec = _radec(pos)
    ra0, dec0 = _radec(phasecentre)
    cart = numpy.array([numpy.cos(dec) * numpy.cos(ra), numpy.cos(dec) * numpy.sin(ra), numpy.sin(dec)])
    x, y, z = _rotation(-dec0, "y") @ (_rotation(ra0, "z") @ cart)
    return numpy.float64(y), numpy.float64(z), numpy.float64(x - 1.0)


def separation(a, b):
    """Angular distance in radians between two directions, as astropy's SkyCoord.separation
    computes it (the Vincenty formula of angle_utilities.angular_separation)."""
    ra1, dec1 = _radec(a)
    ra2, dec2 = _radec(b)
    dl = ra2 - ra1
    n1 = numpy.cos(dec2) * numpy.sin(dl)
    n2 = numpy.cos(dec1) * numpy.sin(dec2) - numpy.sin(dec1) * numpy.cos(dec2) * numpy.cos(dl)
    d = numpy.sin(dec1) * numpy.sin(dec2) + numpy.cos(dec1) * numpy.cos(dec2) * numpy.cos(dl)
    return float(numpy.arctan2(numpy.hypot(n1, n2), d))


def rotation_lmn(newphasecentre, phasecentre):
    """[l, m, sqrt(1 - l^2 - m^2) - 1] (float64) of the new centre seen from the old one, the
    direction whose phasor the rotation applies, or None when |n| < 1e-15 (no rotation is made,
    visibility/base.py:311)."""
    l, m, n = skycoord_to_lmn(newphasecentre, phasecentre)
    if numpy.abs(n) < 1e-15:
        return None
    return numpy.array([l, m, numpy.sqrt(1 - l ** 2 - m ** 2) - 1.0], dtype=numpy.float64)


def phaserotate_visibility(vis, newphasecentre, tangent=True, inverse=False):
    """Phase rotate from the current phase centre to ``newphasecentre`` (base.py:295-381).

    Returns ``vis`` itself when |n| < 1e-15, else a rotated copy: visibilities multiplied by the
    conjugate of the point-source phasor of the new centre (by the phasor for ``inverse``),
    evaluated on the uvw before rotation; with ``tangent=False`` the uvw are also rotated into the
    new centre's frame and the copy's phase centre is set to the new one.  Works on Visibility
    (uvw in wavelengths) and BlockVisibility (uvw in metres, phasor per channel)."""
    if not isinstance(vis, (Visibility, BlockVisibility)):
        raise ValueError("vis argument neither Visibility or BlockVisibility")
    if vis.phasecentre is None:
        raise ValueError("phaserotate_visibility: the visibility has no phase centre")
    lmn = rotation_lmn(newphasecentre, vis.phasecentre)
    if lmn is None:
        return vis
    newvis = vis.copy()
    trig = None
    if not tangent:
        ra0, dec0 = _radec(vis.phasecentre)
        ra1, dec1 = _radec(newphasecentre)
        ha0, ha1 = -ra0, -ra1
        trig = numpy.array([numpy.cos(ha0), numpy.sin(ha0), numpy.cos(dec0), numpy.sin(dec0),
                            numpy.cos(ha1), numpy.sin(ha1), numpy.cos(dec1), numpy.sin(dec1)], dtype=numpy.float64)
    lib = _lib.lib()
    h_lmn = lmn.ctypes.data_as(ctypes.c_void_p)
    h_trig = None if trig is None else trig.ctypes.data_as(ctypes.c_void_p)
    uvw = newvis.data["uvw"]
    visd = newvis.data["vis"]
    if uvw.dtype != torch.float64 or not uvw.is_contiguous() or not visd.is_contiguous():
        raise ValueError("phaserotate_visibility: uvw must be contiguous float64 and vis contiguous")
    if isinstance(vis, Visibility):
        npol = visd.shape[1] if visd.dim() > 1 else 1
        _lib.check(lib.arlmi_phaserotate_rows(ptr(uvw), ptr(visd), visd.shape[0], npol, h_lmn, h_trig,
                                              int(bool(inverse)), newvis.precision, stream_handle()))
    else:
        nt, na, nb, nchan, npol = visd.shape
        freq = torch.as_tensor(numpy.asarray(newvis.frequency, dtype=numpy.float64)).to(visd.device)
        _lib.check(lib.arlmi_phaserotate_blocks(ptr(uvw), ptr(visd), nt * na * nb, nchan, npol, ptr(freq), h_lmn,
                                                h_trig, int(bool(inverse)), newvis.precision, stream_handle()))
    if not tangent:
        newvis.phasecentre = newphasecentre
    return newvis
