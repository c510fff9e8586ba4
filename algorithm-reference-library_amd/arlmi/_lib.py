"""ctypes binding of libarlmi.so (the HIP implementation).

The library is loaded from this package directory.  There is no fallback: if the
shared object is missing, or no GPU is present when a compute entry point is called,
the call raises.
"""

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libarlmi.so")

ARLMI_OK = 0
ERRORS = {
    -1: "ARLMI_E_ARG",
    -2: "ARLMI_E_UV_OFF_GRID",
    -3: "ARLMI_E_W_RANGE",
    -4: "ARLMI_E_FREQ_MAP",
    -5: "ARLMI_E_ODD_SUPPORT",
    -6: "ARLMI_E_SUBSAMPLE",
    -7: "ARLMI_E_HIP",
    -8: "ARLMI_E_UNSUPPORTED",
}
MODE_2D = 0
MODE_WSTACK = 1
FLAG_DROP_OFF_GRID = 1
FLAG_SPARSE_GRIDS = 2
FLAG_OCCUPIED_TILES = 4
FLAG_PRIVATE_GRIDS = 8
FLAG_DETERMINISTIC = 16
KERNEL_SEPARABLE = 0
KERNEL_TABLE = 1

# Names exported by the C ABI (include/arlmi.h).
EXPORTS = (
    "arlmi_version", "arlmi_last_error", "arlmi_wmax", "arlmi_plan_create", "arlmi_plan_update",
    "arlmi_plan_set_chunk", "arlmi_plan_stats_get", "arlmi_plan_plane_w", "arlmi_plan_slab_boxes", "arlmi_plan_destroy", "arlmi_convolution_mapping",
    "arlmi_grid", "arlmi_grid_add", "arlmi_degrid", "arlmi_batch_create", "arlmi_batch_grid", "arlmi_batch_degrid",
    "arlmi_batch_destroy", "arlmi_wslab_sums", "arlmi_plan_fix_plane_w", "arlmi_plan_or_colmask",
    "arlmi_plan_use_colmask", "arlmi_plan_grid_runs",
    "arlmi_grids_to_image", "arlmi_image_to_grids", "arlmi_fft2d", "arlmi_w_screen", "arlmi_dft",
    "arlmi_grid_weights", "arlmi_reweight", "arlmi_block_to_rows", "arlmi_apply_gaintable", "arlmi_apply_gains_rows", "arlmi_apply_gains_rows_adjoint",
    "arlmi_phaserotate_rows", "arlmi_phaserotate_blocks", "arlmi_hogbom", "arlmi_hogbom_noalloc", "arlmi_hogbom_traffic", "arlmi_solve_gains",
    "arlmi_bda_uvdist", "arlmi_bda_average",
    "arlmi_profile_enable", "arlmi_profile_collect", "arlmi_profile_reset",
    "arlmi_profile_report", "arlmi_profile_only",
)


class ArlmiError(RuntimeError):
    """A libarlmi call failed; ``code`` is the ARLMI_E_* value."""

    def __init__(self, code, message):
        super().__init__("%s: %s" % (ERRORS.get(code, str(code)), message))
        self.code = code
        self.name = ERRORS.get(code, str(code))


class UVOffGridError(ArlmiError, AssertionError):
    """ARL asserts when a visibility maps off the grid (gridding.py:43-46)."""


class Geom(ctypes.Structure):
    _fields_ = [
        ("npixel", ctypes.c_int32), ("nchan", ctypes.c_int32), ("support", ctypes.c_int32),
        ("oversampling", ctypes.c_int32), ("nz", ctypes.c_int32), ("nplanes", ctypes.c_int32),
        ("mode", ctypes.c_int32), ("tile", ctypes.c_int32),
        ("du", ctypes.c_double), ("dv", ctypes.c_double), ("ddu", ctypes.c_double), ("ddv", ctypes.c_double),
        ("crpix_uv", ctypes.c_double), ("crpix_d", ctypes.c_double), ("zstep", ctypes.c_double),
        ("crpix_z", ctypes.c_double), ("wstep_grid", ctypes.c_double), ("crpix_wg", ctypes.c_double),
        ("freq0", ctypes.c_double), ("dfreq", ctypes.c_double), ("cellsize", ctypes.c_double),
    ]


class PlanStats(ctypes.Structure):
    _fields_ = [
        ("nrows", ctypes.c_int64), ("nentries", ctypes.c_int64), ("ngrid_items", ctypes.c_int64),
        ("ndegrid_items", ctypes.c_int64), ("nmember", ctypes.c_int64), ("ndropped", ctypes.c_int64),
        ("norphans", ctypes.c_int64), ("ntiles", ctypes.c_int32), ("tile", ctypes.c_int32),
    ]


class Kernel(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("nz", ctypes.c_int32), ("oversampling", ctypes.c_int32),
        ("support", ctypes.c_int32), ("d_table", ctypes.c_void_p),
    ]


class BatchItem(ctypes.Structure):
    _fields_ = [
        ("plan", ctypes.c_void_p), ("d_vis", ctypes.c_void_p), ("d_weight", ctypes.c_void_p), ("kernel", Kernel),
        ("d_grids", ctypes.c_void_p), ("d_sumwt", ctypes.c_void_p), ("d_vis_out", ctypes.c_void_p),
    ]


class Blocks(ctypes.Structure):
    _fields_ = [
        ("ntimes", ctypes.c_int32), ("nants", ctypes.c_int32), ("nchan", ctypes.c_int32),
        ("reserved", ctypes.c_int32), ("d_vis", ctypes.c_void_p), ("d_uvw", ctypes.c_void_p),
        ("d_weight", ctypes.c_void_p), ("d_imaging_weight", ctypes.c_void_p), ("d_time", ctypes.c_void_p),
        ("d_integration_time", ctypes.c_void_p), ("d_frequency", ctypes.c_void_p),
        ("d_channel_bandwidth", ctypes.c_void_p),
    ]


class Rows(ctypes.Structure):
    _fields_ = [
        ("d_uvw", ctypes.c_void_p), ("d_vis", ctypes.c_void_p), ("d_weight", ctypes.c_void_p),
        ("d_imaging_weight", ctypes.c_void_p), ("d_time", ctypes.c_void_p), ("d_integration_time", ctypes.c_void_p),
        ("d_frequency", ctypes.c_void_p), ("d_channel_bandwidth", ctypes.c_void_p), ("d_antenna1", ctypes.c_void_p),
        ("d_antenna2", ctypes.c_void_p),
    ]


_lib = None


def lib():
    """Load libarlmi.so (once) and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libarlmi.so not found at %s; build it with `python -m arlmi._build` "
                          "(or __graft_entry__.build())" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    pp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "arlmi_version": (ctypes.c_int, []),
        "arlmi_last_error": (ctypes.c_char_p, []),
        "arlmi_wmax": (ctypes.c_int, [vp, i64, vp, ctypes.POINTER(dbl)]),
        "arlmi_plan_create": (ctypes.c_int, [ctypes.POINTER(Geom), vp, vp, i64, ctypes.POINTER(dbl), dbl,
                                             ctypes.c_uint32, vp, pp]),
        "arlmi_plan_update": (ctypes.c_int, [vp, vp, vp, i64, ctypes.POINTER(dbl), dbl, vp]),
        "arlmi_plan_set_chunk": (ctypes.c_int, [vp, i32]),
        "arlmi_plan_stats_get": (ctypes.c_int, [vp, ctypes.POINTER(PlanStats)]),
        "arlmi_plan_grid_runs": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), vp]),
        "arlmi_plan_plane_w": (ctypes.c_int, [vp, ctypes.POINTER(dbl), i32]),
        "arlmi_plan_slab_boxes": (ctypes.c_int, [vp, ctypes.POINTER(dbl), i32]),
        "arlmi_plan_destroy": (None, [vp]),
        "arlmi_convolution_mapping": (ctypes.c_int, [ctypes.POINTER(Geom), vp, vp, i64, vp, vp]),
        "arlmi_grid": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.POINTER(Kernel), vp, vp, ctypes.c_int, vp]),
        "arlmi_batch_create": (ctypes.c_int, [ctypes.POINTER(BatchItem), i32, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_void_p)]),
        "arlmi_batch_grid": (ctypes.c_int, [vp, ctypes.c_int, vp]),
        "arlmi_batch_degrid": (ctypes.c_int, [vp, vp]),
        "arlmi_batch_destroy": (None, [vp]),
        "arlmi_grid_add": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.POINTER(Kernel), vp, vp, ctypes.c_int,
                                          vp]),
        "arlmi_wslab_sums": (ctypes.c_int, [vp, i64, ctypes.POINTER(dbl), dbl, i32, ctypes.POINTER(dbl),
                                            ctypes.POINTER(i64), vp]),
        "arlmi_plan_fix_plane_w": (ctypes.c_int, [vp, ctypes.POINTER(dbl), i32]),
        "arlmi_plan_or_colmask": (ctypes.c_int, [vp, vp, ctypes.POINTER(i32), vp]),
        "arlmi_plan_use_colmask": (ctypes.c_int, [vp, vp]),
        "arlmi_degrid": (ctypes.c_int, [vp, vp, ctypes.POINTER(Kernel), vp, ctypes.c_int, vp]),
        "arlmi_grids_to_image": (ctypes.c_int, [vp, vp, vp, vp, ctypes.c_int, vp, vp, ctypes.c_int, vp]),
        "arlmi_image_to_grids": (ctypes.c_int, [vp, vp, vp, vp, ctypes.c_int, vp]),
        "arlmi_fft2d": (ctypes.c_int, [vp, i32, i64, ctypes.c_int, ctypes.c_int, vp]),
        "arlmi_w_screen": (ctypes.c_int, [i32, dbl, dbl, vp, ctypes.c_int, vp]),
        "arlmi_dft": (ctypes.c_int, [vp, vp, i64, vp, i32, i32, vp, ctypes.c_int, vp]),
        "arlmi_grid_weights": (ctypes.c_int, [ctypes.POINTER(Geom), vp, vp, i64, vp, vp, vp, ctypes.c_int, vp]),
        "arlmi_reweight": (ctypes.c_int, [ctypes.POINTER(Geom), vp, vp, i64, vp, vp, ctypes.c_int, vp]),
        "arlmi_block_to_rows": (ctypes.c_int, [ctypes.POINTER(Blocks), ctypes.POINTER(Rows), vp, ctypes.c_int, vp]),
        "arlmi_apply_gaintable": (ctypes.c_int, [vp, i32, i32, i32, vp, vp, vp, i32, ctypes.c_int, ctypes.c_int, vp]),
        "arlmi_apply_gains_rows": (ctypes.c_int, [i64, vp, vp, vp, i64, vp, vp, ctypes.c_int, vp]),
        "arlmi_apply_gains_rows_adjoint": (ctypes.c_int, [i64, vp, vp, vp, i64, vp, vp, vp, vp, ctypes.c_int, vp]),
        "arlmi_phaserotate_rows": (ctypes.c_int, [vp, vp, i64, i32, vp, vp, ctypes.c_int, ctypes.c_int, vp]),
        "arlmi_phaserotate_blocks": (ctypes.c_int, [vp, vp, i64, i32, i32, vp, vp, vp, ctypes.c_int, ctypes.c_int,
                                                    vp]),
        "arlmi_hogbom": (ctypes.c_int, [vp, vp, vp, i32, i32, vp, i32, i32, dbl, dbl, i32, dbl,
                                        ctypes.POINTER(i32), vp]),
        "arlmi_hogbom_noalloc": (ctypes.c_int, [vp, vp, vp, i32, i32, vp, i32, i32, dbl, dbl, i32, dbl,
                                                ctypes.POINTER(i32), vp]),
        "arlmi_solve_gains": (ctypes.c_int, [vp, vp, vp, i32, i32, i32, vp, i32, vp, vp, vp, vp, vp, vp,
                                             ctypes.c_int, i32, dbl, i32, dbl, ctypes.c_int, ctypes.c_int, vp]),
        "arlmi_bda_uvdist": (ctypes.c_int, [vp, vp, i32, i32, i32, vp, vp, ctypes.c_int, vp]),
        "arlmi_bda_average": (ctypes.c_int, [ctypes.POINTER(Blocks), vp, vp, vp, i64, ctypes.POINTER(Rows), vp,
                                             ctypes.c_int, vp]),
        "arlmi_hogbom_traffic": (ctypes.c_int, [ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "arlmi_profile_enable": (ctypes.c_int, [ctypes.c_int]),
        "arlmi_profile_collect": (ctypes.c_int, []),
        "arlmi_profile_reset": (ctypes.c_int, []),
        "arlmi_profile_report": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
        "arlmi_profile_only": (ctypes.c_int, [ctypes.c_char_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    """Raise ArlmiError (UVOffGridError for off-grid rows) when rc is not ARLMI_OK."""
    if rc != ARLMI_OK:
        msg = lib().arlmi_last_error().decode(errors="replace")
        if rc == -2:
            raise UVOffGridError(rc, msg)
        raise ArlmiError(rc, msg)
    return rc


def profile_report():
    """{kernel: (total_ms, launches)} accumulated since the last reset."""
    L = lib()
    L.arlmi_profile_collect()
    need = L.arlmi_profile_report(None, 0)
    buf = ctypes.create_string_buffer(need + 16)
    L.arlmi_profile_report(buf, need + 16)
    out = {}
    for line in buf.value.decode().splitlines():
        name, ms, cnt = line.split()
        out[name] = (float(ms), int(cnt))
    return out


PROBE_LIB_PATH = os.path.join(_PKG, "libarlmi_probe.so")
_probe = None


def probe_lib():
    """The measurement library (libarlmi_probe.so: the LDS microbenchmark), loaded on first use."""
    global _probe
    if _probe is None:
        if not os.path.exists(PROBE_LIB_PATH):
            raise RuntimeError("%s is not built (run arlmi._build.build())" % PROBE_LIB_PATH)
        p = ctypes.CDLL(PROBE_LIB_PATH)
        p.arlmi_probe_lds_add.restype = ctypes.c_int
        p.arlmi_probe_lds_add.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_float)]
        p.arlmi_probe_last_error.restype = ctypes.c_char_p
        p.arlmi_probe_last_error.argtypes = []
        _probe = p
    return _probe


def probe_check(rc):
    if rc != 0:
        raise ArlmiError(rc, probe_lib().arlmi_probe_last_error().decode(errors="replace"))
    return rc
