"""Python host side of the MI355X VERB-4D step library.

Mirrors the reference's Python module ``verb4d_solver``
(src/python_bindings/Define_python_modules.cpp:17-33): the same function
names and argument lists (``Convection_1D``, ``Convection_2D``,
``Diffusion_2D``, plus ``Diffusion_1D``), each returning the updated PSD
array, and the ``BoundaryConditionType`` enum with the reference's values
(src/BoundaryConditionType.hpp:53-88). Errors the reference turns into
``exit(EXIT_FAILURE)`` raise :class:`VerbError` here.

Every call goes through the C-ABI of ``lib/libverb4d_hip.so`` (declared in
include/verb4d_hip.h) and runs the hand-written gfx950 kernels; there is no
CPU fallback. Loading fails loudly when the library has not been built.
"""
import ctypes
import enum
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libverb4d_hip.so")

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)


class BoundaryConditionType(enum.IntEnum):
    Periodic = 0
    ConstantValue = 1
    ConstantDerivative = 2


Periodic = BoundaryConditionType.Periodic
ConstantValue = BoundaryConditionType.ConstantValue
ConstantDerivative = BoundaryConditionType.ConstantDerivative


class VerbError(RuntimeError):
    """A status != 0 returned by the C-ABI (the reference would exit)."""

    def __init__(self, status, message):
        super().__init__("verb4d status %d: %s" % (status, message))
        self.status = status


class StepOptions(ctypes.Structure):
    """verb_step_options: the time loop's clamp, V-lower boundary from
    convection and negative-point count around the step."""
    _fields_ = [("positive_psd", ctypes.c_int), ("vl_from_convection", ctypes.c_void_p),
                ("negative_points", ctypes.POINTER(ctypes.c_longlong)), ("psd_lost", ctypes.c_void_p),
                ("fast_convection", ctypes.c_int), ("vu_from_convection", ctypes.c_void_p)]


class FieldUpdateParams(ctypes.Structure):
    """verb_field_update_params: scaling and per-axis limits of an update."""
    _fields_ = [("scale", ctypes.c_double), ("grid", ctypes.c_void_p * 4), ("from_", ctypes.c_double * 4),
                ("to", ctypes.c_double * 4)]


class ConvectionOptions(ctypes.Structure):
    """verb_convection_options: SAVE_PSD_LOST_CONV accumulator and the
    FAST_CONVECTION sub-step rule."""
    _fields_ = [("psd_lost", ctypes.c_void_p), ("fast_convection", ctypes.c_int)]


class GmresParams(ctypes.Structure):
    """ctypes mirror of ``verb_gmres_params`` (include/verb4d_hip.h)."""

    _fields_ = (("tol", ctypes.c_double), ("restart", ctypes.c_int), ("max_cycles", ctypes.c_int),
                ("inner_tol", ctypes.c_double), ("max_iterations", ctypes.c_int))


class GmresStats(ctypes.Structure):
    """ctypes mirror of ``verb_gmres_stats``."""

    _fields_ = (("nodes", ctypes.c_int), ("converged", ctypes.c_int), ("fallback", ctypes.c_int),
                ("iterations", ctypes.c_longlong), ("max_iterations", ctypes.c_int),
                ("max_residual", ctypes.c_double), ("mean_log10_residual", ctypes.c_double))

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MixedParams(ctypes.Structure):
    """ctypes mirror of ``verb_mixed_params``."""

    _fields_ = (("tol", ctypes.c_double), ("max_sweeps", ctypes.c_int))


class MixedStats(ctypes.Structure):
    """ctypes mirror of ``verb_mixed_stats``."""

    _fields_ = (("nodes", ctypes.c_int), ("converged", ctypes.c_int), ("max_sweeps_used", ctypes.c_int),
                ("mean_sweeps", ctypes.c_double), ("max_residual", ctypes.c_double))

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


TAB_STRLEN = 1024


class TabLine(ctypes.Structure):
    """verb_tab_line: the first line of a .tab file, split as the reference
    splits it."""
    _fields_ = [("has_line", ctypes.c_int), ("data_kind", ctypes.c_int), ("scale_present", ctypes.c_int),
                ("nlimits", ctypes.c_int), ("data", ctypes.c_char * TAB_STRLEN), ("scale", ctypes.c_char * TAB_STRLEN),
                ("from_", (ctypes.c_char * TAB_STRLEN) * 4), ("to", (ctypes.c_char * TAB_STRLEN) * 4)]


class DaParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("model_error", ctypes.c_double), ("observation_error", ctypes.c_double),
                ("p_index", ctypes.c_int)]


class DaConvParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("correlation_time", ctypes.c_double), ("model_error", ctypes.c_double),
                ("observation_error", ctypes.c_double), ("use_log", ctypes.c_int),
                ("growing_Q_at_boundary", ctypes.c_int), ("recompute_model", ctypes.c_int),
                ("plane_begin", ctypes.c_longlong), ("plane_count", ctypes.c_longlong),
                ("process_noise", ctypes.c_void_p)]


class Tuning(ctypes.Structure):
    """ctypes mirror of ``verb_tuning``."""

    _fields_ = (("convection_tq", ctypes.c_int), ("lu_rb", ctypes.c_int), ("lu_cb", ctypes.c_int),
                ("lu_persistent", ctypes.c_int), ("convection_ms", ctypes.c_double), ("local_ms", ctypes.c_double),
                ("convection_kernel", ctypes.c_int))

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class VerbGrid(ctypes.Structure):
    """ctypes mirror of ``verb_grid`` (include/verb4d_hip.h)."""

    _fields_ = (
        [("nP", ctypes.c_int), ("nR", ctypes.c_int), ("nV", ctypes.c_int), ("nK", ctypes.c_int)]
        + [(n, ctypes.c_void_p) for n in ("Pg", "Rg", "Vg", "Kg", "Lg")]
        + [(n, ctypes.c_void_p) for n in ("DLL", "DVV", "DKK", "DVK", "VP", "VR")]
        + [
            (n, ctypes.c_void_p)
            for n in ("G_local", "G_radial", "G_conv", "sources", "losses_local", "losses_radial", "losses_conv")
        ]
        + [(n, ctypes.c_void_p) for n in ("Pl", "Pu", "Rl", "Ru", "Ll", "Lu", "Vl", "Vu", "Kl", "Ku")]
        + [
            (n + "_type", ctypes.c_int)
            for n in ("Pl", "Pu", "Rl", "Ru", "Ll", "Lu", "Vl", "Vu", "Kl", "Ku")
        ]
        + [("dt", ctypes.c_double), ("sub_dt", ctypes.c_double)]
        + [("radial_sources", ctypes.c_double), ("local_sources", ctypes.c_double)]
        + [("include_boundary", ctypes.c_int)]
    )


FIELDS_4D = ("Pg", "Rg", "Vg", "Kg", "Lg", "DLL", "DVV", "DKK", "DVK", "VP", "VR", "G_local", "G_radial",
             "G_conv", "sources", "losses_local", "losses_radial", "losses_conv")
FIELDS_BC = ("Pl", "Pu", "Rl", "Ru", "Ll", "Lu", "Vl", "Vu", "Kl", "Ku")

_lib_handle = None
_lib_lock = threading.Lock()

# Every symbol include/verb4d_hip.h declares: (name, restype, argtypes).
_P = ctypes.c_void_p
_SIGNATURES = [
    ("verb_ctx_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    ("verb_ctx_destroy", ctypes.c_int, [_P]),
    ("verb_ctx_stream", ctypes.c_void_p, [_P]),
    ("verb_ctx_synchronize", ctypes.c_int, [_P]),
    ("verb_ctx_last_error", ctypes.c_char_p, [_P]),
    ("verb_ctx_scratch_bytes", ctypes.c_size_t, [_P]),
    ("verb_ctx_set_timing", ctypes.c_int, [_P, ctypes.c_int]),
    ("verb_ctx_get_timings", ctypes.c_int, [_P, _dp, ctypes.POINTER(ctypes.c_long)]),
    ("verb_diffusion_2d", ctypes.c_int,
     [_P, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp] + [ctypes.c_int] * 4
     + [_dp] * 7 + [ctypes.c_double, ctypes.c_double, _dp, _i64p, ctypes.c_int]),
    ("verb_diffusion_2d_noalloc", ctypes.c_int,
     [_P, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp] + [ctypes.c_int] * 4
     + [_dp] * 7 + [ctypes.c_double, ctypes.c_double, _dp, _i64p, ctypes.c_int]),
    ("verb_diffusion_1d", ctypes.c_int,
     [_P, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
      _dp, ctypes.c_double]),
    ("verb_convection_2d", ctypes.c_int,
     [_P, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp] + [ctypes.c_int] * 4
     + [_dp] * 5 + [ctypes.c_double]),
    ("verb_convection_2d_ex", ctypes.c_int,
     [_P, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp] + [ctypes.c_int] * 4
     + [_dp] * 5 + [ctypes.c_double, ctypes.POINTER(ConvectionOptions)]),
    ("verb_convection_1d_quickest6", ctypes.c_int,
     [_P, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp,
      ctypes.c_double]),
    ("verb_dgbtrf", ctypes.c_int,
     [_P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _dp, ctypes.c_int64, _i64p]),
    ("verb_dgbtrs", ctypes.c_int,
     [_P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _dp, ctypes.c_int64, _i64p, _dp]),
    ("verb_tridag_batched", ctypes.c_int, [_P, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int]),
    ("verb_convection_batched", ctypes.c_int, [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p]),
    ("verb_convection_batched_ex", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p, ctypes.POINTER(ConvectionOptions)]),
    ("verb_radial_diffusion_batched", ctypes.c_int, [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p]),
    ("verb_local_diffusion_batched", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    ("verb_local_diffusion_uneven", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_int]),
    ("verb_step_uneven", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
      ctypes.c_int]),
    ("verb_local_assembly", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("verb_lu_cache_doubles", ctypes.c_size_t, [ctypes.POINTER(VerbGrid)]),
    ("verb_lu_cache_ints", ctypes.c_size_t, [ctypes.POINTER(VerbGrid)]),
    ("verb_step", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_int]),
    ("verb_check_status", ctypes.c_int, [_P]),
    ("verb_field_update", ctypes.c_int, [_P, _P, _P, ctypes.c_size_t, ctypes.POINTER(FieldUpdateParams)]),
    ("verb_check_stability", ctypes.c_int, [_P, ctypes.POINTER(VerbGrid), ctypes.POINTER(ctypes.c_double)]),
    ("verb_remap_batched", ctypes.c_int,
     [_P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("verb_cubic1d", ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P]),
    ("verb_da_radial", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), _P, _P, _P, _P, ctypes.POINTER(DaParams)]),
    ("verb_da_convection", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), _P, _P, _P, _P, _P, _P, ctypes.POINTER(DaConvParams)]),
    ("verb_plt_write", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _P, _P, ctypes.c_char_p, ctypes.c_char_p]),
    ("verb_plt_write_grid", ctypes.c_int,
     [ctypes.c_char_p, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_char_p), _P,
      ctypes.c_char_p]),
    ("verb_plt_sizes", ctypes.c_int, [ctypes.c_char_p, _P]),
    ("verb_plt_read", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _P, _P]),
    ("verb_plt_read_checked", ctypes.c_int,
     [ctypes.c_char_p, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_void_p), _P, ctypes.POINTER(ctypes.c_longlong)]),
    ("verb_pltb_write", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _P, _P]),
    ("verb_pltb_read", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _P, _P, _P]),
    ("verb_grid_read", ctypes.c_int, [ctypes.c_char_p, _P, _P, _P, _P, _P, ctypes.c_size_t]),
    ("verb_psd_lst_append", ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p]),
    ("verb_tab_read", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(TabLine)]),
    ("verb_is_number", ctypes.c_int, [ctypes.c_char_p]),
    ("verb_parse_number", ctypes.c_double, [ctypes.c_char_p]),
    ("verb_queue_read", ctypes.c_int,
     [ctypes.c_char_p, _P, _P, _P, ctypes.c_size_t, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    ("verb_local_diffusion_gmres", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p, ctypes.POINTER(GmresParams), ctypes.POINTER(GmresStats)]),
    ("verb_local_diffusion_gmres_x0", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(GmresParams),
      ctypes.POINTER(GmresStats)]),
    ("verb_step_gmres", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p, ctypes.POINTER(GmresParams), ctypes.POINTER(GmresStats)]),
    ("verb_local_diffusion_mixed", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p, ctypes.POINTER(MixedParams), ctypes.POINTER(MixedStats)]),
    ("verb_step_mixed", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p, ctypes.POINTER(MixedParams), ctypes.POINTER(MixedStats)]),
    ("verb_set_tuning", ctypes.c_int, [_P, ctypes.POINTER(Tuning)]),
    ("verb_get_tuning", ctypes.c_int, [_P, ctypes.POINTER(Tuning)]),
    ("verb_autotune", ctypes.c_int, [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p, ctypes.POINTER(Tuning)]),
    ("verb_step_graph", ctypes.c_int, [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p, ctypes.c_void_p]),
    ("verb_unet_param_count", ctypes.c_size_t, []),
    ("verb_unet_create", ctypes.c_int,
     [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("verb_unet_destroy", ctypes.c_int, [_P]),
    ("verb_unet_set_params", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_float)]),
    ("verb_unet_get_params", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_float)]),
    ("verb_unet_features", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), _P, ctypes.c_int, ctypes.c_int, _P, _P, _P]),
    ("verb_unet_normalize", ctypes.c_int, [_P, ctypes.POINTER(VerbGrid), _P, ctypes.c_int, ctypes.c_int, _P, _P]),
    ("verb_unet_forward", ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]),
    ("verb_unet_gradients", ctypes.c_int,
     [_P, _P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_float)]),
    ("verb_unet_train_step", ctypes.c_int,
     [_P, _P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
    ("verb_local_diffusion_learned", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), _P, _P, ctypes.POINTER(GmresParams), ctypes.POINTER(GmresStats)]),
    ("verb_step_learned", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), _P, _P, ctypes.POINTER(GmresParams), ctypes.POINTER(GmresStats)]),
    ("verb_step_ex", ctypes.c_int,
     [_P, ctypes.POINTER(VerbGrid), ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(StepOptions)]),
]


def lib():
    """The loaded C-ABI library (raises if it has not been built)."""
    global _lib_handle
    with _lib_lock:
        if _lib_handle is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    "libverb4d_hip.so not found at %s: build it with `python verb-4d_amd/build.py`" % LIB_PATH)
            h = ctypes.CDLL(LIB_PATH)
            for name, res, args in _SIGNATURES:
                f = getattr(h, name)
                f.restype = res
                f.argtypes = args
            _lib_handle = h
        return _lib_handle


def exported_symbols():
    return [s[0] for s in _SIGNATURES]


class Context:
    """Owns a ``verb_ctx`` (one HIP stream + scratch) on the current device."""

    def __init__(self):
        self._lib = lib()
        h = ctypes.c_void_p()
        rc = self._lib.verb_ctx_create(ctypes.byref(h))
        if rc != 0:
            raise VerbError(rc, "verb_ctx_create failed (no HIP device?)")
        self.handle = h

    def check(self, rc):
        if rc != 0:
            msg = self._lib.verb_ctx_last_error(self.handle)
            raise VerbError(rc, msg.decode() if msg else "")
        return rc

    @property
    def stream(self):
        return self._lib.verb_ctx_stream(self.handle)

    def synchronize(self):
        self.check(self._lib.verb_ctx_synchronize(self.handle))

    def check_status(self):
        return self.check(self._lib.verb_check_status(self.handle))

    PHASES = ("convection", "radial", "local_assemble", "local_factor", "local_solve", "local_krylov",
              "local_warmstart")

    def set_timing(self, on=True):
        self.check(self._lib.verb_ctx_set_timing(self.handle, 1 if on else 0))

    def timings(self):
        """{phase: (total ms, launches)} since the last call (synchronises)."""
        ms = (ctypes.c_double * len(self.PHASES))()
        cnt = (ctypes.c_long * len(self.PHASES))()
        self.check(self._lib.verb_ctx_get_timings(self.handle, ms, cnt))
        return {p: (ms[i], cnt[i]) for i, p in enumerate(self.PHASES)}

    def close(self):
        if getattr(self, "handle", None):
            self._lib.verb_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context()
    return _default_ctx


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != tuple(shape):
        a = a.reshape(shape)
    return a


def _ptr(a):
    return a.ctypes.data_as(_dp)


# ---------------------------------------------------------------------------
# Mirrors of the reference's Python module (src/python_bindings)
# ---------------------------------------------------------------------------
def Diffusion_2D(psd, x, y, x_size, y_size, x_LBC, x_UBC, y_LBC, y_UBC, x_LBC_type, x_UBC_type, y_LBC_type,
                 y_UBC_type, Dxx, Dyy, Dxy, Dyx, G, Sources, Losses, dt, sub_dt, lu_cache=None, index_cache=None,
                 recompute_lu=True, ctx=None, _entry="verb_diffusion_2d"):
    """Diffusion_2D (src/Diffusion_2D.h:86-111); returns the updated PSD [x][y].

    With ``lu_cache``/``index_cache`` (float64 ``(2kl+ku+1)*n`` and int64 ``n``
    arrays, kl = ku = x_size+1) it behaves as the LU_CACHING build: the band
    factors are written there when ``recompute_lu`` and reused otherwise."""
    ctx = ctx or default_context()
    nx, ny = int(x_size), int(y_size)
    out = _f64(psd, (nx, ny)).copy()
    arrs = [_f64(a, (nx, ny)) for a in (x, y)]
    bcs = [_f64(x_LBC, (ny,)), _f64(x_UBC, (ny,)), _f64(y_LBC, (nx,)), _f64(y_UBC, (nx,))]
    dxy = _f64(Dxy, (nx, ny))
    dyx = dxy if Dyx is Dxy else _f64(Dyx, (nx, ny))
    coefs = [_f64(Dxx, (nx, ny)), _f64(Dyy, (nx, ny)), dxy, dyx, _f64(G, (nx, ny)), _f64(Sources, (nx, ny)),
             _f64(Losses, (nx, ny))]
    lu_p = ctypes.cast(None, _dp)
    piv_p = ctypes.cast(None, _i64p)
    if lu_cache is not None:
        if lu_cache.dtype != np.float64 or index_cache.dtype != np.int64:
            raise TypeError("lu_cache must be float64 and index_cache int64")
        lu_p = lu_cache.ctypes.data_as(_dp)
        piv_p = index_cache.ctypes.data_as(_i64p)
    rc = getattr(lib(), _entry)(
        ctx.handle, _ptr(out), _ptr(arrs[0]), _ptr(arrs[1]), nx, ny, *[_ptr(b) for b in bcs], int(x_LBC_type),
        int(x_UBC_type), int(y_LBC_type), int(y_UBC_type), *[_ptr(c) for c in coefs], float(dt), float(sub_dt),
        lu_p, piv_p, 1 if recompute_lu else 0)
    ctx.check(rc)
    return out


def Diffusion_2D_noalloc(psd, x, y, x_size, y_size, x_LBC, x_UBC, y_LBC, y_UBC, x_LBC_type, x_UBC_type, y_LBC_type,
                         y_UBC_type, Dxx, Dyy, Dxy, Dyx, G, Sources, Losses, dt, sub_dt, lu_cache=None,
                         index_cache=None, recompute_lu=True, ctx=None):
    """Diffusion_2D through verb_diffusion_2d_noalloc (no device allocation:
    a static per-GPU scratch array); same results as Diffusion_2D."""
    return Diffusion_2D(psd, x, y, x_size, y_size, x_LBC, x_UBC, y_LBC, y_UBC, x_LBC_type, x_UBC_type, y_LBC_type,
                        y_UBC_type, Dxx, Dyy, Dxy, Dyx, G, Sources, Losses, dt, sub_dt, lu_cache, index_cache,
                        recompute_lu, ctx, _entry="verb_diffusion_2d_noalloc")


def Diffusion_1D(psd, x, x_size, x_LBC, x_UBC, x_LBC_type, x_UBC_type, Dxx, G, Sources, Losses, dt, ctx=None):
    """Diffusion_1D (src/Diffusion_1D.h:69-77); returns the updated PSD."""
    ctx = ctx or default_context()
    n = int(x_size)
    out = _f64(psd, (n,)).copy()
    a = [_f64(v, (n,)) for v in (x, Dxx, G, Sources, Losses)]
    rc = lib().verb_diffusion_1d(ctx.handle, _ptr(out), _ptr(a[0]), n, float(x_LBC), float(x_UBC), int(x_LBC_type),
                                 int(x_UBC_type), _ptr(a[1]), _ptr(a[2]), _ptr(a[3]), _ptr(a[4]), float(dt))
    ctx.check(rc)
    return out


def Convection_2D(PSD, P, R, P_size, R_size, P_LBC, P_UBC, R_LBC, R_UBC, P_LBC_type, P_UBC_type, R_LBC_type,
                  R_UBC_type, VP, VR, Sources, Losses, G_conv, dt, ctx=None):
    """Convection_2D (src/Convection_2D.h:98-108); returns the updated PSD [P][R]."""
    ctx = ctx or default_context()
    nP, nR = int(P_size), int(R_size)
    out = _f64(PSD, (nP, nR)).copy()
    two = [_f64(v, (nP, nR)) for v in (P, R, VP, VR, Sources, Losses, G_conv)]
    bc = [_f64(P_LBC, (nR,)), _f64(P_UBC, (nR,)), _f64(R_LBC, (nP,)), _f64(R_UBC, (nP,))]
    rc = lib().verb_convection_2d(ctx.handle, _ptr(out), _ptr(two[0]), _ptr(two[1]), nP, nR, *[_ptr(b) for b in bc],
                                  int(P_LBC_type), int(P_UBC_type), int(R_LBC_type), int(R_UBC_type), _ptr(two[2]),
                                  _ptr(two[3]), _ptr(two[4]), _ptr(two[5]), _ptr(two[6]), float(dt))
    ctx.check(rc)
    return out


def Convection_2D_lost(PSD, PSD_lost, P, R, P_size, R_size, P_LBC, P_UBC, R_LBC, R_UBC, P_LBC_type, P_UBC_type,
                       R_LBC_type, R_UBC_type, VP, VR, Sources, Losses, G_conv, dt, fast_convection=False, ctx=None):
    """Convection_2D of a SAVE_PSD_LOST_CONV build (src/Convection_2D.h:84-96);
    returns (PSD, PSD_lost) with PSD_lost accumulated from the given array.
    fast_convection selects FAST_CONVECTION's sub-step rule."""
    ctx = ctx or default_context()
    nP, nR = int(P_size), int(R_size)
    out = _f64(PSD, (nP, nR)).copy()
    lost = None if PSD_lost is None else _f64(PSD_lost, (nP, nR)).copy()
    two = [_f64(v, (nP, nR)) for v in (P, R, VP, VR, Sources, Losses, G_conv)]
    bc = [_f64(P_LBC, (nR,)), _f64(P_UBC, (nR,)), _f64(R_LBC, (nP,)), _f64(R_UBC, (nP,))]
    opt = ConvectionOptions(lost.ctypes.data if lost is not None else None, 1 if fast_convection else 0)
    rc = lib().verb_convection_2d_ex(ctx.handle, _ptr(out), _ptr(two[0]), _ptr(two[1]), nP, nR,
                                     *[_ptr(b) for b in bc], int(P_LBC_type), int(P_UBC_type), int(R_LBC_type),
                                     int(R_UBC_type), _ptr(two[2]), _ptr(two[3]), _ptr(two[4]), _ptr(two[5]),
                                     _ptr(two[6]), float(dt), ctypes.byref(opt))
    ctx.check(rc)
    return out, lost


def Convection_1D(PSD, x, x_size, x_LBC, x_UBC, x_LBC_type, x_UBC_type, Ux, dt, ctx=None):
    """Convection_1D_ULTIMATE_QUICKEST6 (src/Convection_1D_ULTIMATE_QUICKEST6.h:71-79)."""
    ctx = ctx or default_context()
    n = int(x_size)
    out = _f64(PSD, (n,)).copy()
    xx, ux = _f64(x, (n,)), _f64(Ux, (n,))
    rc = lib().verb_convection_1d_quickest6(ctx.handle, _ptr(out), _ptr(xx), n, float(x_LBC), float(x_UBC),
                                            int(x_LBC_type), int(x_UBC_type), _ptr(ux), float(dt))
    ctx.check(rc)
    return out


def dgbtrf(ab, kl, ku, ctx=None):
    """LAPACK dgbtrf on a (ldab, n) column-major band (as scipy's wrapper).
    Returns (lu, ipiv 1-based, info)."""
    ctx = ctx or default_context()
    ab = np.asfortranarray(ab, dtype=np.float64).copy(order="F")
    ldab, n = ab.shape
    ipiv = np.zeros(n, dtype=np.int64)
    info = lib().verb_dgbtrf(ctx.handle, n, n, int(kl), int(ku), ab.ctypes.data_as(_dp), ldab,
                             ipiv.ctypes.data_as(_i64p))
    if info < 0:
        ctx.check(info)
    return ab, ipiv, int(info)


def dgbtrs(lu, kl, ku, b, ipiv, ctx=None):
    """LAPACK dgbtrs('N', nrhs=1) on dgbtrf's output."""
    ctx = ctx or default_context()
    lu = np.asfortranarray(lu, dtype=np.float64)
    ldab, n = lu.shape
    x = np.ascontiguousarray(b, dtype=np.float64).copy()
    piv = np.ascontiguousarray(ipiv, dtype=np.int64)
    ctx.check(lib().verb_dgbtrs(ctx.handle, n, int(kl), int(ku), lu.ctypes.data_as(_dp), ldab,
                                piv.ctypes.data_as(_i64p), x.ctypes.data_as(_dp)))
    return x


def tridag(a, b, c, r, ctx=None):
    """tridag (src/MatrixSolver.cpp:689-720) on one or many lines (last axis)."""
    ctx = ctx or default_context()
    a, b, c, r = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, r))
    n = a.shape[-1]
    lines = a.size // n
    u = np.empty_like(r)
    ctx.check(lib().verb_tridag_batched(ctx.handle, _ptr(a), _ptr(b), _ptr(c), _ptr(r), _ptr(u), n, lines))
    return u


def Cubic1D(old_grid, values, new_grid, ctx=None):
    """Cubic1D (src/Interpolation.cpp:25-51): values on old_grid interpolated
    to new_grid (same size) by the monotone cubic Hermite interpolator over
    the points (0, 0) and (old_grid[i], values[i]). Like the reference,
    `values` (a float64 numpy array) is replaced in place; it is also
    returned."""
    ctx = ctx or default_context()
    xo = np.ascontiguousarray(old_grid, dtype=np.float64)
    xn = np.ascontiguousarray(new_grid, dtype=np.float64)
    if xo.size != xn.size:
        raise ValueError("Only interpolation to the same grid size is supported.")
    v = values if (isinstance(values, np.ndarray) and values.dtype == np.float64 and values.flags.c_contiguous) \
        else np.ascontiguousarray(values, dtype=np.float64)
    if v.size != xo.size:
        raise ValueError("values and grid sizes differ")
    ctx.check(lib().verb_cubic1d(ctx.handle, xo.size, _ptr(xo), _ptr(v), _ptr(xn)))
    return v
