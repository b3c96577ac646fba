"""ctypes binding of ``libaerojax_hip.so`` (C ABI in ``include/aerojax.h``).

The library is required: every compute entry point of this package goes
through it, and there is no CPU fallback.  A missing or stale library raises
``AeroJaxLibraryError`` at the first call.
"""
from __future__ import annotations

import ctypes
import os
import threading

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libaerojax_hip.so")

AJ_OK = 0
AJ_ERR_INVALID = 1
AJ_ERR_GRID = 2
AJ_ERR_HIP = 3
AJ_ERR_NONFINITE = 4
AJ_ERR_UNSUPPORTED = 5

AJ_F32 = 0
AJ_F64 = 1
AJ_MG_REFERENCE = 0
AJ_MG_TEXTBOOK = 1
AJ_MG_MAX_NORMS = 64
AJ_MG_OPT_ADAPTIVE = 0
AJ_MG_OPT_GRAPH = 1
AJ_MG_OPT_GUESS_TEST = 2
AJ_MG_OPT_RELATIVE = 3
AJ_MG_OPT_TEMPORAL = 4
AJ_MG_OPT_OMEGA = 5
AJ_MG_OPT_CYCLES = 6

# Every symbol the header declares; tests check the library exports them.
EXPORTS = (
    "aj_version", "aj_last_error", "aj_rk_step", "aj_rk_step_rhs", "aj_rk_step_nu_field", "aj_les_smagorinsky",
    "aj_les_smagorinsky_total", "aj_les_dynamic_smagorinsky", "aj_divergence_rhs", "aj_divergence_periodic", "aj_project",
    "aj_project_collocated",
    "aj_flow_diagnostics", "aj_flow_diagnostics_noalloc", "aj_geometry", "aj_batch_geometry", "aj_tune_enable", "aj_tune_reset",
    "aj_tune_report", "aj_tune_load",
    "aj_batch_create", "aj_batch_destroy", "aj_batch_cells",
    "aj_batch_num_instances", "aj_batch_offsets", "aj_batch_set_viscosity",
    "aj_batch_set_option", "aj_batch_step", "aj_batch_predict", "aj_batch_solve",
    "aj_batch_project", "aj_batch_get_stats",
    "aj_mg_create", "aj_mg_create_batched", "aj_mg_batch_count", "aj_mg_destroy",
    "aj_mg_workspace_bytes", "aj_mg_num_levels", "aj_mg_set_option", "aj_poisson_mg",
    "aj_poisson_mg_enqueue", "aj_poisson_mg_complete",
    "aj_mg_get_stats", "aj_mg_get_stats_batched",
    "aj_profile_begin", "aj_profile_set_repeat", "aj_profile_end", "aj_profile_read",
)

PROF_TAGS = ("rk", "rhs", "project", "mg_pre0", "mg_post0", "mg_pre", "mg_post", "mg_coarse",
             "mg_resid")


class AeroJaxLibraryError(RuntimeError):
    """The HIP library is missing or could not be loaded."""


class AeroJaxError(RuntimeError):
    """A C ABI call returned an error code."""

    def __init__(self, code, msg):
        super().__init__(f"[aj error {code}] {msg}")
        self.code = code


class RkParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("nu", ctypes.c_double), ("dx", ctypes.c_double),
                ("dy", ctypes.c_double), ("U_inf", ctypes.c_double),
                ("brinkman_eta", ctypes.c_double), ("nu_hyper_ratio", ctypes.c_double),
                ("slip_walls", ctypes.c_int), ("fast_mode", ctypes.c_int)]


class FlowDiag(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("pressure_residual", "rhs_norm",
                                              "relative_residual", "div_rms", "l2_div",
                                              "fluid_cells", "max_speed", "cfl")]

    def as_dict(self):
        return {n: float(getattr(self, n)) for n, _ in self._fields_}


AJ_OBSTACLE_CYLINDER = 0
AJ_OBSTACLE_NACA = 1


class Obstacle(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("reserved", ctypes.c_int), ("x0", ctypes.c_double),
                ("y0", ctypes.c_double), ("radius", ctypes.c_double), ("chord", ctypes.c_double),
                ("aoa_deg", ctypes.c_double), ("m", ctypes.c_double), ("p", ctypes.c_double),
                ("t", ctypes.c_double), ("eps", ctypes.c_double)]


class BatchStepParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("U_inf", ctypes.c_double),
                ("brinkman_eta", ctypes.c_double), ("tolerance", ctypes.c_double),
                ("v_cycles", ctypes.c_int), ("solver", ctypes.c_int),
                ("smagorinsky_cs", ctypes.c_double)]


class MgStats(ctypes.Structure):
    _fields_ = [("cycles", ctypes.c_int), ("outer", ctypes.c_int), ("converged", ctypes.c_int),
                ("nonfinite", ctypes.c_int), ("norm", ctypes.c_double),
                ("norms", ctypes.c_double * AJ_MG_MAX_NORMS),
                ("guess_norm", ctypes.c_double), ("zero_norm", ctypes.c_double),
                ("used_guess", ctypes.c_int), ("reserved", ctypes.c_int)]

    def as_dict(self, textbook=False):
        # reference: one norm per cycle evaluation (outer = evaluations);
        # textbook: initial residual + one norm per outer iteration
        k = self.outer + 1 if textbook else self.outer
        d = {"cycles": self.cycles, "outer": self.outer, "converged": bool(self.converged),
             "nonfinite": bool(self.nonfinite), "norm": self.norm,
             "norms": [self.norms[i] for i in range(min(k, AJ_MG_MAX_NORMS))]}
        if textbook:
            d.update(guess_norm=self.guess_norm, zero_norm=self.zero_norm,
                     used_guess=bool(self.used_guess))
        return d


_lock = threading.Lock()
_lib = None

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double

_SIGS = {
    "aj_version": (ctypes.c_char_p, []),
    "aj_last_error": (ctypes.c_char_p, []),
    "aj_rk_step": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i,
                        ctypes.POINTER(RkParams), _vp]),
    "aj_rk_step_rhs": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i,
                            ctypes.POINTER(RkParams), _vp]),
    "aj_rk_step_nu_field": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i,
                                 ctypes.POINTER(RkParams), _vp]),
    "aj_les_smagorinsky": (_i, [_i, _vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _vp]),
    "aj_les_smagorinsky_total": (_i, [_i, _vp, _vp, _d, _vp, _vp, _i, _i, _d, _d, _d, _d, _vp]),
    "aj_les_dynamic_smagorinsky": (_i, [_i, _vp, _vp, _d, _vp, _vp, _vp, _i, _i, _d, _d, _d, _d,
                                        _vp]),
    "aj_divergence_rhs": (_i, [_i, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _vp]),
    "aj_divergence_periodic": (_i, [_i, _vp, _vp, _vp, _i, _i, _d, _d, _vp]),
    "aj_project": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _vp]),
    "aj_project_collocated": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _d,
                                   _vp]),
    "aj_flow_diagnostics": (_i, [_i, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d,
                                 ctypes.POINTER(FlowDiag), _vp]),
    "aj_flow_diagnostics_noalloc": (_i, [_i, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d,
                                         ctypes.POINTER(FlowDiag), _vp]),
    "aj_geometry": (_i, [_i, ctypes.POINTER(Obstacle), _i, _i, _d, _d, _vp, _vp, _vp]),
    "aj_batch_geometry": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "aj_tune_enable": (_i, [_i]),
    "aj_tune_reset": (_i, []),
    "aj_tune_report": (_i, [ctypes.c_char_p, _i]),
    "aj_tune_load": (_i, [ctypes.c_char_p, ctypes.POINTER(_i)]),
    "aj_batch_create": (_i, [ctypes.POINTER(_vp), _i, _i, _vp, _vp, _vp, _vp, _vp, _i]),
    "aj_batch_destroy": (_i, [_vp]),
    "aj_batch_cells": (ctypes.c_long, [_vp]),
    "aj_batch_num_instances": (_i, [_vp]),
    "aj_batch_offsets": (_i, [_vp, _vp]),
    "aj_batch_set_viscosity": (_i, [_vp, _vp]),
    "aj_batch_set_option": (_i, [_vp, _i, _i]),
    "aj_batch_step": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                           ctypes.POINTER(BatchStepParams), _vp]),
    "aj_batch_predict": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(BatchStepParams),
                              _vp]),
    "aj_batch_solve": (_i, [_vp, _vp, _vp, _vp, ctypes.POINTER(BatchStepParams), _vp]),
    "aj_batch_project": (_i, [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(BatchStepParams), _vp]),
    "aj_batch_get_stats": (_i, [_vp, _vp, _vp]),
    "aj_mg_create": (_i, [ctypes.POINTER(_vp), _i, _i, _i, _i, _d, _d, _i]),
    "aj_mg_create_batched": (_i, [ctypes.POINTER(_vp), _i, _i, _i, _d, _d, _i, _i]),
    "aj_mg_batch_count": (_i, [_vp]),
    "aj_mg_get_stats_batched": (_i, [_vp, _vp, _vp]),
    "aj_mg_destroy": (_i, [_vp]),
    "aj_mg_workspace_bytes": (ctypes.c_size_t, [_vp]),
    "aj_mg_num_levels": (_i, [_vp]),
    "aj_mg_set_option": (_i, [_vp, _i, _i]),
    "aj_poisson_mg": (_i, [_vp, _vp, _vp, _vp, _i, _d, _vp]),
    "aj_poisson_mg_enqueue": (_i, [_vp, _vp, _vp, _vp, _i, _d, _vp]),
    "aj_poisson_mg_complete": (_i, [_vp, ctypes.POINTER(_i)]),
    "aj_mg_get_stats": (_i, [_vp, ctypes.POINTER(MgStats), _vp]),
    "aj_profile_begin": (_i, [_i]),
    "aj_profile_set_repeat": (_i, [_i]),
    "aj_profile_end": (_i, []),
    "aj_profile_read": (_i, [_i, ctypes.POINTER(_i), ctypes.POINTER(_d), ctypes.POINTER(_d)]),
}


def load(path: str | None = None):
    """Load the library once and bind every exported function."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise AeroJaxLibraryError(
                f"{p} not found: build it with `python -m aerojax_amd.build` "
                "(there is no CPU fallback)")
        try:
            lib = ctypes.CDLL(p)
        except OSError as e:  # pragma: no cover - depends on the machine
            raise AeroJaxLibraryError(f"cannot load {p}: {e}") from e
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc: int, what: str):
    if rc != AJ_OK:
        msg = load().aj_last_error().decode(errors="replace")
        if rc == AJ_ERR_GRID:
            raise ValueError(msg)
        raise AeroJaxError(rc, f"{what}: {msg}")


def profile_begin(max_records: int = 20000, repeat: int = 1):
    """Start live event timing; ``repeat`` launches each tunable kernel that
    many times back to back inside its event pair (counted as launches)."""
    check(load().aj_profile_set_repeat(int(repeat)), "aj_profile_set_repeat")
    check(load().aj_profile_begin(int(max_records)), "aj_profile_begin")


def profile_end():
    check(load().aj_profile_end(), "aj_profile_end")


def profile_read() -> dict:
    """{tag: {"count", "ms", "bytes", "avg_us", "gbps"}} of the recorded launches."""
    out = {}
    lib = load()
    for k, name in enumerate(PROF_TAGS):
        c, ms, by = _i(), _d(), _d()
        check(lib.aj_profile_read(k, ctypes.byref(c), ctypes.byref(ms), ctypes.byref(by)),
              "aj_profile_read")
        if c.value:
            out[name] = {"count": c.value, "ms": ms.value, "bytes": by.value,
                         "avg_us": 1e3 * ms.value / c.value,
                         "gbps": by.value / (ms.value * 1e-3) / 1e9 if ms.value > 0 else None}
    return out


def version() -> str:
    return load().aj_version().decode()


_NN_SIGS = {
    "aj_nn_create": (_i, [ctypes.POINTER(_vp), _i, _i, _i]),
    "aj_nn_destroy": (_i, [_vp]),
    "aj_nn_num_params": (ctypes.c_long, [_vp]),
    "aj_nn_get_params": (_i, [_vp, _vp, ctypes.c_long, _i]),
    "aj_nn_set_params": (_i, [_vp, _vp, ctypes.c_long]),
    "aj_nn_forward": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "aj_nn_loss_grad": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp]),
    "aj_nn_adam": (_i, [_vp, ctypes.c_float, ctypes.c_float, _vp]),
    "aj_nn_train_step": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, ctypes.c_float, _vp, _vp]),
    "aj_nn_make_sample": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, ctypes.c_float, ctypes.c_float,
                               _vp, _vp, _vp]),
    "aj_nn_warm_start": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, ctypes.c_float, ctypes.c_float,
                              _vp]),
    "aj_nn_warm_start_batched": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, ctypes.c_float,
                                      ctypes.c_float, _vp]),
    "aj_nn_set_option": (_i, [_vp, _i, _i]),
}
AJ_NN_OPT_FUSED_IO = 0
_SIGS.update(_NN_SIGS)
EXPORTS = EXPORTS + tuple(_NN_SIGS)
PROF_TAGS = PROF_TAGS + ("nn_conv", "nn_dw", "les")


def tune_report() -> str:
    """Autotuner decisions of this process, one "key -> pick [ms ...]" line
    per tuned launch."""
    buf = ctypes.create_string_buffer(1 << 16)
    load().aj_tune_report(buf, len(buf))
    return buf.value.decode()


def tune_load(text: str) -> int:
    """Preset autotuner decisions from ``tune_report`` text; returns how many
    were read."""
    n = _i(0)
    check(load().aj_tune_load(text.encode(), ctypes.byref(n)), "aj_tune_load")
    return n.value
