"""Flow diagnostics of a step (one fused HIP reduction).

Mirrors the debug/viewer metrics of the reference:
``pressure_residual`` / ``relative_residual`` (viewer/simulation_controller.py
:107-121, solver/pure_functions/step.py:229-235), ``div_rms`` / ``l2_div``
over the pure fluid (simulation_controller.py:186-194) and ``max_speed`` /
``cfl`` of check_cfl (advection_schemes/utils.py:14-20).  The multigrid's own
residual norm is ``MultigridPlan.stats()["norm"]``.
"""
from __future__ import annotations

import ctypes

import torch

from . import _arrays as A
from . import _lib


def flow_diagnostics(u, v, p, mask, dx: float, dy: float, dt: float) -> dict:
    """Dict with pressure_residual, rhs_norm, relative_residual, div_rms,
    l2_div, fluid_cells, max_speed, cfl (host floats; synchronises the
    stream)."""
    A.require_gpu()
    dt_ = A.torch_dtype_of(u)
    ut = A.as_device(u, dt_, name="u")
    vt, pt, mt = (A.as_device(x, dt_, ut.device, n) for x, n in ((v, "v"), (p, "p"), (mask, "mask")))
    for n, t in (("v", vt), ("p", pt), ("mask", mt)):
        if tuple(t.shape) != tuple(ut.shape):
            raise ValueError(f"{n} shape {tuple(t.shape)} != u shape {tuple(ut.shape)}")
    nx, ny = ut.shape
    out = _lib.FlowDiag()
    with torch.cuda.device(ut.device):
        rc = _lib.load().aj_flow_diagnostics(A.dtype_code(ut), A.ptr(ut), A.ptr(vt), A.ptr(pt),
                                             A.ptr(mt), nx, ny, float(dx), float(dy), float(dt),
                                             ctypes.byref(out), A.stream_handle(ut.device))
    _lib.check(rc, "aj_flow_diagnostics")
    return out.as_dict()
