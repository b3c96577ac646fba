"""RK predictor with the reference signature.

Mirrors ``advection_schemes.rk3_simple_new.rk_step_unified``
(advection_schemes/rk3_simple_new.py:175-358): same arguments, same meaning,
same return value ``(u_star, v_star)``; the arithmetic runs in the fused HIP
kernel ``aj_rk_step``.
"""
from __future__ import annotations

import ctypes

import torch

from . import _arrays as A
from . import _lib


def rk_step_unified(u, v, dt: float, nu: float, dx: float, dy: float, mask, sdf=None,
                    U_inf: float = 1.0, nu_sgs=None, nu_hyper_ratio: float = 0.0,
                    slip_walls: bool = True, fast_mode: bool = False,
                    brinkman_eta: float = 0.01, flow_type: str = "von_karman", *,
                    out=None):
    """RK3 (or RK2 with ``fast_mode``) predictor with implicit Brinkman
    penalisation, hard mask and von Karman BCs.

    ``u, v, mask, sdf, nu_sgs``: (nx, ny) torch CUDA tensors (float32 or
    float64) or NumPy arrays (then NumPy is returned).  ``nu`` is a scalar,
    or a field as the GUI step passes it (nu + nu_sgs,
    solver/simulation/step_handlers.py:76-90): the Laplacian then uses
    nu + nu_sgs per cell and the biharmonic term nu * nu_hyper_ratio * dx dy
    per cell, as the reference's array arithmetic does.  ``out``: optional
    pair of preallocated output tensors.
    """
    if flow_type != "von_karman":
        raise NotImplementedError(f"flow_type={flow_type!r}: only 'von_karman' is implemented")
    A.require_gpu()
    want_np = A.is_numpy(u)
    dt_ = A.torch_dtype_of(u)
    ut = A.as_device(u, dt_, name="u")
    dev = ut.device
    vt = A.as_device(v, dt_, dev, "v")
    mt = A.as_device(mask, dt_, dev, "mask")
    st = A.as_device(sdf, dt_, dev, "sdf")
    nt = A.as_device(nu_sgs, dt_, dev, "nu_sgs") if isinstance(nu_sgs, (torch.Tensor,)) or (
        nu_sgs is not None and not isinstance(nu_sgs, (int, float))) else None
    field = not isinstance(nu, (int, float))
    nf = A.as_device(nu, dt_, dev, "nu") if field else None
    nu_total = 0.0 if field else float(nu)
    if nu_sgs is not None and isinstance(nu_sgs, (int, float)):
        if field:
            raise ValueError("a scalar nu_sgs with a field nu is not supported")
        nu_total = float(nu) + float(nu_sgs)
    nx, ny = ut.shape
    for name, t in (("v", vt), ("mask", mt), ("sdf", st), ("nu_sgs", nt), ("nu", nf)):
        if t is not None and tuple(t.shape) != (nx, ny):
            raise ValueError(f"{name} shape {tuple(t.shape)} != u shape {(nx, ny)}")
    if out is None:
        uo = torch.empty_like(ut)
        vo = torch.empty_like(ut)
    else:
        uo, vo = out
    prm = _lib.RkParams(float(dt), nu_total, float(dx), float(dy), float(U_inf),
                        float(brinkman_eta), float(nu_hyper_ratio), int(bool(slip_walls)),
                        int(bool(fast_mode)))
    lib = _lib.load()
    with torch.cuda.device(dev):
        if field:
            rc = lib.aj_rk_step_nu_field(A.dtype_code(ut), A.ptr(ut), A.ptr(vt), A.ptr(mt),
                                         A.ptr(st), A.ptr(nf), A.ptr(nt), A.ptr(uo), A.ptr(vo),
                                         nx, ny, ctypes.byref(prm), A.stream_handle(dev))
        else:
            rc = lib.aj_rk_step(A.dtype_code(ut), A.ptr(ut), A.ptr(vt), A.ptr(mt), A.ptr(st),
                                A.ptr(nt), A.ptr(uo), A.ptr(vo), nx, ny, ctypes.byref(prm),
                                A.stream_handle(dev))
    _lib.check(rc, "aj_rk_step")
    return A.out_like(uo, want_np), A.out_like(vo, want_np)
