"""Field operators of the step, each one HIP launch.

Mirrors solver/operators.py (``divergence_nonperiodic`` :81-96),
solver/les_models.py (``constant_smagorinsky`` :179-200) and the projection
tail of solver/pure_functions/step.py (:257-264, 305-349, 364-377, 394-397).
"""
from __future__ import annotations

import torch

from . import _arrays as A
from . import _lib


def _prep(first, *others, names=()):
    A.require_gpu()
    want_np = A.is_numpy(first)
    dt_ = A.torch_dtype_of(first)
    t0 = A.as_device(first, dt_, name=names[0] if names else "field")
    ts = [A.as_device(o, dt_, t0.device, n) for o, n in zip(others, names[1:])]
    for t, n in zip(ts, names[1:]):
        if t is not None and t.shape != t0.shape:
            raise ValueError(f"{n} shape {tuple(t.shape)} != {tuple(t0.shape)}")
    return want_np, t0, ts


def divergence_nonperiodic(u, v, dx: float, dy: float, *, out=None):
    """Central differences on the interior, zero on every boundary cell."""
    want_np, ut, (vt,) = _prep(u, v, names=("u", "v"))
    nx, ny = ut.shape
    o = torch.empty_like(ut) if out is None else out
    with torch.cuda.device(ut.device):
        rc = _lib.load().aj_divergence_rhs(A.dtype_code(ut), A.ptr(ut), A.ptr(vt), None,
                                           A.ptr(o), nx, ny, float(dx), float(dy), 1.0,
                                           A.stream_handle(ut.device))
    _lib.check(rc, "aj_divergence_rhs")
    return A.out_like(o, want_np)


def divergence(u, v, dx: float, dy: float, *, out=None):
    """Periodic divergence (solver/operators.py:76-78)."""
    want_np, ut, (vt,) = _prep(u, v, names=("u", "v"))
    nx, ny = ut.shape
    o = torch.empty_like(ut) if out is None else out
    with torch.cuda.device(ut.device):
        rc = _lib.load().aj_divergence_periodic(A.dtype_code(ut), A.ptr(ut), A.ptr(vt), A.ptr(o),
                                                nx, ny, float(dx), float(dy),
                                                A.stream_handle(ut.device))
    _lib.check(rc, "aj_divergence_periodic")
    return A.out_like(o, want_np)


def pressure_rhs(u_star, v_star, mask, dt: float, dx: float, dy: float, *, out=None):
    """rhs = divergence_nonperiodic(u*, v*) * mask / dt (step.py:194-199);
    ``mask=None`` gives div / dt (step_handlers.py:92-98)."""
    want_np, ut, (vt, mt) = _prep(u_star, v_star, mask, names=("u_star", "v_star", "mask"))
    nx, ny = ut.shape
    o = torch.empty_like(ut) if out is None else out
    with torch.cuda.device(ut.device):
        rc = _lib.load().aj_divergence_rhs(A.dtype_code(ut), A.ptr(ut), A.ptr(vt), A.ptr(mt),
                                           A.ptr(o), nx, ny, float(dx), float(dy), float(dt),
                                           A.stream_handle(ut.device))
    _lib.check(rc, "aj_divergence_rhs")
    return A.out_like(o, want_np)


def constant_smagorinsky(u, v, dx: float, dy: float, delta: float, C_s: float = 0.1, *,
                         out=None):
    """nu_sgs = (C_s delta)^2 sqrt(2(Sxx^2 + Syy^2 + 2 Sxy^2) + 1e-10) with
    periodic central differences (the reference uses jnp.roll even for the
    von Karman channel)."""
    want_np, ut, (vt,) = _prep(u, v, names=("u", "v"))
    nx, ny = ut.shape
    o = torch.empty_like(ut) if out is None else out
    with torch.cuda.device(ut.device):
        rc = _lib.load().aj_les_smagorinsky(A.dtype_code(ut), A.ptr(ut), A.ptr(vt), A.ptr(o), nx,
                                            ny, float(dx), float(dy), float(delta), float(C_s),
                                            A.stream_handle(ut.device))
    _lib.check(rc, "aj_les_smagorinsky")
    return A.out_like(o, want_np)


def project(u_star, v_star, p, mask, dt: float, dx: float, dy: float, U_inf: float, *,
            out=None):
    """Pressure correction with interface-tangential projection and BCs."""
    want_np, ut, (vt, pt, mt) = _prep(u_star, v_star, p, mask,
                                      names=("u_star", "v_star", "p", "mask"))
    nx, ny = ut.shape
    if out is None:
        uo, vo = torch.empty_like(ut), torch.empty_like(ut)
    else:
        uo, vo = out
    with torch.cuda.device(ut.device):
        rc = _lib.load().aj_project(A.dtype_code(ut), A.ptr(ut), A.ptr(vt), A.ptr(pt), A.ptr(mt),
                                    A.ptr(uo), A.ptr(vo), nx, ny, float(dt), float(dx), float(dy),
                                    float(U_inf), A.stream_handle(ut.device))
    _lib.check(rc, "aj_project")
    return A.out_like(uo, want_np), A.out_like(vo, want_np)


def smagorinsky_total(u, v, nu: float, dx: float, dy: float, delta: float, C_s: float = 0.1, *,
                      out=None):
    """``(nu_sgs, nu + nu_sgs)`` in one launch: the LES block of the GUI step
    (solver/simulation/step_handlers.py:67-77, constant model)."""
    want_np, ut, (vt,) = _prep(u, v, names=("u", "v"))
    nx, ny = ut.shape
    if out is None:
        s, t = torch.empty_like(ut), torch.empty_like(ut)
    else:
        s, t = out
    with torch.cuda.device(ut.device):
        rc = _lib.load().aj_les_smagorinsky_total(A.dtype_code(ut), A.ptr(ut), A.ptr(vt),
                                                  float(nu), A.ptr(s), A.ptr(t), nx, ny,
                                                  float(dx), float(dy), float(delta), float(C_s),
                                                  A.stream_handle(ut.device))
    _lib.check(rc, "aj_les_smagorinsky_total")
    return A.out_like(s, want_np), A.out_like(t, want_np)


def dynamic_smagorinsky(u, v, dx: float, dy: float, delta: float, alpha: float = 2.0, *,
                        nu: float = None):
    """``(nu_sgs, C_d)`` of the dynamic Smagorinsky model
    (solver/les_models.py:78-176, collocated).  With ``nu`` given, returns
    ``(nu_sgs, nu + nu_sgs, C_d)`` (the GUI step's viscosity) from the same
    launches."""
    want_np, ut, (vt,) = _prep(u, v, names=("u", "v"))
    nx, ny = ut.shape
    s, c = torch.empty_like(ut), torch.empty_like(ut)
    t = torch.empty_like(ut) if nu is not None else None
    with torch.cuda.device(ut.device):
        rc = _lib.load().aj_les_dynamic_smagorinsky(
            A.dtype_code(ut), A.ptr(ut), A.ptr(vt), float(nu or 0.0), A.ptr(s), A.ptr(t), A.ptr(c),
            nx, ny, float(dx), float(dy), float(delta), float(alpha), A.stream_handle(ut.device))
    _lib.check(rc, "aj_les_dynamic_smagorinsky")
    if nu is None:
        return A.out_like(s, want_np), A.out_like(c, want_np)
    return A.out_like(s, want_np), A.out_like(t, want_np), A.out_like(c, want_np)


def project_collocated(u_star, v_star, p, mask, dt: float, dx: float, dy: float, U_in: float, *,
                       out=None):
    """Pressure correction and BCs of the GUI step (step_handlers.py:131-176):
    no interface projection, dp/dx = 0 on the inlet row, inlet velocity
    ``U_in`` on rows 1..ny-2 of the inlet, no-slip walls, hard mask."""
    want_np, ut, (vt, pt, mt) = _prep(u_star, v_star, p, mask,
                                      names=("u_star", "v_star", "p", "mask"))
    nx, ny = ut.shape
    if out is None:
        uo, vo = torch.empty_like(ut), torch.empty_like(ut)
    else:
        uo, vo = out
    with torch.cuda.device(ut.device):
        rc = _lib.load().aj_project_collocated(A.dtype_code(ut), A.ptr(ut), A.ptr(vt), A.ptr(pt),
                                               A.ptr(mt), A.ptr(uo), A.ptr(vo), nx, ny, float(dt),
                                               float(dx), float(dy), float(U_in),
                                               A.stream_handle(ut.device))
    _lib.check(rc, "aj_project_collocated")
    return A.out_like(uo, want_np), A.out_like(vo, want_np)
