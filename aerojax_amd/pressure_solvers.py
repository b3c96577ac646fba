"""Pressure Poisson solvers with the reference signature.

``poisson_multigrid`` mirrors pressure_solvers/multigrid_solver.py:55-241
(same arguments, defaults, return value and ``ValueError`` for grids the
reference rejects), plus a keyword-only ``p0`` initial guess (default None
= zero, the reference behaviour).  ``poisson_multigrid_to_tolerance`` is this
project's textbook solver (FP32 V-cycles in FP64 defect correction) for the
same outlet-Dirichlet problem, used where ``||div|| <= tolerance`` must be
reached.
"""
from __future__ import annotations

import atexit
import ctypes
import threading

import torch

from . import _arrays as A
from . import _lib

_plans: dict = {}
_plans_lock = threading.Lock()


class MultigridPlan:
    """Owns one ``aj_mg_plan`` (level hierarchy workspace) on one device.
    ``count`` > 1 (textbook only): that many independent instances of the
    grid, stored one after the other in rhs / p0 / p_out, solved together."""

    def __init__(self, nx, ny, dx, dy, dtype=torch.float32, variant="reference", levels=4,
                 device=None, count=1):
        A.require_gpu()
        self.nx, self.ny, self.dx, self.dy = int(nx), int(ny), float(dx), float(dy)
        self.dtype = dtype
        self.variant = variant
        self.levels = int(levels)
        self.count = int(count)
        self.device = torch.device("cuda" if device is None else device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        code = _lib.AJ_F32 if dtype == torch.float32 else _lib.AJ_F64
        var = _lib.AJ_MG_REFERENCE if variant == "reference" else _lib.AJ_MG_TEXTBOOK
        h = ctypes.c_void_p()
        lib = _lib.load()
        with torch.cuda.device(self.device):
            if self.count != 1:
                if variant != "textbook":
                    raise ValueError("batched plans (count > 1) are textbook plans")
                rc = lib.aj_mg_create_batched(ctypes.byref(h), code, self.nx, self.ny, self.dx,
                                              self.dy, self.levels, self.count)
            else:
                rc = lib.aj_mg_create(ctypes.byref(h), code, var, self.nx, self.ny, self.dx,
                                      self.dy, self.levels)
        _lib.check(rc, "aj_mg_create")
        self.handle = h

    @property
    def workspace_bytes(self) -> int:
        return int(_lib.load().aj_mg_workspace_bytes(self.handle))

    def set_option(self, adaptive: bool | None = None, graph: bool | None = None,
                   guess_test: bool | None = None, relative: bool | None = None,
                   temporal: bool | None = None, omega: float | None = None,
                   cycles_per_outer: int | None = None):
        """Textbook plans: ``adaptive`` reads convergence between outer
        iterations (host sync, default) instead of enqueueing all
        ``max_cycles``; ``graph`` replays a captured outer iteration
        (default) instead of launching its kernels one by one;
        ``guess_test`` keeps an initial guess only if its residual is not
        above the zero guess's (default) instead of keeping any finite one;
        ``relative`` makes the tolerance relative to ||b||; ``temporal``
        starts solves without p0 from the previous FP64 results (previous
        solution, then the linear extrapolation of the last two), setting it
        clears that history; ``omega`` sets the over-relaxation of the
        red-black smoother (default 1.2, 1.0 = Gauss-Seidel); ``cycles_per_outer``
        sets the FP32 V-cycles per FP64 defect-correction iteration (default
        2, 1..8)."""
        lib = _lib.load()
        if cycles_per_outer is not None:
            _lib.check(lib.aj_mg_set_option(self.handle, _lib.AJ_MG_OPT_CYCLES,
                                            int(cycles_per_outer)), "aj_mg_set_option")
        if omega is not None:
            _lib.check(lib.aj_mg_set_option(self.handle, _lib.AJ_MG_OPT_OMEGA,
                                            int(round(1000 * float(omega)))), "aj_mg_set_option")
        if temporal is not None:
            _lib.check(lib.aj_mg_set_option(self.handle, _lib.AJ_MG_OPT_TEMPORAL, int(temporal)),
                       "aj_mg_set_option")
        if relative is not None:
            _lib.check(lib.aj_mg_set_option(self.handle, _lib.AJ_MG_OPT_RELATIVE, int(relative)),
                       "aj_mg_set_option")
        if guess_test is not None:
            _lib.check(lib.aj_mg_set_option(self.handle, _lib.AJ_MG_OPT_GUESS_TEST,
                                            int(guess_test)), "aj_mg_set_option")
        if adaptive is not None:
            _lib.check(lib.aj_mg_set_option(self.handle, _lib.AJ_MG_OPT_ADAPTIVE, int(adaptive)),
                       "aj_mg_set_option")
        if graph is not None:
            _lib.check(lib.aj_mg_set_option(self.handle, _lib.AJ_MG_OPT_GRAPH, int(graph)),
                       "aj_mg_set_option")

    @property
    def num_levels(self) -> int:
        return int(_lib.load().aj_mg_num_levels(self.handle))

    def solve(self, rhs: torch.Tensor, p_out: torch.Tensor, p0=None, cycles=5, tolerance=1e-6):
        lib = _lib.load()
        with torch.cuda.device(self.device):
            rc = lib.aj_poisson_mg(self.handle, A.ptr(rhs), A.ptr(p0), A.ptr(p_out), int(cycles),
                                   float(tolerance), A.stream_handle(self.device))
        _lib.check(rc, "aj_poisson_mg")
        return p_out

    def stats(self) -> dict:
        s = _lib.MgStats()
        with torch.cuda.device(self.device):
            rc = _lib.load().aj_mg_get_stats(self.handle, ctypes.byref(s),
                                              A.stream_handle(self.device))
        _lib.check(rc, "aj_mg_get_stats")
        return s.as_dict(textbook=self.variant != "reference")

    def stats_all(self) -> list:
        """Statistics of every instance of the last solve."""
        s = (_lib.MgStats * self.count)()
        with torch.cuda.device(self.device):
            rc = _lib.load().aj_mg_get_stats_batched(self.handle, ctypes.cast(s, ctypes.c_void_p),
                                                      A.stream_handle(self.device))
        _lib.check(rc, "aj_mg_get_stats_batched")
        return [x.as_dict(textbook=self.variant != "reference") for x in s]

    def close(self):
        """Release the plan (waits for an open deferred solve and the plan's
        device first); idempotent."""
        if getattr(self, "handle", None):
            _lib.load().aj_mg_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def get_plan(nx, ny, dx, dy, dtype, variant, levels, device) -> MultigridPlan:
    key = (int(nx), int(ny), float(dx), float(dy), dtype, variant, int(levels), str(device))
    with _plans_lock:
        p = _plans.get(key)
        if p is None:
            p = MultigridPlan(nx, ny, dx, dy, dtype, variant, levels, device)
            _plans[key] = p
        return p


@atexit.register
def _release_plans():
    with _plans_lock:
        for p in _plans.values():
            try:
                p.close()
            except Exception:
                pass
        _plans.clear()


def poisson_multigrid(rhs, mask, dx: float, dy: float, levels: int = 4, v_cycles: int = 5,
                      tolerance: float = 1e-6, flow_type: str = "von_karman", *, p0=None,
                      out=None, return_stats: bool = False):
    """Reference-exact geometric multigrid (Jacobi smoother, fine-grid spacing
    on every level, outlet Dirichlet row, ``v_cycles`` cycles with the
    reference's convergence test).  ``mask`` is accepted and unused, as in
    the reference."""
    if flow_type != "von_karman":
        raise NotImplementedError(f"flow_type={flow_type!r}: only 'von_karman' is implemented")
    A.require_gpu()
    want_np = A.is_numpy(rhs)
    dt_ = A.torch_dtype_of(rhs)
    bt = A.as_device(rhs, dt_, name="rhs")
    nx, ny = bt.shape
    pt = A.as_device(p0, dt_, bt.device, "p0")
    if pt is not None and tuple(pt.shape) != (nx, ny):
        raise ValueError("p0 shape must match rhs")
    plan = get_plan(nx, ny, dx, dy, dt_, "reference", levels, bt.device)
    po = torch.empty_like(bt) if out is None else out
    plan.solve(bt, po, pt, v_cycles, tolerance)
    res = A.out_like(po, want_np)
    if return_stats:
        return res, plan.stats()
    return res


def poisson_multigrid_to_tolerance(rhs, dx: float, dy: float, tolerance: float = 1e-6,
                                   max_cycles: int = 40, *, p0=None, out=None,
                                   return_stats: bool = False, levels: int = 0):
    """Textbook multigrid to ``sqrt(sum_{i<nx-1} (b - A p)^2 dx dy) <=
    tolerance`` on the consistent outlet-Dirichlet operator (FP32 V(2,2)
    cycles, FP64 residual/norm/accumulation, exact FP64 coarsest solve).
    ``levels`` = 0 coarsens down to <= 65536 cells; L > 0 asks for exactly
    L levels."""
    A.require_gpu()
    want_np = A.is_numpy(rhs)
    dt_ = A.torch_dtype_of(rhs)
    bt = A.as_device(rhs, dt_, name="rhs")
    nx, ny = bt.shape
    pt = A.as_device(p0, dt_, bt.device, "p0")
    plan = get_plan(nx, ny, dx, dy, dt_, "textbook", levels, bt.device)
    po = torch.empty_like(bt) if out is None else out
    plan.solve(bt, po, pt, max_cycles, tolerance)
    res = A.out_like(po, want_np)
    if return_stats:
        return res, plan.stats()
    return res
