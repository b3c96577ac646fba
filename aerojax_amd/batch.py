"""Many independent simulations stepped together (``aj_batch_*``).

Each instance is its own von Karman channel (grid, obstacle, viscosity);
one step applies step_pure's composition (RK3 predictor, rhs, reference
multigrid with ``v_cycles`` cycles, projection; solver/pure_functions/
step.py:49-467) to every instance in shared, graph-captured launches.
Instances are grouped by grid shape; fields live in one flat device buffer
per quantity, instance k at ``offsets[k]``.

``solver="textbook"`` solves every pressure to ``tolerance`` instead (one
batched textbook plan per group, at most ``max_cycles`` V-cycles);
``smagorinsky_cs`` > 0 feeds the constant Smagorinsky eddy viscosity to the
RK step; ``warm_start`` (a ``WarmStartNet``) gives every instance a learned
initial guess (one inference per group), kept per instance only if its
residual is not above the zero guess's.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _arrays as A
from . import _lib


class BatchedStepper:
    """``groups``: list of ``(nx, ny, count, dx, dy)``; ``masks`` / ``sdfs``
    / ``nus``: one entry per instance, group by group; ``u0`` / ``v0``:
    per-instance initial fields.  ``obstacles`` (one ``aj_obstacle`` per
    instance, see ``geometry.naca_obstacle`` / ``cylinder_obstacle``) builds
    the masks and SDFs on the GPU instead (pass ``masks=sdfs=None``)."""

    def __init__(self, groups, masks, sdfs, nus, u0, v0, dt, *, U_inf=1.0, brinkman_eta=0.005,
                 v_cycles=3, tolerance=1e-6, levels=4, dtype=torch.float32, device=None,
                 solver="reference", max_cycles=40, smagorinsky_cs=0.0, warm_start=None,
                 obstacles=None):
        A.require_gpu()
        if dtype not in (torch.float32, torch.float64):
            raise TypeError(f"unsupported dtype {dtype} (float32 or float64)")
        if solver not in ("reference", "textbook"):
            raise ValueError(f"solver={solver!r}: 'reference' or 'textbook'")
        if warm_start is not None and (solver != "textbook" or dtype != torch.float32):
            raise ValueError("warm_start needs solver='textbook' and float32 fields")
        self.solver = solver
        self.warm_start = warm_start
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None \
            else torch.device(device)
        self.dtype = dtype
        self.groups = [tuple(g) for g in groups]
        ng = len(self.groups)
        arr = lambda k, t: (t * ng)(*[t(g[k]) for g in self.groups])
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            code = _lib.AJ_F32 if dtype == torch.float32 else _lib.AJ_F64
            rc = _lib.load().aj_batch_create(ctypes.byref(h), code,
                                             ng, arr(0, ctypes.c_int), arr(1, ctypes.c_int),
                                             arr(2, ctypes.c_int), arr(3, ctypes.c_double),
                                             arr(4, ctypes.c_double), int(levels))
        _lib.check(rc, "aj_batch_create")
        self.handle = h
        lib = _lib.load()
        self.n = int(lib.aj_batch_num_instances(h))
        self.cells = int(lib.aj_batch_cells(h))
        off = np.zeros(self.n, dtype=np.int64)
        _lib.check(lib.aj_batch_offsets(h, off.ctypes.data), "aj_batch_offsets")
        self.offsets = off
        self.shapes = [(g[0], g[1]) for g in self.groups for _ in range(g[2])]
        for name, lst in (("masks", masks), ("sdfs", sdfs), ("nus", nus), ("u0", u0), ("v0", v0),
                          ("obstacles", obstacles)):
            if lst is not None and len(lst) != self.n:
                raise ValueError(f"{name}: {len(lst)} entries for {self.n} instances")
        nu = np.ascontiguousarray(np.asarray(nus, dtype=np.float64))
        _lib.check(lib.aj_batch_set_viscosity(h, nu.ctypes.data), "aj_batch_set_viscosity")
        if obstacles is not None:
            if masks is not None or sdfs is not None:
                raise ValueError("give either obstacles or masks / sdfs")
            self.mask = torch.empty(self.cells, dtype=self.dtype, device=self.device)
            self.sdf = torch.empty_like(self.mask)
            arr = (_lib.Obstacle * self.n)(*obstacles)
            with torch.cuda.device(self.device):
                rc = lib.aj_batch_geometry(h, ctypes.cast(arr, ctypes.c_void_p), A.ptr(self.sdf),
                                           A.ptr(self.mask), A.stream_handle(self.device))
            _lib.check(rc, "aj_batch_geometry")
        else:
            self.mask = self._pack(masks)
            self.sdf = self._pack(sdfs) if sdfs is not None else None
        self.u = self._pack(u0)
        self.v = self._pack(v0)
        self.u_next = torch.empty_like(self.u)
        self.v_next = torch.empty_like(self.v)
        self.p = torch.zeros_like(self.u)
        tb = solver == "textbook"
        self.prm = _lib.BatchStepParams(float(dt), float(U_inf), float(brinkman_eta),
                                        float(tolerance), int(max_cycles if tb else v_cycles),
                                        _lib.AJ_MG_TEXTBOOK if tb else _lib.AJ_MG_REFERENCE,
                                        float(smagorinsky_cs))
        if warm_start is not None:
            if self.sdf is None:
                raise ValueError("warm_start needs the instances' sdf")
            self.rhs = torch.empty_like(self.u)
            self.p0 = torch.empty_like(self.u)

    def _pack(self, fields):
        buf = torch.empty(self.cells, dtype=self.dtype, device=self.device)
        for k, f in enumerate(fields):
            nx, ny = self.shapes[k]
            t = A.as_device(f, self.dtype, self.device, "field")
            if tuple(t.shape) != (nx, ny):
                raise ValueError(f"instance {k}: shape {tuple(t.shape)} != {(nx, ny)}")
            buf[self.offsets[k]:self.offsets[k] + nx * ny].copy_(t.reshape(-1))
        return buf

    def view(self, buf, k):
        """Instance k's (nx, ny) view of a flat batch buffer."""
        nx, ny = self.shapes[k]
        return buf[self.offsets[k]:self.offsets[k] + nx * ny].view(nx, ny)

    def set_graph(self, on: bool):
        _lib.check(_lib.load().aj_batch_set_option(self.handle, _lib.AJ_MG_OPT_GRAPH, int(on)),
                   "aj_batch_set_option")

    def step(self):
        """Advance every instance by one step (u, v swapped with the outputs;
        ``p`` holds the pressures)."""
        lib = _lib.load()
        prm = ctypes.byref(self.prm)
        with torch.cuda.device(self.device):
            st = A.stream_handle(self.device)
            if self.warm_start is None:
                rc = lib.aj_batch_step(self.handle, A.ptr(self.u), A.ptr(self.v),
                                       A.ptr(self.mask), A.ptr(self.sdf), A.ptr(self.u_next),
                                       A.ptr(self.v_next), A.ptr(self.p), prm, st)
                _lib.check(rc, "aj_batch_step")
            else:
                rc = lib.aj_batch_predict(self.handle, A.ptr(self.u), A.ptr(self.v),
                                          A.ptr(self.mask), A.ptr(self.sdf), A.ptr(self.rhs),
                                          prm, st)
                _lib.check(rc, "aj_batch_predict")
                k = 0
                for (nx, ny, cnt, _dx, _dy) in self.groups:
                    o, n = int(self.offsets[k]), nx * ny * cnt
                    self.warm_start.warm_start_batched(self.rhs[o:o + n], self.sdf[o:o + n],
                                                       self.p0[o:o + n], cnt, nx, ny)
                    k += cnt
                rc = lib.aj_batch_solve(self.handle, A.ptr(self.rhs), A.ptr(self.p0),
                                        A.ptr(self.p), prm, st)
                _lib.check(rc, "aj_batch_solve")
                rc = lib.aj_batch_project(self.handle, A.ptr(self.p), A.ptr(self.mask),
                                          A.ptr(self.u_next), A.ptr(self.v_next), prm, st)
                _lib.check(rc, "aj_batch_project")
        self.u, self.u_next = self.u_next, self.u
        self.v, self.v_next = self.v_next, self.v

    def stats(self):
        s = (_lib.MgStats * self.n)()
        with torch.cuda.device(self.device):
            rc = _lib.load().aj_batch_get_stats(self.handle, ctypes.cast(s, ctypes.c_void_p),
                                                A.stream_handle(self.device))
        _lib.check(rc, "aj_batch_get_stats")
        return [x.as_dict(textbook=self.solver == "textbook") for x in s]

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().aj_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
