"""ctypes handle of the C++/OpenMP restatement ``oracle/cpu_step.cpp``.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the multi-threaded CPU
baseline of ``bench.py`` and the tests that pin it against the NumPy oracle.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .build import LIB

_lib = None
_vp, _i, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise FileNotFoundError(f"{LIB} missing: run `python -m oracle.build`")
        lib = ctypes.CDLL(LIB)
        lib.ajcpu_threads.restype = _i
        lib.ajcpu_create.restype = _vp
        lib.ajcpu_create.argtypes = [_i, _i, _d, _d]
        lib.ajcpu_destroy.argtypes = [_vp]
        lib.ajcpu_levels.restype = _i
        lib.ajcpu_levels.argtypes = [_vp]
        lib.ajcpu_solve.restype = _i
        lib.ajcpu_solve.argtypes = [_vp, _vp, _vp, _d, _i, ctypes.POINTER(_d)]
        lib.ajcpu_step.restype = _i
        lib.ajcpu_step.argtypes = [_vp, _vp, _vp, _vp, _vp, _d, _d, _d, _d, _d, _i, _vp, _vp,
                                   _vp, _vp, ctypes.POINTER(_d)]
        _lib = lib
    return _lib


def threads() -> int:
    return int(load().ajcpu_threads())


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class CpuStepper:
    """To-tolerance step of one (nx, ny) grid on all OpenMP threads."""

    def __init__(self, nx, ny, dx, dy):
        self.nx, self.ny = int(nx), int(ny)
        self.h = load().ajcpu_create(self.nx, self.ny, float(dx), float(dy))
        if not self.h:
            raise ValueError(f"grid {nx}x{ny} not supported")

    @property
    def levels(self) -> int:
        return int(load().ajcpu_levels(self.h))

    def solve(self, rhs, tol=1e-6, max_outer=40):
        b = _f32(rhs)
        p = np.empty_like(b)
        nrm = _d()
        cyc = load().ajcpu_solve(self.h, b.ctypes.data, p.ctypes.data, float(tol),
                                 int(max_outer), ctypes.byref(nrm))
        return p, cyc, nrm.value

    def step(self, u, v, mask, sdf, nu, dt, U_inf=1.0, eta=0.005, tol=1e-6, max_outer=40,
             parts=False):
        """New (u, v, p, V-cycles, norm); with ``parts`` also u*, v*, rhs."""
        u, v = _f32(u).copy(), _f32(v).copy()
        mask, sdf = _f32(mask), _f32(sdf)
        p = np.empty_like(u)
        extra = [np.empty_like(u) for _ in range(3)] if parts else [None] * 3
        nrm = _d()
        ptr = lambda a: None if a is None else a.ctypes.data  # noqa: E731
        cyc = load().ajcpu_step(self.h, u.ctypes.data, v.ctypes.data, mask.ctypes.data,
                                sdf.ctypes.data, float(nu), float(dt), float(U_inf), float(eta),
                                float(tol), int(max_outer), p.ctypes.data, ptr(extra[0]),
                                ptr(extra[1]), ptr(extra[2]), ctypes.byref(nrm))
        if parts:
            return u, v, p, cyc, nrm.value, extra
        return u, v, p, cyc, nrm.value

    def close(self):
        if getattr(self, "h", None):
            load().ajcpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
