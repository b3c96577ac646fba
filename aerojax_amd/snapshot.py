"""Snapshot records and solver-generated datasets in the reference's formats.

``Snapshot`` carries the fields of the trace viewer's record
(trace_viewer/snapshot.py:14-98: u, v, p, mask, dt, nu, dx, dy, divergence,
Re, CFL, flow_type, solver_metadata, iteration, timestamp) with the same
``to_dict`` / ``from_dict``, so ``reference.Snapshot.from_dict(rec.to_dict())``
builds the reference's object.  Files are written as ``.npz`` (arrays,
scalars, and the metadata as JSON text), never pickled.

``DatasetWriter`` records steps in the key convention of the reference's
training data (neural_operators/training_utils.py:35-51: ``u, v, p, mask,
divergence`` as [steps, nx, ny] arrays, ``dx, dy, dt, nx, ny`` scalars, with
``divergence`` the periodic ``solver.operators.divergence``), plus ``rhs``
and ``sdf`` (the pressure right-hand side and obstacle SDF the warm-start
network takes); ``load_dataset`` reads such a file (no pickle) and
``pairs_from_dataset`` turns it into network training pairs.
"""
from __future__ import annotations

import json
import time
from dataclasses import dataclass, field, fields
from typing import Any, Dict, Optional

import numpy as np
import torch

from . import _arrays as A
from .diagnostics import flow_diagnostics
from .operators import divergence, divergence_nonperiodic

DATASET_KEYS = ("u", "v", "p", "mask", "divergence", "dx", "dy", "dt", "nx", "ny")


@dataclass
class Snapshot:
    """Simulation state at one step (trace_viewer/snapshot.py:14-98)."""
    u: np.ndarray
    v: np.ndarray
    p: np.ndarray
    mask: np.ndarray
    dt: float
    nu: float
    dx: float
    dy: float
    divergence: Optional[np.ndarray] = None
    Re: Optional[float] = None
    CFL: Optional[float] = None
    flow_type: Optional[str] = None
    solver_metadata: Optional[Dict[str, Any]] = field(default=None)
    iteration: int = 0
    timestamp: Optional[float] = None

    def __post_init__(self):
        nx, ny = self.u.shape
        for name in ("v", "p", "mask") + (("divergence",) if self.divergence is not None else ()):
            if getattr(self, name).shape != (nx, ny):
                raise ValueError(f"{name} must have the shape of u {(nx, ny)}")

    @property
    def shape(self):
        return self.u.shape

    @property
    def nx(self) -> int:
        return self.u.shape[0]

    @property
    def ny(self) -> int:
        return self.u.shape[1]

    def to_dict(self) -> Dict[str, Any]:
        return {f.name: getattr(self, f.name) for f in fields(self)}

    @classmethod
    def from_dict(cls, data: Dict[str, Any]) -> "Snapshot":
        names = {f.name for f in fields(cls)}
        return cls(**{k: v for k, v in data.items() if k in names})

    def save(self, path: str) -> None:
        """``.npz``: arrays as arrays, scalars as 0-d arrays, None omitted,
        solver_metadata as JSON text."""
        out = {}
        for k, v in self.to_dict().items():
            if v is None:
                continue
            if k == "solver_metadata":
                out[k] = np.array(json.dumps(v))
            elif k == "flow_type":
                out[k] = np.array(str(v))
            else:
                out[k] = np.asarray(v)
        np.savez(path, **out)

    @classmethod
    def load(cls, path: str) -> "Snapshot":
        with np.load(path, allow_pickle=False) as z:
            d: Dict[str, Any] = {}
            for k in z.files:
                a = z[k]
                if k == "solver_metadata":
                    d[k] = json.loads(str(a))
                elif k == "flow_type":
                    d[k] = str(a)
                elif a.ndim == 0:
                    d[k] = a.item()
                else:
                    d[k] = a
        return cls.from_dict(d)


def take_snapshot(u, v, p, mask, dt, nu, dx, dy, *, Re=None, flow_type="von_karman",
                  iteration=0, solver_metadata=None) -> Snapshot:
    """Snapshot of device fields: divergence_nonperiodic and the CFL number
    (check_cfl) are computed on the GPU, then everything is copied to the
    host once."""
    dt_ = A.torch_dtype_of(u)
    ut = A.as_device(u, dt_, name="u")
    vt, pt, mt = (A.as_device(x, dt_, ut.device, n) for x, n in ((v, "v"), (p, "p"), (mask, "mask")))
    div = divergence_nonperiodic(ut, vt, dx, dy)
    diag = flow_diagnostics(ut, vt, pt, mt, dx, dy, dt)
    h = lambda t: t.detach().cpu().numpy().astype(np.float64)  # noqa: E731
    return Snapshot(u=h(ut), v=h(vt), p=h(pt), mask=h(mt), dt=float(dt), nu=float(nu),
                    dx=float(dx), dy=float(dy), divergence=h(div), Re=Re, CFL=diag["cfl"],
                    flow_type=flow_type, solver_metadata=solver_metadata,
                    iteration=int(iteration), timestamp=time.time())


class DatasetWriter:
    """Accumulates steps of one (nx, ny) simulation on the device and writes
    them in the reference's training-data key convention."""

    def __init__(self, nx, ny, dx, dy, dt, capacity, device=None):
        self.nx, self.ny, self.dx, self.dy, self.dt = int(nx), int(ny), float(dx), float(dy), float(dt)
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else device
        shape = (int(capacity), self.nx, self.ny)
        self.buf = {k: torch.empty(shape, dtype=torch.float32, device=dev)
                    for k in ("u", "v", "p", "mask", "divergence", "rhs", "sdf")}
        self.n = 0

    def record(self, u, v, p, mask, rhs=None, sdf=None):
        """Store one step (device tensors); ``rhs`` / ``sdf`` optional."""
        if self.n >= self.buf["u"].shape[0]:
            raise ValueError("dataset capacity exhausted")
        k = self.n
        self.buf["u"][k].copy_(u)
        self.buf["v"][k].copy_(v)
        self.buf["p"][k].copy_(p)
        self.buf["mask"][k].copy_(mask)
        divergence(u.float(), v.float(), self.dx, self.dy, out=self.buf["divergence"][k])
        for name, t in (("rhs", rhs), ("sdf", sdf)):
            if t is not None:
                self.buf[name][k].copy_(t)
            else:
                self.buf[name][k].fill_(float("nan"))
        self.n += 1

    def save(self, path: str) -> str:
        out = {k: t[:self.n].double().cpu().numpy() for k, t in self.buf.items()
               if k in ("u", "v", "p", "mask", "divergence")}
        for k in ("rhs", "sdf"):
            a = self.buf[k][:self.n]
            if self.n and not torch.isnan(a).any():
                out[k] = a.double().cpu().numpy()
        out.update(dx=self.dx, dy=self.dy, dt=self.dt, nx=self.nx, ny=self.ny)
        np.savez(path, **out)
        return path


def load_dataset(path: str) -> Dict[str, Any]:
    """Arrays and scalars of a dataset file (reference convention; read
    without unpickling)."""
    with np.load(path, allow_pickle=False) as z:
        d = {k: (z[k].item() if z[k].ndim == 0 else z[k]) for k in z.files}
    missing = [k for k in ("p", "mask", "dx", "dy", "nx", "ny") if k not in d]
    if missing:
        raise KeyError(f"{path}: missing keys {missing}")
    return d


def pairs_from_dataset(net, data: Dict[str, Any], eps: float = 0.05, device=None):
    """Network training pairs (x [N, H, W, 2], t [N, H, W]) from a dataset:
    the input is ``rhs`` (else ``divergence / dt``, the reference operator's
    input) with ``sdf`` (else eps * logit(mask), the inverse of
    mask = sigmoid(sdf / eps)); the target is ``p``."""
    dev = net.device if device is None else device
    p = np.asarray(data["p"], np.float32)
    N, nx, ny = p.shape
    if "rhs" in data:
        rhs = np.asarray(data["rhs"], np.float32)
    else:
        rhs = np.asarray(data["divergence"], np.float32) / np.float32(data["dt"])
    if "sdf" in data:
        sdf = np.asarray(data["sdf"], np.float32)
    else:
        m = np.clip(np.asarray(data["mask"], np.float64), 1e-7, 1 - 1e-7)
        sdf = (eps * np.log(m / (1.0 - m))).astype(np.float32)
    H, W = nx >> net.r, ny >> net.r
    X = torch.empty((N, H, W, 2), dtype=torch.float32, device=dev)
    T = torch.empty((N, H, W), dtype=torch.float32, device=dev)
    for k in range(N):
        d = lambda a: torch.from_numpy(np.ascontiguousarray(a[k])).to(dev)  # noqa: E731
        net.make_sample(d(rhs), d(sdf), d(p), X[k], T[k])
    return X, T
