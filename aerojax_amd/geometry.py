"""Grid, obstacle geometry and initial condition (setup, once per geometry).

Setup code evaluated on the host in float64 and handed to the step as
device arrays, the way the reference evaluates it once in
``BaselineSolver.__init__`` and passes ``mask``/``sdf`` into every step.

Follows:
  GridParams.__post_init__        solver/params.py:51-56
  parse_naca_4digit               obstacles/naca_airfoils.py:16-26
  naca_surface_distance           obstacles/naca_airfoils.py:109-178
  cylinder level set              solver/simulation/mask_generator.py:151-160
  mask = 1 - sigmoid(-sdf/eps)    solver/simulation/mask_generator.py:71-72
  von Karman IC (u=U_inf, v=0)    solver/simulation/initializers.py:8-17
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class GridParams:
    """``dx = lx/nx``, ``dy = ly/ny``; coordinates are ``linspace(0, l, n)``
    (spacing l/(n-1)), both exactly as the reference defines them."""
    nx: int
    ny: int
    lx: float
    ly: float
    dx: float = field(init=False)
    dy: float = field(init=False)
    X: np.ndarray = field(init=False, repr=False)
    Y: np.ndarray = field(init=False, repr=False)

    def __post_init__(self):
        self.dx = self.lx / self.nx
        self.dy = self.ly / self.ny
        xs = np.linspace(0.0, self.lx, self.nx)
        ys = np.linspace(0.0, self.ly, self.ny)
        self.X = np.broadcast_to(xs[:, None], (self.nx, self.ny)).copy()
        self.Y = np.broadcast_to(ys[None, :], (self.nx, self.ny)).copy()


def parse_naca_4digit(designation: str):
    """'0012' / 'NACA 2412' -> (m, p, t) = (digit0/100, digit1/10, digits23/100)."""
    d = [c for c in designation if c.isdigit()]
    if len(d) != 4:
        raise ValueError(f"Invalid 4-digit NACA designation: {designation}")
    return int(d[0]) / 100.0, int(d[1]) / 10.0, int(d[2] + d[3]) / 100.0


def _naca_thickness(xc, t):
    """Half thickness y_t(x/c) in chord units (NACA 4-digit, open TE)."""
    return 5.0 * t * (0.2969 * np.sqrt(np.abs(xc)) - 0.1260 * xc - 0.3516 * xc ** 2
                      + 0.2843 * xc ** 3 - 0.1015 * xc ** 4)


def _naca_camber(xc, m, p):
    """Camber line and its slope.  Evaluated with IEEE semantics: a symmetric
    section (m = p = 0) gives 0/0 = NaN on the forward branch x/c <= p,
    which the reference carries too."""
    with np.errstate(divide="ignore", invalid="ignore"):
        fwd_k = np.float64(m) / np.float64(p) ** 2
        aft_k = np.float64(m) / (1.0 - np.float64(p)) ** 2
        fwd = xc <= p
        yc = np.where(fwd, fwd_k * (2 * p * xc - xc ** 2),
                      aft_k * ((1 - 2 * p) + 2 * p * xc - xc ** 2))
        slope = np.where(fwd, 2 * fwd_k * (p - xc), 2 * aft_k * (p - xc))
    return yc, slope


def naca_sdf(X, Y, chord: float, aoa_deg: float, x_le: float, y_le: float, m: float,
             p: float, t: float) -> np.ndarray:
    """Signed pseudo-distance to a NACA 4-digit section whose leading edge is
    at ``(x_le, y_le)``, rotated by ``aoa_deg`` about it (negative inside).

    Inside the chord the distance is vertical (in the rotated frame) to the
    nearer surface; outside it is the distance to the leading or trailing
    edge point.  A point is inside when 0 <= x/c <= 1 and it lies within the
    surfaces widened by 0.01 chord."""
    a = math.radians(aoa_deg)
    ca, sa = math.cos(a), math.sin(a)
    dxp = np.asarray(X, dtype=np.float64) - x_le
    dyp = np.asarray(Y, dtype=np.float64) - y_le
    xr = dxp * ca - dyp * sa
    yr = dxp * sa + dyp * ca
    xc, yc_pt = xr / chord, yr / chord
    within = (xc >= 0.0) & (xc <= 1.0)
    half = _naca_thickness(np.clip(xc, 0.0, 1.0), t)
    ycl, slope = _naca_camber(xc, m, p)
    with np.errstate(invalid="ignore"):
        cth = np.cos(np.arctan(slope))
        upper = ycl + half * cth
        lower = ycl - half * cth
        d_chord = np.minimum(np.abs(yc_pt - upper), np.abs(yc_pt - lower)) * chord
        d_edges = np.minimum(np.hypot(xr, yr), np.hypot(xr - chord, yr))
        dist = np.where(within, d_chord, d_edges)
        inside = within & (yc_pt >= lower - 0.01) & (yc_pt <= upper + 0.01)
    return np.where(inside, -dist, dist)


def cylinder_sdf(X, Y, cx: float, cy: float, r: float) -> np.ndarray:
    return np.sqrt((np.asarray(X, np.float64) - cx) ** 2 + (np.asarray(Y, np.float64) - cy) ** 2) - r


def mask_from_sdf(sdf: np.ndarray, eps: float) -> np.ndarray:
    """Fluid fraction 1 - sigmoid(-sdf/eps) (1 in fluid, 0 in solid)."""
    z = -np.asarray(sdf, np.float64) / eps
    with np.errstate(over="ignore"):
        return 1.0 - 1.0 / (1.0 + np.exp(-z))


def von_karman_ic(nx: int, ny: int, U_inf: float = 1.0, noise_sigma: float = 0.0,
                  seed: int = 0):
    """u = U_inf, v = 0, plus optional seeded Gaussian noise (drawn in float64
    from ``np.random.default_rng(seed)``, u noise first then v)."""
    u = np.full((nx, ny), float(U_inf))
    v = np.zeros((nx, ny))
    if noise_sigma > 0:
        g = np.random.default_rng(seed)
        u += g.normal(0.0, noise_sigma, size=(nx, ny))
        v += g.normal(0.0, noise_sigma, size=(nx, ny))
    return u, v


@dataclass
class Geometry:
    """SDF and mask of one obstacle on one grid (float64)."""
    sdf: np.ndarray
    mask: np.ndarray
    L_char: float


def naca_geometry(grid: GridParams, designation="0012", chord=0.5, aoa_deg=0.0,
                  x_le=1.0, y_le=None, eps=0.05) -> Geometry:
    m, p, t = parse_naca_4digit(designation)
    y_le = grid.ly / 2 if y_le is None else y_le
    sdf = naca_sdf(grid.X, grid.Y, chord, aoa_deg, x_le, y_le, m, p, t)
    return Geometry(sdf, mask_from_sdf(sdf, eps), chord)


def cylinder_geometry(grid: GridParams, cx, cy, r, eps=0.05) -> Geometry:
    sdf = cylinder_sdf(grid.X, grid.Y, cx, cy, r)
    return Geometry(sdf, mask_from_sdf(sdf, eps), 2.0 * r)


def viscosity(U_inf: float, L_char: float, Re: float) -> float:
    """nu = U L / Re (FlowParams.resolve, solver/params.py:143-148)."""
    return U_inf * L_char / Re


# ---------------------------------------------------------------------------
# the same geometry evaluated on the GPU (aj_geometry / aj_batch_geometry)
# ---------------------------------------------------------------------------

def naca_obstacle(designation="0012", chord=0.5, aoa_deg=0.0, x_le=1.0, y_le=0.5, eps=0.05):
    """``aj_obstacle`` of a NACA 4-digit section (leading edge at (x_le, y_le))."""
    from . import _lib
    m, p, t = parse_naca_4digit(designation)
    return _lib.Obstacle(kind=_lib.AJ_OBSTACLE_NACA, x0=float(x_le), y0=float(y_le),
                         chord=float(chord), aoa_deg=float(aoa_deg), m=m, p=p, t=t,
                         eps=float(eps))


def cylinder_obstacle(cx, cy, r, eps=0.05):
    """``aj_obstacle`` of a cylinder of radius r centred at (cx, cy)."""
    from . import _lib
    return _lib.Obstacle(kind=_lib.AJ_OBSTACLE_CYLINDER, x0=float(cx), y0=float(cy),
                         radius=float(r), eps=float(eps))


def device_geometry(grid: GridParams, obstacle, dtype=None, device=None):
    """(sdf, mask) of ``obstacle`` on ``grid`` computed on the GPU."""
    import ctypes

    import torch

    from . import _arrays as A
    from . import _lib
    A.require_gpu()
    dtype = torch.float32 if dtype is None else dtype
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    sdf = torch.empty((grid.nx, grid.ny), dtype=dtype, device=dev)
    mask = torch.empty_like(sdf)
    with torch.cuda.device(dev):
        rc = _lib.load().aj_geometry(A.dtype_code(sdf), ctypes.byref(obstacle), grid.nx, grid.ny,
                                     float(grid.lx), float(grid.ly), A.ptr(sdf), A.ptr(mask),
                                     A.stream_handle(dev))
    _lib.check(rc, "aj_geometry")
    return sdf, mask
