"""Synthetic workloads (seeded inputs) for the benchmark and the tests.

Each function returns a dict with the grid, the obstacle SDF and mask
(float64 host arrays), the initial velocity (float64, seeded noise), the
viscosity and the time step.  Domains the configurations leave open are
fixed here: the 2048x512 learned-warm-start case uses a 4x1 channel
(dx = dy = 1/512).
"""
from __future__ import annotations

import math

import numpy as np

from . import geometry as G


def small_airfoil(noise_sigma=1e-3, seed=0):
    """512x96 on 4x0.75, NACA 0012 chord 0.5 at 10 deg, LE at (1, 0.375),
    mask eps 0.05, Re 1000 (nu = U c / Re = 5e-4), dt 1e-3, 3 V-cycles."""
    grid = G.GridParams(512, 96, 4.0, 0.75)
    geo = G.naca_geometry(grid, "0012", chord=0.5, aoa_deg=10.0, x_le=1.0, y_le=0.375, eps=0.05)
    u, v = G.von_karman_ic(grid.nx, grid.ny, 1.0, noise_sigma, seed)
    return dict(name="naca0012_512x96", grid=grid, sdf=np.nan_to_num(geo.sdf, nan=0.0),
                mask=geo.mask, u=u, v=v, nu=G.viscosity(1.0, 0.5, 1000.0), dt=1e-3,
                U_inf=1.0, v_cycles=3, les=False)


def large_cylinder(noise_sigma=1e-3, seed=0):
    """4096x1024 on 8x2, cylinder r = 0.1 at (1.5, 1.0), eps 0.05,
    Re 1e5 (nu = U 2r / Re = 2e-6), Smagorinsky Cs = 0.17, dt 1e-4."""
    grid = G.GridParams(4096, 1024, 8.0, 2.0)
    geo = G.cylinder_geometry(grid, 1.5, 1.0, 0.1, eps=0.05)
    u, v = G.von_karman_ic(grid.nx, grid.ny, 1.0, noise_sigma, seed)
    return dict(name="cylinder_4096x1024", grid=grid, sdf=geo.sdf, mask=geo.mask, u=u, v=v,
                nu=G.viscosity(1.0, 0.2, 1e5), dt=1e-4, U_inf=1.0, les=True, cs=0.17)


def draw_obstacle(grid, rng):
    """Cylinder with radius in [0.05, 0.15] or NACA 4-digit with thickness
    6-18 % and AoA in [-15, 15] deg (drawn from ``rng``): a dict of the
    parameters."""
    if rng.random() < 0.5:
        r = rng.uniform(0.05, 0.15)
        cx = rng.uniform(0.6, 1.2)
        cy = grid.ly / 2 + rng.uniform(-0.1, 0.1)
        return dict(kind="cylinder", cx=cx, cy=cy, r=r, L=2.0 * r,
                    desc=f"cylinder r={r:.3f} at ({cx:.2f},{cy:.2f})")
    tpct = int(rng.integers(6, 19))
    m = int(rng.integers(0, 7))
    p = int(rng.integers(2, 7)) if m > 0 else 0
    des = f"{m}{p}{tpct:02d}"
    aoa = rng.uniform(-15.0, 15.0)
    chord = rng.uniform(0.3, 0.6)
    x_le = rng.uniform(0.5, 1.0)
    return dict(kind="naca", designation=des, chord=chord, aoa=aoa, x_le=x_le, y_le=grid.ly / 2,
                L=chord, desc=f"NACA {des} c={chord:.2f} aoa={aoa:.1f}")


def obstacle_geometry(grid, ob, eps=0.05):
    """Host SDF and mask of a drawn obstacle."""
    if ob["kind"] == "cylinder":
        return G.cylinder_geometry(grid, ob["cx"], ob["cy"], ob["r"], eps=eps)
    return G.naca_geometry(grid, ob["designation"], chord=ob["chord"], aoa_deg=ob["aoa"],
                           x_le=ob["x_le"], y_le=ob["y_le"], eps=eps)


def obstacle_struct(ob, eps=0.05):
    """The same obstacle as an ``aj_obstacle`` for the GPU geometry."""
    if ob["kind"] == "cylinder":
        return G.cylinder_obstacle(ob["cx"], ob["cy"], ob["r"], eps)
    return G.naca_obstacle(ob["designation"], ob["chord"], ob["aoa"], ob["x_le"], ob["y_le"], eps)


def random_obstacle(grid, rng, eps=0.05):
    """Draw an obstacle (``draw_obstacle``) and build its host geometry."""
    ob = draw_obstacle(grid, rng)
    return obstacle_geometry(grid, ob, eps), ob["desc"]


def warm_start_params(seed=0, nx=2048, ny=512):
    """The draws of ``warm_start_case`` without building the fields:
    (grid, obstacle dict, Re, nu)."""
    rng = np.random.default_rng(seed)
    grid = G.GridParams(nx, ny, 4.0, 1.0)
    ob = draw_obstacle(grid, rng)
    Re = math.exp(rng.uniform(math.log(500.0), math.log(5e4)))
    return grid, ob, Re, G.viscosity(1.0, ob["L"], Re)


def warm_start_case(seed=0, nx=2048, ny=512, noise_sigma=1e-3):
    """2048x512 on a 4x1 channel, geometry drawn per seed, Re log-uniform in
    [500, 5e4], solve to ||div|| <= 1e-6.

    dt = 2.5e-4 gives the advective CFL number U dt/dx = 0.128 of the
    512x96 / dt 1e-3 configuration; at dt = 1e-3 (CFL 0.51) the step with a
    converged pressure blows up for steep airfoils within ~10 steps once the
    local CFL passes ~1 (the CPU restatement does the same)."""
    rng = np.random.default_rng(seed)
    grid = G.GridParams(nx, ny, 4.0, 1.0)
    geo, desc = random_obstacle(grid, rng)
    Re = math.exp(rng.uniform(math.log(500.0), math.log(5e4)))
    u, v = G.von_karman_ic(nx, ny, 1.0, noise_sigma, seed)
    return dict(name=f"warmstart_{nx}x{ny}_seed{seed}", desc=desc, grid=grid,
                sdf=np.nan_to_num(geo.sdf, nan=0.0), mask=geo.mask, u=u, v=v,
                nu=G.viscosity(1.0, geo.L_char, Re), Re=Re, dt=2.5e-4, U_inf=1.0, tol=1e-6)


def batched_sweep_case(n=256, nx=512, ny=128, seed=0, noise_sigma=1e-3):
    """``n`` independent 512x128 channels (4x1, dx = dy = 1/128), NACA 0012
    per instance with AoA ~ U[-15, 15] deg, chord ~ U[0.3, 0.6] (leading edge
    at x = 1, mid-height), Re ~ U[500, 5000], mask eps 0.05, seeded inflow
    noise; dt = 1e-3 (advective CFL 0.128), reference multigrid with 3
    V-cycles per step."""
    rng = np.random.default_rng(seed)
    grid = G.GridParams(nx, ny, 4.0, 1.0)
    masks, sdfs, nus, us, vs, desc = [], [], [], [], [], []
    for k in range(n):
        aoa = rng.uniform(-15.0, 15.0)
        chord = rng.uniform(0.3, 0.6)
        Re = rng.uniform(500.0, 5000.0)
        geo = G.naca_geometry(grid, "0012", chord=chord, aoa_deg=aoa, x_le=1.0,
                              y_le=grid.ly / 2, eps=0.05)
        u, v = G.von_karman_ic(nx, ny, 1.0, noise_sigma, seed * 100003 + k)
        masks.append(geo.mask)
        sdfs.append(np.nan_to_num(geo.sdf, nan=0.0))
        nus.append(G.viscosity(1.0, geo.L_char, Re))
        us.append(u)
        vs.append(v)
        desc.append((aoa, chord, Re))
    return dict(name=f"batched_{n}x{nx}x{ny}_seed{seed}", grid=grid, masks=masks, sdfs=sdfs,
                nus=nus, u=us, v=vs, params=desc, dt=1e-3, U_inf=1.0, v_cycles=3,
                brinkman_eta=0.005)


C5_SIZES = ((256, 64), (512, 128), (1024, 256), (2048, 512))


def uneven_batch_case(per_size=16, sizes=C5_SIZES, seed=0, noise_sigma=1e-3, cs=0.17):
    """``per_size`` channels (4x1, dx = dy = 1/ny) of every size in
    ``sizes``, each with its own obstacle and Reynolds number drawn as in
    ``warm_start_case`` (cylinder or NACA 4-digit, Re log-uniform in
    [500, 5e4]); Smagorinsky LES with Cs = ``cs``; one dt = 2.5e-4 for all
    (advective CFL 0.128 on the finest grid, less on the coarser ones); every
    pressure solved to ||div|| <= 1e-6.  Returns the ``BatchedStepper``
    groups and per-instance lists, group by group."""
    groups, masks, sdfs, nus, us, vs, desc = [], [], [], [], [], [], []
    k = 0
    for (nx, ny) in sizes:
        grid = G.GridParams(nx, ny, 4.0, 1.0)
        groups.append((nx, ny, per_size, grid.dx, grid.dy))
        for _ in range(per_size):
            c = warm_start_case(seed * 100003 + k, nx, ny, noise_sigma)
            masks.append(c["mask"])
            sdfs.append(c["sdf"])
            nus.append(c["nu"])
            us.append(c["u"])
            vs.append(c["v"])
            desc.append((nx, ny, c["desc"], c["Re"]))
            k += 1
    name = "uneven_" + "+".join(f"{per_size}x{nx}x{ny}" for nx, ny in sizes) + f"_seed{seed}"
    return dict(name=name, groups=groups, masks=masks, sdfs=sdfs, nus=nus, u=us, v=vs,
                params=desc, dt=2.5e-4, U_inf=1.0, cs=cs, tol=1e-6, brinkman_eta=0.005)
