"""NumPy restatement of AeroJAX's collocated projection step (CPU oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``; never by the product package.

Every function takes a ``dtype`` (``np.float32`` mirrors the reference's
default ``jax_enable_x64=False``, solver/config.py:25; ``np.float64`` mirrors
the x64 mode).  Scalars that the reference passes as traced jit arguments
(``dt``, ``nu``, ``dx``, ``dy``, ``U_inf``, ``brinkman_eta``, ``tolerance``)
are cast to ``dtype`` first, as JAX does with weakly typed Python floats.
Literal constants stay Python floats, so NumPy 2 (NEP 50) applies them in the
array dtype exactly as jax.numpy does.

Layout: every field is ``(nx, ny)``, axis 0 = x (inlet row 0, outlet row
nx-1), axis 1 = y (walls at columns 0 and ny-1), C order, so y is contiguous
(solver/params.py:54-56, ``indexing='ij'``).

Citations are into the reference tree (arriemeijer-creator/AeroJAX).
"""
from __future__ import annotations

import numpy as np

# ---------------------------------------------------------------------------
# Grid, geometry and initial condition
# ---------------------------------------------------------------------------


def grid_coords(nx, ny, lx, ly, dtype=np.float64):
    """GridParams.__post_init__ (solver/params.py:51-56).

    ``dx = lx/nx`` and ``dy = ly/ny``, but the coordinates are
    ``linspace(0, lx, nx)`` (spacing lx/(nx-1)); the reference uses both, so
    they are restated as they are.
    """
    dx = lx / nx
    dy = ly / ny
    x = np.linspace(0.0, lx, nx, dtype=np.float64)
    y = np.linspace(0.0, ly, ny, dtype=np.float64)
    X, Y = np.meshgrid(x, y, indexing="ij")
    return dx, dy, X.astype(dtype), Y.astype(dtype)


def parse_naca_4digit(designation: str):
    """parse_naca_4digit (obstacles/naca_airfoils.py:16-26)."""
    digits = "".join(ch for ch in designation if ch.isdigit())
    if len(digits) != 4:
        raise ValueError(f"Invalid 4-digit NACA designation: {designation}")
    return int(digits[0]) / 100.0, int(digits[1]) / 10.0, int(digits[2:]) / 100.0


def naca_sdf(X, Y, chord, aoa_deg, pos_x, pos_y, m, p, t):
    """naca_surface_distance (obstacles/naca_airfoils.py:109-178).

    Pseudo-SDF: rotate about the leading edge ``(pos_x, pos_y)`` by +AoA,
    vertical distance to the upper/lower surface inside the chord, distance to
    LE/TE outside, "inside" test with a +0.01 (chord units) tolerance.  For a
    symmetric section (m = p = 0) the camber terms are 0/0 = NaN wherever
    ``x_norm <= 0``; that NaN reaches the result only where ``x_norm == 0``
    exactly, as in the reference.
    """
    dt = X.dtype
    X = X.astype(dt)
    Y = Y.astype(dt)
    with np.errstate(divide="ignore", invalid="ignore"):
        Xl = X - dt.type(pos_x)
        Yl = Y - dt.type(pos_y)
        a = np.radians(dt.type(aoa_deg))
        ca, sa = np.cos(a), np.sin(a)
        Xr = Xl * ca - Yl * sa
        Yr = Xl * sa + Yl * ca
        c = dt.type(chord)
        xn = Xr / c
        yn = Yr / c
        in_chord = (xn >= 0) & (xn <= 1)
        xs = np.clip(xn, 0.0, 1.0)
        tt = dt.type(t)
        yt = 5 * tt * (0.2969 * np.sqrt(np.abs(xs)) - 0.1260 * xs
                       - 0.3516 * xs ** 2 + 0.2843 * xs ** 3 - 0.1015 * xs ** 4)
        mm, pp = dt.type(m), dt.type(p)
        yc = np.where(xn <= pp,
                      mm / pp ** 2 * (2 * pp * xn - xn ** 2),
                      mm / (1 - pp) ** 2 * ((1 - 2 * pp) + 2 * pp * xn - xn ** 2))
        dyc = np.where(xn <= pp,
                       2 * mm / pp ** 2 * (pp - xn),
                       2 * mm / (1 - pp) ** 2 * (pp - xn))
        th = np.arctan(dyc)
        yu = yc + yt * np.cos(th)
        yl = yc - yt * np.cos(th)
        d_up = np.abs(yn - yu) * c
        d_lo = np.abs(yn - yl) * c
        d_surf = np.minimum(d_up, d_lo)
        d_le = np.sqrt(Xr ** 2 + Yr ** 2)
        d_te = np.sqrt((Xr - c) ** 2 + Yr ** 2)
        dist = np.where(in_chord, d_surf, np.minimum(d_le, d_te))
        inside = in_chord & (yn >= yl - 0.01) & (yn <= yu + 0.01)
        return np.where(inside, -dist, dist)


def cylinder_sdf(X, Y, cx, cy, r):
    """Cylinder level set (solver/simulation/mask_generator.py:151-160;
    solver/geometry.py:10-13)."""
    dt = X.dtype
    return np.sqrt((X - dt.type(cx)) ** 2 + (Y - dt.type(cy)) ** 2) - dt.type(r)


def sigmoid(z):
    """jax.nn.sigmoid, 1/(1+exp(-z))."""
    with np.errstate(over="ignore"):
        return 1 / (1 + np.exp(-z))


def mask_from_sdf(sdf, eps):
    """mask = 1 - sigmoid(-sdf/eps) (solver/simulation/mask_generator.py:71-72,
    157-158)."""
    return 1.0 - sigmoid(-sdf / sdf.dtype.type(eps))


def von_karman_ic(nx, ny, U_inf=1.0, noise_sigma=0.0, seed=0, dtype=np.float32):
    """_initialize_von_karman_flow (solver/simulation/initializers.py:8-17):
    u = U_inf, v = 0.  Optional seeded Gaussian noise (a build-defined input;
    the reference has none) is drawn in float64 from
    ``np.random.default_rng(seed)`` (u noise first, then v) and cast."""
    u = np.full((nx, ny), U_inf, dtype=np.float64)
    v = np.zeros((nx, ny), dtype=np.float64)
    if noise_sigma > 0:
        rng = np.random.default_rng(seed)
        u = u + rng.normal(0.0, noise_sigma, size=(nx, ny))
        v = v + rng.normal(0.0, noise_sigma, size=(nx, ny))
    return u.astype(dtype), v.astype(dtype)


# ---------------------------------------------------------------------------
# Differential operators
# ---------------------------------------------------------------------------


def grad_x_periodic(f, dx):
    """grad_x (solver/operators.py:11-14): periodic central difference."""
    return (np.roll(f, -1, axis=0) - np.roll(f, 1, axis=0)) / (2.0 * f.dtype.type(dx))


def grad_y_periodic(f, dy):
    """grad_y (solver/operators.py:35-38)."""
    return (np.roll(f, -1, axis=1) - np.roll(f, 1, axis=1)) / (2.0 * f.dtype.type(dy))


def grad_x_nonperiodic(f, dx):
    """grad_x_nonperiodic (solver/operators.py:17-32): central inside,
    one-sided first order on rows 0 and nx-1."""
    dx = f.dtype.type(dx)
    g = np.zeros_like(f)
    g[1:-1, :] = (f[2:, :] - f[:-2, :]) / (2.0 * dx)
    g[0, :] = (f[1, :] - f[0, :]) / dx
    g[-1, :] = (f[-1, :] - f[-2, :]) / dx
    return g


def grad_y_nonperiodic(f, dy):
    """grad_y_nonperiodic (solver/operators.py:41-56)."""
    dy = f.dtype.type(dy)
    g = np.zeros_like(f)
    g[:, 1:-1] = (f[:, 2:] - f[:, :-2]) / (2.0 * dy)
    g[:, 0] = (f[:, 1] - f[:, 0]) / dy
    g[:, -1] = (f[:, -1] - f[:, -2]) / dy
    return g


def divergence_nonperiodic(u, v, dx, dy):
    """divergence_nonperiodic (solver/operators.py:81-96): central
    differences on ``[1:-1, 1:-1]``, zero on every boundary cell."""
    dx = u.dtype.type(dx)
    dy = u.dtype.type(dy)
    d = np.zeros_like(u)
    d[1:-1, 1:-1] = ((u[2:, 1:-1] - u[:-2, 1:-1]) / (2.0 * dx)
                     + (v[1:-1, 2:] - v[1:-1, :-2]) / (2.0 * dy))
    return d


def constant_smagorinsky(u, v, dx, dy, delta, C_s=0.1):
    """constant_smagorinsky + compute_strain_rate (solver/les_models.py:22-54,
    179-200).  Periodic ``jnp.roll`` gradients even for von Karman flow."""
    dt = u.dtype
    du_dx = grad_x_periodic(u, dx)
    du_dy = grad_y_periodic(u, dy)
    dv_dx = grad_x_periodic(v, dx)
    dv_dy = grad_y_periodic(v, dy)
    Sxx = du_dx
    Syy = dv_dy
    Sxy = 0.5 * (du_dy + dv_dx)
    S = np.sqrt(2.0 * (Sxx ** 2 + Syy ** 2 + 2.0 * Sxy ** 2) + 1e-10)
    return (dt.type(C_s) * dt.type(delta)) ** 2 * S


def strain_rate(u, v, dx, dy):
    """compute_strain_rate (solver/les_models.py:22-54), collocated:
    (|S|, Sxx, Syy, Sxy) with periodic central differences."""
    du_dx = grad_x_periodic(u, dx)
    du_dy = grad_y_periodic(u, dy)
    dv_dx = grad_x_periodic(v, dx)
    dv_dy = grad_y_periodic(v, dy)
    Sxx, Syy, Sxy = du_dx, dv_dy, 0.5 * (du_dy + dv_dx)
    S = np.sqrt(2.0 * (Sxx ** 2 + Syy ** 2 + 2.0 * Sxy ** 2) + 1e-10)
    return S, Sxx, Syy, Sxy


def box_filter_2d(f, filter_ratio=2):
    """box_filter_2d (solver/les_models.py:57-75): mean over the
    (2r+1) x (2r+1) periodic neighbourhood, accumulated in the reference's
    scan order (x offset outer, y offset inner, each -r..r) of
    roll(roll(f, i, 0), j, 1)."""
    acc = np.zeros_like(f)
    for i in range(-filter_ratio, filter_ratio + 1):
        for j in range(-filter_ratio, filter_ratio + 1):
            acc = acc + np.roll(np.roll(f, i, axis=0), j, axis=1)
    return acc / f.dtype.type((2 * filter_ratio + 1) ** 2)


def _local_average(f):
    """3x3 interior average, zero on the border (les_models.py:155-165)."""
    out = np.zeros_like(f)
    out[1:-1, 1:-1] = (f[1:-1, 1:-1] + f[2:, 1:-1] + f[:-2, 1:-1] + f[1:-1, 2:] + f[1:-1, :-2]
                       + f[2:, 2:] + f[2:, :-2] + f[:-2, 2:] + f[:-2, :-2]) / f.dtype.type(9.0)
    return out


def dynamic_smagorinsky(u, v, dx, dy, delta, alpha=2.0):
    """dynamic_smagorinsky (solver/les_models.py:78-176), collocated: Germano
    identity with the 5x5 test filter, L_ij and M_ij, 3x3-averaged least
    squares C_d = LM / (MM + 1e-8) clipped to [0, 0.2],
    nu_sgs = C_d delta^2 |S|.  Returns (nu_sgs, C_d)."""
    T = u.dtype.type
    S, Sxx, Syy, Sxy = strain_rate(u, v, dx, dy)
    ut, vt = box_filter_2d(u), box_filter_2d(v)
    St, Sxxt, Syyt, Sxyt = strain_rate(ut, vt, dx, dy)
    L11 = ut * ut - box_filter_2d(u * u)
    L12 = ut * vt - box_filter_2d(u * v)
    L22 = vt * vt - box_filter_2d(v * v)
    a2 = T(alpha) ** 2
    M11 = a2 * St * Sxxt - box_filter_2d(S * Sxx)
    M12 = a2 * St * Sxyt - box_filter_2d(S * Sxy)
    M22 = a2 * St * Syyt - box_filter_2d(S * Syy)
    LM = _local_average(L11 * M11 + L12 * M12 + L22 * M22)
    MM = _local_average(M11 * M11 + M12 * M12 + M22 * M22)
    C = np.clip(LM / (MM + T(1e-8)), 0.0, 0.2).astype(u.dtype)
    return C * (T(delta) * T(delta)) * S, C


# ---------------------------------------------------------------------------
# rk_step_unified (advection_schemes/rk3_simple_new.py:175-358)
# ---------------------------------------------------------------------------


def _pad_x(f):
    return np.pad(f, ((1, 1), (0, 0)), mode="edge")


def _pad_y(f):
    return np.pad(f, ((0, 0), (1, 1)), mode="edge")


def _laplacian_edge(f, dx, dy):
    """laplacian (rk3_simple_new.py:215-219), edge padding."""
    fp = np.pad(f, ((1, 1), (1, 1)), mode="edge")
    lap_x = (fp[2:, 1:-1] + fp[:-2, 1:-1] - 2 * f) / dx ** 2
    lap_y = (fp[1:-1, 2:] + fp[1:-1, :-2] - 2 * f) / dy ** 2
    return lap_x + lap_y


def _upwind(field, vu, vv, dx, dy):
    """upwind_advection (rk3_simple_new.py:287-298); ``vel > 0`` selects the
    backward difference, otherwise the forward one."""
    fp = _pad_x(field)
    fx = np.where(vu > 0, vu * (field - fp[:-2, :]) / dx, vu * (fp[2:, :] - field) / dx)
    fpy = _pad_y(field)
    fy = np.where(vv > 0, vv * (field - fpy[:, :-2]) / dy, vv * (fpy[:, 2:] - field) / dy)
    return fx + fy


def apply_velocity_bcs(u, v, U_inf, slip_walls=True):
    """apply_boundary_conditions, von Karman branch (rk3_simple_new.py:264-279):
    inlet/outlet rows first, then wall columns."""
    u = u.copy()
    v = v.copy()
    u[0, :] = U_inf
    v[0, :] = 0.0
    u[-1, :] = u[-2, :]
    v[-1, :] = v[-2, :]
    if slip_walls:
        u[:, 0] = u[:, 1]
        u[:, -1] = u[:, -2]
    else:
        u[:, 0] = 0.0
        u[:, -1] = 0.0
    v[:, 0] = 0.0
    v[:, -1] = 0.0
    return u, v


def brinkman_chi(sdf, mask, dx):
    """chi (rk3_simple_new.py:194-200): sigmoid(-sdf/(0.1 dx)) if an SDF is
    given, else 1 - mask."""
    if sdf is not None:
        return sigmoid(-sdf / (0.1 * sdf.dtype.type(dx)))
    return 1.0 - mask


def rk_step_unified(u, v, dt, nu, dx, dy, mask, sdf=None, U_inf=1.0, nu_sgs=None,
                    nu_hyper_ratio=0.0, slip_walls=True, fast_mode=False,
                    brinkman_eta=0.01):
    """rk_step_unified (advection_schemes/rk3_simple_new.py:175-358),
    flow_type='von_karman'.

    RK3 (fast_mode False, :315-337) restated literally: stage 2 and 3 restart
    from ``u`` (``u_star = u + dt*k``), each stage divided by the implicit
    Brinkman factor ``1 + dt*chi/eta``; RK2/Heun for fast_mode (:339-353).
    No BC or masking between stages; ``where(mask > 0.5)`` then the BCs at
    the end.  The sponge field is identically 1 (:282).
    """
    T = u.dtype.type
    dt, dx, dy = T(dt), T(dx), T(dy)
    nu = nu if isinstance(nu, np.ndarray) else T(nu)  # a field in the GUI step
    U_inf, eta = T(U_inf), T(brinkman_eta)
    chi = brinkman_chi(sdf, mask, dx)
    D = 1 + dt * chi / eta

    def rhs(ui, vi):
        adv_u = _upwind(ui, ui, vi, dx, dy)
        adv_v = _upwind(vi, ui, vi, dx, dy)
        lap_u = _laplacian_edge(ui, dx, dy)
        lap_v = _laplacian_edge(vi, dx, dy)
        nu_tot = nu if nu_sgs is None else nu + nu_sgs
        ru = -adv_u + nu_tot * lap_u
        rv = -adv_v + nu_tot * lap_v
        if nu_hyper_ratio > 0:
            nu_h = nu * T(nu_hyper_ratio) * (dx * dy)
            ru = ru - nu_h * 1.0 * _laplacian_edge(lap_u, dx, dy)
            rv = rv - nu_h * 1.0 * _laplacian_edge(lap_v, dx, dy)
        return ru, rv

    if not fast_mode:
        k1u, k1v = rhs(u, v)
        u2 = (u + dt * k1u) / D
        v2 = (v + dt * k1v) / D
        k2u, k2v = rhs(u2, v2)
        us = u + dt * k2u
        vs = v + dt * k2v
        u3 = (0.75 * u + 0.25 * us) / D
        v3 = (0.75 * v + 0.25 * vs) / D
        k3u, k3v = rhs(u3, v3)
        us = u + dt * k3u
        vs = v + dt * k3v
        un = ((1 / 3) * u + (2 / 3) * us) / D
        vn = ((1 / 3) * v + (2 / 3) * vs) / D
    else:
        k1u, k1v = rhs(u, v)
        us = (u + dt * k1u) / D
        vs = (v + dt * k1v) / D
        k2u, k2v = rhs(us, vs)
        un = (u + 0.5 * dt * (k1u + k2u)) / D
        vn = (v + 0.5 * dt * (k1v + k2v)) / D
    un = np.where(mask > 0.5, un, T(0))
    vn = np.where(mask > 0.5, vn, T(0))
    return apply_velocity_bcs(un, vn, U_inf, slip_walls)


# ---------------------------------------------------------------------------
# poisson_multigrid, reference mode (pressure_solvers/multigrid_solver.py:55-241)
# ---------------------------------------------------------------------------


def mg_max_levels(nx, ny, levels):
    """max_levels loop (multigrid_solver.py:96-106)."""
    ml = 1
    tx, ty = nx, ny
    while tx % 2 == 0 and ty % 2 == 0 and ml < levels:
        tx //= 2
        ty //= 2
        ml += 1
    if ml < 2:
        raise ValueError(
            f"Grid dimensions ({nx}, {ny}) not suitable for multigrid with {levels} levels. "
            f"Grid must be divisible by 2^{levels}")
    return ml


def mg_restrict(fine):
    """restrict (multigrid_solver.py:108-132): 2x2 box average; a level whose
    coarse side would be < 2 collapses to ``fine[:1, :1]``."""
    nxf, nyf = fine.shape
    nxc, nyc = nxf // 2, nyf // 2
    if nxc < 2 or nyc < 2:
        return fine[:1, :1].copy()
    fe = fine[:nxc * 2, :nyc * 2]
    return 0.25 * (fe[::2, ::2] + fe[1::2, ::2] + fe[::2, 1::2] + fe[1::2, 1::2])


def mg_prolong(coarse):
    """prolong (multigrid_solver.py:134-156): injection at (even, even),
    linear along odd rows / odd columns, bilinear at interior (odd, odd);
    the last fine row and column stay zero."""
    nxc, nyc = coarse.shape
    nxf, nyf = 2 * nxc, 2 * nyc
    F = np.zeros((nxf, nyf), dtype=coarse.dtype)
    F[::2, ::2] = coarse
    if nxf > 2:
        F[1:-1:2, ::2] = 0.5 * (F[0:-2:2, ::2] + F[2::2, ::2])
    if nyf > 2:
        F[::2, 1:-1:2] = 0.5 * (F[::2, 0:-2:2] + F[::2, 2::2])
    if nxf > 2 and nyf > 2:
        F[1:-1:2, 1:-1:2] = 0.25 * (F[0:-2:2, 0:-2:2] + F[2::2, 0:-2:2]
                                    + F[0:-2:2, 2::2] + F[2::2, 2::2])
    return F


def mg_apply_laplacian(p, ax, ay):
    """apply_laplacian (multigrid_solver.py:173-183), edge padding."""
    pp = np.pad(p, ((1, 1), (1, 1)), mode="edge")
    return (ax * (pp[2:, 1:-1] + pp[:-2, 1:-1] - 2 * p)
            + ay * (pp[1:-1, 2:] + pp[1:-1, :-2] - 2 * p))


def mg_smooth(p, b, ax, ay, nu):
    """smooth (multigrid_solver.py:158-171): ``nu`` Jacobi sweeps (omega=1),
    edge padding, fine-grid ax, ay at every level."""
    for _ in range(nu):
        pp = np.pad(p, ((1, 1), (1, 1)), mode="edge")
        p = (ax * (pp[2:, 1:-1] + pp[:-2, 1:-1]) + ay * (pp[1:-1, 2:] + pp[1:-1, :-2]) - b) \
            / (2.0 * (ax + ay))
    return p


def mg_level_shapes(nx, ny, levels):
    """Shapes visited by the reference recursion, with the shape check that
    the reference only fails at trace time (``p + e`` broadcast)."""
    ml = mg_max_levels(nx, ny, levels)
    shapes = [(nx, ny)]
    for _ in range(ml):
        sx, sy = shapes[-1]
        cx, cy = sx // 2, sy // 2
        if cx < 2 or cy < 2:
            cx, cy = 1, 1
        if (2 * cx, 2 * cy) != (sx, sy):
            raise ValueError(
                f"Grid dimensions ({nx}, {ny}) not suitable for multigrid with {levels} levels. "
                f"Grid must be divisible by 2^{levels}")
        shapes.append((cx, cy))
    return ml, shapes


def poisson_multigrid(rhs, mask, dx, dy, levels=4, v_cycles=5, tolerance=1e-6,
                      p0=None, return_stats=False):
    """poisson_multigrid (multigrid_solver.py:55-241), von Karman branch.

    ``b[-1,:] = 0``; V-cycle: levels < max_levels pre-smooth 2 Jacobi sweeps,
    clamp the outlet row, restrict ``b - A p``, recurse from zero, prolong,
    add, post-smooth 2, clamp; the coarsest level runs 10 sweeps.  The cycle
    loop stops cycling once ``sum(r**2)*dx*dy < tolerance**2`` (residual over
    the whole grid, outlet row included).  ``p0`` (default None = zeros,
    multigrid_solver.py:219) is this project's warm-start superset.
    """
    T = rhs.dtype.type
    dx, dy = T(dx), T(dy)
    nx, ny = rhs.shape
    ml, _ = mg_level_shapes(nx, ny, levels)
    ax = T(1.0) / (dx * dx)
    ay = T(1.0) / (dy * dy)
    b = rhs.copy()
    b[-1, :] = 0.0

    def vcycle(p, bl, level):
        if level >= ml:
            return mg_smooth(p, bl, ax, ay, 10)
        p = mg_smooth(p, bl, ax, ay, 2)
        p = p.copy()
        p[-1, :] = 0.0
        r = bl - mg_apply_laplacian(p, ax, ay)
        rc = mg_restrict(r)
        e = mg_prolong(vcycle(np.zeros_like(rc), rc, level + 1))
        p = p + e
        p = mg_smooth(p, bl, ax, ay, 2)
        p = p.copy()
        p[-1, :] = 0.0
        return p

    p = np.zeros_like(b) if p0 is None else p0.astype(rhs.dtype).copy()
    converged = False
    cycles = 0
    norms = []
    tol2 = T(tolerance) ** 2
    for _ in range(v_cycles):
        if not converged:
            p = vcycle(p, b, 0)
            cycles += 1
        r = b - mg_apply_laplacian(p, ax, ay)
        nsq = np.sum(r.astype(np.float64) ** 2) * float(dx) * float(dy)
        norms.append(nsq)
        converged = bool(nsq < float(tol2))
    if return_stats:
        return p, {"cycles": cycles, "norm_sq": norms, "converged": converged}
    return p


# ---------------------------------------------------------------------------
# step_pure, collocated von Karman branch (solver/pure_functions/step.py:49-467)
# ---------------------------------------------------------------------------


def project(u_star, v_star, p, mask, dt, dx, dy, U_inf):
    """Projection tail of step_pure (step.py:257-264, 305-349, 364-377,
    394-397): one-sided-at-boundary gradients of p; tangential projection of
    grad p at interface cells (|grad mask| > 0.1/dx); dp/dx = 0 on the inlet
    row; ``u = u* - dt grad p``; mask; slip/inlet/outlet BCs; mask."""
    T = u_star.dtype.type
    dt, dx, dy, U_inf = T(dt), T(dx), T(dy), T(U_inf)
    gx = grad_x_nonperiodic(p, dx)
    gy = grad_y_nonperiodic(p, dy)
    mx = grad_x_nonperiodic(mask, dx)
    my = grad_y_nonperiodic(mask, dy)
    gm = np.sqrt(mx ** 2 + my ** 2)
    iface = gm > 0.1 / dx
    nxv = mx / (gm + 1e-8)
    nyv = my / (gm + 1e-8)
    dpn = gx * nxv + gy * nyv
    gxp = gx - dpn * nxv
    gyp = gy - dpn * nyv
    gx = np.where(iface, gxp, gx)
    gy = np.where(iface, gyp, gy)
    gx[0, :] = 0.0
    un = u_star - dt * gx
    vn = v_star - dt * gy
    un = np.where(mask > 0.5, un, T(0))
    vn = np.where(mask > 0.5, vn, T(0))
    un, vn = apply_velocity_bcs(un, vn, U_inf, slip_walls=True)
    un = np.where(mask > 0.5, un, T(0))
    vn = np.where(mask > 0.5, vn, T(0))
    return un, vn


def predictor_bcs(u_star, v_star, U_inf):
    """Inlet + slip walls re-imposed on the predictor (step.py:182-191)."""
    u_star = u_star.copy()
    v_star = v_star.copy()
    u_star[0, :] = U_inf
    v_star[0, :] = 0.0
    u_star[:, 0] = u_star[:, 1]
    u_star[:, -1] = u_star[:, -2]
    v_star[:, 0] = 0.0
    v_star[:, -1] = 0.0
    return u_star, v_star


def pressure_rhs(u_star, v_star, mask, dt, dx, dy):
    """rhs = divergence_nonperiodic(u*, v*) * mask / dt (step.py:194-199)."""
    T = u_star.dtype.type
    return (divergence_nonperiodic(u_star, v_star, dx, dy) * mask) / T(dt)


def step_pure(u, v, mask, sdf, dx, dy, nu, U_inf, dt, use_les=False,
              smagorinsky_constant=0.1, v_cycles=1, fast_mode=False,
              brinkman_eta=0.005, les_coupling="discard", levels=4,
              tolerance=1e-6, p0=None, return_parts=False):
    """step_pure, collocated von Karman branch (step.py:49-467).

    ``les_coupling="discard"`` is the reference composition: nu_sgs is
    computed (step.py:111-122) and not passed on (``nu_sgs=None``,
    step.py:177-180).  ``"apply"`` passes it to ``rk_step_unified`` (the RK
    function's own contract, rk3_simple_new.py:304).
    Returns ``(u_new, v_new, p)``.
    """
    T = u.dtype.type
    nu_sgs = None
    if use_les:
        delta = (dx * dy) ** 0.5
        nu_sgs = constant_smagorinsky(u, v, dx, dy, delta, smagorinsky_constant)
    us, vs = rk_step_unified(u, v, dt, nu, dx, dy, mask, sdf=sdf, U_inf=U_inf,
                             nu_sgs=nu_sgs if les_coupling == "apply" else None,
                             nu_hyper_ratio=0.0, slip_walls=True, fast_mode=fast_mode,
                             brinkman_eta=brinkman_eta)
    us, vs = predictor_bcs(us, vs, T(U_inf))
    rhs = pressure_rhs(us, vs, mask, dt, dx, dy)
    p = poisson_multigrid(rhs, mask, dx, dy, levels=levels, v_cycles=v_cycles,
                          tolerance=tolerance, p0=p0)
    un, vn = project(us, vs, p, mask, dt, dx, dy, U_inf)
    if return_parts:
        return un, vn, p, {"u_star": us, "v_star": vs, "rhs": rhs, "nu_sgs": nu_sgs}
    return un, vn, p


# ---------------------------------------------------------------------------
# GUI step _step_collocated (solver/simulation/step_handlers.py:62-185),
# von Karman branch
# ---------------------------------------------------------------------------


def corner_smooth_pressure_gradient(dp_dx, ny, width=10):
    """apply_corner_smooth_pressure_gradient
    (solver/pure_functions/corner_smoothing.py:39-61): row 0 entries j and
    ny-1-j, j < width, scaled by sin^2(pi/2 * j/width)."""
    T = dp_dx.dtype.type
    dp_dx = dp_dx.copy()
    j = np.arange(min(width, ny))  # narrower grids: in-range entries only
    f = (np.sin(np.pi / 2 * (np.arange(width) / width)) ** 2)[j].astype(T)
    dp_dx[0, j] = dp_dx[0, j] * f
    top = ny - 1 - j
    dp_dx[0, top] = dp_dx[0, top] * f
    return dp_dx


def project_collocated(u_star, v_star, p, mask, dt, dx, dy, U_in):
    """Pressure correction and BCs of _step_collocated (step_handlers.py:
    131-176): dp_dx[0, :] = 0 then the corner taper; u = u* - dt grad p;
    inlet u[0, 1:-1] = U_in, v[0, :] = 0; outlet copies row nx-2; no-slip
    walls (corners included); hard mask.  No interface projection."""
    T = u_star.dtype.type
    dt, dx, dy, U_in = T(dt), T(dx), T(dy), T(U_in)
    ny = p.shape[1]
    gx = grad_x_nonperiodic(p, dx)
    gy = grad_y_nonperiodic(p, dy)
    gx[0, :] = 0.0
    gx = corner_smooth_pressure_gradient(gx, ny, 10)
    un = u_star - dt * gx
    vn = v_star - dt * gy
    un[0, 1:-1] = U_in
    vn[0, :] = 0.0
    un[-1, :] = un[-2, :]
    vn[-1, :] = vn[-2, :]
    un[:, 0] = 0.0
    un[:, -1] = 0.0
    vn[:, 0] = 0.0
    vn[:, -1] = 0.0
    un = np.where(mask > 0.5, un, T(0))
    vn = np.where(mask > 0.5, vn, T(0))
    return un, vn


def step_collocated(u, v, mask, sdf, dx, dy, nu, U_inf, dt, use_les=False,
                    smagorinsky_constant=0.17, nu_hyper_ratio=0.0, slip_walls=True,
                    fast_mode=False, brinkman_eta=0.01, v_cycles=5, solver_dt=None,
                    U_in=None, levels=4, return_parts=False, les_model="smagorinsky"):
    """_step_collocated (step_handlers.py:62-185), constant or dynamic
    Smagorinsky (:71-77), multigrid pressure.  With LES the RK step gets
    nu + nu_sgs as ``nu`` and nu_sgs as ``nu_sgs`` (:76-90); rhs = div / dt
    without the mask (:92-98).  Returns ``(u, v, p)``."""
    T = u.dtype.type
    nu_arg, nu_sgs = nu, None
    if use_les:
        delta = (dx * dy) ** 0.5
        if les_model == "dynamic_smagorinsky":
            nu_sgs, _ = dynamic_smagorinsky(u, v, dx, dy, delta)
        else:
            nu_sgs = constant_smagorinsky(u, v, dx, dy, delta, smagorinsky_constant)
        nu_arg = T(nu) + nu_sgs
    us, vs = rk_step_unified(u, v, dt, nu_arg, dx, dy, mask, sdf=sdf, U_inf=U_inf,
                             nu_sgs=nu_sgs, nu_hyper_ratio=nu_hyper_ratio,
                             slip_walls=slip_walls, fast_mode=fast_mode,
                             brinkman_eta=brinkman_eta)
    rhs = divergence_nonperiodic(us, vs, dx, dy) / T(dt)
    p = poisson_multigrid(rhs, mask, dx, dy, levels=levels, v_cycles=v_cycles)
    un, vn = project_collocated(us, vs, p, mask, dt if solver_dt is None else solver_dt, dx, dy,
                                U_inf if U_in is None else U_in)
    if return_parts:
        return un, vn, p, {"u_star": us, "v_star": vs, "rhs": rhs, "nu_sgs": nu_sgs}
    return un, vn, p


def laplacian_nonperiodic_x(f, dx, dy):
    """laplacian_nonperiodic_x (solver/operators.py:67-72): edge padding in
    x, periodic roll in y."""
    fp = np.concatenate([f[:1], f, f[-1:]], axis=0)
    f_xx = (fp[2:] - 2 * f + fp[:-2]) / (dx * dx)
    f_yy = (np.roll(f, 1, axis=1) + np.roll(f, -1, axis=1) - 2 * f) / (dy * dy)
    return f_xx + f_yy


def check_cfl(u, v, dt, dx, dy):
    """check_cfl (advection_schemes/utils.py:14-20): the largest speed and
    max(speed dt / dx, speed dt / dy).  Returns (max_speed, cfl)."""
    vmax = float(np.max(np.sqrt(np.asarray(u, np.float64) ** 2 + np.asarray(v, np.float64) ** 2)))
    return vmax, max(vmax * dt / dx, vmax * dt / dy)


def flow_diagnostics(u, v, p, mask, dx, dy, dt):
    """Viewer / debug metrics (viewer/simulation_controller.py:107-121,
    186-194): pressure residual of lap_x(p) against div_nonperiodic(u, v)/dt,
    the np.gradient divergence over mask > 0.99, and check_cfl."""
    u, v, p, mask = (np.asarray(a, dtype=np.float64) for a in (u, v, p, mask))
    rhs = divergence_nonperiodic(u, v, dx, dy) / dt
    res = np.linalg.norm(laplacian_nonperiodic_x(p, dx, dy) - rhs)
    rn = np.linalg.norm(rhs)
    div = np.gradient(u, dx, axis=0) + np.gradient(v, dy, axis=1)
    fluid = mask > 0.99
    df = div * fluid
    return {"pressure_residual": res, "rhs_norm": rn, "relative_residual": res / (rn + 1e-10),
            "div_rms": np.sqrt(np.sum(df ** 2) / (np.sum(fluid) + 1e-8)),
            "l2_div": np.sqrt(np.sum(df ** 2) * dx * dy), "fluid_cells": float(np.sum(fluid)),
            **dict(zip(("max_speed", "cfl"), check_cfl(u, v, dt, dx, dy)))}


def rel_l2(a, b):
    """Relative L2 error ||a - b|| / ||b|| in float64 (0 when both vanish)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    na = np.linalg.norm(a - b)
    if nb == 0:
        return na
    return na / nb
