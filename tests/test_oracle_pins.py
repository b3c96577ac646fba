"""Closed-form pins of the oracle's RK stages, upwind branches, interface
projection and step composition (the reference ships no tests or golden
vectors for this path, so each expected value below comes from mathematics
or a hand formula, never from the oracle itself).

RK stages (advection_schemes/rk3_simple_new.py:215-219, 301-353): with
u = U0 + a*phi(j), phi(j) = cos(pi k (j + 1/2) / ny) and v = 0 the upwind
advection vanishes (u does not vary along x, v = 0) and phi is an
eigenvector of the edge-padded 5-point Laplacian along y with eigenvalue
lam = (2 cos(pi k / ny) - 2) / dy^2 (the discrete cosine basis of the
Neumann second difference); lap(lap phi) = lam^2 phi.  Each stage then acts
on the amplitude a by z = dt (nu_tot lam - nu r dx dy lam^2) and on the
constant by nothing, both divided by the implicit Brinkman factor D:
    RK3:  G = (1 + (2/3) z (1 + z (1 + z) / (4 D)) / D) / D
    RK2:  G = (1 + z (1 + (1 + z) / D) / 2) / D
so the interior of u* is U0 / D + G a phi(j).  With D = 1 these are
1 + 2z/3 + z^2/6 + z^3/6 (not SSP-RK3's 1 + z + z^2/2 + z^3/6) and
1 + z + z^2/2.  A swapped stage weight, a wrong biharmonic sign or
coefficient, a dropped nu_sgs, or a Brinkman division in the wrong place
changes G at O(z) or O(z^2); |z| ~ 0.3 here.
"""
import math

import numpy as np
import pytest

from oracle import reference_step as R


def _g_rk3(z, D):
    return (1.0 + (2.0 / 3.0) * z * (1.0 + z * (1.0 + z) / (4.0 * D)) / D) / D


def _g_rk2(z, D):
    return (1.0 + 0.5 * z * (1.0 + (1.0 + z) / D)) / D


def _check_bcs_u(us, U_inf):
    """Von Karman BCs of rk_step_unified (rk3_simple_new.py:264-279, slip)."""
    assert np.all(us[0] == U_inf)
    assert np.array_equal(us[:, 0], us[:, 1]) and np.array_equal(us[:, -1], us[:, -2])


CASES = [
    # fast_mode, D != 1 (uniform chi from a constant sdf), nu_sgs, hyper ratio
    (False, False, 0.0, 0.0),
    (False, True, 0.0, 0.0),
    (False, True, 3e-3, 0.0),
    (False, False, 0.0, 0.2),
    (False, True, 2e-3, 0.15),
    (True, False, 0.0, 0.0),
    (True, True, 2e-3, 0.15),
]


@pytest.mark.parametrize("fast,brinkman,nu_sgs,ratio", CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rk_stages_on_a_y_cosine_mode(fast, brinkman, nu_sgs, ratio, dtype):
    nx, ny, k = 10, 32, 9
    dx, dy = 1.0 / 24, 1.0 / 32
    nu, dt, eta = 1e-2, 4e-2, 0.05
    U0, a = 0.8, 0.3
    j = np.arange(ny)
    phi = np.cos(np.pi * k * (j + 0.5) / ny)
    lam = (2.0 * math.cos(np.pi * k / ny) - 2.0) / dy ** 2
    u = np.broadcast_to(U0 + a * phi, (nx, ny)).astype(dtype)
    v = np.zeros((nx, ny), dtype)
    mask = np.ones((nx, ny), dtype)
    if brinkman:
        s = 2e-3  # chi = sigmoid(-s / (0.1 dx)), uniform
        sdf = np.full((nx, ny), s, dtype)
        chi = 1.0 / (1.0 + math.exp(s / (0.1 * dx)))
    else:
        sdf, chi = None, 0.0  # chi = 1 - mask = 0
    D = 1.0 + dt * chi / eta
    sgs = np.full((nx, ny), nu_sgs, dtype) if nu_sgs else None
    z = dt * ((nu + nu_sgs) * lam - nu * ratio * dx * dy * lam ** 2)
    assert 0.05 < abs(z) < 0.9  # the stage structure is visible at O(z^2)
    G = _g_rk2(z, D) if fast else _g_rk3(z, D)
    us, vs = R.rk_step_unified(u, v, dt, nu, dx, dy, mask, sdf=sdf, U_inf=U0, nu_sgs=sgs,
                               nu_hyper_ratio=ratio, fast_mode=fast, brinkman_eta=eta)
    expect = U0 / D + G * a * phi
    tol = 1e-12 if dtype == np.float64 else 2e-6
    inner = us[1:-1, 1:-1].astype(np.float64)
    assert np.allclose(inner, expect[None, 1:-1], rtol=tol, atol=tol)
    assert np.array_equal(us[-1], us[-2])
    _check_bcs_u(us, dtype(U0))
    assert np.all(vs == 0)
    # the reference scheme is not SSP-RK3 / plain RK3 on the same stage data
    if not fast and not brinkman:
        ssp = 1.0 + z + z * z / 2 + z ** 3 / 6
        assert abs(ssp - G) * abs(a) > 100 * tol


@pytest.mark.parametrize("fast", [False, True])
def test_rk_stages_on_an_x_cosine_mode_with_a_viscosity_field(fast):
    """The x-direction counterpart (edge padding at the inlet/outlet rows)
    on v with u = 0, and the GUI step's viscosity given as a field: the
    Laplacian sees nu_field + nu_sgs, the biharmonic nu_field
    (rk3_simple_new.py:304-309)."""
    nx, ny, k = 40, 6, 11
    dx, dy = 1.0 / 40, 1.0 / 20
    nuf, sgs, ratio, dt = 8e-3, 1e-3, 0.1, 2e-2
    i = np.arange(nx)
    phi = np.cos(np.pi * k * (i + 0.5) / nx)
    lam = (2.0 * math.cos(np.pi * k / nx) - 2.0) / dx ** 2
    a = 0.25
    u = np.zeros((nx, ny))
    v = np.broadcast_to((a * phi)[:, None], (nx, ny)).copy()
    mask = np.ones((nx, ny))
    z = dt * ((nuf + sgs) * lam - nuf * ratio * dx * dy * lam ** 2)
    assert 0.05 < abs(z) < 0.9
    G = _g_rk2(z, 1.0) if fast else _g_rk3(z, 1.0)
    us, vs = R.rk_step_unified(u, v, dt, np.full((nx, ny), nuf), dx, dy, mask, sdf=None,
                               U_inf=0.0, nu_sgs=np.full((nx, ny), sgs), nu_hyper_ratio=ratio,
                               fast_mode=fast)
    assert np.allclose(vs[1:-1, 1:-1], (G * a * phi)[1:-1, None], rtol=1e-12, atol=1e-14)
    assert np.all(vs[0] == 0) and np.all(vs[:, 0] == 0) and np.all(vs[:, -1] == 0)
    assert np.array_equal(vs[-1, 1:-1], vs[-2, 1:-1])
    assert np.all(us == 0)


# ------------------------------------------------------------ upwind branches

def _first_stage_k(u, v, dx, dy):
    """(u* - u)/dt of RK2 at a vanishing dt with nu = 0 and chi = 0: the
    explicit rhs -upwind(u; u, v) (and the same for v) up to O(dt)."""
    dt = 1e-9
    mask = np.ones_like(u)
    us, vs = R.rk_step_unified(u, v, dt, 0.0, dx, dy, mask, sdf=None, U_inf=float(u[0, 0]),
                               fast_mode=True)
    return (us - u) / dt, (vs - v) / dt


def test_upwind_negative_x_velocity_takes_the_forward_difference():
    nx, ny, dx, dy = 11, 6, 0.2, 0.3
    x = np.arange(nx)[:, None] * np.ones((1, ny))
    u = -0.5 - 0.2 * np.sin(0.7 * x)  # < 0 everywhere
    v = np.zeros_like(u)
    ku, kv = _first_stage_k(u, v, dx, dy)
    hand = np.empty_like(u)
    for i in range(1, nx - 1):
        hand[i] = -u[i] * (u[i + 1] - u[i]) / dx  # forward difference
    assert np.allclose(ku[1:-1, 1:-1], hand[1:-1, 1:-1], rtol=1e-5, atol=1e-7)
    assert np.allclose(kv[1:-1, 1:-1], 0.0, atol=1e-7)


def test_upwind_mixed_sign_x_velocity():
    nx, ny, dx, dy = 13, 5, 0.1, 0.25
    x = np.arange(nx)[:, None] * np.ones((1, ny))
    u = 0.4 * np.sin(0.9 * x + 0.3)  # both signs
    v = np.zeros_like(u)
    ku, _ = _first_stage_k(u, v, dx, dy)
    for i in range(1, nx - 1):
        c = u[i, 0]
        d = (u[i, 0] - u[i - 1, 0]) if c > 0 else (u[i + 1, 0] - u[i, 0])
        assert math.isclose(ku[i, 2], -c * d / dx, rel_tol=1e-5, abs_tol=1e-7), i


def test_upwind_y_direction_both_signs_advects_u_and_v_with_v():
    """u = c(j) > 0 and v = w(j) of both signs, neither varying along x:
    only the y terms remain, k_u = -w du/dy and k_v = -w dv/dy with the
    backward difference where w > 0 and the forward one elsewhere."""
    nx, ny, dx, dy = 6, 14, 0.3, 0.1
    j = np.arange(ny)
    c = 1.0 + 0.3 * np.cos(0.8 * j)
    w = 0.5 * np.sin(0.6 * j + 0.4)
    u = np.broadcast_to(c, (nx, ny)).copy()
    v = np.broadcast_to(w, (nx, ny)).copy()
    ku, kv = _first_stage_k(u, v, dx, dy)
    for jj in range(1, ny - 1):
        if w[jj] > 0:
            du, dv = c[jj] - c[jj - 1], w[jj] - w[jj - 1]
        else:
            du, dv = c[jj + 1] - c[jj], w[jj + 1] - w[jj]
        assert math.isclose(ku[2, jj], -w[jj] * du / dy, rel_tol=1e-5, abs_tol=1e-6), jj
        assert math.isclose(kv[2, jj], -w[jj] * dv / dy, rel_tol=1e-5, abs_tol=1e-6), jj
    assert (w[1:-1] > 0).any() and (w[1:-1] < 0).any()


# ---------------------------------------------------------- interface projection

def _bilinear_p(nx, ny, dx, dy, a=1.7, b=-0.9, c=0.6):
    """p = a x + b y + c x y: every one-sided and central difference is exact,
    dp/dx = a + c y, dp/dy = b + c x (x = i dx, y = j dy)."""
    x = np.arange(nx)[:, None] * dx
    y = np.arange(ny)[None, :] * dy
    p = a * x + b * y + c * x * y
    return p, a + c * y + 0 * x, b + c * x + 0 * y


def _expect_bcs(un, vn, U_inf):
    assert np.all(un[0] == U_inf) and np.all(vn[0] == 0)
    assert np.array_equal(un[-1, 1:-1], un[-2, 1:-1]) and np.array_equal(vn[-1, 1:-1], vn[-2, 1:-1])
    assert np.array_equal(un[:, 0], un[:, 1]) and np.array_equal(un[:, -1], un[:, -2])
    assert np.all(vn[:, 0] == 0) and np.all(vn[:, -1] == 0)


def test_project_x_ramp_removes_dpdx_only_above_the_threshold():
    """Mask rising along x by 0.3 per row up to row 5, then by 0.05:
    |grad mask| = 0.3/dx (central), 0.175/dx at row 5 and 0.05/dx beyond;
    the interface test is |grad mask| > 0.1/dx (step.py:311-312), the normal
    is (1, 0), so at interface rows dp/dx is removed and dp/dy kept; rows
    beyond get the full gradient."""
    nx, ny, dx, dy, dt, U = 14, 8, 0.05, 0.05, 0.01, 1.0
    inc = np.array([0.3 if i < 5 else 0.05 for i in range(nx - 1)])
    mrow = 0.6 + np.concatenate([[0.0], np.cumsum(inc)])
    mask = np.broadcast_to(mrow[:, None], (nx, ny)).copy()
    p, gx, gy = _bilinear_p(nx, ny, dx, dy)
    rng = np.random.default_rng(0)
    us = 1.0 + 0.1 * rng.normal(size=(nx, ny))
    vs = 0.1 * rng.normal(size=(nx, ny))
    un, vn = R.project(us, vs, p, mask, dt, dx, dy, U)
    _expect_bcs(un, vn, U)
    for i in range(1, nx - 1):
        central = (mrow[i + 1] - mrow[i - 1]) / (2 * dx)
        iface = central > 0.1 / dx
        assert iface == (i <= 5), i
        du = un[i, 1:-1] - us[i, 1:-1]
        if iface:
            assert np.all(np.abs(du) <= 1e-7 * dt * np.abs(gx[i, 1:-1]) + 1e-15), i
        else:
            assert np.allclose(du, -dt * gx[i, 1:-1], rtol=1e-12), i
        assert np.allclose(vn[i, 1:-1] - vs[i, 1:-1], -dt * gy[i, 1:-1], rtol=1e-12), i


def test_project_y_ramp_threshold_uses_dx():
    """Mask rising along y by 0.08 per column on a grid with dx = 2 dy:
    |grad mask| = 0.08/dy = 1.6 > 0.1/dx = 1.0, so every cell is an interface
    cell (with 0.1/dy = 2.0 as threshold none would be); the normal is (0, 1),
    so dp/dy is removed and dp/dx kept."""
    nx, ny, dx, dy, dt, U = 9, 10, 0.1, 0.05, 0.02, 1.0
    mcol = 0.55 + 0.08 * np.arange(ny)
    mask = np.broadcast_to(mcol[None, :], (nx, ny)).copy()
    p, gx, gy = _bilinear_p(nx, ny, dx, dy)
    rng = np.random.default_rng(1)
    us = 1.0 + 0.1 * rng.normal(size=(nx, ny))
    vs = 0.1 * rng.normal(size=(nx, ny))
    un, vn = R.project(us, vs, p, mask, dt, dx, dy, U)
    _expect_bcs(un, vn, U)
    dv = vn[1:-1, 1:-1] - vs[1:-1, 1:-1]
    assert np.all(np.abs(dv) <= 1e-7 * dt * np.abs(gy[1:-1, 1:-1]) + 1e-15)
    assert np.allclose(un[1:-1, 1:-1] - us[1:-1, 1:-1], -dt * gx[1:-1, 1:-1], rtol=1e-12)


def test_project_diagonal_ramp_keeps_the_tangential_component():
    """Mask rising by 0.3 per cell along x and along y (dx = dy): normal
    (1, 1)/sqrt 2 everywhere, so the applied gradient is the projection of
    grad p on the tangent (1, -1)/sqrt 2: ((gx - gy)/2, (gy - gx)/2)."""
    nx, ny, dx, dt, U = 10, 9, 0.04, 0.01, 1.0
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    mask = 0.6 + 0.3 * (ii + jj)
    p, gx, gy = _bilinear_p(nx, ny, dx, dx)
    rng = np.random.default_rng(2)
    us = 1.0 + 0.1 * rng.normal(size=(nx, ny))
    vs = 0.1 * rng.normal(size=(nx, ny))
    un, vn = R.project(us, vs, p, mask, dt, dx, dx, U)
    _expect_bcs(un, vn, U)
    tx, ty = 0.5 * (gx - gy), 0.5 * (gy - gx)
    sl = (slice(1, -1), slice(1, -1))
    scale = dt * np.abs(gx[sl]) + dt * np.abs(gy[sl])
    assert np.all(np.abs(un[sl] - us[sl] + dt * tx[sl]) <= 1e-7 * scale)
    assert np.all(np.abs(vn[sl] - vs[sl] + dt * ty[sl]) <= 1e-7 * scale)
    # the full gradient would differ by a normal component of order dt |g|
    assert np.max(np.abs(dt * (gx[sl] - tx[sl]))) > 1e-3


def test_project_hard_mask_is_applied_before_and_after_the_bcs():
    """p = 0: the correction is the identity, so the output is the masked
    u*, v* with the BCs, masked again (step.py:368-397).  A solid wall cell
    next to a fluid cell must stay 0 although the slip BC copies the fluid
    value into it; the same at the outlet copy and on the inlet row."""
    nx, ny, dt, dx, U = 8, 7, 0.01, 0.1, 1.0
    rng = np.random.default_rng(3)
    us = 1.0 + 0.1 * rng.normal(size=(nx, ny))
    vs = 0.1 * rng.normal(size=(nx, ny))
    mask = np.ones((nx, ny))
    solid = [(3, 3), (4, 3), (2, 0), (5, ny - 1), (nx - 1, 4), (0, 2)]
    for c in solid:
        mask[c] = 0.3
    un, vn = R.project(us, vs, np.zeros((nx, ny)), mask, dt, dx, dx, U)
    eu, ev = us.copy(), vs.copy()
    for c in solid:
        eu[c] = ev[c] = 0.0
    eu[0, :], ev[0, :] = U, 0.0
    eu[-1, :], ev[-1, :] = eu[-2, :], ev[-2, :]
    eu[:, 0], eu[:, -1] = eu[:, 1], eu[:, -2]
    ev[:, 0], ev[:, -1] = 0.0, 0.0
    for c in solid:
        eu[c] = ev[c] = 0.0
    assert np.array_equal(un, eu) and np.array_equal(vn, ev)
    assert un[2, 1] != 0 and un[2, 0] == 0  # slip copy into a solid wall cell undone


# ------------------------------------------------------------- composition

def test_step_pure_composition_with_uniform_brinkman_factor():
    """step_pure (step.py:174-218, 257-264, 305-349, 364-397) on a uniform
    state U0 with a uniform chi (constant sdf, mask = 1): the RK predictor
    is U0 / D inside, the inlet BC makes row 0 = U_inf, so the only
    divergence is at row 1: rhs = (U0/D - U_inf) / (2 dx) / dt on its
    interior cells and 0 elsewhere.  The pressure is the reference multigrid
    of that rhs, and with no interface (grad mask = 0) the corrected
    interior velocity is u* - dt grad p with the one-sided-at-boundary
    gradients."""
    nx, ny, dx, dy = 64, 32, 1.0 / 64, 1.0 / 32
    dt, nu, U0, U_inf, eta = 1e-3, 1e-3, 0.9, 1.0, 0.005
    s = 2e-3
    sdf = np.full((nx, ny), s)
    mask = np.ones((nx, ny))
    chi = 1.0 / (1.0 + math.exp(s / (0.1 * dx)))
    D = 1.0 + dt * chi / eta
    u = np.full((nx, ny), U0)
    v = np.zeros((nx, ny))
    un, vn, p, parts = R.step_pure(u, v, mask, sdf, dx, dy, nu, U_inf, dt, v_cycles=3,
                                   brinkman_eta=eta, return_parts=True)
    usr = parts["u_star"]
    assert np.allclose(usr[1:, :], U0 / D, rtol=1e-14) and np.all(usr[0] == U_inf)
    rhs = np.zeros((nx, ny))
    rhs[1, 1:-1] = (U0 / D - U_inf) / (2 * dx) / dt
    assert np.allclose(parts["rhs"], rhs, rtol=1e-12, atol=1e-9)
    p_ref = R.poisson_multigrid(rhs, mask, dx, dy, levels=4, v_cycles=3)
    assert np.allclose(p, p_ref, rtol=1e-12, atol=1e-14)
    gx = np.empty_like(p)
    gx[1:-1] = (p[2:] - p[:-2]) / (2 * dx)
    gy = np.empty_like(p)
    gy[:, 1:-1] = (p[:, 2:] - p[:, :-2]) / (2 * dy)
    sl = (slice(1, -1), slice(1, -1))
    assert np.allclose(un[sl], usr[sl] - dt * gx[sl], rtol=1e-12, atol=1e-14)
    assert np.allclose(vn[sl], -dt * gy[sl], rtol=1e-12, atol=1e-14)
    assert np.abs(p).max() > 0
