"""Pins the textbook multigrid oracle: operator vs an independently built
sparse matrix, the exact coarse solve vs a dense solve of the level
operator, convergence to the FP64 sparse direct solve, a manufactured
discrete solution, and the grid-independent convergence factor."""
import numpy as np
import pytest

from oracle import textbook_mg as T


@pytest.mark.parametrize("nx,ny", [(64, 16), (96, 48)])
def test_operator_matches_sparse(nx, ny):
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(0)
    p = rng.normal(size=(nx, ny))
    p[-1] = 0
    A = T.consistent_operator_sparse(nx, ny, dx, dy)
    ref = (A @ p[:-1].reshape(-1)).reshape(nx - 1, ny)
    got = T.apply_A(p, 0, 1 / dx ** 2, 1 / dy ** 2)[:-1]
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-9)


@pytest.mark.parametrize("level,shape", [(0, (8, 6)), (0, (12, 5)), (1, (8, 4)), (3, (6, 8))])
def test_coarse_solve_matches_dense(level, shape):
    """DCT + per-mode Thomas = dense solve of the level operator (with the
    level's outlet ghost) on its unknown rows."""
    ax, ay = 1.0 / 0.01 ** 2, 1.0 / 0.02 ** 2
    n, m = shape
    rows = n - 1 if level == 0 else n
    rng = np.random.default_rng(level + n)
    b = rng.normal(size=shape)
    tab = T.coarse_tables(level, shape, ax, ay)
    C = tab[0]
    assert np.allclose(C @ C.T, np.eye(m), atol=1e-14)
    x = T.coarse_solve(b, level, shape, tab)
    A = T.dense_operator(level, shape, ax, ay)
    xd = np.linalg.solve(A, b[:rows].reshape(-1)).reshape(rows, m)
    assert np.linalg.norm(x[:rows] - xd) / np.linalg.norm(xd) < 1e-12
    if level == 0:
        assert np.all(x[-1] == 0)


@pytest.mark.parametrize("levels", [0, 3])
def test_manufactured_solution_direct_and_mg(levels):
    nx, ny = 128, 32
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(11)
    ps = rng.normal(size=(nx, ny))
    ps[-1] = 0
    b = T.apply_A(ps, 0, 1 / dx ** 2, 1 / dy ** 2)
    pd = T.direct_solve(b, dx, dy)
    assert np.allclose(pd, ps, rtol=0, atol=1e-8)
    mg = T.TextbookMG(nx, ny, dx, dy, levels=levels)
    assert len(mg.shapes) == (1 if levels == 0 else levels)
    p, st = mg.solve(b, tol=1e-9)
    assert st["norms"][-1] <= 1e-9
    assert np.linalg.norm(p - ps) / np.linalg.norm(ps) < 1e-9


@pytest.mark.parametrize("nx,ny,levels", [(256, 64, 4), (512, 128, 4), (512, 128, 6),
                                          (1024, 256, 0)])
def test_convergence_factor_grid_independent(nx, ny, levels):
    dx, dy = 4.0 / nx, 1.0 / ny
    mg = T.TextbookMG(nx, ny, dx, dy, dtype=np.float64, levels=levels)
    rng = np.random.default_rng(0)
    b = rng.normal(size=(nx, ny)) * 500
    b[-1] = 0
    p = np.zeros_like(b)
    norms = []
    for _ in range(8):
        p = mg.vcycle(b, 0, p)
        norms.append(mg.norm(mg.residual64(p, b)))
    rho = norms[-1] / norms[-2]
    assert rho < 0.1, (nx, ny, levels, rho)


def test_defect_correction_reaches_1e6_from_fp32_cycles():
    nx, ny = 256, 64
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(2)
    b = rng.normal(size=(nx, ny)) * 500
    mg = T.TextbookMG(nx, ny, dx, dy, levels=3)  # FP32 V-cycles
    p, st = mg.solve(b, tol=1e-6)
    assert st["norms"][-1] <= 1e-6
    assert st["norms"][0] > 1e2  # ~9 orders below the initial residual: FP32 alone cannot
    pd = T.direct_solve(b, dx, dy)
    bb = b.copy()
    bb[-1] = 0
    r = T.apply_A(pd, 0, 1 / dx ** 2, 1 / dy ** 2) - bb
    assert mg.norm(r) < 1e-8
    assert np.linalg.norm(p - pd) / np.linalg.norm(pd) < 1e-8


def test_hierarchy_and_gamma():
    assert T.hierarchy(2048, 512) == [(2048, 512), (1024, 256), (512, 128)]
    assert T.hierarchy(4096, 1024)[-1] == (512, 128)
    assert T.hierarchy(512, 96) == [(512, 96)]
    assert T.hierarchy(512, 96, 4)[-1] == (64, 12)
    assert T.hierarchy(64, 16, 3) == [(64, 16), (32, 8), (16, 4)]
    assert T.hierarchy(96, 33) == [(96, 33)]  # odd side: no coarsening
    assert T.hierarchy(2048, 32) == [(2048, 32), (1024, 16)]  # rows <= 1024
    with pytest.raises(ValueError, match="not suitable for multigrid"):
        T.hierarchy(96, 33, 2)
    assert np.isclose(T.outlet_gamma(1), -3.0)
    assert -1.0 > T.outlet_gamma(8) > -1.01
