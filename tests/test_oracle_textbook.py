"""Pins the textbook multigrid oracle: operator vs an independently built
sparse matrix, convergence to the FP64 sparse direct solve, a manufactured
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


def test_manufactured_solution_direct_and_mg():
    nx, ny = 128, 32
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(11)
    ps = rng.normal(size=(nx, ny))
    ps[-1] = 0
    b = T.apply_A(ps, 0, 1 / dx ** 2, 1 / dy ** 2)
    pd = T.direct_solve(b, dx, dy)
    assert np.allclose(pd, ps, rtol=0, atol=1e-8)
    mg = T.TextbookMG(nx, ny, dx, dy)
    p, st = mg.solve(b, tol=1e-9)
    assert st["norms"][-1] <= 1e-9
    assert np.linalg.norm(p - ps) / np.linalg.norm(ps) < 1e-9


def test_convergence_factor_grid_independent():
    for (nx, ny) in [(256, 64), (512, 128)]:
        dx, dy = 4.0 / nx, 1.0 / ny
        mg = T.TextbookMG(nx, ny, dx, dy, dtype=np.float64)
        rng = np.random.default_rng(0)
        b = rng.normal(size=(nx, ny)) * 500
        b[-1] = 0
        p = np.zeros_like(b)
        norms = []
        for _ in range(8):
            p = mg.vcycle(b, 0, p)
            norms.append(mg.norm(mg.residual64(p, b)))
        rho = norms[-1] / norms[-2]
        assert rho < 0.1, (nx, ny, rho)


def test_defect_correction_reaches_1e6_from_fp32_cycles():
    nx, ny = 256, 64
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(2)
    b = rng.normal(size=(nx, ny)) * 500
    mg = T.TextbookMG(nx, ny, dx, dy)  # FP32 V-cycles
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
    assert T.hierarchy(2048, 512)[-1] == (16, 4)
    assert T.hierarchy(512, 96)[-1] == (16, 3)
    assert T.hierarchy(64, 16) == [(64, 16), (32, 8), (16, 4)]
    assert T.hierarchy(8, 8) == [(8, 8)]
    assert T.hierarchy(96, 33) == [(96, 33)]  # odd side: no coarsening
    assert np.isclose(T.outlet_gamma(1), -3.0)
    assert -1.0 > T.outlet_gamma(8) > -1.01
