"""The C++/OpenMP restatement of the to-tolerance step (the multi-threaded
CPU baseline) against the NumPy oracle and scipy's sparse direct solve."""
import os

import numpy as np
import pytest

from oracle import reference_step as R
from oracle import textbook_mg as TM


@pytest.fixture(scope="module")
def C():
    from oracle import build, cpu_step
    if not os.path.exists(build.LIB):
        build.build()
    return cpu_step


@pytest.mark.parametrize("nx,ny,seed", [(256, 64, 0), (512, 128, 5)])
def test_cpp_step_matches_numpy_oracle(C, nx, ny, seed):
    from aerojax_amd import workloads
    c = workloads.warm_start_case(seed, nx, ny)
    g = c["grid"]
    f = np.float32
    u, v, mask, sdf = (c[k].astype(f) for k in ("u", "v", "mask", "sdf"))
    us, vs = R.rk_step_unified(u, v, c["dt"], c["nu"], g.dx, g.dy, mask, sdf=sdf, U_inf=1.0,
                               brinkman_eta=0.005)
    us, vs = R.predictor_bcs(us, vs, f(1.0))
    rhs = R.pressure_rhs(us, vs, mask, c["dt"], g.dx, g.dy)
    mg = TM.TextbookMG(nx, ny, g.dx, g.dy)
    p64, st = mg.solve(rhs, tol=1e-6)
    un, vn = R.project(us, vs, p64.astype(f), mask, c["dt"], g.dx, g.dy, 1.0)
    cs = C.CpuStepper(nx, ny, g.dx, g.dy)
    assert cs.levels == len(mg.shapes)
    u2, v2, p2, cyc, nrm, (us2, vs2, rhs2) = cs.step(u, v, mask, sdf, c["nu"], c["dt"],
                                                     parts=True)
    assert cyc == st["cycles"] and nrm <= 1e-6
    for a, b in ((us2, us), (vs2, vs), (rhs2, rhs)):
        assert R.rel_l2(a.astype(np.float64), b.astype(np.float64)) <= 1e-6
    assert R.rel_l2(p2.astype(np.float64), p64) <= 1e-6
    assert R.rel_l2(u2.astype(np.float64), un.astype(np.float64)) <= 1e-6
    assert R.rel_l2(v2.astype(np.float64), vn.astype(np.float64)) <= 1e-6
    cs.close()


@pytest.mark.parametrize("nx,ny", [(128, 64), (512, 256)])
def test_cpp_solve_matches_direct_solve(C, nx, ny):
    dx, dy = 4.0 / nx, 1.0 / ny
    b = np.random.default_rng(9).normal(size=(nx, ny)) * 200
    cs = C.CpuStepper(nx, ny, dx, dy)
    p, cyc, nrm = cs.solve(b, tol=1e-7)
    assert cyc > 0 and nrm <= 1e-7
    pd = TM.direct_solve(b.astype(np.float32).astype(np.float64), dx, dy)
    # the output is rounded to FP32
    assert R.rel_l2(p.astype(np.float64), pd) < 1e-6
    cs.close()
