"""The HIP path against the committed golden fixtures (FP64 oracle outputs
stored in tests/golden/, inputs included): two reference-multigrid steps,
the to-tolerance textbook solve against the exact solution of its operator,
and constant Smagorinsky, all in FP64 to the FP64 bar."""
import os

import numpy as np
import pytest
import torch

from oracle import reference_step as R

from test_gpu_parity import dev, host

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name))


def test_step_pure_golden(gpu):
    import aerojax_amd as aj
    d = load("step_pure_48x32.npz")
    f = torch.float64
    st = aj.initial_state(dev(d["u0"], f), dev(d["v0"], f), float(d["dt"]))
    mask, sdf = dev(d["mask"], f), dev(d["sdf"], f)
    for k in (1, 2):
        st = aj.step_pure(st, mask, sdf, float(d["dx"]), float(d["dy"]), float(d["nu"]),
                          float(d["U_inf"]), v_cycles=3)
        for name, got in (("u", st.u), ("v", st.v), ("p", st.p)):
            e = R.rel_l2(host(got), d[f"{name}{k}"])
            print(f"golden step {k} {name}: rel_l2 {e:.2e}", flush=True)
            assert e <= 1e-12, (k, name, e)


def test_textbook_solve_golden(gpu):
    import aerojax_amd as aj
    d = load("textbook_direct_48x24.npz")
    nx, ny = d["rhs"].shape
    pl = aj.MultigridPlan(nx, ny, float(d["dx"]), float(d["dy"]), torch.float64, "textbook", 0)
    try:
        out = torch.empty((nx, ny), dtype=torch.float64, device="cuda")
        pl.solve(dev(d["rhs"], torch.float64), out, None, 60, 1e-11)
        torch.cuda.synchronize()
        s = pl.stats()
        assert s["converged"], s
    finally:
        pl.close()
    e = R.rel_l2(host(out), d["p"])
    print(f"golden textbook solve: rel_l2 vs exact {e:.2e}", flush=True)
    assert e <= 1e-10, e


def test_smagorinsky_golden(gpu):
    import aerojax_amd as aj
    d = load("smagorinsky_48x24.npz")
    f = torch.float64
    got = aj.constant_smagorinsky(dev(d["u"], f), dev(d["v"], f), float(d["dx"]), float(d["dy"]),
                                  float(d["delta"]), float(d["cs"]))
    e = R.rel_l2(host(got), d["nu_sgs"])
    print(f"golden smagorinsky: rel_l2 {e:.2e}", flush=True)
    assert e <= 1e-12, e
