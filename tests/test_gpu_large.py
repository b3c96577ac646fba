"""Large cylinder workload (4096x1024 on 8x2, cylinder r = 0.1 at (1.5, 1),
Re 1e5, Smagorinsky LES applied with Cs = 0.17, dt 1e-4, textbook multigrid
in FP32 V-cycles inside FP64 defect correction to a relative residual).

The oracle composition of one step is constant_smagorinsky
(les_models.py:179-200) -> rk_step_unified(nu_sgs=...)
(rk3_simple_new.py:175-358, 304) -> predictor BCs (step.py:182-191) ->
rhs = div*mask/dt (step.py:194-199) -> the converged pressure -> projection
(step.py:257-397).  The converged pressure of the oracle is the FP64 defect
correction of oracle/textbook_mg.py run to a relative residual of 1e-13,
checked against scipy's sparse direct solve.

Bars.  FP64: every deterministic piece (nu_sgs, u*, v*, rhs, the
projection of the GPU pressure) within 1e-12; with the solve run to a
relative residual of 1e-12 the whole step (u, v) within 1e-12 and p within
1e-11 (the discrete Poisson problem's conditioning times FP64 epsilon); at
the workload's own 1e-8 the pressure differs from the converged one by what
that tolerance leaves (measured and printed, bounded by 1e-7).  FP32: u, v
within 1e-5 of the FP64 oracle run on the FP32-rounded initial fields; p
is stated: the FP32 rounding of u* (~6e-8 relative) enters
the rhs multiplied by 1/(dt dx), which puts the per-step pressure error at a
few 1e-5 whatever the solver does.
"""
import math

import numpy as np
import pytest
import torch

from oracle import reference_step as R
from oracle import textbook_mg as T

pytestmark = pytest.mark.gpu

CS, DT, ETA = 0.17, 1e-4, 0.005


def _case(nx, ny):
    from aerojax_amd import geometry as G
    grid = G.GridParams(nx, ny, 8.0, 2.0)
    geo = G.cylinder_geometry(grid, 1.5, 1.0, 0.1, eps=0.05)
    u, v = G.von_karman_ic(nx, ny, 1.0, 1e-3, 0)
    return dict(grid=grid, sdf=geo.sdf, mask=geo.mask, u=u.astype(np.float64),
                v=v.astype(np.float64), nu=G.viscosity(1.0, 0.2, 1e5))


def _oracle_pieces(c, inputs=np.float64):
    """FP64 oracle pieces of one step from the initial fields rounded to
    ``inputs`` (the GPU's storage type: the strain of a 1e-3 perturbation of
    a unit flow loses digits to the input rounding alone)."""
    g = c["grid"]
    f = np.float64
    r = lambda a: a.astype(inputs).astype(f)  # noqa: E731
    u, v = r(c["u"]), r(c["v"])
    mask, sdf = r(c["mask"]), r(c["sdf"])
    nsgs = R.constant_smagorinsky(u, v, g.dx, g.dy, math.sqrt(g.dx * g.dy), CS)
    us, vs = R.rk_step_unified(u, v, DT, c["nu"], g.dx, g.dy, mask, sdf=sdf, U_inf=1.0,
                               nu_sgs=nsgs, brinkman_eta=ETA)
    us, vs = R.predictor_bcs(us, vs, f(1.0))
    rhs = R.pressure_rhs(us, vs, mask, DT, g.dx, g.dy)
    return nsgs, us, vs, rhs


def _gpu_step(c, dtype, rel_tol):
    import aerojax_amd as aj
    g = c["grid"]
    stp = aj.CollocatedStepper(g.nx, g.ny, g.dx, g.dy, c["mask"], c["sdf"], c["nu"], 1.0, DT,
                               dtype=dtype, mg_variant="textbook", tolerance=rel_tol,
                               max_cycles=80, use_les=True, les_coupling="apply",
                               smagorinsky_constant=CS, brinkman_eta=ETA)
    stp.plan.set_option(relative=True)
    stp.load(c["u"], c["v"])
    stp.step()
    un, vn = stp.state
    st = stp.plan.stats()
    h = lambda t: t.detach().double().cpu().numpy()  # noqa: E731
    out = dict(nsgs=h(stp.nsgs), us=h(stp.us), vs=h(stp.vs), rhs=h(stp.rhs), p=h(stp.p),
               u=h(un), v=h(vn), stats=st)
    stp.plan.set_option(relative=False)  # the plan may be shared with other callers
    stp.close()
    return out


def _converged_pressure(rhs, g):
    mg = T.TextbookMG(g.nx, g.ny, g.dx, g.dy)
    b = rhs.copy()
    b[-1] = 0.0
    p, st = mg.solve(rhs, tol=1e-13 * mg.norm(b), max_outer=40)
    return p


@pytest.mark.parametrize("nx,ny", [(512, 128), (1024, 256)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_large_cylinder_step_matches_oracle_composition(gpu, nx, ny, dtype):
    c = _case(nx, ny)
    g = c["grid"]
    fp64 = dtype == torch.float64
    nsgs, us, vs, rhs = _oracle_pieces(c, np.float64 if fp64 else np.float32)
    pe = _converged_pressure(rhs, g)
    if nx * ny <= 300000:
        assert R.rel_l2(pe, T.direct_solve(rhs, g.dx, g.dy)) < 1e-11
    mask = c["mask"] if fp64 else c["mask"].astype(np.float32).astype(np.float64)
    ue, ve = R.project(us, vs, pe, mask, DT, g.dx, g.dy, 1.0)
    for tol in ((1e-8, 1e-12) if fp64 else (1e-8,)):
        got = _gpu_step(c, dtype, tol)
        st = got["stats"]
        print(f"{nx}x{ny} {dtype} rel tol {tol:.0e}: outer {st['outer']} cycles {st['cycles']} "
              f"norms {['%.2e' % x for x in st['norms']]}", flush=True)
        assert st["converged"] and not st["nonfinite"]
        assert st["norm"] <= tol * st["norms"][0] * (1 + 1e-12)
        err = {k: R.rel_l2(got[k], ref) for k, ref in
               (("nsgs", nsgs), ("us", us), ("vs", vs), ("rhs", rhs), ("p", pe), ("u", ue),
                ("v", ve))}
        print("  rel l2 vs FP64 oracle: " + ", ".join(f"{k} {e:.2e}" for k, e in err.items()),
              flush=True)
        # projection of the GPU's own pressure, against the oracle's projection
        pu, pv = R.project(got["us"], got["vs"], got["p"], mask, DT, g.dx, g.dy, 1.0)
        ep = max(R.rel_l2(got["u"], pu), R.rel_l2(got["v"], pv))
        if fp64:
            for k in ("nsgs", "us", "vs", "rhs"):
                assert err[k] <= 1e-12, (k, err[k])
            assert ep <= 1e-12, ep
            if tol <= 1e-12:
                assert err["u"] <= 1e-12 and err["v"] <= 1e-12, err
                assert err["p"] <= 1e-11, err
            else:
                assert err["p"] <= 1e-7, err
        else:
            for k in ("nsgs", "us", "vs", "u", "v"):
                assert err[k] <= 1e-5, (k, err[k])
            assert ep <= 1e-6, ep
            assert err["rhs"] <= 1e-4 and err["p"] <= 1e-4, err


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_large_cylinder_full_size_step(gpu, dtype):
    """The workload at its full 4096x1024 size, one step from the seeded
    initial state: the deterministic pieces against the oracle at full size,
    the pressure residual of the GPU's own rhs re-evaluated in FP64 on the
    host with the oracle's operator, and the projection of the GPU pressure."""
    nx, ny = 4096, 1024
    c = _case(nx, ny)
    g = c["grid"]
    got = _gpu_step(c, dtype, 1e-8)
    st = got["stats"]
    print(f"4096x1024 {dtype}: outer {st['outer']} cycles {st['cycles']} "
          f"norms {['%.2e' % x for x in st['norms']]}", flush=True)
    assert st["converged"] and not st["nonfinite"]
    fp64 = dtype == torch.float64
    nsgs, us, vs, rhs = _oracle_pieces(c, np.float64 if fp64 else np.float32)
    mask = c["mask"] if fp64 else c["mask"].astype(np.float32).astype(np.float64)
    err = {k: R.rel_l2(got[k], ref) for k, ref in
           (("nsgs", nsgs), ("us", us), ("vs", vs), ("rhs", rhs))}
    print("  rel l2 vs FP64 oracle: " + ", ".join(f"{k} {e:.2e}" for k, e in err.items()),
          flush=True)
    bar = 1e-12 if dtype == torch.float64 else 1e-5
    for k in ("nsgs", "us", "vs"):
        assert err[k] <= bar, (k, err[k])
    assert err["rhs"] <= (1e-12 if dtype == torch.float64 else 1e-4), err
    mg = T.TextbookMG(nx, ny, g.dx, g.dy)
    b = got["rhs"].copy()
    b[-1] = 0.0
    r = mg.norm(mg.residual64(got["p"], b)) / mg.norm(b)
    print(f"  host FP64 residual of the GPU pressure: {r:.2e} of ||b||", flush=True)
    if dtype == torch.float64:
        assert r <= 1e-8 * (1 + 1e-6)
    pu, pv = R.project(got["us"], got["vs"], got["p"], mask, DT, g.dx, g.dy, 1.0)
    ep = max(R.rel_l2(got["u"], pu), R.rel_l2(got["v"], pv))
    print(f"  projection of the GPU pressure: {ep:.2e}", flush=True)
    assert ep <= (1e-12 if dtype == torch.float64 else 1e-6)
