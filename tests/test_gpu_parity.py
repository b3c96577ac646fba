"""GPU parity of every HIP entry point against the CPU oracle.

Bar: relative L2 <= 1e-5 for FP32 kernels and <= 1e-12 for FP64 kernels,
against the FP64 oracle (FP32 results are also compared with the FP32
oracle); obstacle masks are shared inputs, and the hard-mask zero pattern
must match exactly."""
import math

import numpy as np
import pytest
import torch

from oracle import reference_step as R
from oracle import textbook_mg as T

pytestmark = pytest.mark.gpu

TOL = {torch.float32: 1e-5, torch.float64: 1e-12}
NP = {torch.float32: np.float32, torch.float64: np.float64}


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def host(t):
    return t.detach().cpu().numpy().astype(np.float64)


def case(nx, ny, lx, ly, kind="naca", seed=0, noise=1e-3):
    dx, dy, X, Y = R.grid_coords(nx, ny, lx, ly)
    if kind == "naca":
        m, p, t = R.parse_naca_4digit("0012")
        sdf = R.naca_sdf(X, Y, 0.5 * lx / 4, 10.0, 0.25 * lx, 0.5 * ly, m, p, t)
    else:
        sdf = R.cylinder_sdf(X, Y, 0.3 * lx, 0.5 * ly, 0.1 * ly)
    sdf = np.nan_to_num(sdf, nan=0.0)
    mask = R.mask_from_sdf(sdf, 0.05 * ly)
    u, v = R.von_karman_ic(nx, ny, 1.0, noise, seed=seed, dtype=np.float64)
    # a wake-like perturbation so that advection and diffusion are exercised
    rng = np.random.default_rng(seed + 1)
    u = u + 0.2 * np.sin(3 * X) * np.cos(2 * Y) + 0.01 * rng.normal(size=u.shape)
    v = v + 0.1 * np.cos(5 * X) * np.sin(4 * Y)
    return dict(dx=dx, dy=dy, X=X, Y=Y, sdf=sdf, mask=mask, u=u, v=v)


SHAPES = [(512, 96, 4.0, 0.75), (130, 70, 2.0, 1.0), (2048, 512, 4.0, 1.0)]
DTYPES = [torch.float32, torch.float64]


def check(got, ref64, ref_dt, dtype, what):
    tol = TOL[dtype]
    e64 = R.rel_l2(got, ref64)
    print(f"{what}: rel_l2 vs fp64 oracle = {e64:.3e}", flush=True)
    assert e64 <= tol, (what, e64)
    if ref_dt is not None:
        assert R.rel_l2(got, ref_dt) <= tol, what


# ------------------------------------------------------------------ RK

@pytest.mark.parametrize("shape", SHAPES[:2])
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("variant", ["sdf", "nosdf", "fast", "nsgs", "hyper", "noslip"])
def test_rk_step_parity(gpu, shape, dtype, variant):
    import aerojax_amd as aj
    c = case(*shape)
    dt, nu, eta = 1e-3, 5e-4, 0.005
    kw = dict(U_inf=1.0, brinkman_eta=eta)
    sdf = None if variant == "nosdf" else c["sdf"]
    nsgs = None
    if variant == "nsgs":
        nsgs = R.constant_smagorinsky(c["u"], c["v"], c["dx"], c["dy"], math.sqrt(c["dx"] * c["dy"]), 0.17)
        nsgs = nsgs * 50.0  # make the term matter
    if variant == "fast":
        kw["fast_mode"] = True
    if variant == "hyper":
        kw["nu_hyper_ratio"] = 0.5
    if variant == "noslip":
        kw["slip_walls"] = False
    ref = R.rk_step_unified(c["u"], c["v"], dt, nu, c["dx"], c["dy"], c["mask"], sdf=sdf,
                            nu_sgs=nsgs, **kw)
    f = NP[dtype]
    refd = R.rk_step_unified(c["u"].astype(f), c["v"].astype(f), dt, nu, c["dx"], c["dy"],
                             c["mask"].astype(f), sdf=None if sdf is None else sdf.astype(f),
                             nu_sgs=None if nsgs is None else nsgs.astype(f), **kw)
    out = aj.rk_step_unified(dev(c["u"], dtype), dev(c["v"], dtype), dt, nu, c["dx"], c["dy"],
                             dev(c["mask"], dtype), sdf=None if sdf is None else dev(sdf, dtype),
                             nu_sgs=None if nsgs is None else dev(nsgs, dtype), **kw)
    torch.cuda.synchronize()
    for k, name in enumerate(("u*", "v*")):
        g = host(out[k])
        check(g, ref[k], refd[k], dtype, f"rk {variant} {shape[:2]} {dtype} {name}")
        solid = c["mask"] <= 0.5
        assert np.all(g[1:-1, 1:-1][solid[1:-1, 1:-1]] == 0.0)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("fast", [False, True])
def test_rk_viscosity_field_parity(gpu, dtype, fast):
    """``nu`` as a field with nu_sgs and hyper-viscosity (the GUI step's
    call, step_handlers.py:76-90)."""
    import aerojax_amd as aj
    c = case(256, 64, 4.0, 1.0, kind="cylinder")
    dx, dy, dt = c["dx"], c["dy"], 1e-3
    nsgs = 50.0 * R.constant_smagorinsky(c["u"], c["v"], dx, dy, math.sqrt(dx * dy), 0.17)
    nuf = 5e-4 + nsgs
    kw = dict(U_inf=1.0, brinkman_eta=1e-4, nu_hyper_ratio=0.3, fast_mode=fast)
    ref = R.rk_step_unified(c["u"], c["v"], dt, nuf, dx, dy, c["mask"], sdf=c["sdf"], nu_sgs=nsgs,
                            **kw)
    out = aj.rk_step_unified(dev(c["u"], dtype), dev(c["v"], dtype), dt, dev(nuf, dtype), dx, dy,
                             dev(c["mask"], dtype), sdf=dev(c["sdf"], dtype),
                             nu_sgs=dev(nsgs, dtype), **kw)
    for k, name in enumerate(("u*", "v*")):
        check(host(out[k]), ref[k], None, dtype, f"rk nu-field fast={fast} {dtype} {name}")


@pytest.mark.parametrize("shape", [(512, 96), (130, 70), (2048, 512), (6, 5)])
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("variant", ["sdf", "fast", "nsgs", "nosdf", "unmasked", "noslip"])
def test_rk_step_rhs_equals_separate_launches(gpu, shape, dtype, variant):
    """aj_rk_step_rhs (the RK predictor with the pressure rhs in its
    epilogue) gives the bits of aj_rk_step followed by aj_divergence_rhs."""
    import ctypes
    from aerojax_amd import _lib
    L = _lib.load()
    nx, ny = shape
    c = case(nx, ny, 4.0, 1.0, kind="cylinder" if nx >= 64 else "naca")
    dc = _lib.AJ_F32 if dtype == torch.float32 else _lib.AJ_F64
    u, v, mask, sdf = (dev(c[k], dtype) for k in ("u", "v", "mask", "sdf"))
    if variant == "nosdf":
        sdf = None
    nsgs = None
    if variant == "nsgs":
        nsgs = dev(50.0 * R.constant_smagorinsky(c["u"], c["v"], c["dx"], c["dy"],
                                                 math.sqrt(c["dx"] * c["dy"]), 0.17), dtype)
    dt = 1e-3
    prm = _lib.RkParams(dt, 5e-4, c["dx"], c["dy"], 1.0, 0.005, 0.0,
                        0 if variant == "noslip" else 1, 1 if variant == "fast" else 0)
    masked = 0 if variant == "unmasked" else 1
    s = torch.cuda.current_stream().cuda_stream
    p = lambda t: None if t is None else t.data_ptr()  # noqa: E731
    e = lambda: torch.empty_like(u)  # noqa: E731
    us1, vs1, r1, us2, vs2, r2 = e(), e(), e(), e(), e(), e()
    _lib.check(L.aj_rk_step(dc, p(u), p(v), p(mask), p(sdf), p(nsgs), p(us1), p(vs1), nx, ny,
                            ctypes.byref(prm), s), "aj_rk_step")
    _lib.check(L.aj_divergence_rhs(dc, p(us1), p(vs1), p(mask) if masked else None, p(r1), nx, ny,
                                   c["dx"], c["dy"], dt, s), "aj_divergence_rhs")
    _lib.check(L.aj_rk_step_rhs(dc, p(u), p(v), p(mask), p(sdf), p(nsgs), p(us2), p(vs2), p(r2),
                                masked, nx, ny, ctypes.byref(prm), s), "aj_rk_step_rhs")
    torch.cuda.synchronize()
    assert torch.equal(us1, us2) and torch.equal(vs1, vs2)
    assert torch.equal(r1, r2), float((r1 - r2).abs().max())
    assert float(r1.abs().max()) > 0


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("Re,les,fast", [(800.0, False, False), (5e4, True, False),
                                         (5e3, True, True)])
def test_step_collocated_parity(gpu, dtype, Re, les, fast):
    """The GUI step (_step_collocated) with the Re-dependent parameters."""
    import aerojax_amd as aj
    nx, ny = 256, 64
    c = case(nx, ny, 4.0, 1.0, kind="cylinder")
    dx, dy, dt = c["dx"], c["dy"], 2e-4
    prm = aj.get_re_parameters(Re, nx)
    nu = 1.0 * 0.2 / Re
    kw = dict(use_les=les, smagorinsky_constant=prm["C_s"], nu_hyper_ratio=prm["nu_hyper_ratio"],
              fast_mode=fast, brinkman_eta=prm["brinkman_eta"], v_cycles=5)
    ru, rv, rp = R.step_collocated(c["u"], c["v"], c["mask"], c["sdf"], dx, dy, nu, 1.0, dt,
                                   **kw)
    u, v, p = aj.step_collocated(dev(c["u"], dtype), dev(c["v"], dtype), dev(c["mask"], dtype),
                                 dt, sdf=dev(c["sdf"], dtype), dx=dx, dy=dy, nu=nu, U_inf=1.0,
                                 **kw)
    torch.cuda.synchronize()
    tag = f"step_collocated Re={Re:g} les={les} fast={fast} {dtype}"
    check(host(p), rp, None, dtype, tag + " p")
    check(host(u), ru, None, dtype, tag + " u")
    check(host(v), rv, None, dtype, tag + " v")
    solid = c["mask"] <= 0.5
    assert np.all(host(u)[solid] == 0.0) and np.all(host(v)[solid] == 0.0)


def test_rk_full_size_uniform_flow_fixed_point(gpu):
    import aerojax_amd as aj
    nx, ny = 2048, 512
    u = torch.full((nx, ny), 1.0, device="cuda")
    v = torch.zeros((nx, ny), device="cuda")
    m = torch.ones((nx, ny), device="cuda")
    us, vs = aj.rk_step_unified(u, v, 1e-3, 1e-3, 1 / 512, 1 / 512, m, U_inf=1.0)
    assert torch.equal(us, u) and torch.equal(vs, v)


# ------------------------------------------------------------------ operators

@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dtype", DTYPES)
def test_les_div_project_parity(gpu, shape, dtype):
    import aerojax_amd as aj
    c = case(*shape)
    dx, dy, dt = c["dx"], c["dy"], 1e-3
    f = NP[dtype]
    delta = math.sqrt(dx * dy)
    ref = R.constant_smagorinsky(c["u"], c["v"], dx, dy, delta, 0.17)
    got = aj.constant_smagorinsky(dev(c["u"], dtype), dev(c["v"], dtype), dx, dy, delta, 0.17)
    check(host(got), ref, None, dtype, f"les {shape[:2]}")
    ref = R.pressure_rhs(c["u"], c["v"], c["mask"], dt, dx, dy)
    got = aj.pressure_rhs(dev(c["u"], dtype), dev(c["v"], dtype), dev(c["mask"], dtype), dt, dx, dy)
    check(host(got), ref, None, dtype, f"rhs {shape[:2]}")
    got = aj.divergence_nonperiodic(dev(c["u"], dtype), dev(c["v"], dtype), dx, dy)
    check(host(got), R.divergence_nonperiodic(c["u"], c["v"], dx, dy), None, dtype, "div")
    p = 5.0 * np.sin(2 * c["X"]) * np.cos(3 * c["Y"])
    ref = R.project(c["u"], c["v"], p, c["mask"], dt, dx, dy, 1.0)
    refd = R.project(c["u"].astype(f), c["v"].astype(f), p.astype(f), c["mask"].astype(f), dt, dx,
                     dy, 1.0)
    got = aj.project(dev(c["u"], dtype), dev(c["v"], dtype), dev(p, dtype), dev(c["mask"], dtype),
                     dt, dx, dy, 1.0)
    for k in range(2):
        check(host(got[k]), ref[k], refd[k], dtype, f"project {shape[:2]} {k}")


@pytest.mark.parametrize("shape", [(130, 70, 2.0, 1.0), (96, 8, 2.0, 0.5), (2048, 512, 4.0, 1.0)])
@pytest.mark.parametrize("dtype", DTYPES)
def test_project_pairs_equal_cells(gpu, shape, dtype):
    """An even row length goes by pairs of cells; arrays that start one
    element past a pair boundary take the per-cell kernel.  Same bits, walls
    and outlet row included."""
    import aerojax_amd as aj
    c = case(*shape)
    dx, dy, dt = c["dx"], c["dy"], 1e-3
    p = 5.0 * np.sin(2 * c["X"]) * np.cos(3 * c["Y"])
    args = [c["u"], c["v"], p, c["mask"]]

    def shifted(a):
        t = dev(a, dtype)
        buf = torch.zeros(t.numel() + 1, dtype=t.dtype, device=t.device)
        buf[1:].copy_(t.reshape(-1))
        return buf[1:].view(t.shape)

    pairs = aj.project(*[dev(a, dtype) for a in args], dt, dx, dy, 1.0)
    cells = aj.project(*[shifted(a) for a in args], dt, dx, dy, 1.0)
    ref = R.project(*args, dt, dx, dy, 1.0)
    refd = R.project(*[a.astype(NP[dtype]) for a in args], dt, dx, dy, 1.0)
    for k in range(2):
        assert torch.equal(pairs[k], cells[k]), (shape[:2], k)
        check(host(pairs[k]), ref[k], refd[k], dtype, f"project pairs {shape[:2]} {k}")


# ------------------------------------------------------------------ multigrid

@pytest.mark.parametrize("shape", [(512, 96), (2048, 512), (64, 32)])
@pytest.mark.parametrize("dtype", DTYPES)
def test_reference_multigrid_parity(gpu, shape, dtype):
    import aerojax_amd as aj
    nx, ny = shape
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(3)
    x = np.linspace(0, 1, nx)[:, None]
    y = np.linspace(0, 1, ny)[None, :]
    b = 300 * np.sin(7 * x) * np.cos(5 * y) + 30 * rng.normal(size=(nx, ny))
    ref, st_ref = R.poisson_multigrid(b, None, dx, dy, levels=4, v_cycles=3, return_stats=True)
    got, st = aj.poisson_multigrid(dev(b, dtype), None, dx, dy, levels=4, v_cycles=3,
                                   return_stats=True)
    check(host(got), ref, None, dtype, f"mg-ref {shape} {dtype}")
    assert st["cycles"] == st_ref["cycles"] == 3
    assert not st["nonfinite"]
    assert np.allclose(st["norms"][:3], st_ref["norm_sq"], rtol=1e-4)
    # warm start
    p0 = 0.5 * ref
    ref2 = R.poisson_multigrid(b, None, dx, dy, v_cycles=2, p0=p0)
    got2 = aj.poisson_multigrid(dev(b, dtype), None, dx, dy, v_cycles=2, p0=dev(p0, dtype))
    check(host(got2), ref2, None, dtype, "mg-ref warm")


def test_reference_multigrid_gate_and_errors(gpu):
    import aerojax_amd as aj
    b = torch.zeros((64, 32), device="cuda")
    p, st = aj.poisson_multigrid(b, None, 0.1, 0.1, v_cycles=5, return_stats=True)
    assert st["cycles"] == 1 and st["converged"] and torch.count_nonzero(p) == 0
    with pytest.raises(ValueError, match="not suitable for multigrid"):
        aj.poisson_multigrid(torch.zeros((7, 5), device="cuda"), None, 0.1, 0.1)
    with pytest.raises(ValueError):
        aj.poisson_multigrid(torch.zeros((32, 16), device="cuda"), None, 0.1, 0.1)
    # levels=2 hierarchy (coarsest 4x4 through the LDS coarse kernel)
    rng = np.random.default_rng(9)
    bb = rng.normal(size=(16, 16))
    ref = R.poisson_multigrid(bb, None, 0.1, 0.07, levels=2, v_cycles=2)
    got = aj.poisson_multigrid(dev(bb, torch.float64), None, 0.1, 0.07, levels=2, v_cycles=2)
    check(host(got), ref, None, torch.float64, "mg-ref levels=2")


def test_reference_multigrid_large_coarse_level(gpu):
    # 4096x1024 with levels=4 leaves a 256x64 coarsest grid (multi-launch path)
    import aerojax_amd as aj
    nx, ny = 4096, 1024
    dx = dy = 1.0 / 512
    rng = np.random.default_rng(4)
    b = rng.normal(size=(nx, ny))
    ref = R.poisson_multigrid(b, None, dx, dy, v_cycles=1, tolerance=0.0)
    got = aj.poisson_multigrid(dev(b, torch.float32), None, dx, dy, v_cycles=1, tolerance=0.0)
    check(host(got), ref, None, torch.float32, "mg-ref 4096x1024")


@pytest.mark.parametrize("shape,levels", [((64, 16), 0), ((64, 16), 3), ((512, 96), 0),
                                          ((512, 96), 4), ((2048, 512), 0), ((256, 64), 6),
                                          ((200, 74), 2), ((1000, 300), 0)])
def test_textbook_multigrid_one_outer_iteration_parity(gpu, shape, levels):
    import aerojax_amd as aj
    nx, ny = shape
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(5)
    b = rng.normal(size=(nx, ny)) * 500
    mg = T.TextbookMG(nx, ny, dx, dy, levels=levels)
    ref, st_ref = mg.solve(b, tol=0.0, max_outer=1)
    cpo = 1 if len(mg.shapes) == 1 else 2  # one exact solve, or two V-cycles
    got, st = aj.poisson_multigrid_to_tolerance(dev(b, torch.float64), dx, dy, tolerance=0.0,
                                                max_cycles=cpo, return_stats=True, levels=levels)
    assert st["cycles"] == cpo == st_ref["cycles"] and st["outer"] == 1
    check(host(got), ref, None, torch.float32, f"mg-textbook 1 outer {shape}")
    # initial residual: same FP64 arithmetic on the same b
    assert np.isclose(st["norms"][0], st_ref["norms"][0], rtol=1e-12)
    if len(mg.shapes) == 1:
        # the coarsest level is the fine grid: one exact solve whose FP32
        # output rounding alone sets the residual, differently on each side
        assert st["norms"][1] < 1e-3 * st["norms"][0]
        assert st_ref["norms"][1] < 1e-3 * st_ref["norms"][0]
    else:
        # after two FP32 V-cycles the residual is 1e-4..1e-3 of b: FP32
        # rounding of p (rel ~1e-7) is amplified by ||A|| ||p|| / ||r||, so
        # the two residuals agree only loosely (p itself is checked above)
        r1, r2 = st["norms"][1], st_ref["norms"][1]
        print(f"residual after one outer iteration: {r1:.4e} vs oracle {r2:.4e} "
              f"(of ||b|| {r1 / st['norms'][0]:.2e})", flush=True)
        assert np.isclose(r1, r2, rtol=0.25) and r1 < 2e-3 * st["norms"][0]


@pytest.mark.parametrize("shape", [(512, 96), (2048, 512), (1000, 300)])
def test_textbook_multigrid_to_tolerance(gpu, shape):
    import aerojax_amd as aj
    nx, ny = shape
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(6)
    b = rng.normal(size=(nx, ny)) * 500
    got, st = aj.poisson_multigrid_to_tolerance(dev(b, torch.float64), dx, dy, tolerance=1e-6,
                                                max_cycles=40, return_stats=True)
    print("textbook", shape, st["cycles"], ["%.2e" % x for x in st["norms"]], flush=True)
    assert st["converged"] and st["norm"] <= 1e-6 and not st["nonfinite"]
    p = host(got)
    mgo = T.TextbookMG(nx, ny, dx, dy)
    bb = b.copy()
    bb[-1] = 0
    assert mgo.norm(mgo.residual64(p, bb)) <= 1e-6
    if nx * ny <= 60000:
        pd = T.direct_solve(b, dx, dy)
        assert R.rel_l2(p, pd) < 1e-9


# ------------------------------------------------------------------ full step

@pytest.mark.parametrize("dtype", DTYPES)
def test_step_pure_parity_config_small(gpu, dtype):
    """512x96 on 4x0.75, NACA 0012 chord 0.5 at 10 deg, LE at (1, 0.375),
    mask eps 0.05, Re 1000 (nu = 5e-4), dt 1e-3, 3 reference V-cycles."""
    import aerojax_amd as aj
    from aerojax_amd import geometry as G
    nx, ny, lx, ly = 512, 96, 4.0, 0.75
    grid = G.GridParams(nx, ny, lx, ly)
    geo = G.naca_geometry(grid, "0012", chord=0.5, aoa_deg=10.0, x_le=1.0, y_le=0.375, eps=0.05)
    sdf = np.nan_to_num(geo.sdf, nan=0.0)
    mask = geo.mask
    u, v = G.von_karman_ic(nx, ny, 1.0, 1e-3, seed=0)
    ref = R.step_pure(u, v, mask, sdf, grid.dx, grid.dy, 5e-4, 1.0, 1e-3, v_cycles=3)
    f = NP[dtype]
    refd = R.step_pure(u.astype(f), v.astype(f), mask.astype(f), sdf.astype(f), grid.dx, grid.dy,
                       5e-4, 1.0, 1e-3, v_cycles=3)
    st = aj.initial_state(dev(u, dtype), dev(v, dtype), 1e-3)
    new = aj.step_pure(st, dev(mask, dtype), dev(sdf, dtype), grid.dx, grid.dy, 5e-4, 1.0,
                       v_cycles=3)
    for got, r64, rd, name in ((new.u, ref[0], refd[0], "u"), (new.v, ref[1], refd[1], "v"),
                               (new.p, ref[2], refd[2], "p")):
        check(host(got), r64, rd, dtype, f"step {name} {dtype}")
    assert new.iteration == 1 and new.u_prev is st.u


def test_step_trajectory_error_growth(gpu):
    """20 FP32 steps against the FP64 oracle: error is reported, and must
    stay small (round-off growth, not a gate on bitwise agreement)."""
    import aerojax_amd as aj
    nx, ny = 256, 64
    c = case(nx, ny, 4.0, 1.0, kind="cylinder", noise=1e-3)
    dx, dy = c["dx"], c["dy"]
    u64, v64 = c["u"], c["v"]
    stp = aj.CollocatedStepper(nx, ny, dx, dy, c["mask"], c["sdf"], 5e-4, 1.0, 1e-3, v_cycles=3)
    stp.load(u64, v64)
    errs = []
    for k in range(20):
        u64, v64, _ = R.step_pure(u64, v64, c["mask"], c["sdf"], dx, dy, 5e-4, 1.0, 1e-3,
                                  v_cycles=3)
        stp.step()
        ug, vg = stp.state
        errs.append(R.rel_l2(host(ug), u64))
    print("trajectory rel err:", ["%.1e" % e for e in errs], flush=True)
    assert errs[-1] < 1e-4


@pytest.mark.parametrize("variant,les", [("reference", False), ("textbook", True)])
def test_stepper_equals_public_functions(gpu, variant, les):
    """CollocatedStepper's direct library calls give the bits of the public
    functions composed by hand (LES -> RK -> rhs -> multigrid -> projection)."""
    import math
    import aerojax_amd as aj
    from aerojax_amd.operators import constant_smagorinsky, pressure_rhs, project
    nx, ny = 256, 64
    c = case(nx, ny, 4.0, 1.0, kind="cylinder", noise=1e-3)
    dx, dy, nu, dt = c["dx"], c["dy"], 5e-4, 1e-3
    f = torch.float32
    stp = aj.CollocatedStepper(nx, ny, dx, dy, c["mask"], c["sdf"], nu, 1.0, dt, mg_variant=variant,
                               v_cycles=3, tolerance=1e-6, max_cycles=40, use_les=les,
                               les_coupling="apply", smagorinsky_constant=0.17)
    stp.load(c["u"], c["v"])
    plan = aj.MultigridPlan(nx, ny, dx, dy, f, variant, 4 if variant == "reference" else 0)
    mask, sdf = dev(c["mask"], f), dev(c["sdf"], f)
    u, v = dev(c["u"], f), dev(c["v"], f)
    for _ in range(3):
        stp.step()
        nsgs = constant_smagorinsky(u, v, dx, dy, math.sqrt(dx * dy), 0.17) if les else None
        us, vs = aj.rk_step_unified(u, v, dt, nu, dx, dy, mask, sdf=sdf, U_inf=1.0, nu_sgs=nsgs,
                                    brinkman_eta=0.005)
        rhs = pressure_rhs(us, vs, mask, dt, dx, dy)
        p = torch.empty_like(rhs)
        plan.solve(rhs, p, None, 3 if variant == "reference" else 40, 1e-6)
        u, v = project(us, vs, p, mask, dt, dx, dy, 1.0)
        ug, vg = stp.state
        assert torch.equal(ug, u) and torch.equal(vg, v) and torch.equal(stp.p, p)
    plan.close()


def test_deferred_solve_tops_up_and_matches(gpu):
    """aj_poisson_mg_enqueue + aj_poisson_mg_complete: a solve enqueued with
    fewer iterations than it needs (the plan's previous solve needed one) is
    topped up on completion (more = 1) and gives the bits of the blocking
    solve; a stepper with the deferred convergence read equals the blocking
    stepper step for step."""
    import ctypes
    import aerojax_amd as aj
    from aerojax_amd import _lib
    L = _lib.load()
    nx, ny = 512, 128
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(12)
    b = dev(rng.normal(size=(nx, ny)) * 500, torch.float32)
    pl = aj.MultigridPlan(nx, ny, dx, dy, torch.float32, "textbook", 0)
    ref = aj.MultigridPlan(nx, ny, dx, dy, torch.float32, "textbook", 0)
    p = torch.empty_like(b)
    pl.solve(torch.zeros_like(b), p, None, 40, 1e-6)  # converged at once: one iteration next
    p1, p2 = torch.empty_like(b), torch.empty_like(b)
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(L.aj_poisson_mg_enqueue(pl.handle, b.data_ptr(), None, p1.data_ptr(), 40, 1e-6, s),
               "enqueue")
    with pytest.raises(_lib.AeroJaxError, match="deferred solve is open"):
        pl.solve(b, p, None, 40, 1e-6)
    more = ctypes.c_int(0)
    _lib.check(L.aj_poisson_mg_complete(pl.handle, ctypes.byref(more)), "complete")
    ref.solve(b, p2, None, 40, 1e-6)
    torch.cuda.synchronize()
    assert more.value == 1 and torch.equal(p1, p2)
    s1, s2 = pl.stats(), ref.stats()
    assert s1["converged"] and s1["cycles"] == s2["cycles"] and s1["norm"] == s2["norm"]
    pl.close()
    ref.close()
    c = case(nx, ny, 4.0, 1.0, kind="cylinder", noise=1e-3)
    mk = lambda d: aj.CollocatedStepper(nx, ny, c["dx"], c["dy"], c["mask"], c["sdf"], 5e-4, 1.0,  # noqa
                                        2.5e-4, mg_variant="textbook", tolerance=1e-6,
                                        max_cycles=40, deferred_check=d)
    a, d = mk(False), mk(True)
    a.load(c["u"], c["v"])
    d.load(c["u"], c["v"])
    for _ in range(4):
        a.step()
        d.step()
        ua, va = a.state
        ud, vd = d.state
        assert torch.equal(ua, ud) and torch.equal(va, vd) and torch.equal(a.p, d.p)


@pytest.mark.parametrize("les,network", [(False, False), (True, True)])
def test_deferred_stepper_speculative_predictor_matches(gpu, les, network):
    """Steps with the deferred convergence read and no read in between (each
    step queues its predictor before completing the previous solve) give the
    blocking stepper's bits.  The deferred stepper's plan first solves a zero
    rhs, which settles at once, so its next solve is enqueued with a single
    outer iteration: the following step must top it up, re-project and
    queue the predictor again (counted in ``topups``).  With LES applied and
    a network guess this is the composition of the warm-start arm."""
    import aerojax_amd as aj
    from aerojax_amd.nn import WarmStartNet
    nx, ny = 512, 128
    c = case(nx, ny, 4.0, 1.0, kind="naca", noise=1e-3)
    net = WarmStartNet(base_channels=32, levels=4, seed=7) if network else None
    mk = lambda d: aj.CollocatedStepper(nx, ny, c["dx"], c["dy"], c["mask"], c["sdf"], 5e-4, 1.0,  # noqa
                                        2.5e-4, mg_variant="textbook", tolerance=1e-6,
                                        max_cycles=40, use_les=les, les_coupling="apply",
                                        smagorinsky_constant=0.17, warm_start=net,
                                        deferred_check=d)
    a, d = mk(False), mk(True)
    z = torch.zeros((nx, ny), dtype=torch.float32, device=gpu)
    d.plan.solve(z, torch.empty_like(z), None, 40, 1e-6)  # next enqueue: one iteration
    a.load(c["u"], c["v"])
    d.load(c["u"], c["v"])
    for n in (3, 2):
        for _ in range(n):
            a.step()
            d.step()
        ua, va = a.state
        ud, vd = d.state
        assert torch.equal(ua, ud) and torch.equal(va, vd) and torch.equal(a.p, d.p)
    print(f"deferred stepper top-ups: {d.topups}", flush=True)
    assert d.topups >= 1
    d.close()
    if net is not None:
        net.close()


def test_deferred_stepper_with_accepted_guess_matches(gpu):
    """A warm start that passes the guess test (a loose solve of the step's
    own rhs, residual far below the zero guess's): the guess is kept at every
    step, and the deferred stepper still gives the blocking stepper's bits."""
    import aerojax_amd as aj
    nx, ny = 512, 128
    c = case(nx, ny, 4.0, 1.0, kind="cylinder", noise=1e-3)
    loose = aj.MultigridPlan(nx, ny, c["dx"], c["dy"], torch.float32, "textbook", 0)
    loose.set_option(relative=True)

    def guess(rhs, sdf):
        p0 = torch.empty_like(rhs)
        loose.solve(rhs, p0, None, 2, 1e-2)
        return p0

    mk = lambda d: aj.CollocatedStepper(nx, ny, c["dx"], c["dy"], c["mask"], c["sdf"], 5e-4, 1.0,  # noqa
                                        2.5e-4, mg_variant="textbook", tolerance=1e-6,
                                        max_cycles=40, warm_start=guess, deferred_check=d)
    a, d = mk(False), mk(True)
    a.load(c["u"], c["v"])
    d.load(c["u"], c["v"])
    for _ in range(4):
        a.step()
        d.step()
        d.finish()
        sa, sd = a.plan.stats(), d.plan.stats()
        print(f"guess kept {sa['used_guess']} {sd['used_guess']}, guess/zero residual "
              f"{sd['guess_norm'] / sd['zero_norm']:.3e}", flush=True)
        assert sa["used_guess"] and sd["used_guess"]
        assert sd["guess_norm"] < 0.1 * sd["zero_norm"]
        ua, va = a.state
        ud, vd = d.state
        assert torch.equal(ua, ud) and torch.equal(va, vd) and torch.equal(a.p, d.p)
    d.close()
    loose.close()


def test_closed_steppers_release_their_plans(gpu):
    """Deferred steppers own their multigrid plan; closing them gives the
    workspace back (device memory stays flat over repeated create/close)."""
    import gc
    import aerojax_amd as aj
    nx, ny = 512, 128
    c = case(nx, ny, 4.0, 1.0, kind="cylinder", noise=1e-3)

    def cycle():
        s = aj.CollocatedStepper(nx, ny, c["dx"], c["dy"], c["mask"], c["sdf"], 5e-4, 1.0,
                                 2.5e-4, mg_variant="textbook", tolerance=1e-6, max_cycles=40,
                                 deferred_check=True)
        s.load(c["u"], c["v"])
        s.step()
        s.step()
        ws = s.plan.workspace_bytes
        s.close()
        del s
        gc.collect()
        return ws

    cycle()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(6):
        ws = cycle()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    free1 = torch.cuda.mem_get_info()[0]
    print(f"plan workspace {ws} B; free device memory {free0} -> {free1}", flush=True)
    assert free0 - free1 < ws


def test_step_textbook_reaches_tolerance(gpu):
    import aerojax_amd as aj
    nx, ny = 512, 128
    c = case(nx, ny, 4.0, 1.0, kind="cylinder")
    st = aj.initial_state(dev(c["u"], torch.float32), dev(c["v"], torch.float32), 1e-3)
    new = aj.step_pure(st, dev(c["mask"], torch.float32), dev(c["sdf"], torch.float32), c["dx"],
                       c["dy"], 5e-4, 1.0, mg_variant="textbook", tolerance=1e-6)
    plan = aj.pressure_solvers.get_plan(nx, ny, c["dx"], c["dy"], torch.float32, "textbook", 0,
                                        new.u.device)
    s = plan.stats()
    assert s["converged"] and s["norm"] <= 1e-6
    assert torch.isfinite(new.u).all()


def test_textbook_graph_replays_match_eager(gpu):
    """The captured outer iteration, replayed across solves with eager work
    and synchronisations in between, gives the eager launches' result."""
    import aerojax_amd as aj
    nx, ny = 512, 128
    dx, dy = 4.0 / nx, 1.0 / ny
    plans = {}
    for name, (adaptive, graph) in {"eager": (True, False), "graph": (True, True),
                                    "graph_noadapt": (False, True)}.items():
        pl = aj.MultigridPlan(nx, ny, dx, dy, torch.float32, "textbook", 0)
        pl.set_option(adaptive=adaptive, graph=graph)
        plans[name] = pl
    rng = np.random.default_rng(12)
    for k in range(4):
        b = dev(rng.normal(size=(nx, ny)) * 300 * (1 + k), torch.float32)
        outs, stats = {}, {}
        for name, pl in plans.items():
            o = torch.empty_like(b)
            pl.solve(b, o, None, 40, 1e-6)
            torch.cuda.synchronize()
            outs[name], stats[name] = o.clone(), pl.stats()
            b.mul_(1.0)  # eager work between replays
        for name in ("graph", "graph_noadapt"):
            assert torch.equal(outs[name], outs["eager"]), (k, name)
            assert stats[name]["cycles"] == stats["eager"]["cycles"]
        assert stats["eager"]["converged"]
    for pl in plans.values():
        pl.close()


@pytest.mark.parametrize("dtype", DTYPES)
def test_flow_diagnostics_parity(gpu, dtype):
    """Viewer / debug metrics after one reference step of the 512x96 airfoil."""
    import aerojax_amd as aj
    c = case(512, 96, 4.0, 0.75)
    dx, dy, dt = c["dx"], c["dy"], 1e-3
    un, vn, p = R.step_pure(c["u"], c["v"], c["mask"], c["sdf"], dx, dy, 5e-4, 1.0, dt,
                            v_cycles=3)
    ref = R.flow_diagnostics(un, vn, p, c["mask"], dx, dy, dt)
    got = aj.flow_diagnostics(dev(un, dtype), dev(vn, dtype), dev(p, dtype),
                              dev(c["mask"], dtype), dx, dy, dt)
    tol = 1e-5 if dtype == torch.float32 else 1e-12
    for k, r in ref.items():
        print(f"diag {k}: {got[k]:.6e} vs {r:.6e}", flush=True)
        assert math.isclose(got[k], r, rel_tol=tol, abs_tol=1e-300), k


# ------------------------------------------------------------------ batched


@pytest.mark.parametrize("dtype", DTYPES)
def test_batched_step_matches_oracle_per_instance(gpu, dtype):
    """An uneven batch (3 x 64x32 and 2 x 128x64 channels, each with its own
    obstacle and viscosity) stepped twice in shared launches; every
    instance's step equals the oracle's step_pure (reference multigrid, 3
    V-cycles) from the same state (the seeded fields, then the batch's own
    state after step 1, in the working precision) to the per-step bar."""
    import aerojax_amd as aj
    shapes = [(64, 32, 3), (128, 64, 2)]
    groups, cases, nus = [], [], []
    for (nx, ny, cnt) in shapes:
        for k in range(cnt):
            c = case(nx, ny, 4.0, 1.0, kind="naca" if k % 2 == 0 else "cylinder", seed=k + nx)
            cases.append(c)
            nus.append(2e-3 * (1.0 + 0.5 * k))
        groups.append((nx, ny, cnt, cases[-1]["dx"], cases[-1]["dy"]))
    dt = 1e-3
    f = NP[dtype]
    b = aj.BatchedStepper(groups, [c["mask"] for c in cases], [c["sdf"] for c in cases], nus,
                          [c["u"] for c in cases], [c["v"] for c in cases], dt, v_cycles=3,
                          dtype=dtype)
    # the state each step starts from, as the batch holds it
    prev = [(c["u"].astype(f).astype(np.float64), c["v"].astype(f).astype(np.float64))
            for c in cases]
    mask = [c["mask"].astype(f).astype(np.float64) for c in cases]
    sdf = [c["sdf"].astype(f).astype(np.float64) for c in cases]
    tol = TOL[dtype]
    for step in (1, 2):
        b.step()
        torch.cuda.synchronize()
        for k, (c, nu) in enumerate(zip(cases, nus)):
            ru, rv, rp = R.step_pure(prev[k][0], prev[k][1], mask[k], sdf[k], c["dx"], c["dy"],
                                     nu, 1.0, dt, v_cycles=3)
            for name, got, r in (("u", b.view(b.u, k), ru), ("v", b.view(b.v, k), rv),
                                 ("p", b.view(b.p, k), rp)):
                e = R.rel_l2(host(got), r)
                print(f"batched {dtype} step {step} instance {k} {name}: rel_l2 {e:.2e}",
                      flush=True)
                assert e <= tol, (step, k, name, e)
        prev = [(host(b.view(b.u, k)), host(b.view(b.v, k))) for k in range(len(cases))]
    st = b.stats()
    assert all(s["cycles"] == 3 for s in st)


def test_batched_graph_replays_match_eager(gpu):
    """Graph-captured batched steps (two cached graphs, alternating buffers)
    equal eager launches over several steps with host syncs in between."""
    import aerojax_amd as aj
    from aerojax_amd import workloads
    w = workloads.batched_sweep_case(n=6, seed=3)
    g = w["grid"]
    mk = lambda: aj.BatchedStepper([(g.nx, g.ny, 6, g.dx, g.dy)], w["masks"], w["sdfs"], w["nus"],
                                   w["u"], w["v"], w["dt"], v_cycles=3)
    a, e = mk(), mk()
    e.set_graph(False)
    for _ in range(4):
        a.step()
        e.step()
        torch.cuda.synchronize()
    assert torch.equal(a.u, e.u) and torch.equal(a.v, e.v) and torch.equal(a.p, e.p)


@pytest.mark.parametrize("cpo", [1, 3])
def test_textbook_cycles_per_outer_iteration(gpu, cpo):
    """AJ_MG_OPT_CYCLES: V-cycles per FP64 defect-correction iteration.  The
    solve reaches the tolerance, counts cpo V-cycles per outer iteration, and
    its first outer iteration equals cpo V-cycles of the oracle's FP32
    multigrid on the same rhs (the check of one outer iteration with 2)."""
    import aerojax_amd as aj
    nx, ny = 512, 128
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(12)
    b = rng.normal(size=(nx, ny)) * 300
    pl = aj.MultigridPlan(nx, ny, dx, dy, torch.float64, "textbook", 2)
    pl.set_option(cycles_per_outer=cpo)
    p = torch.empty((nx, ny), dtype=torch.float64, device=gpu)
    pl.solve(dev(b, torch.float64), p, None, 40, 1e-6)
    st = pl.stats()
    print(f"cycles per outer {cpo}: cycles {st['cycles']} outer {st['outer']} "
          f"norms {['%.2e' % x for x in st['norms'][:st['outer'] + 1]]}", flush=True)
    assert st["converged"] and st["norm"] <= 1e-6
    assert st["cycles"] == cpo * st["outer"]
    pl.solve(dev(b, torch.float64), p, None, cpo, 0.0)
    st1 = pl.stats()
    mg = T.TextbookMG(nx, ny, dx, dy, levels=2)
    ref, st_ref = mg.solve(b, tol=0.0, max_outer=1, cycles_per_outer=cpo)
    assert st1["cycles"] == cpo == st_ref["cycles"] and st1["outer"] == 1
    check(host(p), ref, None, torch.float32, f"mg-textbook {cpo} V-cycles")
    with pytest.raises(RuntimeError):
        pl.set_option(cycles_per_outer=9)
    pl.close()


def test_textbook_relative_tolerance(gpu):
    """AJ_MG_OPT_RELATIVE: ||b - A p|| <= tol ||b||, with and without a guess."""
    import aerojax_amd as aj
    nx, ny = 512, 128
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(8)
    b = dev(rng.normal(size=(nx, ny)) * 300, torch.float32)
    pl = aj.MultigridPlan(nx, ny, dx, dy, torch.float32, "textbook", 0)
    pl.set_option(relative=True)
    p = torch.empty_like(b)
    for p0 in (None, torch.zeros_like(b)):
        pl.solve(b, p, p0, 60, 1e-8)
        s = pl.stats()
        print("relative", s["norms"], flush=True)
        assert s["converged"] and s["norm"] <= 1e-8 * s["norms"][0]
        assert s["norms"][-2] > 1e-8 * s["norms"][0]  # the relative bound decided the stop
    pl.close()


def test_autotuned_launches_are_bit_identical(gpu):
    """Tile shapes change which workgroup computes a cell, not the arithmetic:
    the RK step and a textbook solve give identical bits with tuning off
    (first candidate everywhere) and on; the tuner records its decisions."""
    import aerojax_amd as aj
    from aerojax_amd import _lib
    c = case(512, 128, 4.0, 1.0, kind="cylinder")
    dx, dy = c["dx"], c["dy"]
    rng = np.random.default_rng(9)
    b = dev(rng.normal(size=(512, 128)) * 100, torch.float32)
    res = []
    rep = ""
    for on in (0, 1, 2):  # 2: tuning on, decisions preset from the previous pass
        _lib.load().aj_tune_enable(1 if on else 0)
        _lib.load().aj_tune_reset()
        if on == 2:
            assert _lib.tune_load(rep) == len(rep.splitlines())
        us, vs = aj.rk_step_unified(dev(c["u"], torch.float32), dev(c["v"], torch.float32), 1e-3,
                                    5e-4, dx, dy, dev(c["mask"], torch.float32),
                                    sdf=dev(c["sdf"], torch.float32), brinkman_eta=0.005)
        pl = aj.MultigridPlan(512, 128, dx, dy, torch.float32, "textbook", 3)
        p = torch.empty_like(b)
        pl.solve(b, p, None, 40, 1e-6)
        torch.cuda.synchronize()
        res.append((us.clone(), vs.clone(), p.clone()))
        pl.close()
        if on == 1:
            rep = _lib.tune_report()
            print(rep, flush=True)
            assert "rk/f32/rk3/512x128" in rep and "tb_pre/1x512x128" in rep
        if on == 2:  # the preset picks are the ones launched, nothing re-tuned
            picks = lambda t: sorted(l.split(" [")[0] for l in t.splitlines())  # noqa: E731
            assert picks(_lib.tune_report()) == picks(rep)
    _lib.load().aj_tune_enable(1)
    for r in res[1:]:
        for a, bb in zip(res[0], r):
            assert torch.equal(a, bb)


def test_profile_repeat_counts_and_keeps_results(gpu):
    """With a profiling repeat factor the tunable launches run r times inside
    their event pair and count r launches; the outputs (RK predictor, a
    textbook solve) are bit-identical to an unprofiled run."""
    import aerojax_amd as aj
    from aerojax_amd import _lib
    c = case(512, 128, 4.0, 1.0, kind="cylinder")
    dx, dy = c["dx"], c["dy"]
    rng = np.random.default_rng(11)
    b = dev(rng.normal(size=(1024, 256)) * 100, torch.float32)  # two levels: pre0 runs
    pl = aj.MultigridPlan(1024, 256, dx / 2, dy / 2, torch.float32, "textbook", 0)

    def run():
        us, vs = aj.rk_step_unified(dev(c["u"], torch.float32), dev(c["v"], torch.float32), 1e-3,
                                    5e-4, dx, dy, dev(c["mask"], torch.float32),
                                    sdf=dev(c["sdf"], torch.float32), brinkman_eta=0.005)
        p = torch.empty_like(b)
        pl.solve(b, p, None, 40, 1e-6)
        torch.cuda.synchronize()
        return us.clone(), vs.clone(), p.clone()

    ref = run()
    counts = {}
    for r in (1, 4):
        _lib.profile_begin(1000, r)
        out = run()
        _lib.profile_end()
        counts[r] = _lib.profile_read()
        for a, bb in zip(out, ref):
            assert torch.equal(a, bb)
    _lib.profile_begin(1, 1)
    _lib.profile_end()
    pl.close()
    assert counts[4]["rk"]["count"] == 4 * counts[1]["rk"]["count"]
    assert counts[4]["mg_pre0"]["count"] == 4 * counts[1]["mg_pre0"]["count"]
    assert counts[4]["mg_resid"]["count"] == counts[1]["mg_resid"]["count"]  # not a tuned launch


def test_config_small_500_step_trajectory(gpu):
    """The 512x96 airfoil configuration for its full 500 steps (reference
    multigrid, 3 V-cycles): the FP64 GPU trajectory stays on the FP64 oracle's
    (round-off level), the FP32 one is reported against it."""
    import aerojax_amd as aj
    from aerojax_amd import workloads
    w = workloads.small_airfoil()
    g = w["grid"]
    u64, v64 = w["u"], w["v"]
    steppers = {dt: aj.CollocatedStepper(g.nx, g.ny, g.dx, g.dy, w["mask"], w["sdf"], w["nu"],
                                         1.0, w["dt"], dtype=dt, v_cycles=3)
                for dt in (torch.float64, torch.float32)}
    for s in steppers.values():
        s.load(w["u"], w["v"])
    errs = {torch.float64: [], torch.float32: []}
    for k in range(500):
        u64, v64, p64 = R.step_pure(u64, v64, w["mask"], w["sdf"], g.dx, g.dy, w["nu"], 1.0,
                                    w["dt"], v_cycles=3)
        for dt, s in steppers.items():
            s.step()
        if (k + 1) % 100 == 0:
            for dt, s in steppers.items():
                errs[dt].append(R.rel_l2(host(s.state[0]), u64))
    print("C1 500-step u rel err FP64:", ["%.1e" % e for e in errs[torch.float64]],
          "FP32:", ["%.1e" % e for e in errs[torch.float32]], flush=True)
    assert max(errs[torch.float64]) < 1e-10
    assert max(errs[torch.float32]) < 1e-3
    for s in steppers.values():
        assert torch.isfinite(s.state[0]).all()
