"""Batched to-tolerance solving on the GPU: batched textbook plans against
single-grid solves of the same right-hand sides (bit for bit, per-instance
gates and guess tests) and the sparse direct solve; the uneven batched step
with Smagorinsky LES and the textbook solver against the oracle's step
(rk_step_unified with nu_sgs, rhs, direct pressure solve, projection) per
instance; the batched warm start against single-instance warm starts."""
import numpy as np
import pytest
import torch

from oracle import reference_step as R
from oracle import textbook_mg as T

from test_gpu_parity import case, dev, host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(512, 128), (1024, 256)])
def test_batched_textbook_plan_matches_single_solves(gpu, shape):
    """Three instances whose right-hand sides differ by orders of magnitude
    converge after different numbers of outer iterations; each equals the
    single-grid solve of its rhs.  512x128 is one exact level, 1024x256 two."""
    import aerojax_amd as aj
    nx, ny = shape
    B = 3
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(nx)
    bs = [rng.normal(size=(nx, ny)) * 10.0 ** (3 * k) for k in range(B)]
    b = torch.stack([dev(x, torch.float32) for x in bs]).contiguous()
    pb = aj.MultigridPlan(nx, ny, dx, dy, torch.float32, "textbook", 0, count=B)
    ps = aj.MultigridPlan(nx, ny, dx, dy, torch.float32, "textbook", 0)
    assert pb.count == B
    out = torch.empty_like(b)
    pb.solve(b, out, None, 60, 1e-6)
    sts = pb.stats_all()
    outers = []
    for k in range(B):
        o1 = torch.empty_like(b[k])
        ps.solve(b[k].contiguous(), o1, None, 60, 1e-6)
        s1 = ps.stats()
        print(f"batched textbook {shape} instance {k}: outer {sts[k]['outer']} "
              f"norms {['%.2e' % x for x in sts[k]['norms']]}", flush=True)
        assert sts[k]["converged"] and sts[k]["norm"] <= 1e-6 and not sts[k]["nonfinite"]
        assert torch.equal(out[k], o1), k
        assert sts[k]["cycles"] == s1["cycles"] and sts[k]["norms"] == s1["norms"]
        outers.append(sts[k]["outer"])
    assert len(set(outers)) > 1  # the gates really were per instance
    if nx * ny <= 65536:
        pd = T.direct_solve(host(b[0]), dx, dy)
        assert R.rel_l2(host(out[0]), pd) < 1e-6
    # per-instance guess test: the converged solution is kept, a wild guess
    # rejected (restart from zero), a zero guess kept (equal residuals)
    p0 = torch.stack([out[0], 1e6 * torch.randn_like(out[1]), torch.zeros_like(out[2])]).contiguous()
    out2 = torch.empty_like(b)
    pb.solve(b, out2, p0, 60, 1e-6)
    st2 = pb.stats_all()
    assert [s["used_guess"] for s in st2] == [True, False, True]
    for k in range(B):
        o1 = torch.empty_like(b[k])
        ps.solve(b[k].contiguous(), o1, p0[k].contiguous(), 60, 1e-6)
        s1 = ps.stats()
        assert torch.equal(out2[k], o1), k
        assert st2[k]["norms"] == s1["norms"] and st2[k]["used_guess"] == s1["used_guess"]
        assert st2[k]["converged"]
    # the FP32-rounded solution is a better start than zero
    assert st2[0]["outer"] <= sts[0]["outer"]
    pb.close()
    ps.close()


def _oracle_les_textbook_step(c, u, v, nu, dt, cs):
    """One step of the oracle's composition from (u, v): constant
    Smagorinsky fed to the RK step, predictor BCs, rhs, the exact pressure
    of the to-tolerance operator (sparse direct solve), projection."""
    dx, dy = c["dx"], c["dy"]
    nsgs = R.constant_smagorinsky(u, v, dx, dy, (dx * dy) ** 0.5, cs)
    us, vs = R.rk_step_unified(u, v, dt, nu, dx, dy, c["mask"], sdf=c["sdf"], U_inf=1.0,
                               nu_sgs=nsgs, nu_hyper_ratio=0.0, slip_walls=True,
                               fast_mode=False, brinkman_eta=0.005)
    us, vs = R.predictor_bcs(us, vs, 1.0)
    rhs = R.pressure_rhs(us, vs, c["mask"], dt, dx, dy)
    p = T.direct_solve(rhs, dx, dy)
    u, v = R.project(us, vs, p, c["mask"], dt, dx, dy, 1.0)
    return u, v, p


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_batched_textbook_les_step_matches_oracle(gpu, dtype):
    """Uneven batch (2 x 64x32, 2 x 128x64), Smagorinsky Cs = 0.17 fed to the
    RK step, every pressure solved to the tolerance; each of two steps
    against the oracle's composition with a sparse direct pressure solve
    from the same state (the seeded fields, then the batch's own state).
    Velocities to the per-step bar; the pressure to the bar its solve
    tolerance allows (FP64: ||div|| <= 1e-8 of a pressure of norm ~1e2)."""
    import aerojax_amd as aj
    shapes = [(64, 32, 2), (128, 64, 2)]
    groups, cases, nus = [], [], []
    for (nx, ny, cnt) in shapes:
        for k in range(cnt):
            c = case(nx, ny, 4.0, 1.0, kind="cylinder" if k % 2 else "naca", seed=3 * k + nx)
            cases.append(c)
            nus.append(1e-3 * (1.0 + k))
        groups.append((nx, ny, cnt, cases[-1]["dx"], cases[-1]["dy"]))
    dt, cs = 1e-3, 0.17
    tol = 1e-8 if dtype == torch.float64 else 1e-6
    f = np.float64 if dtype == torch.float64 else np.float32
    b = aj.BatchedStepper(groups, [c["mask"] for c in cases], [c["sdf"] for c in cases], nus,
                          [c["u"] for c in cases], [c["v"] for c in cases], dt, dtype=dtype,
                          solver="textbook", tolerance=tol, max_cycles=60, smagorinsky_cs=cs)
    # the oracle runs in FP64 on the batch's inputs as the batch holds them
    rd = lambda a: a.astype(f).astype(np.float64)  # noqa: E731
    oc = [dict(c, mask=rd(c["mask"]), sdf=rd(c["sdf"])) for c in cases]
    prev = [(rd(c["u"]), rd(c["v"])) for c in cases]
    bars = {torch.float64: (1e-12, 1e-9), torch.float32: (1e-5, 1e-5)}[dtype]
    for step in (1, 2):
        b.step()
        torch.cuda.synchronize()
        st = b.stats()
        assert all(s["converged"] and s["norm"] <= tol for s in st), st
        for k, (c, nu) in enumerate(zip(oc, nus)):
            ru, rv, rp = _oracle_les_textbook_step(c, prev[k][0], prev[k][1], nu, dt, cs)
            for name, got, r, bar in (("u", b.view(b.u, k), ru, bars[0]),
                                      ("v", b.view(b.v, k), rv, bars[0]),
                                      ("p", b.view(b.p, k), rp, bars[1])):
                e = R.rel_l2(host(got), r)
                print(f"batched textbook+LES {dtype} step {step} instance {k} {name}: "
                      f"rel_l2 {e:.2e}", flush=True)
                assert e <= bar, (step, k, name, e)
        prev = [(host(b.view(b.u, k)), host(b.view(b.v, k))) for k in range(len(cases))]
    b.close()


def test_batched_warm_start_matches_single(gpu):
    """One inference of batch B equals B single-instance warm starts up to
    bf16 rounding: the launch tuner may split a layer's channel sum
    differently at batch B and at batch 1, and a different FP32 summation
    order can flip the bf16 rounding of an activation (2^-8 relative)."""
    from aerojax_amd.nn import WarmStartNet
    nx, ny, B = 256, 64, 3
    rng = np.random.default_rng(11)
    X = np.linspace(0, 1, nx)[:, None]
    Y = np.linspace(0, 1, ny)[None, :]
    rhs = np.stack([(10 ** k) * np.sin((3 + k) * X) * np.cos(2 * Y) + rng.normal(size=(nx, ny))
                    for k in range(B)]).astype(np.float32)
    sdf = np.stack([np.hypot(X - 0.3 - 0.1 * k, Y - 0.5) - 0.1 for k in range(B)]).astype(np.float32)
    net = WarmStartNet(base_channels=32, levels=4, seed=5)
    rt, st = dev(rhs, torch.float32), dev(sdf, torch.float32)
    p0 = torch.empty_like(rt)
    net.warm_start_batched(rt, st, p0, B, nx, ny)
    for k in range(B):
        one = net.warm_start(rt[k].contiguous(), st[k].contiguous())
        e = R.rel_l2(host(p0[k]), host(one))
        print(f"batched warm start instance {k}: rel diff {e:.2e}", flush=True)
        assert e < 1e-3
        assert torch.all(p0[k, -1] == 0)
    net.close()


def test_uneven_batch_with_warm_start_reaches_tolerance(gpu):
    """The uneven learned configuration at small sizes: LES on, network
    guesses per group (untrained network), every instance solved to
    ||div|| <= 1e-6 each step; the result matches the same batch stepped
    without the network to within the tolerance's reach."""
    import aerojax_amd as aj
    from aerojax_amd import workloads
    from aerojax_amd.nn import WarmStartNet
    w = workloads.uneven_batch_case(per_size=2, sizes=((128, 32), (256, 64), (512, 128)), seed=1)
    net = WarmStartNet(base_channels=32, levels=4, seed=2)
    mk = lambda ws: aj.BatchedStepper(w["groups"], w["masks"], w["sdfs"], w["nus"], w["u"],  # noqa
                                      w["v"], w["dt"], solver="textbook", tolerance=w["tol"],
                                      max_cycles=60, smagorinsky_cs=w["cs"], warm_start=ws)
    a, z = mk(net), mk(None)
    for _ in range(3):
        a.step()
        z.step()
        sa, sz = a.stats(), z.stats()
        assert all(s["converged"] and s["norm"] <= w["tol"] for s in sa + sz)
    torch.cuda.synchronize()
    assert torch.isfinite(a.u).all() and torch.isfinite(a.p).all()
    e = R.rel_l2(host(a.u), host(z.u))
    print("uneven batch: used guesses", [s["used_guess"] for s in sa], "u rel diff", e, flush=True)
    assert e < 1e-5
    a.close()
    z.close()
    net.close()


def test_divergence_snapshot_and_dataset(gpu, tmp_path):
    """Periodic divergence against the oracle; a snapshot's divergence and CFL
    against the oracle; a dataset written from device fields reads back and
    gives the same training pairs as make_sample."""
    import aerojax_amd as aj
    from aerojax_amd.nn import WarmStartNet
    c = case(128, 64, 4.0, 1.0, kind="cylinder")
    dx, dy, dt = c["dx"], c["dy"], 1e-3
    ref = R.grad_x_periodic(c["u"], dx) + R.grad_y_periodic(c["v"], dy)
    for dtype in (torch.float32, torch.float64):
        got = aj.divergence(dev(c["u"], dtype), dev(c["v"], dtype), dx, dy)
        e = R.rel_l2(host(got), ref)
        print(f"periodic divergence {dtype}: rel_l2 {e:.2e}", flush=True)
        assert e <= (1e-5 if dtype == torch.float32 else 1e-12)
    u, v, p = (dev(x, torch.float64) for x in (c["u"], c["v"], c["u"] * 0.5))
    s = aj.take_snapshot(u, v, p, dev(c["mask"], torch.float64), dt, 5e-4, dx, dy, Re=1000.0,
                         iteration=3)
    assert R.rel_l2(s.divergence, R.divergence_nonperiodic(c["u"], c["v"], dx, dy)) <= 1e-12
    assert np.isclose(s.CFL, R.check_cfl(c["u"], c["v"], dt, dx, dy)[1], rtol=1e-12)
    s.save(str(tmp_path / "s.npz"))
    assert aj.Snapshot.load(str(tmp_path / "s.npz")).iteration == 3
    # dataset
    nx, ny = 128, 64
    w = aj.DatasetWriter(nx, ny, dx, dy, dt, capacity=3)
    rng = np.random.default_rng(3)
    fields = []
    for k in range(2):
        f = [dev(rng.normal(size=(nx, ny)), torch.float32) for _ in range(4)]
        rhs, sdf = dev(rng.normal(size=(nx, ny)) * 50, torch.float32), dev(c["sdf"], torch.float32)
        w.record(f[0], f[1], f[2], f[3], rhs=rhs, sdf=sdf)
        fields.append((f, rhs, sdf))
    path = w.save(str(tmp_path / "training_data.npz"))
    d = aj.load_dataset(path)
    assert d["u"].shape == (2, nx, ny) and d["nx"] == nx and d["dt"] == dt
    for k, (f, rhs, sdf) in enumerate(fields):
        assert np.array_equal(d["p"][k], host(f[2]))
        dref = R.grad_x_periodic(host(f[0]), dx) + R.grad_y_periodic(host(f[1]), dy)
        assert R.rel_l2(d["divergence"][k], dref) <= 1e-5
    net = WarmStartNet(base_channels=32, levels=4, seed=1)
    X, T = aj.pairs_from_dataset(net, d)
    H, W = nx >> net.r, ny >> net.r
    for k, (f, rhs, sdf) in enumerate(fields):
        xs = torch.empty((H, W, 2), device="cuda")
        ts = torch.empty((H, W), device="cuda")
        net.make_sample(rhs, sdf, f[2], xs, ts)
        assert torch.equal(X[k], xs) and torch.equal(T[k], ts)
    net.close()


def _oracle_geometry(nx, ny, lx, ly, ob):
    _, _, X, Y = R.grid_coords(nx, ny, lx, ly)
    if ob.kind == 1:
        sdf = R.naca_sdf(X, Y, ob.chord, ob.aoa_deg, ob.x0, ob.y0, ob.m, ob.p, ob.t)
    else:
        sdf = R.cylinder_sdf(X, Y, ob.x0, ob.y0, ob.radius)
    return sdf, R.mask_from_sdf(sdf, ob.eps)


def _check_geometry(got_sdf, got_mask, ref_sdf, ref_mask, dtype, what):
    gs, gm = host(got_sdf), host(got_mask)
    assert np.array_equal(np.isnan(gs), np.isnan(ref_sdf)), what  # x_norm == 0 of 00xx sections
    ok = ~np.isnan(ref_sdf)
    tol = 1e-12 if dtype == torch.float64 else 1e-6
    es, em = R.rel_l2(gs[ok], ref_sdf[ok]), R.rel_l2(gm[ok], ref_mask[ok])
    print(f"geometry {what} {dtype}: sdf {es:.2e} mask {em:.2e}", flush=True)
    assert es <= tol and em <= tol, what


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_device_geometry_matches_oracle(gpu, dtype):
    """SDF and mask of NACA sections (cambered, symmetric, nose up / down) and
    a cylinder computed on the GPU against the oracle's host restatement."""
    from aerojax_amd import geometry as G
    grid = G.GridParams(512, 96, 4.0, 0.75)
    obs = {"naca2412+7": G.naca_obstacle("2412", 0.5, 7.0, 1.0, 0.375, 0.05),
           "naca6315-12": G.naca_obstacle("6315", 0.4, -12.0, 0.8, 0.4, 0.03),
           "naca0012+10": G.naca_obstacle("0012", 0.5, 10.0, 1.0, 0.375, 0.05),
           "cylinder": G.cylinder_obstacle(1.2, 0.37, 0.1, 0.05)}
    for name, ob in obs.items():
        sdf, mask = G.device_geometry(grid, ob, dtype)
        rs, rm = _oracle_geometry(grid.nx, grid.ny, grid.lx, grid.ly, ob)
        _check_geometry(sdf, mask, rs, rm, dtype, name)
    # the host setup path of the product gives the same fields
    geo = G.naca_geometry(grid, "2412", chord=0.5, aoa_deg=7.0, x_le=1.0, y_le=0.375, eps=0.05)
    sdf, mask = G.device_geometry(grid, obs["naca2412+7"], torch.float64)
    assert R.rel_l2(host(sdf), geo.sdf) <= 1e-12 and R.rel_l2(host(mask), geo.mask) <= 1e-12


def test_batched_geometry_matches_per_instance(gpu):
    """Per-instance obstacles of an uneven batch built on the GPU equal the
    oracle's geometry of each instance (lx = nx dx, ly = ny dy)."""
    import aerojax_amd as aj
    from aerojax_amd import geometry as G
    groups = [(64, 32, 2, 4.0 / 64, 1.0 / 32), (128, 64, 3, 4.0 / 128, 1.0 / 64)]
    obs = [G.naca_obstacle("0012", 0.5, 5.0, 1.0, 0.5), G.cylinder_obstacle(1.0, 0.5, 0.12),
           G.naca_obstacle("4415", 0.45, -8.0, 0.9, 0.52), G.cylinder_obstacle(1.3, 0.45, 0.08),
           G.naca_obstacle("2312", 0.3, 14.0, 1.1, 0.5)]
    n = len(obs)
    shapes = [(g[0], g[1], g[3], g[4]) for g in groups for _ in range(g[2])]
    u0 = [np.ones((nx, ny)) for nx, ny, _, _ in shapes]
    v0 = [np.zeros((nx, ny)) for nx, ny, _, _ in shapes]
    b = aj.BatchedStepper(groups, None, None, [1e-3] * n, u0, v0, 1e-3, obstacles=obs,
                          dtype=torch.float64)
    for k, ((nx, ny, dx, dy), ob) in enumerate(zip(shapes, obs)):
        rs, rm = _oracle_geometry(nx, ny, nx * dx, ny * dy, ob)
        _check_geometry(b.view(b.sdf, k), b.view(b.mask, k), rs, rm, torch.float64, f"instance {k}")
    b.close()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_dynamic_smagorinsky_parity(gpu, dtype):
    """dynamic_smagorinsky (nu_sgs, C_d, nu + nu_sgs) against the oracle on a
    wake-like field, and the GUI step with les_model='dynamic_smagorinsky'."""
    import aerojax_amd as aj
    c = case(256, 64, 4.0, 1.0, kind="cylinder", seed=4)
    dx, dy = c["dx"], c["dy"]
    delta = (dx * dy) ** 0.5
    f = {torch.float32: np.float32, torch.float64: np.float64}[dtype]
    rnu, rC = R.dynamic_smagorinsky(c["u"], c["v"], dx, dy, delta)
    rnu_d, rC_d = R.dynamic_smagorinsky(c["u"].astype(f), c["v"].astype(f), dx, dy, delta)
    nu_sgs, nu_tot, C = aj.dynamic_smagorinsky(dev(c["u"], dtype), dev(c["v"], dtype), dx, dy,
                                               delta, nu=1e-3)
    # L_ij = u~u~ - (uu)~ cancels (|u| ~ 1 against a sub-filter variance
    # ~1e-3), so FP32 rounding (6e-8) reaches ~1e-4 in C_d whatever the
    # summation order: the FP32 restatement of the reference formula is itself
    # ~7e-4 from FP64.  FP32 bars: 2e-3 against the FP64 oracle, against the
    # FP64 oracle on the FP32-rounded inputs and against the FP32 oracle;
    # FP64 bar: 1e-12.
    tol = 2e-3 if dtype == torch.float32 else 1e-12
    r32 = lambda a: a.astype(f).astype(np.float64)  # noqa: E731
    rnu_r, rC_r = R.dynamic_smagorinsky(r32(c["u"]), r32(c["v"]), dx, dy, delta)
    for name, got, r64, rdt, rr in (("nu_sgs", nu_sgs, rnu, rnu_d, rnu_r), ("C_d", C, rC, rC_d,
                                                                           rC_r)):
        e64, edt, er = (R.rel_l2(host(got), x) for x in (r64, rdt, rr))
        print(f"dynamic smagorinsky {name} {dtype}: vs fp64 {e64:.2e}, vs same dtype {edt:.2e}, "
              f"vs fp64 on rounded inputs {er:.2e}", flush=True)
        assert e64 <= tol, (name, e64)
        assert edt <= tol, (name, edt)
        assert er <= tol, (name, er)
    assert R.rel_l2(host(nu_tot), 1e-3 + rnu) <= tol
    assert host(C).max() <= 0.2 and host(C).min() >= 0.0
    # GUI step with the dynamic model
    kw = dict(use_les=True, les_model="dynamic_smagorinsky", nu_hyper_ratio=0.0, fast_mode=False,
              brinkman_eta=0.01, v_cycles=5)
    ru, rv, rp = R.step_collocated(c["u"].astype(f), c["v"].astype(f), c["mask"].astype(f),
                                   c["sdf"].astype(f), dx, dy, 1e-3, 1.0, 2e-4, **kw)
    u, v, p = aj.step_collocated(dev(c["u"], dtype), dev(c["v"], dtype), dev(c["mask"], dtype),
                                 2e-4, sdf=dev(c["sdf"], dtype), dx=dx, dy=dy, nu=1e-3, U_inf=1.0,
                                 **kw)
    for name, got, r in (("u", u, ru), ("v", v, rv), ("p", p, rp)):
        e = R.rel_l2(host(got), r)
        print(f"GUI step dynamic LES {dtype} {name}: {e:.2e}", flush=True)
        assert e <= (1e-5 if dtype == torch.float32 else 1e-11), name


def test_nonfinite_inputs_are_flagged_per_instance(gpu):
    """A NaN in the right-hand side sets the non-finite flag (the reference
    propagates NaN silently) and ends the to-tolerance solve; in a batched
    plan only that instance is affected, the others converge as alone."""
    import aerojax_amd as aj
    nx, ny, B = 512, 128, 3
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(31)
    b = torch.stack([dev(rng.normal(size=(nx, ny)) * 100, torch.float32) for _ in range(B)])
    b = b.contiguous()
    bad = b.clone()
    bad[1, 10, 20] = float("nan")
    # reference multigrid: all cycles run, the flag is raised
    _, st = aj.poisson_multigrid(bad[1].contiguous(), None, dx, dy, v_cycles=3, return_stats=True)
    assert st["nonfinite"] and not st["converged"]
    # textbook, single grid: the adaptive driver stops on the flag
    _, st = aj.poisson_multigrid_to_tolerance(bad[1].contiguous(), dx, dy, tolerance=1e-6,
                                              max_cycles=40, return_stats=True)
    assert st["nonfinite"] and not st["converged"]
    # batched: instances 0 and 2 converge and equal the clean solves
    pb = aj.MultigridPlan(nx, ny, dx, dy, torch.float32, "textbook", 0, count=B)
    out_bad, out_ok = torch.empty_like(b), torch.empty_like(b)
    pb.solve(bad, out_bad, None, 40, 1e-6)
    sb = pb.stats_all()
    pb.solve(b, out_ok, None, 40, 1e-6)
    so = pb.stats_all()
    assert sb[1]["nonfinite"] and not sb[1]["converged"]
    for k in (0, 2):
        assert sb[k]["converged"] and not sb[k]["nonfinite"]
        assert torch.equal(out_bad[k], out_ok[k]) and sb[k]["norms"] == so[k]["norms"]
    pb.close()


class _SolvedGuess:
    """A warm start whose guess is the converged pressure of the rhs itself
    (solved with a separate plan), so the guess test keeps it."""

    def __init__(self, nx, ny, dx, dy):
        import aerojax_amd as aj
        self.plan = aj.MultigridPlan(nx, ny, dx, dy, torch.float32, "textbook", 0)

    def warm_start(self, rhs, sdf, p0_out=None):
        p0 = torch.empty_like(rhs) if p0_out is None else p0_out
        self.plan.solve(rhs, p0, None, 60, 1e-7)
        return p0

    __call__ = warm_start


@pytest.mark.parametrize("guess", ["network", "solved"])
def test_overlapped_warm_start_equals_sequential(gpu, guess):
    """The overlapped stepper (zero-guess solve concurrent with the guess and
    its test on a second stream) gives the sequential stepper's bits, with a
    rejected guess (untrained network) and with a kept one."""
    import aerojax_amd as aj
    from aerojax_amd.nn import WarmStartNet
    nx, ny = 512, 128
    c = case(nx, ny, 4.0, 1.0, kind="cylinder", seed=7)
    ws = WarmStartNet(base_channels=32, levels=4, seed=3) if guess == "network" else \
        _SolvedGuess(nx, ny, c["dx"], c["dy"])
    mk = lambda ov: aj.CollocatedStepper(nx, ny, c["dx"], c["dy"], c["mask"], c["sdf"], 1e-3,  # noqa
                                         1.0, 1e-3, mg_variant="textbook", tolerance=1e-6,
                                         max_cycles=60, warm_start=ws, overlap_guess=ov)
    a, b = mk(False), mk(True)
    a.load(c["u"], c["v"])
    b.load(c["u"], c["v"])
    for _ in range(3):
        a.step()
        b.step()
        sa, sb = a.plan.stats(), b.plan.stats()
        assert sa["converged"] and sb["converged"]
        used = b.guess_stats["used_guess"]
        assert used == (guess == "solved")
    torch.cuda.synchronize()
    assert torch.equal(a.state[0], b.state[0]) and torch.equal(a.state[1], b.state[1])
    assert torch.equal(a.p, b.p)


def test_warm_start_graph_replays_match_eager(gpu):
    """The warm start of a repeated call signature is captured once and
    replayed: every replay (new inputs in the same buffers, eager work and
    synchronisations in between) equals the eager launches (forced by the
    live-timing mode, which bypasses graphs)."""
    from aerojax_amd import _lib
    from aerojax_amd.nn import WarmStartNet
    nx, ny = 256, 64
    net = WarmStartNet(base_channels=32, levels=4, seed=5)
    rng = np.random.default_rng(12)
    rhs = torch.empty((nx, ny), device="cuda")
    sdf = dev(rng.normal(size=(nx, ny)), torch.float32)
    p0 = torch.empty_like(rhs)
    p1 = torch.empty_like(rhs)
    for k in range(5):
        rhs.copy_(dev(rng.normal(size=(nx, ny)) * 10 ** k, torch.float32))
        got = net.warm_start(rhs, sdf, p0).clone()  # eager, captured, then replays
        torch.cuda.synchronize()
        _ = (rhs * 2).sum().item()  # unrelated eager work
        _lib.profile_begin(1000)
        ref = net.warm_start(rhs, sdf, p1)
        torch.cuda.synchronize()
        _lib.profile_end()
        assert torch.equal(got, ref), k
    net.close()
