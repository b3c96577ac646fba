"""Initial guesses of the to-tolerance multigrid: the temporal guess kept on
the plan (previous FP64 solution, then the linear extrapolation of the last
two) and what the trained warm-start network gains over the zero guess."""
import numpy as np
import pytest
import torch

from oracle import textbook_mg as T

pytestmark = pytest.mark.gpu


def _host(t):
    return t.detach().double().cpu().numpy()


def test_temporal_guess_uses_the_previous_fp64_results(gpu):
    """Right-hand sides b_k = b0 + k b1: the plan's guess for solve 2 is the
    FP64 result of solve 1 and for solve 3 the extrapolation 2 p2 - p1; the
    guess residual it reports equals the host's FP64 residual of that guess,
    and resetting the option drops the history."""
    import aerojax_amd as aj
    nx, ny = 256, 64
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(4)
    b0 = rng.normal(size=(nx, ny)) * 300.0
    b1 = rng.normal(size=(nx, ny)) * 30.0
    pl = aj.MultigridPlan(nx, ny, dx, dy, torch.float64, "textbook", 0)
    pl.set_option(temporal=True)
    mg = T.TextbookMG(nx, ny, dx, dy)
    ps, stats = [], []
    for k in range(4):
        b = b0 + k * b1
        out = torch.empty((nx, ny), dtype=torch.float64, device=gpu)
        pl.solve(torch.from_numpy(b).to(gpu), out, None, 40, 1e-6)
        st = pl.stats()
        stats.append(st)
        bb = b.copy()
        bb[-1] = 0.0
        if k == 1:
            g = ps[0]
        elif k >= 2:
            g = 2.0 * ps[-1] - ps[-2]
        if k >= 1:
            hg = mg.norm(mg.residual64(g, bb))
            print(f"solve {k}: guess residual {st['guess_norm']:.6e} (host {hg:.6e}), zero "
                  f"{st['zero_norm']:.3e}, kept {st['used_guess']}, cycles {st['cycles']}",
                  flush=True)
            assert st["used_guess"]
            assert np.isclose(st["guess_norm"], hg, rtol=1e-9, atol=1e-12)
        else:
            assert not st["used_guess"] and st["guess_norm"] == 0.0
        assert st["converged"] and st["norm"] <= 1e-6
        ps.append(_host(out))
    # b is linear in k, so the extrapolated guess is already converged
    assert stats[2]["guess_norm"] < 1e-3 * stats[2]["zero_norm"]
    assert stats[3]["cycles"] <= stats[0]["cycles"] - 2
    pl.set_option(temporal=True)
    out = torch.empty((nx, ny), dtype=torch.float64, device=gpu)
    pl.solve(torch.from_numpy(b0).to(gpu), out, None, 40, 1e-6)
    st = pl.stats()
    assert not st["used_guess"] and st["guess_norm"] == 0.0
    pl.close()


def test_temporal_stepper_deferred_equals_blocking_and_matches_plain(gpu):
    """A stepper with the temporal guess gives the same bits with and without
    the deferred convergence read, and the plain stepper's flow to the level
    the 1e-6 tolerance leaves."""
    import aerojax_amd as aj
    from aerojax_amd import workloads
    c = workloads.warm_start_case(3, 1024, 256)  # two levels (512x128 is one exact solve)
    g = c["grid"]
    mk = lambda **kw: aj.CollocatedStepper(g.nx, g.ny, g.dx, g.dy, c["mask"], c["sdf"],  # noqa
                                           c["nu"], 1.0, c["dt"], mg_variant="textbook",
                                           tolerance=1e-6, max_cycles=60, **kw)
    a = mk(temporal_guess=True)
    d = mk(temporal_guess=True, deferred_check=True)
    p = mk()
    for s in (a, d, p):
        s.load(c["u"], c["v"])
    used, cyc_t, cyc_p = [], [], []
    for _ in range(12):
        a.step()
        d.step()
        p.step()
        d.finish()
        sa, sp = a.plan.stats(), p.plan.stats()
        used.append(sa["used_guess"])
        cyc_t.append(sa["cycles"])
        cyc_p.append(sp["cycles"])
        ua, va = a.state
        ud, vd = d.state
        assert torch.equal(ua, ud) and torch.equal(va, vd) and torch.equal(a.p, d.p)
    up, vp = p.state
    ua, va = a.state
    eu = float((ua - up).norm() / up.norm())
    ev = float((va - vp).norm() / vp.norm())
    print(f"temporal guess kept {used}; V-cycles temporal {cyc_t} plain {cyc_p}; "
          f"u/v vs plain {eu:.2e} {ev:.2e}", flush=True)
    assert all(used[1:])
    assert sum(cyc_t) <= sum(cyc_p)
    assert eu < 1e-6 and ev < 1e-5
    for s in (a, d, p):
        s.close()


def test_trained_network_gain_over_zero_guess(gpu):
    """What the trained warm-start network gains: it is trained on solver
    pairs of 40 flows (held-out loss on 8 other flows must drop), then on a
    flow it has not seen the V-cycles to ||div|| <= 1e-6 are counted from the
    zero guess, from the network's guess with the guess test off, from the
    network's guess under the guess test (the warm-start arm) and, for
    scale, from the temporal guess.  The guess test guarantees that a kept
    guess has at most the zero guess's residual; the counts are printed."""
    import aerojax_amd as aj
    from aerojax_amd import workloads
    from aerojax_amd.nn import WarmStartNet
    from aerojax_amd.warmstart import generate_pairs, normalise_targets, train, validation_loss
    nx, ny = 1024, 256  # two levels, like the 2048x512 workload's three
    net = WarmStartNet(base_channels=32, levels=4, seed=0, r=2, device=gpu)
    X, Tt = generate_pairs(net, 400, range(100, 140), nx, ny, steps_per_seed=24, stride=2, skip=4)
    Xv, Tv = generate_pairs(net, 80, range(5000, 5008), nx, ny, steps_per_seed=24, stride=2,
                            skip=4)
    q = normalise_targets(net, Tt)
    Tv.div_(q)
    l0 = validation_loss(net, Xv, Tv)
    hist = train(net, X, Tt, epochs=6, batch=16, lr=5e-4, X_val=Xv, T_val=Tv)
    print(f"pairs {X.shape[0]} / held-out {Xv.shape[0]}; held-out loss {l0:.3e} -> "
          f"{hist['val']} (best epoch {hist['best_epoch']})", flush=True)
    assert min(hist["val"]) < 0.5 * l0
    assert validation_loss(net, Xv, Tv) == pytest.approx(min(hist["val"]), rel=1e-6)

    c = workloads.warm_start_case(777, nx, ny)
    g = c["grid"]
    stp = aj.CollocatedStepper(nx, ny, g.dx, g.dy, c["mask"], c["sdf"], c["nu"], 1.0, c["dt"],
                               mg_variant="textbook", tolerance=1e-6, max_cycles=60)
    stp.load(c["u"], c["v"])
    free = aj.MultigridPlan(nx, ny, g.dx, g.dy, torch.float32, "textbook", 0)
    free.set_option(guess_test=False)
    tested = aj.MultigridPlan(nx, ny, g.dx, g.dy, torch.float32, "textbook", 0)
    temporal = aj.MultigridPlan(nx, ny, g.dx, g.dy, torch.float32, "textbook", 0)
    temporal.set_option(temporal=True)
    out = torch.empty((nx, ny), device=gpu)
    rows = []
    for k in range(14):
        stp.step()
        rhs = stp.rhs.clone()
        temporal.solve(rhs, out, None, 60, 1e-6)
        st_t = temporal.stats()
        if k < 4:
            continue
        zero = stp.plan.stats()
        guess = net.warm_start(rhs, stp.sdf)
        free.solve(rhs, out, guess, 60, 1e-6)
        st_f = free.stats()
        tested.solve(rhs, out, guess, 60, 1e-6)
        st_g = tested.stats()
        for s in (zero, st_f, st_g, st_t):
            assert s["converged"] and s["norm"] <= 1e-6
        if st_g["used_guess"]:
            assert st_g["guess_norm"] <= st_g["zero_norm"]
        rows.append((k, zero["cycles"], st_f["cycles"], st_f["guess_norm"] / st_f["zero_norm"],
                     st_g["cycles"], st_g["used_guess"], st_t["cycles"]))
    print("step | V-cycles zero | net (no test) | net guess/zero residual | net under test "
          "(kept) | temporal", flush=True)
    for r in rows:
        print(f"{r[0]:4d} | {r[1]:13d} | {r[2]:13d} | {r[3]:23.3f} | {r[4]:8d} ({r[5]}) | "
              f"{r[6]:8d}", flush=True)
    saved_net = sum(r[1] - r[4] for r in rows)
    saved_temporal = sum(r[1] - r[6] for r in rows)
    print(f"V-cycles saved over {len(rows)} steps: network {saved_net}, temporal guess "
          f"{saved_temporal}", flush=True)
    assert saved_net >= 0  # a guess under the test never costs cycles here
    for pl in (free, tested, temporal):
        pl.close()
    stp.close()
    net.close()
