"""Benchmark: projection steps/s at 2048x512 solved to ||div|| <= 1e-6,
NN-warm-started multigrid vs plain multigrid.

One step = RK3 predictor (fused HIP kernel) -> pressure rhs -> textbook
multigrid (FP32 V(2,2) cycles in FP64 defect correction) until
sqrt(sum (b - A p)^2 dx dy) <= 1e-6 -> projection, on a 2048x512 FP32 grid
(4x1 channel; obstacle and Re drawn from seed 0).  Arms:
  plain : zero initial guess (the reference's poisson_multigrid start)
  nn    : initial guess from the warm-start U-Net (trained beforehand on
          pairs generated by the same solver on other seeds; zero guess
          whenever the network's guess has the larger residual)
Inputs are resident in HBM before timing starts.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-steps S] [--no-nn]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_arm(aj, case, dev, steps, warmup, warm_start=None, overlap=False, temporal=False):
    """Warmup then `steps` timed steps; returns (steps/s, stepper, per-step stats)."""
    import torch
    g = case["grid"]
    stp = aj.CollocatedStepper(g.nx, g.ny, g.dx, g.dy, case["mask"], case["sdf"], case["nu"],
                               case["U_inf"], case["dt"], dtype=torch.float32, device=dev,
                               mg_variant="textbook", tolerance=case["tol"], max_cycles=60,
                               warm_start=warm_start, overlap_guess=overlap, deferred_check=True,
                               temporal_guess=temporal)
    stp.load(case["u"], case["v"])
    stats = []
    for _ in range(warmup):
        stp.step()
        stp.finish()
        s = stp.plan.stats()
        assert s["converged"], s
        stats.append(s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        stp.step()
    stp.finish()  # the last step's convergence read is part of the timed work
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s = stp.plan.stats()
    assert s["converged"] and s["norm"] <= case["tol"], s
    return steps / elapsed, elapsed, stp, stats


def cycles_per_step(aj, stp, n):
    """Untimed: V-cycles and guess use over n further steps."""
    cyc, used, ratio = [], [], []
    for _ in range(n):
        stp.step()
        stp.finish()
        s = stp.plan.stats()
        cyc.append(s["cycles"])
        g = stp.guess_stats if stp.guess_stats is not None else s
        if "used_guess" in g and g.get("zero_norm", 0) > 0 and \
                (stp.warm_start is not None or stp.temporal):
            used.append(g["used_guess"])
            ratio.append(g["guess_norm"] / g["zero_norm"])
    return cyc, used, ratio


def batched_measure(instances, steps, warmup, seed):
    """Batched uniform sweep: `instances` independent 512x128 channels (NACA
    0012 with per-instance AoA / chord / Re), reference multigrid with 3
    V-cycles, one graph-captured batched step per iteration; then a 3-step
    event-timed pass for the per-kernel figures.  Returns the measurements."""
    import torch

    import aerojax_amd as aj
    from aerojax_amd import _lib, workloads

    t0 = time.perf_counter()
    w = workloads.batched_sweep_case(n=instances, seed=seed)
    g = w["grid"]
    b = aj.BatchedStepper([(g.nx, g.ny, instances, g.dx, g.dy)], w["masks"], w["sdfs"],
                          w["nus"], w["u"], w["v"], w["dt"], v_cycles=w["v_cycles"],
                          brinkman_eta=w["brinkman_eta"])
    log(f"built {instances} instances in {time.perf_counter() - t0:.1f}s")
    for _ in range(warmup):
        b.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    st = b.stats()
    assert all(s["cycles"] == w["v_cycles"] and not s["nonfinite"] for s in st)
    _lib.profile_begin(20000)
    for _ in range(3):
        b.step()
    torch.cuda.synchronize()
    _lib.profile_end()
    prof = _lib.profile_read()
    log("kernel profile (batched): " + ", ".join(f"{k} {v['avg_us']:.1f}us x{v['count']} "
                                                   f"{v['gbps']:.0f} GB/s"
                                                   for k, v in prof.items()))
    b.close()
    dom = max(prof, key=lambda k: prof[k]["ms"])

    def roof(k):
        d = prof[k]
        return {"kernel": k, "bound": "hbm", "achieved": d["gbps"], "peak": 8000.0,
                "unit": "GB/s", "frac": d["gbps"] / 8000.0, "avg_launch_us": d["avg_us"],
                "timing": "HIP events on the stream around each launch, 3 steps"}

    # the level-0 smoother with the largest total time (the multigrid's
    # dominant kernel), reported beside the step's dominant kernel
    smk = [k for k in ("mg_pre0", "mg_post0") if k in prof]
    sm = max(smk, key=lambda k: prof[k]["ms"]) if smk else None
    return {"workload": w["name"], "grid": [g.nx, g.ny], "instances": instances,
            "steps": steps, "seconds": el, "instance_steps_per_s": steps * instances / el,
            "cell_updates_per_s": steps * instances * g.nx * g.ny / el,
            "roofline": roof(dom),
            "roofline_smoother": roof(sm) if sm else None,
            "kernels": {k: {"avg_us": round(v["avg_us"], 2), "gbps": round(v["gbps"], 1)}
                        for k, v in prof.items()}}


def run_batched(args):
    """The batched uniform sweep alone; prints one JSON line."""
    r = batched_measure(args.instances, args.steps, args.warmup, args.seed)
    line = {"metric": f"instance-steps/s, {args.instances} x {r['grid'][0]}x{r['grid'][1]} "
                      "batched step (reference multigrid, 3 V-cycles)",
            "value": r["instance_steps_per_s"], "unit": "instance-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * r["seconds"] / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": r["workload"], "instances": args.instances, "grid": r["grid"],
                       "cell_updates_per_s": r["cell_updates_per_s"]},
            "roofline": {**r["roofline"], "traffic": None},
            "roofline_smoother": r["roofline_smoother"], "kernels": r["kernels"]}
    print(json.dumps(line), flush=True)


def train_net(args, dev, nx=2048, ny=512):
    """Warm-start network trained on solver-generated pairs at nx x ny: 10
    pairs per flow (every second step from step 4 to 22) so that the pairs
    come from many flows; a held-out set from other flows picks the best
    epoch."""
    import torch
    from aerojax_amd.nn import WarmStartNet
    from aerojax_amd.warmstart import generate_pairs, normalise_targets, train, validation_loss
    net = WarmStartNet(base_channels=32, levels=4, seed=0, r=args.nn_r, device=dev)
    t0 = time.perf_counter()
    kw = dict(steps_per_seed=24, stride=2, skip=4)
    X, T = generate_pairs(net, args.nn_pairs, range(1000, 1000000), nx, ny, **kw)
    Xv, Tv = generate_pairs(net, args.nn_val_pairs, range(2000000, 3000000), nx, ny, **kw)
    t_gen = time.perf_counter() - t0
    log(f"generated {X.shape[0]} training + {Xv.shape[0]} held-out pairs at "
        f"{X.shape[1]}x{X.shape[2]} in {t_gen:.1f}s")
    pscale = normalise_targets(net, T)
    Tv.div_(pscale)
    t0 = time.perf_counter()
    hist = train(net, X, T, epochs=args.nn_epochs, batch=16, lr=5e-4, log=log, X_val=Xv,
                 T_val=Tv)
    val_best = validation_loss(net, Xv, Tv)
    t_train = time.perf_counter() - t0
    info = {"train_pairs": int(X.shape[0]), "heldout_pairs": int(Xv.shape[0]),
            "pairs_per_flow": 10, "train_epochs": args.nn_epochs, "train_loss": hist["train"],
            "heldout_loss": hist["val"], "best_epoch": hist["best_epoch"],
            "heldout_loss_kept": val_best, "gen_s": round(t_gen, 1),
            "train_s": round(t_train, 1), "target_scale": pscale, "params": net.num_params}
    del X, T, Xv, Tv
    torch.cuda.empty_cache()
    return net, info


def ablation(aj, case, dev, net, steps=40, at=(10, 20, 30, 39)):
    """Untimed: V-cycles to the tolerance and guess/zero residual ratio from
    initial guesses of the same pressure problems (the rhs of a plain
    trajectory at the listed steps), each solved with the guess test off:
      zero        the reference's start
      net         the trained network's guess (if a network is given)
      exact_rK    the converged pressure restricted K times (2x2 average) and
                  prolonged back (bilinear): the best a network running on
                  the grid restricted K times could return
      exact       the converged FP32 pressure itself: the floor of any FP32
                  guess (its rounding leaves a residual of ~1e-5 of ||b||)
      previous    the previous step's pressure
      extrapolated 2 p_prev - p_prevprev
    Returns {guess: {"vcycles": [...], "residual_ratio": [...]}}."""
    import torch
    import torch.nn.functional as F
    g = case["grid"]
    nx, ny = g.nx, g.ny
    stp = aj.CollocatedStepper(nx, ny, g.dx, g.dy, case["mask"], case["sdf"], case["nu"],
                               case["U_inf"], case["dt"], dtype=torch.float32, device=dev,
                               mg_variant="textbook", tolerance=case["tol"], max_cycles=60)
    stp.load(case["u"], case["v"])
    probe = aj.MultigridPlan(nx, ny, g.dx, g.dy, torch.float32, "textbook", 0, dev)
    probe.set_option(guess_test=False)
    out = torch.empty((nx, ny), device=dev)

    def rp(p, r):
        x = p[None, None]
        for _ in range(r):
            x = F.avg_pool2d(x, 2)
        for _ in range(r):
            x = F.interpolate(x, scale_factor=2, mode="bilinear", align_corners=False)
        return x[0, 0].contiguous()

    res = {}

    def run(name, rhs, p0):
        probe.solve(rhs, out, p0, 60, case["tol"])
        st = probe.stats()
        assert st["converged"], (name, st)
        d = res.setdefault(name, {"vcycles": [], "residual_ratio": []})
        d["vcycles"].append(st["cycles"])
        d["residual_ratio"].append(st["norms"][0] / st["zero_norm"] if p0 is not None else 1.0)

    prev = prev2 = None
    for k in range(steps):
        stp.step()
        rhs, p = stp.rhs.clone(), stp.p.clone()
        if k in at:
            run("zero", rhs, None)
            if net is not None:
                run("net", rhs, net.warm_start(rhs, stp.sdf))
            for r in (1, 2, 3):
                run(f"exact_r{r}", rhs, rp(p, r))
            run("exact", rhs, p)
            run("previous", rhs, prev)
            run("extrapolated", rhs, (2.0 * prev - prev2).contiguous())
        prev2, prev = prev, p
    probe.close()
    stp.close()
    for d in res.values():
        d["mean_vcycles"] = sum(d["vcycles"]) / len(d["vcycles"])
    res["steps"] = list(at)
    return res


def gpu_first_step(aj, case, dev):
    """Untimed: one plain step from the initial state; host copies of
    (u, v, p) for the error against the FP64 oracle."""
    import torch
    g = case["grid"]
    stp = aj.CollocatedStepper(g.nx, g.ny, g.dx, g.dy, case["mask"], case["sdf"], case["nu"],
                               case["U_inf"], case["dt"], dtype=torch.float32, device=dev,
                               mg_variant="textbook", tolerance=case["tol"], max_cycles=60)
    stp.load(case["u"], case["v"])
    stp.step()
    u, v = stp.state
    out = tuple(t.double().cpu().numpy() for t in (u, v, stp.p))
    stp.close()
    return out


def cpu_baseline_leg(case, gpu_step, seconds, numpy_steps):
    """CPU baseline and the error field (the only part of the bench that uses
    oracle/, as the checker and the timed CPU restatement):
      * the C++/OpenMP restatement of the same to-tolerance step on all the
        threads OpenMP is given, steps from the initial condition until
        `seconds` have passed;
      * the NumPy restatement (one core) for `numpy_steps` steps;
      * one FP64 oracle step from the initial state (pressure to a relative
        residual of 1e-13) against the GPU's first step."""
    import numpy as np
    from oracle import cpu_step as C
    from oracle import reference_step as R
    from oracle import textbook_mg as TM
    g = case["grid"]
    nx, ny = g.nx, g.ny
    f = np.float32
    mask, sdf = case["mask"].astype(f), case["sdf"].astype(f)
    out = {}
    cs = C.CpuStepper(nx, ny, g.dx, g.dy)
    uu, vv = case["u"].astype(f), case["v"].astype(f)
    n, cyc = 0, []
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and n < 10000:
        uu, vv, _, c, _ = cs.step(uu, vv, mask, sdf, case["nu"], case["dt"], tol=case["tol"])
        if c < 0:
            raise RuntimeError("CPU restatement did not reach the tolerance")
        cyc.append(c)
        n += 1
    el = time.perf_counter() - t0
    cs.close()
    out["cpp"] = {"value": n / el, "steps": n, "seconds": el, "threads": C.threads(),
                  "vcycles_per_step": cyc[:5] + ["..."] + cyc[-3:] if n > 8 else cyc}
    log(f"cpu baseline (C++/OpenMP, {C.threads()} threads): {n / el:.2f} steps/s over {n} steps")
    if numpy_steps > 0:
        mg = TM.TextbookMG(nx, ny, g.dx, g.dy)
        uu, vv = case["u"].astype(f), case["v"].astype(f)
        t0 = time.perf_counter()
        for _ in range(numpy_steps):
            us, vs = R.rk_step_unified(uu, vv, case["dt"], case["nu"], g.dx, g.dy, mask, sdf=sdf,
                                       U_inf=1.0, brinkman_eta=0.005)
            us, vs = R.predictor_bcs(us, vs, f(1.0))
            rhs = R.pressure_rhs(us, vs, mask, case["dt"], g.dx, g.dy)
            p64, _ = mg.solve(rhs, tol=case["tol"])
            uu, vv = R.project(us, vs, p64.astype(f), mask, case["dt"], g.dx, g.dy, 1.0)
        el = time.perf_counter() - t0
        out["numpy"] = {"value": numpy_steps / el, "steps": numpy_steps, "seconds": el,
                        "threads": 1}
        log(f"cpu baseline (NumPy, 1 core): {numpy_steps / el:.3f} steps/s")
    # FP64 oracle step from the same initial state
    m64, s64 = case["mask"].astype(np.float64), case["sdf"].astype(np.float64)
    u0, v0 = case["u"].astype(f).astype(np.float64), case["v"].astype(f).astype(np.float64)
    us, vs = R.rk_step_unified(u0, v0, case["dt"], case["nu"], g.dx, g.dy, m64, sdf=s64,
                               U_inf=1.0, brinkman_eta=0.005)
    us, vs = R.predictor_bcs(us, vs, 1.0)
    rhs = R.pressure_rhs(us, vs, m64, case["dt"], g.dx, g.dy)
    mg = TM.TextbookMG(nx, ny, g.dx, g.dy)
    b = rhs.copy()
    b[-1] = 0.0
    p64, _ = mg.solve(rhs, tol=1e-13 * mg.norm(b), max_outer=40)
    un, vn = R.project(us, vs, p64, m64, case["dt"], g.dx, g.dy, 1.0)
    gu, gv, gp = gpu_step
    out["error"] = {"u": R.rel_l2(gu, un), "v": R.rel_l2(gv, vn), "p": R.rel_l2(gp, p64),
                    "reference": "FP64 oracle step from the same initial state, pressure to a "
                                 "relative residual of 1e-13; GPU step FP32 to ||div|| <= 1e-6"}
    log(f"GPU first step vs FP64 oracle: {out['error']}")
    return out


def run_uneven(args):
    """Batched uneven + learned: 16 instances of each of 256x64, 512x128,
    1024x256 and 2048x512 (obstacle and Re drawn per instance), Smagorinsky
    LES, every pressure solved to ||div|| <= 1e-6 by the batched textbook
    multigrid; plain arm (zero guesses) and NN arm (one network inference
    per group each step).  Prints one JSON line."""
    import torch

    import aerojax_amd as aj
    from aerojax_amd import _lib, workloads

    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    w = workloads.uneven_batch_case(per_size=args.per_size, seed=args.seed)
    n = len(w["masks"])
    cells = sum(nx * ny * c for nx, ny, c, _, _ in w["groups"])
    log(f"workload {w['name']}: {n} instances, {cells} cells, built in "
        f"{time.perf_counter() - t0:.1f}s")

    def arm(net):
        b = aj.BatchedStepper(w["groups"], w["masks"], w["sdfs"], w["nus"], w["u"], w["v"],
                              w["dt"], solver="textbook", tolerance=w["tol"], max_cycles=60,
                              smagorinsky_cs=w["cs"], brinkman_eta=w["brinkman_eta"],
                              warm_start=net, device=dev)
        for _ in range(args.warmup):
            b.step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            b.step()
        torch.cuda.synchronize()
        el = time.perf_counter() - t1
        st = b.stats()
        assert all(s["converged"] and s["norm"] <= w["tol"] for s in st), \
            [(k, s["norm"]) for k, s in enumerate(st) if not s["converged"]]
        return b, el, st

    bp, el_plain, st_plain = arm(None)
    sps_plain = args.steps / el_plain
    cyc = {}
    k = 0
    for nx, ny, c, _, _ in w["groups"]:
        cyc[f"{nx}x{ny}"] = sorted({s["cycles"] for s in st_plain[k:k + c]})
        k += c
    log(f"plain: {sps_plain * n:.0f} instance-steps/s ({1e3 / sps_plain:.2f} ms per batched "
        f"step), V-cycles per group {cyc}")
    _lib.profile_begin(50000)
    for _ in range(3):
        bp.step()
    torch.cuda.synchronize()
    _lib.profile_end()
    prof = _lib.profile_read()
    log("kernel profile (plain): " + ", ".join(f"{k} {v['avg_us']:.1f}us x{v['count']} "
                                                 f"{v['gbps']:.0f} GB/s" for k, v in prof.items()))
    bp.close()
    del bp
    torch.cuda.empty_cache()
    nn_info, sps_nn, el_nn = None, None, None
    if not args.no_nn:
        net, nn_info = train_net(args, dev)
        bn, el_nn, st_nn = arm(net)
        sps_nn = args.steps / el_nn
        used = [bool(s["used_guess"]) for s in st_nn]
        ratio = [s["guess_norm"] / s["zero_norm"] for s in st_nn if s["zero_norm"] > 0]
        nn_info.update(instance_steps_per_s=sps_nn * n, guesses_used=sum(used),
                       guess_over_zero_residual_median=float(sorted(ratio)[len(ratio) // 2])
                       if ratio else None)
        log(f"NN: {sps_nn * n:.0f} instance-steps/s, guesses kept {sum(used)}/{n}")
        bn.close()
    hbm = {k: v for k, v in prof.items() if not k.startswith("nn_") and k != "mg_coarse"}
    dom = max(hbm, key=lambda k: hbm[k]["ms"])
    d = hbm[dom]
    sps = sps_nn if sps_nn is not None else sps_plain
    el = el_nn if el_nn is not None else el_plain
    line = {"metric": f"instance-steps/s, {n} uneven instances "
                      "({256x64,512x128,1024x256,2048x512}) + LES to ||div||<=1e-6",
            "value": sps * n, "unit": "instance-steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": w["name"], "instances": n, "cells": cells,
                       "arm": "nn_warm_start" if sps_nn is not None else "plain_mg",
                       "plain_instance_steps_per_s": sps_plain * n,
                       "plain_cell_updates_per_s": sps_plain * cells,
                       "plain_vcycles_per_group": cyc, "nn": nn_info},
            "roofline": {"kernel": dom, "bound": "hbm", "achieved": d["gbps"], "peak": 8000.0,
                         "unit": "GB/s", "frac": d["gbps"] / 8000.0, "traffic": None},
            "kernels": {k: {"avg_us": round(v["avg_us"], 2), "gbps": round(v["gbps"], 1)}
                        for k, v in prof.items()}}
    print(json.dumps(line), flush=True)


def run_inference(args):
    """The warm start alone: restriction, network inference at 2048x512 / 2^r,
    head and prolongation on seeded 2048x512 fields with a seeded untrained
    network (the launches and their cost do not depend on the weights);
    `warmup` calls (the first one tunes and runs eagerly, the second
    captures the graph), then `steps` timed calls.  Prints one JSON line."""
    import numpy as np
    import torch

    from aerojax_amd.nn import WarmStartNet

    nx, ny = 2048, 512
    rng = np.random.default_rng(args.seed)
    rhs = torch.from_numpy((rng.normal(size=(nx, ny)) * 30).astype(np.float32)).cuda()
    sdf = torch.from_numpy(rng.normal(size=(nx, ny)).astype(np.float32)).cuda()
    net = WarmStartNet(base_channels=32, levels=4, seed=args.seed, r=args.nn_r)
    p0 = torch.empty_like(rhs)
    for _ in range(max(args.warmup, 2)):
        net.warm_start(rhs, sdf, p0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.warm_start(rhs, sdf, p0)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    net.close()
    H, W = nx >> args.nn_r, ny >> args.nn_r
    line = {"metric": f"warm-start inferences/s, 2048x512 fields, network at {H}x{W} (batch 1)",
            "value": args.steps / el, "unit": "inferences/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
            "data": "synthetic", "config": {"workload": f"warm_start_inference_{nx}x{ny}_r{args.nn_r}"}}
    print(json.dumps(line), flush=True)


def run_large(args):
    """4096x1024 cylinder at Re 1e5 with Smagorinsky LES (Cs 0.17), textbook
    multigrid in FP32 inside FP64 defect correction to a relative residual
    <= 1e-8; the FP32 trajectory is compared with the FP64 one from the same
    initial state.  Prints one JSON line."""
    import torch

    import aerojax_amd as aj
    from aerojax_amd import workloads

    c = workloads.large_cylinder()
    g = c["grid"]
    out = {}
    for dtype in (torch.float32, torch.float64):
        stp = aj.CollocatedStepper(g.nx, g.ny, g.dx, g.dy, c["mask"], c["sdf"], c["nu"], c["U_inf"],
                                   c["dt"], dtype=dtype, mg_variant="textbook", tolerance=1e-8,
                                   max_cycles=60, use_les=True, les_coupling="apply",
                                   smagorinsky_constant=c["cs"])
        stp.plan.set_option(relative=True)
        stp.load(c["u"], c["v"])
        for _ in range(args.warmup):
            stp.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cyc = []
        for _ in range(args.steps):
            stp.step()
            cyc.append(stp.plan.stats()["cycles"])
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        s = stp.plan.stats()
        assert s["converged"] and s["norm"] <= 1e-8 * s["norms"][0] * 1.0000001, s
        u, v = stp.state
        out[dtype] = dict(sps=args.steps / el, u=u.double().cpu(), v=v.double().cpu(),
                          p=stp.p.double().cpu(), cycles=cyc)
        log(f"{dtype}: {args.steps / el:.1f} steps/s, V-cycles/step {cyc[:6]}...")
    rel = lambda a, b: float((a - b).norm() / b.norm())
    f32, f64 = out[torch.float32], out[torch.float64]
    err = {k: rel(f32[k], f64[k]) for k in ("u", "v", "p")}
    log(f"FP32 vs FP64 after {args.warmup + args.steps} steps: {err}")
    line = {"metric": "steps/s, 4096x1024 cylinder Re 1e5 + LES, textbook MG to relative "
                      "residual 1e-8 (FP32 cycles, FP64 defect correction)",
            "value": f32["sps"], "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": 1e3 / f32["sps"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "cylinder_4096x1024", "fp64_steps_per_s": f64["sps"],
                       "vcycles_per_step_fp32": f32["cycles"],
                       "rel_l2_fp32_vs_fp64": err,
                       "cell_updates_per_s": f32["sps"] * g.nx * g.ny}}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="time budget of the C++/OpenMP CPU baseline (steps from the initial "
                         "condition)")
    ap.add_argument("--cpu-numpy-steps", type=int, default=3,
                    help="steps of the single-core NumPy restatement reported beside it")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-nn", action="store_true")
    ap.add_argument("--no-ablation", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the 256 x 512x128 batched measurement of the default run")
    ap.add_argument("--nn-pairs", type=int, default=20000)
    ap.add_argument("--nn-val-pairs", type=int, default=1024)
    ap.add_argument("--nn-epochs", type=int, default=3)
    ap.add_argument("--nn-r", type=int, default=2, help="network runs on the grid restricted r times")
    ap.add_argument("--workload", choices=["warmstart", "batched", "large", "uneven", "inference"],
                    default="warmstart",
                    help="warmstart: 2048x512 to ||div|| <= 1e-6, NN vs plain (default); "
                         "batched: 256 x 512x128 reference steps in shared launches; "
                         "large: 4096x1024 + LES to relative residual 1e-8, FP32 vs FP64; "
                         "uneven: 16 each of 256x64..2048x512 + LES to ||div|| <= 1e-6, "
                         "NN vs plain; inference: the warm start alone (2048x512 fields, "
                         "untrained network), for profiling its launches")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--per-size", type=int, default=16, help="uneven: instances per grid size")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 256 MiB device buffer 5 times (reference traffic for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--tune-file", default=None,
                    help="autotuner decisions: preset from this file when it exists (a "
                         "profiled run then launches the shapes a timed run picked), else "
                         "written to it at the end of the run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(single-GPU hot path; the sharding design is documented, not built)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    tune_preset = bool(args.tune_file) and os.path.exists(args.tune_file)
    if tune_preset:
        from aerojax_amd import _lib
        with open(args.tune_file) as fh:
            log(f"autotuner: {_lib.tune_load(fh.read())} decisions preset from {args.tune_file}")
    if args.tune_file and not tune_preset:
        import atexit
        from aerojax_amd import _lib

        def _save_tune():
            with open(args.tune_file, "w") as fh:
                fh.write(_lib.tune_report())
        atexit.register(_save_tune)

    if args.workload == "batched":
        torch.cuda.set_device(0)
        run_batched(args)
        return
    if args.workload == "large":
        torch.cuda.set_device(0)
        run_large(args)
        return
    if args.workload == "uneven":
        torch.cuda.set_device(0)
        run_uneven(args)
        return
    if args.workload == "inference":
        torch.cuda.set_device(0)
        run_inference(args)
        return
    if args.copy_calibration:
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        log("copied 5 x 256 MiB")
        return

    import aerojax_amd as aj
    from aerojax_amd import _lib, workloads

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    case = workloads.warm_start_case(args.seed)
    g = case["grid"]
    nx, ny = g.nx, g.ny
    log(f"workload {case['name']}: {case['desc']}, Re={case['Re']:.0f}")

    # ---- plain multigrid arm (zero initial guess, the reference's start)
    sps_plain, el_plain, stp_plain, _ = run_arm(aj, case, dev, args.steps, args.warmup)
    cyc_plain, _, _ = cycles_per_step(aj, stp_plain, 5)
    log(f"plain MG: {sps_plain:.1f} steps/s, V-cycles/step {cyc_plain}")

    # ---- temporal arm: the plan starts each solve from the extrapolation of
    # the previous FP64 pressures (not the metric's subject; the measured
    # best practical initial guess, reported beside it)
    sps_tmp, _, stp_tmp, _ = run_arm(aj, case, dev, args.steps, args.warmup, temporal=True)
    cyc_tmp, used_tmp, ratio_tmp = cycles_per_step(aj, stp_tmp, 5)
    log(f"temporal guess: {sps_tmp:.1f} steps/s, V-cycles/step {cyc_tmp}, kept {used_tmp}, "
        f"guess/zero residual {['%.4f' % r for r in ratio_tmp]}")
    stp_tmp.close()
    temporal_info = {"steps_per_s": sps_tmp, "over_plain": sps_tmp / sps_plain,
                     "vcycles_per_step": cyc_tmp, "guess_used": used_tmp,
                     "guess_over_zero_residual": ratio_tmp}

    # ---- NN warm-start arm
    nn_info = None
    sps_nn = None
    net = None
    if not args.no_nn:
        net, tinfo = train_net(args, dev, nx, ny)
        # overlapped variant: the zero-guess solve runs while the network and its
        # guess test run on a second stream (reported, not the headline)
        sps_ov, _, stp_ov, _ = run_arm(aj, case, dev, args.steps, args.warmup, warm_start=net,
                                       overlap=True)
        log(f"NN warm start, overlapped: {sps_ov:.1f} steps/s")
        stp_ov.close()
        del stp_ov
        # sequential (the stepper's default): inference, guess test, then the solve
        # from the kept guess or from zero
        sps_nn, el_nn, stp_nn, _ = run_arm(aj, case, dev, args.steps, args.warmup, warm_start=net)
        cyc_nn, used, ratio = cycles_per_step(aj, stp_nn, 5)
        log(f"NN warm start: {sps_nn:.1f} steps/s, V-cycles/step {cyc_nn}, guess used {used}, "
            f"guess/zero residual {['%.2f' % r for r in ratio]}")
        nn_info = {"steps_per_s": sps_nn, "mode": "sequential",
                   "overlapped_steps_per_s": sps_ov, "vcycles_per_step": cyc_nn,
                   "guess_used": used, "guess_over_zero_residual": ratio,
                   "net_resolution": [nx >> args.nn_r, ny >> args.nn_r], **tinfo}

    # ---- ablation of initial guesses (untimed)
    abl = None
    if not args.no_ablation:
        abl = ablation(aj, case, dev, net)
        log("ablation (mean V-cycles to the tolerance, mean guess/zero residual): " + ", ".join(
            f"{k} {v['mean_vcycles']:.1f} {sum(v['residual_ratio']) / len(v['residual_ratio']):.3g}"
            for k, v in abl.items() if k != "steps"))

    # ---- live per-kernel timing (HIP event pairs on the launching stream):
    # the stencil / multigrid kernels from the plain arm, the network's from
    # the NN arm
    nprof = 10
    repeat = 10

    def timed(stp):
        """Two event-timed passes of nprof steps: launches per step and total
        time from single launches; per-launch time from a pass in which every
        tunable kernel (outputs independent of what it writes) runs `repeat`
        times back to back inside its event pair, which leaves out the
        dispatch gap an event pair around one short launch includes."""
        out = {}
        for r in (1, repeat):
            _lib.profile_begin(20000, r)
            for _ in range(nprof):
                stp.step()
            stp.finish()
            torch.cuda.synchronize()
            _lib.profile_end()
            out[r] = _lib.profile_read()
        _lib.profile_begin(1, 1)  # back to single launches
        _lib.profile_end()
        prof = {}
        for k, v in out[1].items():
            b = out[repeat].get(k, v)
            per_launch = b["bytes"] / b["count"]
            prof[k] = {"count": v["count"], "ms": v["ms"], "bytes": per_launch * v["count"],
                       "avg_us": b["avg_us"], "gbps": b["gbps"], "single_launch_us": v["avg_us"]}
        return prof

    prof = timed(stp_plain)
    if sps_nn is not None:
        prof.update({k: v for k, v in timed(stp_nn).items() if k.startswith("nn_")})
    # roofline kernel: the single HBM kernel with the largest total time
    # ("mg_coarse" times a group of three small-level kernels, "nn_*" are MFMA)
    hbm = {k: v for k, v in prof.items() if not k.startswith("nn_") and k != "mg_coarse"}
    dom = max(hbm, key=lambda k: hbm[k]["ms"])
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "hbm_traffic_2048x512.json")
    if os.path.exists(tfile) and (nx, ny) == (2048, 512):
        with open(tfile) as fh:
            traffic = json.load(fh).get("kernels", {}).get(dom, {}).get("hbm_bytes_per_launch")
    d = hbm[dom]
    log("kernel profile: " + ", ".join(f"{k} {v['avg_us']:.1f}us x{v['count']}"
                                         for k, v in prof.items()))
    log("autotuner:\n" + _lib.tune_report())
    mfma = None
    if "nn_conv" in prof:
        c = prof["nn_conv"]
        mfma = {"kernel": "k_conv3x3 (inference)", "achieved_tflops": c["gbps"] / 1e3,
                "peak_tflops": 2500.0, "frac": c["gbps"] / 1e3 / 2500.0,
                "avg_launch_us": c["avg_us"]}

    # ---- the batched sweep (256 x 512x128, 16.8 M cells: beyond the 256 MiB
    # Infinity Cache, so HBM-bound), for its dominant kernel's roofline
    c4 = None
    if not args.no_batched:
        stp_plain.close()
        if sps_nn is not None:
            stp_nn.close()
        torch.cuda.empty_cache()
        c4 = batched_measure(256, 10, 3, args.seed)
        log(f"batched 256 x 512x128: {c4['instance_steps_per_s']:.0f} instance-steps/s, "
            f"dominant kernel {c4['roofline']['kernel']} at {c4['roofline']['frac']:.3f} of HBM peak")

    # ---- CPU baseline (C++/OpenMP restatement on the host's cores, NumPy on
    # one core) and the first step's error against the FP64 oracle
    gstep = gpu_first_step(aj, case, dev)
    cpu = None
    err = None
    if not args.no_cpu:
        leg = cpu_baseline_leg(case, gstep, args.cpu_seconds, args.cpu_numpy_steps)
        cc = leg["cpp"]
        cpu = {"value": cc["value"], "unit": "steps/s", "cores": cc["threads"], "kind": "port",
               "sample": f"{cc['steps']} steps ({cc['seconds']:.1f} s) from the initial condition "
                         "of the same 2048x512 to-tolerance step (FP32 fields, FP64 defect "
                         "correction, plain multigrid), CPU restatement of the reference algorithm "
                         f"in C++/OpenMP on {cc['threads']} threads "
                         f"(os.sched_getaffinity: {len(os.sched_getaffinity(0))} cores)",
               "numpy_one_core": leg.get("numpy")}
        err = leg["error"]

    value = sps_nn if sps_nn is not None else sps_plain
    elapsed = el_nn if sps_nn is not None else el_plain
    line = {
        "metric": "steps/s at 2048x512 to ||div||<=1e-6: NN-warm-start vs plain MG; HBM GB/s",
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": case["name"], "geometry": case["desc"], "Re": round(case["Re"], 1),
                   "grid": [nx, ny], "tolerance": case["tol"],
                   "arm": "nn_warm_start" if sps_nn is not None else "plain_mg",
                   "cell_updates_per_s": value * nx * ny,
                   "plain_mg_steps_per_s": sps_plain, "plain_vcycles_per_step": cyc_plain,
                   "nn_over_plain": (sps_nn / sps_plain) if sps_nn else None, "nn": nn_info,
                   "temporal_guess": temporal_info, "guess_ablation": abl},
        "error_vs_fp64_oracle": err,
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": d["gbps"], "peak": 8000.0,
                     "unit": "GB/s", "frac": d["gbps"] / 8000.0, "traffic": traffic,
                     "traffic_source": "profiles/hbm_traffic_2048x512.json: rocprofv3 --pmc "
                                       "FETCH_SIZE / WRITE_SIZE of the same kernel, scaled by a "
                                       "copy calibration; read from that file, not counted in "
                                       "this run" if traffic is not None else None,
                     "avg_launch_us": d["avg_us"], "bytes_per_launch": d["bytes"] / d["count"],
                     "single_launch_event_us": d["single_launch_us"],
                     "timing": f"HIP events on the plan's stream around {repeat} back-to-back "
                               f"launches per instrumented launch, {nprof} steps"},
        "mfma": mfma,
        "kernels": {k: {"avg_us": round(v["avg_us"], 2), "launches_per_step": v["count"] / nprof,
                        ("tflops" if k.startswith("nn_") else "gbps"):
                            round(v["gbps"] / (1e3 if k.startswith("nn_") else 1), 2)}
                    for k, v in prof.items()},
        "cpu_baseline": cpu,
        "batched_256x512x128": c4,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
