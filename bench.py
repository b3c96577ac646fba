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


def run_arm(aj, case, dev, steps, warmup, warm_start=None, overlap=False):
    """Warmup then `steps` timed steps; returns (steps/s, stepper, per-step stats)."""
    import torch
    g = case["grid"]
    stp = aj.CollocatedStepper(g.nx, g.ny, g.dx, g.dy, case["mask"], case["sdf"], case["nu"],
                               case["U_inf"], case["dt"], dtype=torch.float32, device=dev,
                               mg_variant="textbook", tolerance=case["tol"], max_cycles=60,
                               warm_start=warm_start, overlap_guess=overlap, deferred_check=True)
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
        if "used_guess" in g and g.get("zero_norm", 0) > 0 and stp.warm_start is not None:
            used.append(g["used_guess"])
            ratio.append(g["guess_norm"] / g["zero_norm"])
    return cyc, used, ratio


def run_batched(args):
    """Batched uniform sweep: 256 independent 512x128 channels (NACA 0012 with
    per-instance AoA / chord / Re), reference multigrid with 3 V-cycles,
    one graph-captured batched step per iteration.  Prints one JSON line."""
    import torch

    import aerojax_amd as aj
    from aerojax_amd import _lib, workloads

    t0 = time.perf_counter()
    w = workloads.batched_sweep_case(n=args.instances, seed=args.seed)
    g = w["grid"]
    b = aj.BatchedStepper([(g.nx, g.ny, args.instances, g.dx, g.dy)], w["masks"], w["sdfs"],
                          w["nus"], w["u"], w["v"], w["dt"], v_cycles=w["v_cycles"],
                          brinkman_eta=w["brinkman_eta"])
    log(f"built {args.instances} instances in {time.perf_counter() - t0:.1f}s")
    for _ in range(args.warmup):
        b.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
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
    log("kernel profile: " + ", ".join(f"{k} {v['avg_us']:.1f}us x{v['count']} "
                                         f"{v['gbps']:.0f} GB/s" for k, v in prof.items()))
    dom = max(prof, key=lambda k: prof[k]["ms"])
    d = prof[dom]
    sps = args.steps / el
    line = {"metric": f"instance-steps/s, {args.instances} x {g.nx}x{g.ny} batched step "
                      "(reference multigrid, 3 V-cycles)",
            "value": sps * args.instances, "unit": "instance-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": w["name"], "instances": args.instances, "grid": [g.nx, g.ny],
                       "cell_updates_per_s": sps * args.instances * g.nx * g.ny},
            "roofline": {"kernel": dom, "bound": "hbm", "achieved": d["gbps"], "peak": 8000.0,
                         "unit": "GB/s", "frac": d["gbps"] / 8000.0, "traffic": None},
            "kernels": {k: {"avg_us": round(v["avg_us"], 2), "gbps": round(v["gbps"], 1)}
                        for k, v in prof.items()}}
    print(json.dumps(line), flush=True)


def train_net(args, dev, nx=2048, ny=512):
    """Warm-start network trained on solver-generated pairs at nx x ny."""
    import torch
    from aerojax_amd.nn import WarmStartNet
    from aerojax_amd.warmstart import generate_pairs, normalise_targets, train
    net = WarmStartNet(base_channels=32, levels=4, seed=0, r=args.nn_r, device=dev)
    t0 = time.perf_counter()
    X, T = generate_pairs(net, args.nn_pairs, range(1000, 100000), nx, ny,
                          steps_per_seed=40, stride=2, skip=2)
    t_gen = time.perf_counter() - t0
    log(f"generated {X.shape[0]} pairs at {X.shape[1]}x{X.shape[2]} in {t_gen:.1f}s")
    pscale = normalise_targets(net, T)
    t0 = time.perf_counter()
    hist = train(net, X, T, epochs=args.nn_epochs, batch=16, lr=5e-4, log=log)
    t_train = time.perf_counter() - t0
    del X, T
    torch.cuda.empty_cache()
    info = {"train_pairs": args.nn_pairs, "train_epochs": args.nn_epochs, "train_loss": hist,
            "gen_s": round(t_gen, 1), "train_s": round(t_train, 1), "target_scale": pscale,
            "params": net.num_params}
    return net, info


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
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-nn", action="store_true")
    ap.add_argument("--nn-pairs", type=int, default=20000)
    ap.add_argument("--nn-epochs", type=int, default=1)
    ap.add_argument("--nn-r", type=int, default=2, help="network runs on the grid restricted r times")
    ap.add_argument("--workload", choices=["warmstart", "batched", "large", "uneven"],
                    default="warmstart",
                    help="warmstart: 2048x512 to ||div|| <= 1e-6, NN vs plain (default); "
                         "batched: 256 x 512x128 reference steps in shared launches; "
                         "large: 4096x1024 + LES to relative residual 1e-8, FP32 vs FP64; "
                         "uneven: 16 each of 256x64..2048x512 + LES to ||div|| <= 1e-6, "
                         "NN vs plain")
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

    # ---- plain multigrid arm
    sps_plain, el_plain, stp_plain, _ = run_arm(aj, case, dev, args.steps, args.warmup)
    cyc_plain, _, _ = cycles_per_step(aj, stp_plain, 5)
    log(f"plain MG: {sps_plain:.1f} steps/s, V-cycles/step {cyc_plain}")

    # ---- NN warm-start arm
    nn_info = None
    sps_nn = None
    if not args.no_nn:
        net, tinfo = train_net(args, dev, nx, ny)
        # overlapped variant: the zero-guess solve runs while the network and its
        # guess test run on a second stream (reported, not the headline)
        sps_ov, _, stp_ov, _ = run_arm(aj, case, dev, args.steps, args.warmup, warm_start=net,
                                       overlap=True)
        log(f"NN warm start, overlapped: {sps_ov:.1f} steps/s")
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

    # ---- CPU baseline: NumPy restatement of the same step, bounded sample
    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        from oracle import reference_step as R
        from oracle import textbook_mg as TM
        f = np.float32
        mg = TM.TextbookMG(nx, ny, g.dx, g.dy)
        uu, vv = case["u"].astype(f), case["v"].astype(f)
        mask, sdf = case["mask"].astype(f), case["sdf"].astype(f)
        t0 = time.perf_counter()
        for _ in range(args.cpu_steps):
            us, vs = R.rk_step_unified(uu, vv, case["dt"], case["nu"], g.dx, g.dy, mask, sdf=sdf,
                                       U_inf=1.0, brinkman_eta=0.005)
            us, vs = R.predictor_bcs(us, vs, f(1.0))
            rhs = R.pressure_rhs(us, vs, mask, case["dt"], g.dx, g.dy)
            p64, _ = mg.solve(rhs, tol=case["tol"])
            uu, vv = R.project(us, vs, p64.astype(f), mask, case["dt"], g.dx, g.dy, 1.0)
        cpu_el = time.perf_counter() - t0
        cpu = {"value": args.cpu_steps / cpu_el, "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the NumPy restatement (FP32 fields, FP64 "
                         "defect correction, plain multigrid) of the same 2048x512 "
                         "to-tolerance step, first steps from the initial condition"}
        log(f"cpu baseline: {cpu['value']:.3f} steps/s")

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
                   "nn_over_plain": (sps_nn / sps_plain) if sps_nn else None, "nn": nn_info},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": d["gbps"], "peak": 8000.0,
                     "unit": "GB/s", "frac": d["gbps"] / 8000.0, "traffic": traffic,
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
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
