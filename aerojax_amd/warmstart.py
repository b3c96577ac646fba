"""Solver-generated training data and training of the warm-start network.

Training pairs are (rhs, sdf) -> p taken from the project's own to-tolerance
solver (textbook multigrid, ||div|| <= 1e-6) while it advances flows with
randomly drawn obstacles and Reynolds numbers, restricted to the network's
resolution and normalised (see ``WarmStartNet.make_sample``).  All of it
runs on the GPU; the host only draws the geometry of each flow.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import workloads
from .geometry import device_geometry
from .nn import WarmStartNet
from .step import CollocatedStepper


def generate_pairs(net: WarmStartNet, n_pairs: int, seeds, nx=2048, ny=512, steps_per_seed=40,
                   stride=2, skip=2, log=None, noise_sigma=1e-3):
    """Advance one flow per seed with the plain solver and record every
    ``stride``-th step after ``skip`` steps.  Each seed draws its obstacle and
    Reynolds number as ``workloads.warm_start_case``; the geometry is built on
    the GPU and the inflow noise is drawn there (torch generator seeded with
    the seed).  Returns device tensors x [N, H, W, 2] and t [N, H, W]
    (float32)."""
    dev = net.device
    H, W = nx >> net.r, ny >> net.r
    X = torch.empty((n_pairs, H, W, 2), dtype=torch.float32, device=dev)
    T = torch.empty((n_pairs, H, W), dtype=torch.float32, device=dev)
    k = 0
    t0 = time.perf_counter()
    gen = torch.Generator(device=dev)
    for seed in seeds:
        if k >= n_pairs:
            break
        g, ob, Re, nu = workloads.warm_start_params(int(seed), nx, ny)
        sdf, mask = device_geometry(g, workloads.obstacle_struct(ob), torch.float32, dev)
        sdf = torch.nan_to_num(sdf, nan=0.0)
        gen.manual_seed(int(seed))
        u = 1.0 + noise_sigma * torch.randn((nx, ny), generator=gen, device=dev)
        v = noise_sigma * torch.randn((nx, ny), generator=gen, device=dev)
        stp = CollocatedStepper(nx, ny, g.dx, g.dy, mask, sdf, nu, 1.0, 2.5e-4, device=dev,
                                mg_variant="textbook", tolerance=1e-6, max_cycles=60)
        stp.load(u, v)
        case = {"desc": ob["desc"], "Re": Re}
        for s in range(steps_per_seed):
            stp.step()
            if s >= skip and (s - skip) % stride == 0 and k < n_pairs:
                st = stp.plan.stats()
                if not st["converged"] or st["nonfinite"]:
                    break  # the flow left the scheme's stability range: drop the seed
                net.make_sample(stp.rhs, stp.sdf, stp.p, X[k], T[k])
                k += 1
        if log:
            log(f"  seed {seed}: {case['desc']}, Re={case['Re']:.0f}; {k}/{n_pairs} pairs, "
                f"{time.perf_counter() - t0:.1f}s")
    torch.cuda.synchronize(dev)
    return X[:k], T[:k]


def normalise_targets(net: WarmStartNet, T):
    """Scale the targets to unit RMS in place and fold the factor into the
    network's output scale (``pscale``), so that the regression target is
    O(1) whatever the pressure's magnitude.  Returns the factor."""
    q = float(torch.sqrt(torch.mean(T.double() ** 2)).item())
    if not math.isfinite(q) or q <= 0.0:
        raise ValueError(f"training targets have RMS {q}")
    T.div_(q)
    net.pscale *= q
    return q


def validation_loss(net: WarmStartNet, X, T, batch=16) -> float:
    """Mean squared error of the network on (X, T) over whole batches of the
    training batch size (the network's activations stay allocated for that
    size); one host sync."""
    n = (X.shape[0] // batch) * batch
    if n == 0:
        raise ValueError(f"held-out set smaller than one batch of {batch}")
    acc = torch.zeros((), dtype=torch.float64, device=net.device)
    for b in range(0, n, batch):
        out = net.forward(X[b:b + batch])
        acc += ((out - T[b:b + batch]).double() ** 2).sum()
    return float(acc.item()) / float(n * T[0].numel())


def train(net: WarmStartNet, X, T, epochs=2, batch=16, lr=5e-4, lr_min=1e-5, seed=0, log=None,
          X_val=None, T_val=None):
    """Adam on shuffled mini-batches with a cosine learning-rate schedule.

    With a held-out set (X_val, T_val), its loss is evaluated after every
    epoch and the parameters of the best epoch are kept at the end (the
    reference keeps its best checkpoint, neural_operators/training_utils.py:
    195-248).  Returns {"train": per-epoch mean training loss, "val": per-epoch
    held-out loss (or []), "best_epoch": epoch whose parameters are kept}."""
    n = X.shape[0]
    steps = max(1, (n // batch) * epochs)
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist, val = [], []
    best, best_ep, best_params = math.inf, epochs - 1, None
    if X_val is not None:
        val0 = validation_loss(net, X_val, T_val)
        if log:
            log(f"  held-out loss before training: {val0:.4e}")
    it = 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(torch.int32).to(net.device)
        acc = torch.zeros((), dtype=torch.float32, device=net.device)
        nb = n // batch
        for b in range(nb):
            idx = perm[b * batch:(b + 1) * batch]
            cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * it / steps))
            acc += net.train_step(X, T, idx, cur)[0]
            it += 1
        hist.append(float(acc.item()) / max(nb, 1))
        msg = f"  epoch {ep}: loss {hist[-1]:.4e}"
        if X_val is not None:
            v = validation_loss(net, X_val, T_val)
            val.append(v)
            msg += f", held-out {v:.4e}"
            if v < best:
                best, best_ep, best_params = v, ep, net.get_params()
        if log:
            log(msg + f" ({time.perf_counter() - t0:.1f}s)")
    if best_params is not None and best_ep != epochs - 1:
        net.set_params(best_params)
    return {"train": hist, "val": val, "best_epoch": best_ep}
