"""Pins the torch-CPU oracle of the warm-start network: central finite
differences of the FP64 loss agree with its autograd gradient at random
coordinates of every layer, and the parameter layout round-trips."""
import numpy as np
import torch

from oracle import nn_reference as NR


def _setup(seed=0, B=1, H=8, W=8, c0=32, levels=2):
    rng = np.random.default_rng(seed)
    shapes = NR.layer_shapes(c0, levels)
    n = sum(co * ci * 9 + co for co, ci in shapes) + c0 + 1
    flat = rng.normal(scale=0.1, size=n)
    x = rng.normal(size=(B, H, W, 2))
    t = rng.normal(size=(B, H, W))
    return flat, x, t, shapes


def test_layout_roundtrip():
    flat, _, _, _ = _setup()
    convs, hw, hb = NR.unflatten(flat, 32, 2)
    assert np.array_equal(NR.flatten(convs, hw, hb).numpy(), flat)


def test_autograd_matches_finite_differences():
    flat, x, t, shapes = _setup()
    _, g, _ = NR.loss_and_grads(flat, x, t, c0=32, levels=2, bf16=False)
    rng = np.random.default_rng(1)
    off = 0
    picks = []
    for co, ci in shapes:
        picks += list(off + rng.integers(0, co * ci * 9, size=2))   # weights
        picks.append(off + co * ci * 9 + int(rng.integers(0, co)))  # a bias
        off += co * ci * 9 + co
    picks += [off, off + 32]  # head weight, head bias
    h = 1e-6
    for k in picks:
        fp = flat.copy()
        fm = flat.copy()
        fp[k] += h
        fm[k] -= h
        lp, _, _ = NR.loss_and_grads(fp, x, t, c0=32, levels=2, bf16=False)
        lm, _, _ = NR.loss_and_grads(fm, x, t, c0=32, levels=2, bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_adam_first_step_closed_form():
    # step 1 with bias correction: p - lr * g / (|g| + eps) (after clipping)
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([3.0, -0.25, 0.0])
    out = NR.adam_step(p, g, 0.01)
    gc = np.clip(g, -1, 1)
    assert np.allclose(out, p - 0.01 * gc / (np.abs(gc) + 1e-8), rtol=1e-12, atol=1e-12)
