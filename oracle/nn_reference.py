"""CPU oracle of the warm-start network (torch on CPU, float64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The same U-Net as the HIP implementation, written with library routines
(``torch.nn.functional.conv2d``, exact GELU, ``avg_pool2d``, nearest
``interpolate``, ``cat``) so that forward values, autograd gradients and
``torch.optim.Adam`` pin the project's own BF16-MFMA forward, backward and
Adam kernels.  Parameters use the flat layout of ``aj_nn_get_params``:
for each conv (encoder levels 0..L-1: a, b; decoder levels L-2..0: a, b)
weight [co][ci][3][3] then bias [co]; then the 1x1 head weight [c0] and
bias [1].  ``bf16=True`` rounds weights and every stored activation to
bfloat16 the way the GPU stores them (accumulation stays float64).
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F


def layer_shapes(c0=32, levels=4, cin=2):
    ch = [c0 << l for l in range(levels)]
    shapes = []
    for l in range(levels):
        shapes.append((ch[l], cin if l == 0 else ch[l - 1]))
        shapes.append((ch[l], ch[l]))
    for l in range(levels - 2, -1, -1):
        shapes.append((ch[l], ch[l + 1] + ch[l]))
        shapes.append((ch[l], ch[l]))
    return shapes


def unflatten(flat, c0=32, levels=4):
    flat = torch.as_tensor(np.asarray(flat, dtype=np.float64))
    off = 0
    convs = []
    for co, ci in layer_shapes(c0, levels):
        w = flat[off:off + co * ci * 9].reshape(co, ci, 3, 3)
        off += co * ci * 9
        b = flat[off:off + co]
        off += co
        convs.append([w, b])
    hw = flat[off:off + c0]
    off += c0
    hb = flat[off:off + 1]
    off += 1
    assert off == flat.numel()
    return convs, hw, hb


def flatten(convs, hw, hb):
    parts = []
    for w, b in convs:
        parts += [w.reshape(-1), b.reshape(-1)]
    parts += [hw.reshape(-1), hb.reshape(-1)]
    return torch.cat(parts)


def to_bf16(t):
    return t.to(torch.bfloat16).to(t.dtype)


def forward(x, convs, hw, hb, levels=4, bf16=True):
    """x: [B, H, W, 2] -> out [B, H, W]."""
    r = to_bf16 if bf16 else (lambda t: t)
    h = r(x.permute(0, 3, 1, 2))

    def conv(i, inp):
        w, b = convs[i]
        pre = F.conv2d(inp, r(w), b, padding=1)
        return r(F.gelu(pre)), pre

    skips = []
    for l in range(levels):
        if l > 0:
            h = r(F.avg_pool2d(skips[-1], 2))
        a, _ = conv(2 * l, h)
        h, _ = conv(2 * l + 1, a)
        if l < levels - 1:
            skips.append(h)
    k = 2 * levels
    for l in range(levels - 2, -1, -1):
        up = F.interpolate(h, scale_factor=2, mode="nearest")
        cat = torch.cat([up, skips[l]], dim=1)
        a, _ = conv(k, cat)
        h, _ = conv(k + 1, a)
        k += 2
    out = torch.einsum("bchw,c->bhw", h, hw) + hb[0]
    return out


def loss_and_grads(flat, x, t, c0=32, levels=4, bf16=True):
    """MSE loss and its gradient w.r.t. the flat parameters (float64)."""
    convs, hw, hb = unflatten(flat, c0, levels)
    leaves = []
    for pair in convs:
        for k in range(2):
            pair[k] = pair[k].clone().requires_grad_(True)
            leaves.append(pair[k])
    hw = hw.clone().requires_grad_(True)
    hb = hb.clone().requires_grad_(True)
    out = forward(torch.as_tensor(x, dtype=torch.float64), convs, hw, hb, levels, bf16)
    loss = ((out - torch.as_tensor(t, dtype=torch.float64)) ** 2).mean()
    loss.backward()
    g = flatten([(w.grad, b.grad) for w, b in convs], hw.grad, hb.grad)
    return float(loss), g.numpy(), out.detach().numpy()


def adam_step(flat, grads, lr, clip=1.0):
    """One torch.optim.Adam step (betas 0.9/0.999, eps 1e-8) on clipped grads."""
    p = torch.nn.Parameter(torch.as_tensor(np.asarray(flat, np.float64)).clone())
    opt = torch.optim.Adam([p], lr=lr, betas=(0.9, 0.999), eps=1e-8)
    p.grad = torch.as_tensor(np.clip(np.asarray(grads, np.float64), -clip, clip))
    opt.step()
    return p.detach().numpy()
