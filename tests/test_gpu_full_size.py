"""Full-size launches of the batched and warm-start paths: the 256-instance
batched step at 512x128 (every instance in shared persistent launches, a
sample of instances checked against the oracle's step_pure) and the warm
start at 2048x512 (the network on the grid restricted twice, 512x128)."""
import numpy as np
import pytest
import torch

from oracle import reference_step as R
from oracle import textbook_mg as T

pytestmark = pytest.mark.gpu


def test_batched_step_256_instances(gpu):
    """One batched step of `batched_sweep_case(256)` (reference multigrid, 3
    V-cycles): instances spread over the batch (first, last, and ones in
    between) equal the oracle's step_pure of their own FP32-rounded fields
    within the FP32 bar, 1e-5 relative L2."""
    import aerojax_amd as aj
    from aerojax_amd import workloads
    w = workloads.batched_sweep_case(n=256, seed=2)
    g = w["grid"]
    b = aj.BatchedStepper([(g.nx, g.ny, 256, g.dx, g.dy)], w["masks"], w["sdfs"], w["nus"],
                          w["u"], w["v"], w["dt"], v_cycles=3)
    b.step()
    torch.cuda.synchronize()
    for k in (0, 1, 77, 128, 200, 255):
        u, v, p = R.step_pure(w["u"][k].astype(np.float32).astype(np.float64),
                              w["v"][k].astype(np.float32).astype(np.float64),
                              w["masks"][k].astype(np.float32).astype(np.float64),
                              w["sdfs"][k].astype(np.float32).astype(np.float64),
                              g.dx, g.dy, w["nus"][k], 1.0, w["dt"], v_cycles=3)
        for name, got, ref in (("u", b.view(b.u, k), u), ("v", b.view(b.v, k), v),
                               ("p", b.view(b.p, k), p)):
            e = R.rel_l2(got.detach().double().cpu().numpy(), ref)
            print(f"256 x 512x128 instance {k} {name}: rel_l2 vs FP64 oracle {e:.2e}", flush=True)
            assert e <= 1e-5, (k, name, e)
    assert all(s["cycles"] == 3 for s in b.stats())


def test_warm_start_full_size(gpu):
    """The warm start of the 2048x512 workload: inputs restricted twice and
    normalised as the oracle does, and the guess equal to the oracle's
    multigrid prolongation of the network output at 512x128 (times the RMS
    scale), zero on the outlet row."""
    import aerojax_amd as aj
    from aerojax_amd import workloads
    from aerojax_amd.nn import WarmStartNet
    c = workloads.warm_start_case(0)
    g = c["grid"]
    nx, ny, r = g.nx, g.ny, 2
    stp = aj.CollocatedStepper(nx, ny, g.dx, g.dy, c["mask"], c["sdf"], c["nu"], 1.0, c["dt"],
                               mg_variant="textbook", tolerance=1e-6, max_cycles=60)
    stp.load(c["u"], c["v"])
    for _ in range(3):
        stp.step()
    rhs = stp.rhs.clone()
    sdf = torch.from_numpy(c["sdf"].astype(np.float32)).to(gpu)
    net = WarmStartNet(base_channels=32, levels=4, seed=5, r=r, device=gpu)
    H, W = nx >> r, ny >> r
    xs = torch.empty((H, W, 2), device=gpu)
    ts = torch.empty((H, W), device=gpu)
    net.make_sample(rhs, sdf, rhs, xs, ts)

    def restrict(f, k):
        for _ in range(k):
            f = 0.25 * ((f[0::2, 0::2] + f[1::2, 0::2]) + (f[0::2, 1::2] + f[1::2, 1::2]))
        return f

    rh = rhs.double().cpu().numpy()
    rc = restrict(rh, r)
    s = np.sqrt(np.mean(rc ** 2))
    assert np.allclose(xs[..., 0].cpu().numpy(), rc / s, rtol=1e-4, atol=1e-5)
    sd = restrict(c["sdf"].astype(np.float32).astype(np.float64), r)
    assert np.allclose(xs[..., 1].cpu().numpy(), np.tanh(sd / 0.1), atol=1e-5)
    p0 = net.warm_start(rhs, sdf).cpu().numpy()
    assert p0.shape == (nx, ny) and np.all(np.isfinite(p0)) and np.all(p0[-1] == 0)
    out = net.forward(xs[None].contiguous())[0].cpu().numpy().astype(np.float64) * s
    e = out
    for lev in range(r - 1, -1, -1):
        e = T.prolong(e, lev, (e.shape[0] * 2, e.shape[1] * 2))
    err = np.abs(p0 - e).max() / np.abs(e).max()
    print(f"2048x512 warm start vs prolonged network output: max rel {err:.2e}", flush=True)
    assert np.allclose(p0, e, rtol=1e-4, atol=1e-4 * np.abs(e).max())
    net.close()
    stp.close()
