"""Every launch candidate of the tuned kernels gives the same bits.

The autotuner picks among LDS tiles and register-column kernels per shape;
the tests force each candidate in turn (decisions preset before the first
launch) and compare with candidate 0, at shapes that exercise several
column strips, tiles touching the inlet and outlet rows, and narrow grids.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import reference_step as R

pytestmark = pytest.mark.gpu


def _case(nx, ny):
    dx, dy, X, Y = R.grid_coords(nx, ny, 4.0, 1.0)
    sdf = R.cylinder_sdf(X, Y, 1.0, 0.5, 0.1)
    mask = R.mask_from_sdf(sdf, 0.05)
    u, v = R.von_karman_ic(nx, ny, 1.0, 1e-3, seed=0, dtype=np.float64)
    return dict(u=u, v=v, mask=mask, sdf=sdf, dx=dx, dy=dy)


@pytest.mark.parametrize("shape", [(512, 96), (2048, 512), (300, 260), (130, 70)])
@pytest.mark.parametrize("variant", ["sdf", "fast_nsgs", "nosdf_noslip"])
def test_rk_candidates_bit_identical(gpu, shape, variant):
    from aerojax_amd import _lib
    L = _lib.load()
    nx, ny = shape
    c = _case(nx, ny)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu, torch.float32)  # noqa: E731
    u, v, mask, sdf = t(c["u"]), t(c["v"]), t(c["mask"]), t(c["sdf"])
    nsgs = None
    if variant == "fast_nsgs":
        nsgs = t(50.0 * R.constant_smagorinsky(c["u"], c["v"], c["dx"], c["dy"],
                                               (c["dx"] * c["dy"]) ** 0.5, 0.17))
    if variant == "nosdf_noslip":
        sdf = None
    fast = variant == "fast_nsgs"
    prm = _lib.RkParams(1e-3, 5e-4, c["dx"], c["dy"], 1.0, 0.005, 0.0,
                        0 if variant == "nosdf_noslip" else 1, 1 if fast else 0)
    s = torch.cuda.current_stream().cuda_stream
    p = lambda x: None if x is None else x.data_ptr()  # noqa: E731
    scheme = "rk2" if fast else "rk3"
    for key, ncand in (("rk", 6), ("rkdiv", 5)):
        outs = {}
        for cand in range(ncand):
            L.aj_tune_reset()
            assert _lib.tune_load(f"{key}/f32/{scheme}/{nx}x{ny} -> {cand}\n") == 1
            us, vs, rr = torch.zeros_like(u), torch.zeros_like(u), torch.zeros_like(u)
            if key == "rkdiv":
                rc = L.aj_rk_step_rhs(_lib.AJ_F32, p(u), p(v), p(mask), p(sdf), p(nsgs), p(us),
                                      p(vs), p(rr), 1, nx, ny, ctypes.byref(prm), s)
            else:
                rc = L.aj_rk_step(_lib.AJ_F32, p(u), p(v), p(mask), p(sdf), p(nsgs), p(us), p(vs),
                                  nx, ny, ctypes.byref(prm), s)
            torch.cuda.synchronize()
            assert rc == 0, (key, cand, L.aj_last_error())
            outs[cand] = (us, vs, rr)
        L.aj_tune_reset()
        for cand, o in outs.items():
            for name, a, b in zip("uvr", o, outs[0]):
                if key == "rk" and name == "r":
                    continue
                diff = int((a != b).sum())
                print(f"{shape} {variant} {key} candidate {cand} {name}: {diff} cells differ",
                      flush=True)
                assert diff == 0, (key, cand, name)


@pytest.mark.parametrize("shape", [(2048, 512), (1000, 300), (1024, 96), (512, 256)])
def test_smoother_candidates_bit_identical(gpu, shape):
    """Textbook plans with every level's pre- and post-smoother forced to
    one candidate (LDS tiles 0-3, register columns 4-8): two V-cycles of
    the same rhs give the same bits."""
    import aerojax_amd as aj
    from aerojax_amd import _lib
    L = _lib.load()
    nx, ny = shape
    dx, dy = 4.0 / nx, 1.0 / ny
    rng = np.random.default_rng(11)
    b = torch.from_numpy(rng.normal(size=(nx, ny)) * 100).to(gpu, torch.float32)
    shapes, n, m = [], nx, ny
    while n * m > 65536 or n > 1024:
        if n % 2 or m % 2:
            break
        shapes.append((n, m))
        n //= 2
        m //= 2
    assert shapes, "no smoothed level"
    res = {}
    for cand in range(9):
        L.aj_tune_reset()
        text = "".join(f"tb_pre/1x{a}x{c} -> {cand}\ntb_post/1x{a}x{c} -> {cand}\n"
                       for a, c in shapes)
        assert _lib.tune_load(text) == 2 * len(shapes)
        pl = aj.MultigridPlan(nx, ny, dx, dy, torch.float32, "textbook", 0)
        out = torch.empty_like(b)
        pl.solve(b, out, None, 2, 0.0)
        torch.cuda.synchronize()
        res[cand] = out.clone()
        pl.close()
    L.aj_tune_reset()
    for cand, r in res.items():
        diff = int((r != res[0]).sum())
        print(f"textbook {shape} levels {shapes} candidate {cand}: {diff} cells differ", flush=True)
        assert diff == 0, cand
