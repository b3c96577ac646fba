"""Allocation-free diagnostics entry point: aj_flow_diagnostics_noalloc,
called through the C ABI, gives the bits of aj_flow_diagnostics on two
calls in a row (the second on a smaller grid, so stale partial sums of the
first call would show), and refuses grids above its fixed maximum before
launching anything."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import reference_step as R

pytestmark = pytest.mark.gpu

FIELDS = ("pressure_residual", "rhs_norm", "relative_residual", "div_rms", "l2_div",
          "fluid_cells", "max_speed", "cfl")


def _fields(nx, ny, dtype, seed):
    dx, dy, X, Y = R.grid_coords(nx, ny, 4.0, 1.0)
    sdf = R.cylinder_sdf(X, Y, 1.0, 0.5, 0.1)
    mask = R.mask_from_sdf(sdf, 0.05)
    rng = np.random.default_rng(seed)
    u = 1.0 + 1e-2 * rng.normal(size=(nx, ny))
    v = 1e-2 * rng.normal(size=(nx, ny))
    p = rng.normal(size=(nx, ny))
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dtype)  # noqa: E731
    return t(u), t(v), t(p), t(mask), dx, dy


def _call(L, name, dc, u, v, p, mask, dx, dy, dt):
    from aerojax_amd import _lib
    out = _lib.FlowDiag()
    nx, ny = u.shape
    rc = getattr(L, name)(dc, u.data_ptr(), v.data_ptr(), p.data_ptr(), mask.data_ptr(), nx, ny,
                          dx, dy, dt, ctypes.byref(out), torch.cuda.current_stream().cuda_stream)
    _lib.check(rc, name)
    return out.as_dict()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_diagnostics_equal_allocating_entry_point(gpu, dtype):
    from aerojax_amd import _lib
    L = _lib.load()
    dc = _lib.AJ_F32 if dtype == torch.float32 else _lib.AJ_F64
    for (nx, ny, seed) in ((1024, 512, 1), (256, 96, 2)):
        print(f"noalloc diagnostics {nx}x{ny} {dtype} seed {seed}", flush=True)
        u, v, p, mask, dx, dy = _fields(nx, ny, dtype, seed)
        ref = _call(L, "aj_flow_diagnostics", dc, u, v, p, mask, dx, dy, 1e-3)
        got = _call(L, "aj_flow_diagnostics_noalloc", dc, u, v, p, mask, dx, dy, 1e-3)
        for k in FIELDS:
            assert got[k] == ref[k], (k, got[k], ref[k])
        assert ref["fluid_cells"] > 0 and ref["max_speed"] > 0


def test_noalloc_diagnostics_rejects_grids_above_the_maximum(gpu):
    from aerojax_amd import _lib
    L = _lib.load()
    dev = torch.device("cuda", 0)
    nx, ny = 4096, 2049  # 8,392,704 cells > 8,388,608; the check precedes any launch
    x = torch.zeros(1, dtype=torch.float32, device=dev)
    out = _lib.FlowDiag()
    rc = L.aj_flow_diagnostics_noalloc(_lib.AJ_F32, x.data_ptr(), x.data_ptr(), x.data_ptr(),
                                       x.data_ptr(), nx, ny, 1.0, 1.0, 1.0, ctypes.byref(out),
                                       torch.cuda.current_stream().cuda_stream)
    assert rc == _lib.AJ_ERR_INVALID
    assert b"exceeds the maximum" in L.aj_last_error()
