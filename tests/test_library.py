"""CPU-side checks of the C ABI library: it exists, loads without a GPU,
exports every function the header declares, and the Python layer refuses to
run without a GPU instead of falling back to the CPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "aerojax.h")
LIB = os.path.join(ROOT, "aerojax_amd", "libaerojax_hip.so")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aj_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_api():
    fns = header_functions()
    for name in ("aj_rk_step", "aj_poisson_mg", "aj_project", "aj_divergence_rhs",
                 "aj_les_smagorinsky", "aj_mg_create"):
        assert name in fns


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from aerojax_amd import build
        build.build()
    return ctypes.CDLL(LIB)


def test_library_exports_every_header_symbol(lib):
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_python_binding_covers_header():
    from aerojax_amd import _lib
    assert sorted(_lib.EXPORTS) == header_functions()


def test_version_and_error_string_without_gpu(lib):
    lib.aj_version.restype = ctypes.c_char_p
    lib.aj_last_error.restype = ctypes.c_char_p
    assert b"gfx950" in lib.aj_version()
    # argument validation happens before any HIP call
    lib.aj_rk_step.restype = ctypes.c_int
    rc = lib.aj_rk_step(0, None, None, None, None, None, None, None, 16, 16, None, None)
    assert rc == 1
    assert b"null" in lib.aj_last_error()


def test_tune_decisions_round_trip_without_gpu():
    """aj_tune_load reads the lines aj_tune_report writes (host-only state),
    ignores lines without an arrow and rejects an arrow without a pick."""
    from aerojax_amd import _lib
    L = _lib.load()
    L.aj_tune_reset()
    text = "tb_pre/1x2048x512 -> 2 [0.0100 0.0090 0.0080 0.0110 ms]\nrk/f32/rk3/2048x512 -> 1\nnoise\n"
    assert _lib.tune_load(text) == 2
    rep = _lib.tune_report().splitlines()
    assert [l.split(" [")[0] for l in rep] == ["rk/f32/rk3/2048x512 -> 1", "tb_pre/1x2048x512 -> 2"]
    assert _lib.tune_load("\n".join(rep)) == 2
    with pytest.raises(_lib.AeroJaxError, match="bad decision line"):
        _lib.tune_load("tb_post/1x8x8 -> x")
    L.aj_tune_reset()
    assert _lib.tune_report() == ""


def test_struct_layouts_match_header():
    from aerojax_amd import _lib
    assert ctypes.sizeof(_lib.RkParams) == 7 * 8 + 2 * 4 + 0  # 64 bytes, no tail padding
    assert ctypes.sizeof(_lib.MgStats) == 4 * 4 + 8 + 64 * 8 + 2 * 8 + 2 * 4


def test_no_cpu_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import numpy as np
    import aerojax_amd
    u = np.ones((8, 8), np.float32)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        aerojax_amd.rk_step_unified(u, u, 1e-3, 1e-3, 0.1, 0.1, u)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        aerojax_amd.poisson_multigrid(u, u, 0.1, 0.1)


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "aerojax_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in re.findall(r"^\s*(?:from|import)\s+(\w+)", src, re.M), f
