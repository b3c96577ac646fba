"""Device-buffer plumbing: torch-ROCm tensors hold the fields, the HIP
library does the arithmetic.  NumPy inputs are uploaded and results are
returned as NumPy (host round trip, for convenience only)."""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("aerojax_amd needs a ROCm GPU: there is no CPU fallback")


def dtype_code(t: torch.Tensor) -> int:
    if t.dtype == torch.float32:
        return _lib.AJ_F32
    if t.dtype == torch.float64:
        return _lib.AJ_F64
    raise TypeError(f"unsupported dtype {t.dtype} (float32 or float64)")


def torch_dtype_of(x):
    if isinstance(x, torch.Tensor):
        return x.dtype
    a = np.asarray(x)
    return torch.float64 if a.dtype == np.float64 else torch.float32


def as_device(x, dtype=None, device=None, name="array"):
    """Contiguous device tensor view/copy of ``x`` (torch tensor or array)."""
    if x is None:
        return None
    if isinstance(x, torch.Tensor):
        t = x
        if not t.is_cuda:
            t = t.to("cuda" if device is None else device)
        elif device is not None and t.device != torch.device(device):
            t = t.to(device)
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(x)))
        t = t.to("cuda" if device is None else device)
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    if not t.is_contiguous():
        t = t.contiguous()
    if t.dim() != 2:
        raise ValueError(f"{name} must be 2-D (nx, ny), got shape {tuple(t.shape)}")
    return t


def scalar_or_none(x):
    return None if x is None else x


def stream_handle(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def is_numpy(x) -> bool:
    return not isinstance(x, torch.Tensor)


def out_like(t: torch.Tensor, want_numpy: bool):
    return t.cpu().numpy() if want_numpy else t
