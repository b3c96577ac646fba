"""Learned multigrid warm start: U-Net (rhs, sdf) -> p on BF16 MFMA.

Python handle of the ``aj_nn_*`` C ABI.  The network follows the
conventions of the reference's learned pressure operators
(neural_operators/training_utils.py:84-253: MSE loss, gradient clipping to
+-1, Adam, per-sample input normalisation) but is used as the initial guess
of the multigrid solve, which then runs to the same tolerance.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _arrays as A
from . import _lib


class WarmStartNet:
    """``levels``-level U-Net with ``base_channels`` channels at the input
    resolution; runs on the fine grid restricted ``r`` times."""

    def __init__(self, base_channels=32, levels=4, seed=0, r=2, sdf_ref=0.1, pscale=1.0,
                 device=None):
        A.require_gpu()
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None \
            else torch.device(device)
        self.c0, self.levels, self.r = int(base_channels), int(levels), int(r)
        self.sdf_ref, self.pscale = float(sdf_ref), float(pscale)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            rc = _lib.load().aj_nn_create(ctypes.byref(h), self.c0, self.levels, int(seed))
        _lib.check(rc, "aj_nn_create")
        self.handle = h
        self._loss = torch.zeros(1, dtype=torch.float32, device=self.device)

    @property
    def num_params(self) -> int:
        return int(_lib.load().aj_nn_num_params(self.handle))

    def set_option(self, fused_io: bool | None = None):
        """``fused_io`` (default True): the warm start builds the network
        input in its first convolution and applies the 1x1 head in its last
        one (two launches fewer, the same bits); False runs them as
        separate launches."""
        if fused_io is not None:
            _lib.check(_lib.load().aj_nn_set_option(self.handle, _lib.AJ_NN_OPT_FUSED_IO,
                                                    int(bool(fused_io))), "aj_nn_set_option")

    def get_params(self, grads=False) -> np.ndarray:
        out = np.empty(self.num_params, dtype=np.float32)
        rc = _lib.load().aj_nn_get_params(self.handle, out.ctypes.data, out.size, int(grads))
        _lib.check(rc, "aj_nn_get_params")
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float32))
        rc = _lib.load().aj_nn_set_params(self.handle, p.ctypes.data, p.size)
        _lib.check(rc, "aj_nn_set_params")

    def _s(self):
        return A.stream_handle(self.device)

    def forward(self, x: torch.Tensor, idx=None) -> torch.Tensor:
        """x: float32 [B, H, W, 2] (or [N, ...] with idx) -> [B, H, W]."""
        B = x.shape[0] if idx is None else idx.shape[0]
        H, W = x.shape[1], x.shape[2]
        out = torch.empty((B, H, W), dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            rc = _lib.load().aj_nn_forward(self.handle, x.data_ptr(), A.ptr(idx), out.data_ptr(),
                                           B, H, W, self._s())
        _lib.check(rc, "aj_nn_forward")
        return out

    def loss_grad(self, x, t, idx=None) -> torch.Tensor:
        B = x.shape[0] if idx is None else idx.shape[0]
        with torch.cuda.device(self.device):
            rc = _lib.load().aj_nn_loss_grad(self.handle, x.data_ptr(), t.data_ptr(), A.ptr(idx),
                                             B, x.shape[1], x.shape[2], self._loss.data_ptr(),
                                             self._s())
        _lib.check(rc, "aj_nn_loss_grad")
        return self._loss

    def adam(self, lr, clip=1.0):
        with torch.cuda.device(self.device):
            rc = _lib.load().aj_nn_adam(self.handle, float(lr), float(clip), self._s())
        _lib.check(rc, "aj_nn_adam")

    def train_step(self, x, t, idx, lr) -> torch.Tensor:
        """One Adam step on samples ``idx`` (device int32) of x / t; returns
        the device loss tensor (no host sync)."""
        with torch.cuda.device(self.device):
            rc = _lib.load().aj_nn_train_step(self.handle, x.data_ptr(), t.data_ptr(),
                                              A.ptr(idx), idx.shape[0], x.shape[1], x.shape[2],
                                              float(lr), self._loss.data_ptr(), self._s())
        _lib.check(rc, "aj_nn_train_step")
        return self._loss

    def make_sample(self, rhs, sdf, p, x_out, t_out):
        nx, ny = rhs.shape
        with torch.cuda.device(self.device):
            rc = _lib.load().aj_nn_make_sample(self.handle, rhs.data_ptr(), sdf.data_ptr(),
                                               A.ptr(p), nx, ny, self.r, self.sdf_ref,
                                               self.pscale, x_out.data_ptr(), A.ptr(t_out),
                                               self._s())
        _lib.check(rc, "aj_nn_make_sample")

    def warm_start(self, rhs, sdf, p0_out=None) -> torch.Tensor:
        """Initial pressure guess on the fine grid (float32)."""
        nx, ny = rhs.shape
        p0 = torch.empty_like(rhs) if p0_out is None else p0_out
        with torch.cuda.device(self.device):
            rc = _lib.load().aj_nn_warm_start(self.handle, rhs.data_ptr(), sdf.data_ptr(),
                                              p0.data_ptr(), nx, ny, self.r, self.sdf_ref,
                                              self.pscale, self._s())
        _lib.check(rc, "aj_nn_warm_start")
        return p0

    def warm_start_batched(self, rhs, sdf, p0_out, count, nx, ny):
        """Initial guesses of ``count`` instances of one (nx, ny) grid stored
        one after the other in the flat tensors rhs / sdf / p0_out (one
        inference of batch ``count``)."""
        with torch.cuda.device(self.device):
            rc = _lib.load().aj_nn_warm_start_batched(self.handle, rhs.data_ptr(), sdf.data_ptr(),
                                                      p0_out.data_ptr(), int(count), int(nx),
                                                      int(ny), self.r, self.sdf_ref, self.pscale,
                                                      self._s())
        _lib.check(rc, "aj_nn_warm_start_batched")
        return p0_out

    def __call__(self, rhs, sdf):
        return self.warm_start(rhs, sdf)

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().aj_nn_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
