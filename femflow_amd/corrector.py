"""Learned grid corrector: a small MLP that lets the solver take dt_big = 8 dt.

The network maps the 3x3x3 node-stencil features of an active grid node
(mass, momentum and the mass-weighted mean deformation gradient of the
particles in each of the 27 neighbouring cells: 27 x 13 = 351 values) after
the big step's P2G to a correction of that node's updated velocity.  It is
trained on rollouts the solver itself generates (targets from 8 plain
substeps), by the library's own BF16-MFMA forward and backward kernels
(`mpm_mlp_*`), with Adam in FP32.  No reference counterpart: the reference
has no learned component.
"""
from __future__ import annotations

import ctypes as ct
import json
import math
import os
from typing import Optional

import numpy as np
import torch

from . import _lib
from ._arrays import stream_ptr

IN, IN_PAD, HIDDEN, OUT, OUT_PAD, DY_PAD = 351, 352, 128, 3, 16, 32


class CorrectorMLP:
    """FP32 master parameters + BF16 operand copies on one device."""

    def __init__(self, device=None, seed: int = 0, params: Optional[torch.Tensor] = None,
                 out_scale: float = 0.1):
        _lib.require_gpu()
        self.L = _lib.load()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        npar, npk = ct.c_int64(), ct.c_int64()
        _lib.check(self.L.mpm_mlp_sizes(ct.byref(npar), ct.byref(npk)), "mpm_mlp_sizes")
        self.n_params, self.n_packed = npar.value, npk.value
        if params is None:
            params = torch.from_numpy(self.init_params(seed, out_scale))
        self.params = params.to(device=self.device, dtype=torch.float32).contiguous().clone()
        if self.params.numel() != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {self.params.numel()}")
        self.packed = torch.empty(self.n_packed, dtype=torch.bfloat16, device=self.device)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self._ws_rows = 0
        self.pack()

    @staticmethod
    def init_params(seed: int = 0, out_scale: float = 0.1) -> np.ndarray:
        """U(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases; the output
        layer scaled by out_scale so an untrained corrector starts near zero."""
        rng = np.random.default_rng(seed)
        parts = []
        for fan_out, fan_in, scale in ((HIDDEN, IN, 1.0), (HIDDEN, HIDDEN, 1.0), (OUT, HIDDEN, out_scale)):
            b = 1.0 / math.sqrt(fan_in)
            parts.append(scale * rng.uniform(-b, b, size=fan_out * fan_in))
            parts.append(scale * rng.uniform(-b, b, size=fan_out))
        return np.concatenate(parts).astype(np.float32)

    def pack(self):
        _lib.check(self.L.mpm_mlp_pack(stream_ptr(self.device), self.params.data_ptr(), self.packed.data_ptr()),
                   "mpm_mlp_pack")

    def _workspace(self, rows: int):
        if rows > self._ws_rows:
            cap = max(rows, 1024)
            d = self.device
            self.H1 = torch.empty((cap, HIDDEN), dtype=torch.bfloat16, device=d)
            self.H2 = torch.empty((cap, HIDDEN), dtype=torch.bfloat16, device=d)
            self.dH1 = torch.empty((cap, HIDDEN), dtype=torch.bfloat16, device=d)
            self.dH2 = torch.empty((cap, HIDDEN), dtype=torch.bfloat16, device=d)
            self.Y = torch.empty((cap, OUT_PAD), dtype=torch.float32, device=d)
            self.dY = torch.empty((cap, DY_PAD), dtype=torch.bfloat16, device=d)
            self._ws_rows = cap

    @staticmethod
    def _check_X(X: torch.Tensor):
        if X.dtype != torch.bfloat16 or X.dim() != 2 or X.shape[1] != IN_PAD or not X.is_contiguous():
            raise ValueError("X must be a contiguous (rows, 352) bfloat16 tensor (column 351 zero)")

    def forward(self, X: torch.Tensor) -> torch.Tensor:
        """Y (rows, 3) FP32 for features X (rows, 352) BF16."""
        self._check_X(X)
        rows = X.shape[0]
        self._workspace(rows)
        _lib.check(self.L.mpm_mlp_forward(stream_ptr(self.device), self.packed.data_ptr(), self.params.data_ptr(),
                                          X.data_ptr(), rows, self.H1.data_ptr(), self.H2.data_ptr(),
                                          self.Y.data_ptr()), "mpm_mlp_forward")
        return self.Y[:rows, :OUT]

    def loss_grads(self, X: torch.Tensor, T: torch.Tensor, w: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Forward + backward; gradients into self.grads (overwritten).
        Returns the loss as a device double tensor (no host sync)."""
        self._check_X(X)
        rows = X.shape[0]
        T = T.to(device=self.device, dtype=torch.float32).contiguous()
        if w is not None:
            w = w.to(device=self.device, dtype=torch.float32).contiguous()
            norm = float(w.sum())
        else:
            norm = float(rows)
        self.forward(X)
        loss = torch.zeros(1, dtype=torch.float64, device=self.device)
        st = stream_ptr(self.device)
        _lib.check(self.L.mpm_mlp_loss(st, self.Y.data_ptr(), T.data_ptr(), w.data_ptr() if w is not None else None,
                                       rows, 1.0 / max(norm, 1e-30), self.dY.data_ptr(), loss.data_ptr()),
                   "mpm_mlp_loss")
        self.grads.zero_()
        _lib.check(self.L.mpm_mlp_backward(st, self.packed.data_ptr(), X.data_ptr(), rows, self.H1.data_ptr(),
                                           self.H2.data_ptr(), self.dY.data_ptr(), self.dH2.data_ptr(),
                                           self.dH1.data_ptr(), self.grads.data_ptr()), "mpm_mlp_backward")
        return loss

    def adam_step(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        self.step += 1
        _lib.check(self.L.mpm_mlp_adam(stream_ptr(self.device), self.params.data_ptr(), self.grads.data_ptr(),
                                       self.m.data_ptr(), self.v.data_ptr(), self.n_params, lr, beta1, beta2, eps,
                                       self.step), "mpm_mlp_adam")
        self.pack()

    def train_step(self, X, T, w=None, lr: float = 1e-3) -> torch.Tensor:
        loss = self.loss_grads(X, T, w)
        self.adam_step(lr)
        return loss

    # -- persistence ---------------------------------------------------------
    def save(self, path: str, meta: Optional[dict] = None):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, params=self.params.detach().cpu().numpy(), meta=np.array(json.dumps(meta or {})))

    @classmethod
    def load(cls, path: str, device=None):
        with np.load(path, allow_pickle=False) as f:
            params = torch.from_numpy(f["params"].astype(np.float32))
            meta = json.loads(str(f["meta"]))
        net = cls(device=device, params=params)
        net.meta = meta
        return net


def to_bf16_features(X: np.ndarray, device) -> torch.Tensor:
    """(rows, 351) float array -> contiguous (rows, 352) bfloat16 device tensor."""
    rows = X.shape[0]
    out = torch.zeros((rows, IN_PAD), dtype=torch.bfloat16, device=device)
    out[:, :IN] = torch.as_tensor(np.asarray(X, dtype=np.float32), device=device).to(torch.bfloat16)
    return out
