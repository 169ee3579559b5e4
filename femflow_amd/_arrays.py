"""Device staging for caller arrays.

The reference mutates numpy arrays in place.  Here a caller may pass numpy
arrays (copied to the GPU, computed there, and copied back into the same
arrays) or CUDA/ROCm torch tensors (used in place, no copies).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


class Staged:
    """A contiguous device tensor for one caller array, written back on commit."""

    __slots__ = ("tensor", "_target", "_shape")

    def __init__(self, a, dtype, device, writable=True):
        if isinstance(a, torch.Tensor):
            if a.is_cuda and a.dtype == dtype and a.is_contiguous():
                self.tensor, self._target = a, None
            else:
                self.tensor = a.detach().to(device=device, dtype=dtype).contiguous()
                self._target = a if writable else None
        else:
            arr = np.asarray(a)
            self.tensor = torch.from_numpy(np.ascontiguousarray(arr, dtype=_np_dtype(dtype))).to(device)
            self._target = a if writable else None
        self._shape = tuple(self.tensor.shape)

    @property
    def ptr(self):
        return self.tensor.data_ptr()

    def commit(self):
        t = self._target
        if t is None:
            return
        if isinstance(t, torch.Tensor):
            with torch.no_grad():
                t.copy_(self.tensor.view(t.shape))
        else:
            np.copyto(t, self.tensor.cpu().numpy().reshape(np.shape(t)).astype(t.dtype, copy=False))


def _np_dtype(dtype):
    return {torch.float32: np.float32, torch.float64: np.float64, torch.int32: np.int32}[dtype]


def compute_dtype(*arrays):
    """FP32 if the state arrays are float32 (numpy or torch), else FP64."""
    for a in arrays:
        if a is None:
            continue
        dt = a.dtype
        if dt in (torch.float32, np.float32):
            return torch.float32
        if dt in (torch.float64, np.float64):
            return torch.float64
    return torch.float64


def mpm_dtype(dtype):
    return _lib.MPM_F32 if dtype == torch.float32 else _lib.MPM_F64


def device_of(*arrays):
    for a in arrays:
        if isinstance(a, torch.Tensor) and a.is_cuda:
            return a.device
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr(device):
    return torch.cuda.current_stream(device).cuda_stream
