"""The allocation-free binning entry point (mpm_bin_noalloc) against mpm_bin,
through the C ABI: bit-identical keys and permutations on two calls in a row,
the second with fewer particles, and the size limit rejected cleanly."""
import ctypes as ct

import numpy as np
import pytest

from femflow_amd import _lib

pytestmark = pytest.mark.gpu


def _bin(fn, x, res, layout):
    import torch

    n = x.shape[0]
    keys = torch.full((n,), 0xDEAD, dtype=torch.int32, device=x.device)
    perm = torch.full((n,), 0xBEEF, dtype=torch.int32, device=x.device)
    rc = fn(None, x.shape[1], res, float(res), layout, x.data_ptr(), n, keys.data_ptr(), perm.data_ptr())
    assert rc == _lib.MPM_OK, _lib.last_error()
    torch.cuda.synchronize()
    return keys.cpu().numpy().view(np.uint32), perm.cpu().numpy().view(np.uint32)


def test_bin_noalloc_matches_bin(gpu):
    import torch

    from oracle import mpm_oracle as O

    L = _lib.load()
    rng = np.random.default_rng(5)
    for n, res, layout, dim in ((300_000, 128, _lib.MPM_KEY_BLOCKED, 3), (70_001, 64, _lib.MPM_KEY_ROWMAJOR, 3),
                                (40_000, 128, _lib.MPM_KEY_ROWMAJOR, 2)):
        x = (rng.random((n, dim)) * 0.7 + 0.15).astype(np.float32)
        x[: n // 4] = x[: n // 4].round(2)  # many equal keys: stability matters
        xt = torch.tensor(x, device=gpu)
        k0, p0 = _bin(L.mpm_bin, xt, res, layout)
        k1, p1 = _bin(L.mpm_bin_noalloc, xt, res, layout)
        assert np.array_equal(k0, k1) and np.array_equal(p0, p1), (n, res, layout)
        ref = O.bin_keys(x, float(res), res, layout="blocked" if layout == _lib.MPM_KEY_BLOCKED else "rowmajor")
        assert np.array_equal(k1, ref) and np.array_equal(p1, O.bin_sort(ref))


def test_bin_noalloc_full_size_and_limit(gpu):
    import torch

    L = _lib.load()
    n = 1 << 20
    x = torch.rand((n, 3), device=gpu) * 0.8 + 0.1
    k0, p0 = _bin(L.mpm_bin, x, 128, _lib.MPM_KEY_BLOCKED)
    k1, p1 = _bin(L.mpm_bin_noalloc, x, 128, _lib.MPM_KEY_BLOCKED)
    assert np.array_equal(k0, k1) and np.array_equal(p0, p1)
    big = torch.zeros(((1 << 20) + 1, 3), device=gpu)
    out = torch.empty((1 << 20) + 1, dtype=torch.int32, device=gpu)
    rc = L.mpm_bin_noalloc(None, 3, 128, 128.0, _lib.MPM_KEY_BLOCKED, big.data_ptr(), big.shape[0], out.data_ptr(),
                           out.data_ptr())
    assert rc == _lib.MPM_ERR_ARG and "exceeds" in _lib.last_error()
