"""NumPy restatement of the corrector MLP's forward and backward passes with
the GPU kernels' rounding points (test infrastructure only).

The GPU stores X, the hidden activations H1, H2, the output gradient dY and
the hidden gradients dH2, dH1 in BF16 and accumulates every product in FP32;
this restatement rounds at the same points (round to nearest even) and
accumulates in FP64, so the two agree to FP32 accumulation-order noise.
Layouts follow femflow_amd/csrc/mpm_mlp.h: W1 [128][351], b1, W2 [128][128],
b2, W3 [3][128], b3 in one FP32 vector.
"""
from __future__ import annotations

import numpy as np

IN, H, OUT = 351, 128, 3
SIZES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)), ("b3", (OUT,))]


def n_params():
    return sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float32).reshape(-1) for name, _ in SIZES])


def bf16(a):
    """Round FP32 values to BF16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(np.asarray(a, dtype=np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _r(a, rounding):
    return bf16(a) if rounding else np.asarray(a, dtype=np.float64)


def forward(p, X, rounding=True):
    """X: (rows, 351) values already representable in BF16.  rounding=False
    gives the exact FP64 network (used to pin the formulas by differences)."""
    w = unpack(p)
    W1, W2, W3 = _r(w["W1"], rounding), _r(w["W2"], rounding), _r(w["W3"], rounding)
    H1 = _r(np.maximum(X @ W1.T + w["b1"], 0.0), rounding)
    H2 = _r(np.maximum(H1 @ W2.T + w["b2"], 0.0), rounding)
    Y = H2 @ W3.T + w["b3"]
    return Y, H1, H2


def loss_and_grads(p, X, T, wt=None, rounding=True):
    """Weighted MSE sum_i w_i |Y_i - T_i|^2 / sum_i w_i and its gradient."""
    w = unpack(p)
    Y, H1, H2 = forward(p, X, rounding)
    n = len(X)
    wt = np.ones(n) if wt is None else np.asarray(wt, dtype=np.float64)
    inv = 1.0 / wt.sum()
    d = (Y.astype(np.float32).astype(np.float64) if rounding else Y) - T
    loss = float((wt[:, None] * d * d).sum() * inv)
    dY = _r(2.0 * wt[:, None] * d * inv, rounding)
    W2, W3 = _r(w["W2"], rounding), _r(w["W3"], rounding)
    g = {"W3": dY.T @ H2, "b3": dY.sum(0)}
    dH2 = _r((dY @ W3) * (H2 > 0), rounding)
    g["W2"], g["b2"] = dH2.T @ H1, dH2.sum(0)
    dH1 = _r((dH2 @ W2) * (H1 > 0), rounding)
    g["W1"], g["b1"] = dH1.T @ X, dH1.sum(0)
    return loss, np.concatenate([np.asarray(g[name]).reshape(-1) for name, _ in SIZES]), Y


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
