"""Per-field error norms used by every parity test.

err(field) = ||a - b||_inf / max(||b||_inf, floor).  The floors are the
magnitude of the terms each field is summed from, so that a field that is
zero up to cancellation (C and the off-diagonals of F in free fall, for
example) is measured against the scale its rounding error actually has:
  x, F, Jp : floor 1e-12 (their norm is O(1))
  v        : floor 1e-6 (m/s)
  C        : floor 4 inv_dx ||v||_inf (C = 4 inv_dx sum w g (ijk - fx)^T)
"""
import numpy as np

FIELDS = ("x", "v", "F", "C", "Jp")


def _np(a):
    try:
        import torch

        if isinstance(a, torch.Tensor):
            return a.detach().double().cpu().numpy()
    except ImportError:
        pass
    return np.asarray(a, dtype=np.float64)


def field_errors(a, b, inv_dx):
    """Relative errors of state dict ``a`` against reference state dict ``b``."""
    out = {}
    n = len(_np(b["x"]))
    vmax = np.abs(_np(b["v"])).max() if n else 0.0
    floors = {"x": 1e-12, "v": 1e-6, "F": 1e-12, "C": 4.0 * inv_dx * max(vmax, 1e-6), "Jp": 1e-12}
    for k in FIELDS:
        ra = _np(a[k]).reshape(n, -1)
        rb = _np(b[k]).reshape(n, -1)
        scale = max(np.abs(rb).max() if n else 0.0, floors[k])
        out[k] = float(np.abs(ra - rb).max() / scale) if n else 0.0
    return out
