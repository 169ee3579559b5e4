"""Scalar helpers of femflow/solvers/mpm/utils.py (host side).

The per-particle versions run inside the HIP kernels; these scalar forms keep
the reference's module surface for callers that compute material parameters."""
from __future__ import annotations

import math


def constant_hardening(mu_0: float, lambda_0: float, e: float):
    """utils.py:7-24: (mu_0 e, lambda_0 e)."""
    return mu_0 * e, lambda_0 * e


def snow_hardening(mu_0: float, lambda_0: float, h: float, Jp: float):
    """utils.py:27-49: e = exp(h (1 - Jp))."""
    jp = float(Jp[0]) if hasattr(Jp, "__len__") else float(Jp)
    return constant_hardening(mu_0, lambda_0, math.exp(h * (1.0 - jp)))


def oob(base, res: int, ijk=(0, 0, 0)) -> bool:
    """utils.py:138-150: any component of base+ijk outside [0, res)."""
    return any(b + o >= res or b + o < 0 for b, o in zip(base, ijk))
