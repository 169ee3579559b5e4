"""Scene generators feeding the rollout driver (femflow/simulation/mpm/primitives.py:9-76,
femflow/numerics/geometry.py:100-116), vectorised with numpy.  Host-side set-up
code: it runs once before a rollout, not on the substep path."""
from __future__ import annotations

from typing import Callable, Tuple, Union

import numpy as np


def grid(res) -> np.ndarray:
    """Lattice vertices with the first coordinate varying fastest; coordinate
    c is sub_c / (res_c - 1), or sub_c / res_c when res_c == 1
    (numerics/geometry.py:100-116)."""
    res = np.asarray(res, dtype=np.int64)
    axes = [np.arange(r) / (r if r == 1 else r - 1) for r in res]
    mesh = np.meshgrid(*axes[::-1], indexing="ij")  # last axis slowest
    return np.stack([m.reshape(-1) for m in mesh[::-1]], axis=1).astype(np.float64)


def _angle(k):
    return (2.0 * np.pi) / k


def primitive(k: float, t: float, pos: np.ndarray) -> np.ndarray:
    """Schwarz P surface minus t (primitives.py:9-13); pos (..., 3)."""
    a = _angle(k)
    x, y, z = np.moveaxis(np.asarray(pos, dtype=np.float64), -1, 0)
    return np.cos(a * x) + np.cos(a * y) + np.cos(a * z) - t


def gyroid(k: float, t: float, pos: np.ndarray) -> np.ndarray:
    """Gyroid minus t (primitives.py:16-25)."""
    a = _angle(k)
    x, y, z = np.moveaxis(np.asarray(pos, dtype=np.float64), -1, 0)
    return (np.sin(a * x) * np.cos(a * y) + np.sin(a * y) * np.cos(a * z) + np.sin(a * z) * np.cos(a * x)
            - t)


def diamond(k: float, t: float, pos: np.ndarray) -> np.ndarray:
    """Schwarz D surface minus t (primitives.py:28-38)."""
    a = _angle(k)
    x, y, z = np.moveaxis(np.asarray(pos, dtype=np.float64), -1, 0)
    sx, sy, sz, cx, cy, cz = np.sin(a * x), np.sin(a * y), np.sin(a * z), np.cos(a * x), np.cos(a * y), np.cos(a * z)
    return sx * sy * sz + sx * cy * cz + cx * sy * cz + cx * cy * sz - t


_IMPLICIT = {"gyroid": gyroid, "diamond": diamond, "primitive": primitive}


def generate_implicit_points(implicit_fn: Union[Callable, str], k: float, t: float, res: int) -> np.ndarray:
    """Lattice points where the implicit function (already minus t) exceeds t
    again -- the reference compares ``inside > t`` (primitives.py:46-61)."""
    if isinstance(implicit_fn, str):
        if implicit_fn not in _IMPLICIT:
            raise ValueError("Invalid implicit function specified")
        implicit_fn = _IMPLICIT[implicit_fn]
    g = grid(np.array((res, res, res)))
    inside = np.asarray(implicit_fn(k, t, g))
    return g[inside > t]


def generate_cube_points(xb: Tuple[float, float], yb: Tuple[float, float], zb: Tuple[float, float],
                         res: int = 10) -> np.ndarray:
    """res^3 lattice in the box; rows are (z, y, x) with x fastest -- the
    reference's column order (primitives.py:64-76)."""
    x = np.linspace(*xb, num=res)
    y = np.linspace(*yb, num=res)
    z = np.linspace(*zb, num=res)
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    return np.stack([zz.reshape(-1), yy.reshape(-1), xx.reshape(-1)], axis=1).astype(np.float64)
