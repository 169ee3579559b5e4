"""ctypes front end for oracle/mpm_oracle.c (test infrastructure only).

Used by tests/ as a second restatement of the reference step and by bench.py's
CPU-baseline leg.  The library is built by ``oracle/Makefile`` (called from
``__graft_entry__.build()``) into ``oracle/_build/libmpm_oracle.so``.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmpm_oracle.so")
_lib = None


class _Particles(ct.Structure):
    _fields_ = [("n", ct.c_int64)] + [(f, ct.c_void_p) for f in (
        "x", "v", "F", "C", "Jp", "mass", "mu0", "lam0", "h", "theta_c", "theta_s", "kind")]


class _Params(ct.Structure):
    _fields_ = [("res", ct.c_int), ("inv_dx", ct.c_double), ("dx", ct.c_double),
                ("dt", ct.c_double), ("volume", ct.c_double), ("gravity", ct.c_double),
                ("textbook", ct.c_int), ("sticky", ct.c_int), ("bound", ct.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ct.CDLL(_LIB_PATH)
        P, Q = ct.POINTER(_Particles), ct.POINTER(_Params)
        dp = ct.c_void_p
        _lib.oracle_run.argtypes = [ct.c_int, P, Q, dp, dp, ct.c_int, ct.c_int]
        _lib.oracle_run.restype = ct.c_int
        _lib.oracle_p2g3d.argtypes = [P, Q, dp, dp, ct.c_int]
        _lib.oracle_p2g3d.restype = ct.c_int
        _lib.oracle_grid_op3d.argtypes = [Q, dp, dp, ct.c_int]
        _lib.oracle_g2p3d.argtypes = [P, Q, dp, ct.c_int]
        _lib.oracle_g2p3d.restype = ct.c_int
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class CState:
    """FP64 contiguous copies of a state dict, mutated in place by the C code."""

    def __init__(self, state):
        n, dim = state["x"].shape
        self.dim = dim
        self.arrays = {
            "x": _c(state["x"], np.float64).copy(), "v": _c(state["v"], np.float64).copy(),
            "F": _c(state["F"], np.float64).reshape(n, dim * dim).copy(),
            "C": _c(state["C"], np.float64).reshape(n, dim * dim).copy(),
            "Jp": _c(state["Jp"], np.float64).reshape(n).copy(),
            "mass": _c(np.broadcast_to(state["mass"], (n,)), np.float64),
            "mu0": _c(np.broadcast_to(state["mu0"], (n,)), np.float64),
            "lam0": _c(np.broadcast_to(state["lam0"], (n,)), np.float64),
            "h": _c(np.broadcast_to(state["h"], (n,)), np.float64),
            "theta_c": _c(np.broadcast_to(state["theta_c"], (n,)), np.float64),
            "theta_s": _c(np.broadcast_to(state["theta_s"], (n,)), np.float64),
            "kind": _c(np.broadcast_to(state["kind"], (n,)), np.int32),
        }
        self.struct = _Particles(n, *[self.arrays[f].ctypes.data for f in (
            "x", "v", "F", "C", "Jp", "mass", "mu0", "lam0", "h", "theta_c", "theta_s", "kind")])

    def to_state(self):
        n = self.arrays["x"].shape[0]
        d = self.dim
        out = dict(self.arrays)
        out["F"] = self.arrays["F"].reshape(n, d, d)
        out["C"] = self.arrays["C"].reshape(n, d, d)
        out["Jp"] = self.arrays["Jp"].reshape(n, 1)
        return out


def make_params(params, quirks="reference"):
    return _Params(params["res"], params["inv_dx"], params["dx"], params["dt"], params["volume"],
                   params["gravity"], 1 if quirks == "textbook" else 0,
                   1 if params.get("boundary", "reference") == "sticky" else 0,
                   params.get("bound", 3))


def run(state, params, nsteps, nthreads=1, quirks="reference", cstate=None):
    """Advance ``nsteps`` substeps in C; returns the CState (FP64 arrays)."""
    L = lib()
    cs = cstate if cstate is not None else CState(state)
    r1 = params["res"] + 1
    dim = cs.dim
    gv = np.zeros(r1 ** dim * dim)
    gm = np.zeros(r1 ** dim)
    prm = make_params(params, quirks)
    rc = L.oracle_run(dim, ct.byref(cs.struct), ct.byref(prm), gv.ctypes.data, gm.ctypes.data,
                      nsteps, nthreads)
    if rc:
        raise RuntimeError("particle outside the grid")
    return cs
