"""ctypes binding of ``libmpm355.so`` (the C ABI in ``include/mpm355.h``).

The library is the only compute path of this package: when it is missing, or
no ROCm GPU is visible, every operation raises instead of falling back to a
CPU implementation.
"""
from __future__ import annotations

import ctypes as ct
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmpm355.so")

MPM_OK, MPM_ERR_OOB, MPM_ERR_NONFINITE, MPM_ERR_ARG, MPM_ERR_HIP = 0, 1, 2, 3, 4
MPM_F32, MPM_F64 = 0, 1
MPM_ELASTIC, MPM_SNOW, MPM_WATER = 0, 1, 2
MPM_QUIRKS_REFERENCE, MPM_QUIRKS_TEXTBOOK = 0, 1
MPM_BC_REFERENCE, MPM_BC_STICKY = 0, 1
MPM_KEY_BLOCKED, MPM_KEY_ROWMAJOR = 0, 1

# every symbol include/mpm355.h declares (checked by the CPU test suite)
EXPORTS = (
    "mpm_last_error_string", "mpm_version", "mpm_device_count", "mpm_p2g", "mpm_grid_op",
    "mpm_g2p", "mpm_ctx_create", "mpm_ctx_destroy", "mpm_ctx_set_stream", "mpm_ctx_tune", "mpm_ctx_set_spread",
    "mpm_ctx_kernel_timing", "mpm_ctx_kernel_times", "mpm_state_upload", "mpm_state_download",
    "mpm_run", "mpm_run_frames", "mpm_run_states", "mpm_run_states_grids", "mpm_state_positions", "mpm_substep", "mpm_ctx_oob_count", "mpm_ctx_fallback_count",
    "mpm_ctx_active_blocks", "mpm_ctx_request_rebin",
    "mpm_particle_sums", "mpm_bin", "mpm_bin_noalloc", "mpm_dense_workspace_bytes", "mpm_dense_substep", "mpm_adjoint_step",
    "mpm_adjoint_step_grids",
    "mpm_ctx_set_corrector", "mpm_corrector_sample", "mpm_grid_velocity",
    "mpm_collide_planes", "mpm_ctx_set_planes",
    "mpm_mlp_sizes", "mpm_mlp_pack", "mpm_mlp_forward", "mpm_mlp_loss", "mpm_mlp_backward", "mpm_mlp_adam",
)


class Material(ct.Structure):
    _fields_ = [("kind", ct.c_int32), ("hardening", ct.c_float), ("theta_c", ct.c_float),
                ("theta_s", ct.c_float)]


class Params(ct.Structure):
    _fields_ = [("dim", ct.c_int32), ("res", ct.c_int32), ("inv_dx", ct.c_double),
                ("dx", ct.c_double), ("dt", ct.c_double), ("volume", ct.c_double),
                ("gravity", ct.c_double), ("boundary", ct.c_int32), ("bound", ct.c_int32),
                ("quirks", ct.c_int32), ("check_oob", ct.c_int32)]


class Particles(ct.Structure):
    _fields_ = [("n", ct.c_int64), ("dtype", ct.c_int32), ("mat_scalar", ct.c_int32),
                ("x", ct.c_void_p), ("v", ct.c_void_p), ("F", ct.c_void_p), ("C", ct.c_void_p),
                ("Jp", ct.c_void_p), ("mass", ct.c_void_p), ("mu0", ct.c_void_p),
                ("lambda0", ct.c_void_p), ("material", ct.c_void_p)]


class Scene(ct.Structure):
    _fields_ = [("res", ct.c_int32), ("boundary", ct.c_int32), ("bound", ct.c_int32),
                ("quirks", ct.c_int32), ("dt", ct.c_double), ("gravity", ct.c_double),
                ("volume", ct.c_double)]


_lib = None
_lock = threading.Lock()


class ExtensionMissing(RuntimeError):
    pass


def load():
    """Load the HIP library (built by ``femflow_amd.build`` / ``__graft_entry__.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ExtensionMissing(
                f"{LIB_PATH} is missing: build it with `python -m femflow_amd.build` "
                "(femflow_amd has no CPU fallback)")
        L = ct.CDLL(LIB_PATH)
        i32, i64, vp, dbl = ct.c_int32, ct.c_int64, ct.c_void_p, ct.c_double
        P, Q, M, S = ct.POINTER(Particles), ct.POINTER(Params), ct.POINTER(Material), ct.POINTER(Scene)
        sig = {
            "mpm_last_error_string": ([], ct.c_char_p),
            "mpm_version": ([], i32),
            "mpm_device_count": ([], i32),
            "mpm_p2g": ([vp, Q, P, i32, dbl, M, i32, vp, vp, vp], i32),
            "mpm_grid_op": ([vp, Q, i32, vp, vp], i32),
            "mpm_collide_planes": ([vp, i32, i32, dbl, i32, vp, vp, i32, vp], i32),
            "mpm_ctx_set_planes": ([vp, i32, i32, vp, vp, i32], i32),
            "mpm_g2p": ([vp, Q, P, i32, dbl, dbl, M, i32, vp, vp], i32),
            "mpm_ctx_create": ([ct.POINTER(vp), i64, i32, S, i32, M, i32], i32),
            "mpm_ctx_destroy": ([vp], i32),
            "mpm_ctx_set_stream": ([vp, vp], i32),
            "mpm_ctx_tune": ([vp, i32, i32, i32, i32], i32),
            "mpm_ctx_set_spread": ([vp, i32], i32),
            "mpm_ctx_kernel_timing": ([vp, i32], i32),
            "mpm_ctx_kernel_times": ([vp, ct.POINTER(dbl), ct.POINTER(i64)], i32),
            "mpm_state_upload": ([vp, P, vp], i32),
            "mpm_state_download": ([vp, P], i32),
            "mpm_run": ([vp, i32, vp], i32),
            "mpm_run_frames": ([vp, i32, vp], i32),
            "mpm_run_states": ([vp, i32, vp, vp, vp, vp, vp], i32),
            "mpm_run_states_grids": ([vp, i32, vp, vp, vp, vp, vp, vp], i32),
            "mpm_state_positions": ([vp, vp], i32),
            "mpm_substep": ([vp, P, vp, i32], i32),
            "mpm_ctx_oob_count": ([vp, ct.POINTER(i64)], i32),
            "mpm_ctx_fallback_count": ([vp, ct.POINTER(i64)], i32),
            "mpm_ctx_active_blocks": ([vp, ct.POINTER(i64)], i32),
            "mpm_ctx_request_rebin": ([vp], i32),
            "mpm_particle_sums": ([vp, ct.POINTER(dbl)], i32),
            "mpm_bin": ([vp, i32, i32, dbl, i32, vp, i64, vp, vp], i32),
            "mpm_bin_noalloc": ([vp, i32, i32, dbl, i32, vp, i64, vp, vp], i32),
            "mpm_dense_workspace_bytes": ([Q, ct.POINTER(i64)], i32),
            "mpm_dense_substep": ([vp, Q, P, dbl, vp, vp], i32),
            "mpm_adjoint_step": ([vp, Q, P, dbl, vp, vp, vp, vp, vp, vp, vp, vp], i32),
            "mpm_adjoint_step_grids": ([vp, Q, P, dbl, vp, vp, vp, vp, vp, vp, vp, vp, vp], i32),
            "mpm_ctx_set_corrector": ([vp, vp, ct.c_float, ct.c_float, i32], i32),
            "mpm_corrector_sample": ([vp, vp, vp, vp, i64, ct.POINTER(i64)], i32),
            "mpm_grid_velocity": ([vp, vp, i64, vp], i32),
            "mpm_mlp_sizes": ([ct.POINTER(i64), ct.POINTER(i64)], i32),
            "mpm_mlp_pack": ([vp, vp, vp], i32),
            "mpm_mlp_forward": ([vp, vp, vp, vp, i64, vp, vp, vp], i32),
            "mpm_mlp_loss": ([vp, vp, vp, vp, i64, ct.c_float, vp, vp], i32),
            "mpm_mlp_backward": ([vp, vp, vp, i64, vp, vp, vp, vp, vp, vp], i32),
            "mpm_mlp_adam": ([vp, vp, vp, vp, vp, i64, ct.c_float, ct.c_float, ct.c_float, ct.c_float, i32], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
        return _lib


def last_error():
    return load().mpm_last_error_string().decode(errors="replace")


def check(rc, what=""):
    """Map a status code to the reference's exception behaviour."""
    if rc == MPM_OK:
        return
    msg = f"{what}: {last_error()}" if what else last_error()
    if rc == MPM_ERR_OOB:
        raise RuntimeError(msg)  # the reference raises a bare RuntimeError (three_d/p2g.py:51-52)
    if rc == MPM_ERR_ARG:
        raise ValueError(msg)
    raise RuntimeError(msg)


def require_gpu():
    """Raise unless the HIP library is loadable and a GPU is visible."""
    import torch

    load()
    if not torch.cuda.is_available():
        raise RuntimeError("femflow_amd needs a ROCm GPU (no CPU fallback)")
