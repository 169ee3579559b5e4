"""Particle-resident MLS-MPM engine (many scenes, block-sparse grids, FP32).

Wraps the ``mpm_ctx`` half of the C ABI.  One ``Engine`` owns a binned copy of
the particle state on one GPU and advances every scene by whole substeps of
the reference step (femflow/solvers/mpm/mls_mpm.py:40-79): P2G, grid update,
G2P.  Between rebinning points the G2P of one substep and the P2G of the next
run fused in one kernel; the sequence is captured once as a hipGraph and
replayed.
"""
from __future__ import annotations

import ctypes as ct
import sys
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib

KINDS = {"elastic": _lib.MPM_ELASTIC, "snow": _lib.MPM_SNOW, "water": _lib.MPM_WATER}


@dataclass
class SceneSpec:
    """One scene: a (res+1)^3 grid on [0,1]^3 with its own time step."""

    res: int
    dt: float
    gravity: float = -9.8
    volume: float = 1.0
    boundary: str = "reference"  # "reference" (three_d/grid_op.py:37-47) or "sticky"
    bound: int = 3
    quirks: str = "reference"     # "reference" (utils.py:129 broadcast) or "textbook"


@dataclass
class MaterialSpec:
    kind: str = "elastic"  # elastic | snow | water
    hardening: float = 1.0
    theta_c: float = 2.5e-2
    theta_s: float = 7.5e-3


class Engine:
    def __init__(self, scenes: Sequence[SceneSpec], materials: Sequence[MaterialSpec], n_max: int,
                 device=None):
        bcs = {"reference": _lib.MPM_BC_REFERENCE, "sticky": _lib.MPM_BC_STICKY}
        qks = {"reference": _lib.MPM_QUIRKS_REFERENCE, "textbook": _lib.MPM_QUIRKS_TEXTBOOK}
        for i, s in enumerate(scenes):
            if s.boundary not in bcs or s.quirks not in qks:
                raise ValueError(f"scene {i}: boundary must be 'reference' or 'sticky' and quirks "
                                 f"'reference' or 'textbook' (got {s.boundary!r}, {s.quirks!r})")
        _lib.require_gpu()
        self._L = _lib.load()
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.scenes = list(scenes)
        self.materials = list(materials)
        sc = (_lib.Scene * len(self.scenes))()
        for i, s in enumerate(self.scenes):
            sc[i] = _lib.Scene(int(s.res), bcs[s.boundary], int(s.bound), qks[s.quirks],
                               float(s.dt), float(s.gravity), float(s.volume))
        mt = (_lib.Material * len(self.materials))()
        for i, m in enumerate(self.materials):
            mt[i] = _lib.Material(KINDS[m.kind] if isinstance(m.kind, str) else int(m.kind),
                                  float(m.hardening), float(m.theta_c), float(m.theta_s))
        h = ct.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self._L.mpm_ctx_create(ct.byref(h), int(n_max), len(self.scenes), sc,
                                              len(self.materials), mt, 0), "mpm_ctx_create")
        self._h = h
        self.n_max = int(n_max)
        self.n = 0
        self._material = None
        self._scene = None
        self.tune()

    # ------------------------------------------------------------------
    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _sync_stream(self):
        _lib.check(self._L.mpm_ctx_set_stream(self._h, self._stream()), "set_stream")

    def tune(self, chunk: int = 512, sort_every: int = 48, grid_mult: int = 8, use_graph: bool = True,
             spread: int = 9):
        """chunk: binned particles per work item (256, 512, 1024); sort_every:
        substeps between rebinning; grid_mult: persistent workgroups per CU;
        use_graph: replay the substep sequence as a captured hipGraph;
        spread: scatter start offsets per cell (9, or 3: the narrower
        rotation, slower where many particles share a cell)."""
        self.config = dict(chunk=chunk, sort_every=sort_every, grid_mult=grid_mult, use_graph=use_graph,
                           spread=spread)
        _lib.check(self._L.mpm_ctx_tune(self._h, chunk, sort_every, grid_mult, 1 if use_graph else 0),
                   "tune")
        _lib.check(self._L.mpm_ctx_set_spread(self._h, int(spread)), "tune")

    def _particles(self, x, v, F, C, Jp, mass=None, mu0=None, lam0=None, material=None):
        def ptr(t):
            return None if t is None else t.data_ptr()

        return _lib.Particles(int(x.shape[0]), _lib.MPM_F32, 0, ptr(x), ptr(v), ptr(F), ptr(C), ptr(Jp),
                              ptr(mass), ptr(mu0), ptr(lam0), ptr(material))

    @staticmethod
    def _check_tensor(t, name, dtype=torch.float32):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype and t.is_contiguous()):
            raise ValueError(f"{name} must be a contiguous CUDA {dtype} tensor")

    def upload(self, x, v, F, C, Jp, mass, mu0, lam0, material=None, scene=None):
        """Load FP32 CUDA tensors in the reference layouts (x, v: (N,3); F, C:
        (N,3,3); Jp: (N,) or (N,1); mass/mu0/lam0: (N,)); material/scene: int32 (N,)."""
        for t, nm in ((x, "x"), (v, "v"), (F, "F"), (C, "C"), (Jp, "Jp"), (mass, "mass"), (mu0, "mu0"),
                      (lam0, "lam0")):
            self._check_tensor(t, nm)
        for t, nm, hi in ((material, "material", len(self.materials)), (scene, "scene", len(self.scenes))):
            if t is not None:
                self._check_tensor(t, nm, torch.int32)
                if t.numel() and (int(t.min()) < 0 or int(t.max()) >= hi):
                    raise ValueError(f"{nm} ids must lie in [0, {hi})")
        if int(x.shape[0]) > self.n_max:
            raise ValueError(f"{x.shape[0]} particles exceed n_max={self.n_max}")
        self._sync_stream()
        p = self._particles(x, v, F, C, Jp, mass, mu0, lam0, material)
        _lib.check(self._L.mpm_state_upload(self._h, ct.byref(p),
                                            None if scene is None else scene.data_ptr()), "upload")
        self.n = int(x.shape[0])

    def download(self, x, v, F, C, Jp):
        for t, nm in ((x, "x"), (v, "v"), (F, "F"), (C, "C"), (Jp, "Jp")):
            self._check_tensor(t, nm)
        self._sync_stream()
        p = self._particles(x, v, F, C, Jp)
        _lib.check(self._L.mpm_state_download(self._h, ct.byref(p)), "download")

    def run(self, nsteps: int, sums: Optional[torch.Tensor] = None, check_oob: bool = True):
        """Advance every scene by ``nsteps`` substeps.  ``sums``: optional FP64
        CUDA tensor (nsteps, 8) receiving per-substep grid conservation sums
        [mass, momentum after P2G (3), momentum after the grid update (3), nodes with mass]."""
        self._sync_stream()
        if sums is not None:
            self._check_tensor(sums, "sums", torch.float64)
            if sums.numel() < nsteps * 8:
                raise ValueError("sums needs nsteps*8 entries")
        _lib.check(self._L.mpm_run(self._h, int(nsteps), None if sums is None else sums.data_ptr()),
                   "run")
        if check_oob:
            n = self.oob_count()
            if n >= 1 << 40:  # the grid update's wait for the FP64 pass ran out
                raise RuntimeError("the FP64 pass for deferred particles did not finish in time")
            if n:
                raise RuntimeError(f"{n} particle stencil(s) left the grid")

    def run_states(self, nsteps: int, x, v, F, C, Jp, check_oob: bool = True, grids=None):
        """Advance ``nsteps`` substeps, writing the full state after substep k
        (caller order) to slot k: x, v (nsteps, N, 3), F, C (nsteps, N, 3, 3),
        Jp (nsteps, N); float32 CUDA tensors.  grids (single-scene engines):
        float32 (nsteps, 7 (res+1)^3) receiving each substep's dense P2G
        momentum, P2G mass and updated velocity."""
        n = self.n
        for t, nm, shape in ((x, "x", (nsteps, n, 3)), (v, "v", (nsteps, n, 3)), (F, "F", (nsteps, n, 3, 3)),
                             (C, "C", (nsteps, n, 3, 3)), (Jp, "Jp", (nsteps, n))):
            self._check_tensor(t, nm)
            if tuple(t.shape) != shape:
                raise ValueError(f"{nm} must have shape {shape}")
        self._sync_stream()
        if grids is None:
            _lib.check(self._L.mpm_run_states(self._h, int(nsteps), x.data_ptr(), v.data_ptr(), F.data_ptr(),
                                              C.data_ptr(), Jp.data_ptr()), "run_states")
        else:
            self._check_tensor(grids, "grids")
            r1 = self.scenes[0].res + 1
            if grids.numel() != nsteps * 7 * r1 ** 3:
                raise ValueError(f"grids must hold {nsteps} x 7 x {r1}^3 floats")
            _lib.check(self._L.mpm_run_states_grids(self._h, int(nsteps), x.data_ptr(), v.data_ptr(),
                                                    F.data_ptr(), C.data_ptr(), Jp.data_ptr(), grids.data_ptr()),
                       "run_states")
        if check_oob:
            k = self.oob_count()
            if k:
                raise RuntimeError(f"{k} particle stencil(s) left the grid")

    def run_frames(self, nsteps: int, frames: torch.Tensor, check_oob: bool = True):
        """Advance ``nsteps`` substeps, writing each substep's positions (caller
        order) into ``frames``, a float32 CUDA tensor (nsteps, N, 3)."""
        self._check_tensor(frames, "frames")
        if frames.numel() < nsteps * self.n * 3:
            raise ValueError("frames needs nsteps * N * 3 entries")
        self._sync_stream()
        _lib.check(self._L.mpm_run_frames(self._h, int(nsteps), frames.data_ptr()), "run_frames")
        if check_oob:
            n = self.oob_count()
            if n:
                raise RuntimeError(f"{n} particle stencil(s) left the grid")

    def positions(self, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Current positions in the caller's order, (N, 3) float32 on the device."""
        if out is None:
            out = torch.empty((self.n, 3), dtype=torch.float32, device=self.device)
        self._check_tensor(out, "out")
        self._sync_stream()
        _lib.check(self._L.mpm_state_positions(self._h, out.data_ptr()), "positions")
        return out

    def oob_count(self) -> int:
        out = ct.c_int64()
        _lib.check(self._L.mpm_ctx_oob_count(self._h, ct.byref(out)), "oob_count")
        return int(out.value)

    def fallback_count(self) -> int:
        out = ct.c_int64()
        _lib.check(self._L.mpm_ctx_fallback_count(self._h, ct.byref(out)), "fallback_count")
        return int(out.value)

    def active_block_count(self) -> int:
        """Grid blocks (4x4x4 nodes) the last substep's P2G reached."""
        out = ct.c_int64()
        _lib.check(self._L.mpm_ctx_active_blocks(self._h, ct.byref(out)), "active_block_count")
        return int(out.value)

    def request_rebin(self):
        """Rebin the particles at the start of the next run (the engine otherwise
        rebins after an upload and then every sort_every substeps, across calls)."""
        _lib.check(self._L.mpm_ctx_request_rebin(self._h), "request_rebin")

    def particle_sums(self):
        out = (ct.c_double * 7)()
        _lib.check(self._L.mpm_particle_sums(self._h, out), "particle_sums")
        return list(out)

    # -- plane colliders ----------------------------------------------------
    def set_planes(self, points, normals, scene: int = 0, quirks: str = "reference"):
        """Sticky half-space colliders for one scene (three_d/grid_op.py:50-67),
        applied by every grid update after the walls; replaces the scene's
        previous planes (empty arrays remove them).  points, normals: (K, 3)
        in the unit-cube coordinates of the grid; at most 16 per engine."""
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
        nrm = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).reshape(-1, 3))
        if pts.shape != nrm.shape:
            raise ValueError("points and normals must have the same number of rows")
        q = _lib.MPM_QUIRKS_TEXTBOOK if quirks == "textbook" else _lib.MPM_QUIRKS_REFERENCE
        self._sync_stream()
        _lib.check(self._L.mpm_ctx_set_planes(self._h, int(scene), len(pts), pts.ctypes.data if len(pts) else None,
                                              nrm.ctypes.data if len(nrm) else None, q), "set_planes")

    # -- learned corrector --------------------------------------------------
    def set_corrector(self, params: Optional[torch.Tensor], m_ref: float = 1.0, v_ref: float = 1.0,
                      enable: bool = True):
        """Turn the learned grid corrector on (params: FP32 CUDA MLP parameters,
        femflow_amd.corrector layout) or off (enable=False)."""
        self._sync_stream()
        ptr = None
        if params is not None:
            self._check_tensor(params, "params")
            ptr = params.data_ptr()
        _lib.check(self._L.mpm_ctx_set_corrector(self._h, ptr, float(m_ref), float(v_ref), 1 if enable else 0),
                   "set_corrector")

    def corrector_sample(self, max_rows: int):
        """Features of every node with mass after one P2G + grid update from the
        current state (particles unchanged): (X (rows, 352) bf16, node ids
        (rows,) int32 view of uint32, updated velocity (rows, 3))."""
        self._sync_stream()
        X = torch.empty((max_rows, 352), dtype=torch.bfloat16, device=self.device)
        nodes = torch.empty(max_rows, dtype=torch.int32, device=self.device)
        vel = torch.empty((max_rows, 3), dtype=torch.float32, device=self.device)
        rows = ct.c_int64()
        _lib.check(self._L.mpm_corrector_sample(self._h, X.data_ptr(), nodes.data_ptr(), vel.data_ptr(),
                                                int(max_rows), ct.byref(rows)), "corrector_sample")
        r = min(rows.value, max_rows)
        return X[:r], nodes[:r], vel[:r], rows.value

    def p2g_grid_only(self):
        """One P2G + grid update from the current state; particles unchanged."""
        self._sync_stream()
        rows = ct.c_int64()
        _lib.check(self._L.mpm_corrector_sample(self._h, None, None, None, 0, ct.byref(rows)), "p2g_grid_only")

    def grid_velocity(self, nodes: torch.Tensor) -> torch.Tensor:
        self._check_tensor(nodes, "nodes", torch.int32)
        out = torch.empty((nodes.numel(), 3), dtype=torch.float32, device=self.device)
        _lib.check(self._L.mpm_grid_velocity(self._h, nodes.data_ptr(), nodes.numel(), out.data_ptr()),
                   "grid_velocity")
        return out

    def kernel_timing(self, enable: bool = True):
        _lib.check(self._L.mpm_ctx_kernel_timing(self._h, 1 if enable else 0), "kernel_timing")

    def kernel_times(self):
        """{name: (total_ms, launches)} since the previous call (timing mode).
        "p2g" is the pass that starts a segment: a run's first P2G, or, after
        a rebin inside a run, the fused pass that moves the particles."""
        ms = (ct.c_double * 4)()
        cnt = (ct.c_int64 * 4)()
        _lib.check(self._L.mpm_ctx_kernel_times(self._h, ms, cnt), "kernel_times")
        names = ("p2g", "g2p2g", "g2p", "grid_update")
        return {nm: (ms[i], int(cnt[i])) for i, nm in enumerate(names)}

    def close(self):
        if getattr(self, "_h", None):
            self._L.mpm_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def bin_particles(x: torch.Tensor, res: int, inv_dx: float, layout: str = "blocked"):
    """Cell keys and their stable ordering for FP32 CUDA positions (N, d)."""
    _lib.require_gpu()
    L = _lib.load()
    if not (x.is_cuda and x.dtype == torch.float32 and x.is_contiguous() and x.dim() == 2):
        raise ValueError("x must be a contiguous CUDA float32 (N, d) tensor")
    n, dim = x.shape
    keys = torch.empty(n, dtype=torch.int32, device=x.device)
    perm = torch.empty(n, dtype=torch.int32, device=x.device)
    torch.cuda.current_stream(x.device).synchronize()
    lay = _lib.MPM_KEY_BLOCKED if layout == "blocked" else _lib.MPM_KEY_ROWMAJOR
    _lib.check(L.mpm_bin(None, int(dim), int(res), float(inv_dx), lay, x.data_ptr(), int(n),
                         keys.data_ptr(), perm.data_ptr()), "mpm_bin")
    return keys, perm
