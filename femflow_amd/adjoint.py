"""Differentiable rollouts of the explicit substep (hand-written adjoint).

``rollout(x0, v0, mu0, lam0, mass, cfg)`` runs ``cfg.nsteps`` elastic
(fixed-corotated) 3D substeps of mls_mpm.py:40-79 in FP32 on the GPU and
returns the final positions and velocities.  It is a
``torch.autograd.Function``: the backward pass is the library's adjoint of
one substep (``mpm_adjoint_step``, on dense (res+1)^3 grids), applied in
reverse order.  The forward substeps run on the particle-resident engine and
keep a checkpoint of the particle state every ``cfg.checkpoint_every``
substeps; the backward replays one checkpoint interval at a time on the
engine, keeping that interval's per-substep states (25 floats per particle
per substep) and dense forward grids (7 floats per node per substep, from
``mpm_run_states_grids``), and reverses it with ``mpm_adjoint_step_grids``.
When the record of the whole rollout fits in ``record_budget_gib`` and in
80% of the free device memory (10 GiB for 2^17 particles, a 64^3 grid and
500 substeps), the forward itself keeps it and the backward replays
nothing.  When no input needs a gradient the forward keeps nothing.  Gradients flow to x0, v0, F0, C0, mu0 and lambda0 (per particle
or scalar); mass is a constant.

The reference has no gradients; the forward it differentiates is its step
(three_d/p2g.py:14-80, three_d/grid_op.py:5-47, three_d/g2p.py:9-59 with the
stress of utils.py:95-135 and constant hardening, utils.py:7-24).
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass
from typing import Optional, Tuple

import torch

from . import _lib
from ._arrays import stream_ptr


@dataclass
class RolloutConfig:
    res: int
    dt: float
    volume: float
    nsteps: int
    gravity: float = -9.8
    hardening: float = 1.0
    checkpoint_every: int = 50
    # when the whole forward record (24 floats per particle and 7 per grid
    # node, per substep) fits in this many GiB of device memory, the forward
    # keeps it and the backward replays nothing; otherwise checkpoints every
    # `checkpoint_every` substeps and one replay per interval
    record_budget_gib: float = 32.0
    boundary: str = "reference"
    bound: int = 3
    quirks: str = "reference"

    def params(self) -> _lib.Params:
        inv_dx = float(self.res)
        return _lib.Params(3, int(self.res), inv_dx, 1.0 / inv_dx, float(self.dt), float(self.volume),
                           float(self.gravity),
                           _lib.MPM_BC_REFERENCE if self.boundary == "reference" else _lib.MPM_BC_STICKY,
                           int(self.bound),
                           _lib.MPM_QUIRKS_REFERENCE if self.quirks == "reference" else _lib.MPM_QUIRKS_TEXTBOOK,
                           1)


def _f32(t: torch.Tensor, device) -> torch.Tensor:
    return t.detach().to(device=device, dtype=torch.float32).contiguous()


class DenseStepper:
    """The dense-grid substep and its adjoint for one particle count."""

    def __init__(self, cfg: RolloutConfig, n: int, device):
        _lib.require_gpu()
        self.L = _lib.load()
        self.cfg = cfg
        self.n = int(n)
        self.device = torch.device(device)
        self.prm = cfg.params()
        nbytes = ct.c_int64()
        _lib.check(self.L.mpm_dense_workspace_bytes(ct.byref(self.prm), ct.byref(nbytes)), "workspace")
        self.work = torch.empty((nbytes.value + 3) // 4, dtype=torch.float32, device=self.device)
        self.oob = torch.zeros(1, dtype=torch.int64, device=self.device)
        self.Jp = torch.ones(max(self.n, 1), dtype=torch.float32, device=self.device)

    def _particles(self, x, v, F, C, mass, mu0, lam0) -> _lib.Particles:
        scalar = mass.numel() == 1 and mu0.numel() == 1 and lam0.numel() == 1
        if not scalar and not (mass.numel() == mu0.numel() == lam0.numel() == self.n):
            raise ValueError("mass, mu0 and lambda0 must all be scalars or all have one entry per particle")
        return _lib.Particles(self.n, _lib.MPM_F32, 1 if scalar else 0, x.data_ptr(), v.data_ptr(),
                              F.data_ptr(), C.data_ptr(), self.Jp.data_ptr(), mass.data_ptr(), mu0.data_ptr(),
                              lam0.data_ptr(), None)

    def step(self, x, v, F, C, mass, mu0, lam0):
        p = self._particles(x, v, F, C, mass, mu0, lam0)
        _lib.check(self.L.mpm_dense_substep(stream_ptr(self.device), ct.byref(self.prm), ct.byref(p),
                                            float(self.cfg.hardening), self.work.data_ptr(),
                                            self.oob.data_ptr()), "mpm_dense_substep")

    def adjoint(self, x, v, F, C, mass, mu0, lam0, xb, vb, Fb, Cb, mub, lamb):
        p = self._particles(x, v, F, C, mass, mu0, lam0)
        _lib.check(self.L.mpm_adjoint_step(stream_ptr(self.device), ct.byref(self.prm), ct.byref(p),
                                           float(self.cfg.hardening), self.work.data_ptr(), xb.data_ptr(),
                                           vb.data_ptr(), Fb.data_ptr(), Cb.data_ptr(), mub.data_ptr(),
                                           lamb.data_ptr(), self.oob.data_ptr()), "mpm_adjoint_step")

    def adjoint_grids(self, x, v, F, C, mass, mu0, lam0, grids, xb, vb, Fb, Cb, mub, lamb):
        """adjoint() with the substep's forward grids given (one slot of
        EngineStepper.replay) instead of recomputed from the state."""
        p = self._particles(x, v, F, C, mass, mu0, lam0)
        _lib.check(self.L.mpm_adjoint_step_grids(stream_ptr(self.device), ct.byref(self.prm), ct.byref(p),
                                                 float(self.cfg.hardening), grids.data_ptr(), self.work.data_ptr(),
                                                 xb.data_ptr(), vb.data_ptr(), Fb.data_ptr(), Cb.data_ptr(),
                                                 mub.data_ptr(), lamb.data_ptr(), self.oob.data_ptr()),
                   "mpm_adjoint_step_grids")

    def check_oob(self):
        n = int(self.oob.item())
        if n:
            self.oob.zero_()
            raise RuntimeError(f"{n} particle stencil(s) left the grid")  # three_d/p2g.py:51-52


class EngineStepper:
    """Forward substeps on the particle-resident engine (the fast path: binned
    particles, LDS arenas); the adjoint of each substep runs on the dense
    grids of DenseStepper from the states this stepper produces."""

    def __init__(self, cfg: RolloutConfig, n: int, device, mass, mu0, lam0):
        from .engine import Engine, MaterialSpec, SceneSpec

        self.device = torch.device(device)
        self.n = int(n)
        self.eng = Engine([SceneSpec(res=cfg.res, dt=cfg.dt, gravity=cfg.gravity, volume=cfg.volume,
                                     boundary=cfg.boundary, bound=cfg.bound, quirks=cfg.quirks)],
                          [MaterialSpec("elastic", cfg.hardening)], n_max=max(self.n, 1), device=self.device)
        self.eng.tune(sort_every=16)
        self.res = int(cfg.res)
        full = lambda t: (t.expand(self.n) if t.numel() == 1 else t.reshape(self.n)).contiguous()  # noqa: E731
        self.mass, self.mu0, self.lam0 = full(mass), full(mu0), full(lam0)
        self.Jp = torch.ones(self.n, dtype=torch.float32, device=self.device)
        self.kind = torch.zeros(self.n, dtype=torch.int32, device=self.device)

    def load(self, x, v, F, C):
        self.eng.upload(x.contiguous(), v.contiguous(), F.contiguous(), C.contiguous(), self.Jp, self.mass,
                        self.mu0, self.lam0, self.kind)

    def run(self, nsteps: int):
        self.eng.run(nsteps, check_oob=False)

    def replay(self, nsteps: int):
        """The next ``nsteps`` substeps in one engine run: the states after
        each substep and each substep's dense forward grids (P2G momentum,
        mass, updated velocity), which the reverse pass reads instead of
        recomputing them."""
        n, d = self.n, self.device
        x, v = torch.empty((nsteps, n, 3), device=d), torch.empty((nsteps, n, 3), device=d)
        F, C = torch.empty((nsteps, n, 3, 3), device=d), torch.empty((nsteps, n, 3, 3), device=d)
        r1 = self.res + 1
        grids = torch.empty((nsteps, 7 * r1 ** 3), device=d)
        self.eng.run_states(nsteps, x, v, F, C, torch.empty((nsteps, n), device=d), check_oob=False, grids=grids)
        return [(x[k], v[k], F[k], C[k]) for k in range(nsteps)], grids

    def state(self):
        n, d = self.n, self.device
        x, v = torch.empty((n, 3), device=d), torch.empty((n, 3), device=d)
        F, C = torch.empty((n, 3, 3), device=d), torch.empty((n, 3, 3), device=d)
        self.eng.download(x, v, F, C, torch.empty(n, device=d))
        return x, v, F, C

    def check_oob(self):
        n = self.eng.oob_count()
        if n:
            raise RuntimeError(f"{n} particle stencil(s) left the grid")  # three_d/p2g.py:51-52

    def close(self):
        self.eng.close()


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, F0, C0, mu0, lam0, mass, cfg: RolloutConfig):
        dev = x0.device
        n = x0.shape[0]
        st = DenseStepper(cfg, n, dev)
        mass_, mu_, lam_ = _f32(mass, dev), _f32(mu0, dev), _f32(lam0, dev)
        fwd = EngineStepper(cfg, n, dev, mass_, mu_, lam_)
        x, v = _f32(x0, dev).clone(), _f32(v0, dev).clone()
        F, C = _f32(F0, dev).clone(), _f32(C0, dev).clone()
        K = max(1, int(cfg.checkpoint_every))
        ckpts = []
        record = None
        fwd.load(x, v, F, C)
        # the replay writes 25 floats per particle (x, v, F, C, Jp) and 7 per
        # grid node per substep; record only when a gradient will be asked
        # for and the record fits both the budget and most of the free memory
        needs_grad = any(ctx.needs_input_grad[:7])
        per_step = 4 * (25 * n + 7 * (cfg.res + 1) ** 3)
        budget = cfg.record_budget_gib * 2 ** 30
        if dev.type == "cuda":
            budget = min(budget, 0.8 * torch.cuda.mem_get_info(dev)[0])
        if not needs_grad:
            fwd.run(cfg.nsteps)
            x, v, _, _ = fwd.state()
        elif cfg.nsteps > 0 and per_step * cfg.nsteps <= budget:
            # one recorded forward: every substep's state and grids stay on
            # the device for the backward
            after, grids = fwd.replay(cfg.nsteps)
            record = ([(x, v, F, C)] + after[:-1], grids)
            x, v = after[-1][0].clone(), after[-1][1].clone()
        else:
            for s0 in range(0, cfg.nsteps, K):
                ckpts.append((x, v, F, C) if s0 == 0 else fwd.state())
                fwd.run(min(K, cfg.nsteps - s0))
            x, v, _, _ = fwd.state()
        fwd.check_oob()
        ctx.stepper, ctx.fwd, ctx.ckpts, ctx.cfg, ctx.record = st, fwd, ckpts, cfg, record
        ctx.mats = (mass_, mu_, lam_)
        ctx.shapes = (mu0.shape, lam0.shape, mu0.dtype, lam0.dtype)
        return x, v

    @staticmethod
    def backward(ctx, gx, gv):
        if ctx.ckpts is None:
            raise RuntimeError("rollout: the backward pass consumes the forward's record and checkpoints, so it "
                               "runs once per forward (call rollout again instead of retain_graph=True)")
        st, cfg = ctx.stepper, ctx.cfg
        mass_, mu_, lam_ = ctx.mats
        dev = st.device
        n = st.n
        xb = _f32(gx, dev).clone() if gx is not None else torch.zeros((n, 3), device=dev)
        vb = _f32(gv, dev).clone() if gv is not None else torch.zeros((n, 3), device=dev)
        Fb = torch.zeros((n, 3, 3), dtype=torch.float32, device=dev)
        Cb = torch.zeros((n, 3, 3), dtype=torch.float32, device=dev)
        mub = torch.zeros_like(mu_)
        lamb = torch.zeros_like(lam_)
        K = max(1, int(cfg.checkpoint_every))
        fwd = ctx.fwd
        if ctx.record is not None:
            states, grids = ctx.record
            for k in range(cfg.nsteps - 1, -1, -1):
                x, v, F, C = states[k]
                st.adjoint_grids(x, v, F, C, mass_, mu_, lam_, grids[k], xb, vb, Fb, Cb, mub, lamb)
            ctx.record = None
            del states, grids
        for seg in range(len(ctx.ckpts) - 1, -1, -1):
            s0 = seg * K
            s1 = min(s0 + K, cfg.nsteps)
            # replay the interval on the engine: each substep's input state and
            # its forward grids
            fwd.load(*ctx.ckpts[seg])
            after, grids = fwd.replay(s1 - s0)
            states = [ctx.ckpts[seg]] + after[:-1]
            for k in range(s1 - s0 - 1, -1, -1):
                x, v, F, C = states[k]
                st.adjoint_grids(x, v, F, C, mass_, mu_, lam_, grids[k], xb, vb, Fb, Cb, mub, lamb)
            del states, after, grids
        fwd.check_oob()
        st.check_oob()
        fwd.close()
        ctx.fwd = None
        mu_shape, lam_shape, mu_dt, lam_dt = ctx.shapes
        ctx.ckpts = None
        return (xb, vb, Fb, Cb, mub.reshape(mu_shape).to(mu_dt), lamb.reshape(lam_shape).to(lam_dt), None, None)


def rollout(x0: torch.Tensor, v0: torch.Tensor, mu0: torch.Tensor, lam0: torch.Tensor, mass: torch.Tensor,
            cfg: RolloutConfig, F0: Optional[torch.Tensor] = None,
            C0: Optional[torch.Tensor] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """Final (x, v) after cfg.nsteps substeps; differentiable in x0, v0, F0,
    C0, mu0 and lam0.  Inputs are CUDA tensors: x0, v0 (n, 3); mu0, lam0,
    mass (n,) or one element; F0, C0 (n, 3, 3), default identity / zero."""
    n = x0.shape[0]
    if F0 is None:
        F0 = torch.eye(3, device=x0.device, dtype=torch.float32).expand(n, 3, 3)
    if C0 is None:
        C0 = torch.zeros((n, 3, 3), device=x0.device, dtype=torch.float32)
    # The scatters aggregate particles of one stencil base within a wave, so
    # the rollout runs in cell order; results (and gradients, through the
    # indexing) come back in the caller's order.
    perm = cell_order(x0.detach(), cfg.res)
    inv = torch.empty_like(perm)
    inv[perm] = torch.arange(n, device=perm.device)
    pick = lambda t: t[perm] if t.numel() > 1 else t  # noqa: E731
    xT, vT = _Rollout.apply(x0[perm], v0[perm], F0[perm], C0[perm], pick(mu0), pick(lam0), pick(mass), cfg)
    return xT[inv], vT[inv]


def cell_order(x: torch.Tensor, res: int) -> torch.Tensor:
    """Stable order of particles by stencil base (row-major cell index of
    trunc(x res - 0.5), three_d/p2g.py:50)."""
    b = torch.clamp((x.double() * res - 0.5).trunc().long(), 0, res)
    key = (b[:, 0] * (res + 1) + b[:, 1]) * (res + 1) + b[:, 2]
    return torch.argsort(key, stable=True)


def lame_from_E(E: torch.Tensor, nu: float):
    """numerics/fem.py:1-6: mu = E / (2 (1 + nu)), lambda = E nu / ((1 + nu)(1 - 2 nu))."""
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
