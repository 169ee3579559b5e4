"""Differentiable FP64 restatement of the elastic 3D substep (test
infrastructure only: the checker for the adjoint, never the thing measured).

``substep`` is the same step as ``mpm_oracle.substep`` for ELASTIC particles
(three_d/p2g.py:14-80 with utils.py:95-135 and constant hardening
utils.py:7-24, three_d/grid_op.py:5-47, three_d/g2p.py:9-59), written with
torch operations on CPU float64 so that torch.autograd yields the reference
gradient.  The polar factor comes from the Newton iteration
R <- (R + R^-T) / 2 started at F (converged to rounding after ~10 steps from
near-rotations, differentiable by unrolling); at the converged R it agrees
with the SVD-based polar factor of linear_algebra.py:119-135.  The stencil
base is held fixed (piecewise differentiation), as in the GPU adjoint.
Pinned by tests/test_adjoint.py against mpm_oracle.substep (forward) and
against central differences (gradient).
"""
from __future__ import annotations

import itertools

import torch


def polar_newton(F: torch.Tensor, iters: int = 20) -> torch.Tensor:
    R = F
    for _ in range(iters):
        R = 0.5 * (R + torch.linalg.inv(R).transpose(-1, -2))
    return R


def substep(x, v, F, C, mass, mu0, lam0, res, dt, volume, gravity=-9.8, hardening=1.0,
            quirks="reference", boundary="reference", bound=3, newton_iters=20):
    """One substep; returns (x', v', F', C').  mass, mu0, lam0: (n,) or scalars."""
    inv_dx = float(res)
    dx = 1.0 / inv_dx
    n = x.shape[0]
    r1 = res + 1
    dev, dt_ = x.device, x.dtype
    s = x * inv_dx
    base = torch.trunc((s - 0.5).detach()).long()
    fx = s - base
    w = [0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2]
    m = torch.as_tensor(mass, dtype=dt_, device=dev).expand(n)
    mu = torch.as_tensor(mu0, dtype=dt_, device=dev).expand(n) * hardening
    lam = torch.as_tensor(lam0, dtype=dt_, device=dev).expand(n) * hardening
    # stress (utils.py:119-135)
    R = polar_newton(F, newton_iters)
    J = torch.linalg.det(F)
    if quirks == "reference":
        M = torch.ones(3, 3, dtype=dt_, device=dev)
    else:
        M = torch.eye(3, dtype=dt_, device=dev)
    PF = 2.0 * mu[:, None, None] * (F - R) @ F.transpose(-1, -2) + (lam * (J - 1.0) * J)[:, None, None] * M
    A = -(dt * volume) * (4.0 * inv_dx * inv_dx) * PF + m[:, None, None] * C
    mv = m[:, None] * v
    # P2G
    gp = torch.zeros((r1 ** 3, 3), dtype=dt_, device=dev)
    gm = torch.zeros(r1 ** 3, dtype=dt_, device=dev)
    offs = list(itertools.product(range(3), repeat=3))
    W, idx, dpos = [], [], []
    for (i, j, k) in offs:
        o = torch.tensor([i, j, k], dtype=dt_, device=dev)
        Wn = w[i][:, 0] * w[j][:, 1] * w[k][:, 2]
        node = base + torch.tensor([i, j, k], device=dev)
        fl = (node[:, 0] * r1 + node[:, 1]) * r1 + node[:, 2]
        dp = o[None, :] - fx
        W.append(Wn)
        idx.append(fl)
        dpos.append(dp)
        contrib = Wn[:, None] * (mv + (A @ (dp * dx)[..., None])[..., 0])
        gp = gp.index_add(0, fl, contrib)
        gm = gm.index_add(0, fl, Wn * m)
    # grid update (three_d/grid_op.py:5-47)
    va = dx * 0.9 / dt
    pos = gm > 0
    safe = torch.where(pos, gm, torch.ones_like(gm))
    u = torch.where(pos[:, None], gp / safe[:, None], gp)
    grav = torch.zeros(3, dtype=dt_, device=dev)
    grav[1] = dt * gravity
    u = torch.where(pos[:, None], torch.clamp(u + grav, -va, va), u)
    ids = torch.arange(r1 ** 3, device=dev)
    I = torch.stack([ids // (r1 * r1), (ids // r1) % r1, ids % r1], dim=1)
    if boundary == "reference":
        keep = ~((I < 1) | (I >= res - 1))
        u = u * keep.to(dt_)
    else:
        wall = ((I < bound) | (I > res - bound)).any(dim=1)
        u = u * (~wall).to(dt_)[:, None]
    # G2P (three_d/g2p.py:29-46)
    vn = torch.zeros((n, 3), dtype=dt_, device=dev)
    Cn = torch.zeros((n, 3, 3), dtype=dt_, device=dev)
    for Wn, fl, dp in zip(W, idx, dpos):
        g = u[fl]
        vn = vn + Wn[:, None] * g
        Cn = Cn + 4.0 * inv_dx * (Wn[:, None] * g)[:, :, None] * dp[:, None, :]
    xn = x + dt * vn
    Fn = (torch.eye(3, dtype=dt_, device=dev)[None] + dt * Cn) @ F
    return xn, vn, Fn, Cn


def rollout(x, v, F, C, mass, mu0, lam0, res, dt, volume, nsteps, **kw):
    for _ in range(nsteps):
        x, v, F, C = substep(x, v, F, C, mass, mu0, lam0, res, dt, volume, **kw)
    return x, v, F, C
