"""One 3D MLS-MPM substep (femflow/solvers/mpm/mls_mpm.py:9-79) on the GPU."""
from __future__ import annotations

import numpy as np
import torch

from ... import _lib
from ...engine import Engine, MaterialSpec, SceneSpec
from . import _phase


def make_mls_mpm_coefficients(lenx: int, dim: int, device=None, dtype=None):
    """v=0, F=I, C=0, Jp=1 (mls_mpm.py:9-15).  numpy float64 by default, or
    torch tensors when ``device`` is given."""
    if device is None:
        v = np.zeros((lenx, dim), dtype=np.float64)
        F = np.array([np.eye(dim, dtype=np.float64) for _ in range(lenx)]).reshape(lenx, dim, dim)
        C = np.zeros((lenx, dim, dim), dtype=np.float64)
        Jp = np.ones((lenx, 1), dtype=np.float64)
        return v, F, C, Jp
    dtype = dtype or torch.float32
    v = torch.zeros((lenx, dim), dtype=dtype, device=device)
    F = torch.eye(dim, dtype=dtype, device=device).expand(lenx, dim, dim).contiguous()
    C = torch.zeros((lenx, dim, dim), dtype=dtype, device=device)
    Jp = torch.ones((lenx, 1), dtype=dtype, device=device)
    return v, F, C, Jp


_ENGINES = {}


def _engine_for(res, dt, gravity, volume, hardening, n, device, boundary="reference", bound=3,
                quirks="reference"):
    key = (int(res), float(dt), float(gravity), float(volume), float(hardening), str(device), boundary,
           int(bound), quirks)
    eng = _ENGINES.get(key)
    if eng is None or eng.n_max < n:
        if eng is not None:
            eng.close()
        eng = Engine([SceneSpec(res=res, dt=dt, gravity=gravity, volume=volume, boundary=boundary,
                                bound=int(bound), quirks=quirks)],
                     [MaterialSpec("elastic", hardening)], n_max=max(n, 1), device=device)
        eng.tune(sort_every=1)
        _ENGINES[key] = eng
    return eng


def solve_mls_mpm_3d(res: int, inv_dx: float, hardening: float, dx: float, dt: float, volume: float,
                     gravity: float, particles, v, F, C, Jp, *, boundary: str = "reference", bound: int = 3,
                     quirks: str = "reference"):
    """mls_mpm.py:40-79: zero grids, P2G ("neo_hookean"), grid update, G2P.

    Mutates the particle positions and v, F, C, Jp in place.  FP64 inputs run
    the FP64 phase kernels on a dense (res+1)^3 grid; FP32 CUDA tensors run
    the FP32 engine (binned, LDS-staged transfers) for one substep.  The
    keyword-only options select the sticky walls (boundary="sticky", `bound`
    nodes thick) and the textbook stress (quirks="textbook"); their defaults
    are the reference's behaviour."""
    bc = _phase.boundary_code(boundary)
    q = _phase.quirks_code(quirks)
    sp = _phase.stage_3d(particles, v, F, C, Jp)
    if sp.dtype == torch.float64:
        dres = res + 1
        gv = torch.zeros((dres, dres, dres, 3), dtype=torch.float64, device=sp.device)
        gm = torch.zeros((dres, dres, dres, 1), dtype=torch.float64, device=sp.device)
        prm = _phase.make_params(3, res, inv_dx, dx, dt, volume, gravity, boundary=bc, bound=int(bound),
                                 quirks=q)
        n_oob = _phase.run_p2g(sp, prm, _lib.MPM_ELASTIC, hardening, gv, gm)
        if n_oob:
            raise RuntimeError(f"{n_oob} particle(s) outside the grid")
        _phase.run_grid_op(prm, gv, gm)
        n_oob = _phase.run_g2p(sp, prm, _lib.MPM_ELASTIC, gv)
        if n_oob:
            raise RuntimeError(f"{n_oob} particle(s) outside the grid")
        sp.commit()
        return
    if abs(inv_dx * dx - 1.0) > 1e-9 or abs(inv_dx - res) > 1e-9 * res:
        raise ValueError("the FP32 engine runs on [0,1]^3 with dx = 1/res")
    eng = _engine_for(res, dt, gravity, volume, hardening, sp.n, sp.device, boundary, bound, quirks)
    jp = sp.Jp.tensor.view(-1)
    eng.upload(sp.x.tensor, sp.v.tensor, sp.F.tensor, sp.C.tensor, jp, sp.mass.tensor.contiguous(),
               sp.mu0.tensor.contiguous(), sp.lam0.tensor.contiguous())
    eng.run(1)
    eng.download(sp.x.tensor, sp.v.tensor, sp.F.tensor, sp.C.tensor, jp)
    sp.commit()
