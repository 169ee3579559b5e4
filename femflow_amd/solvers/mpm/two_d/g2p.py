"""2D G2P gather (femflow/solvers/mpm/two_d/g2p.py:5-47) on the GPU."""
from __future__ import annotations

import torch

from .. import _phase


def g2p(inv_dx, dt, grid_velocity, x, v, F, C, Jp, model: str = "neo_hookean"):
    """Gather, advect, update F; the SVD rebuild and Jp update run for every
    particle, the singular value clamp only for ``model == "snow"``."""
    res = int(grid_velocity.shape[0]) - 1
    dev = _phase.device_of(x, v, F)
    dt_ = _phase.compute_dtype(v, F, x)
    one = torch.ones(1, dtype=dt_, device=dev)
    sp = _phase.StagedParticles(x, v, F, C, Jp, one, one, one, mat_scalar=True)
    prm = _phase.make_params(2, res, inv_dx, 1.0 / inv_dx, dt, check_oob=False)
    _phase.run_g2p(sp, prm, _phase.g2p_kind(model), grid_velocity)
    sp.commit()
