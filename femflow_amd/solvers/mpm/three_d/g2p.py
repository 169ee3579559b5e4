"""3D G2P gather (femflow/solvers/mpm/three_d/g2p.py:9-59) on the GPU."""
from __future__ import annotations

from .. import _phase


def g2p(inv_dx, dt, grid_velocity, particles, v, F, C, Jp, model: str = "neo_hookean"):
    """Gather v and C, advect positions, update F (snow: singular values
    clamped to [1-2.5e-2, 1+7.5e-3] and Jp updated), all in place."""
    res = int(grid_velocity.shape[0]) - 1
    sp = _phase.stage_3d(particles, v, F, C, Jp)
    prm = _phase.make_params(3, res, inv_dx, 1.0 / inv_dx, dt)
    n_oob = _phase.run_g2p(sp, prm, _phase.g2p_kind(model), grid_velocity)
    sp.commit()
    if n_oob:
        raise RuntimeError(f"{n_oob} particle(s) outside the grid")
