"""3D P2G scatter (femflow/solvers/mpm/three_d/p2g.py:14-80) on the GPU."""
from __future__ import annotations

from .. import _phase


def p2g(inv_dx, hardening, dx, dt, volume, grid_velocity, grid_mass, particles, v, F, C, Jp,
        model: str = "neo_hookean", *, quirks: str = "reference"):
    """Scatter momentum and mass of every particle to the (res+1)^3 grid.

    Same arguments and in-place contract as the reference; raises
    RuntimeError when a particle's stencil leaves the grid (p2g.py:51-52).
    quirks="textbook" adds the lambda term of the stress to the diagonal only
    (the reference broadcasts it to every entry, utils.py:129)."""
    res = int(grid_velocity.shape[0]) - 1
    q = _phase.quirks_code(quirks)
    sp = _phase.stage_3d(particles, v, F, C, Jp)
    prm = _phase.make_params(3, res, inv_dx, dx, dt, volume, quirks=q)
    n_oob = _phase.run_p2g(sp, prm, _phase.p2g_kind(model), hardening, grid_velocity, grid_mass)
    if n_oob:
        raise RuntimeError(f"{n_oob} particle(s) outside the grid")
