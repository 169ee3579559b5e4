"""2D grid update (femflow/solvers/mpm/two_d/grid_op.py:5-24) on the GPU."""
from __future__ import annotations

from .. import _phase


def grid_op(grid_resolution, dt, gravity, grid_velocity, grid_mass, *, boundary: str = "reference",
            bound: int = 3):
    """Nodes with mass: v = p/m, gravity on axis 1, sticky side and top walls
    (5% of the domain), floor keeps v_y >= 0 (in place).  boundary="sticky"
    instead zeroes nodes with mass within `bound` nodes of any wall."""
    bc = _phase.boundary_code(boundary)
    prm = _phase.make_params(2, grid_resolution, float(grid_resolution), 1.0 / grid_resolution, dt,
                             0.0, gravity, boundary=bc, bound=int(bound), check_oob=False)
    _phase.run_grid_op(prm, grid_velocity, grid_mass)
