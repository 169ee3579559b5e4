"""3D grid update (femflow/solvers/mpm/three_d/grid_op.py:5-47) on the GPU."""
from __future__ import annotations

from .. import _phase


def grid_op(grid_resolution, dx, dt, gravity, grid_velocity, grid_mass):
    """Momentum -> velocity, gravity on axis 1, clamp to 0.9 dx/dt, slip walls
    one node thick at the low side and two at the high side (in place)."""
    prm = _phase.make_params(3, grid_resolution, 1.0 / dx if dx else 0.0, dx, dt, 0.0, gravity)
    _phase.run_grid_op(prm, grid_velocity, grid_mass)
