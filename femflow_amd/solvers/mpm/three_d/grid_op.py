"""3D grid update (femflow/solvers/mpm/three_d/grid_op.py:5-47) and the
plane colliders of the same file (:50-67) on the GPU."""
from __future__ import annotations

from .. import _phase


def grid_op(grid_resolution, dx, dt, gravity, grid_velocity, grid_mass, *, boundary: str = "reference",
            bound: int = 3):
    """Momentum -> velocity, gravity on axis 1, clamp to 0.9 dx/dt, slip walls
    one node thick at the low side and two at the high side (in place).
    boundary="sticky" zeroes every component within `bound` nodes of a wall
    instead of the slip walls."""
    bc = _phase.boundary_code(boundary)
    prm = _phase.make_params(3, grid_resolution, 1.0 / dx if dx else 0.0, dx, dt, 0.0, gravity,
                             boundary=bc, bound=int(bound))
    _phase.run_grid_op(prm, grid_velocity, grid_mass)


def check_collision_points(points, normals, grid_resolution, dx, grid_velocity, *, quirks="reference"):
    """Sticky half-space colliders (three_d/grid_op.py:50-67), in place: for
    each (point, normal), every node I with dot(I dx - point, n) < 0 gets
    zero velocity.  quirks="reference" uses n = normal + 1/|normal| (the
    reference adds the scalar to every component); "textbook" uses the unit
    normal.  points, normals: (K, 3), K <= 16; grid_velocity: the dense
    (R+1, R+1, R+1, 3) grid (numpy or a CUDA tensor, FP32 or FP64)."""
    _phase.run_collide_planes(points, normals, grid_resolution, dx, grid_velocity, quirks)
