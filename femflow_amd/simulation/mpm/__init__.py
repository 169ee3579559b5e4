from .primitives import (diamond, generate_cube_points, generate_implicit_points, grid, gyroid,
                         primitive)
from .simulation import MPMSimulation, PointCloud
