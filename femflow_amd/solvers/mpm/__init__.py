"""Drop-in GPU versions of femflow.solvers.mpm."""
from . import three_d, two_d
from .mls_mpm import make_mls_mpm_coefficients, solve_mls_mpm_3d
from .particle import Particle, ParticleSet, map_particles_to_pos
