"""Particle records (femflow/solvers/mpm/particle.py:10-33).

``Particle`` keeps the reference's fields (a plain class instead of a numba
jitclass).  ``ParticleSet`` is the structure-of-arrays form the GPU path uses
directly: positions (N, 3) and per-particle mass, mu_0, lambda_0 and an
optional material id, as numpy arrays or CUDA tensors.
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np


class Particle:
    """particle.py:10-27: pos (float64[3]), force, mass, lambda_0, mu_0."""

    __slots__ = ("pos", "force", "mass", "lambda_0", "mu_0")

    def __init__(self, pos, force: float, mass: float, lambda_: float, mu: float):
        self.pos = pos
        self.force = force
        self.mass = mass
        self.lambda_0 = lambda_
        self.mu_0 = mu


class ParticleSet:
    """SoA particles: ``pos`` (N, d), ``mass``/``mu_0``/``lambda_0`` (N) and
    optional int32 ``material`` ids (N) into a material table."""

    def __init__(self, pos, mass, mu_0, lambda_0, material=None):
        self.pos = pos
        n = len(pos)
        self.mass = _per_particle(mass, n)
        self.mu_0 = _per_particle(mu_0, n)
        self.lambda_0 = _per_particle(lambda_0, n)
        self.material = material

    def __len__(self):
        return len(self.pos)


def _per_particle(a, n):
    try:
        import torch

        if isinstance(a, torch.Tensor):
            return a.expand(n) if a.dim() == 0 else a
    except ImportError:  # pragma: no cover
        pass
    a = np.asarray(a, dtype=np.float64)
    return np.full(n, float(a)) if a.ndim == 0 else a


def particles_to_set(particles: Sequence[Particle]) -> ParticleSet:
    """Stack a list of ``Particle`` records into a ``ParticleSet`` (copies)."""
    pos = np.array([p.pos for p in particles], dtype=np.float64).reshape(len(particles), -1)
    return ParticleSet(pos, np.array([p.mass for p in particles], dtype=np.float64),
                       np.array([p.mu_0 for p in particles], dtype=np.float64),
                       np.array([p.lambda_0 for p in particles], dtype=np.float64))


def scatter_positions(particles: Sequence[Particle], pos) -> None:
    """Write updated positions back into each record's ``pos`` in place."""
    pos = np.asarray(pos)
    for p, x in zip(particles, pos):
        p.pos[...] = x


def map_particles_to_pos(particles, coeff: float):
    """Flat positions divided by ``coeff`` (particle.py:30-33)."""
    if isinstance(particles, ParticleSet):
        pos = particles.pos
        try:
            import torch

            if isinstance(pos, torch.Tensor):
                pos = pos.detach().double().cpu().numpy()
        except ImportError:  # pragma: no cover
            pass
        return (np.asarray(pos, dtype=np.float64) / coeff).reshape(-1)
    return np.array([p.pos.copy() / coeff for p in particles]).reshape(-1)
