"""Shared implementation of the phase functions (P2G, grid update, G2P).

Each call stages the caller's arrays on the GPU (zero-copy for CUDA tensors of
the compute dtype), launches the HIP phase kernel through the C ABI, and
writes results back in place, matching the reference's mutate-in-place
contract (femflow/solvers/mpm/three_d/*.py, two_d/*.py).
"""
from __future__ import annotations

import ctypes as ct

import numpy as np
import torch

from ... import _lib
from ..._arrays import Staged, compute_dtype, device_of, mpm_dtype, stream_ptr
from .particle import Particle, ParticleSet, particles_to_set, scatter_positions

THETA_C, THETA_S = 2.5e-2, 7.5e-3  # three_d/g2p.py:50


def p2g_kind(model: str) -> int:
    # three_d/p2g.py:57-61: "neo_hookean" -> constant hardening, anything else -> snow law
    return _lib.MPM_ELASTIC if model == "neo_hookean" else _lib.MPM_SNOW


def g2p_kind(model: str) -> int:
    # three_d/g2p.py:48: only "snow" projects the singular values
    return _lib.MPM_SNOW if model == "snow" else _lib.MPM_ELASTIC


def boundary_code(boundary: str) -> int:
    """Grid wall rule: "reference" (three_d/grid_op.py:37-47 slip walls,
    two_d/grid_op.py:13-24 sticky sides and floor) or "sticky" (every
    component zero within `bound` nodes of any wall)."""
    codes = {"reference": _lib.MPM_BC_REFERENCE, "sticky": _lib.MPM_BC_STICKY}
    if boundary not in codes:
        raise ValueError(f"boundary must be 'reference' or 'sticky', got {boundary!r}")
    return codes[boundary]


def quirks_code(quirks: str) -> int:
    """Stress rule: "reference" adds lambda (J-1) J to every entry of PF
    (utils.py:129) or "textbook" adds it to the diagonal."""
    codes = {"reference": _lib.MPM_QUIRKS_REFERENCE, "textbook": _lib.MPM_QUIRKS_TEXTBOOK}
    if quirks not in codes:
        raise ValueError(f"quirks must be 'reference' or 'textbook', got {quirks!r}")
    return codes[quirks]


def make_params(dim, res, inv_dx=0.0, dx=0.0, dt=1.0, volume=0.0, gravity=0.0,
                boundary=_lib.MPM_BC_REFERENCE, bound=3, quirks=_lib.MPM_QUIRKS_REFERENCE,
                check_oob=None):
    return _lib.Params(dim, int(res), float(inv_dx), float(dx), float(dt), float(volume),
                       float(gravity), int(boundary), int(bound), int(quirks),
                       int(dim == 3 if check_oob is None else check_oob))


def materials_array(materials):
    if materials is None:
        return None, 0
    arr = (_lib.Material * len(materials))()
    for i, m in enumerate(materials):
        kind, h, tc, ts = m
        arr[i] = _lib.Material(int(kind), float(h), float(tc), float(ts))
    return arr, len(materials)


class StagedParticles:
    """Device views of (x, v, F, C, Jp, mass, mu0, lambda0[, material])."""

    def __init__(self, pos, v, F, C, Jp, mass, mu0, lam0, material=None, mat_scalar=False,
                 records=None):
        self.dtype = compute_dtype(v, F, pos)
        self.device = device_of(pos, v, F, C, Jp)
        dt = self.dtype
        self.dim = int(v.shape[-1])
        self.n = int(v.shape[0])
        self.x = Staged(pos, dt, self.device)
        self.v = Staged(v, dt, self.device)
        self.F = Staged(F, dt, self.device)
        self.C = Staged(C, dt, self.device)
        self.Jp = Staged(Jp, dt, self.device)
        self.mass = Staged(mass, dt, self.device, writable=False)
        self.mu0 = Staged(mu0, dt, self.device, writable=False)
        self.lam0 = Staged(lam0, dt, self.device, writable=False)
        self.material = None if material is None else Staged(material, torch.int32, self.device,
                                                              writable=False)
        self.mat_scalar = mat_scalar
        self.records = records

    def struct(self):
        return _lib.Particles(
            self.n, mpm_dtype(self.dtype), 1 if self.mat_scalar else 0, self.x.ptr, self.v.ptr,
            self.F.ptr, self.C.ptr, self.Jp.ptr, self.mass.ptr, self.mu0.ptr, self.lam0.ptr,
            self.material.ptr if self.material is not None else None)

    def commit(self, state=True, positions=True):
        if positions:
            if self.records is not None:
                scatter_positions(self.records, self.x.tensor.cpu().numpy())
            else:
                self.x.commit()
        if state:
            for s in (self.v, self.F, self.C, self.Jp):
                s.commit()


def stage_3d(particles, v, F, C, Jp):
    """Stage a reference-style 3D particle list (or ParticleSet) with its state."""
    if isinstance(particles, ParticleSet):
        return StagedParticles(particles.pos, v, F, C, Jp, particles.mass, particles.mu_0,
                               particles.lambda_0, particles.material)
    records = list(particles)
    ps = particles_to_set(records)
    return StagedParticles(ps.pos, v, F, C, Jp, ps.mass, ps.mu_0, ps.lambda_0, records=records)


def run_p2g(sp, prm, kind_override, hardening, grid_velocity, grid_mass, materials=None):
    _lib.require_gpu()
    L = _lib.load()
    gv = Staged(grid_velocity, sp.dtype, sp.device)
    gm = Staged(grid_mass, sp.dtype, sp.device)
    oob = torch.zeros(1, dtype=torch.int64, device=sp.device)
    mats, nmat = materials_array(materials)
    p = sp.struct()
    _lib.check(L.mpm_p2g(stream_ptr(sp.device), ct.byref(prm), ct.byref(p), int(kind_override),
                         float(hardening), mats, nmat, gv.ptr, gm.ptr, oob.data_ptr()), "p2g")
    n_oob = int(oob.item())
    gv.commit()
    gm.commit()
    return n_oob


def run_grid_op(prm, grid_velocity, grid_mass):
    _lib.require_gpu()
    L = _lib.load()
    dt = compute_dtype(grid_velocity, grid_mass)
    dev = device_of(grid_velocity, grid_mass)
    gv = Staged(grid_velocity, dt, dev)
    gm = Staged(grid_mass, dt, dev, writable=False)
    _lib.check(L.mpm_grid_op(stream_ptr(dev), ct.byref(prm), mpm_dtype(dt), gv.ptr, gm.ptr), "grid_op")
    gv.commit()


def plane_arrays(points, normals):
    """(K, 3) float64 host arrays of plane points and normals, K <= 16."""
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    nrm = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).reshape(-1, 3))
    if pts.shape != nrm.shape:
        raise ValueError("points and normals must have the same number of rows")
    return pts, nrm


def run_collide_planes(points, normals, res, dx, grid_velocity, quirks):
    _lib.require_gpu()
    L = _lib.load()
    pts, nrm = plane_arrays(points, normals)
    dt = compute_dtype(grid_velocity)
    dev = device_of(grid_velocity)
    gv = Staged(grid_velocity, dt, dev)
    q = _lib.MPM_QUIRKS_TEXTBOOK if quirks == "textbook" else _lib.MPM_QUIRKS_REFERENCE
    _lib.check(L.mpm_collide_planes(stream_ptr(dev), mpm_dtype(dt), int(res), float(dx), len(pts),
                                    pts.ctypes.data, nrm.ctypes.data, q, gv.ptr), "check_collision_points")
    gv.commit()


def run_g2p(sp, prm, kind_override, grid_velocity, theta_c=THETA_C, theta_s=THETA_S, materials=None):
    _lib.require_gpu()
    L = _lib.load()
    gv = Staged(grid_velocity, sp.dtype, sp.device, writable=False)
    oob = torch.zeros(1, dtype=torch.int64, device=sp.device)
    mats, nmat = materials_array(materials)
    p = sp.struct()
    _lib.check(L.mpm_g2p(stream_ptr(sp.device), ct.byref(prm), ct.byref(p), int(kind_override),
                         float(theta_c), float(theta_s), mats, nmat, gv.ptr, oob.data_ptr()), "g2p")
    return int(oob.item())
