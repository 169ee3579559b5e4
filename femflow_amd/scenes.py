"""Seeded synthetic scenes (numpy, float64) for the benchmark configurations.

State dicts hold per-particle arrays in the reference's layouts:
x, v (N, d); F, C (N, d, d); Jp (N, 1); mass, mu0, lam0, h, theta_c, theta_s
(N,); kind (N,) int32 in {0 elastic, 1 snow, 2 water}.  Params hold res, dx,
inv_dx, dt, volume, gravity, boundary, bound.
"""
from __future__ import annotations

import numpy as np

from .numerics.fem import Ev_to_lambda, Ev_to_mu

ELASTIC, SNOW, WATER = 0, 1, 2
THETA_C_REF, THETA_S_REF = 2.5e-2, 7.5e-3

MATERIALS_C2 = (
    # kind, E, nu, h, theta_c, theta_s   (water: the E slot holds the bulk modulus K)
    (ELASTIC, 5e3, 0.2, 1.0, THETA_C_REF, THETA_S_REF),
    (SNOW, 5e3, 0.2, 10.0, 2.5e-2, 7.5e-3),
    (WATER, 1e4, 0.0, 1.0, THETA_C_REF, THETA_S_REF),
)

CUBE_CORNERS_C2 = ((0.15, 0.45, 0.40), (0.42, 0.20, 0.40), (0.65, 0.45, 0.40))


def _coefficients(n, dim):
    v = np.zeros((n, dim))
    F = np.broadcast_to(np.eye(dim), (n, dim, dim)).copy()
    C = np.zeros((n, dim, dim))
    Jp = np.ones((n, 1))
    return v, F, C, Jp


def scene_2d_c1(n=8192, seed=0, res=128):
    """2D: n particles (seed 0) uniform in [0.4,0.6]x[0.6,0.8]; vol=(0.5 dx)^2
    (the 4-particles-per-cell convention); rho=1; E=1e3, nu=0.2; g=-9.8;
    dt=1e-4; the reference 2D walls."""
    rng = np.random.default_rng(seed)
    u = rng.random((n, 2))
    x = np.empty((n, 2))
    x[:, 0] = 0.4 + 0.2 * u[:, 0]
    x[:, 1] = 0.6 + 0.2 * u[:, 1]
    dx = 1.0 / res
    vol = (0.5 * dx) ** 2
    E, nu = 1e3, 0.2
    v, F, C, Jp = _coefficients(n, 2)
    state = dict(x=x, v=v, F=F, C=C, Jp=Jp, mass=np.full(n, vol * 1.0), mu0=np.full(n, Ev_to_mu(E, nu)),
                 lam0=np.full(n, Ev_to_lambda(E, nu)), kind=np.zeros(n, np.int32), h=np.ones(n),
                 theta_c=np.full(n, THETA_C_REF), theta_s=np.full(n, THETA_S_REF))
    params = dict(res=res, dx=dx, inv_dx=float(res), dt=1e-4, volume=vol, gravity=-9.8,
                  boundary="reference", bound=3)
    return state, params


def scene_3d_cubes(n=1 << 20, res=128, seed=0, side=0.2, corners=CUBE_CORNERS_C2,
                   materials=MATERIALS_C2, vel_sigma=0.5, dt=5e-5, gravity=-9.8, rho=1.0):
    """Cubes of side ``side`` (one per corner) filled uniformly, one material
    each, one N(0, vel_sigma^2) velocity per cube, all from default_rng(seed)
    in that order.  Counts split n evenly (the first cube takes the
    remainder); one scalar particle volume side^3 / (n // cubes)."""
    rng = np.random.default_rng(seed)
    k = len(corners)
    counts = [n // k] * k
    counts[0] += n - sum(counts)
    xs, vs, kinds, mus, lams, hs, tcs, tss, masses = [], [], [], [], [], [], [], [], []
    for corner, cnt in zip(corners, counts):
        xs.append(np.asarray(corner)[None, :] + side * rng.random((cnt, 3)))
    for cnt in counts:
        vs.append(np.broadcast_to(rng.normal(0.0, vel_sigma, size=3), (cnt, 3)))
    vol_p = side ** 3 / max(counts[-1], 1)
    for c, cnt in enumerate(counts):
        kind, E, nu, h, tc, ts = materials[c % len(materials)]
        masses.append(np.full(cnt, rho * vol_p))
        kinds.append(np.full(cnt, kind, np.int32))
        if kind == WATER:
            mus.append(np.zeros(cnt))
            lams.append(np.full(cnt, E))
        else:
            mus.append(np.full(cnt, Ev_to_mu(E, nu)))
            lams.append(np.full(cnt, Ev_to_lambda(E, nu)))
        hs.append(np.full(cnt, h))
        tcs.append(np.full(cnt, tc))
        tss.append(np.full(cnt, ts))
    x = np.concatenate(xs)
    v = np.ascontiguousarray(np.concatenate(vs))
    _, F, C, Jp = _coefficients(n, 3)
    state = dict(x=x, v=v, F=F, C=C, Jp=Jp, mass=np.concatenate(masses), mu0=np.concatenate(mus),
                 lam0=np.concatenate(lams), kind=np.concatenate(kinds), h=np.concatenate(hs),
                 theta_c=np.concatenate(tcs), theta_s=np.concatenate(tss))
    dx = 1.0 / res
    params = dict(res=res, dx=dx, inv_dx=float(res), dt=dt, volume=vol_p, gravity=gravity,
                  boundary="reference", bound=3)
    return state, params


def batched_scene_specs(n_scenes=64, seed=0, n_lo=4096, n_hi=262144, res_lo=64, res_hi=128,
                        side_lo=0.1, side_hi=0.3):
    """Per-scene (n, res, material index, side, corner) for the batched
    configuration: n log-uniform in [n_lo, n_hi], res uniform integer in
    [res_lo, res_hi], one random material of the three, a box of random side
    placed at a random corner away from the walls."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n_scenes):
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        res = int(rng.integers(res_lo, res_hi + 1))
        mat = int(rng.integers(0, 3))
        side = float(rng.uniform(side_lo, side_hi))
        lo, hi = 0.1, 0.9 - side
        corner = tuple(float(c) for c in rng.uniform(lo, hi, size=3))
        specs.append(dict(n=n, res=res, material=mat, side=side, corner=corner))
    return specs


def batched_scene(spec, seed, dt=5e-5):
    """One scene of the batched configuration (a single box of one material)."""
    mat = MATERIALS_C2[spec["material"]]
    return scene_3d_cubes(n=spec["n"], res=spec["res"], seed=seed, side=spec["side"],
                          corners=(spec["corner"],), materials=(mat,), dt=dt)
