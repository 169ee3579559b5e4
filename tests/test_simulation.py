"""Rollout driver and scene generators (femflow/simulation/mpm/*.py)."""
import os

import numpy as np
import pytest

from femflow_amd.simulation.mpm import primitives as P


# ------------------------------------------------------------------ CPU ----
def test_grid_reference_kat():
    """numerics/tests/test_geometry.py:6-40: res (3, 3, 3) gives the 27 points
    of {0, 0.5, 1}^3 with the first coordinate varying fastest."""
    h = (0.0, 0.5, 1.0)
    exp = np.array([[x, y, z] for z in h for y in h for x in h])
    assert np.array_equal(P.grid(np.array((3, 3, 3))), exp)


@pytest.mark.parametrize("res", [(3, 2, 4), (1, 3, 2), (5, 5, 5)])
def test_grid_mixed_radix(res):
    """Row i holds the mixed-radix digits of i (first axis least significant),
    each divided by res_c - 1, or by res_c on an axis of one point."""
    g = P.grid(np.array(res))
    i = np.arange(int(np.prod(res)))
    digits = (i % res[0], (i // res[0]) % res[1], i // (res[0] * res[1]))
    assert g.shape == (len(i), 3)
    for c in range(3):
        assert np.array_equal(g[:, c], digits[c] / (res[c] if res[c] == 1 else res[c] - 1))


def test_cube_points_column_order():
    pts = P.generate_cube_points((0.0, 1.0), (0.0, 2.0), (0.0, 3.0), res=3)
    exp = []
    for layer in np.linspace(0, 3, 3):
        for row in np.linspace(0, 2, 3):
            for col in np.linspace(0, 1, 3):
                exp.append([layer, row, col])
    assert np.array_equal(pts, np.array(exp))


def test_implicit_functions_closed_form():
    # at the origin: gyroid = -t, primitive = 3 - t, diamond = 0*... + 1*1*0... = -t
    o = np.zeros((1, 3))
    assert np.isclose(P.gyroid(0.2, 0.3, o)[0], -0.3)
    assert np.isclose(P.primitive(0.2, 0.3, o)[0], 3 - 0.3)
    assert np.isclose(P.diamond(0.2, 0.3, o)[0], -0.3)
    # period k: f(x + k) = f(x)
    x = np.random.default_rng(1).random((20, 3))
    for f in (P.gyroid, P.primitive, P.diamond):
        assert np.allclose(f(0.2, 0.1, x), f(0.2, 0.1, x + 0.2))
    pts = P.generate_implicit_points("gyroid", 0.2, 0.3, 12)
    g = P.grid(np.array((12, 12, 12)))
    assert len(pts) == int((P.gyroid(0.2, 0.3, g) > 0.3).sum())
    with pytest.raises(ValueError):
        P.generate_implicit_points("nope", 0.2, 0.3, 4)


# ------------------------------------------------------------------ GPU ----
def _paper_scene(mesh_res=16):
    """Experiment 0 of paper_1.py:75-96 at a reduced lattice resolution: a
    gyroid lifted by 0.1 and a collider cube over its bounding box lifted by 3,
    scaled by tightening_coeff = 0.05 inside load()."""
    from femflow_amd.simulation.mpm import PointCloud

    g = P.generate_implicit_points("gyroid", 0.2, 0.3, mesh_res)
    g = g + np.array([0.0, 0.1, 0.0])
    lo, hi = g.min(0), g.max(0)
    c = P.generate_cube_points((lo[0], hi[0]), (lo[1], hi[1]), (lo[2], hi[2]), mesh_res)
    c = c + np.array([0.0, 3.0, 0.0])
    return [PointCloud(g), PointCloud(c)], [(1.0, 140.0, 0.2), (10.0, 1000.0, 0.4)]


@pytest.mark.gpu
def test_rollout_matches_oracle_and_writes_reference_frames(gpu, tmp_path):
    from femflow_amd.simulation.mpm import MPMSimulation
    from oracle import mpm_oracle as O
    from tests._parity import field_errors

    meshes, params = _paper_scene()
    steps = 30
    sim = MPMSimulation(str(tmp_path / "out"), steps, 1e-4, 1.0, 10.0, 1.0, -9.8, 140.0, 1000.0, 0.2, 0.4,
                        0.7, 64, 0.05, device=gpu, frames_per_chunk=7, sort_every=4)
    sim.load(meshes=meshes, params=params)
    x0 = sim.positions.copy()
    sim.start()
    sim.join(timeout=120)
    assert not sim.running and sim.error is None
    files = sorted(os.listdir(tmp_path / "out"), key=lambda s: int(s.split(".")[0]))
    assert len(files) == steps + 1
    f0 = np.load(tmp_path / "out" / "0.npy")
    assert f0.dtype == np.float64 and f0.shape == (3 * len(x0),)
    assert np.allclose(f0, x0.reshape(-1) / 0.05, rtol=0, atol=1e-6)
    # oracle: FP64 solve_mls_mpm_3d per substep from the same FP32 start
    n = len(x0)
    mass = np.concatenate([np.full(len(m.vertices) // 3, p[0]) for m, p in zip(meshes, params)])
    mu = np.concatenate([np.full(len(m.vertices) // 3, O.Ev_to_mu(p[1], p[2])) for m, p in zip(meshes, params)])
    lam = np.concatenate([np.full(len(m.vertices) // 3, O.Ev_to_lambda(p[1], p[2])) for m, p in zip(meshes, params)])
    x = x0.copy()
    v, F, C, Jp = O.make_mls_mpm_coefficients(n, 3)
    worst = 0.0
    for s in range(1, steps + 1):
        O.solve_mls_mpm_3d(64, 64.0, 0.7, 1 / 64, 1e-4, 1.0, -9.8, x, mass, mu, lam, v, F, C, Jp)
        fs = np.load(tmp_path / "out" / f"{s}.npy").reshape(n, 3) * 0.05
        worst = max(worst, np.abs(fs - x).max() / np.abs(x).max())
    print("rollout position drift vs FP64 oracle over", steps, "steps:", worst, flush=True)
    assert worst < 1e-4
    got = dict(x=sim.positions, v=sim.v, F=sim.F, C=sim.C, Jp=sim.Jp)
    errs = field_errors(got, dict(x=x, v=v, F=F, C=C, Jp=Jp), 64.0)
    print("final state errors", errs, flush=True)
    assert errs["x"] < 1e-4 and errs["v"] < 1e-3 and errs["F"] < 1e-3
    # the reference's particle records (simulation.py:78-89)
    parts = sim.particles
    assert len(parts) == n
    p0, plast = parts[0], parts[-1]
    assert np.array_equal(p0.pos, got["x"][0]) and np.array_equal(plast.pos, got["x"][-1])
    assert p0.pos.dtype == np.float64 and p0.force == -9.8
    assert p0.mass == 1.0 and np.isclose(p0.mu_0, mu[0]) and np.isclose(p0.lambda_0, lam[0])
    assert np.isclose(plast.mass, 10.0) and np.isclose(plast.mu_0, mu[-1]) and np.isclose(plast.lambda_0, lam[-1])
