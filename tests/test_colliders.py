"""Plane colliders (three_d/grid_op.py:50-67) on the GPU against the oracle.

The node test runs in FP64 on both sides with the same operation order, so
the set of zeroed nodes must match exactly (bit-exact grids); the engine's
rollouts with planes follow the FP32-vs-FP64 bound of the engine parity
tests (1e-4 over several substeps).
"""
import numpy as np
import pytest
import torch

from femflow_amd.solvers.mpm import three_d
from oracle import mpm_oracle as O
from tests._parity import field_errors
from tests.gpu_helpers import engine_for, engine_roundtrip, oracle_copy_from, to_torch_state

pytestmark = pytest.mark.gpu


def _planes(seed, k=3):
    rng = np.random.default_rng(seed)
    return rng.random((k, 3)), rng.normal(size=(k, 3))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("quirks", ["reference", "textbook"])
def test_check_collision_points_matches_oracle(gpu, dtype, quirks):
    res = 24
    dx = 1.0 / res
    pts, nrm = _planes(7)
    g = torch.randn((res + 1,) * 3 + (3,), dtype=dtype, device=gpu)
    ref = g.double().cpu().numpy()
    O.collide_planes(pts, nrm, res, dx, ref, quirks=quirks)
    three_d.check_collision_points(pts, nrm, res, dx, g, quirks=quirks)
    out = g.double().cpu().numpy()
    print("collide", dtype, quirks, "zeroed", int((out == 0).all(-1).sum()), flush=True)
    assert np.array_equal(out, ref)
    assert 0 < int((out == 0).all(-1).sum()) < (res + 1) ** 3


def test_check_collision_points_numpy_and_errors(gpu):
    res = 8
    g = np.ones((res + 1,) * 3 + (3,))
    three_d.check_collision_points([(0.5, 0.5, 0.5)], [(0, 0, 2)], res, 1.0 / res, g)
    ref = np.ones_like(g)
    O.collide_planes([(0.5, 0.5, 0.5)], [(0, 0, 2)], res, 1.0 / res, ref)
    assert np.array_equal(g, ref)
    with pytest.raises(ValueError):
        three_d.check_collision_points([(0.5, 0.5, 0.5)], [(0, 0, 0)], res, 1.0 / res, g)
    with pytest.raises(ValueError):
        three_d.check_collision_points(np.zeros((17, 3)), np.ones((17, 3)), res, 1.0 / res, g)


@pytest.mark.parametrize("quirks", ["reference", "textbook"])
def test_engine_planes_multistep(gpu, quirks):
    st, pr = O.scene_3d_cubes(n=30000, res=64)
    # a floor under the first cube's upper half and a tilted wall through the third
    pts = np.array([[0.25, 0.55, 0.5], [0.75, 0.5, 0.5]])
    nrm = np.array([[0.0, 1.0, 0.0], [-1.0, 0.2, 0.1]])
    ts = to_torch_state(st, gpu)
    ref = oracle_copy_from(ts, st)
    eng = engine_for(st, pr, gpu, sort_every=4, use_graph=True, chunk=512)
    eng.set_planes(pts, nrm, quirks=quirks)
    nsteps = 8
    out = engine_roundtrip(eng, ts, nsteps)
    prp = dict(pr, planes=(pts, nrm, quirks))
    free = O.copy_state(ref)
    for _ in range(nsteps):
        O.substep(ref, prp)
        O.substep(free, pr)
    errs = field_errors(out, ref, pr["inv_dx"])
    moved = float(np.abs(ref["x"] - free["x"]).max())
    print("engine planes", quirks, errs, "effect", moved, flush=True)
    assert moved > 1e-6  # the planes change the motion
    for k in ("x", "v", "F"):
        assert errs[k] < 1e-4, errs
    assert errs["Jp"] < 1e-4, errs
    # removing the planes restores the free motion
    eng.set_planes(np.zeros((0, 3)), np.zeros((0, 3)))
    out2 = engine_roundtrip(eng, to_torch_state(st, gpu), nsteps)
    free2 = oracle_copy_from(to_torch_state(st, gpu), st)
    for _ in range(nsteps):
        O.substep(free2, pr)
    errs2 = field_errors(out2, free2, pr["inv_dx"])
    assert errs2["x"] < 1e-4 and errs2["v"] < 1e-4, errs2
    with pytest.raises(ValueError):
        eng.set_planes(pts, nrm, scene=1)
    eng.close()
