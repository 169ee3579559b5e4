"""Sticky walls: every velocity component zero within `bound` nodes of a wall.
This is the boundary option that the dense phase kernels and the engine
carry beside the reference's rules, the 3D slip walls (three_d/grid_op.py:37-47)
and the 2D floor and sticky sides (two_d/grid_op.py:13-24).  The oracle's rule
is pinned by node counts in test_oracle.py."""
import numpy as np
import pytest
import torch

from femflow_amd import _lib
from femflow_amd.solvers.mpm import _phase
from oracle import mpm_oracle as O
from tests._parity import field_errors
from tests.gpu_helpers import engine_for, engine_roundtrip, oracle_copy_from, to_torch_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("dim", [2, 3])
def test_dense_grid_op_sticky(gpu, dim, dtype):
    rng = np.random.default_rng(40 + dim)
    res, bound, dt, g = 24, 3, 1e-3, -9.8
    shape = (res + 1,) * dim
    gv = rng.standard_normal(shape + (dim,))
    gm = np.where(rng.random(shape) < 0.7, rng.random(shape) + 0.1, 0.0)
    tv = torch.tensor(gv, dtype=dtype, device=gpu)
    tm = torch.tensor(gm if dim == 2 else gm[..., None], dtype=dtype, device=gpu)
    ref_v = tv.double().cpu().numpy()
    ref_m = tm.double().cpu().numpy()
    prm = _phase.make_params(dim, res, float(res), 1.0 / res, dt, 0.0, g, boundary=_lib.MPM_BC_STICKY, bound=bound)
    _phase.run_grid_op(prm, tv, tm)
    if dim == 3:
        O.grid_op_3d(res, 1.0 / res, dt, g, ref_v, ref_m, boundary="sticky", bound=bound)
    else:
        O.grid_op_2d(res, dt, g, ref_v, ref_m, boundary="sticky", bound=bound)
    out = tv.double().cpu().numpy()
    tol = 1e-12 if dtype == torch.float64 else 1e-6
    assert np.abs(out - ref_v).max() <= tol * np.abs(ref_v).max()
    assert np.array_equal(out == 0.0, ref_v == 0.0)  # the same nodes zeroed


def test_engine_sticky_walls(gpu):
    """A cube hugging the low x wall: its stencils reach the three sticky node
    layers, where every component is zeroed (the slip rule would keep v_y, v_z)."""
    res = 32
    dx = 1.0 / res
    st, pr = O.scene_3d_cubes(n=6000, res=res, seed=44, side=4.0 * dx, corners=((1.6 * dx, 0.45, 0.45),),
                              vel_sigma=0.0)
    st["v"] = np.ascontiguousarray(np.tile([-1.0, 0.5, 0.8], (len(st["x"]), 1)))
    pr = dict(pr, boundary="sticky", bound=3)
    ts = to_torch_state(st, gpu)
    ref = oracle_copy_from(ts, st)
    eng = engine_for(st, pr, gpu, sort_every=2)
    nsteps = 3
    out = engine_roundtrip(eng, ts, nsteps)
    for _ in range(nsteps):
        O.substep(ref, pr)
    slip = oracle_copy_from(ts, st)
    for _ in range(nsteps):
        O.substep(slip, dict(pr, boundary="reference"))
    assert np.abs(slip["v"] - ref["v"]).max() > 1e-2  # the rules differ on this scene
    errs = field_errors(out, ref, pr["inv_dx"])
    print("engine sticky", errs, flush=True)
    for k in ("x", "v", "F"):
        assert errs[k] < 1e-5, errs
    assert errs["C"] < 1e-4 and errs["Jp"] < 1e-5, errs
    eng.close()
