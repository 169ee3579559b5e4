"""Sticky walls: every velocity component zero within `bound` nodes of a wall.
This is the boundary option that the dense phase kernels and the engine
carry beside the reference's rules, the 3D slip walls (three_d/grid_op.py:37-47)
and the 2D floor and sticky sides (two_d/grid_op.py:13-24).  The oracle's rule
is pinned by node counts in test_oracle.py."""
import numpy as np
import pytest
import torch

from femflow_amd.solvers.mpm import ParticleSet, solve_mls_mpm_3d, three_d, two_d
from oracle import mpm_oracle as O
from tests._parity import field_errors
from tests.gpu_helpers import engine_for, engine_roundtrip, oracle_copy_from, to_numpy_state, to_torch_state

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
    if dim == 3:
        three_d.grid_op(res, 1.0 / res, dt, g, tv, tm, boundary="sticky", bound=bound)
        O.grid_op_3d(res, 1.0 / res, dt, g, ref_v, ref_m, boundary="sticky", bound=bound)
    else:
        two_d.grid_op(res, dt, g, tv, tm, boundary="sticky", bound=bound)
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


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_solve_mls_mpm_3d_sticky(gpu, dtype):
    """The substep driver with boundary="sticky": FP64 runs the dense phase
    kernels, FP32 the engine."""
    res = 32
    dx = 1.0 / res
    st, pr = O.scene_3d_cubes(n=3000, res=res, seed=45, side=3.0 * dx, corners=((1.4 * dx, 0.45, 0.45),),
                              materials=(O.MATERIALS_C2[0],), vel_sigma=0.0)
    st["v"] = np.ascontiguousarray(np.tile([-1.0, 0.5, 0.8], (len(st["x"]), 1)))
    ts = to_torch_state(st, gpu, dtype)
    ref = oracle_copy_from(ts, st)
    ps = ParticleSet(ts["x"], ts["mass"], ts["mu0"], ts["lam0"])
    Jp = ts["Jp"].view(-1, 1)
    for _ in range(2):
        solve_mls_mpm_3d(res, float(res), 1.0, dx, pr["dt"], pr["volume"], pr["gravity"], ps, ts["v"], ts["F"],
                         ts["C"], Jp, boundary="sticky", bound=3)
        O.substep(ref, dict(pr, boundary="sticky", bound=3))
    torch.cuda.synchronize()
    errs = field_errors(to_numpy_state(ts), ref, pr["inv_dx"])
    print("solve sticky", dtype, errs, flush=True)
    lim = 1e-9 if dtype == torch.float64 else 1e-5
    for k in ("x", "v", "F"):
        assert errs[k] < lim, errs
    assert errs["C"] < 10 * lim and errs["Jp"] < 10 * lim, errs


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_three_d_p2g_textbook_stress(gpu, dtype):
    """quirks="textbook": lambda (J-1) J on the diagonal of PF only; the two
    modes differ wherever J != 1 and F is not diagonal."""
    st, pr = O.scene_3d_cubes(n=4000, res=32, seed=46, materials=(O.MATERIALS_C2[0],))
    rng = np.random.default_rng(47)
    st["F"] = st["F"] + 0.05 * rng.standard_normal(st["F"].shape)
    ts = to_torch_state(st, gpu, dtype)
    ref = oracle_copy_from(ts, st)
    r1 = pr["res"] + 1
    grids = {}
    for quirks in ("textbook", "reference"):
        gv = torch.zeros((r1, r1, r1, 3), dtype=dtype, device=gpu)
        gm = torch.zeros((r1, r1, r1, 1), dtype=dtype, device=gpu)
        ps = ParticleSet(ts["x"], ts["mass"], ts["mu0"], ts["lam0"])
        three_d.p2g(pr["inv_dx"], 1.0, pr["dx"], pr["dt"], pr["volume"], gv, gm, ps, ts["v"], ts["F"], ts["C"],
                    ts["Jp"].view(-1, 1), quirks=quirks)
        rgv, rgm = np.zeros((r1, r1, r1, 3)), np.zeros((r1, r1, r1, 1))
        O.p2g_3d(pr["inv_dx"], 1.0, pr["dx"], pr["dt"], pr["volume"], rgv, rgm, ref["x"], ref["mass"], ref["mu0"],
                 ref["lam0"], ref["v"], ref["F"], ref["C"], ref["Jp"], "neo_hookean", quirks=quirks)
        tol = 1e-9 if dtype == torch.float64 else 2e-5
        assert np.abs(gv.double().cpu().numpy() - rgv).max() <= tol * np.abs(rgv).max(), quirks
        grids[quirks] = rgv
    assert np.abs(grids["textbook"] - grids["reference"]).max() > 1e-3 * np.abs(grids["reference"]).max()
