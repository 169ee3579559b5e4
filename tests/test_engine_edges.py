"""Engine edge cases against the oracle: an empty engine, thousands of
particles sharing one cell (many chunks per block, the fixed-point bound
under heavy overlap), and particles within half a cell of the low walls,
where the reference's truncated base gives fx < 0.5 and a negative middle
weight (three_d/p2g.py:50-55, quirk Q1)."""
import numpy as np
import pytest
import torch

from oracle import mpm_oracle as O
from tests._parity import field_errors
from tests.gpu_helpers import engine_for, engine_roundtrip, oracle_copy_from, to_torch_state

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_engine_empty_state(gpu):
    st, pr = O.scene_3d_cubes(n=64, res=32)
    eng = engine_for(st, pr, gpu, sort_every=4)
    z3 = torch.zeros((0, 3), device=gpu)
    z33 = torch.zeros((0, 3, 3), device=gpu)
    z = torch.zeros((0,), device=gpu)
    eng.upload(z3, z3, z33, z33, z, z, z, z, torch.zeros((0,), dtype=torch.int32, device=gpu))
    sums = torch.zeros((6, 8), dtype=torch.float64, device=gpu)
    eng.run(6, sums=sums)
    eng.download(z3.clone(), z3.clone(), z33.clone(), z33.clone(), z.clone())
    assert eng.oob_count() == 0
    assert float(sums.abs().sum()) == 0.0
    assert eng.positions().shape == (0, 3)
    eng.close()


@pytest.mark.parametrize("chunk,spread", [(256, 3), (512, 3), (512, 9)])
def test_engine_thousands_in_one_cell(gpu, chunk, spread):
    res = 32
    elastic = (O.MATERIALS_C2[0],)
    # 6000 particles in one cell plus a sparse cube around it
    st, pr = O.scene_3d_cubes(n=6000, res=res, seed=5, side=1.0 / res, corners=((0.5, 0.5, 0.5),),
                              materials=elastic, vel_sigma=0.3)
    st2, _ = O.scene_3d_cubes(n=1500, res=res, seed=6, side=0.15, corners=((0.44, 0.44, 0.44),),
                              materials=elastic, vel_sigma=0.3)
    for k in st:
        st[k] = np.concatenate([st[k], st2[k]])
    ts = to_torch_state(st, gpu)
    ref = oracle_copy_from(ts, st)
    eng = engine_for(st, pr, gpu, sort_every=2, chunk=chunk, use_graph=True, spread=spread)
    nsteps = 4
    out = engine_roundtrip(eng, ts, nsteps)
    for _ in range(nsteps):
        O.substep(ref, pr)
    errs = field_errors(out, ref, pr["inv_dx"])
    _say("one cell", chunk, spread, errs)
    for k in ("x", "v", "F"):
        assert errs[k] < 1e-5, errs
    assert errs["C"] < 1e-4 and errs["Jp"] < 1e-5, errs
    eng.close()


def test_engine_low_wall_truncated_base(gpu):
    res = 32
    dx = 1.0 / res
    elastic = (O.MATERIALS_C2[0],)
    # a slab hugging the x = 0 and z = 0 walls: x, z in [0, 1.5 dx)
    st, pr = O.scene_3d_cubes(n=4000, res=res, seed=9, side=1.5 * dx, corners=((0.0, 0.5, 0.0),),
                              materials=elastic, vel_sigma=0.0)
    st["x"][:, 1] = 0.5 + 0.2 * np.random.default_rng(1).random(len(st["x"]))
    fx0 = st["x"][:, 0] * res
    assert (fx0 < 0.134).sum() > 100  # middle weight negative for these (w1 = 0.75 - (fx - 1)^2)
    ts = to_torch_state(st, gpu)
    ref = oracle_copy_from(ts, st)
    eng = engine_for(st, pr, gpu, sort_every=1)
    out = engine_roundtrip(eng, ts, 1)
    O.substep(ref, pr)
    errs = field_errors(out, ref, pr["inv_dx"])
    _say("low wall", errs)
    for k in ("x", "v", "F"):
        assert errs[k] < 1e-5, errs
    assert errs["C"] < 1e-4 and errs["Jp"] < 1e-5, errs
    eng.close()


def test_engine_run_states_matches_frames_and_download(gpu):
    st, pr = O.scene_3d_cubes(n=20000, res=48)
    ts = to_torch_state(st, gpu)
    n, k = len(st["x"]), 7
    outs = []
    for mode in ("states", "frames"):
        eng = engine_for(st, pr, gpu, sort_every=3)
        eng.upload(ts["x"], ts["v"], ts["F"], ts["C"], ts["Jp"], ts["mass"], ts["mu0"], ts["lam0"], ts["kind"])
        if mode == "states":
            x, v = torch.empty((k, n, 3), device=gpu), torch.empty((k, n, 3), device=gpu)
            F, C = torch.empty((k, n, 3, 3), device=gpu), torch.empty((k, n, 3, 3), device=gpu)
            Jp = torch.empty((k, n), device=gpu)
            eng.run_states(k, x, v, F, C, Jp)
            outs.append((x, v, F, C, Jp))
        else:
            frames = torch.empty((k, n, 3), device=gpu)
            eng.run_frames(k, frames)
            last = {f: torch.empty_like(ts[f]) for f in ("x", "v", "F", "C", "Jp")}
            eng.download(last["x"], last["v"], last["F"], last["C"], last["Jp"])
            outs.append((frames, last))
        eng.close()
    (x, v, F, C, Jp), (frames, last) = outs
    torch.cuda.synchronize()
    assert torch.equal(x, frames)
    assert torch.equal(v[-1], last["v"]) and torch.equal(F[-1], last["F"])
    assert torch.equal(C[-1], last["C"]) and torch.equal(Jp[-1], last["Jp"])
    assert not torch.equal(x[0], x[-1])
    # each recorded substep against a run that ends there: rebins every 3
    # substeps put the G2P of substeps 2 and 5 into the fused pass that moves
    # the particles, whose per-substep state output is checked here
    eng = engine_for(st, pr, gpu, sort_every=3)
    for j in range(k):
        eng.upload(ts["x"], ts["v"], ts["F"], ts["C"], ts["Jp"], ts["mass"], ts["mu0"], ts["lam0"], ts["kind"])
        eng.run(j + 1)
        got = {f: torch.empty_like(ts[f]) for f in ("x", "v", "F", "C", "Jp")}
        eng.download(got["x"], got["v"], got["F"], got["C"], got["Jp"])
        torch.cuda.synchronize()
        for f, rec in (("x", x), ("v", v), ("F", F), ("C", C), ("Jp", Jp)):
            a, b = rec[j].reshape(got[f].shape).double(), got[f].double()
            d = float((a - b).abs().max()) / max(float(b.abs().max()), 1e-30)
            assert d <= 1e-6, (j, f, d)
    eng.close()


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def test_engine_deferred_fp64_path(gpu):
    """Near-degenerate deformation gradients take the deferred FP64 pass: snow
    whose smallest singular value is below 1e-2 of the largest (FP64 SVD and
    clamp) and elastic particles whose |det F| is below the FP32 polar guard
    (FP64 polar).  One substep against the FP64 oracle."""
    rng = np.random.default_rng(11)
    mats = (O.MATERIALS_C2[0], O.MATERIALS_C2[1])
    st, pr = O.scene_3d_cubes(n=20000, res=32, seed=12, side=0.25, corners=((0.3, 0.35, 0.3), (0.55, 0.35, 0.55)),
                              materials=mats, vel_sigma=0.2)
    n = len(st["x"])
    F = st["F"].reshape(n, 3, 3)
    picks = rng.choice(n, size=600, replace=False)
    for p in picks:
        if st["kind"][p] == O.SNOW:
            s = np.diag([1.0, 0.95, 0.004])       # s_min / s_max = 0.004 < 1e-2
        else:
            s = np.diag([1.02, 0.98, 2e-4])       # |det F| far below 1e-3 (|F|/sqrt 3)^3
        F[p] = _rotation(rng) @ s @ _rotation(rng)
    st["F"] = F.reshape(st["F"].shape)
    ts = to_torch_state(st, gpu)
    ref = oracle_copy_from(ts, st)
    eng = engine_for(st, pr, gpu, sort_every=1)
    out = engine_roundtrip(eng, ts, 1)
    fb = eng.fallback_count()
    O.substep(ref, pr)
    errs = field_errors(out, ref, pr["inv_dx"])
    _say("deferred", fb, errs)
    assert fb >= 500  # the degenerate particles went through the FP64 pass
    for k in ("x", "v", "F"):
        assert errs[k] < 1e-5, errs
    assert errs["C"] < 1e-4 and errs["Jp"] < 1e-5, errs
    eng.close()


def test_engine_deferred_in_grid_update_matches_separate_pass(gpu):
    """A plain run finishes the FP64-deferred particles inside the grid update
    (the first-started workgroup, the others waiting for it); a run that
    writes every substep's state finishes them in their own launch.  Over 4
    substeps (both deferred-count parities, graph replay) with ~600
    near-degenerate particles the two agree to float-atomic ordering, and
    both defer in every substep."""
    rng = np.random.default_rng(13)
    st, pr = O.scene_3d_cubes(n=20000, res=32, seed=14, side=0.25, corners=((0.3, 0.35, 0.3),),
                              materials=(O.MATERIALS_C2[0],), vel_sigma=0.2)
    n = len(st["x"])
    F = st["F"].reshape(n, 3, 3)
    for p in rng.choice(n, size=600, replace=False):
        F[p] = _rotation(rng) @ np.diag([1.02, 0.98, 2e-4]) @ _rotation(rng)
    st["F"] = F.reshape(st["F"].shape)
    ts = to_torch_state(st, gpu)
    k = 4
    eng = engine_for(st, pr, gpu, sort_every=48)
    fused = engine_roundtrip(eng, ts, k)
    fb_fused = eng.fallback_count()
    # the same run again: the captured graph replays
    again = engine_roundtrip(eng, ts, k)
    fb_again = eng.fallback_count()
    eng.upload(ts["x"], ts["v"], ts["F"], ts["C"], ts["Jp"], ts["mass"], ts["mu0"], ts["lam0"], ts["kind"])
    rec = {f: torch.empty((k,) + tuple(ts[f].shape), device=gpu) for f in ("x", "v", "F", "C", "Jp")}
    eng.run_states(k, rec["x"], rec["v"], rec["F"], rec["C"], rec["Jp"])
    fb_sep = eng.fallback_count()
    torch.cuda.synchronize()
    _say("deferred fused/again/separate", fb_fused, fb_again, fb_sep)
    assert min(fb_fused, fb_again, fb_sep) >= 4 * 500, (fb_fused, fb_again, fb_sep)
    assert abs(fb_fused - fb_sep) <= 0.01 * fb_sep and abs(fb_again - fb_sep) <= 0.01 * fb_sep
    # the bars of the FP64-oracle check above: C (a velocity gradient, times
    # inv_dx) shows the float-atomic order of the deferred scatter the most
    bars = {"x": 1e-5, "v": 1e-5, "F": 1e-5, "C": 1e-4, "Jp": 1e-5}
    for f in ("x", "v", "F", "C", "Jp"):
        sep = rec[f][-1]
        scale = float(sep.abs().max()) + 1e-30
        for got in (fused[f], again[f]):
            d = float((got - sep).abs().max()) / scale
            _say("fused vs separate", f, d)
            assert d < bars[f], (f, d)
    eng.close()


def test_engine_drift_between_rebins(gpu):
    """Fast particles and rare rebinning: over 24 substeps without a rebin the
    particles travel ~2 cells, so many leave the 8^3 arena of the block they
    were binned in and take the global gather / scatter paths (and the arena's
    eighth node layer its global atomics).  Particles that drift across a
    block face both ways reach the arena's outer node layers 0 and 7, which
    fresh binning never touches; an LDS layout that aliased node (e, 7, 6)
    with (e+1, 0, 0) showed up here as a 3e-3 velocity error after 6 steps."""
    elastic = (O.MATERIALS_C2[0],)
    st, pr = O.scene_3d_cubes(n=20000, res=64, seed=21, side=0.15, corners=((0.42, 0.42, 0.42),),
                              materials=elastic, vel_sigma=30.0)
    rng = np.random.default_rng(22)
    st["v"] = np.ascontiguousarray(st["v"] + rng.normal(0.0, 10.0, size=st["v"].shape))  # spread the speeds
    ts = to_torch_state(st, gpu)
    ref = oracle_copy_from(ts, st)
    eng = engine_for(st, pr, gpu, sort_every=24, use_graph=False)
    nsteps = 24
    out = engine_roundtrip(eng, ts, nsteps)
    x0 = ts["x"].double().cpu().numpy()
    for _ in range(nsteps):
        O.substep(ref, pr)
    moved = np.abs(ref["x"] - x0).max(axis=1) * pr["inv_dx"]
    errs = field_errors(out, ref, pr["inv_dx"])
    _say("drift", float(np.mean(moved > 1.0)), errs)
    assert np.mean(moved > 1.0) > 0.2  # a fifth of the particles crossed more than a cell
    for k in ("x", "v", "F"):
        assert errs[k] < 1e-5, errs
    assert errs["C"] < 1e-4 and errs["Jp"] < 1e-5, errs
    eng.close()


def test_engine_sparse_scene_beyond_staging(gpu):
    """About one particle per 4^3-cell block over most of a 128^3 grid: far
    more chunks (one per occupied block) than the engine stages, which it
    sizes at 2 ceil(n / 256) + 8192 arenas, so the chunks past that capacity
    flush through global atomics and the grid update sums the staged
    remainder with them.  The overlapping cubes move at different velocities
    (the three materials interpenetrate) and the particles drift between the
    rebins.  The scene is stiff for its particle mass (wave speed times
    dt / dx = 0.45): FP32 rounding grows about 3x per substep here, equally in
    the dense reference-order FP32 phases, so it runs two substeps."""
    res = 128
    st, pr = O.scene_3d_cubes(n=36000, res=res, seed=31, side=0.6,
                              corners=((0.08, 0.08, 0.08), (0.32, 0.32, 0.32), (0.08, 0.32, 0.2)),
                              vel_sigma=4.0)
    base = np.trunc(st["x"] * res - 0.5).astype(np.int64) >> 2
    occupied = len(np.unique(base[:, 0] * 4096 + base[:, 1] * 64 + base[:, 2]))
    staged = 2 * ((len(st["x"]) + 255) // 256) + 8192
    assert occupied > staged + 2000, (occupied, staged)
    ts = to_torch_state(st, gpu)
    ref = oracle_copy_from(ts, st)
    eng = engine_for(st, pr, gpu, sort_every=2, use_graph=True)
    nsteps = 2
    out = engine_roundtrip(eng, ts, nsteps)
    for _ in range(nsteps):
        O.substep(ref, pr)
    errs = field_errors(out, ref, pr["inv_dx"])
    _say("sparse", occupied, staged, errs)
    for k in ("x", "v", "F"):
        assert errs[k] < 1e-5, errs
    assert errs["C"] < 1e-4 and errs["Jp"] < 1e-5, errs
    eng.close()


def test_engine_high_walls(gpu):
    """A slab against the high x and z walls: stencils reach node res, the
    last grid node, which sits in a partly filled 4^3 block (res + 1 nodes per
    axis), and the walls' two-node layer zeroes the normal velocity
    (three_d/grid_op.py:38-47); particles moving outward pile up there."""
    res = 32
    dx = 1.0 / res
    elastic = (O.MATERIALS_C2[0],)
    st, pr = O.scene_3d_cubes(n=4000, res=res, seed=12, side=1.7 * dx,
                              corners=((1.0 - 2.35 * dx, 0.5, 1.0 - 2.35 * dx),), materials=elastic,
                              vel_sigma=0.0)
    st["x"][:, 1] = 0.5 + 0.2 * np.random.default_rng(2).random(len(st["x"]))
    st["v"] = np.ascontiguousarray(np.tile([3.0, 0.0, 2.0], (len(st["x"]), 1)))  # toward the walls
    t = st["x"] * res - 0.5
    assert t[:, [0, 2]].max() < res - 1 and (np.trunc(t[:, 0]) == res - 2).sum() > 1000
    ts = to_torch_state(st, gpu)
    ref = oracle_copy_from(ts, st)
    eng = engine_for(st, pr, gpu, sort_every=2)
    nsteps = 3
    out = engine_roundtrip(eng, ts, nsteps)
    for _ in range(nsteps):
        O.substep(ref, pr)
    errs = field_errors(out, ref, pr["inv_dx"])
    _say("high walls", errs)
    for k in ("x", "v", "F"):
        assert errs[k] < 1e-5, errs
    assert errs["C"] < 1e-4 and errs["Jp"] < 1e-5, errs
    eng.close()


def test_engine_run_splits_are_invariant(gpu):
    """The same 60 substeps as one run, as two runs of 30 and as eight runs of
    7 plus one of 4.  The rebin schedule (every 16 substeps) counts across
    runs, so only the run boundaries differ: a split run ends with a G2P that
    stores v and C and starts with a P2G that reads them back, where one run
    keeps them in registers through the fused pass.  With no FP64-deferred
    particles every split gives the same state to FP32 rounding: the
    fixed-point arena sums do not depend on the order of the scatter, and
    only the few contributions that reach a node through global float
    atomics (a particle that crossed a block face since its binning) may
    add in another order."""
    st, pr = O.scene_3d_cubes(n=20000, res=64, seed=51)
    ts = to_torch_state(st, gpu)
    eng = engine_for(st, pr, gpu, sort_every=16)
    outs = []
    for splits in ([60], [30, 30], [7] * 8 + [4]):
        eng.upload(ts["x"], ts["v"], ts["F"], ts["C"], ts["Jp"], ts["mass"], ts["mu0"], ts["lam0"],
                   ts["kind"])
        eng.fallback_count()  # reading resets it
        for k in splits:
            eng.run(k)
        out = {k: torch.empty_like(ts[k]) for k in ("x", "v", "F", "C", "Jp")}
        eng.download(out["x"], out["v"], out["F"], out["C"], out["Jp"])
        torch.cuda.synchronize()
        fb = eng.fallback_count()
        _say("splits", splits, "fallbacks", fb)
        assert fb == 0  # the comparison below is of the FP32 path only
        outs.append(out)
    for other in outs[1:]:
        for k in ("x", "v", "F", "C", "Jp"):
            a, b = outs[0][k].double(), other[k].double()
            d = float((a - b).abs().max()) / max(float(a.abs().max()), 1e-30)
            _say("split difference", k, d)
            assert d <= 1e-6, (k, d)
    eng.close()
