"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Tolerances (relative, per field, norms of tests/_parity.py):
  * FP64 GPU kernels vs the FP64 oracle: 1e-9 (only summation order and
    atomic order differ).
  * FP32 GPU vs FP64 oracle, one substep from identical inputs: 1e-5 for x, v
    and F, 1e-4 for C and Jp.
  * several substeps: the bound grows with the number of steps (reported).
Binning keys and permutations: bit-exact.
"""
import os

import numpy as np
import pytest
import torch

from oracle import c_oracle as CO
from oracle import mpm_oracle as O
from tests._parity import field_errors
from tests.gpu_helpers import (engine_for, engine_roundtrip, oracle_copy_from, to_numpy_state,
                               to_torch_state)

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print(*a, flush=True)


# ------------------------------------------------------------- binning ----
@pytest.mark.parametrize("n,res,layout", [(1, 8, "blocked"), (4095, 32, "blocked"), (4097, 32, "rowmajor"),
                                          (100_000, 128, "blocked"), (100_000, 128, "rowmajor"),
                                          (300_001, 100, "blocked"),
                                          # keys past 22 bits: three 11-bit radix passes
                                          (200_003, 256, "blocked"), (60_000, 300, "rowmajor")])
def test_bin_bit_exact(gpu, n, res, layout):
    from femflow_amd.engine import bin_particles

    _say("bin", n, res, layout)
    rng = np.random.default_rng(n + res)
    x = (rng.random((n, 3)) * 0.6 + 0.2).astype(np.float32)
    # clustered duplicates: many particles in few cells exercise stability
    x[: n // 3] = x[: n // 3].round(2)
    xt = torch.tensor(x, device=gpu)
    keys, perm = bin_particles(xt, res, float(res), layout)
    ref_keys = O.bin_keys(x, float(res), res, layout=layout)
    ref_perm = O.bin_sort(ref_keys)
    assert np.array_equal(keys.cpu().numpy().view(np.uint32), ref_keys)
    assert np.array_equal(perm.cpu().numpy().view(np.uint32), ref_perm)


def test_bin_2d_bit_exact(gpu):
    from femflow_amd.engine import bin_particles

    rng = np.random.default_rng(11)
    x = rng.random((50_000, 2)).astype(np.float32)
    keys, perm = bin_particles(torch.tensor(x, device=gpu), 128, 128.0, "rowmajor")
    ref = O.bin_keys(x, 128.0, 128, layout="rowmajor")
    assert np.array_equal(keys.cpu().numpy().view(np.uint32), ref)
    assert np.array_equal(perm.cpu().numpy().view(np.uint32), O.bin_sort(ref))


# ---------------------------------------------- drop-in step functions ----
def _particle_list(x, mass, mu0, lam0):
    from femflow_amd.solvers.mpm import Particle

    return [Particle(x[i].copy(), -9.8, mass[i], lam0[i], mu0[i]) for i in range(len(x))]


def test_solve_mls_mpm_3d_fp64_numpy(gpu):
    """Reference signature with a Particle list and float64 numpy state."""
    from femflow_amd.solvers.mpm import solve_mls_mpm_3d

    st, pr = O.scene_3d_cubes(n=3000, res=32)
    st["kind"][:] = O.ELASTIC
    parts = _particle_list(st["x"], st["mass"], st["mu0"], st["lam0"])
    v, F, C, Jp = (st[k].copy() for k in ("v", "F", "C", "Jp"))
    solve_mls_mpm_3d(pr["res"], pr["inv_dx"], 1.0, pr["dx"], pr["dt"], pr["volume"], pr["gravity"],
                     parts, v, F, C, Jp)
    ref = O.copy_state(st)
    O.solve_mls_mpm_3d(pr["res"], pr["inv_dx"], 1.0, pr["dx"], pr["dt"], pr["volume"], pr["gravity"],
                       ref["x"], ref["mass"], ref["mu0"], ref["lam0"], ref["v"], ref["F"], ref["C"], ref["Jp"])
    got = dict(x=np.array([p.pos for p in parts]), v=v, F=F, C=C, Jp=Jp)
    errs = field_errors(got, ref, pr["inv_dx"])
    _say("fp64 solve errs", errs)
    assert max(errs.values()) < 1e-9, errs


def test_solve_mls_mpm_3d_fp32_engine(gpu):
    from femflow_amd.solvers.mpm import ParticleSet, solve_mls_mpm_3d

    st, pr = O.scene_3d_cubes(n=20000, res=64)
    st["kind"][:] = O.ELASTIC
    ts = to_torch_state(st, gpu)
    ps = ParticleSet(ts["x"], ts["mass"], ts["mu0"], ts["lam0"])
    ref = oracle_copy_from(ts, st)
    Jp = ts["Jp"].view(-1, 1)
    solve_mls_mpm_3d(pr["res"], pr["inv_dx"], 1.0, pr["dx"], pr["dt"], pr["volume"], pr["gravity"], ps,
                     ts["v"], ts["F"], ts["C"], Jp)
    O.solve_mls_mpm_3d(pr["res"], pr["inv_dx"], 1.0, pr["dx"], pr["dt"], pr["volume"], pr["gravity"],
                       ref["x"], ref["mass"], ref["mu0"], ref["lam0"], ref["v"], ref["F"], ref["C"], ref["Jp"])
    errs = field_errors(to_numpy_state(ts), ref, pr["inv_dx"])
    _say("fp32 engine solve errs", errs)
    for k in ("x", "v", "F"):
        assert errs[k] < 1e-5, errs
    assert errs["C"] < 1e-4 and errs["Jp"] < 1e-5, errs


def test_solve_mls_mpm_3d_fp32_fewer_then_more_particles(gpu):
    """Successive FP32 calls whose particle counts change (2000, then 1000,
    then 3000) reuse the cached engine; each must match the FP64 oracle.  A
    graph captured for one count must never be replayed for another."""
    from femflow_amd.solvers.mpm import ParticleSet, solve_mls_mpm_3d

    for n in (2000, 1000, 3000):
        st, pr = O.scene_3d_cubes(n=n, res=32, seed=n)
        st["kind"][:] = O.ELASTIC
        ts = to_torch_state(st, gpu)
        ps = ParticleSet(ts["x"], ts["mass"], ts["mu0"], ts["lam0"])
        ref = oracle_copy_from(ts, st)
        solve_mls_mpm_3d(pr["res"], pr["inv_dx"], 1.0, pr["dx"], pr["dt"], pr["volume"], pr["gravity"], ps,
                         ts["v"], ts["F"], ts["C"], ts["Jp"].view(-1, 1))
        O.solve_mls_mpm_3d(pr["res"], pr["inv_dx"], 1.0, pr["dx"], pr["dt"], pr["volume"], pr["gravity"],
                           ref["x"], ref["mass"], ref["mu0"], ref["lam0"], ref["v"], ref["F"], ref["C"],
                           ref["Jp"])
        errs = field_errors(to_numpy_state(ts), ref, pr["inv_dx"])
        _say("fp32 solve, n =", n, errs)
        for k in ("x", "v", "F"):
            assert errs[k] < 1e-5, (n, errs)


def test_engine_reupload_changes_count(gpu):
    """One engine, uploads of 4000 then 1500 then 4000 particles, several
    graph-replayed substeps each, against the oracle; the downloads land only
    in the caller's arrays of the current count."""
    st, pr = O.scene_3d_cubes(n=4000, res=32)
    eng = engine_for(st, pr, gpu, sort_every=4)
    for n in (4000, 1500, 4000):
        sub = {k: (a[:n].copy() if isinstance(a, np.ndarray) and len(a) == 4000 else a) for k, a in st.items()}
        ts = to_torch_state(sub, gpu)
        guard = torch.full((n + 1000, 3), 7.0, device=gpu)
        x_out = guard[:n]
        eng.upload(ts["x"], ts["v"], ts["F"], ts["C"], ts["Jp"], ts["mass"], ts["mu0"], ts["lam0"], ts["kind"])
        eng.run(6)
        out = {k: torch.empty_like(ts[k]) for k in ("v", "F", "C", "Jp")}
        eng.download(x_out, out["v"], out["F"], out["C"], out["Jp"])
        torch.cuda.synchronize()
        out["x"] = x_out
        ref = oracle_copy_from(ts, sub)
        for _ in range(6):
            O.substep(ref, pr)
        errs = field_errors(to_numpy_state(out), ref, pr["inv_dx"])
        _say("reupload", n, errs)
        # about ten times the error measured after 6 substeps (x 2e-7,
        # v 5e-6, F 2.3e-7)
        for k, b in (("x", 5e-6), ("v", 5e-5), ("F", 5e-6)):
            assert errs[k] < b, (n, k, errs)
        assert bool((guard[n:] == 7.0).all()), "download wrote past the current particle count"
    eng.close()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("model", ["neo_hookean", "snow"])
def test_three_d_phases(gpu, dtype, model):
    from femflow_amd.solvers.mpm import ParticleSet, three_d

    st, pr = O.scene_3d_cubes(n=5000, res=32)
    rng = np.random.default_rng(12)
    st["F"] = st["F"] + 0.02 * rng.standard_normal(st["F"].shape)
    st["C"] = 0.5 * rng.standard_normal(st["C"].shape)
    ts = to_torch_state(st, gpu, dtype)
    ref = oracle_copy_from(ts, st)
    h = 10.0 if model == "snow" else 1.0
    r1 = pr["res"] + 1
    gv = torch.zeros((r1, r1, r1, 3), dtype=dtype, device=gpu)
    gm = torch.zeros((r1, r1, r1, 1), dtype=dtype, device=gpu)
    ps = ParticleSet(ts["x"], ts["mass"], ts["mu0"], ts["lam0"])
    Jp = ts["Jp"].view(-1, 1)
    three_d.p2g(pr["inv_dx"], h, pr["dx"], pr["dt"], pr["volume"], gv, gm, ps, ts["v"], ts["F"], ts["C"], Jp,
                model)
    rgv, rgm = np.zeros((r1, r1, r1, 3)), np.zeros((r1, r1, r1, 1))
    O.p2g_3d(pr["inv_dx"], h, pr["dx"], pr["dt"], pr["volume"], rgv, rgm, ref["x"], ref["mass"], ref["mu0"],
             ref["lam0"], ref["v"], ref["F"], ref["C"], ref["Jp"], model)
    tol = 1e-9 if dtype == torch.float64 else 2e-5
    gm_np = gm.double().cpu().numpy()
    gv_np = gv.double().cpu().numpy()
    assert np.abs(gm_np - rgm).max() <= tol * np.abs(rgm).max()
    assert np.abs(gv_np - rgv).max() <= tol * np.abs(rgv).max()
    three_d.grid_op(pr["res"], pr["dx"], pr["dt"], pr["gravity"], gv, gm)
    O.grid_op_3d(pr["res"], pr["dx"], pr["dt"], pr["gravity"], rgv, rgm)
    gv_np = gv.double().cpu().numpy()
    assert np.abs(gv_np - rgv).max() <= tol * max(np.abs(rgv).max(), 1e-6)
    three_d.g2p(pr["inv_dx"], pr["dt"], gv, ps, ts["v"], ts["F"], ts["C"], Jp, model)
    O.g2p_3d(pr["inv_dx"], pr["dt"], rgv, ref["x"], ref["v"], ref["F"], ref["C"], ref["Jp"], model)
    errs = field_errors(to_numpy_state(ts), ref, pr["inv_dx"])
    _say("three_d phases", dtype, model, errs)
    lim = 1e-9 if dtype == torch.float64 else 1e-5
    for k in ("x", "v", "F"):
        assert errs[k] < lim, errs
    assert errs["C"] < lim * 10 and errs["Jp"] < lim * 10, errs


def test_three_d_p2g_raises_oob(gpu):
    from femflow_amd.solvers.mpm import ParticleSet, three_d

    st, pr = O.scene_3d_cubes(n=100, res=32)
    st["x"][5] = [0.5, 0.999, 0.5]
    ts = to_torch_state(st, gpu, torch.float64)
    r1 = pr["res"] + 1
    gv = torch.zeros((r1, r1, r1, 3), dtype=torch.float64, device=gpu)
    gm = torch.zeros((r1, r1, r1, 1), dtype=torch.float64, device=gpu)
    ps = ParticleSet(ts["x"], ts["mass"], ts["mu0"], ts["lam0"])
    with pytest.raises(RuntimeError):
        three_d.p2g(pr["inv_dx"], 1.0, pr["dx"], pr["dt"], pr["volume"], gv, gm, ps, ts["v"], ts["F"],
                    ts["C"], ts["Jp"].view(-1, 1))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_two_d_config1_steps(gpu, dtype):
    """Config-1 scene (8,192 particles, res 128) through two_d.{p2g,grid_op,g2p}."""
    from femflow_amd.solvers.mpm import two_d

    st, pr = O.scene_2d_c1()
    n = len(st["x"])
    dev = gpu
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)  # noqa: E731
    x, v, F, C, Jp = t(st["x"]), t(st["v"]), t(st["F"]), t(st["C"]), t(st["Jp"])
    ref = O.copy_state(st)
    ref["x"] = x.double().cpu().numpy()
    r1 = pr["res"] + 1
    mu0, lam0, mass = st["mu0"][0], st["lam0"][0], st["mass"][0]
    worst = {}
    for step in range(10):
        gv = torch.zeros((r1, r1, 2), dtype=dtype, device=dev)
        gm = torch.zeros((r1, r1), dtype=dtype, device=dev)
        two_d.p2g(pr["inv_dx"], 1.0, mu0, lam0, mass, pr["dx"], pr["dt"], pr["volume"], gv, gm, x, v, F, C, Jp)
        two_d.grid_op(pr["res"], pr["dt"], pr["gravity"], gv, gm)
        two_d.g2p(pr["inv_dx"], pr["dt"], gv, x, v, F, C, Jp)
        O.substep(ref, pr)
        errs = field_errors(dict(x=x, v=v, F=F, C=C, Jp=Jp), ref, pr["inv_dx"])
        worst = {k: max(worst.get(k, 0), e) for k, e in errs.items()}
        if dtype == torch.float32:  # restart from the GPU state: per-substep error
            for k, a in (("x", x), ("v", v), ("F", F), ("C", C), ("Jp", Jp)):
                ref[k] = a.double().cpu().numpy().reshape(ref[k].shape)
    _say("two_d", dtype, worst)
    if dtype == torch.float64:
        g = np.load(os.path.join(GOLDEN, "oracle_c1_2d.npz"))
        got = dict(x=x, v=v, F=F, C=C, Jp=Jp)
        gold = {k: g[f"s10_{k}"] for k in ("x", "v", "F", "C", "Jp")}
        e2 = field_errors(got, gold, pr["inv_dx"])
        assert max(e2.values()) < 1e-8, e2
        assert max(worst.values()) < 1e-8, worst
    else:
        for k in ("x", "v", "F"):
            assert worst[k] < 1e-5, worst
        assert worst["Jp"] < 1e-5, worst


# ------------------------------------------------------------- engine ----
@pytest.mark.parametrize("quirks", ["reference", "textbook"])
def test_engine_one_substep_three_materials(gpu, quirks):
    st, pr = O.scene_3d_cubes(n=30000, res=64)
    ts = to_torch_state(st, gpu)
    ref = oracle_copy_from(ts, st)
    from femflow_amd.engine import Engine, SceneSpec
    from tests.gpu_helpers import material_table_c2

    eng = Engine([SceneSpec(res=pr["res"], dt=pr["dt"], gravity=pr["gravity"], volume=pr["volume"],
                            quirks=quirks)], material_table_c2(), n_max=len(st["x"]), device=gpu)
    out = engine_roundtrip(eng, ts, 1)
    O.substep(ref, pr, quirks=quirks)
    errs = field_errors(out, ref, pr["inv_dx"])
    _say("engine 1 step", quirks, errs)
    for k in ("x", "v", "F"):
        assert errs[k] < 1e-5, errs
    assert errs["C"] < 1e-4 and errs["Jp"] < 1e-5, errs
    eng.close()


@pytest.mark.parametrize("sort_every,use_graph,chunk,spread", [(1, False, 512, 3), (4, False, 256, 3),
                                                                (4, True, 512, 3), (8, True, 1024, 3),
                                                                (16, True, 512, 3), (4, True, 512, 9),
                                                                (8, False, 1024, 9)])
def test_engine_fused_multistep(gpu, sort_every, use_graph, chunk, spread):
    st, pr = O.scene_3d_cubes(n=30000, res=64)
    ts = to_torch_state(st, gpu)
    ref = oracle_copy_from(ts, st)
    eng = engine_for(st, pr, gpu, sort_every=sort_every, use_graph=use_graph, chunk=chunk, spread=spread)
    nsteps = 10
    out = engine_roundtrip(eng, ts, nsteps)
    for _ in range(nsteps):
        O.substep(ref, pr)
    errs = field_errors(out, ref, pr["inv_dx"])
    _say("engine fused", sort_every, use_graph, chunk, spread, errs)
    # about ten times the accumulated FP32 error measured after 10 substeps
    # (x 3.3e-7, v 1.8e-6, F 2e-8, C 4e-7, Jp 1e-9, the same for every launch
    # shape here)
    bars = {"x": 5e-6, "v": 2e-5, "F": 1e-6, "C": 5e-6, "Jp": 1e-7}
    for k, b in bars.items():
        assert errs[k] < b, (k, errs)
    eng.close()


def test_engine_graph_replays_match_eager(gpu):
    """Replays of the captured graph (with eager work and syncs between them)
    agree with the eager sequence on every replay."""
    st, pr = O.scene_3d_cubes(n=20000, res=64)
    ts = to_torch_state(st, gpu)
    e_eager = engine_for(st, pr, gpu, sort_every=4, use_graph=False)
    e_graph = engine_for(st, pr, gpu, sort_every=4, use_graph=True)
    for e in (e_eager, e_graph):
        e.upload(ts["x"], ts["v"], ts["F"], ts["C"], ts["Jp"], ts["mass"], ts["mu0"], ts["lam0"], ts["kind"])
    outs = {}
    for rep in range(4):
        for name, e in (("eager", e_eager), ("graph", e_graph)):
            e.run(4)
            o = {k: torch.empty_like(ts[k]) for k in ("x", "v", "F", "C", "Jp")}
            e.download(o["x"], o["v"], o["F"], o["C"], o["Jp"])
            torch.cuda.synchronize()
            outs[name] = o
            _ = torch.ones(1000, device=gpu).sum().item()  # eager work between replays
        errs = field_errors(outs["graph"], outs["eager"], pr["inv_dx"])
        _say("graph vs eager replay", rep, errs)
        assert max(errs[k] for k in ("x", "v", "F", "Jp")) < 1e-5, errs
    e_eager.close()
    e_graph.close()


def test_engine_conservation_sums(gpu):
    st, pr = O.scene_3d_cubes(n=30000, res=64)
    ts = to_torch_state(st, gpu)
    eng = engine_for(st, pr, gpu, sort_every=4)
    nsteps = 8
    sums = torch.zeros((nsteps, 8), dtype=torch.float64, device=gpu)
    eng.upload(ts["x"], ts["v"], ts["F"], ts["C"], ts["Jp"], ts["mass"], ts["mu0"], ts["lam0"], ts["kind"])
    p0 = eng.particle_sums()
    eng.run(nsteps, sums=sums)
    s = sums.cpu().numpy()
    mtot = ts["mass"].double().sum().item()
    _say("sums", s[:, :4], p0)
    assert np.allclose(s[:, 0], mtot, rtol=1e-5)
    # APIC P2G conserves momentum: first step's grid momentum = sum m v
    assert np.allclose(s[0, 1:4], p0[1:4], rtol=1e-4, atol=1e-6 * mtot)
    # walls are far away: the grid update adds exactly dt g sum m on axis 1
    assert np.allclose(s[:, 5] - s[:, 2], pr["dt"] * pr["gravity"] * mtot, rtol=1e-3)
    assert np.allclose(s[:, 4], s[:, 1], rtol=1e-5, atol=1e-9)
    eng.close()


def test_engine_batched_scenes_match_single(gpu):
    """Three scenes of different resolutions in one launch sequence agree with
    each scene run alone against the oracle."""
    from femflow_amd.engine import Engine, SceneSpec
    from tests.gpu_helpers import material_table_c2

    specs = [(32, 3000, 1), (48, 5000, 2), (64, 8000, 3)]
    scenes, states, params = [], [], []
    for res, n, seed in specs:
        st, pr = O.scene_3d_cubes(n=n, res=res, seed=seed)
        scenes.append(SceneSpec(res=res, dt=pr["dt"], gravity=pr["gravity"], volume=pr["volume"]))
        states.append(st)
        params.append(pr)
    ntot = sum(len(s["x"]) for s in states)
    eng = Engine(scenes, material_table_c2(), n_max=ntot, device=gpu)
    eng.tune(sort_every=4)
    tss = [to_torch_state(s, gpu) for s in states]
    cat = {k: torch.cat([t[k] for t in tss]) for k in tss[0]}
    scene_id = torch.cat([torch.full((len(s["x"]),), i, dtype=torch.int32, device=gpu)
                          for i, s in enumerate(states)])
    eng.upload(cat["x"], cat["v"], cat["F"], cat["C"], cat["Jp"], cat["mass"], cat["mu0"], cat["lam0"],
               cat["kind"], scene=scene_id)
    eng.run(5)
    out = {k: torch.empty_like(cat[k]) for k in ("x", "v", "F", "C", "Jp")}
    eng.download(out["x"], out["v"], out["F"], out["C"], out["Jp"])
    off = 0
    for ts, st, pr in zip(tss, states, params):
        n = len(st["x"])
        ref = oracle_copy_from(ts, st)
        for _ in range(5):
            O.substep(ref, pr)
        got = {k: v[off:off + n] for k, v in out.items()}
        errs = field_errors(got, ref, pr["inv_dx"])
        _say("batched scene", pr["res"], errs)
        # about ten times the error measured after 5 substeps (x 1.7e-7,
        # v 4.2e-6, F 2.2e-7)
        for k, b in (("x", 5e-6), ("v", 5e-5), ("F", 5e-6)):
            assert errs[k] < b, (k, errs)
        off += n
    eng.close()


def test_engine_raises_oob(gpu):
    st, pr = O.scene_3d_cubes(n=1000, res=32)
    st["x"][10] = [0.5, 0.5, 0.999]
    ts = to_torch_state(st, gpu)
    eng = engine_for(st, pr, gpu)
    eng.upload(ts["x"], ts["v"], ts["F"], ts["C"], ts["Jp"], ts["mass"], ts["mu0"], ts["lam0"], ts["kind"])
    with pytest.raises(RuntimeError):
        eng.run(1)
    eng.close()


def test_engine_full_size_one_substep_vs_c_oracle(gpu):
    """Config 2 at its real size (2^20 particles, 128^3): one substep against
    the C restatement (FP64, OpenMP)."""
    st, pr = O.scene_3d_cubes(n=1 << 20, res=128)
    ts = to_torch_state(st, gpu)
    ref_in = oracle_copy_from(ts, st)
    eng = engine_for(st, pr, gpu)
    out = engine_roundtrip(eng, ts, 1)
    ref = CO.run(ref_in, pr, 1, nthreads=8).to_state()
    errs = field_errors(out, ref, pr["inv_dx"])
    _say("full size errs", errs)
    for k in ("x", "v", "F"):
        assert errs[k] < 1e-5, errs
    assert errs["C"] < 1e-4 and errs["Jp"] < 1e-5, errs
    eng.close()


def test_engine_256_grid_vs_c_oracle(gpu):
    """A 256^3 grid: 65^3 blocks, cell keys of 25 bits (three radix passes in
    the engine's own binning) and block indices past 2^18; three
    three-material cubes, 4 substeps with a rebin every 2, against the C
    restatement (FP64)."""
    st, pr = O.scene_3d_cubes(n=30000, res=256, seed=41, side=0.08)
    ts = to_torch_state(st, gpu)
    ref_in = oracle_copy_from(ts, st)
    eng = engine_for(st, pr, gpu, sort_every=2)
    out = engine_roundtrip(eng, ts, 4)
    ref = CO.run(ref_in, pr, 4, nthreads=8).to_state()
    errs = field_errors(out, ref, pr["inv_dx"])
    _say("256 grid errs", errs)
    for k in ("x", "v", "F"):
        assert errs[k] < 1e-5, errs
    assert errs["C"] < 1e-4 and errs["Jp"] < 1e-5, errs
    eng.close()
