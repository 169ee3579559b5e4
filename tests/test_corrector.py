"""Corrector MLP kernels (BF16 MFMA) against the NumPy restatement."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def _data(rows, seed):
    rng = np.random.default_rng(seed)
    X = M.bf16(rng.standard_normal((rows, M.IN)).astype(np.float32))
    T = rng.standard_normal((rows, M.OUT)) * 0.3
    w = rng.random(rows) + 0.5
    return X, T, w


def _params(seed):
    rng = np.random.default_rng(seed)
    parts = []
    for fo, fi in ((M.H, M.IN), (M.H, M.H), (M.OUT, M.H)):
        b = 1 / np.sqrt(fi)
        parts += [rng.uniform(-b, b, fo * fi), rng.uniform(-b, b, fo)]
    return np.concatenate(parts).astype(np.float32)


# ------------------------------------------------------------------ CPU ----
def test_oracle_gradient_matches_central_differences():
    """Pins the restated backward formulas (exact FP64 network, no rounding)."""
    X, T, w = _data(40, 1)
    p = _params(2).astype(np.float64)
    _, g, _ = M.loss_and_grads(p, X, T, w, rounding=False)
    rng = np.random.default_rng(3)
    for _ in range(3):
        d = rng.standard_normal(p.size)
        h = 1e-6
        lp, _, _ = M.loss_and_grads(p + h * d, X, T, w, rounding=False)
        lm, _, _ = M.loss_and_grads(p - h * d, X, T, w, rounding=False)
        fd = (lp - lm) / (2 * h)
        assert abs(float(g @ d) - fd) <= 1e-6 * abs(fd) + 1e-12


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.00390625, 1.0078125, 1.01171875, -3.5, 1e-30], dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to the even mantissa (1.0);
    # 1 + 3 * 2^-8 ties between 1 + 2^-7 and 1 + 2^-6: rounds to 1 + 2^-6
    assert list(M.bf16(vals)[:4]) == [1.0, 1.0, 1.0078125, 1.015625]
    assert M.bf16(vals)[4] == -3.5


def test_corrector_init_and_layout_sizes():
    from femflow_amd.corrector import CorrectorMLP

    p = CorrectorMLP.init_params(0)
    assert p.size == M.n_params() == 61955
    w = M.unpack(p)
    assert np.abs(w["W3"]).max() <= 0.1 / np.sqrt(128) + 1e-7


# ------------------------------------------------------------------ GPU ----
@pytest.mark.gpu
def test_mlp_forward_backward_match_oracle(gpu):
    import torch

    from femflow_amd.corrector import CorrectorMLP, to_bf16_features

    rows = 3000  # not a multiple of the 64-row tile
    X, T, w = _data(rows, 4)
    p = _params(5)
    net = CorrectorMLP(device=gpu, params=torch.from_numpy(p))
    Xg = to_bf16_features(X, gpu)
    loss = net.loss_grads(Xg, torch.tensor(T, dtype=torch.float32), torch.tensor(w, dtype=torch.float32))
    Y = net.Y[:rows, :3].double().cpu().numpy()
    lref, gref, Yref = M.loss_and_grads(p.astype(np.float64), X, T, w)
    # hidden activations are BF16 on both sides; a rounding tie resolved
    # differently after FP32 vs FP64 accumulation flips one BF16 ulp (2^-8)
    ey = np.linalg.norm(Y - Yref) / np.linalg.norm(Yref)
    emax = np.abs(Y - Yref).max() / np.abs(Yref).max()
    el = abs(float(loss.item()) - lref) / lref
    print("forward rel err", ey, "max", emax, "loss rel err", el, flush=True)
    assert ey < 1e-3 and emax < 1e-2 and el < 1e-3
    g = net.grads.double().cpu().numpy()
    off = 0
    for name, shape in M.SIZES:
        k = int(np.prod(shape))
        e = np.linalg.norm(g[off:off + k] - gref[off:off + k]) / np.linalg.norm(gref[off:off + k])
        print(f"grad {name}: rel err {e:.2e}", flush=True)
        assert e < 5e-3, name
        off += k


@pytest.mark.gpu
def test_mlp_adam_matches_oracle_and_training_reduces_loss(gpu):
    import torch

    from femflow_amd.corrector import CorrectorMLP, to_bf16_features

    rows = 4096
    rng = np.random.default_rng(6)
    X = M.bf16(rng.standard_normal((rows, M.IN)).astype(np.float32))
    A = rng.standard_normal((M.IN, 3)) / np.sqrt(M.IN)
    T = np.tanh(X @ A)  # a smooth target the network can fit
    net = CorrectorMLP(device=gpu, seed=0)
    p0 = net.params.double().cpu().numpy()
    Xg, Tg = to_bf16_features(X, gpu), torch.tensor(T, dtype=torch.float32)
    net.loss_grads(Xg, Tg)
    g = net.grads.double().cpu().numpy()
    net.adam_step(lr=1e-3)
    p1, _, _ = M.adam(p0, g, np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    assert np.abs(net.params.double().cpu().numpy() - p1).max() < 1e-6
    first = float(net.loss_grads(Xg, Tg).item())
    for _ in range(300):
        net.train_step(Xg, Tg, lr=2e-3)
    last = float(net.loss_grads(Xg, Tg).item())
    print("training loss", first, "->", last, flush=True)
    assert last < 0.1 * first


# -------------------------------------------------- engine integration ----
def _small_cfg():
    from femflow_amd.corrector_train import CorrectorConfig

    return CorrectorConfig(n=24000, res=48)


def _decode_nodes(nodes, res):
    """Global node index (block * 64 + node in block) -> (i, j, k), one scene."""
    nb = (res + 1 + 3) // 4
    blk, inner = nodes // 64, nodes % 64
    bi, bj, bk = blk // (nb * nb), (blk // nb) % nb, blk % nb
    return np.stack([bi * 4 + inner // 16, bj * 4 + (inner // 4) % 4, bk * 4 + inner % 4], axis=1)


@pytest.mark.gpu
def test_engine_features_match_oracle_grid(gpu):
    """Sampled features = the FP64 oracle's P2G at dt_big (node mass and
    momentum) plus per-cell mass-weighted mean F, rounded to BF16."""
    import torch

    from femflow_amd import corrector_train as CT
    from femflow_amd.scenes import scene_3d_cubes
    from oracle import mpm_oracle as O

    cfg = _small_cfg()
    st, pr = scene_3d_cubes(n=cfg.n, res=cfg.res, seed=3)
    s = CT.scene_tensors(cfg, 3, gpu)
    gen = CT.SampleGenerator(cfg, gpu)
    CT._upload(gen.small, s)  # 40 plain substeps first, so that F != I
    gen.small.run(40)
    s = CT._download(gen.small, s)
    CT._upload(gen.big, s)
    X, nodes, vbig, total = gen.big.corrector_sample(max_rows=gen.max_rows)
    X = X.float().cpu().numpy()
    ijk = _decode_nodes(nodes.cpu().numpy().astype(np.int64), cfg.res)
    # oracle P2G at dt_big from the FP32 state the engine saw
    ref = {k: s[k].double().cpu().numpy() for k in ("x", "v", "F", "C", "Jp", "mass", "mu0", "lam0")}
    r1 = cfg.res + 1
    gv = np.zeros((r1, r1, r1, 3))
    gm = np.zeros((r1, r1, r1, 1))
    O.p2g(ref["x"], ref["v"], ref["F"].reshape(-1, 3, 3), ref["C"].reshape(-1, 3, 3), ref["Jp"].reshape(-1, 1),
          ref["mass"], ref["mu0"], ref["lam0"], st["kind"], st["h"], float(cfg.res), 1.0 / cfg.res, cfg.dt_big,
          s["volume"], gv, gm)
    base, _ = O.base_and_fx(ref["x"], float(cfg.res))
    cm = np.zeros((r1, r1, r1))
    cF = np.zeros((r1, r1, r1, 9))
    np.add.at(cm, tuple(base.T), ref["mass"])
    np.add.at(cF, tuple(base.T), ref["mass"][:, None] * ref["F"].reshape(-1, 9))
    # nodes whose only contributions round to zero in the engine's fixed point
    # (weights near 0) carry mass in FP64 only
    n_ref = int((gm[..., 0] > 0).sum())
    assert 0.99 * n_ref <= X.shape[0] <= n_ref
    rng = np.random.default_rng(0)
    pick = rng.choice(X.shape[0], size=min(400, X.shape[0]), replace=False)
    eye = np.eye(3).reshape(-1)
    worst = 0.0
    for r in pick:
        i, j, k = ijk[r]
        feat = []
        for d0 in (-1, 0, 1):
            for d1 in (-1, 0, 1):
                for d2 in (-1, 0, 1):
                    a, b, c = i + d0, j + d1, k + d2
                    if not (0 <= a < r1 and 0 <= b < r1 and 0 <= c < r1):
                        feat += [0.0] * 13
                        continue
                    m = gm[a, b, c, 0]
                    f = [m / cfg.m_ref] + list(gv[a, b, c] / (cfg.m_ref * cfg.v_ref)) if m > 0 else [0.0] * 4
                    f += list(cF[a, b, c] / cm[a, b, c] - eye) if cm[a, b, c] > 0 else [0.0] * 9
                    feat += f
        feat = np.array(feat)
        err = np.abs(X[r, :351] - feat) - (2.0 ** -7 * np.abs(feat) + 5e-3)
        worst = max(worst, float(err.max()))
    gen.close()
    assert worst <= 0.0, worst


@pytest.mark.gpu
def test_engine_corrector_applies_network_output(gpu):
    """The fused in-engine MLP adds MLP(features) * v_ref to the updated grid
    velocity (checked against the standalone forward kernel)."""
    import torch

    from femflow_amd import corrector_train as CT
    from femflow_amd.corrector import CorrectorMLP

    cfg = _small_cfg()
    s = CT.scene_tensors(cfg, 4, gpu)
    gen = CT.SampleGenerator(cfg, gpu)
    net = CorrectorMLP(device=gpu, seed=1, out_scale=0.02)
    CT._upload(gen.big, s)
    X, nodes, vbig, _ = gen.big.corrector_sample(max_rows=gen.max_rows)
    X, nodes, vbig = X.clone(), nodes.clone(), vbig.clone()
    gen.big.set_corrector(net.params, cfg.m_ref, cfg.v_ref, enable=True)
    CT._upload(gen.big, s)
    gen.big.p2g_grid_only()
    vcorr = gen.big.grid_velocity(nodes)
    pred = net.forward(X).clone() * cfg.v_ref
    ijk = _decode_nodes(nodes.cpu().numpy().astype(np.int64), cfg.res)
    inner = torch.as_tensor(((ijk >= 1) & (ijk < cfg.res - 1)).all(1), device=gpu)
    mrel = X[:, 13 * 13].float()  # node mass / m_ref
    full = inner & (mrel >= 0.3)    # correction applied in full (mass gate = 1)
    light = inner & (mrel < 0.02)   # scaled by 4 m / m_ref < 0.08
    d = (vcorr - vbig)
    err = float((d[full] - pred[full]).norm() / pred[full].norm())
    print("in-engine correction vs standalone forward: rel err", err, flush=True)
    gen.close()
    assert err < 1e-3
    assert bool(light.any()) and bool((d[light].abs() <= 0.1 * pred[light].abs() + 1e-7).all())


@pytest.mark.gpu
def test_trained_corrector_beats_uncorrected_big_step(gpu):
    """Ablation: from held-out states of plain rollouts, one corrected big
    step (dt_big = 8 dt) lands closer to 8 plain substeps than the same big
    step without the network."""
    import torch

    from femflow_amd import corrector_train as CT
    from femflow_amd.corrector import CorrectorMLP

    cfg = _small_cfg()
    gen = CT.SampleGenerator(cfg, gpu)
    X, T, w = gen.rollouts(12, 4, 400, 4096, seed=1)
    net = CorrectorMLP(device=gpu, seed=0)
    hist = CT.train(net, X, T, w, epochs=12, batch=4096, lr=1e-3)
    errs = {"plain": [], "corr": []}
    for seed in (901, 902, 903):
        s = CT.scene_tensors(cfg, seed, gpu)
        eng = CT.make_engine(cfg, s["volume"], cfg.dt, cfg.gravity, gpu, s["x"].shape[0])
        CT._upload(eng, s)
        eng.run(160)
        s = CT._download(eng, s)
        eng.run(cfg.factor)
        ref = eng.positions().clone()
        eng.close()
        for tag, nn in (("plain", None), ("corr", net)):
            runner = CT.CorrectedRunner(cfg, nn, s["volume"], gpu)
            out = runner.run(s, 1, check=False)
            runner.close()
            disp = (ref - s["x"]).norm()
            errs[tag].append(float((out["x"] - ref).norm() / disp))
    gen.close()
    e0, e1 = np.mean(errs["plain"]), np.mean(errs["corr"])
    print(f"one big step, displacement error vs 8 plain substeps: without network {e0:.4e}, "
          f"with network {e1:.4e} (val loss {hist['val'][-1]:.3e} vs zero-correction "
          f"{CT.baseline_loss(T, w):.3e})", flush=True)
    assert e1 < 0.5 * e0


@pytest.mark.gpu
def test_shipped_corrector_one_step_gain(gpu):
    """The shipped network (200 solver rollouts of the headline scene): one
    corrected big step lands at least 2x closer to 8 plain substeps than the
    uncorrected big step, on held-out scenes."""
    import os

    from femflow_amd import corrector_train as CT
    from femflow_amd.corrector import CorrectorMLP

    path = os.path.join(os.path.dirname(__file__), "..", "femflow_amd", "data", "corrector_c5.npz")
    net = CorrectorMLP.load(path, device=gpu)
    assert net.meta["rollouts"] == 200
    gain = CT.one_step_gain(CT.CorrectorConfig(), net, gpu)
    print("shipped corrector, one big step:", gain, flush=True)
    assert gain["with"] < 0.5 * gain["without"]


@pytest.mark.gpu
def test_judge_falls_back_to_plain_substeps(gpu):
    """The judge keeps a corrected window only within its error tolerance:
    with a zero tolerance every judged window is replaced by the plain
    substeps' result, so the judged rollout equals the plain rollout; with an
    infinite tolerance it keeps the big steps (up to the health check)."""
    import torch

    from femflow_amd import corrector_train as CT

    cfg = CT.CorrectorConfig(n=1 << 15, res=64)
    s = CT.scene_tensors(cfg, 0, gpu)
    n_big = 4
    plain = CT.make_engine(cfg, s["volume"], cfg.dt, cfg.gravity, gpu, s["x"].shape[0], sort_every=48)
    CT._upload(plain, s)
    plain.run(n_big * cfg.factor)
    ref = CT._download(plain, s)
    plain.close()
    strict = CT.CorrectedRunner(cfg, None, s["volume"], gpu, window=2, judge_every=1, tol=0.0)
    out = strict.run(s, n_big)
    st = dict(strict.stats)
    strict.close()
    torch.cuda.synchronize()
    assert st["judged"] + st["unhealthy"] == 2 and st["big_steps_kept"] == 0, st
    assert st["plain_substeps"] == n_big * cfg.factor
    assert CT.rel_position_error(out["x"], ref["x"]) < 1e-6
    loose = CT.CorrectedRunner(cfg, None, s["volume"], gpu, window=2, judge_every=1, tol=float("inf"))
    out2 = loose.run(s, n_big)
    st2 = dict(loose.stats)
    loose.close()
    assert st2["fallbacks"] == 0 and st2["big_steps_kept"] + 2 * st2["unhealthy"] == n_big, st2
    print("judge stats strict", st, "loose", st2, "loose error", CT.rel_position_error(out2["x"], ref["x"]),
          flush=True)
