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
