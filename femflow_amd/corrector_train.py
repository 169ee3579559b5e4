"""Training the learned corrector on the solver's own rollouts, and running
corrected rollouts with the plain substep as fallback.

Training data.  A rollout of the three-cube scene (three materials, seeded
cube velocities) runs plain substeps at dt.  At sampled times t (multiples
of factor = dt_big / dt) its state s(t) gives one training sample:
  * features and the big step's updated grid velocity v_big: one P2G + grid
    update of s(t) at dt_big (engine sampling mode);
  * target grid velocity v*: the P2G (mass-weighted average) of each
    particle's mean velocity over the big step, (x(t + dt_big) - x(t)) /
    dt_big, from `factor` plain substeps, with the wall conditions applied;
  * target correction (v* - v_big) / v_ref, weighted by node mass.
The applied correction is scaled by min(1, 4 m / m_ref): nodes with little
mass (surfaces, where the few contributing particles make the features
noisy) keep the plain big step's velocity.
A corrected big step then replaces `factor` plain substeps: P2G + grid update
at dt_big, plus the network's correction, then G2P.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional

import numpy as np
import torch

from .corrector import CorrectorMLP
from .engine import Engine, MaterialSpec, SceneSpec
from .scenes import MATERIALS_C2, scene_3d_cubes

MATERIAL_TABLE = [MaterialSpec(("elastic", "snow", "water")[k], h, tc, ts) for k, _, _, h, tc, ts in MATERIALS_C2]
CENTER = 13 * 13  # column of the node's own mass feature (neighbour (0,0,0))


@dataclass
class CorrectorConfig:
    n: int = 1 << 20
    res: int = 128
    dt: float = 5e-5
    factor: int = 8
    gravity: float = -9.8
    v_ref: float = 1.0
    rho: float = 1.0

    @property
    def dt_big(self):
        return self.dt * self.factor

    @property
    def m_ref(self):
        return self.rho / float(self.res) ** 3  # mass of a full node cell


def scene_tensors(cfg: CorrectorConfig, seed: int, device) -> Dict[str, torch.Tensor]:
    st, _ = scene_3d_cubes(n=cfg.n, res=cfg.res, seed=seed, dt=cfg.dt, gravity=cfg.gravity, rho=cfg.rho)
    n = len(st["x"])
    f = lambda a, shape=None: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32,  # noqa: E731
                                           device=device).reshape(shape if shape else np.shape(a))
    return dict(x=f(st["x"]), v=f(st["v"]), F=f(st["F"], (n, 3, 3)), C=f(st["C"], (n, 3, 3)),
                Jp=f(st["Jp"], (n,)), mass=f(st["mass"]), mu0=f(st["mu0"]), lam0=f(st["lam0"]),
                kind=torch.tensor(st["kind"], dtype=torch.int32, device=device),
                volume=float(st["mass"][0] / cfg.rho))


def make_engine(cfg: CorrectorConfig, volume: float, dt: float, gravity: float, device, n: int,
                chunk: int = 512, sort_every: int = 16) -> Engine:
    eng = Engine([SceneSpec(res=cfg.res, dt=dt, gravity=gravity, volume=volume)], MATERIAL_TABLE, n_max=n,
                 device=device)
    eng.tune(chunk=chunk, sort_every=sort_every, grid_mult=8)
    return eng


def _upload(eng: Engine, s: Dict[str, torch.Tensor], **over):
    g = lambda k: over.get(k, s[k])  # noqa: E731
    eng.upload(g("x"), g("v"), g("F"), g("C"), g("Jp"), s["mass"], s["mu0"], s["lam0"], s["kind"])


def _download(eng: Engine, s: Dict[str, torch.Tensor]) -> Dict[str, torch.Tensor]:
    out = dict(s)
    for k in ("x", "v", "F", "C", "Jp"):
        out[k] = torch.empty_like(s[k])
    eng.download(out["x"], out["v"], out["F"], out["C"], out["Jp"])
    return out


class SampleGenerator:
    """Engines for the plain rollout (dt), the big-step features (dt_big) and
    the target P2G (dt_big, no gravity, zero stress)."""

    def __init__(self, cfg: CorrectorConfig, device):
        self.cfg = cfg
        self.device = torch.device(device)
        s = scene_tensors(cfg, 0, device)
        self.n = s["x"].shape[0]
        vol = s["volume"]
        self.volume = vol
        self.small = make_engine(cfg, vol, cfg.dt, cfg.gravity, device, self.n)
        self.big = make_engine(cfg, vol, cfg.dt_big, cfg.gravity, device, self.n, sort_every=1)
        self.tgt = make_engine(cfg, vol, cfg.dt_big, 0.0, device, self.n, sort_every=1)
        dummy = torch.zeros(CorrectorMLP.init_params().size, dtype=torch.float32, device=device)
        self.big.set_corrector(dummy, cfg.m_ref, cfg.v_ref, enable=False)
        self.tgt.set_corrector(dummy, cfg.m_ref, cfg.v_ref, enable=False)
        self.eye = torch.eye(3, device=device).expand(self.n, 3, 3).contiguous()
        self.zero33 = torch.zeros((self.n, 3, 3), device=device)
        self.ones = torch.ones(self.n, device=device)
        self.max_rows = min(2 * self.n, (cfg.res + 1) ** 3, 1 << 19)

    def sample(self, s: Dict[str, torch.Tensor], rows: int, gen: torch.Generator, noise_v: float = 0.0,
               noise_F: float = 0.0):
        """Training rows from state s (caller order); advances self.small by
        `factor` substeps from s.  Returns X, T, w (row weights).
        noise_v, noise_F > 0 perturb the particles' v and F (Gaussian, per
        entry) for the big step only, while the target comes from the clean
        state: the network then also learns to pull perturbed states back."""
        cfg = self.cfg
        if noise_v > 0 or noise_F > 0:
            g = torch.Generator(device=self.device).manual_seed(int(torch.randint(0, 2 ** 31, (1,), generator=gen)))
            vn = s["v"] + noise_v * torch.randn(s["v"].shape, generator=g, device=self.device)
            Fn = s["F"] + noise_F * torch.randn(s["F"].shape, generator=g, device=self.device)
            _upload(self.big, s, v=vn.contiguous(), F=Fn.contiguous())
        else:
            _upload(self.big, s)
        X, nodes, vbig, total = self.big.corrector_sample(max_rows=self.max_rows)
        if total > self.max_rows:
            raise RuntimeError(f"{total} nodes with mass exceed the sampling buffer ({self.max_rows})")
        _upload(self.small, s)
        self.small.run(cfg.factor)
        x8 = self.small.positions()
        u = (x8 - s["x"]) / cfg.dt_big
        _upload(self.tgt, s, v=u.contiguous(), F=self.eye, C=self.zero33, Jp=self.ones)
        self.tgt.p2g_grid_only()
        vstar = self.tgt.grid_velocity(nodes)
        T = (vstar - vbig) / cfg.v_ref
        # weight = node mass / m_ref: the engine scales the applied correction
        # by min(1, 4 m / m_ref), so light (surface) nodes matter less
        w = X[:, CENTER].float()
        if X.shape[0] > rows:
            idx = torch.randperm(X.shape[0], generator=gen, device="cpu")[:rows].to(self.device)
            X, T, w = X[idx], T[idx], w[idx]
        return X.contiguous(), T.contiguous(), w.contiguous()

    def rollouts(self, n_rollouts: int, samples_per_rollout: int, steps: int, rows: int, seed: int = 0,
                 log: Optional[Callable[[str], None]] = None, noise_v: float = 0.0, noise_F: float = 0.0):
        """Rows from n_rollouts seeded rollouts of `steps` plain substeps."""
        cfg = self.cfg
        rng = np.random.default_rng(seed)
        gen = torch.Generator().manual_seed(seed)
        Xs, Ts, ws = [], [], []
        t0 = time.perf_counter()
        for r in range(n_rollouts):
            s = scene_tensors(cfg, seed * 100003 + r + 1, self.device)
            _upload(self.small, s)
            slots = np.arange(0, steps - cfg.factor + 1, cfg.factor)
            times = np.sort(rng.choice(slots, size=min(samples_per_rollout, len(slots)), replace=False))
            t = 0
            cur = s
            for ts in times:
                if ts > t:
                    _upload(self.small, cur)
                    self.small.run(int(ts - t))
                    cur = _download(self.small, cur)
                    t = int(ts)
                X, T, w = self.sample(cur, rows, gen, noise_v, noise_F)
                Xs.append(X)
                Ts.append(T)
                ws.append(w)
                cur = _download(self.small, cur)  # the state factor substeps later
                t += cfg.factor
            if log and (r + 1) % max(1, n_rollouts // 10) == 0:
                log(f"rollout {r + 1}/{n_rollouts}: {sum(x.shape[0] for x in Xs)} rows, "
                    f"{time.perf_counter() - t0:.1f} s")
        return torch.cat(Xs), torch.cat(Ts), torch.cat(ws)

    def policy_rollouts(self, net: CorrectorMLP, n_rollouts: int, samples_per_rollout: int, n_big: int,
                        rows: int, seed: int = 0, log: Optional[Callable[[str], None]] = None):
        """Rows from states visited by corrected rollouts (the network's own
        state distribution), with the same plain-substep targets."""
        cfg = self.cfg
        rng = np.random.default_rng(seed + 7)
        gen = torch.Generator().manual_seed(seed + 7)
        corr = make_engine(cfg, self.volume, cfg.dt_big, cfg.gravity, self.device, self.n)
        corr.set_corrector(net.params, cfg.m_ref, cfg.v_ref, enable=True)
        Xs, Ts, ws = [], [], []
        vmax = 0.9 * (1.0 / cfg.res) / cfg.dt_big
        try:
            for r in range(n_rollouts):
                s = scene_tensors(cfg, seed * 100003 + 50000 + r, self.device)
                _upload(corr, s)
                times = np.sort(rng.choice(np.arange(1, n_big + 1), size=samples_per_rollout, replace=False))
                t = 0
                cur = s
                for tb in times:
                    try:
                        corr.run(int(tb - t), check_oob=True)
                    except RuntimeError:
                        break
                    t = int(tb)
                    cur = _download(corr, cur)
                    if not bool(torch.isfinite(cur["x"]).all()) or float(cur["v"].norm(dim=1).max()) >= vmax:
                        break
                    X, T, w = self.sample(cur, rows, gen)
                    Xs.append(X)
                    Ts.append(T)
                    ws.append(w)
                if log and (r + 1) % max(1, n_rollouts // 5) == 0:
                    log(f"policy rollout {r + 1}/{n_rollouts}: {sum(x.shape[0] for x in Xs)} rows")
        finally:
            corr.close()
        if not Xs:
            return None
        return torch.cat(Xs), torch.cat(Ts), torch.cat(ws)

    def close(self):
        for e in (self.small, self.big, self.tgt):
            e.close()


def train(net: CorrectorMLP, X, T, w, epochs: int = 10, batch: int = 32768, lr: float = 1e-3,
          val_frac: float = 0.1, seed: int = 0, log: Optional[Callable[[str], None]] = None) -> Dict:
    """Adam on mini-batches with a cosine learning-rate decay; returns the
    train/validation loss history (weighted MSE of the scaled correction)."""
    n = X.shape[0]
    nval = int(n * val_frac)
    gen = torch.Generator().manual_seed(seed)
    perm = torch.randperm(n, generator=gen).to(X.device)
    val, tr = perm[:nval], perm[nval:]
    Xv, Tv, wv = X[val].contiguous(), T[val].contiguous(), w[val].contiguous()
    steps_per_epoch = max(1, len(tr) // batch)
    total = epochs * steps_per_epoch
    hist = {"train": [], "val": []}
    step = 0
    for ep in range(epochs):
        order = tr[torch.randperm(len(tr), generator=gen).to(X.device)]
        acc = torch.zeros(1, dtype=torch.float64, device=X.device)
        for b in range(steps_per_epoch):
            idx = order[b * batch:(b + 1) * batch]
            cur_lr = lr * 0.5 * (1 + math.cos(math.pi * step / total))
            acc += net.train_step(X[idx].contiguous(), T[idx].contiguous(), w[idx].contiguous(), lr=cur_lr)
            step += 1
        trl = float(acc.item()) / steps_per_epoch
        vl = float(net.loss_grads(Xv, Tv, wv).item()) if nval else float("nan")
        hist["train"].append(trl)
        hist["val"].append(vl)
        if log:
            log(f"epoch {ep + 1}/{epochs}: train {trl:.4e} val {vl:.4e}")
    return hist


def baseline_loss(T, w) -> float:
    """Weighted MSE of predicting zero correction (the untrained reference)."""
    return float(((T ** 2).sum(1) * w).sum() / w.sum())


class RolloutDiverged(RuntimeError):
    """The plain substeps that judge or replace a window left the grid or
    produced non-finite positions: the window's start state (a corrected
    state kept without judging) is beyond what the fallback can recover."""

    def __init__(self, msg: str, big_steps_done: int):
        super().__init__(msg)
        self.big_steps_done = big_steps_done


class CorrectedRunner:
    """Corrected big steps, judged by the plain substep, which is also the
    fallback.

    The rollout advances in windows of ``window`` big steps (``window *
    factor`` plain substeps of simulated time).  A window is *judged* when it
    is the ``judge_every``-th since the last judged one, or while the runner
    is on probation after a fallback: the same window is then also run with
    plain substeps from the window's start state, and the corrected result is
    kept only if its relative errors against the plain one (|x - x_plain| /
    |x_plain|, and the same for v and F) are at most ``tol``; otherwise the
    plain result replaces it (a fallback) and the next window is judged too.
    Every window, judged or not, must also be healthy -- no stencil left the
    grid, positions finite and inside the unit domain, speeds below the big
    step's one-cell-per-step limit -- or it is rerun with plain substeps.  If
    the plain substeps themselves fail that check the rollout has diverged
    beyond recovery and ``run`` raises ``RolloutDiverged``.
    ``judge_every = 1`` judges every window (the
    error bound holds window by window, at the cost of the plain substeps);
    larger values trade the bound for throughput.  ``stats`` counts the big
    steps kept, the plain substeps run and the fallbacks."""

    def __init__(self, cfg: CorrectorConfig, net: Optional[CorrectorMLP], volume: float, device,
                 window: int = 4, judge_every: int = 1, tol: float = 1e-2, check_every: Optional[int] = None):
        self.cfg = cfg
        self.device = torch.device(device)
        self.net = net
        self.volume = volume
        self.window = int(check_every or window)
        self.judge_every = max(1, int(judge_every))
        self.tol = float(tol)
        self._engines = {}  # particle count -> (big, plain), kept across run() calls
        self.reset_stats()

    def reset_stats(self):
        self.stats = dict(windows=0, judged=0, fallbacks=0, unhealthy=0, big_steps_kept=0, plain_substeps=0,
                          worst_judged_error=0.0)

    def _engines_for(self, n: int):
        if n not in self._engines:
            cfg = self.cfg
            big = make_engine(cfg, self.volume, cfg.dt_big, cfg.gravity, self.device, n)
            plain = make_engine(cfg, self.volume, cfg.dt, cfg.gravity, self.device, n, sort_every=48)
            if self.net is not None:
                big.set_corrector(self.net.params, cfg.m_ref, cfg.v_ref, enable=True)
            self._engines[n] = (big, plain)
        return self._engines[n]

    def close(self):
        for big, plain in self._engines.values():
            big.close()
            plain.close()
        self._engines = {}

    @property
    def fallbacks(self):
        return self.stats["fallbacks"] + self.stats["unhealthy"]

    def _healthy(self, eng, cur, dt: float, tag: str = "big") -> bool:
        dx = 1.0 / self.cfg.res
        vmax = 0.9 * dx / dt
        x = cur["x"]
        if eng.oob_count():  # reading resets the engine's counter
            why = "left_grid"
        elif not bool(torch.isfinite(x).all()):
            why = "non_finite"
        elif float(x.min()) < 0.0 or float(x.max()) > 1.0:
            why = "outside_domain"
        elif not float(cur["v"].norm(dim=1).max()) < vmax:
            why = "too_fast"
        else:
            return True
        key = f"unhealthy_{tag}_{why}"
        self.stats[key] = self.stats.get(key, 0) + 1
        return False

    def run(self, s: Dict[str, torch.Tensor], n_big: int, check: bool = True) -> Dict[str, torch.Tensor]:
        """n_big big steps of simulated time from state s (caller order);
        the engines are created on the first call and reused."""
        cfg = self.cfg
        n = s["x"].shape[0]
        big, plain = self._engines_for(n)
        st = self.stats
        cur = s
        done, since_judged, probation = 0, 0, False
        if not check:
            _upload(big, cur)
            big.run(n_big, check_oob=False)
            st["big_steps_kept"] += n_big
            return _download(big, cur)
        while done < n_big:
            k = min(self.window, n_big - done)
            start = cur
            _upload(big, start)
            big.run(k, check_oob=False)
            cand = _download(big, start)
            st["windows"] += 1
            since_judged += 1
            ok = self._healthy(big, cand, cfg.dt_big)
            if not ok:
                st["unhealthy"] += 1
            judge = ok and (probation or since_judged >= self.judge_every)
            ref = None
            if judge or not ok:
                _upload(plain, start)
                plain.run(k * cfg.factor, check_oob=False)
                st["plain_substeps"] += k * cfg.factor
                ref = _download(plain, start)
                if not self._healthy(plain, ref, cfg.dt, "plain"):
                    raise RolloutDiverged(f"plain substeps from the state after big step {done} left the "
                                          f"domain or went non-finite", done)
            if judge:
                st["judged"] += 1
                since_judged = 0
                err = judged_error(cand, ref)
                st["worst_judged_error"] = max(st["worst_judged_error"], err if math.isfinite(err) else 1.0)
                ok = math.isfinite(err) and err <= self.tol
                if not ok:
                    st["fallbacks"] += 1
                probation = not ok
            if ok:
                cur = cand
                st["big_steps_kept"] += k
            else:
                cur = ref
                probation = True
            done += k
        return cur


def judged_error(cand: Dict[str, torch.Tensor], ref: Dict[str, torch.Tensor]) -> float:
    """The judge's error of a corrected window against the plain substeps
    from the same start: the largest relative error of positions,
    velocities and deformation gradients.  Positions alone let a window
    through whose F has drifted, and the plain substeps that continue from
    it later blow up."""
    return max(rel_position_error(cand[k], ref[k]) for k in ("x", "v", "F"))


def rel_position_error(x, ref) -> float:
    return float((x - ref).norm() / ref.norm())


def one_step_gain(cfg: CorrectorConfig, net: CorrectorMLP, device, seeds=(901, 902, 903), warm: int = 160) -> Dict:
    """From plain-rollout states at substep `warm` of held-out scenes: the
    displacement error of one big step, with and without the network,
    against `factor` plain substeps."""
    errs = {"without": [], "with": []}
    for seed in seeds:
        s = scene_tensors(cfg, seed, device)
        eng = make_engine(cfg, s["volume"], cfg.dt, cfg.gravity, device, s["x"].shape[0])
        _upload(eng, s)
        eng.run(warm)
        s = _download(eng, s)
        eng.run(cfg.factor)
        ref = eng.positions().clone()
        eng.close()
        disp = (ref - s["x"]).norm()
        for tag, nn in (("without", None), ("with", net)):
            runner = CorrectedRunner(cfg, nn, s["volume"], device)
            out = runner.run(s, 1, check=False)
            runner.close()
            errs[tag].append(float((out["x"] - ref).norm() / disp))
    return {k: float(np.mean(v)) for k, v in errs.items()}


def rollout_report(cfg: CorrectorConfig, net: Optional[CorrectorMLP], device, seed: int = 0,
                   checkpoints=(1, 2, 4, 8, 16, 32, 64, 125)) -> Dict:
    """Big-step rollout from the scene's t = 0 against the plain substep: the
    relative position error |x - x_plain| / |x_plain| at each checkpoint (in
    big steps), until a stencil leaves the grid."""
    s = scene_tensors(cfg, seed, device)
    n = s["x"].shape[0]
    plain = make_engine(cfg, s["volume"], cfg.dt, cfg.gravity, device, n)
    _upload(plain, s)
    ref, t = {}, 0
    for k in checkpoints:
        plain.run(k * cfg.factor - t)
        t = k * cfg.factor
        ref[k] = plain.positions().clone()
    plain.close()
    eng = make_engine(cfg, s["volume"], cfg.dt_big, cfg.gravity, device, n)
    if net is not None:
        eng.set_corrector(net.params, cfg.m_ref, cfg.v_ref, enable=True)
    _upload(eng, s)
    out, t = {}, 0
    try:
        for k in checkpoints:
            try:
                eng.run(k - t, check_oob=True)
            except RuntimeError:
                out[str(k)] = "left the grid"
                break
            t = k
            x = eng.positions()
            e = float((x - ref[k]).norm() / ref[k].norm())
            out[str(k)] = e if math.isfinite(e) else "non-finite"
    finally:
        eng.close()
    return out


def judged_rollout(cfg: CorrectorConfig, net: Optional[CorrectorMLP], device, n_substeps: int = 1000,
                   settings=((1, 1), (1, 4), (2, 1), (1, 1000000), (4, 1000000)), tol: float = 1e-2,
                   seed: int = 0) -> Dict:
    """The corrected form against the plain substep over ``n_substeps`` of
    simulated time from the scene's t = 0: for each (window, judge_every) the
    relative position error of the final state against the plain rollout's,
    the wall time (fallbacks and judging included) and the simulated-time
    speedup over the plain rollout."""
    s = scene_tensors(cfg, seed, device)
    n = s["x"].shape[0]
    n_big = n_substeps // cfg.factor
    plain = make_engine(cfg, s["volume"], cfg.dt, cfg.gravity, device, n, sort_every=48)
    _upload(plain, s)
    plain.run(n_substeps, check_oob=False)  # captures the graphs
    _upload(plain, s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain.run(n_substeps, check_oob=True)
    ref = plain.positions().clone()
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    plain.close()
    out = {"substeps": n_substeps, "big_steps": n_big, "tol": tol, "plain_seconds": t_plain, "settings": []}
    for window, every in settings:
        runner = CorrectedRunner(cfg, net, s["volume"], device, window=window, judge_every=every, tol=tol)
        try:
            runner.run(s, min(n_big, 2 * window))  # warm-up: engines, graphs
            runner.reset_stats()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fin = runner.run(s, n_big)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            err = rel_position_error(fin["x"], ref)
            out["settings"].append(dict(window=window, judge_every=every, seconds=t, speedup=t_plain / t,
                                        rel_position_error=err if math.isfinite(err) else "non-finite",
                                        within_tol=bool(math.isfinite(err) and err <= tol), **runner.stats))
        except RolloutDiverged as e:
            torch.cuda.synchronize()
            out["settings"].append(dict(window=window, judge_every=every, seconds=None, speedup=None,
                                        rel_position_error="diverged", diverged_after_big_steps=e.big_steps_done,
                                        within_tol=False, **runner.stats))
        finally:
            runner.close()
    return out


def main(argv=None):
    import argparse
    import json

    ap = argparse.ArgumentParser(description="train the grid corrector on solver rollouts")
    ap.add_argument("--rollouts", type=int, default=200)
    ap.add_argument("--samples", type=int, default=4, help="sampled big steps per rollout")
    ap.add_argument("--steps", type=int, default=1000, help="plain substeps per rollout")
    ap.add_argument("--rows", type=int, default=8192, help="nodes kept per sample")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=32768)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--noise-v", type=float, default=0.02)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--res", type=int, default=128)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default="corrector.npz")
    args = ap.parse_args(argv)
    dev = torch.device("cuda", 0)
    cfg = CorrectorConfig(n=args.n, res=args.res)
    log = lambda m: print(m, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    gen = SampleGenerator(cfg, dev)
    X, T, w = gen.rollouts(args.rollouts, args.samples, args.steps, args.rows, seed=args.seed, log=log,
                           noise_v=args.noise_v)
    gen.close()
    log(f"{X.shape[0]} rows from {args.rollouts} rollouts in {time.perf_counter() - t0:.1f} s; "
        f"zero-correction loss {baseline_loss(T, w):.4e}")
    net = CorrectorMLP(device=dev, seed=args.seed)
    t1 = time.perf_counter()
    hist = train(net, X, T, w, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed, log=log)
    torch.cuda.synchronize()
    meta = dict(rollouts=args.rollouts, samples=args.samples, steps=args.steps, rows=int(X.shape[0]),
                epochs=args.epochs, batch=args.batch, lr=args.lr, noise_v=args.noise_v, n=args.n, res=args.res,
                dt=cfg.dt, factor=cfg.factor, m_ref=cfg.m_ref, v_ref=cfg.v_ref, seed=args.seed,
                train_loss=hist["train"], val_loss=hist["val"], zero_loss=baseline_loss(T, w),
                train_seconds=time.perf_counter() - t1)
    net.save(args.out, meta)
    log(json.dumps({k: v for k, v in meta.items() if k not in ("train_loss",)}))


if __name__ == "__main__":
    main()
