"""Benchmark: explicit MLS-MPM substeps on one MI355X.

Workload (the headline configuration): 2^20 particles in three seeded cubes of
side 0.2 (elastic fixed-corotated E=5e3 nu=0.2; snow E=5e3 nu=0.2 h=10,
theta_c=2.5e-2, theta_s=7.5e-3; weakly compressible water K=1e4), one
N(0, 0.5^2) velocity per cube, 128^3 grid on [0,1]^3, dt=5e-5, gravity -9.8,
FP32, the reference's 3D walls.  A step is one substep of the whole scene
(P2G, grid update, G2P with the constitutive and F update), with the state
resident in HBM.

Prints one JSON line.  `value` = particle-substeps/s over the timed steps,
which are substeps 0 .. steps of the configuration: the warm-up captures
every graph the window replays, then the initial state is uploaded again
(the synthetic particles in cell order, the order the engine keeps).
The window starts with a rebin (the engine rebins every `sort_every`
substeps), so it holds ceil(steps / sort_every) rebins and no capture;
`steady_state` times whole rebin periods from t = 0 for comparison.
`roofline` is for the dominant kernel (the fused G2P+P2G transfer): its
algorithmic bytes (212 B per particle-substep plus 28 B per grid node with
mass) over its mean duration on the same substeps, replayed with each
kernel timed by HIP events that take the kernel's own start and end on the
engine's stream.  `cpu_baseline` times the C restatement of the reference
step (FP64, one thread, the reference's loop order) on a bounded sample of
the same scene.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MPM particle-substeps/s (1M p, 128³ grid), %HBM peak; corrector speedup"
HBM_PEAK_GBS = 8000.0
BYTES_PER_PARTICLE = 212.0
BYTES_PER_NODE_TRANSFER = 28.0  # G2P reads 12 B of velocity + P2G writes 16 B per active node


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--res", type=int, default=128)
    ap.add_argument("--chunk", type=int, default=512)
    ap.add_argument("--sort-every", type=int, default=48)
    ap.add_argument("--grid-mult", type=int, default=8)
    ap.add_argument("--spread", type=int, default=9, choices=(3, 9),
                    help="scatter start offsets per cell (see Engine.tune)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget of the CPU-baseline sample (0 skips it)")
    ap.add_argument("--batched-steps", type=int, default=500,
                    help="substeps of the 64-scene batched workload (0 skips it)")
    ap.add_argument("--adjoint", type=int, default=1,
                    help="run the differentiable-rollout leg (0 skips it)")
    ap.add_argument("--corrector", type=int, default=1,
                    help="run the learned-corrector leg (0 skips it)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="first run 5 plain device copies of this many MiB (calibrates HBM byte counters)")
    ap.add_argument("--steady-periods", type=int, default=10,
                    help="rebin periods timed for the steady-state figure (0 skips it)")
    ap.add_argument("--config1", type=int, default=1,
                    help="time the 2D config (8,192 particles, 128^2, 200 FP64 substeps) on GPU and CPU")
    return ap.parse_args()


def host_threads():
    """Threads this process may run on (affinity), capped by OMP_NUM_THREADS."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(avail, int(env))) if env.isdigit() else max(1, avail)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(state, params, budget_s, nthreads=1):
    """The C restatement of the reference step on the host, on `nthreads`
    threads (OpenMP over particles, atomic grid adds, for more than one)."""
    from oracle import c_oracle

    cs = c_oracle.CState(state)
    t0 = time.perf_counter()
    c_oracle.run(state, params, 1, nthreads=nthreads, cstate=cs)
    t1 = time.perf_counter() - t0
    steps = 1
    extra = int(max(0, min(29, budget_s / max(t1, 1e-6) - 1)))
    if extra:
        t0 = time.perf_counter()
        c_oracle.run(state, params, extra, nthreads=nthreads, cstate=cs)
        t1 += time.perf_counter() - t0
        steps += extra
    n = len(state["x"])
    how = "one thread" if nthreads == 1 else f"{nthreads} OpenMP threads"
    return {"value": n * steps / t1, "unit": "particle-substeps/s", "cores": nthreads, "kind": "port",
            "sample": f"C restatement of the reference step (FP64, {how}, reference loop order), "
                      f"{steps} substep(s) of the full {n}-particle scene from t=0",
            "cpu_model": cpu_model(), "nproc": os.cpu_count()}


def cpu_drift_baseline(dev, nthreads, n=1 << 17, res=128, steps=1000, every=100):
    """The C restatement of the reference step on `nthreads` host threads,
    timed over the configuration's whole rollout (1000 substeps) of the
    three-material scene at 128^3 with n particles, with the FP32 engine's
    state compared against it every `every` substeps (relative inf-norm
    errors; v floored at 1e-6 m/s): the drift over the rollout."""
    import numpy as np
    import torch

    from femflow_amd.engine import Engine, MaterialSpec, SceneSpec
    from femflow_amd.scenes import MATERIALS_C2, scene_3d_cubes
    from oracle import c_oracle

    st, pr = scene_3d_cubes(n=n, res=res)
    n = len(st["x"])
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=dev)  # noqa: E731
    ts = {k: t(st[k]) for k in ("x", "v", "F", "C", "mass", "mu0", "lam0")}
    ts["Jp"] = t(st["Jp"].reshape(n))
    ref = dict(st)
    for k in ("x", "v", "F", "C", "Jp", "mass", "mu0", "lam0"):
        ref[k] = ts[k].double().cpu().numpy().reshape(np.shape(st[k]))
    cs = c_oracle.CState(ref)
    mats = [MaterialSpec(("elastic", "snow", "water")[k], h, tc, tsn) for k, _, _, h, tc, tsn in MATERIALS_C2]
    eng = Engine([SceneSpec(res=pr["res"], dt=pr["dt"], gravity=pr["gravity"], volume=pr["volume"])], mats,
                 n_max=n, device=dev)
    eng.tune(chunk=512, sort_every=48, grid_mult=8)
    eng.upload(ts["x"], ts["v"], ts["F"], ts["C"], ts["Jp"], ts["mass"], ts["mu0"], ts["lam0"],
               torch.tensor(st["kind"], dtype=torch.int32, device=dev))
    out = {k: torch.empty_like(ts[k]) for k in ("x", "v", "F", "C", "Jp")}
    cpu_s, table = 0.0, {}
    for k in range(1, steps // every + 1):
        eng.run(every, check_oob=True)
        eng.download(out["x"], out["v"], out["F"], out["C"], out["Jp"])
        t0 = time.perf_counter()
        c_oracle.run(ref, pr, every, nthreads=nthreads, cstate=cs)
        cpu_s += time.perf_counter() - t0
        torch.cuda.synchronize()
        errs = {}
        for f, floor in (("x", 1e-12), ("v", 1e-6), ("F", 1e-12)):
            a = out[f].double().cpu().numpy().reshape(n, -1)
            b = cs.arrays[f].reshape(n, -1)
            errs[f] = float(np.abs(a - b).max() / max(np.abs(b).max(), floor))
        table[str(k * every)] = errs
    eng.close()
    print(f"[bench] CPU restatement on {nthreads} threads: {steps} substeps of {n} particles in {cpu_s:.1f} s; "
          f"drift after {steps}: {table[str(steps)]}", file=sys.stderr, flush=True)
    how = "one thread" if nthreads == 1 else f"{nthreads} OpenMP threads (node slabs, deterministic)"
    return {"value": n * steps / cpu_s, "unit": "particle-substeps/s", "cores": nthreads, "kind": "port",
            "sample": f"C restatement of the reference step (FP64, {how}), the whole {steps}-substep rollout of "
                      f"the three-material scene at {res}^3 with {n} particles",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "drift_fp32_engine_vs_this": table}


def batched_bench(steps, dev):
    """64 scenes (log-uniform 4k-256k particles, 64^3-128^3 grids, random
    material and box) in one launch sequence per substep, autotuned; the
    timed substeps start from the configuration's initial state."""
    import torch

    from femflow_amd.batched import BatchedRun, build_batched

    work = build_batched(n_scenes=64, seed=0)
    run = BatchedRun(work, device=dev)
    tuned = run.autotune(steps=32, reps=2, cache_path=None)
    run.upload()
    for _ in range(2):  # captures the timed shape's graphs (both buffer parities)
        run.engine.request_rebin()
        run.run(steps, check_oob=False)
    # the timed substeps are substeps 0 .. steps of the configuration (the
    # upload makes the first of them rebin)
    run.upload()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    run.run(steps, check_oob=False)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    oob = run.engine.oob_count()
    run.close()
    print(f"[bench] batched: {run.n} particles, {steps} substeps in {ms:.1f} ms", file=sys.stderr, flush=True)
    return {"workload": "64 independent 3D scenes in one launch per substep: particle counts log-uniform in "
                        "[4k, 256k], grids 64^3..128^3, one random material and box side in [0.1, 0.3] each, "
                        "dt=5e-5, hipGraph replay, autotuned launch shape",
            "particles": run.n, "steps": steps, "value": run.n * steps / (ms / 1e3),
            "unit": "particle-substeps/s", "ms_per_step": ms / steps, "autotuned": tuned["config"],
            "oob": oob}


def corrector_bench(dev):
    """Learned corrector (dt_big = 8 dt) on the headline scene with the
    shipped network (trained on 200 solver rollouts by the library's own
    BF16-MFMA kernels): one-step accuracy gain, rollout error against the
    plain substep, and wall time per simulated time, corrected vs plain."""
    import torch

    from femflow_amd import corrector_train as CT
    from femflow_amd.corrector import CorrectorMLP

    path = os.path.join(ROOT, "femflow_amd", "data", "corrector_c5.npz")
    net = CorrectorMLP.load(path, device=dev)
    cfg = CT.CorrectorConfig()
    gain = CT.one_step_gain(cfg, net, dev)
    roll_with = CT.rollout_report(cfg, net, dev)
    roll_without = CT.rollout_report(cfg, None, dev)
    # wall time for the same simulated time (16 dt = 2 big steps, inside the
    # horizon where the big-step rollout is still accurate), from the t = 0
    # state, 5 repetitions each
    s = CT.scene_tensors(cfg, 0, dev)
    n = s["x"].shape[0]
    times = {}
    for tag, dt, steps, use in (("plain", cfg.dt, 2 * cfg.factor, False), ("corrected", cfg.dt_big, 2, True)):
        eng = CT.make_engine(cfg, s["volume"], dt, cfg.gravity, dev, n)
        if use:
            eng.set_corrector(net.params, cfg.m_ref, cfg.v_ref, enable=True)
        CT._upload(eng, s)
        eng.run(steps, check_oob=False)  # captures the graph
        tot = 0.0
        for _ in range(5):
            CT._upload(eng, s)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            eng.run(steps, check_oob=False)
            e1.record()
            torch.cuda.synchronize()
            tot += e0.elapsed_time(e1)
        times[tag] = tot / 5
        eng.close()
    meta = net.meta
    print(f"[bench] corrector: one-step {gain}, plain {times['plain']:.3f} ms vs corrected {times['corrected']:.3f} ms "
          f"per {16 * cfg.dt * 1e3:.2f} ms simulated", file=sys.stderr, flush=True)
    # the judged form over the configuration's 1000 substeps: the plain
    # substep judges windows of big steps and replaces any window whose
    # position error exceeds 1% (fallback and judging cost included)
    judged = CT.judged_rollout(cfg, net, dev, n_substeps=1000)
    met = any(st["within_tol"] and st["speedup"] is not None and st["speedup"] >= 4.0
              for st in judged["settings"])
    print(f"[bench] corrector judged 1000 substeps: " + "; ".join(
        f"window {st['window']} judge every {st['judge_every']}: speedup "
        f"{'-' if st['speedup'] is None else format(st['speedup'], '.2f')} err "
        f"{st['rel_position_error']} fallbacks {st['fallbacks']}" for st in judged["settings"]),
        file=sys.stderr, flush=True)
    return {"workload": "headline scene, dt_big = 8 dt = 4e-4, network 351-128-128-3 trained on "
                        f"{meta.get('rollouts')} solver rollouts ({meta.get('rows')} node samples)",
            "one_step_displacement_error": gain,
            "rollout_rel_position_error_by_big_step": {"with_network": roll_with, "without_network": roll_without},
            "ms_per_simulated_0.8ms": {"plain": times["plain"], "corrected": times["corrected"]},
            "speedup_simulated_time_2_big_steps": times["plain"] / times["corrected"],
            "judged_1000_substeps": judged,
            "target_met": met,
            "note": "one corrected big step is ~4x closer to 8 plain substeps than the uncorrected big step, but "
                    "at 8x the explicit step the big-step rollouts go unstable within 4-8 big steps, so the "
                    "judged settings fall back to plain substeps (or, unjudged, diverge); target_met says "
                    "whether any setting kept 1000 substeps within 1% at >= 4x"}


def config1_bench(dev, cpu=True):
    """The 2D configuration in full: 8,192 particles, 128^2 grid, 200 FP64
    substeps through the reference-signature phase functions on device tensors
    (grids zeroed each substep as the reference does), and the C restatement
    of the reference step on one host thread."""
    import numpy as np
    import torch

    from femflow_amd.scenes import scene_2d_c1
    from femflow_amd.solvers.mpm import two_d

    st, pr = scene_2d_c1()
    n, steps = len(st["x"]), 200
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)  # noqa: E731
    r1 = pr["res"] + 1
    mu0, lam0, mass = float(st["mu0"][0]), float(st["lam0"][0]), float(st["mass"][0])

    def roll(k):
        x, v, F, C, Jp = t(st["x"]), t(st["v"]), t(st["F"]), t(st["C"]), t(st["Jp"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            gv = torch.zeros((r1, r1, 2), dtype=torch.float64, device=dev)
            gm = torch.zeros((r1, r1), dtype=torch.float64, device=dev)
            two_d.p2g(pr["inv_dx"], 1.0, mu0, lam0, mass, pr["dx"], pr["dt"], pr["volume"], gv, gm,
                      x, v, F, C, Jp)
            two_d.grid_op(pr["res"], pr["dt"], pr["gravity"], gv, gm)
            two_d.g2p(pr["inv_dx"], pr["dt"], gv, x, v, F, C, Jp)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, x

    roll(5)  # warm-up
    gpu_s, x_gpu = roll(steps)
    out = {"workload": "2D MLS-MPM, 8,192 particles in [0.4,0.6]x[0.6,0.8] (seed 0), 128^2 grid, neo-Hookean "
                       "E=1e3 nu=0.2, dt=1e-4, gravity -9.8, the reference's 2D walls, FP64, 200 substeps",
           "gpu": {"seconds": gpu_s, "value": n * steps / gpu_s, "unit": "particle-substeps/s",
                   "path": "two_d.p2g / grid_op / g2p on FP64 device tensors, one call per phase"}}
    if cpu:
        from oracle import c_oracle

        cs = c_oracle.CState(st)
        t0 = time.perf_counter()
        c_oracle.run(st, pr, steps, nthreads=1, cstate=cs)
        cpu_s = time.perf_counter() - t0
        dx_max = float(np.abs(x_gpu.cpu().numpy() - cs.arrays["x"]).max())
        out["cpu_baseline"] = {"seconds": cpu_s, "value": n * steps / cpu_s, "unit": "particle-substeps/s",
                               "cores": 1, "kind": "port",
                               "sample": "C restatement of the reference step, FP64, one thread, all 200 substeps",
                               "cpu_model": cpu_model()}
        out["max_abs_position_difference_gpu_vs_cpu"] = dx_max
    print(f"[bench] config 1: GPU {gpu_s * 1e3:.1f} ms for {steps} substeps", file=sys.stderr, flush=True)
    return out


def binning_bench(x, res, reps=7):
    """The standalone binning entry point (mpm_bin: cell keys and their stable
    LSD radix sort) on the headline positions; median of `reps` calls, each
    including the entry point's own workspace allocation."""
    import torch

    from femflow_amd.engine import bin_particles

    bin_particles(x, res, float(res))
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bin_particles(x, res, float(res))
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    ms = sorted(times)[len(times) // 2] * 1e3
    n = int(x.shape[0])
    return {"workload": f"mpm_bin on the {n} headline positions, {res}^3 grid, blocked keys, stable sort; "
                        "wall time per call including its workspace allocation",
            "ms": ms, "value": n / (ms / 1e3), "unit": "particles/s"}


def adjoint_bench(dev):
    """Gradient of a 500-substep rollout (2^17 particles, 64^3 grid, E per
    particle) through the hand-written adjoint, checkpoints every 50."""
    import numpy as np
    import torch

    from femflow_amd.adjoint import RolloutConfig, lame_from_E, rollout

    n, res, dt, steps = 1 << 17, 64, 1e-4, 500
    rng = np.random.default_rng(0)
    x0 = 0.4 + 0.2 * rng.random((n, 3))
    v0 = np.tile(0.5 * rng.standard_normal(3), (n, 1))
    v0[:, 1] -= 3.0 * (x0[:, 1] - 0.5)
    E = 1e3 + 9e3 * rng.random(n)
    sub = torch.as_tensor(np.nonzero(x0[:, 1] > 0.5)[0], device=dev)
    target = torch.tensor(x0[x0[:, 1] > 0.5].mean(0) + 0.05 * rng.standard_normal(3), dtype=torch.float32,
                          device=dev)
    vol = 0.2 ** 3 / n
    cfg = RolloutConfig(res=res, dt=dt, volume=vol, nsteps=steps, checkpoint_every=50)
    f32 = lambda a: torch.tensor(np.asarray(a), dtype=torch.float32, device=dev)  # noqa: E731
    mass = f32(np.full(n, vol))
    v0t, Et = f32(v0).requires_grad_(True), f32(E).requires_grad_(True)

    def once():
        mu0, lam0 = lame_from_E(Et, 0.2)
        xT, _ = rollout(f32(x0), v0t, mu0, lam0, mass, cfg)
        L = ((xT[sub].mean(0) - target) ** 2).sum()
        return torch.autograd.grad(L, (v0t, Et))

    once()  # warm-up
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    gv, gE = once()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    print(f"[bench] adjoint: forward+backward of {steps} substeps in {ms:.1f} ms", file=sys.stderr, flush=True)
    return {"workload": "2^17 particles in one elastic cube, E ~ U[1e3, 1e4] per particle, 64^3 grid, dt = 1e-4, "
                        "500 substeps, checkpoints every 50; loss |centroid of the upper half at T - target|^2; "
                        "gradient w.r.t. v0 and per-particle E",
            "ms_forward_backward": ms, "value": n * steps / (ms / 1e3),
            "unit": "particle-substeps/s (forward + backward)", "dtype": "f32",
            "grad_norms": {"v0": float(gv.norm()), "E": float(gE.norm())}}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "this build has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from femflow_amd.engine import Engine, MaterialSpec, SceneSpec
    from femflow_amd.scenes import MATERIALS_C2, scene_3d_cubes

    dev = torch.device("cuda", 0)
    if args.copy_mb > 0:
        src = torch.ones(args.copy_mb << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] calibration copies: 5 x {args.copy_mb} MiB read + {args.copy_mb} MiB written",
              file=sys.stderr, flush=True)
        del src, dst
    st, pr = scene_3d_cubes(n=args.n, res=args.res)
    n = len(st["x"])
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=dev)  # noqa: E731
    x, v, F, C, Jp = t(st["x"]), t(st["v"]), t(st["F"]), t(st["C"]), t(st["Jp"].reshape(n))
    mass, mu0, lam0 = t(st["mass"]), t(st["mu0"]), t(st["lam0"])
    kind = torch.tensor(st["kind"], dtype=torch.int32, device=dev)
    x_caller = x
    # The synthetic particles in cell order (4x4x4 blocks, cells inside a
    # block): the input's order is arbitrary, and the engine keeps its state
    # in this order, so a restart from t = 0 does not time a one-off move of
    # the particles out of the generator's random order
    cell = torch.clamp(torch.floor(x.double() * pr["inv_dx"] - 0.5), 0, args.res).to(torch.int64)
    nb = (args.res + 4) // 4
    key = (((cell[:, 0] >> 2) * nb + (cell[:, 1] >> 2)) * nb + (cell[:, 2] >> 2)) * 64 + \
        ((cell[:, 0] & 3) << 4) + ((cell[:, 1] & 3) << 2) + (cell[:, 2] & 3)
    order = torch.argsort(key, stable=True)
    x, v, F, C, Jp, mass, mu0, lam0, kind = (a[order].contiguous()
                                            for a in (x, v, F, C, Jp, mass, mu0, lam0, kind))
    del cell, key, order
    mats = [MaterialSpec(("elastic", "snow", "water")[k], h, tc, ts) for k, _, _, h, tc, ts in MATERIALS_C2]
    eng = Engine([SceneSpec(res=pr["res"], dt=pr["dt"], gravity=pr["gravity"], volume=pr["volume"])], mats,
                 n_max=n, device=dev)
    eng.tune(chunk=args.chunk, sort_every=args.sort_every, grid_mult=args.grid_mult,
             use_graph=not args.no_graph, spread=args.spread)
    def restart():
        # the configuration's initial state (t = 0); graphs already captured
        # for this particle count are kept, and the next run starts with a
        # rebin
        eng.upload(x, v, F, C, Jp, mass, mu0, lam0, kind)

    restart()
    eng.run(args.warmup, check_oob=False)
    # the timed window's shape from a rebin, once per state-buffer parity, so
    # that its graphs are captured before the clock starts
    for _ in range(2):
        eng.request_rebin()
        eng.run(args.steps, check_oob=False)
    torch.cuda.synchronize()
    print(f"[bench] warmup {args.warmup} (+ graph capture of the timed shape) done", file=sys.stderr, flush=True)

    # ---- timed region: substeps 0 .. steps of the configuration ----------
    stream = torch.cuda.current_stream(dev)

    def timed(k):
        restart()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        ev0.record(stream)
        eng.run(k, check_oob=False)
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1), time.perf_counter() - w0

    ms, wall = timed(args.steps)
    oob = eng.oob_count()
    if oob:
        raise RuntimeError(f"{oob} particle stencils left the grid during the benchmark")
    value = n * args.steps / (ms / 1e3)
    print(f"[bench] timed {args.steps} substeps: {ms:.2f} ms (wall {wall * 1e3:.2f} ms)",
          file=sys.stderr, flush=True)
    steady = None
    if args.steady_periods > 0:
        # whole rebin periods from t = 0 (their graphs captured first)
        ks = args.steady_periods * args.sort_every
        for _ in range(2):
            eng.request_rebin()
            eng.run(ks, check_oob=False)
        sms, _ = timed(ks)
        steady = {"steps": ks, "rebins": args.steady_periods, "ms_per_step": sms / ks,
                  "value": n * ks / (sms / 1e3)}
        print(f"[bench] steady state: {ks} substeps in {sms:.2f} ms", file=sys.stderr, flush=True)

    # ---- per-kernel durations over the timed region's substeps -----------
    # The same substeps again from t = 0 (at least one rebin period), eager,
    # each kernel timed by HIP events that take its own start and end on the
    # engine's stream
    nprof = max(args.steps, args.sort_every)
    restart()
    eng.kernel_timing(True)
    eng.run(nprof, check_oob=False)
    kt = eng.kernel_times()
    eng.kernel_timing(False)
    grid_blocks = eng.active_block_count()  # blocks the last grid update swept
    # the same substeps with the FP64 per-substep sums on: sums[:, 7] counts
    # the grid nodes with mass (the active nodes), sums[:, 0] the grid mass
    sums = torch.zeros((nprof, 8), dtype=torch.float64, device=dev)
    restart()
    eng.run(nprof, sums=sums, check_oob=False)
    g_ms, g_cnt = kt["g2p2g"]
    dom = "g2p2g" if g_cnt else "p2g"
    d_ms, d_cnt = kt[dom]
    avg_s = d_ms / 1e3 / max(d_cnt, 1)
    s = sums.cpu().numpy()
    active_nodes = float(s[:, 7].mean()) if nprof else 0.0
    alg_bytes = BYTES_PER_PARTICLE * n + BYTES_PER_NODE_TRANSFER * active_nodes
    achieved = alg_bytes / avg_s / 1e9
    mass_drift = float(np.abs(s[:, 0] / s[0, 0] - 1).max()) if nprof else 0.0
    # HBM bytes per launch of the fused kernel from the committed rocprofv3
    # PMC summary of this workload (FETCH_SIZE, WRITE_SIZE, calibrated by a
    # plain copy); counters cannot be read inside this run
    traffic = grid_traffic = None
    prof = os.path.join(ROOT, "profiles", "g2p2g_c2_1M_128_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            tj = json.load(f)
        traffic = tj.get("bytes_per_launch")
        grid_traffic = tj.get("grid_update_bytes_per_launch")

    eng.close()
    batched = batched_bench(args.batched_steps, dev) if args.batched_steps > 0 else None
    corrector = corrector_bench(dev) if args.corrector else None
    adjoint = adjoint_bench(dev) if args.adjoint else None
    binning = binning_bench(x_caller, args.res)
    config1 = config1_bench(dev, cpu=args.cpu_seconds > 0) if args.config1 else None

    cpu = cpu_omp = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(st, pr, args.cpu_seconds)
        nt = host_threads()
        if nt > 1:
            cpu_omp = cpu_drift_baseline(dev, nt)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "3D MLS-MPM substep, 2^20 particles in three seeded cubes (elastic fixed-corotated, "
                        "snow, weakly compressible water), 128^3 grid, dt=5e-5",
            "particles": n, "grid": f"{args.res + 1}^3 nodes", "chunk": args.chunk,
            "sort_every": args.sort_every, "spread": args.spread, "hip_graph": not args.no_graph,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_source": "profiles/g2p2g_c2_1M_128_traffic.json (rocprofv3 PMC run of this workload)",
            # the HBM rate the kernel actually sustains: measured bytes (the
            # fused pass neither reads nor writes v and C, which the next G2P
            # recomputes) over the same mean duration
            "traffic_gbs": (traffic / avg_s / 1e9) if traffic else None,
            "kernel": dom, "kernel_ms": avg_s * 1e3,
            "alg_bytes_per_launch": alg_bytes, "active_nodes": active_nodes,
            "grid_blocks_processed": grid_blocks,
            "grid_update_traffic": grid_traffic,
            "grid_update_alg_bytes": BYTES_PER_NODE_TRANSFER * 64 * grid_blocks,
        },
        "rebins_in_window": -(-args.steps // args.sort_every),
        "steady_state": steady,
        "step_hbm_frac": (BYTES_PER_PARTICLE * n / (ms / 1e3 / args.steps)) / 1e9 / HBM_PEAK_GBS,
        "kernel_ms_per_step": {k: (vv[0] / vv[1] if vv[1] else 0.0) for k, vv in kt.items()},
        "grid_mass_drift_fp64": mass_drift,
        "batched": batched,
        "corrector": corrector,
        "adjoint": adjoint,
        "binning": binning,
        "config1": config1,
        "cpu_baseline": cpu,
        "cpu_baseline_omp": cpu_omp,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
