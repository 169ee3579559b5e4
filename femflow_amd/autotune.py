"""Autotuner for the engine's launch shapes.

For a loaded ``Engine`` it times a short substep sequence for every candidate
(particles per work item, persistent workgroups per CU, substeps between
rebinning, scatter spread) and keeps the fastest.  The engine state is saved before and
restored after the search, so tuning does not advance the simulation.
Results are cached in a JSON file keyed by the workload's signature (device
name, particle count, scene resolutions) so later runs skip the search.
"""
from __future__ import annotations

import hashlib
import itertools
import json
import os
from typing import Dict, Iterable, Optional, Tuple

import torch

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "femflow_amd", "autotune.json")

CHUNKS = (256, 512, 1024)
GRID_MULTS = (4, 8)
SORT_EVERY = (16, 32, 48)
SPREADS = (3, 9)


def signature(engine) -> str:
    dev = torch.cuda.get_device_name(engine.device)
    res = ",".join(str(s.res) for s in engine.scenes)
    raw = f"{dev}|n={engine.n}|scenes={len(engine.scenes)}|res={res}"
    return hashlib.sha1(raw.encode()).hexdigest()[:16]


def _snapshot(engine):
    n = engine.n
    dev = engine.device
    st = dict(x=torch.empty((n, 3), device=dev), v=torch.empty((n, 3), device=dev),
              F=torch.empty((n, 3, 3), device=dev), C=torch.empty((n, 3, 3), device=dev),
              Jp=torch.empty(n, device=dev))
    engine.download(st["x"], st["v"], st["F"], st["C"], st["Jp"])
    return st


def _time(engine, steps: int, reps: int) -> float:
    """Mean substep time over ``steps`` substeps that start with a rebin
    (``steps`` is a multiple of sort_every, so rebins are counted at their
    steady-state rate)."""
    for _ in range(2):  # warm-up: captures the graphs of both buffer parities
        engine.request_rebin()
        engine.run(steps, check_oob=False)
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        engine.request_rebin()
        torch.cuda.synchronize(engine.device)
        e0.record()
        engine.run(steps, check_oob=False)
        e1.record()
        torch.cuda.synchronize(engine.device)
        best = min(best, e0.elapsed_time(e1))
    return best / steps


def autotune(engine, restore: Tuple, steps: int = 32, reps: int = 2,
             candidates: Optional[Iterable[Tuple[int, ...]]] = None,
             cache_path: Optional[str] = DEFAULT_CACHE, verbose: bool = False) -> Dict:
    """Pick (chunk, grid_mult, sort_every, spread) for ``engine``; a candidate
    given as (chunk, grid_mult, sort_every) takes spread 9.

    ``restore`` = (mass, mu0, lam0, material, scene) tensors that were passed to
    ``upload`` (the engine keeps only its binned copy), used to put the state
    back after timing.  Returns {"config": {...}, "ms_per_substep": t, "table": [...]}."""
    key = signature(engine)
    if cache_path and os.path.exists(cache_path):
        try:
            with open(cache_path) as f:
                cached = json.load(f).get(key)
        except (OSError, ValueError):
            cached = None
        if cached:
            engine.tune(**cached["config"])
            return dict(cached, cached=True)
    snap = _snapshot(engine)
    mass, mu0, lam0, material, scene = restore
    cands = (list(candidates) if candidates is not None
             else list(itertools.product(CHUNKS, GRID_MULTS, SORT_EVERY, SPREADS)))
    cands = [tuple(c) + (9,) * (4 - len(c)) for c in cands]
    table = []
    best = None
    for chunk, gm, se, sp in cands:
        engine.tune(chunk=chunk, grid_mult=gm, sort_every=se, spread=sp)
        t = _time(engine, se * max(1, -(-steps // se)), reps)
        table.append({"chunk": chunk, "grid_mult": gm, "sort_every": se, "spread": sp, "ms_per_substep": t})
        if verbose:
            print("autotune", chunk, gm, se, sp, f"{t:.4f} ms", flush=True)
        if best is None or t < best[0]:
            best = (t, dict(chunk=chunk, grid_mult=gm, sort_every=se, spread=sp))
        engine.upload(snap["x"], snap["v"], snap["F"], snap["C"], snap["Jp"], mass, mu0, lam0, material, scene)
    engine.tune(**best[1])
    result = {"config": best[1], "ms_per_substep": best[0], "table": table}
    if cache_path:
        try:
            os.makedirs(os.path.dirname(cache_path), exist_ok=True)
            data = {}
            if os.path.exists(cache_path):
                with open(cache_path) as f:
                    data = json.load(f)
            data[key] = {"config": best[1], "ms_per_substep": best[0]}
            with open(cache_path, "w") as f:
                json.dump(data, f, indent=1)
        except (OSError, ValueError):
            pass
    return dict(result, cached=False)
