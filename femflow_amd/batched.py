"""Many independent scenes in one engine (one launch sequence per substep).

``build_batched`` assembles the batched configuration: ``n_scenes`` scenes with
log-uniform particle counts, per-scene grid resolutions, one of the three
materials each and a box of random size, concatenated into one particle set
with per-particle scene ids.  ``BatchedRun`` owns the engine, autotunes its
launch shape and runs whole-batch substeps under the captured hipGraph.
"""
from __future__ import annotations

from typing import Dict, List, Optional

import numpy as np
import torch

from .autotune import autotune
from .engine import Engine, MaterialSpec, SceneSpec
from .scenes import MATERIALS_C2, batched_scene, batched_scene_specs

MATERIAL_TABLE = [MaterialSpec(("elastic", "snow", "water")[k], h, tc, ts) for k, _, _, h, tc, ts in MATERIALS_C2]


def build_batched(n_scenes: int = 64, seed: int = 0, count_scale: float = 1.0, dt: float = 5e-5,
                  keep_states: bool = False, n_lo: int = 4096, n_hi: int = 262144,
                  res_lo: int = 64, res_hi: int = 128) -> Dict:
    """Seeded batched workload.  ``count_scale`` shrinks every scene's particle
    count (tests); the default is the full configuration."""
    specs = batched_scene_specs(n_scenes=n_scenes, seed=seed, n_lo=n_lo, n_hi=n_hi, res_lo=res_lo,
                                res_hi=res_hi)
    scenes: List[SceneSpec] = []
    states, params = [], []
    for i, sp in enumerate(specs):
        sp = dict(sp, n=max(64, int(round(sp["n"] * count_scale))))
        st, pr = batched_scene(sp, seed=seed * 1000 + i + 1, dt=dt)
        scenes.append(SceneSpec(res=pr["res"], dt=pr["dt"], gravity=pr["gravity"], volume=pr["volume"]))
        states.append(st)
        params.append(pr)
    cat = {k: np.concatenate([st[k].reshape(len(st["x"]), -1) for st in states]) for k in
           ("x", "v", "F", "C", "Jp", "mass", "mu0", "lam0")}
    kind = np.concatenate([st["kind"] for st in states]).astype(np.int32)
    scene = np.concatenate([np.full(len(st["x"]), i, np.int32) for i, st in enumerate(states)])
    out = dict(scenes=scenes, cat=cat, kind=kind, scene=scene, counts=[len(st["x"]) for st in states],
               specs=specs)
    if keep_states:
        out["states"] = states
        out["params"] = params
    return out


class BatchedRun:
    """Engine over a batched workload, with device tensors and autotuning."""

    def __init__(self, work: Dict, device=None):
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        c = work["cat"]
        n = len(c["x"])
        f32 = lambda a, shape=None: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32,  # noqa: E731
                                                 device=dev).reshape(shape if shape else a.shape)
        self.t = dict(x=f32(c["x"]), v=f32(c["v"]), F=f32(c["F"], (n, 3, 3)), C=f32(c["C"], (n, 3, 3)),
                      Jp=f32(c["Jp"], (n,)), mass=f32(c["mass"], (n,)), mu0=f32(c["mu0"], (n,)),
                      lam0=f32(c["lam0"], (n,)))
        self.material = torch.tensor(work["kind"], dtype=torch.int32, device=dev)
        self.scene = torch.tensor(work["scene"], dtype=torch.int32, device=dev)
        self.engine = Engine(work["scenes"], MATERIAL_TABLE, n_max=n, device=dev)
        self.n = n
        self.upload()

    def upload(self):
        t = self.t
        self.engine.upload(t["x"], t["v"], t["F"], t["C"], t["Jp"], t["mass"], t["mu0"], t["lam0"],
                           self.material, self.scene)

    def autotune(self, **kw):
        t = self.t
        return autotune(self.engine, (t["mass"], t["mu0"], t["lam0"], self.material, self.scene), **kw)

    def run(self, nsteps: int, **kw):
        self.engine.run(nsteps, **kw)

    def download(self) -> Dict[str, torch.Tensor]:
        out = {k: torch.empty_like(self.t[k]) for k in ("x", "v", "F", "C", "Jp")}
        self.engine.download(out["x"], out["v"], out["F"], out["C"], out["Jp"])
        return out

    def close(self):
        self.engine.close()
