"""Helpers shared by the GPU tests: oracle state dicts <-> CUDA tensors and
engine construction for the seeded scenes."""
import numpy as np
import torch

from femflow_amd.engine import Engine, MaterialSpec, SceneSpec
from oracle import mpm_oracle as O


def to_torch_state(st, device, dtype=torch.float32):
    n, d = st["x"].shape
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=device)  # noqa: E731
    return dict(x=t(st["x"]), v=t(st["v"]), F=t(st["F"].reshape(n, d, d)), C=t(st["C"].reshape(n, d, d)),
                Jp=t(st["Jp"].reshape(n)), mass=t(st["mass"]), mu0=t(st["mu0"]), lam0=t(st["lam0"]),
                kind=torch.tensor(np.ascontiguousarray(st["kind"]), dtype=torch.int32, device=device))


def to_numpy_state(ts):
    return {k: v.detach().double().cpu().numpy() for k, v in ts.items() if k in ("x", "v", "F", "C", "Jp")}


def oracle_copy_from(ts, st):
    """Oracle state (FP64) built from the FP32 values the GPU actually sees."""
    out = O.copy_state(st)
    n, d = st["x"].shape
    for k in ("x", "v", "F", "C", "Jp", "mass", "mu0", "lam0"):
        out[k] = ts[k].detach().double().cpu().numpy().reshape(st[k].shape)
    return out


def material_table_c2():
    """The three cube materials; ids equal the oracle's kinds (0 elastic, 1 snow, 2 water)."""
    tab = []
    for kind, E, nu, h, tc, ts in O.MATERIALS_C2:
        name = {O.ELASTIC: "elastic", O.SNOW: "snow", O.WATER: "water"}[kind]
        tab.append(MaterialSpec(name, h, tc, ts))
    return tab


def engine_for(st, pr, device, n_max=None, **tune):
    eng = Engine([SceneSpec(res=pr["res"], dt=pr["dt"], gravity=pr["gravity"], volume=pr["volume"],
                            boundary=pr.get("boundary", "reference"), bound=pr.get("bound", 3))],
                 material_table_c2(), n_max=n_max or len(st["x"]), device=device)
    if tune:
        eng.tune(**tune)
    return eng


def engine_roundtrip(eng, ts, nsteps, sums=None):
    eng.upload(ts["x"], ts["v"], ts["F"], ts["C"], ts["Jp"], ts["mass"], ts["mu0"], ts["lam0"],
               ts["kind"])
    eng.run(nsteps, sums=sums)
    out = {k: torch.empty_like(ts[k]) for k in ("x", "v", "F", "C", "Jp")}
    eng.download(out["x"], out["v"], out["F"], out["C"], out["Jp"])
    torch.cuda.synchronize()
    return out
