"""2D P2G scatter (femflow/solvers/mpm/two_d/p2g.py:11-76) on the GPU."""
from __future__ import annotations

from .. import _phase


def p2g(inv_dx, hardening, mu_0, lambda_0, mass, dx, dt, volume, grid_velocity, grid_mass, x, v, F,
        C, Jp, model: str = "neo_hookean", *, quirks: str = "reference"):
    """Scalar material (mu_0, lambda_0, mass), positions ``x`` (N, 2).

    The reference has no bounds check in 2D; particles whose stencil would
    leave the grid are skipped here instead of indexing out of range.
    quirks="textbook" adds the lambda term of the stress to the diagonal only
    (the reference broadcasts it to every entry, utils.py:86)."""
    res = int(grid_velocity.shape[0]) - 1
    q = _phase.quirks_code(quirks)
    import torch

    dev = _phase.device_of(x, v, F)
    dt_ = _phase.compute_dtype(v, F, x)
    scal = lambda s: torch.tensor([float(s)], dtype=dt_, device=dev)  # noqa: E731
    sp = _phase.StagedParticles(x, v, F, C, Jp, scal(mass), scal(mu_0), scal(lambda_0),
                                mat_scalar=True)
    prm = _phase.make_params(2, res, inv_dx, dx, dt, volume, quirks=q, check_oob=False)
    _phase.run_p2g(sp, prm, _phase.p2g_kind(model), hardening, grid_velocity, grid_mass)
