"""FP64 NumPy restatement of FEMFlow's explicit MLS-MPM substep.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP path: it is
imported by ``tests/``, by ``__graft_entry__.smoke()`` and by the CPU-baseline
leg of ``bench.py`` -- never by the ``femflow_amd`` package, whose product path
runs only the HIP kernels.

Every function names the reference lines it restates (paths relative to the
FEMFlow source tree).  The reference itself is not importable here (its hot
path imports numba, which is absent), so the restatement is pinned by the
reference's own known-answer tests (``femflow/solvers/mpm/tests``), by closed
forms (partition of unity, linear-field reproduction, conservation, rigid
rotations) and by library routines (LAPACK SVD through numpy, scipy's polar
decomposition).  Parity status per function is listed in ``PARITY`` below.

Conventions
-----------
* ``quirks="reference"`` (default) reproduces the reference exactly, including
  the lambda-term broadcast to all d*d entries of P F^T
  (``utils.py:86,129``) and the eps=1e-10 of the 2D polar
  (``linear_algebra.py:110``).  ``quirks="textbook"`` uses lambda (J-1) J * I.
* ``svd_gauge="reference"`` rebuilds a plastic F as ``U diag(s) Vh^T`` exactly
  as ``three_d/g2p.py:57`` and ``two_d/g2p.py:327`` do (numpy's third SVD
  output is already Vh); ``svd_gauge="consistent"`` uses ``U diag(s) Vh``.
  The two differ by a right rotation of F and agree on F F^T, J and every
  mechanical quantity.
* Materials are per particle: ``kind`` in {ELASTIC, SNOW, WATER} with a
  hardening ``h`` and snow clamp ``theta_c``/``theta_s``.  The reference's
  ``model`` strings map onto them in the phase wrappers (``p2g_3d`` etc.).
"""
from __future__ import annotations

import numpy as np

ELASTIC, SNOW, WATER = 0, 1, 2
THETA_C_REF, THETA_S_REF = 2.5e-2, 7.5e-3  # three_d/g2p.py:50, two_d/g2p.py:323

# Which oracle functions are pinned by what (checked in tests/test_oracle.py).
PARITY = {
    "weights": "KAT test_utils.py:6-15; partition of unity; first/second moments",
    "lame": "KAT test_particle_to_grid.py:30-38 (hardening at Jp=1); snow e=exp(h(1-Jp)) at Jp=0.9 and "
            "1.05 against e^1 and e^-1/2 (tests/test_oracle_pins.py::test_pin_snow_hardening, which fails "
            "for a negated exponent or a dropped h)",
    "stress_2d": "KAT test_particle_to_grid.py:41-50 off-diagonals; pure dilation closed form",
    "stress_3d": "rotation -> zero stress; pure dilation closed form; scipy polar",
    "polar_3d": "scipy.linalg.polar (independent library routine)",
    "p2g_2d": "KAT ground_truth_grid.txt multiset; mass and momentum conservation; APIC affine "
              "reproduction (node velocity a + B x_node, tests/test_oracle_pins.py)",
    "p2g_3d": "mass/momentum conservation; APIC affine reproduction: particles on v = a + B x with C = B, "
              "F = I give node velocity a + B x_node to 1e-12 and the G2P returns v, C "
              "(tests/test_oracle_pins.py::test_pin_affine_scatter, which fails for a flipped dpos, a dropped "
              "dx, a negated or transposed affine term); C++ loop restatement (oracle/mpm_oracle.c)",
    "grid_op_2d": "per-node closed form on a brute-force grid",
    "grid_op_3d": "per-node closed form on a brute-force grid",
    "g2p_2d": "KAT test_grid_to_particle.py:31-103 (v, x); linear field reproduction",
    "g2p_3d": "linear field reproduction (v = a + B x, C = B); C++ restatement",
    "snow_plasticity": "F = s I: clamp to c I, Jp <- clip(Jp s^3 / (c^3 + 1e-10), 0.6, 20) in closed form "
                       "(tests/test_oracle_pins.py::test_pin_snow_plasticity, which fails for an inverted "
                       "ratio, swapped thetas, a wrong clip or a lost J_old); C restatement pinned the same way",
    "water": "P F^T = K (J-1) J I and F <- cbrt(J) I in closed form (tests/test_oracle_pins.py::test_pin_water, "
             "which fails for the all-entries broadcast, a dropped J, a lost cube root or kept shear)",
    "bin_keys": "brute-force integer arithmetic; truncation toward zero",
    "bin_sort": "numpy.argsort(kind='stable') (library routine)",
    "collide_planes": "closed-form node counts (axis-aligned textbook planes; a hand-counted "
                      "reference-mode plane); per-node loop restatement in plain Python floats",
}


# --------------------------------------------------------------------------
# numerics/fem.py:1-6
# --------------------------------------------------------------------------
def Ev_to_mu(E, nu):
    """Lame mu from Young's modulus and Poisson ratio (numerics/fem.py:1-2)."""
    return E / (2 * (1 + nu))


def Ev_to_lambda(E, nu):
    """Lame lambda (numerics/fem.py:5-6)."""
    return E * nu / ((1 + nu) * (1 - 2 * nu))


# --------------------------------------------------------------------------
# Quadratic B-spline (three_d/p2g.py:50-55, two_d/p2g.py:50-53)
# --------------------------------------------------------------------------
def base_and_fx(x, inv_dx):
    """base = trunc(x*inv_dx - 0.5) (astype(int64) truncates toward zero,
    three_d/p2g.py:50); fx = x*inv_dx - base (three_d/p2g.py:53)."""
    s = x * inv_dx
    base = (s - 0.5).astype(np.int64)
    fx = s - base
    return base, fx


def weights(fx):
    """w[i] = [0.5(1.5-fx)^2, 0.75-(fx-1)^2, 0.5(fx-0.5)^2] (three_d/p2g.py:55).

    Returns an array of shape (3,) + fx.shape."""
    return np.stack([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1) ** 2, 0.5 * (fx - 0.5) ** 2])


# --------------------------------------------------------------------------
# Hardening (utils.py:7-49)
# --------------------------------------------------------------------------
def lame(kind, h, mu0, lam0, Jp):
    """Per-particle (mu, lambda).

    ELASTIC: constant_hardening -> (mu0*h, lam0*h)          (utils.py:7-24)
    SNOW:    snow_hardening e=exp(h(1-Jp)) -> (mu0*e, lam0*e) (utils.py:27-49)
    WATER:   (0, lam0) -- lam0 holds the bulk modulus K (not in the reference).
    """
    kind = np.asarray(kind)
    e = np.where(kind == SNOW, np.exp(h * (1.0 - Jp)), h)
    mu = np.where(kind == WATER, 0.0, mu0 * e)
    lam = np.where(kind == WATER, lam0, lam0 * e)
    return mu, lam


# --------------------------------------------------------------------------
# Polar decompositions (numerics/linear_algebra.py:96-135)
# --------------------------------------------------------------------------
def polar_decomp_2d(m, eps=1e-10):
    """Closed-form 2x2 rotation (linear_algebra.py:96-116), batched over
    leading axes.  eps=1e-10 is the reference's regulariser (:110)."""
    x = m[..., 0, 0] + m[..., 1, 1]
    y = m[..., 1, 0] - m[..., 0, 1]
    denom = np.sqrt(x * x + y * y) + eps
    scale = 1.0 / denom
    c = x * scale
    s = y * scale
    r = np.empty(m.shape, dtype=np.float64)
    r[..., 0, 0] = c
    r[..., 0, 1] = -s
    r[..., 1, 0] = s
    r[..., 1, 1] = c
    return r


def polar_decomp_3d(m):
    """R = U Vh from LAPACK SVD, no det fix (linear_algebra.py:119-135)."""
    w, _s, vh = np.linalg.svd(m, full_matrices=False)
    return w @ vh


# --------------------------------------------------------------------------
# Fixed-corotated stress fused with APIC (utils.py:52-135)
# --------------------------------------------------------------------------
def affine_momentum(F, C, mass, mu, lam, kind, inv_dx, dt, volume, quirks="reference"):
    """stress + m*C with stress = -(dt*vol) * 4*inv_dx^2 * PF  (utils.py:77-92,
    120-135).  PF = 2 mu (F - R) F^T + lam (J-1) J [broadcast to all entries in
    reference mode, *I in textbook mode]; WATER uses lam (J-1) J * I."""
    dim = F.shape[-1]
    J = np.linalg.det(F)
    if dim == 2:
        r = polar_decomp_2d(F, eps=1e-10 if quirks == "reference" else 0.0)
    else:
        r = polar_decomp_3d(F)
    D_inv = 4 * inv_dx * inv_dx
    mu = np.asarray(mu, dtype=np.float64)
    lam = np.asarray(lam, dtype=np.float64)
    dev = (2 * mu)[..., None, None] * (F - r) @ np.swapaxes(F, -1, -2)
    vol_term = (lam * (J - 1) * J)[..., None, None]
    eye = np.eye(dim)
    kind = np.broadcast_to(np.asarray(kind), J.shape)
    if quirks == "reference":
        shape = np.where((kind == WATER)[..., None, None], eye, np.ones((dim, dim)))
    else:
        shape = np.broadcast_to(eye, F.shape)
    PF = dev + vol_term * shape
    stress = -(dt * volume) * (D_inv * PF)
    return stress + np.asarray(mass, dtype=np.float64)[..., None, None] * C


# --------------------------------------------------------------------------
# P2G (three_d/p2g.py:14-80, two_d/p2g.py:11-76)
# --------------------------------------------------------------------------
def _stencil(dim):
    if dim == 3:
        return np.array([(i, j, k) for i in range(3) for j in range(3) for k in range(3)])
    return np.array([(i, j) for i in range(3) for j in range(3)])


def oob_mask(base, res_plus_1):
    """utils.py:138-150 applied to base and base+2 (the largest stencil offset)."""
    return np.any((base < 0) | (base + 2 >= res_plus_1), axis=-1)


def p2g(x, v, F, C, Jp, mass, mu0, lam0, kind, h, inv_dx, dx, dt, volume,
        grid_v, grid_m, quirks="reference", check_oob=None):
    """Scatter momentum and mass to the (res+1)^d grid, in place.

    Accumulates in the reference's order: particle-major, then lexicographic
    (i, j[, k]) (three_d/p2g.py:49,67-80).  ``np.add.at`` is unbuffered and
    walks the index array in order, so each node receives its contributions in
    exactly the sequence the reference's loop adds them."""
    n, dim = x.shape
    if check_oob is None:
        check_oob = dim == 3
    base, fx = base_and_fx(x, inv_dx)
    if check_oob and n and oob_mask(base, grid_m.shape[0]).any():
        raise RuntimeError("particle outside the grid")  # three_d/p2g.py:51-52
    w = weights(fx)  # (3, n, dim)
    mu, lam = lame(kind, h, mu0, lam0, Jp.reshape(n))
    A = affine_momentum(F, C, mass, mu, lam, kind, inv_dx, dt, volume, quirks)
    mv = v * np.asarray(mass, dtype=np.float64)[:, None]
    offs = _stencil(dim)
    res1 = grid_m.shape[0]
    nn = len(offs)
    contrib_v = np.empty((n, nn, dim))
    contrib_m = np.empty((n, nn))
    flat = np.empty((n, nn), dtype=np.int64)
    mass_b = np.broadcast_to(np.asarray(mass, dtype=np.float64), (n,))
    for o, ijk in enumerate(offs):
        dpos = (ijk[None, :] - fx) * dx
        if dim == 3:
            weight = w[ijk[0], :, 0] * w[ijk[1], :, 1] * w[ijk[2], :, 2]
        else:
            weight = w[ijk[0], :, 0] * w[ijk[1], :, 1]
        adpos = np.einsum("nij,nj->ni", A, dpos)
        contrib_v[:, o] = weight[:, None] * (mv + adpos)
        contrib_m[:, o] = weight * mass_b
        node = base + ijk[None, :]
        if dim == 3:
            flat[:, o] = (node[:, 0] * res1 + node[:, 1]) * res1 + node[:, 2]
        else:
            flat[:, o] = node[:, 0] * res1 + node[:, 1]
    gv = grid_v.reshape(-1, dim)
    gm = grid_m.reshape(-1)
    np.add.at(gv, flat.reshape(-1), contrib_v.reshape(-1, dim))
    np.add.at(gm, flat.reshape(-1), contrib_m.reshape(-1))


# --------------------------------------------------------------------------
# Grid update (three_d/grid_op.py:5-47, two_d/grid_op.py:5-24)
# --------------------------------------------------------------------------
def grid_op_3d(res, dx, dt, gravity, grid_v, grid_m, boundary="reference", bound=3):
    """3D: m>0 -> v/=m, v_y += dt g, clip to +-0.9 dx/dt; then for every node
    and axis d zero v_d if I_d < 1 or I_d >= res-1 (grid_op.py:25-47).
    ``boundary="sticky"``: all components zero within ``bound`` nodes of a wall."""
    m = grid_m.reshape(grid_m.shape[:3])
    v_allowed = dx * 0.9 / dt
    pos = m > 0
    gv = grid_v
    gv[pos] = gv[pos] / m[pos][:, None]
    gv[pos, 1] += dt * gravity
    gv[pos] = np.clip(gv[pos], -v_allowed, v_allowed)
    idx = np.arange(res + 1)
    if boundary == "reference":
        wall = (idx < 1) | (idx >= res - 1)
        gv[wall, :, :, 0] = 0
        gv[:, wall, :, 1] = 0
        gv[:, :, wall, 2] = 0
    else:
        wall = (idx < bound) | (idx > res - bound)
        gv[wall] = 0
        gv[:, wall] = 0
        gv[:, :, wall] = 0


def plane_normal(normal, quirks="reference"):
    """three_d/grid_op.py:59-60: n + 1/|n| (the scalar added to every
    component; reference) or n/|n| (textbook)."""
    n = np.asarray(normal, dtype=np.float64)
    denom = np.sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2])
    return n + 1.0 / denom if quirks == "reference" else n / denom


def collide_planes(points, normals, res, dx, grid_v, quirks="reference"):
    """Sticky half-space colliders, three_d/grid_op.py:50-67: for each plane,
    every node I with dot(I dx - point, n) < 0 gets v = 0 (all components).
    The dot product is taken as ((o0 n0 + o1 n1) + o2 n2), each operation
    rounded once (the reference's np.dot may order its three terms
    differently, which can only matter for nodes on the plane itself)."""
    idx = np.arange(res + 1, dtype=np.float64) * dx
    for point, normal in zip(np.asarray(points, np.float64).reshape(-1, 3),
                             np.asarray(normals, np.float64).reshape(-1, 3)):
        n = plane_normal(normal, quirks)
        o0 = (idx - point[0])[:, None, None]
        o1 = (idx - point[1])[None, :, None]
        o2 = (idx - point[2])[None, None, :]
        d = (o0 * n[0] + o1 * n[1]) + o2 * n[2]
        grid_v[d < 0] = 0.0


def grid_op_2d(res, dt, gravity, grid_v, grid_m, boundary="reference", bound=3):
    """2D: only nodes with m>0: v/=m, v_y += dt g; sticky if x<0.05, x>0.95 or
    y>0.95; floor y<0.05 keeps v_y>=0 (two_d/grid_op.py:13-24).
    ``boundary="sticky"``: all components zero within ``bound`` nodes of a wall."""
    m = grid_m.reshape(grid_m.shape[:2])
    pos = m > 0
    gv = grid_v
    gv[pos] = gv[pos] / m[pos][:, None]
    gv[pos, 1] += dt * gravity
    i = np.arange(res + 1)[:, None] * np.ones((1, res + 1))
    j = np.ones((res + 1, 1)) * np.arange(res + 1)[None, :]
    if boundary == "reference":
        xs, ys = i / res, j / res
        sticky = pos & ((xs < 0.05) | (xs > 1 - 0.05) | (ys > 1 - 0.05))
        gv[sticky] = 0.0
        floor = pos & (ys < 0.05)
        gv[floor, 1] = np.maximum(0.0, gv[floor, 1])
    else:
        wall = (i < bound) | (i > res - bound) | (j < bound) | (j > res - bound)
        gv[pos & wall] = 0.0


# --------------------------------------------------------------------------
# G2P (three_d/g2p.py:9-59, two_d/g2p.py:5-47)
# --------------------------------------------------------------------------
def g2p(x, v, F, C, Jp, kind, theta_c, theta_s, inv_dx, dt, grid_v,
        svd_gauge="consistent", always_svd=None, check_oob=None):
    """Gather v and C, advect, update F, apply plasticity; all in place.

    The 27 (9) stencil terms are added in the reference's lexicographic order
    (three_d/g2p.py:29-46).  SNOW clamps singular values to [1-theta_c,
    1+theta_s] and updates Jp <- clip(Jp*J_old/(J_new+1e-10), 0.6, 20)
    (three_d/g2p.py:48-58).  In 2D the SVD, rebuild and Jp update run for every
    particle and only the clamp is SNOW-only (two_d/g2p.py:321-330).  WATER
    resets F to cbrt(J) I (not in the reference)."""
    n, dim = x.shape
    if check_oob is None:
        check_oob = dim == 3
    if always_svd is None:
        always_svd = dim == 2
    base, fx = base_and_fx(x, inv_dx)
    if check_oob and n and oob_mask(base, grid_v.shape[0]).any():
        raise RuntimeError("particle outside the grid")  # three_d/g2p.py:23-24
    w = weights(fx)
    res1 = grid_v.shape[0]
    gv = grid_v.reshape(-1, dim)
    vn = np.zeros((n, dim))
    Cn = np.zeros((n, dim, dim))
    for ijk in _stencil(dim):
        dpos = ijk[None, :] - fx
        node = base + ijk[None, :]
        if dim == 3:
            flat = (node[:, 0] * res1 + node[:, 1]) * res1 + node[:, 2]
            weight = w[ijk[0], :, 0] * w[ijk[1], :, 1] * w[ijk[2], :, 2]
        else:
            flat = node[:, 0] * res1 + node[:, 1]
            weight = w[ijk[0], :, 0] * w[ijk[1], :, 1]
        g = gv[flat]
        vn += weight[:, None] * g
        Cn += 4 * inv_dx * ((weight[:, None] * g)[:, :, None] * dpos[:, None, :])
    x += dt * vn
    v[...] = vn
    C[...] = Cn
    Fn = (np.eye(dim)[None] + dt * Cn) @ F
    kind = np.broadcast_to(np.asarray(kind), (n,))
    theta_c = np.broadcast_to(np.asarray(theta_c, dtype=np.float64), (n,))
    theta_s = np.broadcast_to(np.asarray(theta_s, dtype=np.float64), (n,))
    snow = kind == SNOW
    sel = np.ones(n, bool) if always_svd else snow
    if sel.any():
        Fs = Fn[sel]
        U, sig, Vh = np.linalg.svd(Fs)
        sig = np.where(snow[sel][:, None],
                       np.clip(sig, 1.0 - theta_c[sel][:, None], 1.0 + theta_s[sel][:, None]),
                       sig)
        old_J = np.linalg.det(Fs)
        if svd_gauge == "reference":
            Fr = U @ (sig[:, :, None] * np.swapaxes(Vh, -1, -2))
        else:
            Fr = U @ (sig[:, :, None] * Vh)
        det = np.linalg.det(Fr) + 1e-10
        Jsel = Jp.reshape(n)[sel]
        Jp.reshape(n)[sel] = np.clip(Jsel * old_J / det, 0.6, 20.0)
        Fn[sel] = Fr
    water = kind == WATER
    if water.any():
        Jw = np.linalg.det(Fn[water])
        Fn[water] = np.cbrt(Jw)[:, None, None] * np.eye(dim)[None]
    F[...] = Fn


# --------------------------------------------------------------------------
# Reference-signature wrappers (phase APIs and the step driver)
# --------------------------------------------------------------------------
def make_mls_mpm_coefficients(lenx, dim):
    """v=0, F=I, C=0, Jp=1 (mls_mpm.py:9-15)."""
    v = np.zeros((lenx, dim), dtype=np.float64)
    F = np.array([np.eye(dim, dtype=np.float64) for _ in range(lenx)]).reshape(lenx, dim, dim)
    C = np.zeros((lenx, dim, dim), dtype=np.float64)
    Jp = np.ones((lenx, 1), dtype=np.float64)
    return v, F, C, Jp


def _p2g_kind(model):
    # three_d/p2g.py:57-61: "neo_hookean" -> constant hardening, any other string -> snow law
    return ELASTIC if model == "neo_hookean" else SNOW


def _g2p_kind(model):
    # three_d/g2p.py:48: only "snow" clamps
    return SNOW if model == "snow" else ELASTIC


def p2g_3d(inv_dx, hardening, dx, dt, volume, grid_velocity, grid_mass, x, mass, mu0, lam0,
           v, F, C, Jp, model="neo_hookean", quirks="reference"):
    """three_d.p2g (three_d/p2g.py:14-80) with the particle list as arrays."""
    n = len(x)
    p2g(x, v, F, C, Jp, np.broadcast_to(mass, (n,)), np.broadcast_to(mu0, (n,)),
        np.broadcast_to(lam0, (n,)), _p2g_kind(model), hardening, inv_dx, dx, dt, volume,
        grid_velocity, grid_mass, quirks=quirks)


def g2p_3d(inv_dx, dt, grid_velocity, x, v, F, C, Jp, model="neo_hookean", svd_gauge="consistent"):
    """three_d.g2p (three_d/g2p.py:9-59)."""
    g2p(x, v, F, C, Jp, _g2p_kind(model), THETA_C_REF, THETA_S_REF, inv_dx, dt, grid_velocity,
        svd_gauge=svd_gauge)


def p2g_2d(inv_dx, hardening, mu_0, lambda_0, mass, dx, dt, volume, grid_velocity, grid_mass,
           x, v, F, C, Jp, model="neo_hookean", quirks="reference"):
    """two_d.p2g (two_d/p2g.py:11-76): scalar material, no bounds check."""
    n = len(x)
    p2g(x, v, F, C, Jp, np.full(n, mass, dtype=np.float64), np.full(n, mu_0, dtype=np.float64),
        np.full(n, lambda_0, dtype=np.float64), _p2g_kind(model), hardening, inv_dx, dx, dt,
        volume, grid_velocity, grid_mass, quirks=quirks, check_oob=False)


def g2p_2d(inv_dx, dt, grid_velocity, x, v, F, C, Jp, model="neo_hookean", svd_gauge="consistent"):
    """two_d.g2p (two_d/g2p.py:5-47)."""
    g2p(x, v, F, C, Jp, _g2p_kind(model), THETA_C_REF, THETA_S_REF, inv_dx, dt, grid_velocity,
        svd_gauge=svd_gauge, always_svd=True, check_oob=False)


def solve_mls_mpm_3d(res, inv_dx, hardening, dx, dt, volume, gravity, x, mass, mu0, lam0,
                     v, F, C, Jp, quirks="reference"):
    """mls_mpm.py:40-79: fresh zero grids, p2g('neo_hookean'), grid_op, g2p."""
    dres = res + 1
    grid_velocity = np.zeros((dres, dres, dres, 3))
    grid_mass = np.zeros((dres, dres, dres, 1))
    p2g_3d(inv_dx, hardening, dx, dt, volume, grid_velocity, grid_mass, x, mass, mu0, lam0,
           v, F, C, Jp, "neo_hookean", quirks)
    grid_op_3d(res, dx, dt, gravity, grid_velocity, grid_mass)
    g2p_3d(inv_dx, dt, grid_velocity, x, v, F, C, Jp, "neo_hookean")
    return grid_velocity, grid_mass


def substep(state, params, quirks="reference", svd_gauge="consistent", return_grid=False):
    """One multi-material substep on a state dict (any dimension).

    state: x, v, F, C, Jp, mass, mu0, lam0, kind (per particle) and per-particle
    h, theta_c, theta_s.  params: res, dx, inv_dx, dt, volume, gravity,
    boundary ("reference"|"sticky"), bound (sticky thickness in nodes),
    optional planes = (points, normals, quirks) applied after the 3D grid
    update (collide_planes)."""
    x = state["x"]
    n, dim = x.shape
    res = params["res"]
    dres = res + 1
    shape = (dres,) * dim
    gv = np.zeros(shape + (dim,))
    gm = np.zeros(shape + (1,))
    p2g(x, state["v"], state["F"], state["C"], state["Jp"], state["mass"], state["mu0"],
        state["lam0"], state["kind"], state["h"], params["inv_dx"], params["dx"], params["dt"],
        params["volume"], gv, gm, quirks=quirks, check_oob=(dim == 3))
    grid_mass = gm.copy() if return_grid else None
    grid_mom = gv.copy() if return_grid else None
    if dim == 3:
        grid_op_3d(res, params["dx"], params["dt"], params["gravity"], gv, gm,
                   boundary=params.get("boundary", "reference"), bound=params.get("bound", 3))
        if params.get("planes"):
            pts, nrm, pq = params["planes"]
            collide_planes(pts, nrm, res, params["dx"], gv, quirks=pq)
    else:
        grid_op_2d(res, params["dt"], params["gravity"], gv, gm,
                   boundary=params.get("boundary", "reference"), bound=params.get("bound", 3))
    g2p(x, state["v"], state["F"], state["C"], state["Jp"], state["kind"], state["theta_c"],
        state["theta_s"], params["inv_dx"], params["dt"], gv, svd_gauge=svd_gauge,
        always_svd=(dim == 2), check_oob=(dim == 3))
    if return_grid:
        return {"mom": grid_mom, "mass": grid_mass, "vel": gv}
    return None


# --------------------------------------------------------------------------
# Binning (no reference counterpart: the reference walks particles in list
# order, three_d/p2g.py:49).  Keys derive from the reference's base (Q1 trunc).
# --------------------------------------------------------------------------
def bin_keys(x32, inv_dx, res, layout="blocked", block=4):
    """Cell key of each particle from FP32 positions, computed exactly as the
    GPU does: s = fl32(x*inv_dx), t = fl32(s-0.5), base = trunc(t) clamped to
    [0, res].  ``layout="rowmajor"``: (bx*(res+1)+by)*(res+1)+bz, the grid's
    own [i,j,k] order.  ``layout="blocked"``: block id of (base // 4) in
    row-major block order times block^dim plus the row-major cell inside the
    block."""
    x32 = np.ascontiguousarray(x32, dtype=np.float32)
    s = x32 * np.float32(inv_dx)
    t = s - np.float32(0.5)
    base = np.clip(t.astype(np.int64), 0, res)
    dim = x32.shape[1]
    r1 = res + 1
    if layout == "rowmajor":
        key = base[:, 0]
        for d in range(1, dim):
            key = key * r1 + base[:, d]
        return key.astype(np.uint32)
    nb = (r1 + block - 1) // block
    blk = base // block
    cell = base % block
    bid = blk[:, 0]
    cid = cell[:, 0]
    for d in range(1, dim):
        bid = bid * nb + blk[:, d]
        cid = cid * block + cell[:, d]
    return (bid * block ** dim + cid).astype(np.uint32)


def bin_sort(keys):
    """Stable ordering of particles by key: the permutation the HIP radix sort
    must reproduce bit for bit."""
    return np.argsort(np.asarray(keys), kind="stable").astype(np.uint32)


# --------------------------------------------------------------------------
# Conservation sums (FP64; closed forms used by the drift report)
# --------------------------------------------------------------------------
def particle_sums(x, v, mass):
    """[sum m, sum m v_0.., sum m x_0..] in FP64."""
    m = np.broadcast_to(np.asarray(mass, dtype=np.float64), (len(x),))
    return np.concatenate([[m.sum()], (m[:, None] * v).sum(0), (m[:, None] * x).sum(0)])


# --------------------------------------------------------------------------
# Scene builders shared by tests and bench (seeded synthetic inputs)
# --------------------------------------------------------------------------
# Seeded synthetic scenes: the same generator the package uses for its inputs
# (input generation only; nothing in the checking logic comes from there).
from femflow_amd.scenes import (CUBE_CORNERS_C2, MATERIALS_C2, batched_scene,  # noqa: E402,F401
                                batched_scene_specs, scene_2d_c1, scene_3d_cubes)


def copy_state(state):
    return {k: np.array(v, copy=True) for k, v in state.items()}
