"""Closed-form pins for the oracle pieces that conservation sums and the
reference's known-answer tests leave open, each shown to catch its slip.

A pin is a function ``pin_*(impl)`` that raises AssertionError when ``impl``
(an oracle function, or a copy of it with one planted mistake) violates a
closed form of the reference's algorithm:

* P2G affine scatter, ``weight * (v*m + affine @ dpos)`` with
  ``dpos = (ijk - fx) * dx`` (three_d/p2g.py:67-77, two_d/p2g.py:65-76):
  particles on a linear velocity field v = a + B x carrying C = B, F = I
  (zero stress) must give node velocity a + B x_node at every node with mass
  (APIC reproduces affine fields exactly), and the G2P that follows must
  return v_p and C_p.  A flipped ``dpos`` or a dropped ``dx`` moves the node
  velocities while leaving mass and momentum sums unchanged.
* Snow hardening e = exp(h (1 - Jp)) (utils.py:27-49, exp at :48) at Jp != 1.
* The snow clamp and Jp <- clip(Jp * det(F_old) / (det(F_new) + 1e-10), 0.6, 20)
  (three_d/g2p.py:48-58) on F = s I, where every singular value is s.
* The weakly compressible water model (no reference counterpart; the
  configuration names it): P F^T = K (J - 1) J I and F <- cbrt(J) I.

The mutants are built on the CPU from the oracle's own source text with one
substitution; the real oracle must pass every pin and each mutant must fail
its pin.
"""
import inspect
import textwrap

import numpy as np
import pytest

from oracle import mpm_oracle as O


def mutant(func, old, new, **overrides):
    """``func`` recompiled from its source with ``old`` replaced by ``new``
    (exactly one occurrence), in a copy of its module namespace in which
    ``overrides`` replace other functions."""
    src = textwrap.dedent(inspect.getsource(func))
    assert src.count(old) == 1, (func.__name__, old)
    ns = dict(func.__globals__)
    ns.update(overrides)
    exec(compile(src.replace(old, new), f"<mutant of {func.__name__}>", "exec"), ns)
    return ns[func.__name__]


# ------------------------------------------------------ P2G affine scatter ----
def _affine_field_scene(dim, seed=3, n=300, res=16):
    rng = np.random.default_rng(seed)
    x = 0.3 + 0.4 * rng.random((n, dim))  # stencils well inside the walls
    a = rng.standard_normal(dim)
    B = rng.standard_normal((dim, dim))
    v = a + x @ B.T
    C = np.broadcast_to(B, (n, dim, dim)).copy()
    F = np.broadcast_to(np.eye(dim), (n, dim, dim)).copy()
    mass = 0.5 + rng.random(n)
    return x, v, C, F, mass, a, B, res


def pin_affine_scatter(p2g, g2p=O.g2p, dim=3):
    x, v, C, F, mass, a, B, res = _affine_field_scene(dim)
    n = len(x)
    inv_dx, dx, dt, vol = float(res), 1.0 / res, 1e-4, 1e-3
    r1 = res + 1
    gv = np.zeros((r1,) * dim + (dim,))
    gm = np.zeros((r1,) * dim + (1,) if dim == 3 else (r1,) * dim)
    Jp = np.ones((n, 1))
    kind = np.full(n, O.ELASTIC)
    # F = I: R = I and J = 1, so P F^T = 0 in both stress conventions and
    # the affine momentum is m C = m B
    p2g(x, v, F, C, Jp, mass, np.full(n, 300.0), np.full(n, 200.0), kind, np.ones(n), inv_dx, dx, dt, vol,
        gv, gm)
    m = gm.reshape((r1,) * dim)
    has = m > 0
    vel = np.zeros_like(gv)
    vel[has] = gv[has] / m[has][:, None]
    idx = np.stack(np.meshgrid(*[np.arange(r1)] * dim, indexing="ij"), axis=-1) * dx
    expect = a + idx @ B.T
    scale = np.abs(expect[has]).max()
    err = np.abs(vel[has] - expect[has]).max() / scale
    assert err < 1e-12, f"node velocity is not a + B x_node: {err:.3e}"
    # G2P on that grid returns the particle's own v and C
    xg, vg, Cg, Fg, Jg = x.copy(), np.zeros_like(v), np.zeros_like(C), F.copy(), Jp.copy()
    g2p(xg, vg, Fg, Cg, Jg, kind, O.THETA_C_REF, O.THETA_S_REF, inv_dx, 0.0, vel)
    ev = np.abs(vg - v).max() / np.abs(v).max()
    eC = np.abs(Cg - C).max() / np.abs(B).max()
    # C sums 4 inv_dx w g (ijk - fx) over nodes whose g carries a, so it
    # cancels ~1e4 of its magnitude: 1e-10 is that rounding, far below the
    # O(1) error of any slip
    assert ev < 1e-12 and eC < 1e-10, f"G2P does not return v, C: {ev:.3e} {eC:.3e}"


@pytest.mark.parametrize("dim", [3, 2])
def test_pin_affine_scatter(dim):
    pin_affine_scatter(O.p2g, dim=dim)


@pytest.mark.parametrize("dim", [3, 2])
@pytest.mark.parametrize("old,new", [
    ("dpos = (ijk[None, :] - fx) * dx", "dpos = (fx - ijk[None, :]) * dx"),   # flipped offset
    ("dpos = (ijk[None, :] - fx) * dx", "dpos = (ijk[None, :] - fx)"),        # dropped dx
    ("contrib_v[:, o] = weight[:, None] * (mv + adpos)", "contrib_v[:, o] = weight[:, None] * (mv - adpos)"),
    ('adpos = np.einsum("nij,nj->ni", A, dpos)', 'adpos = np.einsum("nji,nj->ni", A, dpos)'),  # A^T
])
def test_pin_affine_scatter_catches(dim, old, new):
    with pytest.raises(AssertionError):
        pin_affine_scatter(mutant(O.p2g, old, new), dim=dim)


def test_pin_affine_scatter_catches_g2p_slips():
    bad = mutant(O.g2p, "Cn += 4 * inv_dx * ((weight[:, None] * g)[:, :, None] * dpos[:, None, :])",
                 "Cn += 4 * inv_dx * ((weight[:, None] * g)[:, None, :] * dpos[:, :, None])")  # C^T
    with pytest.raises(AssertionError):
        pin_affine_scatter(O.p2g, g2p=bad)


def test_affine_scatter_conservation_blind_spot():
    """Why the pin above is needed: the flipped-offset mutant still conserves
    mass and momentum exactly (sum_i w_i (x_i - x_p) = 0)."""
    bad = mutant(O.p2g, "dpos = (ijk[None, :] - fx) * dx", "dpos = (fx - ijk[None, :]) * dx")
    x, v, C, F, mass, a, B, res = _affine_field_scene(3)
    n = len(x)
    out = []
    for f in (O.p2g, bad):
        gv, gm = np.zeros((res + 1,) * 3 + (3,)), np.zeros((res + 1,) * 3 + (1,))
        f(x, v, F, C, np.ones((n, 1)), mass, np.full(n, 300.0), np.full(n, 200.0), np.zeros(n, int), np.ones(n),
          float(res), 1.0 / res, 1e-4, 1e-3, gv, gm)
        out.append((gm.sum(), gv.reshape(-1, 3).sum(0)))
    assert np.isclose(out[0][0], out[1][0], rtol=1e-13)
    assert np.allclose(out[0][1], out[1][1], rtol=1e-11)


# ------------------------------------------------------- snow hardening ----
def pin_snow_hardening(lame):
    mu0, lam0 = 416.0 + 2.0 / 3.0, 277.0 + 7.0 / 9.0
    # e = exp(10 (1 - 0.9)) = e^1
    mu, lam = lame(O.SNOW, 10.0, mu0, lam0, 0.9)
    assert abs(mu / mu0 - np.e) < 1e-14 and abs(lam / lam0 - np.e) < 1e-14, (mu / mu0, lam / lam0)
    # Jp = 1.05: e = exp(10 * -0.05) = e^-1/2
    mu, lam = lame(O.SNOW, 10.0, mu0, lam0, 1.05)
    assert abs(mu / mu0 - np.exp(-0.5)) < 1e-14 and abs(lam / lam0 - np.exp(-0.5)) < 1e-14
    # constant hardening ignores Jp
    mu, lam = lame(O.ELASTIC, 2.0, mu0, lam0, 0.9)
    assert mu == 2.0 * mu0 and lam == 2.0 * lam0


def test_pin_snow_hardening():
    pin_snow_hardening(O.lame)


@pytest.mark.parametrize("old,new", [
    ("np.exp(h * (1.0 - Jp))", "np.exp(h * (Jp - 1.0))"),   # negated exponent
    ("np.exp(h * (1.0 - Jp))", "np.exp(1.0 - Jp)"),          # dropped h
    ("lam = np.where(kind == WATER, lam0, lam0 * e)", "lam = np.where(kind == WATER, lam0, lam0)"),
])
def test_pin_snow_hardening_catches(old, new):
    with pytest.raises(AssertionError):
        pin_snow_hardening(mutant(O.lame, old, new))


# ------------------------------------------------- snow clamp and Jp ----
def pin_snow_plasticity(g2p):
    tc, ts = 2.5e-2, 7.5e-3
    cases = [(1.02, 1.0), (1.02, 0.7), (0.9, 1.0), (0.9, 0.61), (1.3, 19.0), (1.003, 1.2)]
    n = len(cases)
    s = np.array([c[0] for c in cases])
    Jp0 = np.array([c[1] for c in cases])
    x = np.full((n, 3), 0.5)
    v, C = np.zeros((n, 3)), np.zeros((n, 3, 3))
    F = s[:, None, None] * np.eye(3)[None]
    Jp = Jp0.reshape(n, 1).copy()
    res = 8
    # zero grid velocity: C = 0, so F_new before plasticity is F itself
    g2p(x, v, F, C, Jp, np.full(n, O.SNOW), tc, ts, float(res), 1e-4, np.zeros((res + 1,) * 3 + (3,)))
    c = np.clip(s, 1.0 - tc, 1.0 + ts)
    assert np.abs(F - c[:, None, None] * np.eye(3)[None]).max() < 1e-14, "clamped F is not c I"
    expect = np.clip(Jp0 * s ** 3 / (c ** 3 + 1e-10), 0.6, 20.0)
    assert np.abs(Jp.reshape(n) - expect).max() < 1e-13, (Jp.reshape(n), expect)


def test_pin_snow_plasticity():
    pin_snow_plasticity(O.g2p)


@pytest.mark.parametrize("old,new", [
    ("np.clip(Jsel * old_J / det, 0.6, 20.0)", "np.clip(Jsel * det / old_J, 0.6, 20.0)"),   # inverted ratio
    ("np.clip(Jsel * old_J / det, 0.6, 20.0)", "np.clip(Jsel * old_J / det, 0.6, 2.0)"),    # wrong upper clip
    ("np.clip(sig, 1.0 - theta_c[sel][:, None], 1.0 + theta_s[sel][:, None])",
     "np.clip(sig, 1.0 - theta_s[sel][:, None], 1.0 + theta_c[sel][:, None])"),            # swapped thetas
    ("old_J = np.linalg.det(Fs)", "old_J = 1.0"),                                           # lost J_old
])
def test_pin_snow_plasticity_catches(old, new):
    with pytest.raises(AssertionError):
        pin_snow_plasticity(mutant(O.g2p, old, new))


# ---------------------------------------------------------------- water ----
def _rot(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def pin_water(affine_momentum, g2p, lame=O.lame):
    rng = np.random.default_rng(7)
    n = 5
    F = np.stack([_rot(rng) @ np.diag(0.9 + 0.2 * rng.random(3)) @ _rot(rng) for _ in range(n)])
    C = 0.1 * rng.standard_normal((n, 3, 3))
    mass, K = 0.3 + rng.random(n), 1e4
    inv_dx, dt, vol = 64.0, 5e-5, 2e-6
    kind = np.full(n, O.WATER)
    mu, lam = lame(kind, np.ones(n), np.full(n, 123.0), np.full(n, K), np.full(n, 0.8))
    assert np.all(mu == 0.0) and np.all(lam == K)
    A = affine_momentum(F, C, mass, mu, lam, kind, inv_dx, dt, vol, quirks="reference")
    J = np.linalg.det(F)
    PF = (K * (J - 1) * J)[:, None, None] * np.eye(3)[None]
    expect = -(dt * vol) * 4 * inv_dx * inv_dx * PF + mass[:, None, None] * C
    assert np.abs(A - expect).max() / np.abs(expect).max() < 1e-13, "water stress is not K (J-1) J I"
    # G2P on a zero grid: F <- cbrt(J) I, Jp untouched
    x = np.full((n, 3), 0.5)
    Fg, Jp = F.copy(), np.full((n, 1), 0.8)
    g2p(x, np.zeros((n, 3)), Fg, np.zeros((n, 3, 3)), Jp, kind, 2.5e-2, 7.5e-3, 8.0, 1e-4, np.zeros((9, 9, 9, 3)))
    assert np.abs(Fg - np.cbrt(J)[:, None, None] * np.eye(3)[None]).max() < 1e-14, "water F is not cbrt(J) I"
    assert np.all(Jp == 0.8)


def test_pin_water():
    pin_water(O.affine_momentum, O.g2p)


@pytest.mark.parametrize("which,old,new", [
    ("am", "shape = np.where((kind == WATER)[..., None, None], eye, np.ones((dim, dim)))",
     "shape = np.ones((dim, dim))"),                                                        # broadcast for water
    ("am", "vol_term = (lam * (J - 1) * J)[..., None, None]", "vol_term = (lam * (J - 1))[..., None, None]"),
    ("g2p", "Fn[water] = np.cbrt(Jw)[:, None, None] * np.eye(dim)[None]",
     "Fn[water] = Jw[:, None, None] * np.eye(dim)[None]"),                                  # lost cube root
    ("lame", "mu = np.where(kind == WATER, 0.0, mu0 * e)", "mu = mu0 * e"),                 # shear kept
])
def test_pin_water_catches(which, old, new):
    am, g2p, lame = O.affine_momentum, O.g2p, O.lame
    if which == "am":
        am = mutant(am, old, new)
    elif which == "g2p":
        g2p = mutant(g2p, old, new)
    else:
        lame = mutant(lame, old, new)
    with pytest.raises(AssertionError):
        pin_water(am, g2p, lame)


# --------------------------------------- the C restatement inherits the pins ----
def _cstate(x, v, F, C, Jp, mass, kind, h=1.0, tc=2.5e-2, ts=7.5e-3, mu0=300.0, lam0=200.0):
    from oracle import c_oracle as CO

    n = len(x)
    st = dict(x=x, v=v, F=F, C=C, Jp=Jp, mass=mass, mu0=np.full(n, mu0), lam0=np.full(n, lam0),
              h=np.full(n, h), theta_c=np.full(n, tc), theta_s=np.full(n, ts), kind=np.asarray(kind, np.int32))
    return CO, CO.CState(st)


def test_c_p2g_matches_pinned_affine_scatter():
    import ctypes as ct

    x, v, C, F, mass, a, B, res = _affine_field_scene(3)
    n = len(x)
    CO, cs = _cstate(x, v, F, C, np.ones((n, 1)), mass, np.zeros(n))
    prm = CO.make_params(dict(res=res, inv_dx=float(res), dx=1.0 / res, dt=1e-4, volume=1e-3, gravity=0.0))
    r1 = res + 1
    gv_c, gm_c = np.zeros(r1 ** 3 * 3), np.zeros(r1 ** 3)
    assert CO.lib().oracle_p2g3d(ct.byref(cs.struct), ct.byref(prm), gv_c.ctypes.data, gm_c.ctypes.data, 1) == 0
    gv, gm = np.zeros((r1, r1, r1, 3)), np.zeros((r1, r1, r1, 1))
    O.p2g(x, v, F, C, np.ones((n, 1)), mass, np.full(n, 300.0), np.full(n, 200.0), np.zeros(n, int), np.ones(n),
          float(res), 1.0 / res, 1e-4, 1e-3, gv, gm)
    assert np.abs(gv_c - gv.reshape(-1)).max() <= 1e-13 * np.abs(gv).max()
    assert np.abs(gm_c - gm.reshape(-1)).max() <= 1e-13 * np.abs(gm).max()


def test_c_g2p_matches_pinned_snow_and_water():
    import ctypes as ct

    tc, ts = 2.5e-2, 7.5e-3
    s = np.array([1.02, 1.02, 0.9, 0.9, 1.3, 1.003])
    Jp0 = np.array([1.0, 0.7, 1.0, 0.61, 19.0, 1.2])
    n = len(s)
    rng = np.random.default_rng(7)
    Fw = np.stack([_rot(rng) @ np.diag(0.9 + 0.2 * rng.random(3)) @ _rot(rng) for _ in range(3)])
    F = np.concatenate([s[:, None, None] * np.eye(3)[None], Fw])
    kind = np.array([O.SNOW] * n + [O.WATER] * 3)
    Jp = np.concatenate([Jp0, np.full(3, 0.8)]).reshape(-1, 1)
    m = len(kind)
    x = np.full((m, 3), 0.5)
    CO, cs = _cstate(x, np.zeros((m, 3)), F, np.zeros((m, 3, 3)), Jp, np.ones(m), kind, tc=tc, ts=ts)
    prm = CO.make_params(dict(res=8, inv_dx=8.0, dx=0.125, dt=1e-4, volume=1e-3, gravity=0.0))
    assert CO.lib().oracle_g2p3d(ct.byref(cs.struct), ct.byref(prm), np.zeros(9 ** 3 * 3).ctypes.data, 1) == 0
    got = cs.to_state()
    c = np.clip(s, 1.0 - tc, 1.0 + ts)
    assert np.abs(got["F"][:n] - c[:, None, None] * np.eye(3)[None]).max() < 1e-14
    assert np.abs(got["Jp"][:n, 0] - np.clip(Jp0 * s ** 3 / (c ** 3 + 1e-10), 0.6, 20.0)).max() < 1e-13
    J = np.linalg.det(Fw)
    assert np.abs(got["F"][n:] - np.cbrt(J)[:, None, None] * np.eye(3)[None]).max() < 1e-14
