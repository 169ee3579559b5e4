"""Pins the CPU oracle (oracle/mpm_oracle.py, oracle/mpm_oracle.c) before it is
trusted as the checker of the HIP path.

Each check is chosen so that a plausible slip (a dropped factor, a wrong sign
or index, a transposed operand) fails: the reference's own known-answer
vectors, closed forms (partition of unity, linear-field reproduction,
conservation, rigid rotations, pure dilations) and independent library
routines (scipy's polar decomposition, numpy's stable argsort).
"""
import json
import os
from collections import Counter

import numpy as np
import pytest
import scipy.linalg

from oracle import c_oracle as CO
from oracle import mpm_oracle as O

from tests._parity import field_errors  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def kats():
    with open(os.path.join(GOLDEN, "ref_kats.json")) as f:
        return json.load(f)


# ---------------------------------------------------------------- KATs ----
def test_weights_kat(kats):
    w = O.weights(np.array(kats["weights"]["fx"]))
    assert np.allclose(w, np.array(kats["weights"]["expected"]), rtol=1e-5, atol=1e-8)


def test_lame_kat(kats):
    p = kats["params"]
    mu0, lam0 = O.Ev_to_mu(p["E"], p["nu"]), O.Ev_to_lambda(p["E"], p["nu"])
    # snow hardening at Jp = 1 has factor exp(h*0) = 1 (utils.py:48)
    mu, lam = O.lame(O.SNOW, p["hardening"], mu0, lam0, 1.0)
    assert np.isclose(mu, kats["lame"]["mu"]) and np.isclose(lam, kats["lame"]["lambda"])
    # constant hardening multiplies by h (utils.py:24)
    mu, lam = O.lame(O.ELASTIC, 3.0, mu0, lam0, 1.0)
    assert np.isclose(mu, 3 * mu0) and np.isclose(lam, 3 * lam0)
    mu, lam = O.lame(O.WATER, 3.0, 0.0, 1e4, 0.5)
    assert mu == 0.0 and lam == 1e4


def _kat_stress(kats, quirks):
    p = kats["params"]
    mu0, lam0 = O.Ev_to_mu(p["E"], p["nu"]), O.Ev_to_lambda(p["E"], p["nu"])
    F = np.array(kats["stress_2d"]["F"], dtype=np.float64)
    C = np.array(kats["stress_2d"]["C"], dtype=np.float64)
    return O.affine_momentum(F, C, p["mass"], mu0, lam0, O.ELASTIC, 1.0 / p["dx"], p["dt"],
                             p["volume"], quirks=quirks)


def test_stress_2d_kat_textbook(kats):
    A = _kat_stress(kats, "textbook")
    assert np.allclose(A, np.array(kats["stress_2d"]["expected"]), rtol=1e-4, atol=1e-12)


def test_stress_2d_kat_reference_quirk(kats):
    """Reference mode reproduces the off-diagonals; the diagonal carries the
    eps=1e-10 of polar_decomp_2d (linear_algebra.py:110), whose closed form is
    -(dt vol 4 inv_dx^2) 2 mu (1 - 2/(2+eps))."""
    p = kats["params"]
    A = _kat_stress(kats, "reference")
    exp = np.array(kats["stress_2d"]["expected"])
    assert np.allclose(A[0, 1], exp[0, 1], rtol=1e-4) and np.allclose(A[1, 0], exp[1, 0], rtol=1e-4)
    mu0 = O.Ev_to_mu(p["E"], p["nu"])
    inv_dx = 1.0 / p["dx"]
    eps_term = -(p["dt"] * p["volume"]) * 4 * inv_dx ** 2 * 2 * mu0 * (1 - 2 / (2 + 1e-10))
    assert np.isclose(A[0, 0], eps_term, rtol=1e-3)
    assert np.isclose(A[1, 1], eps_term + exp[1, 1], rtol=1e-3)


def _kat_scene(kats):
    p = kats["params"]
    x = np.array(kats["p2g_2d"]["positions"], dtype=np.float64)
    v, F, C, Jp = O.make_mls_mpm_coefficients(len(x), 2)
    res = p["grid_resolution"]
    gv = np.zeros((res + 1, res + 1, 2))
    gm = np.zeros((res + 1, res + 1))
    mu0, lam0 = O.Ev_to_mu(p["E"], p["nu"]), O.Ev_to_lambda(p["E"], p["nu"])
    return p, x, v, F, C, Jp, gv, gm, mu0, lam0


def test_p2g_2d_ground_truth_grid(kats):
    p, x, v, F, C, Jp, gv, gm, mu0, lam0 = _kat_scene(kats)
    O.p2g_2d(1 / p["dx"], p["hardening"], mu0, lam0, p["mass"], p["dx"], p["dt"], p["volume"], gv, gm,
             x, v, F, C, Jp, model="snow")
    got = Counter(gm[gm != 0].tolist())
    assert got == Counter(kats["p2g_2d"]["mass_multiset"])
    # the file on disk says the same (ground_truth_grid.txt third column)
    assert sorted(got.elements()) == sorted(kats["p2g_2d"]["mass_multiset"])


@pytest.mark.parametrize("quirks", ["reference", "textbook"])
def test_g2p_2d_one_step_kat(kats, quirks):
    p, x, v, F, C, Jp, gv, gm, mu0, lam0 = _kat_scene(kats)
    res = p["grid_resolution"]
    O.p2g_2d(1 / p["dx"], p["hardening"], mu0, lam0, p["mass"], p["dx"], p["dt"], p["volume"], gv, gm,
             x, v, F, C, Jp, model="snow", quirks=quirks)
    O.grid_op_2d(res, p["dt"], p["gravity"], gv, gm)
    O.g2p_2d(1 / p["dx"], p["dt"], gv, x, v, F, C, Jp, model="snow")
    g = kats["g2p_2d"]
    assert np.allclose(x, np.array(g["expected_positions"]))
    assert np.allclose(v, np.array(g["expected_velocity"])[None, :])
    if quirks == "textbook":
        assert np.allclose(C, np.array(g["expected_C"])[None], atol=g["expected_C_atol"])


# ------------------------------------------------------- closed forms ----
def test_weight_moments():
    fx = np.linspace(0.5, 1.5, 101)[:-1]
    w = O.weights(fx)
    i = np.arange(3)[:, None]
    assert np.allclose(w.sum(0), 1.0)
    assert np.allclose((w * (i - fx)).sum(0), 0.0, atol=1e-14)
    assert np.allclose((w * (i - fx) ** 2).sum(0), 0.25)
    # Q1: base truncated toward zero next to the low wall still sums to one
    fx_low = np.linspace(0.0, 0.5, 11)
    assert np.allclose(O.weights(fx_low).sum(0), 1.0)
    assert (O.weights(np.array([0.0]))[1] < 0).all()


def test_polar_matches_scipy():
    rng = np.random.default_rng(3)
    F = np.eye(3) + 0.3 * rng.standard_normal((50, 3, 3))
    R = O.polar_decomp_3d(F)
    for f, r in zip(F, R):
        u, _ = scipy.linalg.polar(f)
        assert np.allclose(r, u, atol=1e-12)


def test_stress_rotation_is_zero_and_dilation_closed_form():
    rng = np.random.default_rng(4)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    q *= np.sign(np.linalg.det(q))
    C = rng.standard_normal((3, 3))
    A = O.affine_momentum(q[None], C[None], np.array([2.0]), np.array([100.0]), np.array([50.0]),
                          O.ELASTIC, 16.0, 1e-3, 0.01)
    assert np.allclose(A[0], 2.0 * C, atol=1e-10)
    s, mu, lam, inv_dx, dt, vol = 1.01, 100.0, 50.0, 16.0, 1e-3, 0.01
    F = s * np.eye(3)
    A = O.affine_momentum(F[None], np.zeros((1, 3, 3)), np.array([1.0]), np.array([mu]),
                          np.array([lam]), O.ELASTIC, inv_dx, dt, vol)
    J = s ** 3
    PF = 2 * mu * (s - 1) * s * np.eye(3) + lam * (J - 1) * J * np.ones((3, 3))
    assert np.allclose(A[0], -(dt * vol) * 4 * inv_dx ** 2 * PF, rtol=1e-12)
    A = O.affine_momentum(F[None], np.zeros((1, 3, 3)), np.array([1.0]), np.array([mu]),
                          np.array([lam]), O.ELASTIC, inv_dx, dt, vol, quirks="textbook")
    PF = (2 * mu * (s - 1) * s + lam * (J - 1) * J) * np.eye(3)
    assert np.allclose(A[0], -(dt * vol) * 4 * inv_dx ** 2 * PF, rtol=1e-12)


def test_p2g_conserves_mass_and_momentum():
    st, pr = O.scene_3d_cubes(n=3000, res=32)
    rng = np.random.default_rng(5)
    st["F"] = st["F"] + 0.01 * rng.standard_normal(st["F"].shape)
    st["C"] = rng.standard_normal(st["C"].shape)
    r1 = pr["res"] + 1
    gv = np.zeros((r1, r1, r1, 3))
    gm = np.zeros((r1, r1, r1, 1))
    O.p2g(st["x"], st["v"], st["F"], st["C"], st["Jp"], st["mass"], st["mu0"], st["lam0"], st["kind"],
          st["h"], pr["inv_dx"], pr["dx"], pr["dt"], pr["volume"], gv, gm)
    assert np.isclose(gm.sum(), st["mass"].sum(), rtol=1e-12)
    assert np.allclose(gv.reshape(-1, 3).sum(0), (st["mass"][:, None] * st["v"]).sum(0), rtol=1e-10,
                       atol=1e-12)


def test_g2p_reproduces_linear_field():
    rng = np.random.default_rng(6)
    res = 16
    r1 = res + 1
    dx = 1.0 / res
    a = rng.standard_normal(3)
    B = rng.standard_normal((3, 3))
    I = np.stack(np.meshgrid(*(np.arange(r1),) * 3, indexing="ij"), -1) * dx
    gv = a + I @ B.T
    x = 0.2 + 0.6 * rng.random((64, 3))
    v, F, C, Jp = O.make_mls_mpm_coefficients(64, 3)
    x0 = x.copy()
    O.g2p_3d(float(res), 1e-3, gv, x, v, F, C, Jp)
    assert np.allclose(v, a + x0 @ B.T, atol=1e-12)
    assert np.allclose(C, np.broadcast_to(B, C.shape), atol=1e-10)
    assert np.allclose(x, x0 + 1e-3 * v, atol=1e-15)
    assert np.allclose(F, np.eye(3) + 1e-3 * B, atol=1e-12)


def test_grid_op_3d_brute_force():
    rng = np.random.default_rng(7)
    res, dx, dt, g = 6, 1 / 6, 1e-2, -9.8
    gv = rng.standard_normal((7, 7, 7, 3))
    gm = np.where(rng.random((7, 7, 7, 1)) < 0.7, rng.random((7, 7, 7, 1)), 0.0)
    exp = gv.copy()
    lim = dx * 0.9 / dt
    for i in range(7):
        for j in range(7):
            for k in range(7):
                m = gm[i, j, k, 0]
                if m > 0:
                    exp[i, j, k] = np.clip(exp[i, j, k] / m + np.array([0, dt * g, 0]), -lim, lim)
                for d, I in enumerate((i, j, k)):
                    if I < 1 or I >= res - 1:
                        exp[i, j, k, d] = 0
    O.grid_op_3d(res, dx, dt, g, gv, gm)
    assert np.array_equal(gv, exp)


def test_grid_op_2d_brute_force():
    rng = np.random.default_rng(8)
    res, dt, g = 40, 1e-3, -9.8
    gv = rng.standard_normal((41, 41, 2))
    gm = np.where(rng.random((41, 41)) < 0.7, rng.random((41, 41)), 0.0)
    exp = gv.copy()
    for i in range(41):
        for j in range(41):
            if gm[i, j] > 0:
                e = exp[i, j] / gm[i, j] + np.array([0, dt * g])
                xs, ys = i / res, j / res
                if xs < 0.05 or xs > 0.95 or ys > 0.95:
                    e[:] = 0
                if ys < 0.05:
                    e[1] = max(0.0, e[1])
                exp[i, j] = e
    O.grid_op_2d(res, dt, g, gv, gm)
    assert np.array_equal(gv, exp)


@pytest.mark.parametrize("res,bound", [(16, 3), (20, 1), (12, 5)])
def test_sticky_walls_node_counts(res, bound):
    """Sticky walls of thickness `bound`: the untouched nodes are exactly the
    (res - 2 bound + 1)^d interior lattice (counted, not re-derived per node),
    and in 2D only nodes with mass are touched."""
    n1 = res + 1
    inner = res - 2 * bound + 1
    gv = np.ones((n1, n1, n1, 3))
    O.grid_op_3d(res, 1.0 / res, 1e-4, 0.0, gv, np.ones((n1, n1, n1, 1)), boundary="sticky", bound=bound)
    zero = np.all(gv == 0.0, axis=-1)
    assert zero.sum() == n1 ** 3 - inner ** 3
    core = slice(bound, res - bound + 1)
    assert np.all(gv[core, core, core] == 1.0) and not zero[core, core, core].any()
    gv2 = np.full((n1, n1, 2), 2.0)
    gm2 = np.full((n1, n1), 2.0)
    gm2[0, :] = 0.0  # massless nodes on a wall keep their value
    O.grid_op_2d(res, 1e-4, 0.0, gv2, gm2, boundary="sticky", bound=bound)
    zero2 = np.all(gv2 == 0.0, axis=-1)
    assert zero2.sum() == n1 ** 2 - inner ** 2 - n1
    assert np.all(gv2[0, :] == 2.0) and np.all(gv2[core, core] == 1.0)


def test_snow_gauges_agree_on_mechanics():
    st, pr = O.scene_3d_cubes(n=2000, res=32)
    st["kind"][:] = O.SNOW
    rng = np.random.default_rng(9)
    st["F"] = st["F"] + 0.03 * rng.standard_normal(st["F"].shape)
    a, b = O.copy_state(st), O.copy_state(st)
    for _ in range(3):
        O.substep(a, pr, svd_gauge="reference")
        O.substep(b, pr, svd_gauge="consistent")
    for k in ("x", "v", "C", "Jp"):
        assert np.allclose(a[k], b[k], rtol=1e-9, atol=1e-11), k
    FFa = a["F"] @ np.swapaxes(a["F"], -1, -2)
    FFb = b["F"] @ np.swapaxes(b["F"], -1, -2)
    assert np.allclose(FFa, FFb, atol=1e-12)
    s = np.linalg.svd(b["F"], compute_uv=False)
    assert (s >= 1 - 2.5e-2 - 1e-9).all() and (s <= 1 + 7.5e-3 + 1e-9).all()


def test_oob_raises_in_3d():
    st, pr = O.scene_3d_cubes(n=100, res=32)
    st["x"][7] = [0.5, 0.999, 0.5]  # base = 31 -> base+2 = 33 > res
    with pytest.raises(RuntimeError):
        O.substep(st, pr)
    st, pr = O.scene_3d_cubes(n=100, res=32)
    st["x"][3] = [-0.02, 0.5, 0.5]  # x*inv_dx - 0.5 = -1.14 -> base -1
    with pytest.raises(RuntimeError):
        O.substep(st, pr)


# ------------------------------------------------------------ binning ----
def test_bin_keys_brute_force():
    rng = np.random.default_rng(10)
    res = 37
    x = rng.random((500, 3)).astype(np.float32) * np.float32(1.02) - np.float32(0.01)
    x[:5] = np.float32(0.0)
    x[5] = np.float32(0.5 / res)  # exactly on a cell boundary in float32 arithmetic
    keys_b = O.bin_keys(x, float(res), res, layout="blocked")
    keys_r = O.bin_keys(x, float(res), res, layout="rowmajor")
    nb = (res + 1 + 3) // 4
    inv = np.float32(res)
    for p in range(len(x)):
        b = []
        for d in range(3):
            t = float(np.float32(np.float32(x[p, d]) * inv) - np.float32(0.5))
            c = int(t)  # Python int() truncates toward zero
            b.append(min(max(c, 0), res))
        assert keys_r[p] == (b[0] * (res + 1) + b[1]) * (res + 1) + b[2]
        blk = ((b[0] // 4) * nb + b[1] // 4) * nb + b[2] // 4
        assert keys_b[p] == blk * 64 + (b[0] % 4) * 16 + (b[1] % 4) * 4 + b[2] % 4
    perm = O.bin_sort(keys_b)
    assert np.all(np.diff(keys_b[perm].astype(np.int64)) >= 0)
    # stability: equal keys keep index order
    same = keys_b[perm][1:] == keys_b[perm][:-1]
    assert np.all(perm[1:][same] > perm[:-1][same])


# ------------------------------------------------- two restatements ----
@pytest.mark.parametrize("dim", [2, 3])
def test_c_restatement_matches_numpy(dim):
    if dim == 3:
        st, pr = O.scene_3d_cubes(n=3000, res=32)
        steps = 4
    else:
        st, pr = O.scene_2d_c1(n=3000)
        steps = 4
    a = O.copy_state(st)
    for _ in range(steps):
        O.substep(a, pr)
    b = CO.run(st, pr, steps).to_state()
    errs = field_errors(b, a, pr["inv_dx"])
    assert max(errs.values()) <= 1e-9, errs


def test_c_restatement_oob():
    st, pr = O.scene_3d_cubes(n=100, res=32)
    st["x"][0] = [0.999, 0.5, 0.5]
    with pytest.raises(RuntimeError):
        CO.run(st, pr, 1)


def test_golden_fixtures_match_oracle():
    st, pr = O.scene_3d_cubes(n=6000, res=32)
    s = O.copy_state(st)
    O.substep(s, pr)
    g = np.load(os.path.join(GOLDEN, "oracle_3d_cubes.npz"))
    for k in ("x", "v", "F", "C", "Jp"):
        assert np.allclose(s[k], g[f"s1_{k}"], rtol=1e-12, atol=1e-14), k


# ------------------------------------------------ plane colliders ----
# three_d/grid_op.py:50-67.  Pinned by node counts derived by hand (integer
# lattice arguments), not by re-evaluating the oracle's formula.
@pytest.mark.parametrize("axis,sign,c,scale", [(0, 1, 0.33, 1.0), (1, 1, 0.6, 3.0), (2, -1, 0.71, 0.5),
                                               (1, -1, 0.1, 7.0)])
def test_collide_planes_textbook_axis_counts(axis, sign, c, scale):
    res, dx = 16, 1.0 / 16
    point = np.zeros(3)
    point[axis] = c
    normal = np.zeros(3)
    normal[axis] = sign * scale
    gv = np.ones((res + 1,) * 3 + (3,))
    O.collide_planes([point], [normal], res, dx, gv, quirks="textbook")
    # sign * (i dx - c) < 0: i < c/dx (sign +) or i > c/dx (sign -); c is off the lattice
    rows = sum(1 for i in range(res + 1) if (i < c / dx if sign > 0 else i > c / dx))
    assert int((gv[..., 0] == 0).sum()) == rows * (res + 1) ** 2
    assert np.all((gv == 0).all(-1) | (gv == 1).all(-1))  # whole nodes only


def test_collide_planes_reference_quirk_hand_counts():
    res, dx = 4, 0.25
    # normal (0, 0, 2): n' = (0.5, 0.5, 2.5); dot < 0  <=>  i + j + 5k < 14 -> 25 + 25 + 10 = 60 nodes
    gv = np.ones((5, 5, 5, 3))
    O.collide_planes([(0.5, 0.5, 0.5)], [(0, 0, 2)], res, dx, gv)
    assert int((gv[..., 0] == 0).sum()) == 60
    # normal (0, -2, 0): n' = (0.5, -1.5, 0.5); dot < 0  <=>  3j > i + k + 2 -> 58 nodes
    gv = np.ones((5, 5, 5, 3))
    O.collide_planes([(0.5, 0.5, 0.5)], [(0, -2, 0)], res, dx, gv)
    assert int((gv[..., 0] == 0).sum()) == 58
    # normal (0, -1, 0): n' = (1, 0, 1); dot = (i + k) dx >= 0 everywhere -> nothing is zeroed
    gv = np.ones((5, 5, 5, 3))
    O.collide_planes([(0.0, 0.5, 0.0)], [(0, -1, 0)], res, dx, gv)
    assert int((gv == 0).sum()) == 0


def test_collide_planes_matches_node_loop():
    rng = np.random.default_rng(3)
    res, dx = 8, 1.0 / 8
    pts, nrm = rng.random((3, 3)), rng.normal(size=(3, 3))
    for quirks in ("reference", "textbook"):
        gv = rng.normal(size=(res + 1,) * 3 + (3,))
        ref = gv.copy()
        O.collide_planes(pts, nrm, res, dx, gv, quirks=quirks)
        for p, n in zip(pts, nrm):
            d2 = float(n[0]) * float(n[0]) + float(n[1]) * float(n[1]) + float(n[2]) * float(n[2])
            den = d2 ** 0.5
            nn = [float(v) + 1.0 / den if quirks == "reference" else float(v) / den for v in n]
            for i in range(res + 1):
                for j in range(res + 1):
                    for k in range(res + 1):
                        o = (i * dx - float(p[0]), j * dx - float(p[1]), k * dx - float(p[2]))
                        if (o[0] * nn[0] + o[1] * nn[1]) + o[2] * nn[2] < 0:
                            ref[i, j, k] = 0.0
        assert np.array_equal(gv, ref)


def test_config1_full_rollout_fixture():
    """All 200 substeps of the 2D configuration: the C restatement (reference
    loop order) lands on the NumPy oracle's committed final state."""
    st, pr = O.scene_2d_c1()
    g = np.load(os.path.join(GOLDEN, "oracle_c1_2d.npz"))
    b = CO.run(st, pr, 200).to_state()
    errs = field_errors(b, {k: g[f"s200_{k}"] for k in ("x", "v", "F", "C", "Jp")}, pr["inv_dx"])
    assert max(errs.values()) <= 1e-9, errs


def test_c_restatement_threads_bit_identical():
    """The C restatement's parallel P2G owns node slabs and walks particles in
    list order, so any thread count reproduces the one-thread result bit for
    bit (a deterministic FP64 reference for finite differences)."""
    st, pr = O.scene_3d_cubes(n=20000, res=48)
    runs = [CO.run(st, pr, 4, nthreads=nt).arrays for nt in (1, 2, 5)]
    for other in runs[1:]:
        for k in ("x", "v", "F", "C", "Jp"):
            assert np.array_equal(runs[0][k], other[k]), k
