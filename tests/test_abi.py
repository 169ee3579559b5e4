"""CPU checks of the C ABI: the library loads without a GPU and exports every
function include/mpm355.h declares; argument validation fails cleanly."""
import ctypes as ct
import os
import re

import pytest

from femflow_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "mpm355.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert sorted(_lib.EXPORTS) == declared_functions()


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for name in declared_functions():
        assert hasattr(L, name), name
    assert L.mpm_version() >= 1


def test_device_count_without_gpu_is_safe():
    assert _lib.load().mpm_device_count() >= 0


def test_argument_errors_do_not_touch_the_gpu():
    L = _lib.load()
    assert L.mpm_run(None, 1, None) == _lib.MPM_ERR_ARG
    assert "null context" in _lib.last_error()
    assert L.mpm_ctx_tune(None, 512, 1, 1, 1) == _lib.MPM_ERR_ARG
    assert L.mpm_ctx_set_spread(None, 3) == _lib.MPM_ERR_ARG
    bad = _lib.Params(4, 16, 16.0, 1 / 16, 1e-4, 1.0, -9.8, 0, 3, 0, 1)
    assert L.mpm_grid_op(None, ct.byref(bad), _lib.MPM_F32, None, None) == _lib.MPM_ERR_ARG
    # plane colliders: null grid, too many planes, bad quirks mode, null context
    pts = (ct.c_double * 3)(0.5, 0.5, 0.5)
    nrm = (ct.c_double * 3)(0.0, 1.0, 0.0)
    assert L.mpm_collide_planes(None, _lib.MPM_F32, 8, 0.125, 1, pts, nrm, 0, None) == _lib.MPM_ERR_ARG
    assert L.mpm_collide_planes(None, _lib.MPM_F32, 8, 0.125, 17, pts, nrm, 0, ct.c_void_p(16)) == _lib.MPM_ERR_ARG
    assert L.mpm_collide_planes(None, _lib.MPM_F32, 8, 0.125, 1, pts, nrm, 7, ct.c_void_p(16)) == _lib.MPM_ERR_ARG
    assert L.mpm_ctx_set_planes(None, 0, 1, pts, nrm, 0) == _lib.MPM_ERR_ARG
    # replayed-grid adjoint: null context, bad parameters, null grids
    assert L.mpm_run_states_grids(None, 1, None, None, None, None, None, None) == _lib.MPM_ERR_ARG
    pr3 = _lib.Params(3, 16, 16.0, 1 / 16, 1e-4, 1.0, -9.8, 0, 3, 0, 1)
    parts = _lib.Particles(0, _lib.MPM_F32, 0, None, None, None, None, None, None, None, None, None)
    assert L.mpm_adjoint_step_grids(None, ct.byref(bad), ct.byref(parts), 1.0, None, ct.c_void_p(16), None, None,
                                    None, None, None, None, None) == _lib.MPM_ERR_ARG
    assert L.mpm_adjoint_step_grids(None, ct.byref(pr3), ct.byref(parts), 1.0, None, ct.c_void_p(16), None, None,
                                    None, None, None, None, None) == _lib.MPM_ERR_ARG
    assert "null grid" in _lib.last_error()
    with pytest.raises(ValueError):
        _lib.check(_lib.MPM_ERR_ARG)
    with pytest.raises(RuntimeError):
        _lib.check(_lib.MPM_ERR_OOB)


def test_model_string_mapping():
    from femflow_amd.solvers.mpm import _phase

    # three_d/p2g.py:57-61 and three_d/g2p.py:48
    assert _phase.p2g_kind("neo_hookean") == _lib.MPM_ELASTIC
    assert _phase.p2g_kind("snow") == _lib.MPM_SNOW
    assert _phase.p2g_kind("anything") == _lib.MPM_SNOW
    assert _phase.g2p_kind("snow") == _lib.MPM_SNOW
    assert _phase.g2p_kind("anything") == _lib.MPM_ELASTIC


def test_product_path_has_no_cpu_fallback():
    import torch

    from femflow_amd.solvers.mpm import three_d

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import numpy as np

    with pytest.raises(RuntimeError):
        three_d.grid_op(8, 1 / 8, 1e-3, -9.8, np.zeros((9, 9, 9, 3)), np.zeros((9, 9, 9, 1)))


def test_package_never_imports_oracle():
    pkg = os.path.join(ROOT, "femflow_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".h", ".cpp")):
                with open(os.path.join(dirpath, fn)) as f:
                    src = f.read()
                assert "import oracle" not in src and "from oracle" not in src, fn


def test_option_strings_are_validated():
    import numpy as np

    from femflow_amd.engine import Engine, MaterialSpec, SceneSpec
    from femflow_amd.solvers.mpm import _phase, three_d, two_d

    assert _phase.boundary_code("sticky") == _lib.MPM_BC_STICKY
    assert _phase.quirks_code("textbook") == _lib.MPM_QUIRKS_TEXTBOOK
    g3, m3 = np.zeros((9, 9, 9, 3)), np.zeros((9, 9, 9, 1))
    with pytest.raises(ValueError):
        three_d.grid_op(8, 1 / 8, 1e-3, -9.8, g3, m3, boundary="slippery")
    with pytest.raises(ValueError):
        two_d.grid_op(8, 1e-3, -9.8, np.zeros((9, 9, 2)), np.zeros((9, 9)), boundary="Sticky")
    with pytest.raises(ValueError):
        _phase.quirks_code("numba")
    with pytest.raises(ValueError):
        Engine([SceneSpec(res=16, dt=1e-4, boundary="open")], [MaterialSpec("elastic", 1.0)], n_max=8,
               device="cpu")


@pytest.mark.gpu
def test_mpm_substep_with_changing_counts(gpu):
    """mpm_substep on one context, called with 2000 particles, then 1000, then
    2000 again, straight through ctypes: every call must match the FP64 oracle
    and write only the caller's arrays of that call's count."""
    import numpy as np
    import torch

    from oracle import mpm_oracle as O
    from tests._parity import field_errors
    from tests.gpu_helpers import oracle_copy_from, to_numpy_state, to_torch_state

    L = _lib.load()
    st, pr = O.scene_3d_cubes(n=2000, res=32)
    sc = (_lib.Scene * 1)(_lib.Scene(pr["res"], _lib.MPM_BC_REFERENCE, 3, _lib.MPM_QUIRKS_REFERENCE, pr["dt"],
                                     pr["gravity"], pr["volume"]))
    mats = (_lib.Material * 3)(*[_lib.Material(k, h, tc, ts) for k, _, _, h, tc, ts in O.MATERIALS_C2])
    h = ct.c_void_p()
    assert L.mpm_ctx_create(ct.byref(h), 2000, 1, sc, 3, mats, 0) == _lib.MPM_OK
    L.mpm_ctx_set_stream(h, torch.cuda.current_stream(gpu).cuda_stream)
    try:
        for n in (2000, 1000, 2000):
            sub = {k: (a[:n].copy() if isinstance(a, np.ndarray) and len(a) == 2000 else a) for k, a in st.items()}
            ts = to_torch_state(sub, gpu)
            ref = oracle_copy_from(ts, sub)
            xbuf = torch.full((2000 * 3 + 300,), 5.0, device=gpu)
            xbuf[: 3 * n] = ts["x"].reshape(-1)
            p = _lib.Particles(n, _lib.MPM_F32, 0, xbuf.data_ptr(), ts["v"].data_ptr(), ts["F"].data_ptr(),
                               ts["C"].data_ptr(), ts["Jp"].data_ptr(), ts["mass"].data_ptr(),
                               ts["mu0"].data_ptr(), ts["lam0"].data_ptr(), ts["kind"].data_ptr())
            assert L.mpm_substep(h, ct.byref(p), None, 3) == _lib.MPM_OK, _lib.last_error()
            torch.cuda.synchronize()
            for _ in range(3):
                O.substep(ref, pr)
            got = dict(ts)
            got["x"] = xbuf[: 3 * n].reshape(n, 3)
            errs = field_errors(to_numpy_state(got), ref, pr["inv_dx"])
            for k in ("x", "v", "F"):
                assert errs[k] < 1e-4, (n, errs)
            assert bool((xbuf[3 * n:] == 5.0).all()), "mpm_substep wrote past the caller's count"
    finally:
        L.mpm_ctx_destroy(h)
