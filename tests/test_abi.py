"""The C-ABI library loads without a GPU and exports every entry point include/*.h declares;
the reference's own test programs pass where the reference was compiled."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dsmc_c.h")
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]
LIB = os.path.join(ROOT, "dsmc_erlend_amd", "libdsmc_hip.so")


def declared_functions(headers=None):
    names = set()
    for h in headers or HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(dsmc_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_step_api():
    names = declared_functions([HEADER])
    for n in ("dsmc_create", "dsmc_move", "dsmc_boundary", "dsmc_index", "dsmc_collide",
              "dsmc_sample", "dsmc_step", "dsmc_get_cells", "dsmc_set_particles",
              "dsmc_get_particles", "dsmc_get_index", "dsmc_last_error", "dsmc_destroy",
              "dsmc_init_freestream", "dsmc_init_from_fields", "dsmc_write_field"):
        assert n in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build with python -m dsmc_erlend_amd.build"
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert "dsmc_net_train" in names and "dsmc_net_predict" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_net_config_layout_and_host_normalisation():
    """dsmc_net_normalize runs on the host only: the library's target normalisation against the
    oracle's, on a field with a shock-like jump, the obstacle and empty cells."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import surrogate_oracle as so
    from dsmc_erlend_amd import default_params
    from dsmc_erlend_amd import surrogate as S
    cfg = S.default_config()
    assert (cfg.height, cfg.width, cfg.base_channels, cfg.max_batch) == (256, 256, 32, 8)
    cfg = S.default_config(height=32, width=24)
    cases = [default_params(attack_angle_deg=a, mach=m, wall_model=w)
             for a, m, w in ((15.0, 4.0, 0), (-20.0, 7.0, 2))]
    rng = np.random.default_rng(0)
    f = np.abs(rng.normal(size=(2, 4, 32, 24))).astype(np.float32)
    f[:, 0] *= 1e17
    f[:, 3] *= 500
    f[:, 0, :5] = 0.0          # empty cells
    f[1, 0, 10:, :] *= 4.0     # a jump
    t, w = S.normalize(cfg, cases, f)
    for i, p in enumerate(cases):
        tg, wg = so.normalize(p, f[i])
        np.testing.assert_allclose(t[i], tg, rtol=2e-6, atol=1e-6)
        np.testing.assert_allclose(w[i], wg, rtol=1e-6, atol=0)


def test_params_struct_layout_matches_header():
    from dsmc_erlend_amd._abi import Params, State, Constants
    lib = ctypes.CDLL(LIB)
    p = Params()
    lib.dsmc_params_default(ctypes.byref(p))
    # the reference's compile-time defaults (src/Geometry.cpp, src/Initialization.cpp,
    # src/Boundary.cpp); a mismatched struct layout scrambles these
    assert (p.nx, p.ny, p.dt, p.n_step, p.first_sampling_step) == (100, 100, 1e-5, 100000, 10000)
    assert (p.box_width, p.box_height, p.mach, p.n_free, p.t_free) == (4.0, 4.0, 4.0, 1e17, 200.0)
    assert (p.wall_model, p.acc_n, p.acc_t, p.t_wall) == (2, 0.8, 0.8, 300.0)
    assert (p.d_ref, p.viscosity_index, p.t_ref, p.seed) == (4.17e-10, 0.81, 273.0, 1)
    assert ctypes.sizeof(State) == 15 * 8 and ctypes.sizeof(Constants) == 16 * 8


def test_enums_match_header():
    from dsmc_erlend_amd import _abi
    text = open(HEADER).read()
    enum = dict((k, int(v)) for k, v in re.findall(r"\b(DSMC_[A-Z_]+)\s*=\s*(\d+)", text))
    assert enum["DSMC_SCHEDULE_LARGE"] == _abi.SCHEDULE_LARGE
    assert enum["DSMC_SCHEDULE_SMALL"] == _abi.SCHEDULE_SMALL
    assert enum["DSMC_NUM_STAGES"] == len(_abi.STAGES)
    assert (enum["DSMC_WALL_DIFFUSE"], enum["DSMC_WALL_SPECULAR"], enum["DSMC_WALL_CLL"]) == (
        _abi.WALL_DIFFUSE, _abi.WALL_SPECULAR, _abi.WALL_CLL)
    assert enum["DSMC_E_CLL_COEFF"] == _abi.E_CLL_COEFF
    assert ctypes.sizeof(_abi.Tuning) == 12
    assert "12288" in text and _abi.SMALL_MAX_CELLS == 12288


def test_python_mirror_fails_loudly_without_library(tmp_path):
    import dsmc_erlend_amd as D
    with pytest.raises(RuntimeError):
        D.load_library(str(tmp_path / "missing.so"))


@pytest.mark.parametrize("prog", ["geometry_tests", "cll_tests"])
def test_reference_own_tests_pass(prog):
    path = os.path.join(ROOT, "oracle", "_ref", prog)
    if not os.path.exists(path):
        pytest.skip("reference not compiled here")
    r = subprocess.run([path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
