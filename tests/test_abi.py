"""The C ABI library: loads, exports every declared symbol, validates its
arguments before touching a device (so these run without a GPU)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import steps_amd
from steps_amd import _abi

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    txt = (ROOT / "include" / "steps_force.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(steps_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = steps_amd.load_library()
    decl = declared_symbols()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(decl) == sorted(_abi.EXPORTED)


def test_version_and_desc_defaults():
    lib = steps_amd.load_library()
    assert b"gfx950" in lib.steps_version()
    d = _abi.ForceDesc()
    lib.steps_force_desc_init(C.byref(d))
    assert d.topology == 0 and d.precision == 0 and d.s1r2_lookup == 1 and d.device == 0


def _eval_rc(desc, N, id_min, id_max, x=None):
    lib = steps_amd.load_library()
    x = np.zeros(3 * max(N, 1)) if x is None else x
    M = np.ones(max(N, 1))
    h = np.ones(max(N, 1))
    F = np.zeros(3 * max(N, 1) + 3)
    dp = _abi._ptr
    return lib.steps_force_eval(C.byref(desc), N, dp(x), dp(M), dp(h), dp(F), id_min, id_max)


def test_argument_errors_before_device():
    d, _ = steps_amd.make_desc(0)
    assert _eval_rc(d, -1, 0, 0) == _abi.STEPS_ERR_ARG
    assert _eval_rc(d, 10, 5, 10) == _abi.STEPS_ERR_RANGE     # id_max >= N
    assert _eval_rc(d, 10, -1, 3) == _abi.STEPS_ERR_RANGE     # id_min < 0
    assert _eval_rc(d, 10, 5, 2) == _abi.STEPS_ERR_RANGE      # id_max < id_min - 1
    assert _eval_rc(d, 10, 5, 4) == _abi.STEPS_OK             # empty range is a no-op
    dt, _ = steps_amd.make_desc(1, is_periodic=2, L=10.0)     # T3 Ewald without a table
    assert _eval_rc(dt, 10, 0, 9) == _abi.STEPS_ERR_TABLE
    dt1, _ = steps_amd.make_desc(1, is_periodic=0, L=10.0)    # T3 needs IS_PERIODIC >= 1
    assert _eval_rc(dt1, 10, 0, 9) == _abi.STEPS_ERR_ARG
    ds, _ = steps_amd.make_desc(2, is_periodic=2, L=10.0, Rsim=5.0)
    assert _eval_rc(ds, 10, 0, 9) == _abi.STEPS_ERR_TABLE
    dr, _ = steps_amd.make_desc(2, is_periodic=1, L=10.0, Rsim=5.0, cosmology=1, comoving=1)
    assert _eval_rc(dr, 10, 0, 9) == _abi.STEPS_ERR_TABLE     # radial table needed
    dm, _ = steps_amd.make_desc(2, is_periodic=2, L=10.0, Rsim=5.0, s1r2_lookup=0, precision=1)
    assert _eval_rc(dm, 10, 0, 9) == _abi.STEPS_ERR_ARG       # no mixed kernel for the S1R2 image sum
    bad = np.zeros(30)
    bad[4] = np.nan
    assert _eval_rc(d, 10, 0, 9, x=bad) == _abi.STEPS_ERR_NONFINITE
    assert b"non-finite" in steps_amd.load_library().steps_last_error()


def test_no_device_is_an_error_not_a_fallback():
    if steps_amd.device_count() > 0:
        pytest.skip("a HIP device is present")
    d, _ = steps_amd.make_desc(0)
    assert _eval_rc(d, 10, 0, 9) == _abi.STEPS_ERR_DEVICE
    with pytest.raises(steps_amd.StepsError):
        steps_amd.Sim(d, 10)


def test_softening_lengths_host_entry(oracle):
    rng = np.random.default_rng(1)
    M = rng.uniform(0.1, 10.0, 1000)
    h, mmin, rp = steps_amd.softening_lengths(M, 0.05)
    ho, mmino, rpo = oracle.softening_lengths(M, 0.05)
    assert mmin == mmino and rp == pytest.approx(rpo, rel=1e-15)
    assert np.allclose(h, ho, rtol=1e-15, atol=0)


def test_server_without_device_or_with_bad_config():
    import ctypes as C
    lib = steps_amd.load_library()
    cfg = _abi.ServerConfig()
    lib.steps_server_config_init(C.byref(cfg))
    assert cfg.streams == 4 and cfg.use_graphs == 1 and cfg.hbm_budget_bytes == 200 << 30
    assert cfg.pack_max_n == 65536
    h = C.c_void_p()
    cfg.pack_max_n = -1
    assert lib.steps_server_create(C.byref(cfg), C.byref(h)) == _abi.STEPS_ERR_ARG
    cfg.pack_max_n = 65536
    cfg.streams = 0
    assert lib.steps_server_create(C.byref(cfg), C.byref(h)) == _abi.STEPS_ERR_ARG
    cfg.streams = 4
    cfg.hbm_budget_bytes = 0
    assert lib.steps_server_create(C.byref(cfg), C.byref(h)) == _abi.STEPS_ERR_ARG
    if steps_amd.device_count() == 0:
        cfg.hbm_budget_bytes = 1 << 30
        assert lib.steps_server_create(C.byref(cfg), C.byref(h)) == _abi.STEPS_ERR_DEVICE
