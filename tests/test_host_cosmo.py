"""The expansion history of a run (include/steps_host.h, steps_amd/csrc/host/
cosmo.cc; friedmann_solver.cc), pinned by closed forms and independent
quadrature:
* matter only: a(t) = (a0^(3/2) + 3/2 H0 t)^(2/3) and t(a) = 2/(3 H0) a^(3/2);
* dark energy alone with w0 = -1/3: da/dt = H0, so the RK4 steps are exact;
* wCDM with w0 = -1 is LCDM and w0-wa with wa = 0 is wCDM, bit for bit;
* the w0-wa dark-energy density against exp(3 int_a^1 (1 + w(x)) / x dx) by
  quadrature of the equation of state;
* LCDM: the step against the oracle's restatement, the start time against
  quadrature of da / (a H);
* the tabulated history: linear, quadratic and cubic data come back at
  off-node times and scale factors for interpolation orders 1, 2, 3, and the
  start time is the reference's fall-through value (the row before the first
  one past a_start)."""
import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest
from scipy import integrate

ROOT = Path(__file__).resolve().parent.parent
UNIT_T, UNIT_V = 47.14829951063323, 20.738652969925447


class Expansion(C.Structure):
    _fields_ = [("cosmoparam", C.c_int32), ("H0", C.c_double), ("Omega_m", C.c_double), ("Omega_r", C.c_double),
                ("Omega_lambda", C.c_double), ("Omega_k", C.c_double), ("w0", C.c_double), ("wa", C.c_double),
                ("table", C.POINTER(C.c_double)), ("rows", C.c_int64), ("order", C.c_int32), ("index", C.c_int64)]


@pytest.fixture(scope="module")
def host():
    p = ROOT / "steps_amd" / "lib" / "libsteps_host.so"
    if not p.exists():
        from steps_amd import build
        build.build_host()
    lib = C.CDLL(str(p))
    E = C.POINTER(Expansion)
    lib.steps_hubble.restype = C.c_double
    lib.steps_hubble.argtypes = [E, C.c_double]
    lib.steps_friedmann_step.restype = C.c_double
    lib.steps_friedmann_step.argtypes = [E, C.c_double, C.c_double, C.c_double]
    lib.steps_friedmann_start.restype = C.c_double
    lib.steps_friedmann_start.argtypes = [E, C.c_double, C.c_double]
    lib.steps_read_expansion_history.restype = C.c_int64
    lib.steps_read_expansion_history.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.c_int64]
    return lib


def model(param=0, H0=67.66 / UNIT_V, Om=0.3111, Or=0.0, OL=0.6889, w0=-1.0, wa=0.0):
    e = Expansion()
    e.cosmoparam, e.H0, e.Omega_m, e.Omega_r, e.Omega_lambda = param, H0, Om, Or, OL
    e.Omega_k = 1.0 - Om - OL - Or
    e.w0, e.wa = w0, wa
    return e


def run(host, e, a, h, n, T0=0.0):
    T = T0
    for _ in range(n):
        T += h
        a = host.steps_friedmann_step(C.byref(e), a, h, T)
    return a


def test_matter_only_closed_forms(host):
    e = model(Om=1.0, OL=0.0)
    H0, a0 = e.H0, 0.05
    want = (a0 ** 1.5 + 1.5 * H0 * 0.04) ** (2.0 / 3.0)
    # fourth order: halving the step divides the error by ~16
    err = [abs(run(host, e, a0, 0.04 / n, n) - want) for n in (200, 400)]
    assert err[0] < 1e-9 * want and 14.0 < err[0] / err[1] < 18.0, err
    # the start time stops when two estimates agree to 1e-10 (absolute): ~1e-8 here
    t = host.steps_friedmann_start(C.byref(e), a0, 1e-6)
    assert t == pytest.approx(2.0 / (3.0 * H0) * a0 ** 1.5, rel=5e-8)
    assert host.steps_hubble(C.byref(e), 0.25) == pytest.approx(H0 * 0.25 ** -1.5, rel=1e-15)


def test_dark_energy_w_minus_third_grows_linearly(host):
    # Omega_de = 1, w0 = -1/3: H = H0 / a, da/dt = H0 exactly
    e = model(param=1, Om=0.0, OL=1.0, w0=-1.0 / 3.0)
    a = run(host, e, 0.1, 2e-3, 500)
    assert a == pytest.approx(0.1 + 500 * 2e-3 * e.H0, rel=1e-13)
    # the exponent's sign: w0 = -2/3 gives H = H0 a^(-1/2), not a^(-5/2)
    e2 = model(param=1, Om=0.0, OL=1.0, w0=-2.0 / 3.0)
    assert host.steps_hubble(C.byref(e2), 0.3) == pytest.approx(e2.H0 * 0.3 ** -0.5, rel=1e-14)


def test_wcdm_and_cpl_reduce_to_lcdm_bitwise(host):
    lcdm = model(0)
    wcdm = model(1, w0=-1.0)
    cpl0 = model(2, w0=-0.9, wa=0.0)
    wcdm9 = model(1, w0=-0.9)
    rng = np.random.default_rng(3)
    for a, h in zip(rng.uniform(0.02, 1.0, 50), rng.uniform(1e-5, 1e-2, 50)):
        assert host.steps_friedmann_step(C.byref(wcdm), a, h, 0.0) == host.steps_friedmann_step(C.byref(lcdm), a, h, 0.0)
        assert host.steps_hubble(C.byref(wcdm), a) == host.steps_hubble(C.byref(lcdm), a)
        assert host.steps_friedmann_step(C.byref(cpl0), a, h, 0.0) == host.steps_friedmann_step(C.byref(wcdm9), a, h, 0.0)


def test_cpl_density_against_the_equation_of_state(host):
    w0, wa = -0.85, 0.4
    e = model(2, w0=w0, wa=wa, Om=0.3, OL=0.7)
    for a in (0.05, 0.2, 0.5, 0.9, 1.3):
        I, _ = integrate.quad(lambda x: (1.0 + w0 + wa * (1.0 - x)) / x, a, 1.0, epsabs=0, epsrel=1e-13)
        want = e.H0 * math.sqrt(e.Omega_m * a ** -3 + e.Omega_lambda * math.exp(3.0 * I) + e.Omega_k * a ** -2)
        assert host.steps_hubble(C.byref(e), a) == pytest.approx(want, rel=1e-12)


def test_lcdm_step_matches_oracle_and_start_time_matches_quadrature(host, oracle):
    e = model(0, Or=8.5e-5, OL=0.6889 - 8.5e-5)
    c = oracle.OracleCosmo(e.H0, e.Omega_m, e.Omega_lambda, e.Omega_r, e.Omega_k)
    lib = oracle.lib()
    lib.oracle_friedmann_step.restype = C.c_double
    lib.oracle_friedmann_step.argtypes = [C.POINTER(oracle.OracleCosmo), C.c_double, C.c_double]
    for a, h in ((0.03125, 1e-6), (0.5, 1e-3), (0.9, 2e-2)):
        assert host.steps_friedmann_step(C.byref(e), a, h, 0.0) == lib.oracle_friedmann_step(C.byref(c), a, h)
    a0 = 0.03125
    t = host.steps_friedmann_start(C.byref(e), a0, 0.00002 / UNIT_T * 0.05)
    want, _ = integrate.quad(lambda a: 1.0 / (a * host.steps_hubble(C.byref(e), a)), 0.0, a0, epsabs=0,
                             epsrel=1e-13, limit=200)
    # the algorithm's own accuracy with radiation in the last steps (7.8e-8);
    # a missed power-law extrapolation to a = 0 would be off by ~1e-5
    assert t == pytest.approx(want, rel=2e-7)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_tabulated_history_interpolates_polynomials(host, order):
    # rows (t, a(t), H(a)) with a and H polynomials of the interpolation's degree
    t = np.linspace(0.01, 1.0, 40)
    pa = [0.3, 0.9, -0.2, 0.05][: order + 1]
    ph = [2.0, -0.7, 0.3, 0.1][: order + 1]
    a = np.polyval(pa[::-1], t)
    H = np.polyval(ph[::-1], a)
    tab = np.ascontiguousarray(np.stack([t, a, H], axis=1).reshape(-1))
    e = model(-1)
    e.table = tab.ctypes.data_as(C.POINTER(C.c_double))
    e.rows = len(t)
    e.order = order
    # the start time: the row before the first one whose a exceeds a_start (at least row 1)
    a_start = 0.5 * (a[5] + a[6])
    T = host.steps_friedmann_start(C.byref(e), a_start, 1e-3)
    assert T == t[5] and e.index == 6
    # steps of a quarter row: a(T) and H(a) come back off the nodes
    h = 0.25 * (t[1] - t[0])
    aa = a_start
    for k in range(100):
        T += h
        aa = host.steps_friedmann_step(C.byref(e), aa, h, T)
        assert aa == pytest.approx(np.polyval(pa[::-1], T), rel=1e-12, abs=1e-13)
        assert host.steps_hubble(C.byref(e), aa) == pytest.approx(np.polyval(ph[::-1], aa), rel=1e-11, abs=1e-13)


def test_read_expansion_history_units(host, tmp_path):
    f = tmp_path / "hist.dat"
    f.write_text("0.1 0.01 5000.0\n1.0 0.1 300.0\n13.8 1.0 67.66\n")
    assert host.steps_read_expansion_history(str(f).encode(), None, 0) == 3
    out = np.zeros(9)
    assert host.steps_read_expansion_history(str(f).encode(), out.ctypes.data_as(C.POINTER(C.c_double)), 3) == 3
    want = np.array([[0.1 / UNIT_T, 0.01, 5000.0 / UNIT_V], [1.0 / UNIT_T, 0.1, 300.0 / UNIT_V],
                     [13.8 / UNIT_T, 1.0, 67.66 / UNIT_V]]).reshape(-1)
    assert np.array_equal(out, want)
