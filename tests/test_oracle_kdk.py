"""Pins the oracle's integrator pieces (step.cc:35-312, friedmann_solver.cc:100-164)
against closed forms: a circular two-body orbit and the analytic flat-LCDM a(t)."""
import ctypes as C
import math

import numpy as np
import pytest


def test_two_body_circular_orbit_returns(oracle):
    lib = oracle.lib()
    m, d = 1.0, 2.0
    v = math.sqrt(m / (2 * d))                      # G = 1
    period = 2 * math.pi * (d / 2) / v
    x = np.array([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    vel = np.array([[0.0, -v, 0.0], [0.0, v, 0.0]])
    M = np.array([m, m])
    h = np.array([1e-4, 1e-4])
    p = oracle.params(0)
    cosmo = oracle.OracleCosmo(0.0, 0.0, 0.0, 0.0, 0.0)
    nsteps = 4000
    st = oracle.OracleStepState(1.0, 0.0, period / nsteps, 1.0, 1.0)
    F = oracle.forces(p, x, M, h).reshape(-1).copy()
    xf, vf = x.reshape(-1).copy(), vel.reshape(-1).copy()
    e0 = 0.5 * m * (vf ** 2).sum() - m * m / d
    for _ in range(nsteps):
        lib.oracle_kdk_step(C.byref(p), C.byref(cosmo), C.byref(st), 2, oracle.dp(xf), oracle.dp(vf),
                            oracle.dp(F), oracle.dp(M), oracle.dp(h))
    assert np.allclose(xf.reshape(2, 3), x, atol=2e-4)
    dd = np.linalg.norm(xf[3:] - xf[:3])
    e1 = 0.5 * m * (vf ** 2).sum() - m * m / dd
    assert abs(e1 - e0) < 1e-6 * abs(e0)


def test_errmax_formula(oracle):
    lib = oracle.lib()
    p = oracle.params(0)
    N = 5
    rng = np.random.default_rng(1)
    x = rng.normal(size=3 * N)
    v = rng.normal(size=3 * N)
    F = rng.normal(size=3 * N)
    M = rng.uniform(1, 3, N)
    st = oracle.OracleStepState(0.5, 0.7, 0.01, 2.0, 1.0)
    e = lib.oracle_initial_errmax(C.byref(p), C.byref(st), N, oracle.dp(x), oracle.dp(v), oracle.dp(F),
                                  oracle.dp(M))
    A = F.reshape(N, 3) * 0.5 ** -3 - 2 * 0.7 * v.reshape(N, 3)
    expect = (np.linalg.norm(A, axis=1) / np.cbrt(M * 3 / (4 * math.pi * 2.0))).max()
    assert e == pytest.approx(expect, rel=1e-14)


def test_friedmann_rk4_matches_flat_lcdm(oracle):
    lib = oracle.lib()
    H0 = 67.66 / 20.738652969925447
    Om, OL = 0.3111, 0.6889
    c = oracle.OracleCosmo(H0, Om, OL, 0.0, 1.0 - Om - OL)
    a_exact = lambda t: (Om / OL) ** (1 / 3) * np.sinh(1.5 * math.sqrt(OL) * H0 * t) ** (2 / 3)
    t_of_a = lambda a: 2 / (3 * math.sqrt(OL) * H0) * math.asinh(math.sqrt(OL / Om * a ** 3))
    a = 1.0 / 32.0
    t = t_of_a(a)
    h = 1e-4
    for _ in range(2000):
        a = lib.oracle_friedmann_step(C.byref(c), a, h)
        t += h
    assert a == pytest.approx(a_exact(t), rel=1e-10)
    assert lib.oracle_hubble(C.byref(c), 0.5) == pytest.approx(H0 * math.sqrt(Om * 8 + OL), rel=1e-14)
