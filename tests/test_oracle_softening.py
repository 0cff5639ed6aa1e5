"""Pins the oracle's softening law (forces.cc:52-87) independently.

The reference's piecewise polynomial is the force of the Gadget-2 cubic-spline
density kernel of radius beta: w(r) = M_enc(r) / r^3 with
W(u) = 8/(pi b^3) (1 - 6u^2 + 6u^3) for u <= 1/2 and 16/(pi b^3) (1-u)^3 for
1/2 < u <= 1.  Integrating that density numerically (scipy quad) pins every
coefficient; a dropped term or wrong power fails it.
"""
import math

import numpy as np
import pytest
from scipy import integrate


def spline_density(r, b):
    u = r / b
    if u <= 0.5:
        return 8.0 / (math.pi * b ** 3) * (1.0 - 6.0 * u * u + 6.0 * u ** 3)
    if u <= 1.0:
        return 16.0 / (math.pi * b ** 3) * (1.0 - u) ** 3
    return 0.0


def enclosed_mass(r, b):
    f = lambda s: 4.0 * math.pi * s * s * spline_density(s, b)
    if r <= 0.5 * b:
        m, _ = integrate.quad(f, 0.0, r, epsabs=0, epsrel=1e-13, limit=200)
    else:
        m1, _ = integrate.quad(f, 0.0, 0.5 * b, epsabs=0, epsrel=1e-13, limit=200)
        m2, _ = integrate.quad(f, 0.5 * b, min(r, b), epsabs=0, epsrel=1e-13, limit=200)
        m = m1 + m2
    return m


@pytest.mark.parametrize("beta", [0.37, 1.0, 12.5])
def test_softening_is_spline_enclosed_mass(oracle, beta):
    lib = oracle.lib()
    # total kernel mass is one
    assert enclosed_mass(beta, beta) == pytest.approx(1.0, rel=1e-12)
    for u in np.linspace(0.02, 0.999, 60):
        r = u * beta
        expect = enclosed_mass(r, beta) / r ** 3
        got = lib.oracle_force_softening(r, beta)
        assert got == pytest.approx(expect, rel=2e-12), (u, got, expect)


def test_softening_outer_and_continuity(oracle):
    lib = oracle.lib()
    beta = 0.8
    for r in [beta, 1.3 * beta, 10.0, 1e4]:
        assert lib.oracle_force_softening(r, beta) == pytest.approx(r ** -3, rel=1e-15)
    eps = 1e-9
    for edge in [0.5 * beta, beta]:
        lo = lib.oracle_force_softening(edge * (1 - eps), beta)
        hi = lib.oracle_force_softening(edge * (1 + eps), beta)
        assert lo == pytest.approx(hi, rel=1e-7)
    # value at r=0 is the finite central limit 32/(3 beta^3)
    assert lib.oracle_force_softening(0.0, beta) == pytest.approx(32.0 / (3.0 * beta ** 3), rel=1e-14)


def test_gpu_form_equals_cpu_form(oracle):
    lib = oracle.lib()
    rng = np.random.default_rng(5)
    for _ in range(2000):
        beta = 10 ** rng.uniform(-3, 2)
        r = beta * rng.uniform(0.0, 2.0)
        a = lib.oracle_force_softening(r, beta)
        b = lib.oracle_force_softening_gpu(r, beta)
        assert b == pytest.approx(a, rel=5e-13, abs=0)
