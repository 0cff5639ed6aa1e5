"""Pins the oracle's R^3 direct sum (forces.cc:510-577 / forces_cuda.cu:522-563).

Checks that do not restate the oracle: a long-double brute force (different
precision, branch formulation in u = r/beta), Newton's third law, the shell
theorem, the two-body closed form, and the exact external terms.
"""
import math

import numpy as np
import pytest

from conftest import rel_err, report
from steps_amd import workloads as wl


def small_cloud(N, seed=3, close=True):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(N, 3)) * 10.0
    if close:  # plant close pairs inside the softening radius
        k = N // 5
        x[1:2 * k:2] = x[0:2 * k - 1:2] + rng.normal(size=(k, 3)) * 0.05
    M = rng.uniform(0.5, 2.0, N)
    h = 0.1 * np.cbrt(M)
    return x, M, h


def test_r3_matches_longdouble_bruteforce(oracle):
    x, M, h = small_cloud(400)
    p = oracle.params(0)
    F = oracle.forces(p, x, M, h)
    Fl = oracle.forces_longdouble(p, x, M, h).astype(np.float64)
    e = rel_err(F, Fl)
    report("oracle R3 vs long double", e)
    assert np.median(e) < 1e-14
    assert np.max(e) < 1e-11


def test_r3_compactified_small_matches_longdouble(oracle):
    w = wl.r3_compactified(N=1024, seed=11)
    p = oracle.params(0, cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere)
    F = oracle.forces(p, w.x, w.M, w.soft)
    Fl = oracle.forces_longdouble(p, w.x, w.M, w.soft).astype(np.float64)
    assert np.median(rel_err(F, Fl)) < 1e-14


def test_newton_third_law(oracle):
    x, M, h = small_cloud(600, seed=7)
    p = oracle.params(0)
    F = oracle.forces(p, x, M, h)
    mom = (M[:, None] * F).sum(axis=0)
    scale = (M[:, None] * np.abs(F)).sum(axis=0)
    assert np.all(np.abs(mom) < 1e-13 * scale)


def test_two_body_closed_form(oracle):
    x = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0]])
    M = np.array([2.0, 5.0])
    h = np.array([0.1, 0.2])
    F = oracle.forces(oracle.params(0), x, M, h)
    d = np.array([3.0, 4.0, 0.0])
    assert np.allclose(F[0], 5.0 * d / 125.0, rtol=1e-15, atol=0)
    assert np.allclose(F[1], -2.0 * d / 125.0, rtol=1e-15, atol=0)
    # inside the softening radius: w = u-polynomial / beta^3
    x2 = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0]])
    M2 = np.array([1.0, 1.0])
    h2 = np.array([0.1, 0.1])  # beta = 0.2, u = 0.5
    F2 = oracle.forces(oracle.params(0), x2, M2, h2)
    b = 0.2
    w = (32 * 0.125 - 38.4 * 0.25 + 32.0 / 3.0) / b ** 3
    assert F2[0, 0] == pytest.approx(w * 0.1, rel=1e-13)


def test_shell_theorem(oracle):
    # Fibonacci sphere of unit radius, test particle of zero mass
    n = 4000
    k = np.arange(n) + 0.5
    th = np.arccos(1 - 2 * k / n)
    ph = math.pi * (1 + 5 ** 0.5) * k
    shell = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
    for r_test, expect_inside in [(3.0, False), (5.0, False), (0.2, True)]:
        x = np.vstack([shell, [[r_test, 0.0, 0.0]]])
        M = np.concatenate([np.full(n, 1.0 / n), [0.0]])
        h = np.full(n + 1, 1e-6)
        F = oracle.forces(oracle.params(0), x, M, h, id_min=n, id_max=n)[0]
        if expect_inside:
            assert np.linalg.norm(F) < 2e-3
        else:
            assert F[0] == pytest.approx(-1.0 / r_test ** 2, rel=2e-4)
            assert abs(F[1]) < 1e-4 / r_test ** 2 and abs(F[2]) < 1e-4 / r_test ** 2


def test_external_terms_exact(oracle):
    x, M, h = small_cloud(200, seed=9)
    base = oracle.forces(oracle.params(0), x, M, h)
    misph = 0.7
    Fc = oracle.forces(oracle.params(0, cosmology=1, comoving=1, mass_in_unit_sphere=misph), x, M, h)
    assert np.allclose(Fc - base, misph * x, rtol=1e-12, atol=1e-12 * np.abs(base).max())
    H0, OL = 3.26, 0.69
    Fn = oracle.forces(oracle.params(0, cosmology=1, comoving=0, H0=H0, Omega_lambda=OL), x, M, h)
    assert np.allclose(Fn - base, H0 * H0 * OL * x, rtol=1e-12, atol=1e-12 * np.abs(base).max())


def test_row_ranges_and_variants_agree(oracle):
    w = wl.r3_compactified(N=2048, seed=4)
    p = oracle.params(0, cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere)
    full = oracle.forces(p, w.x, w.M, w.soft)
    sub = oracle.forces(p, w.x, w.M, w.soft, id_min=100, id_max=357)
    assert np.array_equal(sub, full[100:358])
    rows = np.arange(0, 2048, 37, dtype=np.int32)
    for kind in ("oracle", "faithful", "fast"):
        Fr = oracle.forces_rows(p, w.x, w.M, w.soft, rows, kind=kind)
        assert np.median(rel_err(Fr, full[rows])) < 1e-13, kind


def test_softening_setup_closed_form(oracle):
    M = np.array([2.0, 16.0, 54.0, 2.0])
    h, mmin, rp = oracle.softening_lengths(M, 0.3)
    assert mmin == 2.0
    assert rp == pytest.approx(2.0 / (4.0 * math.pi * 0.3 ** 3 / 3.0), rel=1e-15)
    assert np.allclose(h, 0.3 * np.cbrt(M / 2.0), rtol=1e-15)


def test_fp32_stand_in_two_body_and_branches(oracle):
    """The all-float stand-in of the reference's single-precision build: a
    two-body force is the FP64 one to float rounding, in every softening
    branch, and other topologies are refused."""
    rows = np.array([0, 1], dtype=np.int32)
    for sep in (5.0, 0.15, 0.05):  # beta = 0.2: r >= beta, beta/2 < r < beta, r <= beta/2
        x = np.array([[0.25, -0.5, 1.0], [0.25 + 0.6 * sep, -0.5 + 0.8 * sep, 1.0]])
        M = np.array([1.5, 0.75])
        h = np.array([0.1, 0.1])
        F64 = oracle.forces(oracle.params(0), x, M, h)
        F32 = oracle.forces_rows(oracle.params(0), x, M, h, rows, kind="fp32")
        assert np.all(rel_err(F32, F64) <= 2e-6)
        assert not np.array_equal(F32, F64)  # the float path really ran
    with pytest.raises(AssertionError):
        oracle.forces_rows(oracle.params(1, is_periodic=1, L=10.0), x, M, h, rows, kind="fp32")


def test_fp32_stand_in_error_grows_with_n(oracle):
    """Uncompensated float accumulation: against the FP64 oracle the error is
    bounded by float rounding times the number of terms and grows with N."""
    med = []
    for N in (512, 16384):
        w = wl.r3_compactified(N=N, seed=1000)
        p = oracle.params(0, cosmology=0)
        rows = np.linspace(0, N - 1, 48).astype(np.int32)
        e = rel_err(oracle.forces_rows(p, w.x, w.M, w.soft, rows, kind="fp32"),
                    oracle.forces_rows(p, w.x, w.M, w.soft, rows))
        report(f"fp32 stand-in vs fp64 N={N}", e)
        assert np.median(e) <= 6e-8 * N
        med.append(np.median(e))
    assert med[1] > med[0]
