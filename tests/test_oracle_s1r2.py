"""Pins the oracle's S^1 x R^2 path (ewald_space.cc:539-1048,
forces_cuda.cu:177-455, :652-864, utils.cc:162-228) independently.

* The Tornberg 1P Ewald table against a direct real-space sum over 2*10^5
  z-images written here in numpy, to 1e-6 of max|D| (the finer pins, an
  exact-erfc numpy Ewald at 1e-11 max|D| and the A&S fit's bound, are in
  test_oracle_pins.py).
* The reference's real-space table mode (PERIODIC_Z_RSPACELOOKUP) against the
  same numpy sum truncated identically (exact).
* NGP/CIC reproduce table nodes; TSC reproduces linear data.
* The radial (finite-cylinder) table against adaptive quadrature; its
  correction of order 1-3 (forces.cc:279-316) against polynomial data.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate

LZ, RSIM = 200.0, 100.0


def direct_image_force(rho, z, Lz, nmax):
    n = np.arange(-nmax, nmax + 1, dtype=np.float64)
    zz = z + n * Lz
    r2 = rho * rho + zz * zz
    ok = r2 > 1e-18
    inv3 = 1.0 / (r2[ok] * np.sqrt(r2[ok]))
    return np.array([-(rho * inv3).sum(), -(zz[ok] * inv3).sum()])


def nearest_newton(rho, z):
    r2 = rho * rho + z * z
    if r2 == 0:
        return np.zeros(2)
    return np.array([-rho, -z]) / r2 ** 1.5


@pytest.fixture(scope="module")
def table(oracle):
    T, shape = oracle.s1r2_table(2, LZ, RSIM)
    return T, shape


def test_table_shape_follows_main_cc(oracle, table):
    T, (nrho, nz) = table
    assert nz == 128 and nrho == int(math.floor(128 * 2.25 * RSIM / LZ))
    alpha, rel, rec, nz2, nrho2 = oracle.s1r2_params(2, LZ, RSIM)
    assert (rel, rec) == (4, 10) and alpha == pytest.approx(0.787875 / LZ, rel=1e-15)


def test_ewald_table_matches_direct_image_sum(table):
    T, (nrho, nz) = table
    T = T.reshape(nrho, nz, 2)
    rho_max = 2.25 * RSIM
    drho, dz = rho_max / (nrho - 1), LZ / nz
    scale = np.abs(T).max()
    rng = np.random.default_rng(0)
    for _ in range(30):
        ir, iz = rng.integers(0, nrho), rng.integers(0, nz)
        rho, z = ir * drho, (iz + 0.5) * dz - 0.5 * LZ
        D = direct_image_force(rho, z, LZ, 100000) - nearest_newton(rho, z)
        assert np.abs(T[ir, iz] - D).max() < 1e-6 * scale, (ir, iz, T[ir, iz], D)


def test_rspace_table_equals_truncated_image_sum(oracle):
    Lz, Rs = 50.0, 20.0
    nrho, nz = 40, 16
    T = np.zeros(nrho * nz * 2)
    oracle.lib().oracle_s1r2_table(nrho, nz, 2.25 * Rs, Lz, 0.0, 300, 0, 1, oracle.dp(T))
    T = T.reshape(nrho, nz, 2)
    drho, dz = 2.25 * Rs / (nrho - 1), Lz / nz
    for ir in range(0, nrho, 7):
        for iz in range(0, nz, 5):
            rho, z = ir * drho, (iz + 0.5) * dz - 0.5 * Lz
            D = direct_image_force(rho, z, Lz, 300) - nearest_newton(rho, z)
            assert np.allclose(T[ir, iz], D, rtol=1e-11, atol=1e-14 * np.abs(D).max())


@pytest.mark.parametrize("order", [0, 2])
def test_ngp_cic_reproduce_nodes(oracle, table, order):
    T, (nrho, nz) = table
    Tr = T.reshape(nrho, nz, 2)
    rho_max = 2.25 * RSIM
    drho, dz = rho_max / (nrho - 1), LZ / nz
    rng = np.random.default_rng(3)
    for _ in range(20):
        ir, iz = rng.integers(1, nrho - 2), rng.integers(1, nz - 1)
        rho, z = ir * drho, (iz + 0.5) * dz - 0.5 * LZ
        phi = rng.uniform(0, 2 * np.pi)
        d = np.array([rho * np.cos(phi), rho * np.sin(phi), z])
        D = oracle.s1r2_interp(T, (nrho, nz), rho_max, LZ, d, order)
        assert D[2] == pytest.approx(Tr[ir, iz, 1], rel=1e-9, abs=1e-12 * np.abs(Tr).max())
        assert math.hypot(D[0], D[1]) == pytest.approx(abs(Tr[ir, iz, 0]), rel=1e-9)


def test_tsc_reproduces_linear_data(oracle):
    nrho, nz, Lz, rho_max = 30, 24, 10.0, 9.0
    drho, dz = rho_max / (nrho - 1), Lz / nz
    ir, iz = np.meshgrid(np.arange(nrho), np.arange(nz), indexing="ij")
    T = np.stack([0.3 + 2.0 * ir * drho, 1.5 - 0.7 * ((iz + 0.5) * dz - 0.5 * Lz)], -1).reshape(-1)
    rng = np.random.default_rng(2)
    for _ in range(30):
        rho = rng.uniform(2 * drho, rho_max - 2 * drho)
        z = rng.uniform(-0.5 * Lz + 2 * dz, 0.5 * Lz - 2 * dz)
        D = oracle.s1r2_interp(T, (nrho, nz), rho_max, Lz, [rho, 0.0, z], 4)
        assert D[0] == pytest.approx(0.3 + 2.0 * rho, rel=1e-12)
        assert D[2] == pytest.approx(1.5 - 0.7 * z, rel=1e-12)


def test_lookup_pair_force_is_periodic_force(oracle, table):
    T, shape = table
    p = oracle.params(2, is_periodic=2, L=LZ, Rsim=RSIM, s1r2_lookup=1, interp_order=4, s1r2_table=T,
                      s1r2_shape=shape)
    rng = np.random.default_rng(8)
    for _ in range(20):
        x0 = np.array([rng.uniform(-50, 50), rng.uniform(-50, 50), rng.uniform(0, LZ)])
        d = np.array([rng.uniform(-60, 60), rng.uniform(-60, 60), rng.uniform(-0.45, 0.45) * LZ])
        x = np.array([x0, x0 + d])
        x[1, 2] %= LZ
        F = oracle.forces(p, x, np.array([1.0, 1.0]), np.array([1e-3, 1e-3]))[0]
        rho = math.hypot(d[0], d[1])
        Fd = direct_image_force(rho, -d[2], LZ, 100000)   # force on 0 (at -d from the source)
        Fc = np.array([Fd[0] * -d[0] / rho, Fd[0] * -d[1] / rho, Fd[1]])
        Fn = np.linalg.norm(d) ** -2
        assert np.linalg.norm(F - Fc) < 1e-4 * Fn, (F, Fc)  # TSC smoothing error ~1e-5


def test_image_sum_force_matches_explicit_images(oracle):
    Lz = 30.0
    for is_periodic in (2, 3):
        emax = is_periodic + 1
        cut = (emax - 0.4) * Lz
        x = np.array([[1.0, 2.0, 3.0], [4.0, -1.5, 27.5]])
        M = np.array([1.3, 0.7])
        h = np.array([0.01, 0.01])
        for cut_incl in (0, 1):
            p = oracle.params(2, is_periodic=is_periodic, L=Lz, s1r2_lookup=0, cut_inclusive=cut_incl)
            F = oracle.forces(p, x, M, h)
            d = x[1] - x[0]
            expect = np.zeros(3)
            for m in range(-emax, emax + 1):
                dz = d[2] + m * Lz
                if abs(dz) < cut:
                    v = np.array([d[0], d[1], dz])
                    expect += M[1] * v / np.linalg.norm(v) ** 3
            assert np.allclose(F[0], expect, rtol=1e-13)


def radial_integrand(R, Lz, a, Y):
    s = math.sqrt(max(R * R - Y * Y, 0.0))
    f1 = -2 * math.log((math.sqrt(Lz ** 2 + R ** 2 + a ** 2 + 2 * a * s) + Lz) /
                       (math.sqrt(Lz ** 2 + R ** 2 + a ** 2 - 2 * a * s) + Lz))
    f2 = math.log((R * R + a * a + 2 * a * s) / (R * R + a * a - 2 * a * s))
    return f1 + f2


def test_radial_table_against_quadrature(oracle):
    R, Lz, size, acc = 500.0, 50.0, 200, 7500
    T = oracle.radial_table(R, Lz, size, acc)
    assert np.all(np.isfinite(T))
    for i in (5, 40, 120, 180):
        a = R / (size - 1) * i
        val, _ = integrate.quad(lambda Y: radial_integrand(R, Lz, a, Y), 0.0, R, limit=400, epsrel=1e-12,
                                points=[a] if a < R else None)
        expect = 2 * val / (2 * math.pi * a)
        assert T[i] == pytest.approx(expect, rel=2e-3), (i, T[i], expect)
    # node 0 is the linear extrapolation of nodes 1 and 2 (utils.cc:221-227)
    X1, X2 = 2 * R / size, 4 * R / size
    A = (T[2] - T[1]) / (X2 - X1)
    assert T[0] == pytest.approx(T[1] - A * X1, rel=1e-12)


def test_radial_correction_interpolation(oracle):
    T = np.linspace(0.5, 1.5, 11)
    lib = oracle.lib()
    # nodes assumed at R/size * i (forces_cuda.cu:58-69)
    R = 10.0
    for r in [0.0, 3.3, 7.77, 9.99, 12.0]:
        c = lib.oracle_cylindrical_force_correction(r, R, oracle.dp(T), T.size)
        i = int(math.floor(r / R * (T.size - 1)))
        step = R / T.size
        if i < T.size - 1:
            expect = T[i] + (T[i + 1] - T[i]) / step * (r - step * i)
        else:
            expect = T[-1]
        assert c == pytest.approx(expect, rel=1e-13)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_radial_correction_orders_reproduce_polynomials(oracle, order):
    """forces.cc:279-316: the order-k correction is Lagrange interpolation on
    the nodes R/size * i, so it reproduces polynomials of degree k exactly;
    past its last full stencil it shifts the stencil down (order 2 at
    i = size-2, order 3 at i = size-3) or returns the last entry."""
    lib = oracle.lib()
    fn = lib.oracle_cylindrical_force_correction_order
    fn.restype = C.c_double
    fn.argtypes = [C.c_double, C.c_double, oracle._DP, C.c_int, C.c_int]
    R, size = 50.0, 40
    step = R / size
    coef = [0.7, -0.03, 0.004, -1.1e-4][: order + 1]
    poly = lambda r: sum(c * r ** k for k, c in enumerate(coef))
    T = np.array([poly(step * i) for i in range(size)])
    rng = np.random.default_rng(order)
    last_full = {1: size - 2, 2: size - 2, 3: size - 3}[order]   # highest i with a stencil
    for r in rng.uniform(0.0, R, 400):
        i = int(math.floor(r / R * (size - 1)))
        c = fn(r, R, oracle.dp(T), size, order)
        if i <= last_full:
            assert c == pytest.approx(poly(r), rel=1e-11, abs=1e-13), (r, i)
        else:
            assert c == T[-1], (r, i)
    # the order-1 function is the one the force paths use
    for r in (0.0, 3.3, 49.0, 60.0):
        assert fn(r, R, oracle.dp(T), size, 1) == lib.oracle_cylindrical_force_correction(r, R, oracle.dp(T), size)
