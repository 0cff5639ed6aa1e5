"""Pins the oracle's T^3 Ewald table and interpolation (ewald_space.cc:118-523,
forces_cuda.cu:93-167) independently of its own code.

Independent checks: the image index sets against a brute-force enumeration; an
Ewald sum written here in numpy with a different splitting parameter and wider
cutoffs; the table's O_h symmetry and antisymmetry; node reproduction of the
Catmull-Rom / trilinear interpolators; and the full T^3 pair force against the
numpy Ewald force at random separations (interpolation error only).
"""
import math

import numpy as np
import pytest
from scipy.special import erfc

L_BOX = 100.0 / 0.6766


def ewald_force_numpy(r, L, alpha, nreal=5, nrec=12):
    """Periodic force on a unit mass at r from a unit mass at 0 (with the
    neutralising background), standard Ewald split, numpy."""
    r = np.asarray(r, dtype=np.float64)
    rng = np.arange(-nreal, nreal + 1)
    n = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
    R = r[None, :] + n * L
    R2 = (R * R).sum(1)
    ok = R2 > 0
    R, R2 = R[ok], R2[ok]
    Rn = np.sqrt(R2)
    coeff = (erfc(alpha * Rn) + 2 * alpha / math.sqrt(math.pi) * Rn * np.exp(-alpha ** 2 * R2)) / Rn ** 3
    F = -(R * coeff[:, None]).sum(0)
    kr = np.arange(-nrec, nrec + 1)
    kn = np.stack(np.meshgrid(kr, kr, kr, indexing="ij"), -1).reshape(-1, 3)
    kn = kn[(kn * kn).sum(1) > 0]
    kv = kn * (2 * math.pi / L)
    k2 = (kv * kv).sum(1)
    c = 4 * math.pi / L ** 3 * np.exp(-k2 / (4 * alpha ** 2)) / k2 * np.sin(kv @ r)
    F += -(kv * c[:, None]).sum(0)
    return F


def test_ewald_space_index_sets(oracle):
    lib = oracle.lib()
    for R in [3.6, 4.6, 10.0, 12.0, 14.0]:
        m = int(math.ceil(R))
        rr = np.arange(-m, m + 1)
        n = np.stack(np.meshgrid(rr, rr, rr, indexing="ij"), -1).reshape(-1, 3)
        brute = int(((np.abs(n) < R).all(1) & ((n * n).sum(1) < R * R)).sum())
        cnt = lib.oracle_ewald_space_count(R)
        assert cnt == brute, (R, cnt, brute)
        idx = np.zeros(4 * cnt, dtype=np.int32)
        lib.oracle_ewald_space(R, oracle.ip(idx))
        idx = idx.reshape(-1, 4)
        assert len({tuple(v) for v in idx[:, :3]}) == cnt           # no duplicates
        assert np.array_equal(idx[:, 3], (idx[:, :3] ** 2).sum(1))  # stored |n|^2


@pytest.fixture(scope="module")
def t3_table(oracle):
    return oracle.t3_table(2, L_BOX)


def test_t3_table_symmetries(oracle, t3_table):
    ng = 63
    T = t3_table.reshape(ng, ng, ng, 3)
    c = ng // 2
    assert np.all(T[c, c, c] == 0.0)                       # D(0) = 0
    # inversion: D(-r) = -D(r)
    assert np.allclose(T[::-1, ::-1, ::-1], -T, atol=1e-14 * np.abs(T).max())
    # mirror x -> -x flips the x component only
    Tm = T[::-1, :, :].copy()
    Tm[..., 0] *= -1
    assert np.allclose(Tm, T, atol=1e-14 * np.abs(T).max())
    # axis swap x <-> y swaps the components
    Ts = np.transpose(T, (1, 0, 2, 3))[..., [1, 0, 2]]
    assert np.allclose(Ts, T, atol=1e-14 * np.abs(T).max())


def test_t3_table_matches_independent_ewald(oracle, t3_table):
    ng = 63
    T = t3_table.reshape(ng, ng, ng, 3)
    gs = L_BOX / ng
    rng = np.random.default_rng(2)
    scale = np.abs(T).max()
    for _ in range(25):
        i, j, k = rng.integers(0, ng, 3)
        r = (np.array([i, j, k]) + 0.5) * gs - L_BOX / 2
        if not np.any(r):
            continue
        F = ewald_force_numpy(r, L_BOX, alpha=2.6 / L_BOX)
        D = F - (-r / np.linalg.norm(r) ** 3)
        assert np.allclose(T[i, j, k], D, atol=2e-10 * scale, rtol=0), (i, j, k, T[i, j, k], D)


def test_interpolators_reproduce_nodes(oracle, t3_table):
    ng = 63
    T = t3_table.reshape(ng, ng, ng, 3)
    gs = L_BOX / ng
    rng = np.random.default_rng(4)
    for _ in range(20):
        i, j, k = rng.integers(3, ng - 3, 3)
        r = (np.array([i, j, k]) + 0.5) * gs - L_BOX / 2
        for order in (4, 2):
            D = oracle.t3_interp(t3_table, L_BOX, r, order)
            assert np.allclose(D, T[i, j, k], rtol=1e-12, atol=1e-13 * np.abs(T).max())


def test_catmull_rom_weights_partition_of_unity(oracle):
    # interpolating a constant table returns the constant anywhere
    T = np.full(7 ** 3 * 3, 2.5)
    for d in np.random.default_rng(1).uniform(-L_BOX / 2, L_BOX / 2, (20, 3)):
        assert np.allclose(oracle.t3_interp(T, L_BOX, d, 4), 2.5, rtol=1e-14)


def test_t3_pair_force_against_independent_ewald(oracle, t3_table):
    """Two unit masses at random separations (|d_c| < 0.4 L): the oracle's T3
    force (nearest image + interpolated correction) equals the independent
    periodic force up to the 63^3 interpolation error."""
    rng = np.random.default_rng(6)
    p = oracle.params(1, is_periodic=2, L=L_BOX, t3_table=t3_table)
    worst = 0.0
    for _ in range(30):
        d = rng.uniform(-0.4, 0.4, 3) * L_BOX
        x = np.array([[50.0, 60.0, 70.0], (np.array([50.0, 60.0, 70.0]) + d) % L_BOX])
        M = np.array([1.0, 1.0])
        h = np.array([1e-3, 1e-3])
        F = oracle.forces(p, x, M, h)[0]
        Fp = ewald_force_numpy(-d, L_BOX, alpha=2.6 / L_BOX)  # particle 0 sits at -d from its source
        Fn = d / np.linalg.norm(d) ** 3
        worst = max(worst, np.linalg.norm(F - Fp) / np.linalg.norm(Fn))
    assert worst < 2e-5, worst


def test_perfect_lattice_has_zero_force(oracle, t3_table):
    n = 3
    g = (np.arange(n) + 0.5) * L_BOX / n
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    M = np.ones(n ** 3)
    h = np.full(n ** 3, 0.01)
    F = oracle.forces(oracle.params(1, is_periodic=2, L=L_BOX, t3_table=t3_table), X, M, h)
    assert np.abs(F).max() < 1e-13 * (1.0 / (L_BOX / n) ** 2)
