"""GPU parity of the R^3 force path against the CPU oracle (through the C ABI).

Tolerances: FP64 path median relative force error <= 1e-12 (summation order
differs from the oracle), p99 <= 1e-10; mixed path p99 <= 1e-5 against FP64.
"""
import numpy as np
import pytest

from conftest import rel_err, report
from steps_amd import workloads as wl

pytestmark = pytest.mark.gpu


def r3_descs(steps, oracle, w, cosmology=1, precision=0):
    kw = dict(cosmology=cosmology, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere)
    d, keep = steps.make_desc(steps.TOPO_R3, precision=precision, **kw)
    p = oracle.params(0, **kw)
    return d, keep, p


@pytest.mark.parametrize("cosmology", [1, 0])
def test_r3_fp64_compactified_16k(steps, oracle, cosmology):
    print(f"case r3 fp64 N=16384 seed=1000 cosmology={cosmology}", flush=True)
    w = wl.r3_compactified(N=16384, seed=1000)
    d, keep, p = r3_descs(steps, oracle, w, cosmology)
    F = steps.force_eval(d, w.x, w.M, w.soft)
    Fo = oracle.forces(p, w.x, w.M, w.soft)
    e = rel_err(F, Fo)
    report(f"R3 fp64 N=16384 cosmology={cosmology}", e)
    assert np.median(e) <= 1e-12
    assert np.percentile(e, 99) <= 1e-10


def test_r3_fp64_vs_longdouble_close_pairs(steps, oracle):
    rng = np.random.default_rng(3)
    N = 777
    x = rng.normal(size=(N, 3)) * 10.0
    x[1:300:2] = x[0:299:2] + rng.normal(size=(150, 3)) * 0.05   # inside softening
    x[5] = x[4]                                                    # coincident pair (r = 0)
    M = rng.uniform(0.5, 2.0, N)
    h = 0.1 * np.cbrt(M)
    d, keep = steps.make_desc(steps.TOPO_R3)
    F = steps.force_eval(d, x, M, h)
    Fl = oracle.forces_longdouble(oracle.params(0), x, M, h).astype(np.float64)
    e = rel_err(F, Fl)
    report("R3 fp64 vs long double", e)
    assert np.all(np.isfinite(F))
    assert np.median(e) <= 1e-13
    assert np.max(e) <= 1e-10


@pytest.mark.parametrize("N,id_min,id_max", [(1, 0, 0), (2, 0, 1), (255, 0, 254), (257, 3, 256),
                                              (1000, 511, 513), (5000, 0, 4999), (5000, 4321, 4321)])
def test_r3_row_ranges_and_odd_sizes(steps, oracle, N, id_min, id_max):
    w = wl.r3_compactified(N=N, seed=17)
    d, keep, p = r3_descs(steps, oracle, w)
    F = steps.force_eval(d, w.x, w.M, w.soft, id_min, id_max)
    Fo = oracle.forces(p, w.x, w.M, w.soft, id_min, id_max)
    assert F.shape == Fo.shape
    assert np.median(rel_err(F, Fo)) <= 1e-12


def test_r3_sim_matches_oneshot_and_is_deterministic(steps):
    w = wl.r3_compactified(N=6000, seed=23)
    d, keep = steps.make_desc(steps.TOPO_R3, cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere)
    F1 = steps.force_eval(d, w.x, w.M, w.soft)
    sim = steps.Sim(d, w.N)
    sim.set_state(w.x, w.M, w.soft, w.v)
    sim.forces()
    F2 = sim.get_forces()
    sim.forces()
    F3 = sim.get_forces()
    assert np.array_equal(F2, F3)          # same plan, fixed reduction order
    assert np.array_equal(F1, F2)
    pair_ms, total_ms = sim.last_kernel_ms()
    assert 0 < pair_ms <= total_ms
    sim.forces(100, 199)
    assert np.array_equal(sim.get_forces(), F1[100:200])
    x, v = sim.get_state()
    assert np.array_equal(x, w.x) and np.array_equal(v, w.v)
    sim.close()


def test_r3_mixed_precision_error(steps, oracle):
    w = wl.r3_compactified(N=16384, seed=1000)
    for cosmology in (1, 0):
        d64, k1, _ = r3_descs(steps, oracle, w, cosmology, precision=0)
        dmx, k2, _ = r3_descs(steps, oracle, w, cosmology, precision=1)
        F64 = steps.force_eval(d64, w.x, w.M, w.soft)
        Fmx = steps.force_eval(dmx, w.x, w.M, w.soft)
        e = rel_err(Fmx, F64)
        report(f"R3 mixed vs fp64 N=16384 cosmology={cosmology}", e)
        assert np.percentile(e, 99) <= 1e-5
    # the comparison point: the reference's all-float build (stand-in), same
    # rows, pair sum only; the mixed path must be the more accurate
    rows = np.arange(0, w.N, 8, dtype=np.int32)
    e32 = rel_err(oracle.forces_rows(oracle.params(0, cosmology=0), w.x, w.M, w.soft, rows, kind="fp32"),
                  F64[rows])
    emx = rel_err(Fmx[rows], F64[rows])
    report("R3 reference all-float stand-in vs fp64 N=16384", e32)
    assert np.median(emx) < np.median(e32) and np.percentile(emx, 99) < np.percentile(e32, 99)


def test_r3_full_size_2m(steps, oracle):
    """The N = 2^21 compactified R^3 configuration (seed 1003) at full size:
    64 sampled rows against the oracle; Newton's third law over all rows
    (external term off, sum_i m_i F_i = 0 up to rounding); the mixed path's
    pair-sum error against FP64; two evaluations bitwise equal."""
    print("case r3 N=2^21 full size", flush=True)
    w = wl.r3_compactified(N=2 ** 21, seed=1003)
    d, keep, p = r3_descs(steps, oracle, w, cosmology=0)
    sim = steps.Sim(d, w.N)
    sim.set_state(w.x, w.M, w.soft)
    sim.forces()
    F = sim.get_forces()
    sim.forces()
    assert np.array_equal(F, sim.get_forces())
    rows = np.linspace(0, w.N - 1, 64).astype(np.int32)
    e = rel_err(F[rows], oracle.forces_rows(p, w.x, w.M, w.soft, rows))
    report("R3 N=2^21 sampled rows", e)
    assert np.median(e) <= 1e-12 and e.max() <= 1e-10
    mF = w.M[:, None] * F
    net = np.abs(mF.sum(axis=0)).max() / np.abs(mF).sum(axis=0).max()
    print(f"[R3 N=2^21] |sum m F| / sum |m F| = {net:.3e}", flush=True)
    assert net <= 1e-12
    sim.set_precision(steps.PREC_MIXED)
    sim.forces()
    em = rel_err(sim.get_forces(), F)
    report("R3 N=2^21 mixed pair sum vs fp64", em)
    assert np.percentile(em, 99) <= 1e-5
    e32 = rel_err(oracle.forces_rows(p, w.x, w.M, w.soft, rows, kind="fp32"), F[rows])
    report("R3 N=2^21 reference all-float stand-in vs fp64 (sampled rows)", e32)
    assert np.median(em[rows]) * 10 < np.median(e32)
    sim.close()
