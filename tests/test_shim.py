"""The C++ drop-in shim: the reference's own symbols (forces, forces_periodic,
forces_periodic_z, recalculate_softening) over the C ABI, driven the way
StePS's step.cc:191-197 drives them, with its globals defined by the harness
in tests/shim (standing in for StePS's main.cc)."""
import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest

import steps_amd
from conftest import rel_err, report
from steps_amd import workloads as wl

LIB = Path(__file__).resolve().parent.parent / "steps_amd" / "lib"
DP = C.POINTER(C.c_double)


def shim(variant):
    lib = C.CDLL(str(LIB / f"libsteps_shim_{variant}.so"))
    lib.harness_set_particles.argtypes = [C.c_int, DP, DP]
    lib.harness_set_scalars.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 5
    lib.harness_set_tables.argtypes = [DP, C.c_int, DP, C.c_int, C.c_int, DP, C.c_int]
    lib.harness_set_softening.argtypes = [C.c_double, C.c_double]
    lib.harness_call.argtypes = [DP, DP, C.c_int, C.c_int]
    lib.harness_call.restype = C.c_int
    lib.harness_recalculate_softening.argtypes = [DP, DP]
    return lib


def p(a):
    return a.ctypes.data_as(DP) if a is not None else None


def test_shim_variants_export_reference_symbols():
    import subprocess
    expect = {"r3": "forces(double*, double*, int, int)",
              "t3": "forces_periodic(double*, double*, int, int)",
              "s1r2": "forces_periodic_z(double*, double*, int, int)"}
    for v, sym in expect.items():
        out = subprocess.run(["nm", "-DC", "--defined-only", str(LIB / f"libsteps_shim_{v}.so")],
                             capture_output=True, text=True, check=True).stdout
        assert sym in out, (v, sym)
        assert "recalculate_softening()" in out


def test_recalculate_softening_matches_reference_formula():
    lib = shim("r3")
    lib.harness_set_scalars(0, 1, 1, 0, 0, 0, 0, 0)
    lib.harness_set_softening(0.25, 3.0)
    b, rp = C.c_double(), C.c_double()
    lib.harness_recalculate_softening(C.byref(b), C.byref(rp))
    assert b.value == 0.25
    assert rp.value == pytest.approx(3.0 / (4.0 * math.pi * 0.25 ** 3 / 3.0), rel=1e-15)


def test_failure_sets_force_error_without_device():
    if steps_amd.device_count() > 0:
        pytest.skip("a HIP device is present")
    lib = shim("r3")
    x = np.zeros(30)
    M = np.ones(10)
    h = np.ones(10)
    F = np.zeros(30)
    lib.harness_set_particles(10, p(M), p(h))
    lib.harness_set_scalars(0, 0, 0, 0, 0, 0, 0, 0)
    assert lib.harness_call(p(x), p(F), 0, 9) == 1        # STEPS_ERR_DEVICE -> ForceError


def test_bad_row_range_sets_force_error():
    lib = shim("r3")
    x = np.zeros(30)
    M = np.ones(10)
    h = np.ones(10)
    F = np.zeros(30)
    lib.harness_set_particles(10, p(M), p(h))
    lib.harness_set_scalars(0, 0, 0, 0, 0, 0, 0, 0)
    assert lib.harness_call(p(x), p(F), 3, 12) == 1       # id_max >= N


@pytest.mark.gpu
def test_shim_r3_matches_abi_and_oracle(steps, oracle):
    w = wl.r3_compactified(N=4096, seed=77)
    lib = shim("r3")
    M, h = w.M.copy(), w.soft.copy()
    x = w.x.reshape(-1).copy()
    lib.harness_set_particles(w.N, p(M), p(h))
    lib.harness_set_scalars(0, 1, 1, 0.0, 0.0, w.mass_in_unit_sphere, wl.H0, wl.OMEGA_L)
    F = np.zeros(3 * 1000)
    assert lib.harness_call(p(x), p(F), 1000, 1999) == 0
    d, keep = steps.make_desc(steps.TOPO_R3, cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere,
                              H0=wl.H0, Omega_lambda=wl.OMEGA_L)
    assert np.array_equal(F.reshape(-1, 3), steps.force_eval(d, w.x, w.M, w.soft, 1000, 1999))
    Fo = oracle.forces(oracle.params(0, cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere),
                       w.x, w.M, w.soft, 1000, 1999)
    assert np.median(rel_err(F, Fo)) <= 1e-12


@pytest.mark.gpu
def test_shim_t3_and_s1r2(steps, oracle):
    # T3 (-DPERIODIC), Ewald table from the reference-equivalent builder
    w = wl.t3_lattice(n_side=10, seed=3)
    T = oracle.t3_table(2, w.L)
    lib = shim("t3")
    M, h, x = w.M.copy(), w.soft.copy(), w.x.reshape(-1).copy()
    lib.harness_set_particles(w.N, p(M), p(h))
    lib.harness_set_scalars(2, 1, 1, w.L, 0.0, 0.0, 0.0, 0.0)
    lib.harness_set_tables(p(T), 63, None, 0, 0, None, 0)
    F = np.zeros(3 * w.N)
    assert lib.harness_call(p(x), p(F), 0, w.N - 1) == 0
    Fo = oracle.forces(oracle.params(1, is_periodic=2, L=w.L, t3_table=T, interp_order=4), w.x, w.M, w.soft)
    report("shim T3", rel_err(F, Fo))
    assert np.median(rel_err(F, Fo)) <= 1e-12
    # S1R2 (-DPERIODIC_Z, EWALD_INTERPOLATION_ORDER=4)
    c = wl.s1r2_cylinder(n_omega=16, n_phi=8, n_z=8, seed=4)
    S, (nrho, nz) = oracle.s1r2_table(2, c.L, c.Rsim)
    lib = shim("s1r2")
    M, h, x = c.M.copy(), c.soft.copy(), c.x.reshape(-1).copy()
    lib.harness_set_particles(c.N, p(M), p(h))
    lib.harness_set_scalars(2, 1, 1, c.L, c.Rsim, c.mass_in_unit_sphere, 0.0, 0.0)
    R = np.zeros(1000)
    lib.harness_set_tables(None, 0, p(S), nrho, nz, p(R), R.size)
    F = np.zeros(3 * c.N)
    assert lib.harness_call(p(x), p(F), 0, c.N - 1) == 0
    kw = dict(is_periodic=2, L=c.L, Rsim=c.Rsim, cosmology=1, comoving=1, mass_in_unit_sphere=c.mass_in_unit_sphere,
              s1r2_lookup=1, interp_order=4, s1r2_table=S, s1r2_shape=(nrho, nz))
    Fo = oracle.forces(oracle.params(2, **kw), c.x, c.M, c.soft)
    report("shim S1R2", rel_err(F, Fo))
    assert np.median(rel_err(F, Fo)) <= 1e-12


@pytest.mark.gpu
def test_shim_barnes_hut_builds(steps, oracle):
    """-DUSE_BH: the reference's forces() / forces_periodic() with the global
    THETA evaluate the octree; bitwise the C ABI's octree evaluation."""
    w = wl.r3_compactified(N=4096, seed=78)
    lib = shim("r3_bh")
    lib.harness_set_theta.argtypes = [C.c_double]
    M, h, x = w.M.copy(), w.soft.copy(), w.x.reshape(-1).copy()
    lib.harness_set_particles(w.N, p(M), p(h))
    lib.harness_set_scalars(0, 1, 1, 0.0, 0.0, w.mass_in_unit_sphere, wl.H0, wl.OMEGA_L)
    lib.harness_set_softening(w.particle_radius, float(w.M.min()))
    lib.harness_set_theta(0.5)
    F = np.zeros(3 * w.N)
    assert lib.harness_call(p(x), p(F), 0, w.N - 1) == 0
    d, keep = steps.make_desc(steps.TOPO_R3, cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere,
                              H0=wl.H0, Omega_lambda=wl.OMEGA_L, bh_theta=0.5, bh_particle_radius=w.particle_radius)
    assert np.array_equal(F.reshape(-1, 3), steps.force_eval(d, w.x, w.M, w.soft))
    Fo = oracle.forces_bh(oracle.params(0, cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere),
                          0.5, w.particle_radius, w.x, w.M, w.soft)
    assert rel_err(F, Fo).max() <= 1e-12
    # T3 (-DPERIODIC -DUSE_BH): the CPU build's trilinear table interpolation
    t = wl.t3_lattice(n_side=10, seed=3)
    T = oracle.t3_table(2, t.L)
    lib = shim("t3_bh")
    lib.harness_set_theta.argtypes = [C.c_double]
    M, h, x = t.M.copy(), t.soft.copy(), t.x.reshape(-1).copy()
    lib.harness_set_particles(t.N, p(M), p(h))
    lib.harness_set_scalars(2, 1, 1, t.L, 0.0, 0.0, 0.0, 0.0)
    lib.harness_set_tables(p(T), 63, None, 0, 0, None, 0)
    lib.harness_set_softening(t.particle_radius, float(t.M.min()))
    lib.harness_set_theta(0.5)
    F = np.zeros(3 * t.N)
    assert lib.harness_call(p(x), p(F), 0, t.N - 1) == 0
    Fo = oracle.forces_bh(oracle.params(1, is_periodic=2, L=t.L, t3_table=T, interp_order=2), 0.5, t.particle_radius,
                          t.x, t.M, t.soft)
    report("shim T3 octree", rel_err(F, Fo))
    assert rel_err(F, Fo).max() <= 1e-12


def shim_randomised_r3_errors():
    """Three calls of the randomised R3 shim after srand(7), the third with the
    radial correction; the largest per-row relative error of each against the
    oracle with the draws rand() gives after srand(7).  Run in a process of
    its own by the test below: the HIP runtime draws from rand() as it starts,
    which must not shift the draws."""
    import oracle_lib as oracle
    libc = C.CDLL("libc.so.6")
    libc.rand.restype = C.c_int
    libc.srand.argtypes = [C.c_uint]
    w = wl.r3_compactified(N=4096, seed=79)
    Rsim = float(np.sqrt((w.x ** 2).sum(axis=1)).max())
    tab = np.linspace(-2.0, 3.0, 16) * 1e-3 * w.mass_in_unit_sphere * Rsim
    lib = shim("r3_bh_rand")
    lib.harness_set_theta.argtypes = [C.c_double]
    lib.harness_set_radial_bh.argtypes = [DP, C.c_int, C.c_int]
    M, h, x = w.M.copy(), w.soft.copy(), w.x.reshape(-1).copy()
    lib.harness_set_particles(w.N, p(M), p(h))
    lib.harness_set_scalars(0, 1, 1, 0.0, Rsim, w.mass_in_unit_sphere, wl.H0, wl.OMEGA_L)
    lib.harness_set_softening(w.particle_radius, float(w.M.min()))
    lib.harness_set_theta(0.5)
    libc.srand(7)
    u = [[libc.rand() / 2147483647 for _ in range(6)] for _ in range(3)]  # glibc RAND_MAX
    libc.srand(7)
    op = oracle.params(0, cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere, Rsim=Rsim)
    errs = []
    for k in range(3):
        lib.harness_set_radial_bh(p(tab), tab.size, 1 if k == 2 else 0)
        F = np.zeros(3 * w.N)
        assert lib.harness_call(p(x), p(F), 0, w.N - 1) == 0
        Fo = oracle.forces_bh_ex(op, 0.5, w.particle_radius, w.x, w.M, w.soft, u=u[k],
                                 radial_bh=tab if k == 2 else None)
        errs.append(float(rel_err(F, Fo).max()))
    assert np.array_equal(x, w.x.reshape(-1))  # positions are left as they were
    return errs


@pytest.mark.gpu
def test_shim_randomised_octree_and_radial_correction(steps, oracle):
    """-DUSE_BH -DRANDOMIZE_BH: each call of the reference's force routine
    draws its frame with rand() in the reference's order (six draws in R3,
    four in S1R2); the harness seeds rand() as main.cc:310 does.  Once
    USE_RADIAL_BH_CORRECTION is set, the table's correction is subtracted."""
    import os
    import subprocess
    import sys
    # a fresh process, whose first library call starts the HIP runtime
    here = Path(__file__).resolve().parent
    env = dict(os.environ, PYTHONPATH=f"{here.parent}{os.pathsep}{here}")
    r = subprocess.run([sys.executable, "-c", "import test_shim; print(test_shim.shim_randomised_r3_errors())"],
                       cwd=str(here), env=env, capture_output=True, text=True, timeout=300)
    print(r.stdout[-500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    errs = [float(v) for v in r.stdout.strip().splitlines()[-1].strip("[]").split(",")]
    report("shim randomised octree R3 (fresh process), max per call", np.array(errs))
    assert max(errs) <= 5e-12
    libc = C.CDLL("libc.so.6")
    libc.rand.restype = C.c_int
    libc.srand.argtypes = [C.c_uint]
    rand_max = 2147483647
    # S1R2 with the CIC table: four draws per call
    c = wl.s1r2_cylinder(n_omega=12, n_phi=12, n_z=12, seed=6, is_periodic=2)
    T, shape = oracle.s1r2_table(2, c.L, c.Rsim)
    lib = shim("s1r2_bh_rand")
    lib.harness_set_theta.argtypes = [C.c_double]
    M, h, x = c.M.copy(), c.soft.copy(), c.x.reshape(-1).copy()
    lib.harness_set_particles(c.N, p(M), p(h))
    lib.harness_set_scalars(2, 1, 1, c.L, c.Rsim, c.mass_in_unit_sphere, 0.0, 0.0)
    lib.harness_set_tables(None, 0, p(T), shape[0], shape[1], None, 0)
    lib.harness_set_softening(c.particle_radius, float(c.M.min()))
    lib.harness_set_theta(0.5)
    libc.srand(11)
    u = [libc.rand() / rand_max for _ in range(4)]
    libc.srand(11)
    F = np.zeros(3 * c.N)
    assert lib.harness_call(p(x), p(F), 0, c.N - 1) == 0
    Fo = oracle.forces_bh_ex(oracle.params(2, is_periodic=2, L=c.L, Rsim=c.Rsim, cosmology=1, comoving=1,
                                           mass_in_unit_sphere=c.mass_in_unit_sphere, s1r2_table=T,
                                           s1r2_shape=shape, interp_order=2),
                             0.5, c.particle_radius, c.x, c.M, c.soft, u=u)
    report("shim randomised octree S1R2 CIC", rel_err(F, Fo))
    assert rel_err(F, Fo).max() <= 5e-12
