"""steps_run, the driver that runs a StePS parameter file on the device: a
compactified R3 comoving run from an ASCII initial-condition file, checked
against the oracle advanced the same way (the same step lengths from the
same errmax rule, the scale factor by RK4), and the snapshot's units."""
import ctypes as C
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from steps_amd import workloads as wl

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
UNIT_T, UNIT_V = 47.14829951063323, 20.738652969925447


def test_steps_run_r3_comoving_matches_oracle(steps, oracle, tmp_path):
    print("case steps_run R3 comoving N=2048 from an ASCII IC", flush=True)
    w = wl.r3_compactified(N=2048, seed=21)
    a0, a_max, acc = 1.0 / 32, 0.033, 0.0125
    h_min, h_max = 2e-5 / UNIT_T, 0.03125 / UNIT_T
    ic = tmp_path / "ic.dat"
    # the IC convention: velocities in km/s * sqrt(a) (main.cc:1147-1154)
    with open(ic, "w") as f:
        for i in range(w.N):
            row = list(w.x[i]) + list(w.v[i] * math.sqrt(a0) * UNIT_V) + [w.M[i]]
            f.write("\t".join("%.17g" % c for c in row) + "\n")
    out = tmp_path / "out"
    out.mkdir()
    par = tmp_path / "run.param"
    par.write_text("\n".join([
        "Omega_b\t0.0", f"Omega_lambda\t{wl.OMEGA_L!r}", f"Omega_m\t{wl.OMEGA_M!r}", "Omega_r\t0.0",
        f"HubbleConstant\t{wl.H0 * UNIT_V!r}", f"a_start\t{a0!r}", f"a_max\t{a_max!r}",
        "COSMOLOGY\t1", "IS_PERIODIC\t0", "COMOVING_INTEGRATION\t1", "L_BOX\t100.0", "R_SIM\t1000.0",
        f"IC_FILE\t{ic}", "IC_FORMAT\t0", f"OUT_DIR\t{out}", "OUTPUT_TIME_VARIABLE\t0", "OUTPUT_FORMAT\t0",
        f"ACC_PARAM\t{acc!r}", "STEP_MIN\t0.00002", "STEP_MAX\t0.03125", f"PARTICLE_RADII\t{w.particle_radius!r}",
        "FIRST_T_OUT\t13.0", "H_OUT\t13.0", ""]))
    exe = ROOT / "steps_amd" / "lib" / "steps_run"
    r = subprocess.run([str(exe), str(par), "--max-steps", "40"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-600:], r.stderr[-600:], flush=True)
    assert r.returncode == 0
    log = np.loadtxt(out / "Logfile.dat")
    nsteps = log.shape[0]
    assert 2 <= nsteps < 40

    # the oracle, the same run
    kw = dict(cosmology=1, comoving=1, mass_in_unit_sphere=4 * math.pi / 3 * 3 * wl.H0 ** 2 / (8 * math.pi) * wl.OMEGA_M)
    p = oracle.params(0, **kw)
    cosmo = oracle.OracleCosmo(wl.H0, wl.OMEGA_M, wl.OMEGA_L, 0.0, 1.0 - wl.OMEGA_M - wl.OMEGA_L)
    lib = oracle.lib()
    soft, M_min, rho_part = steps.softening_lengths(w.M, w.particle_radius)   # utils.cc:59-82, as the driver does
    x = w.x.reshape(-1).copy()
    v = (w.v.reshape(-1) * math.sqrt(a0) * UNIT_V) / math.sqrt(a0) / UNIT_V
    F = oracle.forces(p, w.x, w.M, soft).reshape(-1).copy()
    st = oracle.OracleStepState(a0, lib.oracle_hubble(C.byref(cosmo), a0), 0.0, rho_part, 1.0)
    e = lib.oracle_initial_errmax(C.byref(p), C.byref(st), w.N, oracle.dp(x.copy()), oracle.dp(v), oracle.dp(F),
                                  oracle.dp(w.M))
    hs = []
    while st.a < a_max and len(hs) < 40:
        h = min(max(math.sqrt(2 * acc / e), h_min), h_max)
        hs.append(h)
        st.h = h
        e = lib.oracle_kdk_step(C.byref(p), C.byref(cosmo), C.byref(st), w.N, oracle.dp(x), oracle.dp(v),
                                oracle.dp(F), oracle.dp(w.M), oracle.dp(soft))
    assert len(hs) == nsteps
    assert np.allclose(log[:, 5], hs, rtol=1e-9)
    assert log[-1, 2] == pytest.approx(st.a, rel=1e-12)
    snap = np.loadtxt(out / "t1300.dat")
    assert snap.shape == (w.N, 7)
    assert np.allclose(snap[:, :3], x.reshape(-1, 3), rtol=0, atol=1e-10 * np.abs(x).max())
    vs = v.reshape(-1, 3) * math.sqrt(st.a) * UNIT_V
    assert np.allclose(snap[:, 3:6], vs, rtol=1e-8, atol=1e-10 * np.abs(vs).max())
    assert np.allclose(snap[:, 6], w.M, rtol=1e-15)


def _write_run(tmp_path, w, a0, a_max, acc, tag):
    ic = tmp_path / f"ic_{tag}.dat"
    with open(ic, "w") as f:
        for i in range(w.N):
            row = list(w.x[i]) + list(w.v[i] * math.sqrt(a0) * UNIT_V) + [w.M[i]]
            f.write("\t".join("%.17g" % c for c in row) + "\n")
    out = tmp_path / f"out_{tag}"
    out.mkdir()
    par = tmp_path / f"run_{tag}.param"
    par.write_text("\n".join([
        "Omega_b\t0.0", f"Omega_lambda\t{wl.OMEGA_L!r}", f"Omega_m\t{wl.OMEGA_M!r}", "Omega_r\t0.0",
        f"HubbleConstant\t{wl.H0 * UNIT_V!r}", f"a_start\t{a0!r}", f"a_max\t{a_max!r}",
        "COSMOLOGY\t1", "IS_PERIODIC\t0", "COMOVING_INTEGRATION\t1", "L_BOX\t100.0", "R_SIM\t1000.0",
        f"IC_FILE\t{ic}", "IC_FORMAT\t0", f"OUT_DIR\t{out}", "OUTPUT_TIME_VARIABLE\t0", "OUTPUT_FORMAT\t0",
        f"ACC_PARAM\t{acc!r}", "STEP_MIN\t0.00002", "STEP_MAX\t0.03125", f"PARTICLE_RADII\t{w.particle_radius!r}",
        "FIRST_T_OUT\t13.0", "H_OUT\t13.0", ""]))
    return par, out


def test_steps_run_barnes_hut(tmp_path):
    """--bh THETA (the reference's USE_BH=THETA build): with an opening angle
    too small to take any internal node the octree run is the direct-sum run
    (same step lengths, same final snapshot to rounding); at theta = 0.5 it
    runs the same schedule with forces within the octree's error."""
    print("case steps_run R3 comoving N=2048, direct vs octree", flush=True)
    w = wl.r3_compactified(N=2048, seed=23)
    exe = ROOT / "steps_amd" / "lib" / "steps_run"
    snaps = {}
    for tag, extra in (("direct", []), ("bh_tiny", ["--bh", "1e-9"]), ("bh_half", ["--bh", "0.5"])):
        par, out = _write_run(tmp_path, w, 1.0 / 32, 0.0318, 0.0125, tag)
        r = subprocess.run([str(exe), str(par), "--max-steps", "30", *extra], capture_output=True, text=True,
                           timeout=120)
        print(tag, r.stdout[-300:], r.stderr[-300:], flush=True)
        assert r.returncode == 0
        snaps[tag] = (np.loadtxt(out / "Logfile.dat"), np.loadtxt(out / "t1300.dat"))
    (log_d, s_d), (log_t, s_t), (log_h, s_h) = snaps["direct"], snaps["bh_tiny"], snaps["bh_half"]
    assert 2 <= log_d.shape[0] < 30
    assert log_t.shape == log_d.shape and np.allclose(log_t[:, 5], log_d[:, 5], rtol=1e-9)
    assert np.allclose(s_t[:, :3], s_d[:, :3], rtol=0, atol=1e-9 * np.abs(s_d[:, :3]).max())
    assert np.allclose(s_t[:, 3:6], s_d[:, 3:6], rtol=1e-6, atol=1e-9 * np.abs(s_d[:, 3:6]).max())
    assert log_h.shape[0] >= 2 and np.allclose(s_h[:, :3], s_d[:, :3], rtol=0, atol=1e-4 * np.abs(s_d[:, :3]).max())


def test_steps_run_randomised_octree_radial_table(tmp_path, oracle, steps):
    """--bh 0.5 --bh-seed 5 with RADIAL_BH_FORCE_CORRECTION 1 (the reference's
    USE_BH + RANDOMIZE_BH build, main.cc:791-1091): the table written to
    OUT_DIR/radial_bh_force_table.dat is the per-shell mean radial octree
    force of the IC (GLASS_FILE None) over RADIAL_BH_FORCE_TABLE_ITERATION
    evaluations, each with its own draws of rand() seeded by the seed -- the
    oracle's table from the same draws; the run then completes."""
    print("case steps_run R3 randomised octree + radial table N=2048", flush=True)
    w = wl.r3_compactified(N=2048, seed=24)
    Rsim = float(np.sqrt((w.x ** 2).sum(axis=1)).max()) * 1.01
    par, out = _write_run(tmp_path, w, 1.0 / 32, 0.0318, 0.0125, "bh_rand")
    with open(par, "a") as f:
        f.write(f"R_SIM\t{Rsim!r}\nRADIAL_BH_FORCE_CORRECTION\t1\nGLASS_FILE_FOR_BH_FORCE_CORRECTION\tNone\n"
                "RADIAL_BH_FORCE_TABLE_SIZE\t16\nRADIAL_BH_FORCE_TABLE_ITERATION\t3\n")
    exe = ROOT / "steps_amd" / "lib" / "steps_run"
    r = subprocess.run([str(exe), str(par), "--bh", "0.5", "--bh-seed", "5", "--max-steps", "4"], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-400:], r.stderr[-300:], flush=True)
    assert r.returncode == 0
    got = np.loadtxt(out / "radial_bh_force_table.dat", comments="#")
    assert got.shape == (16, 2)
    assert np.allclose(got[:, 0], np.round((np.arange(16) + 0.5) * Rsim / 16, 6), rtol=0, atol=1e-9)
    # the oracle's table from the same draws
    libc = C.CDLL("libc.so.6")
    libc.rand.restype = C.c_int
    libc.srand.argtypes = [C.c_uint]
    libc.srand(5)
    H0 = float(repr(wl.H0 * UNIT_V)) / UNIT_V
    misph = 4.0 * math.pi * (3.0 * H0 * H0 / (8.0 * math.pi)) * wl.OMEGA_M / 3.0
    soft = steps.softening_lengths(w.M, w.particle_radius)[0]
    p = oracle.params(0, cosmology=1, comoving=1, mass_in_unit_sphere=misph, Rsim=Rsim)
    acc, cnt = np.zeros(16), np.zeros(16, dtype=np.int32)
    for _ in range(3):
        u = [libc.rand() / 2147483647 for _ in range(6)]
        F = oracle.forces_bh_ex(p, 0.5, w.particle_radius, w.x, w.M, soft, u=u)
        oracle.radial_bh_table_accumulate(0, Rsim, w.x, F, acc, cnt)
    want = np.where(cnt > 0, acc / np.maximum(cnt, 1), 0.0)
    assert (cnt > 0).sum() >= 12
    assert np.allclose(got[:, 1], want, rtol=1e-10, atol=1e-15 + 1e-12 * np.abs(want).max())
    log = np.loadtxt(out / "Logfile.dat")
    assert log.shape[0] >= 2


@pytest.mark.parametrize("cosmoparam,extra", [(1, "w0\t-0.9\n"), (2, "w0\t-0.9\nwa\t0.3\n")])
def test_steps_run_dark_energy_models(tmp_path, cosmoparam, extra):
    """--cosmo-param 1 / 2 (the reference's COSMOPARAM builds, friedmann_solver.cc:
    174-333): the logged scale factors and Hubble parameters are the host
    library's wCDM / w0-wa expansion history replayed over the logged steps,
    from its start time of a_start."""
    print(f"case steps_run R3 comoving N=2048 COSMOPARAM={cosmoparam}", flush=True)
    import test_host_cosmo as hc
    w = wl.r3_compactified(N=2048, seed=25)
    par, out = _write_run(tmp_path, w, 1.0 / 32, 0.0318, 0.0125, f"de{cosmoparam}")
    with open(par, "a") as f:
        f.write(extra)
    exe = ROOT / "steps_amd" / "lib" / "steps_run"
    r = subprocess.run([str(exe), str(par), "--cosmo-param", str(cosmoparam), "--max-steps", "12"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:], flush=True)
    assert r.returncode == 0
    log = np.loadtxt(out / "Logfile.dat")
    host = C.CDLL(str(ROOT / "steps_amd" / "lib" / "libsteps_host.so"))
    E = C.POINTER(hc.Expansion)
    host.steps_hubble.restype = C.c_double
    host.steps_hubble.argtypes = [E, C.c_double]
    host.steps_friedmann_step.restype = C.c_double
    host.steps_friedmann_step.argtypes = [E, C.c_double, C.c_double, C.c_double]
    host.steps_friedmann_start.restype = C.c_double
    host.steps_friedmann_start.argtypes = [E, C.c_double, C.c_double]
    H0 = float(repr(wl.H0 * UNIT_V)) / UNIT_V
    e = hc.model(cosmoparam, H0=H0, Om=wl.OMEGA_M, OL=wl.OMEGA_L, w0=-0.9, wa=0.3 if cosmoparam == 2 else 0.0)
    T = host.steps_friedmann_start(C.byref(e), 1.0 / 32, 0.00002 / UNIT_T * 0.05)
    a = 1.0 / 32
    for row in log:
        h = row[5]
        T += h
        a = host.steps_friedmann_step(C.byref(e), a, h, T)
        assert row[1] == pytest.approx(T, rel=1e-13)
        assert row[2] == pytest.approx(a, rel=1e-13)
        assert row[4] == pytest.approx(host.steps_hubble(C.byref(e), a), rel=1e-13)
    # and it is not LCDM
    lc = hc.model(0, H0=H0, Om=wl.OMEGA_M, OL=wl.OMEGA_L)
    assert abs(host.steps_hubble(C.byref(lc), log[-1, 2]) - log[-1, 4]) > 1e-9 * log[-1, 4]


def test_steps_run_tabulated_expansion_history(tmp_path):
    """--cosmo-param -1 (friedmann_solver.cc:336-466): the run reads
    EXPANSION_FILE (t [Gy], a, H [km/s/Mpc]) and takes a and H by cubic
    interpolation (INTERPOLATION_ORDER 3) at each step's end time, from the
    start time of the row before the first past a_start; the log is the host
    library's replay of the same table over the logged steps."""
    print("case steps_run R3 comoving N=2048 tabulated expansion history", flush=True)
    import test_host_cosmo as hc
    host = C.CDLL(str(ROOT / "steps_amd" / "lib" / "libsteps_host.so"))
    E = C.POINTER(hc.Expansion)
    host.steps_hubble.restype = C.c_double
    host.steps_hubble.argtypes = [E, C.c_double]
    host.steps_friedmann_step.restype = C.c_double
    host.steps_friedmann_step.argtypes = [E, C.c_double, C.c_double, C.c_double]
    host.steps_friedmann_start.restype = C.c_double
    host.steps_friedmann_start.argtypes = [E, C.c_double, C.c_double]
    host.steps_read_expansion_history.restype = C.c_int64
    host.steps_read_expansion_history.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.c_int64]
    H0 = float(repr(wl.H0 * UNIT_V)) / UNIT_V
    lc = hc.model(0, H0=H0, Om=wl.OMEGA_M, OL=wl.OMEGA_L)
    # an LCDM history from a = 0.02 on a fine grid, written in the file's units
    a, t, h = 0.02, host.steps_friedmann_start(C.byref(lc), 0.02, 1e-6), 2e-5
    rows = []
    while a < 0.05:
        rows.append((t * UNIT_T, a, host.steps_hubble(C.byref(lc), a) * UNIT_V))
        t += h
        a = host.steps_friedmann_step(C.byref(lc), a, h, t)
    hist = tmp_path / "expansion.dat"
    hist.write_text("".join("%.17g %.17g %.17g\n" % r for r in rows))
    w = wl.r3_compactified(N=2048, seed=26)
    par, out = _write_run(tmp_path, w, 1.0 / 32, 0.0318, 0.0125, "tab")
    with open(par, "a") as f:
        f.write(f"EXPANSION_FILE\t{hist}\nINTERPOLATION_ORDER\t3\n")
    exe = ROOT / "steps_amd" / "lib" / "steps_run"
    r = subprocess.run([str(exe), str(par), "--cosmo-param", "-1", "--max-steps", "12"], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:], flush=True)
    assert r.returncode == 0
    log = np.loadtxt(out / "Logfile.dat")
    n = host.steps_read_expansion_history(str(hist).encode(), None, 0)
    assert n == len(rows)
    tab = np.zeros(3 * n)
    assert host.steps_read_expansion_history(str(hist).encode(), tab.ctypes.data_as(C.POINTER(C.c_double)), n) == n
    e = hc.model(-1, H0=H0, Om=wl.OMEGA_M, OL=wl.OMEGA_L)
    e.table = tab.ctypes.data_as(C.POINTER(C.c_double))
    e.rows = n
    e.order = 3
    T = host.steps_friedmann_start(C.byref(e), 1.0 / 32, 0.00002 / UNIT_T * 0.05)
    assert tab[3 * (e.index - 1) + 1] <= 1.0 / 32 < tab[3 * e.index + 1]
    aa = 1.0 / 32
    for row in log:
        T += row[5]
        aa = host.steps_friedmann_step(C.byref(e), aa, row[5], T)
        assert row[1] == pytest.approx(T, rel=1e-13)
        assert row[2] == pytest.approx(aa, rel=1e-13)
        assert row[4] == pytest.approx(host.steps_hubble(C.byref(e), aa), rel=1e-13)
    # the tabulated run follows the LCDM history it was sampled from
    assert log[-1, 2] == pytest.approx(host.steps_friedmann_step(C.byref(lc), log[-2, 2], log[-1, 5], 0.0), rel=1e-6)
