"""The host side of a StePS run (include/steps_host.h), on the CPU: the
parameter-file parser against the reference's own example files, the
strstr key semantics and clamps of read_paramfile.cc, and the ASCII / GADGET
initial-condition and snapshot formats (round trips, a synthetic GADGET file
with shuffled IDs and two particle types)."""
import ctypes as C
import math
import struct
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLD = Path(__file__).resolve().parent / "golden" / "params"
UNIT_T, UNIT_V = 47.14829951063323, 20.738652969925447
R3, T3, S1R2 = 0, 1, 2


class RunParams(C.Structure):
    _fields_ = [("Omega_b", C.c_double), ("Omega_lambda", C.c_double), ("Omega_m", C.c_double),
                ("Omega_r", C.c_double), ("H0", C.c_double), ("IC_FILE", C.c_char * 1024), ("L", C.c_double),
                ("IS_PERIODIC", C.c_int32), ("OUTPUT_FORMAT", C.c_int32), ("OUT_DIR", C.c_char * 1024),
                ("a_max", C.c_double), ("N_snapshot", C.c_uint32), ("ACC_PARAM", C.c_double),
                ("h_min", C.c_double), ("h_max", C.c_double), ("COSMOLOGY", C.c_int32), ("a_start", C.c_double),
                ("COMOVING_INTEGRATION", C.c_int32), ("TIME_LIMIT_IN_MINS", C.c_double),
                ("FIRST_T_OUT", C.c_double), ("IC_FORMAT", C.c_int32), ("OUTPUT_TIME_VARIABLE", C.c_int32),
                ("OUT_LST", C.c_char * 1024), ("H_OUT", C.c_double), ("ParticleRadi", C.c_double),
                ("MIN_REDSHIFT", C.c_double), ("REDSHIFT_CONE", C.c_int32), ("H0_INDEPENDENT_UNITS", C.c_int32),
                ("Rsim", C.c_double), ("RADIAL_FORCE_ACCURACY", C.c_int32), ("RADIAL_FORCE_TABLE_SIZE", C.c_int32),
                ("RADIAL_BH_FORCE_CORRECTION", C.c_int32), ("GLASS_FILE_FOR_BH_FORCE_CORRECTION", C.c_char * 1024),
                ("RADIAL_BH_FORCE_TABLE_SIZE", C.c_int32), ("RADIAL_BH_FORCE_TABLE_ITERATION", C.c_int32),
                ("w0", C.c_double), ("wa", C.c_double), ("EXPANSION_FILE", C.c_char * 1024),
                ("INTERPOLATION_ORDER", C.c_int32)]


@pytest.fixture(scope="module")
def host():
    p = ROOT / "steps_amd" / "lib" / "libsteps_host.so"
    if not p.exists():
        from steps_amd import build
        build.build_host()
    lib = C.CDLL(str(p))
    dp = C.POINTER(C.c_double)
    lib.steps_read_params.argtypes = [C.c_char_p, C.c_int32, C.POINTER(RunParams)]
    lib.steps_host_last_error.restype = C.c_char_p
    lib.steps_ascii_count_particles.restype = C.c_int64
    lib.steps_ascii_count_particles.argtypes = [C.c_char_p]
    lib.steps_read_ascii_ic.argtypes = [C.c_char_p, C.c_int64, dp, dp, dp]
    lib.steps_read_gadget_ic.argtypes = [C.c_char_p, C.POINTER(C.c_int64), dp, dp, dp, dp, dp]
    lib.steps_read_output_list.restype = C.c_int64
    lib.steps_read_output_list.argtypes = [C.c_char_p, dp, C.c_int64]
    lib.steps_write_ascii_snapshot.argtypes = [C.c_char_p, C.c_int64, dp, dp, dp, C.c_double, C.c_double]
    return lib


def ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def read(host, path, topology):
    p = RunParams()
    rc = host.steps_read_params(str(path).encode(), topology, C.byref(p))
    assert rc == 0, host.steps_host_last_error()
    return p


def test_example_cylindrical_file(host):
    p = read(host, GOLD / "LCDM_cylindrical_Lz100_D1000.param", S1R2)
    assert (p.Omega_b, p.Omega_lambda, p.Omega_m, p.Omega_r) == (0.0, 0.6889, 0.3111, 0.0)
    assert p.H0 == pytest.approx(67.66 / UNIT_V, rel=1e-15)
    assert (p.a_start, p.a_max, p.COSMOLOGY, p.COMOVING_INTEGRATION) == (0.03125, 1.0, 1, 1)
    assert p.IS_PERIODIC == 15                      # not clamped in the S1xR2 build
    assert (p.L, p.Rsim, p.IC_FORMAT, p.OUTPUT_TIME_VARIABLE) == (100.0, 500.0, 2, 1)
    assert p.OUTPUT_FORMAT == 0                     # HDF5 output needs HAVE_HDF5: set to ASCII
    assert p.IC_FILE == b"./examples/ic/lcdm_cylindrical_Lx1000_Ly1000_Lz100_R3D500_D4D75_z31.hdf5"
    assert p.OUT_DIR == b"./examples/LCDM_cylindrical_Lz100_D1000/"
    assert p.OUT_LST == b"./examples/outredshifts.txt"
    assert (p.RADIAL_FORCE_ACCURACY, p.RADIAL_FORCE_TABLE_SIZE) == (7500, 500)
    assert p.h_min == pytest.approx(0.00002 / UNIT_T, rel=1e-15)     # Gy -> internal units
    assert p.h_max == pytest.approx(0.03125 / UNIT_T, rel=1e-15)
    assert (p.ACC_PARAM, p.ParticleRadi, p.MIN_REDSHIFT) == (0.0125, 0.108326, 0.0003012504)
    assert (p.FIRST_T_OUT, p.H_OUT, p.REDSHIFT_CONE, p.H0_INDEPENDENT_UNITS) == (1.0, 1.0, 0, 0)
    # the octree build's keys (read_paramfile.cc:130-141, 348-379)
    assert (p.RADIAL_BH_FORCE_CORRECTION, p.RADIAL_BH_FORCE_TABLE_SIZE, p.RADIAL_BH_FORCE_TABLE_ITERATION) == (1, 128, 16)
    assert p.GLASS_FILE_FOR_BH_FORCE_CORRECTION == b"../../StePS_IC/examples/Glass_cylindrical_L100_D500.hdf5"
    q = read(host, GOLD / "LCDM_SP_1860_com_VOI100.param", R3)
    assert (q.RADIAL_BH_FORCE_CORRECTION, q.RADIAL_BH_FORCE_TABLE_SIZE, q.RADIAL_BH_FORCE_TABLE_ITERATION) == (1, 64, 7)
    assert q.GLASS_FILE_FOR_BH_FORCE_CORRECTION == b"None"


def test_example_files_parse_in_every_build(host):
    for f in sorted(GOLD.glob("*.param")):
        txt = f.read_text()
        for topo in (R3, T3, S1R2):
            p = read(host, f, topo)
            ip = int([ln.split()[1] for ln in txt.splitlines() if "IS_PERIODIC" in ln][0])
            assert p.IS_PERIODIC == (ip if topo == S1R2 else min(ip, 4))
            if topo != S1R2:                        # the RADIAL_FORCE_* keys exist only for S1xR2
                assert p.RADIAL_FORCE_ACCURACY == 0 and p.RADIAL_FORCE_TABLE_SIZE == 0
    p = read(host, GOLD / "LCDM_SP_1860_noncom_VOI100.param", R3)
    assert p.COMOVING_INTEGRATION == 0 and p.L == 1860.053144376


def test_strstr_semantics_and_clamps(host, tmp_path):
    f = tmp_path / "t.param"
    f.write_text("\n".join([
        "HubbleConstant  70.0",
        "COSMOLOGY 0",                   # non-cosmological: step limits stay in their own units
        "STEP_MIN 0.5",
        "COMOVING_INTEGRATION 7",        # clamped to 1
        "IS_PERIODIC 9",                 # clamped to 4 outside S1xR2
        "H_OUT_FIRST_T_OUT 2.5",         # one line, two keys: both set
        "RADIAL_FORCE_TABLE_SIZE 2",     # at least 5 (S1xR2)
        "RADIAL_FORCE_ACCURACY 3",       # at least 10
        "IC_FILE ./data/a_max_run/ic.dat",  # 'a_max' inside a value: its %lf read fails, a_max unchanged
        "OUT_DIR    out/run",            # a missing '/' is appended
        "# Omega_m 0.9",                 # the second token is not a number: unchanged
        "H_INDEPENDENT_UNITS 1",         # dropped without cosmology
        "SNAPSHOT_START_NUMBER 12",
        "ACC_PARAM 0x10",                # %lf reads hex floats too
        "REDSHIFT_CONE 010",             # %i reads octal
    ]) + "\n")
    p = read(host, f, T3)
    assert p.H0 == pytest.approx(70.0 / UNIT_V) and p.COSMOLOGY == 0 and p.h_min == 0.5
    assert p.COMOVING_INTEGRATION == 1 and p.IS_PERIODIC == 4
    assert p.H_OUT == 2.5 and p.FIRST_T_OUT == 2.5
    assert p.IC_FILE == b"./data/a_max_run/ic.dat" and p.a_max == 0.0
    assert p.OUT_DIR == b"out/run/" and p.Omega_m == 0.0 and p.H0_INDEPENDENT_UNITS == 0
    assert p.N_snapshot == 12 and p.ACC_PARAM == 16.0 and p.REDSHIFT_CONE == 8
    s = read(host, f, S1R2)
    assert s.IS_PERIODIC == 9 and s.RADIAL_FORCE_TABLE_SIZE == 5 and s.RADIAL_FORCE_ACCURACY == 10
    # octree keys: defaults, then the clamps
    assert (p.RADIAL_BH_FORCE_CORRECTION, p.RADIAL_BH_FORCE_TABLE_SIZE, p.RADIAL_BH_FORCE_TABLE_ITERATION) == (0, 64, 4)
    assert p.GLASS_FILE_FOR_BH_FORCE_CORRECTION == b"None"
    g = tmp_path / "bh.param"
    g.write_text("RADIAL_BH_FORCE_CORRECTION 2\nRADIAL_BH_FORCE_TABLE_SIZE 3\nRADIAL_BH_FORCE_TABLE_ITERATION 0\n"
                 "GLASS_FILE_FOR_BH_FORCE_CORRECTION glass.dat\n")
    b = read(host, g, R3)
    assert (b.RADIAL_BH_FORCE_CORRECTION, b.RADIAL_BH_FORCE_TABLE_SIZE, b.RADIAL_BH_FORCE_TABLE_ITERATION) == (0, 5, 1)
    assert b.GLASS_FILE_FOR_BH_FORCE_CORRECTION == b"glass.dat"
    # the COSMOPARAM keys; INTERPOLATION_ORDER outside 1..3 reads as 1 (read_paramfile.cc:627-631)
    assert p.INTERPOLATION_ORDER == 1 and p.w0 == 0.0 and p.wa == 0.0
    g.write_text("w0 -0.9\nwa 0.25\nEXPANSION_FILE hist.dat\nINTERPOLATION_ORDER 3\n")
    c = read(host, g, R3)
    assert (c.w0, c.wa, c.EXPANSION_FILE, c.INTERPOLATION_ORDER) == (-0.9, 0.25, b"hist.dat", 3)
    g.write_text("INTERPOLATION_ORDER 7\n")
    assert read(host, g, R3).INTERPOLATION_ORDER == 1
    # H0-independent units convert L and PARTICLE_RADII with cosmology on
    f.write_text("COSMOLOGY 1\nHubbleConstant 67.66\nL_BOX 100\nPARTICLE_RADII 0.5\nH_INDEPENDENT_UNITS 1\n")
    q = read(host, f, T3)
    assert q.L == pytest.approx(100 / 0.6766, rel=1e-14) and q.ParticleRadi == pytest.approx(0.5 / 0.6766, rel=1e-14)
    assert host.steps_read_params(str(tmp_path / "missing.param").encode(), R3, C.byref(RunParams())) < 0


def test_ascii_table_writer(host, tmp_path):
    # save_function_to_ascii_table (inputoutput.cc:1038-1061): header line, then
    # "%.6f\t%.15f" rows while the argument does not exceed x_max
    dp = C.POINTER(C.c_double)
    host.steps_write_ascii_table.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_double, dp, C.c_int32,
                                             C.c_char_p]
    v = np.array([0.1, -2.5, 3.25, 7.0])
    f = tmp_path / "table.dat"
    assert host.steps_write_ascii_table(str(f).encode(), 0.5, 3.0, 1.0, ptr(v), 4, b"# h1\n# h2") == 0
    assert f.read_text() == ("# h1\n# h2\n0.500000\t0.100000000000000\n1.500000\t-2.500000000000000\n"
                             "2.500000\t3.250000000000000\n")


def test_ascii_ic_and_snapshot_round_trip(host, tmp_path):
    rng = np.random.default_rng(5)
    N = 257
    x, v, M = rng.normal(size=(N, 3)) * 50, rng.normal(size=(N, 3)), rng.uniform(0.5, 2.0, N)
    f = tmp_path / "snap.dat"
    a, hd = 0.25, 0.7
    assert host.steps_write_ascii_snapshot(str(f).encode(), N, ptr(x), ptr(v), ptr(M), a, hd) == 0
    rows = [ln.split("\t") for ln in f.read_text().splitlines()]
    assert len(rows) == N and all(len(r) == 8 and r[7] == "" for r in rows)   # "%.16f\t" x 7, newline
    assert rows[0][0] == "%.16f" % (x[0, 0] * hd)
    assert host.steps_ascii_count_particles(str(f).encode()) == N
    xr, vr, Mr = np.zeros(3 * N), np.zeros(3 * N), np.zeros(N)
    assert host.steps_read_ascii_ic(str(f).encode(), N, ptr(xr), ptr(vr), ptr(Mr)) == 0
    assert np.allclose(xr.reshape(N, 3), x * hd, rtol=1e-15, atol=1e-13)
    assert np.allclose(vr.reshape(N, 3), v * math.sqrt(a) * UNIT_V, rtol=1e-15, atol=1e-13)
    assert np.allclose(Mr, M * hd, rtol=1e-15, atol=1e-15)
    # too few particles in the file: an error, not a partial read
    assert host.steps_read_ascii_ic(str(f).encode(), N + 1, ptr(np.zeros(3 * N + 3)), ptr(np.zeros(3 * N + 3)),
                                    ptr(np.zeros(N + 1))) < 0
    # a last line without a newline is not counted (measure_N_part_from_ascii_snapshot)
    g = tmp_path / "two.dat"
    g.write_text("1 2 3 4 5 6 7\n1 2 3 4 5 6 7")
    assert host.steps_ascii_count_particles(str(g).encode()) == 1


def gadget_file(path, pos, vel, ids, types, masstab, box, h, time):
    npart = [int((types == k).sum()) for k in range(6)]
    order = np.argsort(types, kind="stable")                      # blocks are in type order
    pos, vel, ids, types = pos[order], vel[order], ids[order], types[order]
    hdr = struct.pack("<6i6d2d2i6i2i4d", *npart, *masstab, time, 1 / time - 1, 0, 0, *npart, 0, 1, box, 0.3, 0.7, h)
    hdr += b"\0" * (256 - len(hdr))
    var = np.concatenate([np.full(npart[k], 1.0 + k) for k in range(6) if masstab[k] == 0]).astype("<f4")

    def rec(b):
        return struct.pack("<i", len(b)) + b + struct.pack("<i", len(b))
    with open(path, "wb") as f:
        f.write(rec(hdr) + rec(pos.astype("<f4").tobytes()) + rec(vel.astype("<f4").tobytes()) +
                rec(ids.astype("<i4").tobytes()) + (rec(var.tobytes()) if var.size else b""))


def test_gadget_ic(host, tmp_path):
    rng = np.random.default_rng(3)
    n = 50
    types = np.array([0] * 7 + [1] * 30 + [2] * 13)
    pos = rng.uniform(0, 100000, size=(n, 3))         # kpc/h
    vel = rng.normal(size=(n, 3)) * 100
    ids = rng.permutation(n) + 1
    f = tmp_path / "ic.gadget"
    for masstab in ([0.0, 0.0, 0.0, 0, 0, 0], [0.0, 2.5, 0.0, 0, 0, 0]):
        gadget_file(f, pos, vel, ids, types, masstab, 100000.0, 0.7, 0.02)
        N, L, a = C.c_int64(), C.c_double(), C.c_double()
        assert host.steps_read_gadget_ic(str(f).encode(), C.byref(N), None, None, None, C.byref(L), C.byref(a)) == 0
        assert N.value == 30 and L.value == pytest.approx(100.0 / 0.7) and a.value == pytest.approx(0.02)
        x, v, M = np.zeros(90), np.zeros(90), np.zeros(30)
        assert host.steps_read_gadget_ic(str(f).encode(), C.byref(N), ptr(x), ptr(v), ptr(M), C.byref(L),
                                         C.byref(a)) == 0
        sel = np.where(types == 1)[0]
        sel = sel[np.argsort(ids[sel])]               # type 1, in ID order
        assert np.allclose(x.reshape(-1, 3), pos[sel].astype(np.float32).astype(np.float64) / 1000.0 / 0.7, rtol=1e-15)
        assert np.array_equal(v.reshape(-1, 3), vel[sel].astype(np.float32).astype(np.float64))
        assert np.all(M == (2.0 if masstab[1] == 0 else 2.5))
    (tmp_path / "bad").write_bytes(b"\x10\x00\x00\x00" + b"\0" * 20)
    assert host.steps_read_gadget_ic(str(tmp_path / "bad").encode(), C.byref(N), None, None, None, None, None) < 0


def test_output_list(host):
    f = str(GOLD / "outredshifts.txt").encode()
    n = host.steps_read_output_list(f, None, 0)
    out = np.zeros(n)
    assert n == 5 and host.steps_read_output_list(f, ptr(out), n) == 5
    assert list(out) == [15.0, 3.0, 1.0, 0.5, 0.0]
