"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

Test infrastructure: the oracle is the checker, never the thing measured.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_DIR = ROOT / "oracle"
LIB = ORACLE_DIR / "liboracle.so"

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)


class OracleParams(C.Structure):
    _fields_ = [
        ("topology", C.c_int),
        ("is_periodic", C.c_int),
        ("cosmology", C.c_int),
        ("comoving", C.c_int),
        ("L", C.c_double),
        ("Rsim", C.c_double),
        ("mass_in_unit_sphere", C.c_double),
        ("H0", C.c_double),
        ("Omega_lambda", C.c_double),
        ("s1r2_lookup", C.c_int),
        ("interp_order", C.c_int),
        ("cut_inclusive", C.c_int),
        ("t3_table", _DP),
        ("t3_ngrid", C.c_int),
        ("s1r2_table", _DP),
        ("s1r2_nrho", C.c_int),
        ("s1r2_nz", C.c_int),
        ("radial_table", _DP),
        ("radial_table_size", C.c_int),
    ]


class OracleBhOpts(C.Structure):
    _fields_ = [("randomize", C.c_int), ("u", C.c_double * 6), ("radial_bh_table", _DP),
                ("radial_bh_table_size", C.c_int), ("com_from_children", C.c_int)]


class OracleCosmo(C.Structure):
    _fields_ = [("H0", C.c_double), ("Omega_m", C.c_double), ("Omega_lambda", C.c_double),
                ("Omega_r", C.c_double), ("Omega_k", C.c_double)]


class OracleStepState(C.Structure):
    _fields_ = [("a", C.c_double), ("hubble", C.c_double), ("h", C.c_double),
                ("rho_part", C.c_double), ("G", C.c_double)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    L = C.CDLL(str(LIB))
    L.oracle_force_softening.restype = C.c_double
    L.oracle_force_softening.argtypes = [C.c_double, C.c_double]
    L.oracle_force_softening_gpu.restype = C.c_double
    L.oracle_force_softening_gpu.argtypes = [C.c_double, C.c_double]
    for name in ("oracle_forces",):
        getattr(L, name).restype = C.c_int
        getattr(L, name).argtypes = [C.POINTER(OracleParams), C.c_int, _DP, _DP, _DP, _DP, C.c_int, C.c_int]
    for name in ("oracle_forces_rows", "oracle_forces_faithful_rows", "oracle_forces_fast_rows",
                 "oracle_forces_r3_fp32_rows"):
        getattr(L, name).restype = C.c_int
        getattr(L, name).argtypes = [C.POINTER(OracleParams), C.c_int, _DP, _DP, _DP, _DP, _IP, C.c_int]
    L.oracle_max_threads.restype = C.c_int
    L.oracle_bh_last_counts.restype = None
    L.oracle_bh_last_counts.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.oracle_forces_bh.restype = C.c_int
    L.oracle_forces_bh.argtypes = [C.POINTER(OracleParams), C.c_double, C.c_double, C.c_int, _DP, _DP, _DP, _DP,
                                   C.c_int, C.c_int]
    L.oracle_forces_bh_ex.restype = C.c_int
    L.oracle_forces_bh_ex.argtypes = [C.POINTER(OracleParams), C.c_double, C.c_double, C.POINTER(OracleBhOpts),
                                      C.c_int, _DP, _DP, _DP, _DP, C.c_int, C.c_int]
    L.oracle_bh_last_root.restype = None
    L.oracle_bh_last_root.argtypes = [_DP]
    L.oracle_radial_bh_correction.restype = C.c_double
    L.oracle_radial_bh_correction.argtypes = [C.c_double, C.c_double, _DP, C.c_int]
    L.oracle_radial_bh_table_accumulate.restype = None
    L.oracle_radial_bh_table_accumulate.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, _DP, _DP, _DP, _IP]
    L.oracle_forces_r3_longdouble.restype = None
    L.oracle_forces_r3_longdouble.argtypes = [C.POINTER(OracleParams), C.c_int, _DP, _DP, _DP,
                                              C.POINTER(C.c_longdouble)]
    L.oracle_ewald_space_count.restype = C.c_int
    L.oracle_ewald_space_count.argtypes = [C.c_double]
    L.oracle_ewald_space.restype = C.c_int
    L.oracle_ewald_space.argtypes = [C.c_double, _IP]
    L.oracle_t3_ewald_correction.restype = None
    L.oracle_t3_ewald_correction.argtypes = [_DP, C.c_double, C.c_double, C.c_double, C.c_double, _DP]
    L.oracle_t3_ewald_table.restype = None
    L.oracle_t3_ewald_table.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _DP]
    L.oracle_t3_params.restype = None
    L.oracle_t3_params.argtypes = [C.c_int, C.c_double, _DP, _DP, _DP, _IP]
    L.oracle_t3_interpolate.restype = None
    L.oracle_t3_interpolate.argtypes = [C.c_int, C.c_double, _DP, C.c_double, C.c_double, C.c_double, _DP,
                                        C.c_int]
    L.oracle_fast_erfcx.restype = C.c_double
    L.oracle_fast_erfcx.argtypes = [C.c_double]
    L.oracle_s1r2_force_pair.restype = None
    L.oracle_s1r2_force_pair.argtypes = [C.c_double] * 6 + [C.c_int, C.c_int, C.c_int, _DP]
    L.oracle_s1r2_force_pair.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_int, C.c_int, C.c_int, _DP]
    L.oracle_s1r2_table.restype = None
    L.oracle_s1r2_table.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                    C.c_int, _DP]
    L.oracle_s1r2_params.restype = None
    L.oracle_s1r2_params.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, _DP, _IP, _IP, _IP, _IP]
    L.oracle_s1r2_interpolate.restype = None
    L.oracle_s1r2_interpolate.argtypes = [_DP, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_double, C.c_int, _DP]
    L.oracle_cylindrical_force_table.restype = None
    L.oracle_cylindrical_force_table.argtypes = [_DP, C.c_double, C.c_double, C.c_int, C.c_int]
    L.oracle_cylindrical_force_correction.restype = C.c_double
    L.oracle_cylindrical_force_correction.argtypes = [C.c_double, C.c_double, _DP, C.c_int]
    L.oracle_softening_lengths.restype = None
    L.oracle_softening_lengths.argtypes = [_DP, C.c_int, C.c_double, _DP, _DP, _DP]
    L.oracle_friedmann_step.restype = C.c_double
    L.oracle_friedmann_step.argtypes = [C.POINTER(OracleCosmo), C.c_double, C.c_double]
    L.oracle_hubble.restype = C.c_double
    L.oracle_hubble.argtypes = [C.POINTER(OracleCosmo), C.c_double]
    L.oracle_kdk_step.restype = C.c_double
    L.oracle_kdk_step.argtypes = [C.POINTER(OracleParams), C.POINTER(OracleCosmo), C.POINTER(OracleStepState),
                                  C.c_int, _DP, _DP, _DP, _DP, _DP]
    L.oracle_initial_errmax.restype = C.c_double
    L.oracle_initial_errmax.argtypes = [C.POINTER(OracleParams), C.POINTER(OracleStepState), C.c_int, _DP, _DP,
                                        _DP, _DP]
    _lib = L
    return L


def dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


def ip(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_IP)


def params(topology=0, *, is_periodic=0, cosmology=0, comoving=0, L=0.0, Rsim=0.0, mass_in_unit_sphere=0.0,
           H0=0.0, Omega_lambda=0.0, s1r2_lookup=1, interp_order=4, cut_inclusive=0, t3_table=None,
           s1r2_table=None, s1r2_shape=(0, 0), radial_table=None):
    p = OracleParams()
    p.topology = topology
    p.is_periodic = is_periodic
    p.cosmology = cosmology
    p.comoving = comoving
    p.L = L
    p.Rsim = Rsim
    p.mass_in_unit_sphere = mass_in_unit_sphere
    p.H0 = H0
    p.Omega_lambda = Omega_lambda
    p.s1r2_lookup = s1r2_lookup
    p.interp_order = interp_order
    p.cut_inclusive = cut_inclusive
    keep = []
    if t3_table is not None:
        t = np.ascontiguousarray(t3_table, dtype=np.float64)
        keep.append(t)
        p.t3_table = dp(t)
        p.t3_ngrid = round((t.size // 3) ** (1 / 3))
    if s1r2_table is not None:
        t = np.ascontiguousarray(s1r2_table, dtype=np.float64)
        keep.append(t)
        p.s1r2_table = dp(t)
        p.s1r2_nrho, p.s1r2_nz = s1r2_shape
    if radial_table is not None:
        t = np.ascontiguousarray(radial_table, dtype=np.float64)
        keep.append(t)
        p.radial_table = dp(t)
        p.radial_table_size = t.size
    p._keep = keep
    return p


def forces(p: OracleParams, x, M, h, id_min=0, id_max=None) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    M = np.ascontiguousarray(M, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    N = M.size
    if id_max is None:
        id_max = N - 1
    n = max(0, id_max - id_min + 1)
    F = np.zeros(3 * max(n, 1))
    rc = lib().oracle_forces(C.byref(p), N, dp(x), dp(M), dp(h), dp(F), id_min, id_max)
    assert rc == 0, rc
    return F[: 3 * n].reshape(n, 3)


def forces_bh(p: OracleParams, theta: float, particle_radius: float, x, M, h, id_min=0, id_max=None) -> np.ndarray:
    """Barnes-Hut octree forces of the reference's USE_BH build (free boundary)."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    M = np.ascontiguousarray(M, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    N = M.size
    if id_max is None:
        id_max = N - 1
    n = max(0, id_max - id_min + 1)
    F = np.zeros(3 * max(n, 1))
    rc = lib().oracle_forces_bh(C.byref(p), theta, particle_radius, N, dp(x), dp(M), dp(h), dp(F), id_min, id_max)
    assert rc == 0, rc
    return F[: 3 * n].reshape(n, 3)


def forces_bh_ex(p: OracleParams, theta: float, particle_radius: float, x, M, h, *, u=None, radial_bh=None,
                 id_min=0, id_max=None, com_from_children=False) -> np.ndarray:
    """Octree forces with the RANDOMIZE_BH draws u (six values in [0, 1], the
    reference's order) and/or the radial correction table radial_bh;
    com_from_children sums node centres of mass from the children instead of
    the reference's running averages (same tree, same walk)."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    M = np.ascontiguousarray(M, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    N = M.size
    if id_max is None:
        id_max = N - 1
    n = max(0, id_max - id_min + 1)
    o = OracleBhOpts()
    o.com_from_children = 1 if com_from_children else 0
    if u is not None:
        o.randomize = 1
        for k, v in enumerate(list(u) + [0.0] * (6 - len(u))):
            o.u[k] = float(v)
    tab = None
    if radial_bh is not None:
        tab = np.ascontiguousarray(radial_bh, dtype=np.float64)
        o.radial_bh_table = dp(tab)
        o.radial_bh_table_size = tab.size
    F = np.zeros(3 * max(n, 1))
    rc = lib().oracle_forces_bh_ex(C.byref(p), theta, particle_radius, C.byref(o), N, dp(x), dp(M), dp(h), dp(F),
                                   id_min, id_max)
    assert rc == 0, rc
    return F[: 3 * n].reshape(n, 3)


def bh_last_root() -> np.ndarray:
    """(centre x, y, z, side) of the last octree call's root cube."""
    out = np.zeros(4)
    lib().oracle_bh_last_root(dp(out))
    return out


def radial_bh_correction(r: float, R: float, table) -> float:
    t = np.ascontiguousarray(table, dtype=np.float64)
    return lib().oracle_radial_bh_correction(float(r), float(R), dp(t), t.size)


def radial_bh_table_accumulate(topology: int, Rsim: float, x, F, table, count) -> None:
    """Adds one evaluation's per-bin radial force sums and counts (in place)."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1)
    assert table.dtype == np.float64 and count.dtype == np.int32 and table.flags.c_contiguous
    lib().oracle_radial_bh_table_accumulate(topology, Rsim, table.size, x.size // 3, dp(x), dp(F), dp(table),
                                            ip(count))


def bh_last_counts() -> tuple[int, int]:
    """(nodes examined, nodes taken) summed over the rows of the last forces_bh call."""
    v, a = C.c_longlong(), C.c_longlong()
    lib().oracle_bh_last_counts(C.byref(v), C.byref(a))
    return v.value, a.value


def forces_rows(p: OracleParams, x, M, h, rows, kind: str = "oracle") -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    M = np.ascontiguousarray(M, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    F = np.zeros(3 * max(rows.size, 1))
    fn = {"oracle": lib().oracle_forces_rows, "faithful": lib().oracle_forces_faithful_rows,
          "fast": lib().oracle_forces_fast_rows, "fp32": lib().oracle_forces_r3_fp32_rows}[kind]
    rc = fn(C.byref(p), M.size, dp(x), dp(M), dp(h), dp(F), ip(rows), rows.size)
    assert rc == 0, rc
    return F[: 3 * rows.size].reshape(-1, 3)


def forces_longdouble(p: OracleParams, x, M, h) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    M = np.ascontiguousarray(M, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    N = M.size
    F = np.zeros(3 * N, dtype=np.longdouble)
    lib().oracle_forces_r3_longdouble(C.byref(p), N, dp(x), dp(M), dp(h),
                                      F.ctypes.data_as(C.POINTER(C.c_longdouble)))
    return F.reshape(N, 3)


def t3_params(is_periodic: int, L: float):
    a, rc, kc = C.c_double(), C.c_double(), C.c_double()
    ng = C.c_int()
    lib().oracle_t3_params(is_periodic, L, C.byref(a), C.byref(rc), C.byref(kc), C.byref(ng))
    return a.value, rc.value, kc.value, ng.value


def t3_table(is_periodic: int, L: float) -> np.ndarray:
    alpha, rel, rec, ng = t3_params(is_periodic, L)
    T = np.zeros(ng ** 3 * 3)
    lib().oracle_t3_ewald_table(ng, L, alpha, rel, rec, dp(T))
    return T


def t3_correction(r, L, alpha, rel_cut, rec_cut) -> np.ndarray:
    r = np.ascontiguousarray(r, dtype=np.float64)
    D = np.zeros(3)
    lib().oracle_t3_ewald_correction(dp(r), L, alpha, rel_cut, rec_cut, dp(D))
    return D


def t3_interp(T, L, d, order=4) -> np.ndarray:
    T = np.ascontiguousarray(T, dtype=np.float64)
    ng = round((T.size // 3) ** (1 / 3))
    D = np.zeros(3)
    lib().oracle_t3_interpolate(ng, L, dp(T), float(d[0]), float(d[1]), float(d[2]), dp(D), order)
    return D


def s1r2_params(is_periodic, L, Rsim, rspace=0):
    a = C.c_double()
    rc, kc, nz, nrho = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    lib().oracle_s1r2_params(is_periodic, L, Rsim, rspace, C.byref(a), C.byref(rc), C.byref(kc), C.byref(nz),
                             C.byref(nrho))
    return a.value, rc.value, kc.value, nz.value, nrho.value


def s1r2_table(is_periodic, L, Rsim, rspace=0):
    alpha, rel, rec, nz, nrho = s1r2_params(is_periodic, L, Rsim, rspace)
    T = np.zeros(nrho * nz * 2)
    lib().oracle_s1r2_table(nrho, nz, 2.25 * Rsim, L, alpha, rel, rec, rspace, dp(T))
    return T, (nrho, nz)


def s1r2_force_pair(d, Lz, alpha, nmax, mmax, rspace=0) -> np.ndarray:
    F = np.zeros(3)
    lib().oracle_s1r2_force_pair(float(d[0]), float(d[1]), float(d[2]), Lz, alpha, nmax, mmax, rspace, dp(F))
    return F


def s1r2_interp(T, shape, rho_max, Lz, d, order) -> np.ndarray:
    T = np.ascontiguousarray(T, dtype=np.float64)
    D = np.zeros(3)
    lib().oracle_s1r2_interpolate(dp(T), shape[0], shape[1], rho_max, Lz, float(d[0]), float(d[1]),
                                  float(d[2]), order, dp(D))
    return D


def radial_table(R, Lz, size, accuracy) -> np.ndarray:
    T = np.zeros(size)
    lib().oracle_cylindrical_force_table(dp(T), R, Lz, size, accuracy)
    return T


def softening_lengths(M, R0):
    M = np.ascontiguousarray(M, dtype=np.float64)
    h = np.zeros_like(M)
    mmin, rp = C.c_double(), C.c_double()
    lib().oracle_softening_lengths(dp(M), M.size, R0, dp(h), C.byref(mmin), C.byref(rp))
    return h, mmin.value, rp.value
