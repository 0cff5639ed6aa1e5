"""ctypes binding of include/steps_force.h (libsteps_amd.so).

This is plumbing for tests, smoke() and bench.py: every call goes straight to
the HIP library.  There is no CPU fallback -- if the library is missing or no
HIP device is present, calls raise StepsError.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIBPATH = Path(__file__).resolve().parent / "lib" / "libsteps_amd.so"

STEPS_OK = 0
STEPS_ERR_ARG = -1
STEPS_ERR_RANGE = -2
STEPS_ERR_TABLE = -3
STEPS_ERR_DEVICE = -4
STEPS_ERR_NOMEM = -5
STEPS_ERR_ADMISSION = -6
STEPS_ERR_STATE = -7
STEPS_ERR_NONFINITE = -8

TOPO_R3, TOPO_T3, TOPO_S1R2 = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1

# every symbol include/steps_force.h declares (checked by tests/test_abi.py)
EXPORTED = [
    "steps_force_desc_init", "steps_device_count", "steps_last_error", "steps_version",
    "steps_force_eval", "steps_force_eval_noalloc", "steps_sim_create", "steps_sim_destroy", "steps_sim_set_precision",
    "steps_sim_set_bh_random", "steps_sim_set_state",
    "steps_sim_set_positions", "steps_sim_get_state", "steps_sim_forces", "steps_sim_get_forces",
    "steps_sim_synchronize", "steps_sim_last_kernel_ms", "steps_sim_step", "steps_sim_errmax",
    "steps_t3_table", "steps_s1r2_table", "steps_radial_table", "steps_radial_bh_table_accumulate",
    "steps_radial_bh_table_finish", "steps_softening_lengths",
    "steps_server_config_init", "steps_server_create", "steps_server_destroy", "steps_server_add_sim",
    "steps_server_remove_sim", "steps_server_sim_bytes", "steps_server_submit", "steps_server_wait", "steps_server_drain",
    "steps_server_get_state", "steps_server_get_stats", "steps_server_clock",
]

REQ_FORCE, REQ_STEP = 0, 1


class StepsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"steps error {code}: {msg}")
        self.code = code


class ForceDesc(C.Structure):
    _fields_ = [
        ("topology", C.c_int32),
        ("is_periodic", C.c_int32),
        ("cosmology", C.c_int32),
        ("comoving", C.c_int32),
        ("L", C.c_double),
        ("Rsim", C.c_double),
        ("mass_in_unit_sphere", C.c_double),
        ("H0", C.c_double),
        ("Omega_lambda", C.c_double),
        ("s1r2_lookup", C.c_int32),
        ("interp_order", C.c_int32),
        ("t3_table", C.POINTER(C.c_double)),
        ("t3_ngrid", C.c_int32),
        ("s1r2_table", C.POINTER(C.c_double)),
        ("s1r2_nrho", C.c_int32),
        ("s1r2_nz", C.c_int32),
        ("radial_table", C.POINTER(C.c_double)),
        ("radial_table_size", C.c_int32),
        ("precision", C.c_int32),
        ("device", C.c_int32),
        ("bh_theta", C.c_double),
        ("bh_particle_radius", C.c_double),
        ("bh_randomize", C.c_int32),
        ("bh_random", C.c_double * 6),
        ("radial_bh_table", C.POINTER(C.c_double)),
        ("radial_bh_table_size", C.c_int32),
    ]


class StepParams(C.Structure):
    _fields_ = [
        ("h", C.c_double),
        ("a_kick1", C.c_double),
        ("hubble1", C.c_double),
        ("a_kick2", C.c_double),
        ("hubble2", C.c_double),
        ("G", C.c_double),
        ("rho_part", C.c_double),
    ]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("streams", C.c_int32), ("use_graphs", C.c_int32),
                ("hbm_budget_bytes", C.c_int64), ("pack_max_n", C.c_int64)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("kind", C.c_int32), ("sim", C.c_int64), ("ticket", C.c_int64),
                ("errmax", C.c_double), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_end", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("submitted", "completed", "failed", "admissions", "evictions",
                                        "graph_captures", "resident_sims", "bytes_resident", "max_backlog",
                                        "packed_launches", "packed_requests", "tickets_held", "admitting")]


_lib = None
_DP = C.POINTER(C.c_double)


def load_library(path: Path | str | None = None) -> C.CDLL:
    """Load libsteps_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIBPATH
    if not p.exists():
        raise StepsError(STEPS_ERR_DEVICE, f"{p} not built; run python -m steps_amd.build")
    lib = C.CDLL(str(p))
    i64, i32, vp = C.c_int64, C.c_int32, C.c_void_p
    sig = {
        "steps_force_desc_init": (None, [C.POINTER(ForceDesc)]),
        "steps_device_count": (C.c_int, []),
        "steps_last_error": (C.c_char_p, []),
        "steps_version": (C.c_char_p, []),
        "steps_force_eval": (C.c_int, [C.POINTER(ForceDesc), i64, _DP, _DP, _DP, _DP, i64, i64]),
        "steps_force_eval_noalloc": (C.c_int, [C.POINTER(ForceDesc), i64, _DP, _DP, _DP, _DP, i64, i64]),
        "steps_sim_create": (C.c_int, [C.POINTER(ForceDesc), i64, C.POINTER(vp)]),
        "steps_sim_set_precision": (C.c_int, [vp, C.c_int32]),
        "steps_sim_set_bh_random": (C.c_int, [vp, _DP]),
        "steps_sim_destroy": (C.c_int, [vp]),
        "steps_sim_set_state": (C.c_int, [vp, _DP, _DP, _DP, _DP]),
        "steps_sim_set_positions": (C.c_int, [vp, _DP]),
        "steps_sim_get_state": (C.c_int, [vp, _DP, _DP]),
        "steps_sim_forces": (C.c_int, [vp, i64, i64]),
        "steps_sim_get_forces": (C.c_int, [vp, _DP]),
        "steps_sim_synchronize": (C.c_int, [vp]),
        "steps_sim_last_kernel_ms": (C.c_int, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "steps_sim_step": (C.c_int, [vp, C.POINTER(StepParams), _DP]),
        "steps_sim_errmax": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double, _DP]),
        "steps_t3_table": (C.c_int, [i32, C.c_double, _DP, C.POINTER(i32)]),
        "steps_s1r2_table": (C.c_int, [i32, C.c_double, C.c_double, _DP, C.POINTER(i32), C.POINTER(i32)]),
        "steps_radial_table": (C.c_int, [C.c_double, C.c_double, i32, i32, _DP]),
        "steps_radial_bh_table_accumulate": (C.c_int, [i32, C.c_double, i32, i64, _DP, _DP, _DP,
                                                       C.POINTER(C.c_int32)]),
        "steps_radial_bh_table_finish": (C.c_int, [i32, _DP, C.POINTER(C.c_int32)]),
        "steps_softening_lengths": (C.c_int, [_DP, i64, C.c_double, _DP, _DP, _DP]),
        "steps_server_config_init": (None, [C.POINTER(ServerConfig)]),
        "steps_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(vp)]),
        "steps_server_destroy": (C.c_int, [vp]),
        "steps_server_add_sim": (C.c_int, [vp, C.POINTER(ForceDesc), i64, _DP, _DP, _DP, _DP, C.POINTER(i64)]),
        "steps_server_remove_sim": (C.c_int, [vp, i64]),
        "steps_server_sim_bytes": (C.c_int64, [vp, i64]),
        "steps_server_submit": (C.c_int, [vp, i64, i32, i32, C.POINTER(StepParams), C.POINTER(i64)]),
        "steps_server_wait": (C.c_int, [vp, i64, C.POINTER(RequestResult)]),
        "steps_server_drain": (C.c_int, [vp]),
        "steps_server_get_state": (C.c_int, [vp, i64, _DP, _DP, _DP]),
        "steps_server_get_stats": (C.c_int, [vp, C.POINTER(ServerStats)]),
        "steps_server_clock": (C.c_double, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != STEPS_OK:
        msg = load_library().steps_last_error()
        raise StepsError(rc, msg.decode() if msg else "")


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "float64 C-contiguous array required"
    return a.ctypes.data_as(_DP)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def device_count() -> int:
    return int(load_library().steps_device_count())


def make_desc(topology=TOPO_R3, *, is_periodic=0, cosmology=0, comoving=0, L=0.0, Rsim=0.0,
              mass_in_unit_sphere=0.0, H0=0.0, Omega_lambda=0.0, s1r2_lookup=1, interp_order=2,
              t3_table=None, s1r2_table=None, s1r2_shape=(0, 0), radial_table=None,
              precision=PREC_FP64, device=0, bh_theta=0.0, bh_particle_radius=0.0, bh_random=None,
              radial_bh_table=None):
    """Build a ForceDesc; returns (desc, keepalive) -- keep the second alive
    as long as the desc is used (it owns the table arrays).  bh_theta > 0
    selects the Barnes-Hut octree (USE_BH) with that opening angle;
    bh_random (six draws in [0, 1]) randomises it (RANDOMIZE_BH) and
    radial_bh_table adds its radial correction."""
    d = ForceDesc()
    load_library().steps_force_desc_init(C.byref(d))
    d.topology = topology
    d.is_periodic = is_periodic
    d.cosmology = cosmology
    d.comoving = comoving
    d.L = L
    d.Rsim = Rsim
    d.mass_in_unit_sphere = mass_in_unit_sphere
    d.H0 = H0
    d.Omega_lambda = Omega_lambda
    d.s1r2_lookup = s1r2_lookup
    d.interp_order = interp_order
    d.precision = precision
    d.device = device
    d.bh_theta = bh_theta
    d.bh_particle_radius = bh_particle_radius
    keep = []
    if t3_table is not None:
        t = f64(t3_table)
        keep.append(t)
        ng = round(((t.size // 3)) ** (1.0 / 3.0))
        assert ng ** 3 * 3 == t.size
        d.t3_table = _ptr(t)
        d.t3_ngrid = ng
    if s1r2_table is not None:
        t = f64(s1r2_table)
        keep.append(t)
        d.s1r2_table = _ptr(t)
        d.s1r2_nrho, d.s1r2_nz = s1r2_shape
        assert d.s1r2_nrho * d.s1r2_nz * 2 == t.size
    if radial_table is not None:
        t = f64(radial_table)
        keep.append(t)
        d.radial_table = _ptr(t)
        d.radial_table_size = t.size
    if bh_random is not None:
        d.bh_randomize = 1
        for k, u in enumerate(list(bh_random) + [0.0] * (6 - len(bh_random))):
            d.bh_random[k] = float(u)
    if radial_bh_table is not None:
        t = f64(radial_bh_table)
        keep.append(t)
        d.radial_bh_table = _ptr(t)
        d.radial_bh_table_size = t.size
    return d, keep


def force_eval(desc: ForceDesc, x, M, soft, id_min: int = 0, id_max: int | None = None) -> np.ndarray:
    """One-shot evaluation through steps_force_eval (host buffers)."""
    x, M, soft = f64(x), f64(M), f64(soft)
    N = M.size
    if id_max is None:
        id_max = N - 1
    nrows = max(0, id_max - id_min + 1)
    F = np.zeros(3 * max(nrows, 1), dtype=np.float64)
    _check(load_library().steps_force_eval(C.byref(desc), N, _ptr(x.reshape(-1)), _ptr(M), _ptr(soft),
                                           _ptr(F), id_min, id_max))
    return F[: 3 * nrows].reshape(nrows, 3)


class Sim:
    """Resident simulation handle (steps_sim)."""

    def __init__(self, desc: ForceDesc, N: int):
        self._lib = load_library()
        h = C.c_void_p()
        _check(self._lib.steps_sim_create(C.byref(desc), N, C.byref(h)))
        self.h = h
        self.N = N
        self._last_rows = 0

    def close(self):
        if self.h:
            self._lib.steps_sim_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, precision: int):
        _check(self._lib.steps_sim_set_precision(self.h, precision))

    def set_state(self, x, M, soft, v=None):
        x, M, soft = f64(x).reshape(-1), f64(M), f64(soft)
        vv = None if v is None else f64(v).reshape(-1)
        _check(self._lib.steps_sim_set_state(self.h, _ptr(x), _ptr(vv), _ptr(M), _ptr(soft)))

    def set_bh_random(self, u):
        uu = f64(list(u) + [0.0] * (6 - len(u)))
        _check(self._lib.steps_sim_set_bh_random(self.h, _ptr(uu)))

    def set_positions(self, x):
        x = f64(x).reshape(-1)
        _check(self._lib.steps_sim_set_positions(self.h, _ptr(x)))

    def get_state(self):
        x = np.empty(3 * self.N)
        v = np.empty(3 * self.N)
        _check(self._lib.steps_sim_get_state(self.h, _ptr(x), _ptr(v)))
        return x.reshape(-1, 3), v.reshape(-1, 3)

    def forces(self, id_min: int = 0, id_max: int | None = None):
        if id_max is None:
            id_max = self.N - 1
        _check(self._lib.steps_sim_forces(self.h, id_min, id_max))
        self._last_rows = max(0, id_max - id_min + 1)

    def get_forces(self) -> np.ndarray:
        F = np.empty(3 * max(self._last_rows, 1))
        _check(self._lib.steps_sim_get_forces(self.h, _ptr(F)))
        return F[: 3 * self._last_rows].reshape(-1, 3)

    def synchronize(self):
        _check(self._lib.steps_sim_synchronize(self.h))

    def last_kernel_ms(self):
        a, b = C.c_float(), C.c_float()
        _check(self._lib.steps_sim_last_kernel_ms(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def step(self, h, a_kick1, hubble1, a_kick2, hubble2, rho_part, G=1.0) -> float:
        p = StepParams(h, a_kick1, hubble1, a_kick2, hubble2, G, rho_part)
        e = C.c_double()
        _check(self._lib.steps_sim_step(self.h, C.byref(p), C.byref(e)))
        self._last_rows = self.N  # F now holds the step's full-range forces
        return e.value

    def errmax(self, a, hubble, rho_part, G=1.0) -> float:
        e = C.c_double()
        _check(self._lib.steps_sim_errmax(self.h, a, hubble, G, rho_part, C.byref(e)))
        return e.value


def t3_table(is_periodic: int, L: float) -> np.ndarray:
    lib = load_library()
    ng = C.c_int32()
    _check(lib.steps_t3_table(is_periodic, L, None, C.byref(ng)))
    T = np.empty(ng.value ** 3 * 3)
    _check(lib.steps_t3_table(is_periodic, L, _ptr(T), C.byref(ng)))
    return T


def s1r2_table(is_periodic: int, L: float, Rsim: float):
    lib = load_library()
    a, b = C.c_int32(), C.c_int32()
    _check(lib.steps_s1r2_table(is_periodic, L, Rsim, None, C.byref(a), C.byref(b)))
    T = np.empty(a.value * b.value * 2)
    _check(lib.steps_s1r2_table(is_periodic, L, Rsim, _ptr(T), C.byref(a), C.byref(b)))
    return T, (a.value, b.value)


def radial_table(R: float, Lz: float, size: int, accuracy: int) -> np.ndarray:
    T = np.empty(size)
    _check(load_library().steps_radial_table(R, Lz, size, accuracy, _ptr(T)))
    return T


def radial_bh_table(topology: int, Rsim: float, size: int, evaluations) -> np.ndarray:
    """The radial octree correction table from (x, F) pairs of octree force
    evaluations of a glass (steps_radial_bh_table_accumulate, _finish)."""
    lib = load_library()
    T = np.zeros(size)
    n = np.zeros(size, dtype=np.int32)
    for x, F in evaluations:
        x, F = f64(x).reshape(-1), f64(F).reshape(-1)
        _check(lib.steps_radial_bh_table_accumulate(topology, Rsim, size, x.size // 3, _ptr(x), _ptr(F), _ptr(T),
                                                    n.ctypes.data_as(C.POINTER(C.c_int32))))
    _check(lib.steps_radial_bh_table_finish(size, _ptr(T), n.ctypes.data_as(C.POINTER(C.c_int32))))
    return T


def softening_lengths(M, particle_radius: float):
    M = f64(M)
    h = np.empty_like(M)
    mmin, rp = C.c_double(), C.c_double()
    _check(load_library().steps_softening_lengths(_ptr(M), M.size, particle_radius, _ptr(h), C.byref(mmin),
                                                  C.byref(rp)))
    return h, mmin.value, rp.value


class Server:
    """Continuous-batching request loop (steps_server)."""

    def __init__(self, device=0, streams=4, use_graphs=True, hbm_budget_bytes=200 << 30, pack_max_n=None):
        self._lib = load_library()
        cfg = ServerConfig()
        self._lib.steps_server_config_init(C.byref(cfg))
        cfg.device, cfg.streams, cfg.use_graphs, cfg.hbm_budget_bytes = device, streams, int(use_graphs), \
            int(hbm_budget_bytes)
        if pack_max_n is not None:
            cfg.pack_max_n = int(pack_max_n)
        h = C.c_void_p()
        _check(self._lib.steps_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.sizes = {}

    def close(self):
        if self.h:
            self._lib.steps_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_sim(self, desc: ForceDesc, x, M, soft, v=None) -> int:
        x, M, soft = f64(x).reshape(-1), f64(M), f64(soft)
        vv = None if v is None else f64(v).reshape(-1)
        sid = C.c_int64()
        _check(self._lib.steps_server_add_sim(self.h, C.byref(desc), M.size, _ptr(x), _ptr(vv), _ptr(M),
                                              _ptr(soft), C.byref(sid)))
        self.sizes[sid.value] = M.size
        return sid.value

    def sim_bytes(self, sid: int) -> int:
        return int(self._lib.steps_server_sim_bytes(self.h, sid))

    def remove_sim(self, sid: int):
        _check(self._lib.steps_server_remove_sim(self.h, sid))

    def submit(self, sid: int, kind: int, precision: int = PREC_FP64, step: StepParams | None = None) -> int:
        t = C.c_int64()
        _check(self._lib.steps_server_submit(self.h, sid, kind, precision,
                                             C.byref(step) if step is not None else None, C.byref(t)))
        return t.value

    def wait(self, ticket: int) -> RequestResult:
        r = RequestResult()
        _check(self._lib.steps_server_wait(self.h, ticket, C.byref(r)))
        return r

    def drain(self):
        _check(self._lib.steps_server_drain(self.h))

    def get_state(self, sid: int):
        n = self.sizes[sid]
        x, v, F = np.empty(3 * n), np.empty(3 * n), np.empty(3 * n)
        _check(self._lib.steps_server_get_state(self.h, sid, _ptr(x), _ptr(v), _ptr(F)))
        return x.reshape(-1, 3), v.reshape(-1, 3), F.reshape(-1, 3)

    def stats(self) -> dict:
        s = ServerStats()
        _check(self._lib.steps_server_get_stats(self.h, C.byref(s)))
        return {n: getattr(s, n) for n, _ in ServerStats._fields_}

    def clock(self) -> float:
        return float(self._lib.steps_server_clock(self.h))
