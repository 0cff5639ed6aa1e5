#!/usr/bin/env python3
"""Benchmark of the direct-summation force path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload r3_2m]

A step is one force evaluation of all N rows against all N sources (N^2
pair interactions, self pair included as in the reference) on particles that
are already resident in HBM.  The default workload is the compactified R^3
configuration at N = 2^21 (seed 1003): the FP64 path is the headline value;
the mixed-precision path, its accuracy against FP64, and the CPU baseline
(the reference-faithful OpenMP restatement on this host's cores, on a bounded
sample of rows) are reported on the same JSON line, beside the Barnes-Hut
octree on the same particles, the periodic variants, the N = 16384
force-plus-KDK-steps case, the serving loop and the host-buffer entry point.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import steps_amd  # noqa: E402
from steps_amd import workloads as wl  # noqa: E402

# Algorithmic FLOP per pair interaction on the dominant r >= beta branch, as
# written in the reference (forces_cuda.cu:533-542, :482-484): beta 1,
# d = xj - xi 3, r = sqrt(dx^2+dy^2+dz^2) 6, r^2, r^3, 1/r^3 3, *M[j] 1,
# F += w*d 6  ->  20.  sqrt and divide count as 1, loop invariants excluded.
FLOP_PER_PAIR_R3 = 20
PEAK_FP64_VECTOR_TFLOPS = 78.6    # MI355X published FP64 vector peak
PEAK_FP32_VECTOR_TFLOPS = 157.3   # MI355X published FP32 vector peak (packed)

# The same count for the periodic variants (reference lines forces_cuda.cu
# :616-635 with :93-167 for T3, :826-843 with :188-455 for S1xR2):
#   S1xR2 CIC  61 = pair 13 + rho 4 + CIC map and weights 28 + Cartesian 4 + combine 12
#   S1xR2 TSC 104 = pair 13 + rho 4 + TSC map and 9 weights 25 + 18 taps x 2 comps 54
#                   - CIC's 28 -> 104 with Cartesian and combine as for CIC
#   S1xR2 NGP  39 = pair 13 + rho 4 + NGP map 6 + Cartesian 4 + combine 12
#   T3 tricubic 555 = pair 13 + grid map 12 + weights 54 + tensor 464 + combine 12
FLOP_PER_PAIR = {"t3_ewald": 555, "s1r2_ngp": 39, "s1r2_cic": 61, "s1r2_tsc": 104}
T3_TAP_BYTES_PER_PAIR = 64 * 3 * 8   # 4^3 table nodes x 3 components x FP64
# The T3 kernel reads its table taps from a window staged in LDS (128 B per
# clock per CU x 256 CUs x 2.4 GHz); taps outside the window go through the
# vector L1 (64 B per clock per CU).
LDS_PEAK_BYTES_PER_S = 128 * 256 * 2.4e9

WORKLOADS = {
    # name: (generator kwargs, description)
    "r3_2m": dict(N=2 ** 21, seed=1003),
    "r3_16k": dict(N=16384, seed=1000),
}


def cpu_baseline(w, target_s: float = 12.0) -> dict:
    """Reference-faithful OpenMP CPU direct sum on a bounded row sample."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib  # the oracle: timed here as the CPU baseline only

    lib = oracle_lib.lib()
    threads = int(lib.oracle_max_threads())
    p = oracle_lib.params(0, cosmology=w.cosmology, comoving=w.comoving,
                          mass_in_unit_sphere=w.mass_in_unit_sphere)
    N = w.N

    def run(nrows, kind):
        rows = (np.arange(nrows, dtype=np.int64) * N // nrows).astype(np.int32)
        t0 = time.perf_counter()
        oracle_lib.forces_rows(p, w.x, w.M, w.soft, rows, kind=kind)
        return time.perf_counter() - t0

    # grow the row sample until it takes about target_s (thread start-up
    # makes small samples look slower per row than they are)
    nrows = max(16, threads * 2)
    t = run(nrows, "faithful")
    for _ in range(4):
        if t >= 0.6 * target_s or nrows >= N:
            break
        nrows = int(min(N, max(nrows + 1, nrows * target_s / max(t, 1e-3))))
        t = run(nrows, "faithful")
    rate = nrows * N / t
    tf = run(nrows, "fast")
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": rate / 1e9,
        "unit": "G pair-interactions/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{nrows} strided target rows x all {N} sources (R3, FP64, omp dynamic + atomics, "
                  f"-O3 -ffast-math), {t:.1f} s",
        "cpu_model": cpu_model,
        "nproc": os.cpu_count(),
        "optimised_variant_value": nrows * N / tf / 1e9,
        "configs": cpu_configs(lib, oracle_lib),
    }


def cpu_configs(lib, oracle_lib, target_s: float = 3.0) -> dict:
    """The same faithful OpenMP restatement on the other configurations:
    the R3 N = 16384 case in full (median of three runs), and strided row
    samples of the T3 64^3 lattice (63^3 Ewald table, the CPU build's
    default trilinear interpolation, EWALD_INTERPOLATION_ORDER 2) and of the
    S1xR2 cylinder (N = 524288, CIC table), each sized to about target_s."""
    threads = int(lib.oracle_max_threads())
    out = {}
    w0 = wl.r3_compactified(**WORKLOADS["r3_16k"])
    p0 = oracle_lib.params(0, cosmology=w0.cosmology, comoving=w0.comoving,
                           mass_in_unit_sphere=w0.mass_in_unit_sphere)
    rows = np.arange(w0.N, dtype=np.int32)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle_lib.forces_rows(p0, w0.x, w0.M, w0.soft, rows, kind="faithful")
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    out["r3_16k"] = {"value": w0.N * w0.N / t / 1e9, "unit": "G pair-interactions/s", "cores": threads,
                     "seconds_per_evaluation": t, "sample": f"all {w0.N} rows, median of 3"}
    t3 = wl.t3_lattice()
    T = steps_amd.t3_table(2, t3.L)
    c = wl.s1r2_cylinder()
    Ts, shape = steps_amd.s1r2_table(2, c.L, c.Rsim)
    cases = [("t3_64cubed", t3, oracle_lib.params(1, is_periodic=2, L=t3.L, t3_table=T, interp_order=2),
              "T3 Ewald (63^3 table, trilinear)"),
             ("s1r2_cic_524288", c, oracle_lib.params(2, is_periodic=2, L=c.L, Rsim=c.Rsim, cosmology=1, comoving=1,
                                                      mass_in_unit_sphere=c.mass_in_unit_sphere, interp_order=2,
                                                      s1r2_table=Ts, s1r2_shape=shape),
              "S1xR2 CIC table")]
    for name, w, p, what in cases:
        nrows = max(16, threads * 2)
        for _ in range(4):
            r = (np.arange(nrows, dtype=np.int64) * w.N // nrows).astype(np.int32)
            t0 = time.perf_counter()
            oracle_lib.forces_rows(p, w.x, w.M, w.soft, r, kind="faithful")
            t = time.perf_counter() - t0
            if t >= 0.6 * target_s or nrows >= w.N:
                break
            nrows = int(min(w.N, max(nrows + 1, nrows * target_s / max(t, 1e-3))))
        out[name] = {"value": nrows * w.N / t / 1e9, "unit": "G pair-interactions/s", "cores": threads,
                     "sample": f"{nrows} strided rows x all {w.N} sources, {what}, {t:.1f} s"}
    return out


def time_path(sim, steps: int, warmup: int):
    for _ in range(warmup):
        sim.forces()
        sim.synchronize()
    sim.synchronize()
    pair_ms = []
    total_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.forces()
        a, b = sim.last_kernel_ms()   # HIP events on the sim's own stream
        pair_ms.append(a)
        total_ms.append(b)
    sim.synchronize()
    wall = time.perf_counter() - t0
    return wall, float(np.mean(pair_ms)), float(np.mean(total_ms))


def copy_calibration() -> int:
    """One elementwise copy of 2^27 doubles (1 GiB read, 1 GiB written): under
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE its counter values per byte scale
    the pair kernel's counters to HBM bytes."""
    import torch

    n = 1 << 27
    a = torch.full((n,), 1.0, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "check": float(b[n - 1].item())}))
    return 0


def load_traffic(name: str):
    f = ROOT / "profiles" / f"{name}_traffic.json"
    if f.exists():
        try:
            return json.loads(f.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            return None
    return None


def serving_leg(nreq: int) -> dict:
    """64 resident simulations (R^3 and S^1 x R^2 alternating, N log-uniform in
    [8192, 524288]) receive a seeded Poisson stream (200 req/s) of KDK step
    requests, half FP64 and half mixed precision;
    served by the continuous-batching server and, for comparison, one
    request at a time on resident simulations.  The final (x, v) of every
    simulation must be bitwise equal between the two."""
    from steps_amd import serving as sv
    sims = sv.make_ensemble()
    reqs = sv.make_requests(sims, nreq=nreq)
    print(f"[bench] serving: {len(sims)} sims, {nreq} requests", file=sys.stderr, flush=True)
    srv, fin_s = sv.run_server(sims, reqs)
    print(f"[bench] serving: server done {srv}", file=sys.stderr, flush=True)
    seq, fin_q = sv.run_sequential(sims, reqs)
    print(f"[bench] serving: sequential done {seq}", file=sys.stderr, flush=True)
    same = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) for a, b in zip(fin_s, fin_q))
    pairs = float(sum(sims[r.sim][0].N ** 2 for r in reqs))
    # At 200 req/s the stream offers ~15x what the device serves, so latency
    # is queueing time and grows with the stream's length: the same stream cut
    # at shorter lengths shows by how much.  Latency under a load the device
    # can sustain comes from the same ensemble at 70% of the measured
    # capacity, over 20 s of arrivals.
    by_len = []
    for n in (nreq // 4, nreq // 2):
        r, _ = sv.run_server(sims, reqs[:n])
        by_len.append({k: r[k] for k in ("requests", "wall_s", "p50_ms", "p99_ms", "max_backlog")})
        print(f"[bench] serving: {n} requests {by_len[-1]}", file=sys.stderr, flush=True)
    by_len.append({k: srv[k] for k in ("requests", "wall_s", "p50_ms", "p99_ms", "max_backlog")})
    rate = 0.7 * srv["requests_per_s"]
    sus_reqs = sv.make_requests(sims, nreq=max(1, int(rate * 20.0)), rate=rate, seed=1005)
    sus, _ = sv.run_server(sims, sus_reqs)
    print(f"[bench] serving: sustainable {rate:.1f} req/s {sus}", file=sys.stderr, flush=True)
    sus["offered_req_per_s"] = rate
    sus["arrivals_s"] = sus_reqs[-1].t_arrival
    # how much any schedule could gain: the requests' kernel time alone,
    # against the sequential wall clock
    dev = sv.device_time_alone(sims, reqs)
    dev["sequential_wall_ms"] = seq["wall_s"] * 1e3
    dev["speedup_bound_if_kernels_do_not_overlap"] = seq["wall_s"] * 1e3 / dev["kernel_ms_sum"]
    print(f"[bench] serving: kernel time alone {dev}", file=sys.stderr, flush=True)
    return {"workload": "64 sims R3/S1R2, N log-uniform [8192, 524288], Poisson 200 req/s seed 1004, "
                        "KDK steps, FP64 / mixed 50/50",
            "requests": nreq, "arrivals_s": reqs[-1].t_arrival, "pair_interactions": pairs,
            "steps_per_s": srv["requests_per_s"], "p50_ms": srv["p50_ms"], "p99_ms": srv["p99_ms"],
            "server": srv, "sequential": seq,
            "speedup_vs_sequential": srv["requests_per_s"] / seq["requests_per_s"],
            "results_identical": bool(same),
            "latency_vs_stream_length": by_len,
            "sustainable_load": sus,
            "device_time": dev}


def variants_leg(reps: int) -> dict:
    """The periodic force variants on their measurement configurations: T3
    with the tricubic Ewald table and with the nearest image only (64^3
    perturbed lattice, seed 1001); S1xR2 with CIC and TSC table
    interpolation, with the nearest z-image and with the image sum
    (N = 524288, seed 1002); FP64, plus the mixed path where one exists; one
    warm-up, then `reps` evaluations timed with HIP events."""
    out = {}
    t3 = wl.t3_lattice()
    c = wl.s1r2_cylinder()
    # table builds on the device (each the first of its parameters in this
    # process: the library caches built tables by parameters)
    tables = {}
    for ip in (2, 3, 4):
        t0 = time.perf_counter()
        Tt = steps_amd.t3_table(ip, t3.L)
        tables[f"t3_is_periodic_{ip}"] = {"seconds": time.perf_counter() - t0,
                                          "ngrid": round((Tt.size // 3) ** (1 / 3)), "bytes": Tt.nbytes}
        if ip == 2:
            T = Tt
        elif ip == 3:
            T127 = Tt
        t0 = time.perf_counter()
        Tc, shp = steps_amd.s1r2_table(ip, c.L, c.Rsim)
        tables[f"s1r2_is_periodic_{ip}"] = {"seconds": time.perf_counter() - t0, "nrho_nz": list(shp),
                                            "bytes": Tc.nbytes}
        if ip == 2:
            Ts, shape = Tc, shp
    t0 = time.perf_counter()
    steps_amd.radial_table(c.Rsim, 0.5 * c.L, 1000, 1000)
    tables["radial_1000"] = {"seconds": time.perf_counter() - t0}
    out["table_builds"] = tables
    print(f"[bench] table builds: {tables}", file=sys.stderr, flush=True)
    d, keep = steps_amd.make_desc(steps_amd.TOPO_T3, is_periodic=2, L=t3.L, t3_table=T)
    cases = [("t3_ewald", t3, d, [T, keep])]
    for name, order in (("s1r2_cic", 2), ("s1r2_tsc", 4)):
        ds, ks = steps_amd.make_desc(steps_amd.TOPO_S1R2, is_periodic=2, L=c.L, Rsim=c.Rsim, cosmology=1,
                                     comoving=1, mass_in_unit_sphere=c.mass_in_unit_sphere, s1r2_lookup=1,
                                     interp_order=order, s1r2_table=Ts, s1r2_shape=shape)
        cases.append((name, c, ds, [Ts, ks]))
    # the other modes of the same rows: nearest image only (IS_PERIODIC = 1)
    # and the S1xR2 image sum without a table; both use the cylindrical
    # radial correction table with the reference's parameters
    # (main.cc:1302 and :1282: R = Rsim, Lz = L/2 or (P + 0.6) L, 1000 nodes)
    d1, k1 = steps_amd.make_desc(steps_amd.TOPO_T3, is_periodic=1, L=t3.L)
    cases.append(("t3_nearest", t3, d1, [k1]))
    rad1 = steps_amd.radial_table(c.Rsim, 0.5 * c.L, 1000, 1000)
    dn, kn = steps_amd.make_desc(steps_amd.TOPO_S1R2, is_periodic=1, L=c.L, Rsim=c.Rsim, cosmology=1, comoving=1,
                                 mass_in_unit_sphere=c.mass_in_unit_sphere, radial_table=rad1)
    cases.append(("s1r2_nearest", c, dn, [rad1, kn]))
    rad2 = steps_amd.radial_table(c.Rsim, c.L * (2 + 1 - 0.4), 1000, 1000)
    di, ki = steps_amd.make_desc(steps_amd.TOPO_S1R2, is_periodic=2, L=c.L, Rsim=c.Rsim, cosmology=1, comoving=1,
                                 mass_in_unit_sphere=c.mass_in_unit_sphere, s1r2_lookup=0, radial_table=rad2)
    cases.append(("s1r2_images", c, di, [rad2, ki]))
    for name, w, desc, _keep in cases:
        sim = steps_amd.Sim(desc, w.N)
        sim.set_state(w.x, w.M, w.soft, w.v)
        wall, pair_ms, total_ms = time_path(sim, reps, 1)
        F64 = sim.get_forces()
        pairs = float(w.N) ** 2
        r = {"N": w.N, "kernel_ms": pair_ms, "eval_ms": total_ms,
             "g_pairs_per_s": pairs / (pair_ms * 1e-3) / 1e9}
        flop = FLOP_PER_PAIR.get(name)
        if flop is not None:
            r.update({"flop_per_pair": flop, "achieved_tflops": flop * pairs / (pair_ms * 1e-3) / 1e12,
                      "frac_fp64_peak": flop * pairs / (pair_ms * 1e-3) / 1e12 / PEAK_FP64_VECTOR_TFLOPS})
        if name == "s1r2_images":
            r["images_per_pair"] = 2 * 3 + 1  # m in [-(P+1), P+1], P = IS_PERIODIC = 2 (forces_cuda.cu:659-660)
        if name == "t3_ewald":
            r["tap_bytes_per_s"] = T3_TAP_BYTES_PER_PAIR * pairs / (pair_ms * 1e-3)
            r["tap_frac_of_lds_peak"] = r["tap_bytes_per_s"] / LDS_PEAK_BYTES_PER_S
        if name in ("t3_ewald", "t3_nearest", "s1r2_cic", "s1r2_tsc", "s1r2_nearest"):
            # FP32-compensated path on the same state: speed and error against FP64
            sim.set_precision(steps_amd.PREC_MIXED)
            _, pair_m, _ = time_path(sim, reps, 1)
            err = np.linalg.norm(sim.get_forces() - F64, axis=1) / np.linalg.norm(F64, axis=1)
            r["mixed"] = {"kernel_ms": pair_m, "g_pairs_per_s": pairs / (pair_m * 1e-3) / 1e9,
                          "speedup_vs_fp64": pair_ms / pair_m,
                          "rel_err_vs_fp64_p50": float(np.median(err)),
                          "rel_err_vs_fp64_p99": float(np.percentile(err, 99)),
                          "rel_err_vs_fp64_max": float(np.max(err))}
        sim.close()
        out[name] = r
        print(f"[bench] variant {name}: {r}", file=sys.stderr, flush=True)
    # T3 with the 127^3 table of IS_PERIODIC = 3 (the global-table kernel; the
    # LDS window needs Ng <= 119), on a 40^3 lattice to bound the run time
    w40 = wl.t3_lattice(n_side=40, seed=1001, is_periodic=3)
    T40 = steps_amd.t3_table(3, w40.L) if abs(w40.L - t3.L) > 0 else T127
    d3, k3 = steps_amd.make_desc(steps_amd.TOPO_T3, is_periodic=3, L=w40.L, t3_table=T40)
    sim = steps_amd.Sim(d3, w40.N)
    sim.set_state(w40.x, w40.M, w40.soft, w40.v)
    _, pair_ms, _ = time_path(sim, 1, 1)
    sim.close()
    out["t3_ewald_127"] = {"N": w40.N, "kernel_ms": pair_ms, "g_pairs_per_s": float(w40.N) ** 2 / (pair_ms * 1e-3) / 1e9}
    print(f"[bench] variant t3_ewald_127: {out['t3_ewald_127']}", file=sys.stderr, flush=True)
    return out


def c0_kdk_leg(nsteps: int = 10) -> dict:
    """Compactified R3 at N = 16384 (seed 1000): one force evaluation, then
    `nsteps` KDK steps on the resident simulation (step.cc:100-312: kick,
    drift, force, kick; each step returns errmax to the host, which the next
    time step needs), FP64 and mixed; wall clock per step."""
    w = wl.r3_compactified(N=16384, seed=1000)
    kw = dict(cosmology=w.cosmology, comoving=w.comoving, mass_in_unit_sphere=w.mass_in_unit_sphere)
    hub = lambda a: wl.H0 * math.sqrt(wl.OMEGA_M * a ** -3 + wl.OMEGA_L)
    out = {"N": w.N, "kdk_steps": nsteps}
    for name, prec in (("fp64", steps_amd.PREC_FP64), ("mixed", steps_amd.PREC_MIXED)):
        d, keep = steps_amd.make_desc(steps_amd.TOPO_R3, precision=prec, **kw)
        sim = steps_amd.Sim(d, w.N)
        sim.set_state(w.x, w.M, w.soft, w.v)
        sim.forces()
        sim.synchronize()
        t0 = time.perf_counter()
        sim.forces()
        sim.synchronize()
        t_force = time.perf_counter() - t0
        t0 = time.perf_counter()
        for k in range(nsteps):
            a1 = w.a_start * (1 + 1e-3 * k)
            a2 = w.a_start * (1 + 1e-3 * (k + 1))
            sim.step(1e-4, a1, hub(a1), a2, hub(a2), w.rho_part)
        dt = time.perf_counter() - t0
        sim.close()
        out[name] = {"force_ms": t_force * 1e3, "ms_per_step": dt * 1e3 / nsteps, "steps_per_s": nsteps / dt,
                     "g_pairs_per_s": float(w.N) ** 2 * nsteps / dt / 1e9}
    # the mixed path's error against FP64 on the pair sum (external term off)
    F = {}
    for prec in (steps_amd.PREC_FP64, steps_amd.PREC_MIXED):
        d, keep = steps_amd.make_desc(steps_amd.TOPO_R3, precision=prec)
        F[prec] = steps_amd.force_eval(d, w.x, w.M, w.soft)
    e = (np.linalg.norm(F[steps_amd.PREC_MIXED] - F[steps_amd.PREC_FP64], axis=1) /
         np.linalg.norm(F[steps_amd.PREC_FP64], axis=1))
    out["mixed"].update({"rel_err_pairsum_vs_fp64_p50": float(np.median(e)),
                         "rel_err_pairsum_vs_fp64_p99": float(np.percentile(e, 99))})
    print(f"[bench] c0 KDK: {out}", file=sys.stderr, flush=True)
    return out


def bh_leg(w, kw, F64, direct_ms, reps: int, cpu: bool = True, periodic: bool = True) -> dict:
    """Barnes-Hut octree forces (the reference's USE_BH engine, forces.cc:
    378-506, which the reference runs on the CPU only) on the headline
    workload: octree build plus traversal per evaluation (HIP events), the
    error against the device's FP64 direct sum, and the oracle's
    single-thread restatement of the reference's pointer octree on the same
    particles (a full tree build, then a row sample; the traversal time is
    scaled to all rows)."""
    out = {"N": w.N}
    N = w.N
    for theta in (0.5, 0.8):
        d, keep = steps_amd.make_desc(steps_amd.TOPO_R3, bh_theta=theta, bh_particle_radius=w.particle_radius, **kw)
        sim = steps_amd.Sim(d, N)
        sim.set_state(w.x, w.M, w.soft, w.v)
        _, ms, _ = time_path(sim, reps, 1)
        F = sim.get_forces()
        sim.close()
        out[f"theta_{theta}"] = {"eval_ms": ms, "evaluations_per_s": 1e3 / ms}
        if F64 is not None:
            err = np.linalg.norm(F - F64, axis=1) / np.linalg.norm(F64, axis=1)
            out[f"theta_{theta}"].update({"speedup_vs_direct_fp64": direct_ms / ms,
                                          "rel_err_vs_direct_p50": float(np.median(err)),
                                          "rel_err_vs_direct_p99": float(np.percentile(err, 99))})
    if not cpu:
        return out
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib  # the oracle: timed here as the CPU baseline only

    p = oracle_lib.params(0, **kw)
    t0 = time.perf_counter()
    oracle_lib.forces_bh(p, 0.5, w.particle_radius, w.x, w.M, w.soft, 0, 0)
    t_build = time.perf_counter() - t0
    rows = 65536
    t0 = time.perf_counter()
    oracle_lib.forces_bh(p, 0.5, w.particle_radius, w.x, w.M, w.soft, 0, rows - 1)
    t_rows = time.perf_counter() - t0
    visited, accepted = oracle_lib.bh_last_counts()
    t_eval = t_build + max(t_rows - t_build, 0.0) * N / rows
    # node visits and interactions per row at theta 0.5 (the same for every
    # engine that makes the reference's decisions), from the sampled rows
    g = out["theta_0.5"]
    g["nodes_examined_per_row"] = visited / rows
    g["interactions_per_row"] = accepted / rows
    g["g_interactions_per_s"] = accepted / rows * N / (g["eval_ms"] * 1e-3) / 1e9
    threads = int(oracle_lib.lib().oracle_max_threads())
    out["cpu_baseline"] = {"value": 1.0 / t_eval, "unit": "evaluations/s", "cores": threads, "kind": "port",
                           "seconds_per_evaluation": t_eval,
                           "sample": f"theta 0.5: pointer-octree build over all {N} particles, one thread "
                                     f"({t_build:.2f} s), rows 0..{rows - 1} traversed on {threads} OpenMP threads "
                                     f"({t_rows - t_build:.2f} s), scaled to {N} rows"}
    out["theta_0.5"]["speedup_vs_cpu_port"] = t_eval * 1e3 / out["theta_0.5"]["eval_ms"]
    if periodic:
        out["periodic"] = bh_periodic(reps)
    print(f"[bench] barnes-hut: {out}", file=sys.stderr, flush=True)
    return out


def bh_periodic(reps: int) -> dict:
    """The octree on the periodic measurement configurations at theta 0.5
    (forces.cc:582-772 T3, 881-1217 S1R2): T3 Ewald with the Catmull-Rom
    table interpolation (the CPU build's order 4, the direct GPU kernel's
    interpolation) on the 64^3 lattice, and S1xR2 with the CIC table on the
    N = 524288 cylinder; evaluation time and error against the FP64 direct
    sum of the same configuration."""
    out = {}
    t3 = wl.t3_lattice()
    T = steps_amd.t3_table(2, t3.L)
    c = wl.s1r2_cylinder()
    Ts, shape = steps_amd.s1r2_table(2, c.L, c.Rsim)
    kw = dict(L=c.L, Rsim=c.Rsim, cosmology=1, comoving=1, mass_in_unit_sphere=c.mass_in_unit_sphere)
    cases = [("t3_ewald_order4", t3, steps_amd.TOPO_T3, dict(is_periodic=2, L=t3.L, t3_table=T, interp_order=4)),
             ("s1r2_cic", c, steps_amd.TOPO_S1R2, dict(is_periodic=2, s1r2_table=Ts, s1r2_shape=shape,
                                                        interp_order=2, **kw))]
    for name, w, topo, dkw in cases:
        res = {"N": w.N}
        for tag, extra in (("direct", {}), ("bh", dict(bh_theta=0.5, bh_particle_radius=w.particle_radius))):
            d, keep = steps_amd.make_desc(topo, **dkw, **extra)
            sim = steps_amd.Sim(d, w.N)
            sim.set_state(w.x, w.M, w.soft, w.v)
            _, ms, _ = time_path(sim, reps if tag == "bh" else 1, 1)
            res[tag + "_ms"] = ms
            res[tag + "_F"] = sim.get_forces()
            sim.close()
        err = np.linalg.norm(res.pop("bh_F") - res["direct_F"], axis=1) / np.linalg.norm(res.pop("direct_F"), axis=1)
        res.update({"speedup_vs_direct_fp64": res["direct_ms"] / res["bh_ms"],
                    "rel_err_vs_direct_p50": float(np.median(err)),
                    "rel_err_vs_direct_p99": float(np.percentile(err, 99))})
        out[name] = res
    return out


def host_path_leg(w, desc, calls: int) -> dict:
    """The host-buffer entry point (steps_force_eval, the reference routine's
    boundary): host positions, masses and softening lengths in, forces out,
    PCIe copies included; the first call allocates and uploads, then `calls`
    calls are timed with the wall clock."""
    steps_amd.force_eval(desc, w.x, w.M, w.soft)
    t0 = time.perf_counter()
    for _ in range(calls):
        steps_amd.force_eval(desc, w.x, w.M, w.soft)
    dt = (time.perf_counter() - t0) / calls
    return {"N": w.N, "calls": calls, "ms_per_call": dt * 1e3,
            "g_pairs_per_s": float(w.N) ** 2 / dt / 1e9}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="r3_2m", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-variants", action="store_true", help="skip the T3 / S1xR2 variant leg")
    ap.add_argument("--no-bh", action="store_true", help="skip the Barnes-Hut octree leg")
    ap.add_argument("--no-host-path", action="store_true",
                    help="skip the host-buffer entry-point leg (profiling runs: its N = 16384 launches share kernel names)")
    ap.add_argument("--serving-requests", type=int, default=160,
                    help="length of the seeded Poisson request stream of the serving leg")
    ap.add_argument("--only", choices=["variants", "serving", "barnes-hut"],
                    help="run only that leg and print its JSON (profiling runs); skips the headline line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run one plain 1 GiB device copy (scale reference for rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the engine runs the full N x N interaction on a single "
              "MI355X; sharding across GPUs (row ranges per GPU, positions all-gathered each step) "
              "is not implemented", file=sys.stderr)
        return 2

    if args.only == "variants":
        print(json.dumps({"variants": variants_leg(max(1, min(args.steps, 3)))}), flush=True)
        return 0
    if args.only == "serving":
        print(json.dumps({"serving": serving_leg(args.serving_requests)}), flush=True)
        return 0
    if args.only == "barnes-hut":  # the octree kernels alone (profiling runs): no direct sum, no CPU leg
        w = wl.r3_compactified(**WORKLOADS[args.workload])
        kw = dict(cosmology=w.cosmology, comoving=w.comoving, mass_in_unit_sphere=w.mass_in_unit_sphere)
        print(json.dumps({"barnes_hut": bh_leg(w, kw, None, None, max(1, args.steps), cpu=False,
                                               periodic=False)}), flush=True)
        return 0

    gen = WORKLOADS[args.workload]
    w = wl.r3_compactified(**gen)
    N = w.N
    kw = dict(cosmology=w.cosmology, comoving=w.comoving, mass_in_unit_sphere=w.mass_in_unit_sphere)
    d64, k64 = steps_amd.make_desc(steps_amd.TOPO_R3, precision=steps_amd.PREC_FP64, **kw)
    sim = steps_amd.Sim(d64, N)
    sim.set_state(w.x, w.M, w.soft, w.v)
    wall, pair_ms, total_ms = time_path(sim, args.steps, args.warmup)
    ms_per_step = wall * 1e3 / args.steps
    pairs = float(N) * float(N)
    value = pairs * args.steps / wall / 1e9
    achieved = FLOP_PER_PAIR_R3 * pairs / (pair_ms * 1e-3) / 1e12
    F64 = sim.get_forces()
    sim.close()

    out = {
        "metric": "G pair-interactions/s + % vector peak at N=2^21; serving p50/p99 step latency",
        "value": value,
        "unit": "G pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"r3_compactified N={N} seed={gen['seed']} (open R^3, L=100 Mpc, "
                               f"omega_max=2 atan 10, m~(1+r^2/L^2)^3, h=0.05 m^(1/3), comoving LCDM)",
                   "N": N, "precision": "fp64", "pair_interactions_per_step": pairs},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_VECTOR_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_VECTOR_TFLOPS,
                     "traffic": load_traffic(f"r3_fp64_n{N}"),
                     "flop_per_pair": FLOP_PER_PAIR_R3, "kernel_ms": pair_ms, "eval_ms": total_ms},
    }

    # host-buffer entry point (PCIe included): never `value`
    if not args.no_host_path:
        w0 = wl.r3_compactified(**WORKLOADS["r3_16k"])
        d0, k0 = steps_amd.make_desc(steps_amd.TOPO_R3, precision=steps_amd.PREC_FP64, cosmology=w0.cosmology,
                                     comoving=w0.comoving, mass_in_unit_sphere=w0.mass_in_unit_sphere)
        out["host_path"] = {"r3_headline": host_path_leg(w, d64, 1), "r3_16k": host_path_leg(w0, d0, 20)}

    if not args.no_mixed:
        dmx, kmx = steps_amd.make_desc(steps_amd.TOPO_R3, precision=steps_amd.PREC_MIXED, **kw)
        sim = steps_amd.Sim(dmx, N)
        sim.set_state(w.x, w.M, w.soft, w.v)
        wall_m, pair_m, total_m = time_path(sim, args.steps, args.warmup)
        Fm = sim.get_forces()
        sim.close()
        err = np.linalg.norm(Fm - F64, axis=1) / np.linalg.norm(F64, axis=1)
        # the comoving external term dominates |F| for this mass normalisation,
        # so the error of the pair sum alone (cosmology off) is reported too
        errs = {}
        for prec in (steps_amd.PREC_FP64, steps_amd.PREC_MIXED):
            d0, k0 = steps_amd.make_desc(steps_amd.TOPO_R3, precision=prec)
            s0 = steps_amd.Sim(d0, N)
            s0.set_state(w.x, w.M, w.soft, w.v)
            s0.forces()
            errs[prec] = s0.get_forces()
            s0.close()
        e0 = (np.linalg.norm(errs[steps_amd.PREC_MIXED] - errs[steps_amd.PREC_FP64], axis=1) /
              np.linalg.norm(errs[steps_amd.PREC_FP64], axis=1))
        ach_m = FLOP_PER_PAIR_R3 * pairs / (pair_m * 1e-3) / 1e12
        out["mixed"] = {
            "value": pairs * args.steps / wall_m / 1e9,
            "ms_per_step": wall_m * 1e3 / args.steps,
            "kernel_ms": pair_m,
            "speedup_vs_fp64": (pairs / wall_m) / (pairs / wall),
            "rel_err_vs_fp64_p50": float(np.median(err)),
            "rel_err_vs_fp64_p99": float(np.percentile(err, 99)),
            "rel_err_pairsum_vs_fp64_p50": float(np.median(e0)),
            "rel_err_pairsum_vs_fp64_p99": float(np.percentile(e0, 99)),
            "rel_err_pairsum_vs_fp64_max": float(np.max(e0)),
            "roofline": {"bound": "valu", "achieved": ach_m, "peak": PEAK_FP32_VECTOR_TFLOPS,
                         "unit": "TFLOP/s", "frac": ach_m / PEAK_FP32_VECTOR_TFLOPS},
        }

    if not args.no_bh:
        out["barnes_hut"] = bh_leg(w, kw, F64, pair_ms, max(1, min(args.steps, 3)))

    if not args.no_variants:
        out["variants"] = variants_leg(max(1, min(args.steps, 3)))
        out["c0_kdk"] = c0_kdk_leg()

    if not args.no_serving:
        out["serving"] = serving_leg(args.serving_requests)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
    out["host"] = platform.node()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
