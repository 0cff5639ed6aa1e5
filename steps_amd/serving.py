"""Serving workload: an ensemble of independent simulations receiving a seeded
Poisson stream of force / step requests, served either by the continuous-
batching server (steps_server) or sequentially, one simulation per launch.

Ensemble (seed 1004): `nsims` simulations alternating R^3 and S^1 x R^2, N
log-uniform in [n_min, n_max] from the R^3 / S^1 x R^2 generators; each
request picks a simulation uniformly, is a KDK step (or, with probability
1 - step_frac, a force evaluation), and asks for FP64 or mixed precision with
equal probability (mixed only where the topology has a mixed path).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from . import _abi as abi
from . import workloads as wl


@dataclass
class Req:
    t_arrival: float
    sim: int
    kind: int          # abi.REQ_FORCE / abi.REQ_STEP
    precision: int
    step: int          # per-simulation step counter (selects a, H)


def a_of_t(t: float) -> float:
    """Flat LCDM scale factor at cosmic time t (code units), closed form."""
    return (wl.OMEGA_M / wl.OMEGA_L) ** (1 / 3) * math.sinh(1.5 * math.sqrt(wl.OMEGA_L) * wl.H0 * t) ** (2 / 3)


def t_of_a(a: float) -> float:
    return 2.0 / (3.0 * math.sqrt(wl.OMEGA_L) * wl.H0) * math.asinh(math.sqrt(wl.OMEGA_L / wl.OMEGA_M * a ** 3))


def hubble(a: float) -> float:
    return wl.H0 * math.sqrt(wl.OMEGA_M * a ** -3 + wl.OMEGA_L)


def step_params(w, k: int, h: float = 2e-5) -> abi.StepParams:
    """The k-th step of a simulation that started at a_start with step size h."""
    t0 = t_of_a(w.a_start)
    a1, a2 = a_of_t(t0 + k * h), a_of_t(t0 + (k + 1) * h)
    return abi.StepParams(h, a1, hubble(a1), a2, hubble(a2), 1.0, w.rho_part)


def make_ensemble(nsims: int = 64, seed: int = 1004, n_min: int = 8192, n_max: int = 524288):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(n_min), math.log(n_max), nsims)).astype(int)
    sims = []
    s1r2_tab = None
    for k, N in enumerate(sizes):
        if k % 2 == 0:
            w = wl.r3_compactified(N=int(N), seed=seed * 100 + k)
            d, keep = abi.make_desc(abi.TOPO_R3, cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere)
        else:
            w = wl.s1r2_sized(int(N), seed=seed * 100 + k)
            if s1r2_tab is None:
                s1r2_tab = abi.s1r2_table(2, w.L, w.Rsim)
            T, shape = s1r2_tab
            d, keep = abi.make_desc(abi.TOPO_S1R2, is_periodic=2, L=w.L, Rsim=w.Rsim, cosmology=1, comoving=1,
                                    mass_in_unit_sphere=w.mass_in_unit_sphere, s1r2_lookup=1, interp_order=2,
                                    s1r2_table=T, s1r2_shape=shape)
            keep = keep + [T]
        sims.append((w, d, keep))
    return sims


def make_requests(sims, nreq: int = 200, rate: float = 200.0, seed: int = 1004, step_frac: float = 1.0,
                  mixed_ok=None):
    rng = np.random.default_rng(seed + 1)
    t = np.cumsum(rng.exponential(1.0 / rate, nreq))
    steps_done = [0] * len(sims)
    reqs = []
    for k in range(nreq):
        s = int(rng.integers(len(sims)))
        kind = abi.REQ_STEP if rng.random() < step_frac else abi.REQ_FORCE
        prec = abi.PREC_MIXED if rng.random() < 0.5 else abi.PREC_FP64
        if mixed_ok is not None and not mixed_ok(sims[s]):
            prec = abi.PREC_FP64
        reqs.append(Req(float(t[k]), s, kind, prec, steps_done[s]))
        if kind == abi.REQ_STEP:
            steps_done[s] += 1
    return reqs


def _latency_stats(lat, t_first, t_last, n):
    lat = np.asarray(lat)
    return {"requests": n, "wall_s": t_last - t_first, "requests_per_s": n / (t_last - t_first),
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "max_ms": float(lat.max() * 1e3)}


def run_server(sims, reqs, streams: int = 4, use_graphs: bool = True, budget: int = 200 << 30):
    """Submit each request at its arrival time (relative to the server clock)
    from a feeder thread; returns (stats, final (x, v) of every simulation)."""
    srv = abi.Server(streams=streams, use_graphs=use_graphs, hbm_budget_bytes=budget)
    ids = [srv.add_sim(d, w.x, w.M, w.soft, w.v) for (w, d, _) in sims]
    # every simulation needs forces of its initial positions before its first step
    first = [srv.submit(i, abi.REQ_FORCE, abi.PREC_FP64) for i in ids]
    srv.drain()
    for t in first:
        assert srv.wait(t).status == 0
    tickets = [None] * len(reqs)
    base = srv.clock()

    def feeder():
        for k, r in enumerate(reqs):
            delay = base + r.t_arrival - srv.clock()
            if delay > 0:
                time.sleep(delay)
            sp = step_params(sims[r.sim][0], r.step) if r.kind == abi.REQ_STEP else None
            tickets[k] = srv.submit(ids[r.sim], r.kind, r.precision, sp)

    feeder()
    srv.drain()
    res = [srv.wait(t) for t in tickets]
    bad = [r.status for r in res if r.status != 0]
    lat = [r.t_end - (base + q.t_arrival) for r, q in zip(res, reqs)]
    out = _latency_stats(lat, base + reqs[0].t_arrival, max(r.t_end for r in res), len(reqs))
    out.update({"failed": len(bad), "streams": streams, "graphs": bool(use_graphs)})
    st = srv.stats()
    out.update({k: st[k] for k in ("admissions", "evictions", "graph_captures", "max_backlog")})
    finals = [srv.get_state(i)[:2] for i in ids]
    srv.close()
    return out, finals


def device_time_alone(sims, reqs) -> dict:
    """Kernel time of the stream's requests one at a time: each (simulation,
    precision) force evaluation timed once with HIP events on its own stream
    (a step's kick/drift kernels add a few microseconds), summed over the
    requests.  The large simulations' kernels fill the device, so no schedule
    serves the stream in much less than this."""
    need = sorted({(r.sim, r.precision) for r in reqs})
    ms = {}
    for si in sorted({k[0] for k in need}):
        w, d, _ = sims[si]
        s = abi.Sim(d, w.N)
        s.set_state(w.x, w.M, w.soft, w.v)
        for (sj, prec) in need:
            if sj != si:
                continue
            s.set_precision(prec)
            s.forces()
            s.synchronize()
            s.forces()
            ms[(si, prec)] = s.last_kernel_ms()[1]
        s.close()
    total = sum(ms[(r.sim, r.precision)] for r in reqs)
    return {"kernel_ms_sum": total, "kernel_ms_by_largest_sims": sorted(ms.values())[-8:]}


def run_sequential(sims, reqs):
    """The same stream served one request at a time: resident simulations,
    one launch sequence per request, no graphs, no concurrency."""
    objs = []
    for (w, d, _) in sims:
        s = abi.Sim(d, w.N)
        s.set_state(w.x, w.M, w.soft, w.v)
        s.forces()
        s.synchronize()
        objs.append(s)
    t0 = time.perf_counter()
    lat = []
    for r in reqs:
        now = time.perf_counter() - t0
        if now < r.t_arrival:
            time.sleep(r.t_arrival - now)
        s = objs[r.sim]
        s.set_precision(r.precision)
        if r.kind == abi.REQ_STEP:
            p = step_params(sims[r.sim][0], r.step)
            s.step(p.h, p.a_kick1, p.hubble1, p.a_kick2, p.hubble2, p.rho_part)
        else:
            s.forces()
            s.synchronize()
        lat.append(time.perf_counter() - t0 - r.t_arrival)
    t_end = time.perf_counter() - t0
    out = _latency_stats(lat, reqs[0].t_arrival, t_end, len(reqs))
    finals = []
    for s in objs:
        finals.append(s.get_state())
        s.close()
    return out, finals
