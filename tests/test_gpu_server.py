"""The serving loop: every request's result is the same alone and in the mix
(bitwise -- same kernels, plan and reduction order), with and without graph
replay, under memory pressure (admission + LRU eviction), and a failing
request does not disturb the others."""
import ctypes as C
import math

import numpy as np
import pytest

import steps_amd
from steps_amd import workloads as wl

pytestmark = pytest.mark.gpu


def step_params(w, k):
    a1 = w.a_start * (1 + 1e-3 * k)
    a2 = w.a_start * (1 + 1e-3 * (k + 1))
    H = lambda a: wl.H0 * math.sqrt(wl.OMEGA_M * a ** -3 + wl.OMEGA_L)
    return steps_amd.StepParams(1e-4, a1, H(a1), a2, H(a2), 1.0, w.rho_part)


def make_sims(oracle):
    sims = []
    for N, seed in [(3000, 1), (20000, 2), (777, 3), (9000, 4)]:
        w = wl.r3_compactified(N=N, seed=seed)
        d, keep = steps_amd.make_desc(steps_amd.TOPO_R3, cosmology=1, comoving=1,
                                      mass_in_unit_sphere=w.mass_in_unit_sphere)
        sims.append((w, d, keep))
    w = wl.t3_lattice(n_side=10, seed=5, is_periodic=1)
    d, keep = steps_amd.make_desc(steps_amd.TOPO_T3, is_periodic=1, L=w.L, cosmology=1, comoving=1)
    sims.append((w, d, keep))
    c = wl.s1r2_cylinder(n_omega=16, n_phi=8, n_z=8, seed=6)
    T, shape = oracle.s1r2_table(2, c.L, c.Rsim)
    d, keep = steps_amd.make_desc(steps_amd.TOPO_S1R2, is_periodic=2, L=c.L, Rsim=c.Rsim, cosmology=1, comoving=1,
                                  mass_in_unit_sphere=c.mass_in_unit_sphere, s1r2_lookup=1, interp_order=2,
                                  s1r2_table=T, s1r2_shape=shape)
    sims.append((c, d, keep + [T]))
    # TSC: the table-window kernels (FP64, and mixed with input-order sources)
    d, keep = steps_amd.make_desc(steps_amd.TOPO_S1R2, is_periodic=2, L=c.L, Rsim=c.Rsim, cosmology=1, comoving=1,
                                  mass_in_unit_sphere=c.mass_in_unit_sphere, s1r2_lookup=1, interp_order=4,
                                  s1r2_table=T, s1r2_shape=shape)
    sims.append((c, d, keep + [T]))
    return sims


NSTEPS = 3


def alone(w, d, precision=0):
    d.precision = precision
    s = steps_amd.Sim(d, w.N)
    s.set_state(w.x, w.M, w.soft, w.v)
    s.forces()
    F0 = s.get_forces()
    errs = []
    for k in range(NSTEPS):
        p = step_params(w, k)
        errs.append(s.step(p.h, p.a_kick1, p.hubble1, p.a_kick2, p.hubble2, p.rho_part))
    x, v = s.get_state()
    s.forces()
    F = s.get_forces()
    s.close()
    d.precision = 0
    return F0, errs, x, v, F


@pytest.fixture(scope="module")
def ensemble(oracle):
    sims = make_sims(oracle)
    ref = [alone(w, d) for (w, d, _) in sims]
    return sims, ref


def run_mix(sims, server, precisions=None):
    ids = [server.add_sim(d, w.x, w.M, w.soft, w.v) for (w, d, _) in sims]
    tickets = {i: [] for i in range(len(sims))}
    # round-robin interleaving: force, then NSTEPS steps, per simulation
    for k in range(NSTEPS + 1):
        for i, (w, d, _) in enumerate(sims):
            prec = precisions[i] if precisions else 0
            if k == 0:
                tickets[i].append(server.submit(ids[i], steps_amd.REQ_FORCE, prec))
            else:
                tickets[i].append(server.submit(ids[i], steps_amd.REQ_STEP, prec, step_params(w, k - 1)))
    server.drain()
    results = {i: [server.wait(t) for t in tickets[i]] for i in tickets}
    states = {i: server.get_state(ids[i]) for i in tickets}
    return ids, results, states


@pytest.mark.parametrize("use_graphs,pack_max_n,streams", [(True, 0, 4), (False, 0, 4), (True, 65536, 4),
                                                            (True, 65536, 1)])
def test_mix_equals_alone(ensemble, use_graphs, pack_max_n, streams):
    print(f"case server mix use_graphs={use_graphs} pack_max_n={pack_max_n} streams={streams}", flush=True)
    sims, ref = ensemble
    srv = steps_amd.Server(streams=streams, use_graphs=use_graphs, pack_max_n=pack_max_n)
    ids, results, states = run_mix(sims, srv)
    for i, (F0, errs, x, v, F) in enumerate(ref):
        assert all(r.status == 0 for r in results[i]), [r.status for r in results[i]]
        assert [r.errmax for r in results[i][1:]] == errs, i
        xs, vs, Fs = states[i]
        assert np.array_equal(xs, x) and np.array_equal(vs, v) and np.array_equal(Fs, F), i
        # per-simulation FIFO: each request starts after the previous one ended
        for a, b in zip(results[i], results[i][1:]):
            assert b.t_start >= a.t_end
    st = srv.stats()
    print(f"[server mix] {st}", flush=True)
    assert st["completed"] == len(sims) * (NSTEPS + 1) and st["failed"] == 0
    if use_graphs and pack_max_n == 0:
        assert st["graph_captures"] == 2 * len(sims)   # one force and one step graph per simulation
    if pack_max_n == 0:
        assert st["packed_launches"] == 0
    if streams == 1 and pack_max_n > 0:
        # one slot: while it is busy the other simulations' requests queue up,
        # so the next issue finds several small ones ready together
        assert st["packed_launches"] >= 1 and st["packed_requests"] >= 2
    srv.close()


def test_admission_and_eviction_keep_results(ensemble):
    print("case server small budget", flush=True)
    sims, ref = ensemble
    probe = steps_amd.Server(streams=1)
    sizes = [probe.sim_bytes(probe.add_sim(d, w.x, w.M, w.soft)) for (w, d, _) in sims]
    probe.close()
    # room for the largest simulation plus a little: the others have to take turns
    budget = max(sizes) + min(sizes)
    srv = steps_amd.Server(streams=2, hbm_budget_bytes=budget)
    ids, results, states = run_mix(sims, srv)
    st = srv.stats()
    print(f"[server small budget] sizes={sizes} budget={budget} {st}", flush=True)
    assert st["evictions"] > 0 and st["admissions"] > len(sims)
    assert st["bytes_resident"] <= budget
    for i, (F0, errs, x, v, F) in enumerate(ref):
        assert all(r.status == 0 for r in results[i])
        assert [r.errmax for r in results[i][1:]] == errs, i
        xs, vs, Fs = states[i]
        assert np.array_equal(xs, x) and np.array_equal(vs, v) and np.array_equal(Fs, F), i
    srv.close()


def test_mixed_precision_requests_match_alone(ensemble):
    sims, ref = ensemble
    capable = sims                                          # every variant here has a mixed kernel
    ref_mixed = [alone(w, d, precision=1) for (w, d, _) in capable]
    srv = steps_amd.Server(streams=3)
    ids, results, states = run_mix(capable, srv, precisions=[1] * len(capable))
    for i, (F0, errs, x, v, F) in enumerate(ref_mixed):
        assert [r.errmax for r in results[i][1:]] == errs
        assert np.array_equal(states[i][2], F)
    srv.close()


def test_failing_request_is_isolated(ensemble):
    sims, ref = ensemble
    srv = steps_amd.Server(streams=4)
    c = sims[5][0]                # S1R2 image sum: it has no mixed-precision path
    d, keep = steps_amd.make_desc(steps_amd.TOPO_S1R2, is_periodic=2, L=c.L, Rsim=c.Rsim, s1r2_lookup=0)
    w0, d0, _ = sims[0]
    bad = srv.add_sim(d, c.x, c.M, c.soft)
    good = srv.add_sim(d0, w0.x, w0.M, w0.soft, w0.v)
    t_bad = srv.submit(bad, steps_amd.REQ_FORCE, steps_amd.PREC_MIXED)
    t_good = srv.submit(good, steps_amd.REQ_FORCE, steps_amd.PREC_FP64)
    assert srv.wait(t_bad).status != 0
    assert srv.wait(t_good).status == 0
    assert np.array_equal(srv.get_state(good)[2], ref[0][0])
    with pytest.raises(steps_amd.StepsError):
        srv.submit(99, steps_amd.REQ_FORCE)
    with pytest.raises(steps_amd.StepsError):
        srv.submit(good, steps_amd.REQ_STEP)       # a step needs parameters
    srv.close()


def test_server_builds_missing_tables():
    """Descriptors without tables: the server builds the reference's tables
    (and caches them by parameters); results equal those with the same tables
    passed in explicitly."""
    print("case server builds missing tables", flush=True)
    c = wl.s1r2_cylinder(n_omega=16, n_phi=8, n_z=8, seed=6)
    t = wl.t3_lattice(n_side=10, seed=5)
    T, shape = steps_amd.s1r2_table(2, c.L, c.Rsim)
    radial = steps_amd.radial_table(c.Rsim, 0.5 * c.L, 1000, 1000)
    T3 = steps_amd.t3_table(2, t.L)
    base = dict(L=c.L, Rsim=c.Rsim, cosmology=1, comoving=1, mass_in_unit_sphere=c.mass_in_unit_sphere)
    cases = [
        (c, dict(is_periodic=2, s1r2_lookup=1, interp_order=2, **base), dict(s1r2_table=T, s1r2_shape=shape),
         steps_amd.TOPO_S1R2),
        (c, dict(is_periodic=1, **base), dict(radial_table=radial), steps_amd.TOPO_S1R2),
        (t, dict(is_periodic=2, L=t.L), dict(t3_table=T3), steps_amd.TOPO_T3),
    ]
    srv = steps_amd.Server(streams=2)
    for w, kw, tables, topo in cases:
        d0, k0 = steps_amd.make_desc(topo, **kw)
        d1, k1 = steps_amd.make_desc(topo, **kw, **tables)
        a = srv.add_sim(d0, w.x, w.M, w.soft)
        b = srv.add_sim(d1, w.x, w.M, w.soft)
        ta = srv.submit(a, steps_amd.REQ_FORCE)
        tb = srv.submit(b, steps_amd.REQ_FORCE)
        assert srv.wait(ta).status == 0 and srv.wait(tb).status == 0
        Fa, Fb = srv.get_state(a)[2], srv.get_state(b)[2]
        assert np.array_equal(Fa, Fb), kw
        assert np.all(np.isfinite(Fa)) and np.abs(Fa).max() > 0
    srv.close()


def test_collected_tickets_are_forgotten():
    """A long-running server keeps only the tickets whose results were not
    collected; a collected ticket cannot be waited for again."""
    print("case server ticket window", flush=True)
    w = wl.r3_compactified(N=1000, seed=9)
    d, keep = steps_amd.make_desc(steps_amd.TOPO_R3)
    srv = steps_amd.Server(streams=2)
    sid = srv.add_sim(d, w.x, w.M, w.soft)
    tickets = [srv.submit(sid, steps_amd.REQ_FORCE) for _ in range(40)]
    assert srv.stats()["tickets_held"] <= 40
    for t in tickets[:20]:
        assert srv.wait(t).status == 0
    assert srv.stats()["tickets_held"] <= 20
    for t in tickets[20:]:
        assert srv.wait(t).status == 0
    assert srv.stats()["tickets_held"] == 0
    with pytest.raises(steps_amd.StepsError):
        srv.wait(tickets[0])
    t = srv.submit(sid, steps_amd.REQ_FORCE)
    assert t == tickets[-1] + 1 and srv.wait(t).status == 0
    srv.close()


def test_submit_and_wait_return_during_an_admission():
    """The scheduler does device work (allocation, state upload) after
    releasing the server lock: while a large simulation is being admitted,
    get_stats, submit and wait return at once."""
    print("case server admission outside the lock (N = 2^24)", flush=True)
    srv = steps_amd.Server(streams=2)
    w = wl.r3_compactified(N=2000, seed=9)
    d, keep = steps_amd.make_desc(steps_amd.TOPO_R3)
    small = srv.add_sim(d, w.x, w.M, w.soft)
    t_done = srv.submit(small, steps_amd.REQ_FORCE)
    srv.drain()                                       # done, not collected
    n = 1 << 24
    rng = np.random.default_rng(3)
    xb = rng.normal(size=(n, 3)) * 100.0
    Mb, hb = np.ones(n), np.full(n, 0.01)
    # the S1R2 image-sum variant has no mixed-precision kernel: the request is
    # admitted (device copies made) and then fails before any pair work
    db, kb = steps_amd.make_desc(steps_amd.TOPO_S1R2, is_periodic=2, L=200.0, Rsim=1000.0, s1r2_lookup=0)
    big = srv.add_sim(db, xb, Mb, hb)
    del xb
    t_big = srv.submit(big, steps_amd.REQ_FORCE, steps_amd.PREC_MIXED)
    seen = False
    for _ in range(2_000_000):
        if srv.stats()["admitting"] == 1:
            seen = True
            break
        if srv.stats()["failed"] > 0:
            break
    assert seen, "the admission was never observed in progress"
    t_new = srv.submit(small, steps_amd.REQ_FORCE)
    r = srv.wait(t_done)
    still = srv.stats()["admitting"]
    print(f"[server admission] admitting after submit + wait: {still}", flush=True)
    assert r.status == 0 and still == 1
    assert srv.wait(t_big).status != 0
    assert srv.wait(t_new).status == 0
    srv.close()


def test_step_before_force_and_removal():
    """A step request before any force evaluation of its simulation fails
    (the kick needs the forces of the current positions); removing a
    simulation fails its pending requests and frees its device copy."""
    print("case server step-before-force / removal", flush=True)
    w = wl.r3_compactified(N=65536, seed=1)          # ~2 ms a step: the later requests are still queued
    d, keep = steps_amd.make_desc(steps_amd.TOPO_R3, cosmology=1, comoving=1,
                                  mass_in_unit_sphere=w.mass_in_unit_sphere)
    srv = steps_amd.Server(streams=1)
    sid = srv.add_sim(d, w.x, w.M, w.soft, w.v)
    r = srv.wait(srv.submit(sid, steps_amd.REQ_STEP, 0, step_params(w, 0)))
    assert r.status == steps_amd.STEPS_ERR_STATE
    assert srv.wait(srv.submit(sid, steps_amd.REQ_FORCE)).status == 0
    assert srv.wait(srv.submit(sid, steps_amd.REQ_STEP, 0, step_params(w, 0))).status == 0
    tickets = [srv.submit(sid, steps_amd.REQ_STEP, 0, step_params(w, k)) for k in range(1, 6)]
    srv.remove_sim(sid)
    res = [srv.wait(t) for t in tickets]
    assert res[-1].status == steps_amd.STEPS_ERR_STATE      # never issued
    st = srv.stats()
    assert st["resident_sims"] == 0 and st["bytes_resident"] == 0, st
    srv.close()


def test_c4_ensemble_server_equals_sequential():
    """The serving configuration itself: 64 simulations (R3 / S1R2 CIC
    alternating, N log-uniform in [8192, 524288], seed 1004) and a seeded
    Poisson stream of 48 KDK-step requests (200 req/s, FP64 / mixed 50/50),
    served by the server -- with the 200 GB budget, and with a budget that
    holds a third of the ensemble, so that simulations are evicted and
    re-admitted -- and one request at a time: every simulation's final
    (x, v) is bitwise the same."""
    from steps_amd import serving as sv
    print("case server c4 ensemble 64 sims, 48 requests", flush=True)
    sims = sv.make_ensemble()
    reqs = sv.make_requests(sims, nreq=48)
    seq, fin_q = sv.run_sequential(sims, reqs)
    print(f"[c4 sequential] {seq}", flush=True)
    srv, fin_s = sv.run_server(sims, reqs)
    print(f"[c4 server] {srv}", flush=True)
    assert srv["failed"] == 0 and srv["admissions"] == len(sims) and srv["evictions"] == 0
    for i, (a, b) in enumerate(zip(fin_s, fin_q)):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), i
    probe = steps_amd.Server(streams=1)
    total = sum(probe.sim_bytes(probe.add_sim(d, w.x, w.M, w.soft)) for (w, d, _) in sims)
    probe.close()
    small, fin_b = sv.run_server(sims, reqs, budget=total // 3)
    print(f"[c4 server, budget {total // 3} of {total}] {small}", flush=True)
    assert small["failed"] == 0 and small["evictions"] > 0
    for i, (a, b) in enumerate(zip(fin_b, fin_q)):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), i
