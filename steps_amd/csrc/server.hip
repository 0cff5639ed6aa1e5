// server.hip -- continuous-batching request loop over resident simulations.
//
// There is no counterpart in the reference (it runs one simulation per MPI
// job, main.cc:1676-1864).  Requests are force evaluations or KDK steps of
// independent simulations of uneven N.  One scheduler thread owns all device
// work:
//   * admission: a simulation becomes resident (state + scratch in HBM) when
//     it has work and fits the memory budget; idle resident simulations are
//     evicted least-recently-used to host memory to make room, and re-admitted
//     with their state (x, v, F) when they get work again;
//   * up to `streams` requests are on the device at once, each on its own
//     stream, at most one per simulation (a simulation's requests run in
//     submission order); requests that arrive while others run are issued as
//     slots and memory free up, so their kernels overlap on the device;
//   * each simulation's force and step sequences are captured once per
//     precision in a hipGraph and replayed (one launch per request); the
//     per-step scalars travel through a pinned block the graph copies in;
//   * ready requests of small simulations (N <= pack_max_n) with the same
//     pair variant and precision are packed: their pair kernels run as one
//     launch (a flat grid over all their row blocks and splits), the cheap
//     per-simulation kernels around it are launched directly.
// A request's result does not depend on the mix: its kernels, plan and
// reduction order are those of the simulation alone.
#include "sim_internal.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

namespace steps {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t ticket = 0, sim = 0;
    int kind = 0, precision = 0;
    steps_step_params sp{};
    double t_submit = 0, t_start = 0, t_end = 0;
    int status = STEPS_OK;
    double errmax = 0;
    bool done = false;
    bool collected = false;  // its result was returned by steps_server_wait
};

struct SimEntry {
    steps_force_desc desc{};
    int64_t N = 0;
    int variant = 0;                        // pair-kernel variant (make_plan)
    std::vector<double> x, v, M, soft, F;   // host copy, authoritative while not resident
    std::vector<double> t3, s1r2, radial;   // host tables, needed to (re)admit
    steps_sim *res = nullptr;
    size_t bytes = 0;
    bool busy = false, has_F = false, removed = false;
    double last_used = 0;
    std::deque<int64_t> pending;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    int64_t ticket = -1;          // the request on the slot, or the first of a packed batch
    std::vector<int64_t> batch;   // all requests of a packed launch (empty otherwise)
    DevBuf dseg;                  // packed-launch descriptors: device copy ...
    void *hseg = nullptr;         // ... and pinned host staging
};

// Device bytes a resident simulation takes (what sim_create allocates).
size_t footprint(const steps_force_desc &d, int64_t N) {
    EvalPlan p;
    steps_force_desc dd = d;
    if (desc_mixed_capable(d)) dd.precision = STEPS_PREC_MIXED;
    // tables only have to look present for the size computation
    static const double dummy = 0;
    if (!dd.t3_table) dd.t3_table = &dummy;
    if (!dd.s1r2_table) dd.s1r2_table = &dummy;
    if (!dd.radial_table) dd.radial_table = &dummy;
    size_t b = (size_t)N * (32 + 8 + 24 + 24 + 24) + 4096;
    if (make_plan(dd, N, 0, N - 1, p) == STEPS_OK) b += scratch_bytes(p);
    b += (size_t)(kdk_reduce_blocks(N) + 1) * 8;
    if (d.topology == STEPS_TOPO_T3 && d.is_periodic >= 2 && d.t3_ngrid > 0)
        b += t3_device_entries(d.t3_ngrid, 0) * sizeof(double);  // unfolded layout: the larger one
    if (d.topology == STEPS_TOPO_S1R2 && d.is_periodic >= 2 && d.s1r2_lookup)
        b += s1r2_device_nodes(std::max(d.s1r2_nrho, 1), std::max(d.s1r2_nz, 1)) * 16;
    if (d.topology == STEPS_TOPO_S1R2) b += (size_t)std::max(0, d.radial_table_size) * 8;
    return b;
}

// T3 Ewald table (main.cc:422-448), S1R2 Ewald table (main.cc:569-602) and
// the cylindrical radial table (main.cc:1266-1302: R = Rsim, Lz = L/2 with
// IS_PERIODIC == 1, L * (IS_PERIODIC + 1 - 0.4) for the image sum;
// RADIAL_FORCE_TABLE_SIZE and RADIAL_FORCE_ACCURACY 1000, read_paramfile.cc:
// 128-129) for every table the descriptor needs and does not carry.
int build_missing_tables(steps_force_desc &d, std::vector<double> &t3, std::vector<double> &s1r2,
                         std::vector<double> &radial) {
    int rc = STEPS_OK;
    if (d.topology == STEPS_TOPO_T3 && d.is_periodic >= 2 && !d.t3_table) {
        int ng = 0;
        if ((rc = build_t3_table(d.is_periodic, d.L, nullptr, &ng))) return rc;
        t3.resize((size_t)ng * ng * ng * 3);
        if ((rc = build_t3_table(d.is_periodic, d.L, t3.data(), &ng))) return rc;
        d.t3_table = t3.data();
        d.t3_ngrid = ng;
    }
    if (d.topology == STEPS_TOPO_S1R2) {
        if (d.is_periodic >= 2 && d.s1r2_lookup && !d.s1r2_table) {
            int a = 0, b = 0;
            if ((rc = build_s1r2_table(d.is_periodic, d.L, d.Rsim, nullptr, &a, &b))) return rc;
            s1r2.resize((size_t)std::max(a, 0) * b * 2);
            if ((rc = build_s1r2_table(d.is_periodic, d.L, d.Rsim, s1r2.data(), &a, &b))) return rc;
            d.s1r2_table = s1r2.data();
            d.s1r2_nrho = a;
            d.s1r2_nz = b;
        }
        const bool radial_used = d.cosmology == 1 && d.comoving == 1 && (d.is_periodic == 1 || !d.s1r2_lookup);
        if (radial_used && !d.radial_table && d.is_periodic >= 1) {
            const int size = d.radial_table_size > 0 ? d.radial_table_size : 1000;
            const double Lz = (d.is_periodic == 1) ? 0.5 * d.L : d.L * ((double)(d.is_periodic + 1) - 0.4);
            radial.resize((size_t)size);
            if ((rc = build_radial_table(d.Rsim, Lz, size, 1000, radial.data()))) return rc;
            d.radial_table = radial.data();
            d.radial_table_size = size;
        }
    }
    return rc;
}

}  // namespace
}  // namespace steps

using namespace steps;

struct steps_server {
    steps_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread thr;
    bool stop = false;
    std::vector<SimEntry> sims;
    // Requests by ticket: reqs[t - base].  Requests that are done and whose
    // result was collected leave the front, so a long-running server keeps
    // only the window of tickets still in use.
    std::deque<Request> reqs;
    int64_t base = 0;
    Request &req(int64_t t) { return reqs[(size_t)(t - base)]; }
    bool known(int64_t t) const { return t >= base && t < base + (int64_t)reqs.size(); }
    void prune() {
        while (!reqs.empty() && reqs.front().done && reqs.front().collected) {
            reqs.pop_front();
            ++base;
        }
    }
    std::vector<Slot> slots;
    size_t used = 0;
    steps_server_stats st{};
    Clock::time_point t0 = Clock::now();

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    int backlog() const {
        int64_t n = 0;
        for (const auto &s : sims) n += (int64_t)s.pending.size();
        return (int)n;
    }

    void evict(SimEntry &e) {
        // state back to host, then free the device copy
        if (sim_download(e.res, e.x.data(), e.v.data(), e.F.data()) == STEPS_OK) e.has_F = true;
        sim_destroy(e.res);
        e.res = nullptr;
        used -= e.bytes;
        st.evictions++;
    }

    // 0: resident; 1: cannot fit now (retry later); <0: permanent error
    int ensure_resident(SimEntry &e, int64_t self) {
        if (e.res) return 0;
        if ((int64_t)e.bytes > cfg.hbm_budget_bytes) return STEPS_ERR_ADMISSION;
        while ((int64_t)(used + e.bytes) > cfg.hbm_budget_bytes) {
            SimEntry *lru = nullptr;
            for (size_t k = 0; k < sims.size(); ++k) {
                SimEntry &o = sims[k];
                if ((int64_t)k == self || !o.res || o.busy) continue;
                if (!lru || o.last_used < lru->last_used) lru = &o;
            }
            if (!lru) return 1;
            evict(*lru);
        }
        steps_force_desc d = e.desc;
        d.t3_table = e.t3.empty() ? nullptr : e.t3.data();
        d.s1r2_table = e.s1r2.empty() ? nullptr : e.s1r2.data();
        d.radial_table = e.radial.empty() ? nullptr : e.radial.data();
        steps_sim *s = nullptr;
        int rc = sim_create(d, e.N, true, &s);
        if (rc) return rc;
        rc = sim_upload(s, e.x.data(), e.v.data(), e.M.data(), e.soft.data(), e.has_F ? e.F.data() : nullptr);
        if (rc) {
            sim_destroy(s);
            return rc;
        }
        e.res = s;
        used += e.bytes;
        st.admissions++;
        return 0;
    }

    void finish(Request &r, int status) {
        r.status = status;
        r.t_end = now();
        r.done = true;
        if (status == STEPS_OK) st.completed++;
        else st.failed++;
        cv_done.notify_all();
    }

    // Launch request r of entry e on slot sl; returns a status (not thrown).
    int issue(SimEntry &e, Request &r, Slot &sl) {
        steps_sim *s = e.res;
        EvalPlan p;
        int rc = sim_plan(s, 0, e.N - 1, r.precision, p);
        if (rc) return rc;
        if (r.kind == STEPS_REQ_STEP && (rc = sim_fill_params(s, r.sp))) return rc;
        hipError_t he;
        hipGraphExec_t &g = s->graph[r.kind][r.precision];
        if (cfg.use_graphs) {
            if (!g) {
                hipGraph_t graph = nullptr;
                STEPS_HIP(hipStreamBeginCapture(sl.stream, hipStreamCaptureModeRelaxed), "begin capture");
                he = (r.kind == STEPS_REQ_STEP) ? sim_enqueue_step(s, p, sl.stream)
                                                : sim_enqueue_forces(s, p, sl.stream, false);
                hipError_t he2 = hipStreamEndCapture(sl.stream, &graph);
                if (he != hipSuccess) return hip_fail(he, "capture");
                if (he2 != hipSuccess) return hip_fail(he2, "end capture");
                he = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (he != hipSuccess) {
                    g = nullptr;
                    return hip_fail(he, "instantiate");
                }
                st.graph_captures++;
            }
            STEPS_HIP(hipGraphLaunch(g, sl.stream), "graph launch");
        } else {
            he = (r.kind == STEPS_REQ_STEP) ? sim_enqueue_step(s, p, sl.stream)
                                            : sim_enqueue_forces(s, p, sl.stream, false);
            if (he != hipSuccess) return hip_fail(he, "enqueue");
        }
        STEPS_HIP(hipEventRecord(sl.done, sl.stream), "record");
        return STEPS_OK;
    }

    // Ready requests that can share one packed pair launch with the head of
    // ready[qi] (resident, small, same variant and precision, valid plan); the
    // head itself first.  Members are kept from eviction while the group forms.
    std::vector<int64_t> form_pack(const std::vector<int64_t> &ready, size_t qi, const std::vector<char> &taken) {
        const int64_t k = ready[qi];
        std::vector<int64_t> group{k};
        SimEntry &e = sims[k];
        if (cfg.pack_max_n <= 0 || e.N > cfg.pack_max_n) return group;
        const int prec = req(e.pending.front()).precision;
        EvalPlan p;
        if (sim_plan(e.res, 0, e.N - 1, prec, p) != STEPS_OK) return group;
        e.busy = true;
        for (size_t qj = qi + 1; qj < ready.size() && (int)group.size() < kMaxPack; ++qj) {
            const int64_t k2 = ready[qj];
            SimEntry &e2 = sims[k2];
            if (taken[k2] || e2.removed || e2.busy || e2.N > cfg.pack_max_n || e2.variant != e.variant) continue;
            if (req(e2.pending.front()).precision != prec) continue;
            if (ensure_resident(e2, k2) != 0) continue;
            EvalPlan p2;
            if (sim_plan(e2.res, 0, e2.N - 1, prec, p2) != STEPS_OK) continue;
            e2.busy = true;
            group.push_back(k2);
        }
        for (const int64_t g : group) sims[g].busy = false;
        return group;
    }

    // One packed launch for the head requests of `group`: per simulation the
    // step's first half and the source preparation, one pair kernel over all,
    // per simulation the reduction and the step's second half.  No graphs
    // (the composition changes from launch to launch).
    void issue_batch(const std::vector<int64_t> &group, Slot &sl) {
        const size_t n = group.size();
        std::vector<EvalPlan> plans(n);
        std::vector<int64_t> tickets(n);
        const double t0 = now();
        for (size_t i = 0; i < n; ++i) {
            SimEntry &e = sims[group[i]];
            tickets[i] = e.pending.front();
            e.pending.pop_front();
            req(tickets[i]).t_start = t0;
        }
        auto fail_all = [&](int code) {
            for (const int64_t t : tickets) finish(req(t), code);
        };
        for (size_t i = 0; i < n; ++i) {
            SimEntry &e = sims[group[i]];
            const Request &r = req(tickets[i]);
            int rc = sim_plan(e.res, 0, e.N - 1, r.precision, plans[i]);
            if (!rc && r.kind == STEPS_REQ_STEP) rc = sim_fill_params(e.res, r.sp);
            if (rc) return fail_all(rc);
        }
        const hipStream_t qs = sl.stream;
        hipError_t he = hipSuccess;
        for (size_t i = 0; i < n && he == hipSuccess; ++i)
            if (req(tickets[i]).kind == STEPS_REQ_STEP) he = sim_enqueue_step_before(sims[group[i]].res, qs);
        for (size_t i = 0; i < n && he == hipSuccess; ++i) {
            steps_sim *s = sims[group[i]].res;
            he = launch_forces_prepare(plans[i], s->dp, s->part.as<double>(), qs);
        }
        int64_t blocks = 0;
        for (size_t i = 0; i < n && he == hipSuccess; ++i) {
            steps_sim *s = sims[group[i]].res;
            blocks += pack_describe(plans[i], s->dp, s->part.as<double>(), sl.hseg, (int)i, blocks);
        }
        if (he == hipSuccess)
            he = hipMemcpyAsync(sl.dseg.ptr, sl.hseg, n * pack_desc_bytes(), hipMemcpyHostToDevice, qs);
        if (he == hipSuccess) he = launch_pair_packed(plans[0].variant, plans[0].precision, sl.dseg.ptr, (int)n, blocks, qs);
        for (size_t i = 0; i < n && he == hipSuccess; ++i) {
            steps_sim *s = sims[group[i]].res;
            he = launch_forces_finish(plans[i], s->dp, s->part.as<double>(), s->F.as<double>(), qs);
        }
        for (size_t i = 0; i < n && he == hipSuccess; ++i)
            if (req(tickets[i]).kind == STEPS_REQ_STEP) he = sim_enqueue_step_after(sims[group[i]].res, qs);
        if (he == hipSuccess) he = hipEventRecord(sl.done, qs);
        if (he != hipSuccess) {
            // work already queued on this stream must drain before the
            // simulations can be used again
            (void)hipStreamSynchronize(qs);
            return fail_all(hip_fail(he, "packed launch"));
        }
        for (const int64_t g : group) sims[g].busy = true;
        sl.batch = tickets;
        sl.ticket = tickets[0];
        st.packed_launches++;
        st.packed_requests += (int64_t)n;
    }

    void loop() {
        (void)hipSetDevice(cfg.device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            bool progressed = false;
            int inflight = 0;
            // completions
            for (Slot &sl : slots) {
                if (sl.ticket < 0) continue;
                const hipError_t q = hipEventQuery(sl.done);
                if (q == hipErrorNotReady) {
                    inflight++;
                    continue;
                }
                if (sl.batch.empty()) sl.batch.push_back(sl.ticket);
                for (const int64_t t : sl.batch) {
                    Request &r = req(t);
                    SimEntry &e = sims[r.sim];
                    if (q == hipSuccess) {
                        if (r.kind == STEPS_REQ_STEP) r.errmax = *e.res->herrmax;
                        e.has_F = true;
                        finish(r, STEPS_OK);
                    } else {
                        finish(r, hip_fail(q, "request"));
                    }
                    e.busy = false;
                    e.last_used = now();
                    if (e.removed && e.pending.empty() && e.res) {
                        sim_destroy(e.res);
                        e.res = nullptr;
                        used -= e.bytes;
                    }
                }
                sl.batch.clear();
                sl.ticket = -1;
                progressed = true;
            }
            // issue: oldest head request first, while slots are free
            std::vector<int64_t> ready;
            for (size_t k = 0; k < sims.size(); ++k)
                if (!sims[k].busy && !sims[k].pending.empty()) ready.push_back((int64_t)k);
            std::sort(ready.begin(), ready.end(),
                      [&](int64_t a, int64_t b) { return sims[a].pending.front() < sims[b].pending.front(); });
            std::vector<char> taken(sims.size(), 0);
            for (size_t qi = 0; qi < ready.size(); ++qi) {
                const int64_t k = ready[qi];
                if (taken[k]) continue;
                Slot *free_slot = nullptr;
                for (Slot &sl : slots)
                    if (sl.ticket < 0) { free_slot = &sl; break; }
                if (!free_slot) break;
                SimEntry &e = sims[k];
                const int64_t t = e.pending.front();
                Request &r = req(t);
                if (e.removed) {
                    e.pending.pop_front();
                    finish(r, fail(STEPS_ERR_STATE, "simulation removed"));
                    progressed = true;
                    continue;
                }
                const int adm = ensure_resident(e, k);
                if (adm == 1) continue;  // no room until something finishes
                if (adm < 0) {
                    e.pending.pop_front();
                    progressed = true;
                    finish(r, adm);
                    continue;
                }
                std::vector<int64_t> group = form_pack(ready, qi, taken);
                for (const int64_t g : group) taken[g] = 1;
                progressed = true;
                if (group.size() >= 2) {
                    issue_batch(group, *free_slot);
                    if (free_slot->ticket >= 0) inflight++;
                    continue;
                }
                e.pending.pop_front();
                r.t_start = now();
                const int rc = issue(e, r, *free_slot);
                if (rc != STEPS_OK) {
                    finish(r, rc);
                    continue;
                }
                e.busy = true;
                free_slot->ticket = t;
                inflight++;
            }
            if (stop && inflight == 0) break;
            if (!progressed) {
                if (inflight > 0) {
                    lk.unlock();
                    std::this_thread::sleep_for(std::chrono::microseconds(20));
                    lk.lock();
                } else {
                    cv_work.wait_for(lk, std::chrono::milliseconds(50));
                }
            }
        }
    }
};

extern "C" {

void steps_server_config_init(steps_server_config *c) {
    if (!c) return;
    c->device = 0;
    c->streams = 4;
    c->use_graphs = 1;
    c->hbm_budget_bytes = (int64_t)200 << 30;
    c->pack_max_n = 65536;
}

int steps_server_create(const steps_server_config *cfg, steps_server **out) {
    if (!cfg || !out) return fail(STEPS_ERR_ARG, "null argument");
    *out = nullptr;
    if (cfg->streams < 1 || cfg->streams > 16) return fail(STEPS_ERR_ARG, "streams must be in [1, 16]");
    if (cfg->hbm_budget_bytes <= 0) return fail(STEPS_ERR_ARG, "budget must be positive");
    if (cfg->pack_max_n < 0) return fail(STEPS_ERR_ARG, "pack_max_n must be >= 0");
    int rc = set_device(cfg->device);
    if (rc) return rc;
    steps_server *S = new (std::nothrow) steps_server();
    if (!S) return fail(STEPS_ERR_NOMEM, "host allocation");
    S->cfg = *cfg;
    S->slots.resize((size_t)cfg->streams);
    for (Slot &sl : S->slots) {
        hipError_t e;
        if ((e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming)) != hipSuccess ||
            (e = hipHostMalloc(&sl.hseg, kMaxPack * pack_desc_bytes(), hipHostMallocDefault)) != hipSuccess) {
            rc = hip_fail(e, "server streams");
            steps_server_destroy(S);
            return rc;
        }
        if ((rc = sl.dseg.reserve(kMaxPack * pack_desc_bytes(), "pack descriptors"))) {
            steps_server_destroy(S);
            return rc;
        }
    }
    S->thr = std::thread([S] { S->loop(); });
    *out = S;
    return STEPS_OK;
}

int steps_server_destroy(steps_server *S) {
    if (!S) return STEPS_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        for (auto &e : S->sims) {  // pending work is abandoned
            while (!e.pending.empty()) {
                Request &r = S->req(e.pending.front());
                e.pending.pop_front();
                S->finish(r, fail(STEPS_ERR_STATE, "server destroyed"));
            }
        }
    }
    S->cv_work.notify_all();
    if (S->thr.joinable()) S->thr.join();
    (void)hipSetDevice(S->cfg.device);
    for (auto &e : S->sims)
        if (e.res) sim_destroy(e.res);
    for (Slot &sl : S->slots) {
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.hseg) (void)hipHostFree(sl.hseg);
        sl.dseg.release();
    }
    delete S;
    return STEPS_OK;
}

int steps_server_add_sim(steps_server *S, const steps_force_desc *d, int64_t N, const double *x, const double *v,
                         const double *M, const double *soft, int64_t *sim_id) {
    if (!S || !d || !x || !M || !soft || !sim_id || N <= 0) return fail(STEPS_ERR_ARG, "bad add_sim arguments");
    if (d->device != S->cfg.device) return fail(STEPS_ERR_ARG, "simulation device differs from the server's");
    // Tables the descriptor leaves out are built here with the reference's
    // parameters (build_* keeps them cached by parameters across simulations).
    steps_force_desc dd = *d;
    std::vector<double> t3, s1r2, radial;
    int rc = build_missing_tables(dd, t3, s1r2, radial);
    if (rc) return rc;
    d = &dd;
    EvalPlan p;
    rc = make_plan(*d, N, 0, N - 1, p);  // validates the descriptor and its tables
    if (rc) return rc;
    if ((rc = check_finite(x, 3 * N))) return rc;
    SimEntry e;
    e.desc = *d;
    e.N = N;
    e.x.assign(x, x + 3 * N);
    e.v = v ? std::vector<double>(v, v + 3 * N) : std::vector<double>(3 * N, 0.0);
    e.M.assign(M, M + N);
    e.soft.assign(soft, soft + N);
    e.F.assign(3 * N, 0.0);
    if (p.variant == PV_T3_EWALD) e.t3.assign(d->t3_table, d->t3_table + (size_t)d->t3_ngrid * d->t3_ngrid * d->t3_ngrid * 3);
    if (p.variant >= PV_S1R2_NGP && p.variant <= PV_S1R2_TSC)
        e.s1r2.assign(d->s1r2_table, d->s1r2_table + (size_t)d->s1r2_nrho * d->s1r2_nz * 2);
    if (p.ext_mode == 3) e.radial.assign(d->radial_table, d->radial_table + d->radial_table_size);
    e.desc.t3_table = e.desc.s1r2_table = e.desc.radial_table = nullptr;
    e.variant = p.variant;
    e.bytes = footprint(*d, N);
    if ((int64_t)e.bytes > S->cfg.hbm_budget_bytes)
        return fail(STEPS_ERR_ADMISSION, "simulation larger than the whole memory budget");
    std::lock_guard<std::mutex> lk(S->mu);
    *sim_id = (int64_t)S->sims.size();
    S->sims.push_back(std::move(e));
    return STEPS_OK;
}

int steps_server_remove_sim(steps_server *S, int64_t sim) {
    if (!S) return fail(STEPS_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (sim < 0 || sim >= (int64_t)S->sims.size()) return fail(STEPS_ERR_ARG, "unknown simulation");
    SimEntry &e = S->sims[sim];
    e.removed = true;
    if (!e.busy && e.pending.empty() && e.res) {
        sim_destroy(e.res);
        e.res = nullptr;
        S->used -= e.bytes;
    }
    S->cv_work.notify_all();
    return STEPS_OK;
}

int64_t steps_server_sim_bytes(steps_server *S, int64_t sim) {
    if (!S) return fail(STEPS_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (sim < 0 || sim >= (int64_t)S->sims.size()) return fail(STEPS_ERR_ARG, "unknown simulation");
    return (int64_t)S->sims[sim].bytes;
}

int steps_server_submit(steps_server *S, int64_t sim, int32_t kind, int32_t precision, const steps_step_params *sp,
                        int64_t *ticket) {
    if (!S || !ticket) return fail(STEPS_ERR_ARG, "null argument");
    if (kind != STEPS_REQ_FORCE && kind != STEPS_REQ_STEP) return fail(STEPS_ERR_ARG, "unknown request kind");
    if (precision != STEPS_PREC_FP64 && precision != STEPS_PREC_MIXED) return fail(STEPS_ERR_ARG, "unknown precision");
    if (kind == STEPS_REQ_STEP && !sp) return fail(STEPS_ERR_ARG, "step request without parameters");
    std::lock_guard<std::mutex> lk(S->mu);
    if (sim < 0 || sim >= (int64_t)S->sims.size() || S->sims[sim].removed)
        return fail(STEPS_ERR_ARG, "unknown simulation");
    if (S->stop) return fail(STEPS_ERR_STATE, "server stopping");
    Request r;
    r.ticket = S->base + (int64_t)S->reqs.size();
    r.sim = sim;
    r.kind = kind;
    r.precision = precision;
    if (sp) r.sp = *sp;
    r.t_submit = S->now();
    S->reqs.push_back(r);
    S->sims[sim].pending.push_back(r.ticket);
    S->st.submitted++;
    S->st.max_backlog = std::max<int64_t>(S->st.max_backlog, S->backlog());
    *ticket = r.ticket;
    S->cv_work.notify_one();
    return STEPS_OK;
}

int steps_server_wait(steps_server *S, int64_t ticket, steps_request_result *out) {
    if (!S) return fail(STEPS_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    if (!S->known(ticket)) return fail(STEPS_ERR_ARG, "unknown ticket (or its result was already collected)");
    S->cv_done.wait(lk, [&] { return !S->known(ticket) || S->req(ticket).done; });
    if (!S->known(ticket)) return fail(STEPS_ERR_ARG, "the result of this ticket was already collected");
    Request &r = S->req(ticket);
    if (out) {
        out->status = r.status;
        out->kind = r.kind;
        out->sim = r.sim;
        out->ticket = r.ticket;
        out->errmax = r.errmax;
        out->t_submit = r.t_submit;
        out->t_start = r.t_start;
        out->t_end = r.t_end;
    }
    r.collected = true;
    S->prune();
    return STEPS_OK;
}

int steps_server_drain(steps_server *S) {
    if (!S) return fail(STEPS_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] {
        for (const auto &r : S->reqs)
            if (!r.done) return false;
        return true;
    });
    return STEPS_OK;
}

int steps_server_get_state(steps_server *S, int64_t sim, double *x, double *v, double *F) {
    if (!S) return fail(STEPS_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    if (sim < 0 || sim >= (int64_t)S->sims.size()) return fail(STEPS_ERR_ARG, "unknown simulation");
    S->cv_done.wait(lk, [&] { return !S->sims[sim].busy && S->sims[sim].pending.empty(); });
    SimEntry &e = S->sims[sim];
    if (e.res) {
        int rc = sim_download(e.res, x, v, F);
        return rc;
    }
    const size_t n3 = 3 * (size_t)e.N;
    if (x) std::copy(e.x.begin(), e.x.begin() + n3, x);
    if (v) std::copy(e.v.begin(), e.v.begin() + n3, v);
    if (F) std::copy(e.F.begin(), e.F.begin() + n3, F);
    return STEPS_OK;
}

int steps_server_get_stats(steps_server *S, steps_server_stats *out) {
    if (!S || !out) return fail(STEPS_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
    out->bytes_resident = (int64_t)S->used;
    out->tickets_held = (int64_t)S->reqs.size();
    int64_t n = 0;
    for (const auto &e : S->sims) n += e.res ? 1 : 0;
    out->resident_sims = n;
    return STEPS_OK;
}

double steps_server_clock(steps_server *S) { return S ? S->now() : 0.0; }

}  // extern "C"
