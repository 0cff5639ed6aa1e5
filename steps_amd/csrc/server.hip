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
// The scheduler takes its decisions under the server mutex and does the
// device work they imply (allocation, state copies, graph capture, launches)
// after releasing it, so submit, wait, get_stats and get_state never wait for
// device work; a simulation whose device state is in use is marked busy.
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
    // busy: its device state is in use (a request in flight, an admission
    // or eviction in progress, a get_state copy); has_F: the host F copy is
    // the forces of the host positions; lost: an eviction failed to copy the
    // state back, so its requests fail
    bool busy = false, has_F = false, removed = false, lost = false;
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
    // tables only have to look present for the size computation
    static const double dummy = 0;
    if (!dd.t3_table) dd.t3_table = &dummy;
    if (!dd.s1r2_table) dd.s1r2_table = &dummy;
    if (!dd.radial_table) dd.radial_table = &dummy;
    size_t b = (size_t)N * (32 + 8 + 24 + 24 + 24) + 4096;
    // the scratch of either precision (sim_create sizes it for both)
    size_t scratch = 0;
    for (int prec : {STEPS_PREC_FP64, STEPS_PREC_MIXED}) {
        if (prec == STEPS_PREC_MIXED && !desc_mixed_capable(d)) continue;
        dd.precision = prec;
        if (make_plan(dd, N, 0, N - 1, p) == STEPS_OK) scratch = std::max(scratch, scratch_bytes(p));
    }
    b += scratch;
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
    // Simulations by id.  A deque: add_sim appends while the scheduler works
    // on entries through pointers taken under the lock, and appending to a
    // deque leaves existing elements where they are.
    std::deque<SimEntry> sims;
    // Requests by ticket: reqs[t - base].  Requests that are done and whose
    // result was collected leave the front, so a long-running server keeps
    // only the window of tickets still in use.  Erasing at the front and
    // appending leave the other elements in place, so the scheduler may hold
    // a pointer to a request it has taken until it publishes the outcome.
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
    std::vector<Slot> slots;  // the scheduler thread's own (destroy touches them after the join)
    size_t used = 0;          // device bytes of resident simulations, admissions in progress included
    int64_t admitting = 0;    // admissions between decision and publication
    steps_server_stats st{};
    Clock::time_point t0 = Clock::now();

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    int backlog() const {
        int64_t n = 0;
        for (const auto &s : sims) n += (int64_t)s.pending.size();
        return (int)n;
    }

    void finish(Request &r, int status) {
        r.status = status;
        r.t_end = now();
        r.done = true;
        if (status == STEPS_OK) st.completed++;
        else st.failed++;
        cv_done.notify_all();
    }

    // ---- the scheduler's decisions (taken under mu) and the device work they
    // imply (done without it).  Every simulation a decision names is marked
    // busy until the outcome is published, so submit / wait / get_state /
    // get_stats never wait for device work, and no other thread touches its
    // device state meanwhile.
    struct Evict {
        SimEntry *e;
        steps_sim *res;
        int rc = STEPS_OK;
    };
    struct Action {
        Slot *slot = nullptr;
        std::vector<SimEntry *> es;   // one, or the members of a packed launch
        std::vector<Request *> rs;    // their head requests, taken off `pending`
        std::vector<Evict> victims;   // made room for an admission
        bool admit = false;           // es[0] has to become resident first
        steps_sim *admitted = nullptr;
        int rc = STEPS_OK;            // admission or launch status
    };

    // Room for e (not resident) within the budget: least-recently-used idle
    // residents are detached as victims (their device state is downloaded and
    // freed with the action).  0: reserved; 1: no room now (victims already
    // detached still go); <0: it can never fit.
    int plan_admission(SimEntry &e, std::vector<Evict> &victims) {
        if ((int64_t)e.bytes > cfg.hbm_budget_bytes) return fail(STEPS_ERR_ADMISSION, "simulation larger than the budget");
        while ((int64_t)(used + e.bytes) > cfg.hbm_budget_bytes) {
            SimEntry *lru = nullptr;
            for (SimEntry &o : sims) {
                if (&o == &e || !o.res || o.busy) continue;
                if (!lru || o.last_used < lru->last_used) lru = &o;
            }
            if (!lru) return 1;
            victims.push_back(Evict{lru, lru->res});
            lru->res = nullptr;
            lru->busy = true;
            used -= lru->bytes;
        }
        used += e.bytes;
        return 0;
    }

    bool step_without_forces(const SimEntry &e) {
        return req(e.pending.front()).kind == STEPS_REQ_STEP && !e.has_F;
    }

    // Ready resident requests that can share one packed pair launch with the
    // head of ready[qi] (small, same variant and precision, valid plan).
    std::vector<SimEntry *> form_pack(const std::vector<SimEntry *> &ready, size_t qi) {
        SimEntry *e = ready[qi];
        std::vector<SimEntry *> group{e};
        if (cfg.pack_max_n <= 0 || e->N > cfg.pack_max_n || !e->res) return group;
        const int prec = req(e->pending.front()).precision;
        EvalPlan p;
        if (sim_plan(e->res, 0, e->N - 1, prec, p) != STEPS_OK) return group;
        for (size_t qj = qi + 1; qj < ready.size() && (int)group.size() < kMaxPack; ++qj) {
            SimEntry *e2 = ready[qj];
            if (e2->busy || !e2->res || e2->removed || e2->lost || e2->N > cfg.pack_max_n || e2->variant != e->variant)
                continue;
            if (req(e2->pending.front()).precision != prec || step_without_forces(*e2)) continue;
            EvalPlan p2;
            if (sim_plan(e2->res, 0, e2->N - 1, prec, p2) != STEPS_OK) continue;
            e2->busy = true;  // taken: the loop over `ready` skips it
            group.push_back(e2);
        }
        return group;
    }

    // Under mu: requests of removed or lost simulations fail, removed idle
    // residents are detached for destruction, and each free slot gets the
    // oldest ready head request (with its admission) or a packed group.
    bool plan(std::vector<Action> &acts, std::vector<steps_sim *> &reap) {
        bool progressed = false;
        for (SimEntry &e : sims) {
            if (!(e.removed || e.lost)) continue;
            while (!e.pending.empty()) {
                Request &r = req(e.pending.front());
                e.pending.pop_front();
                finish(r, e.removed ? fail(STEPS_ERR_STATE, "simulation removed")
                                    : fail(STEPS_ERR_DEVICE, "simulation state lost in an eviction"));
                progressed = true;
            }
            if (e.removed && !e.busy && e.res) {
                reap.push_back(e.res);
                e.res = nullptr;
                used -= e.bytes;
                progressed = true;
            }
        }
        std::vector<SimEntry *> ready;
        for (SimEntry &e : sims)
            if (!e.busy && !e.pending.empty()) ready.push_back(&e);
        std::sort(ready.begin(), ready.end(),
                  [](const SimEntry *a, const SimEntry *b) { return a->pending.front() < b->pending.front(); });
        for (size_t qi = 0; qi < ready.size(); ++qi) {
            SimEntry &e = *ready[qi];
            if (e.busy) continue;  // packed with an earlier head
            if (step_without_forces(e)) {
                // step.cc:128-140 kicks with the forces of the current
                // positions: a step needs a force request before it
                Request &r = req(e.pending.front());
                e.pending.pop_front();
                finish(r, fail(STEPS_ERR_STATE, "step request before any force evaluation of the simulation"));
                progressed = true;
                continue;
            }
            Slot *free_slot = nullptr;
            for (Slot &sl : slots)
                if (sl.ticket < 0) { free_slot = &sl; break; }
            if (!free_slot) break;
            Action a;
            a.slot = free_slot;
            if (!e.res) {
                const int adm = plan_admission(e, a.victims);
                if (adm != 0) {
                    if (adm < 0) {
                        Request &r = req(e.pending.front());
                        e.pending.pop_front();
                        finish(r, adm);
                        progressed = true;
                    }
                    // no room until something finishes; evictions already
                    // decided still run (an action without a slot)
                    if (!a.victims.empty()) {
                        a.slot = nullptr;
                        acts.push_back(std::move(a));
                        progressed = true;
                    }
                    continue;
                }
                a.admit = true;
                admitting++;
            }
            e.busy = true;
            a.es = a.admit ? std::vector<SimEntry *>{&e} : form_pack(ready, qi);
            const double ts = now();
            for (SimEntry *m : a.es) {
                Request &r = req(m->pending.front());
                m->pending.pop_front();
                r.t_start = ts;
                a.rs.push_back(&r);
            }
            free_slot->ticket = a.rs[0]->ticket;  // reserved; the launch is published later
            acts.push_back(std::move(a));
            progressed = true;
        }
        return progressed;
    }

    // Launch request r on slot sl for resident simulation s (graph captured
    // on first use); returns a status.
    int issue(steps_sim *s, const Request &r, Slot &sl, int64_t &captures) {
        EvalPlan p;
        int rc = sim_plan(s, 0, s->N - 1, r.precision, p);
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
                captures++;
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

    // One packed launch for the head requests of the group: per simulation
    // the step's first half and the source preparation, one pair kernel over
    // all, per simulation the reduction and the step's second half.  No graphs
    // (the composition changes from launch to launch).
    int issue_batch(const Action &a) {
        const size_t n = a.es.size();
        std::vector<EvalPlan> plans(n);
        for (size_t i = 0; i < n; ++i) {
            steps_sim *s = a.es[i]->res;
            const Request &r = *a.rs[i];
            int rc = sim_plan(s, 0, s->N - 1, r.precision, plans[i]);
            if (!rc && r.kind == STEPS_REQ_STEP) rc = sim_fill_params(s, r.sp);
            if (rc) return rc;
        }
        Slot &sl = *a.slot;
        const hipStream_t qs = sl.stream;
        hipError_t he = hipSuccess;
        for (size_t i = 0; i < n && he == hipSuccess; ++i)
            if (a.rs[i]->kind == STEPS_REQ_STEP) he = sim_enqueue_step_before(a.es[i]->res, qs);
        for (size_t i = 0; i < n && he == hipSuccess; ++i) {
            steps_sim *s = a.es[i]->res;
            he = launch_forces_prepare(plans[i], s->dp, s->part.as<double>(), qs);
        }
        int64_t blocks = 0;
        for (size_t i = 0; i < n && he == hipSuccess; ++i) {
            steps_sim *s = a.es[i]->res;
            blocks += pack_describe(plans[i], s->dp, s->part.as<double>(), sl.hseg, (int)i, blocks);
        }
        if (he == hipSuccess)
            he = hipMemcpyAsync(sl.dseg.ptr, sl.hseg, n * pack_desc_bytes(), hipMemcpyHostToDevice, qs);
        if (he == hipSuccess) he = launch_pair_packed(plans[0].variant, plans[0].precision, sl.dseg.ptr, (int)n, blocks, qs);
        for (size_t i = 0; i < n && he == hipSuccess; ++i) {
            steps_sim *s = a.es[i]->res;
            he = launch_forces_finish(plans[i], s->dp, s->part.as<double>(), s->F.as<double>(), qs);
        }
        for (size_t i = 0; i < n && he == hipSuccess; ++i)
            if (a.rs[i]->kind == STEPS_REQ_STEP) he = sim_enqueue_step_after(a.es[i]->res, qs);
        if (he == hipSuccess) he = hipEventRecord(sl.done, qs);
        if (he != hipSuccess) {
            // work already queued on this stream must drain before the
            // simulations can be used again
            (void)hipStreamSynchronize(qs);
            return hip_fail(he, "packed launch");
        }
        return STEPS_OK;
    }

    // Without mu: evictions (state to host, device copy freed), the
    // admission (device copy created from the host state), the launch.
    // Touches only the simulations and requests the action names.
    void execute(Action &a, int64_t &captures) {
        for (Evict &v : a.victims) {
            SimEntry &o = *v.e;
            v.rc = sim_download(v.res, o.x.data(), o.v.data(), o.F.data());
            sim_destroy(v.res);
        }
        if (!a.slot) return;
        SimEntry &e = *a.es[0];
        if (a.admit) {
            steps_force_desc d = e.desc;
            d.t3_table = e.t3.empty() ? nullptr : e.t3.data();
            d.s1r2_table = e.s1r2.empty() ? nullptr : e.s1r2.data();
            d.radial_table = e.radial.empty() ? nullptr : e.radial.data();
            steps_sim *s = nullptr;
            a.rc = sim_create(d, e.N, true, &s);
            if (!a.rc) {
                a.rc = sim_upload(s, e.x.data(), e.v.data(), e.M.data(), e.soft.data(), e.has_F ? e.F.data() : nullptr);
                if (a.rc) sim_destroy(s);
            }
            if (a.rc) return;
            a.admitted = s;
        }
        a.rc = (a.es.size() >= 2) ? issue_batch(a) : issue(a.admit ? a.admitted : e.res, *a.rs[0], *a.slot, captures);
    }

    // Under mu: outcomes of the actions.
    void publish(std::vector<Action> &acts) {
        for (Action &a : acts) {
            for (Evict &v : a.victims) {
                SimEntry &o = *v.e;
                o.busy = false;
                if (v.rc == STEPS_OK) o.has_F = true;
                else o.lost = true;  // its requests fail from now on
                st.evictions++;
            }
            if (!a.slot) continue;
            if (a.admit) {
                admitting--;
                SimEntry &e = *a.es[0];
                if (a.admitted) {
                    e.res = a.admitted;
                    st.admissions++;
                } else {
                    used -= e.bytes;
                }
            }
            if (a.rc == STEPS_OK) {
                if (a.es.size() >= 2) {
                    st.packed_launches++;
                    st.packed_requests += (int64_t)a.es.size();
                }
                a.slot->batch.clear();
                for (Request *r : a.rs) a.slot->batch.push_back(r->ticket);
                a.slot->ticket = a.rs[0]->ticket;
            } else {
                for (size_t i = 0; i < a.es.size(); ++i) {
                    finish(*a.rs[i], a.rc);
                    a.es[i]->busy = false;
                }
                a.slot->ticket = -1;
                a.slot->batch.clear();
            }
        }
        cv_done.notify_all();  // get_state waits for !busy
    }

    // Under mu: requests whose slot event completed.
    bool complete(const std::vector<hipError_t> &q) {
        bool progressed = false;
        for (size_t i = 0; i < slots.size(); ++i) {
            Slot &sl = slots[i];
            if (sl.ticket < 0 || q[i] == hipErrorNotReady) continue;
            for (const int64_t t : sl.batch) {
                Request &r = req(t);
                SimEntry &e = sims[(size_t)r.sim];
                if (q[i] == hipSuccess) {
                    if (r.kind == STEPS_REQ_STEP) r.errmax = *e.res->herrmax;
                    e.has_F = true;
                    finish(r, STEPS_OK);
                } else {
                    finish(r, hip_fail(q[i], "request"));
                }
                e.busy = false;
                e.last_used = now();
            }
            sl.batch.clear();
            sl.ticket = -1;
            progressed = true;
        }
        return progressed;
    }

    void loop() {
        (void)hipSetDevice(cfg.device);
        std::unique_lock<std::mutex> lk(mu, std::defer_lock);
        std::vector<hipError_t> q(slots.size());
        for (;;) {
            // the slots are this thread's: their events are queried unlocked
            for (size_t i = 0; i < slots.size(); ++i)
                q[i] = (slots[i].ticket >= 0) ? hipEventQuery(slots[i].done) : hipErrorNotReady;
            lk.lock();
            bool progressed = complete(q);
            std::vector<Action> acts;
            std::vector<steps_sim *> reap;
            progressed = plan(acts, reap) || progressed;
            lk.unlock();
            for (steps_sim *s : reap) sim_destroy(s);
            int64_t captures = 0;
            for (Action &a : acts) execute(a, captures);
            lk.lock();
            publish(acts);
            st.graph_captures += captures;
            int inflight = 0;
            for (const Slot &sl : slots) inflight += sl.ticket >= 0;
            if (stop && inflight == 0) break;
            if (!progressed) {
                if (inflight > 0) {
                    lk.unlock();
                    std::this_thread::sleep_for(std::chrono::microseconds(20));
                    continue;
                }
                cv_work.wait_for(lk, std::chrono::milliseconds(50));
            }
            lk.unlock();
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
    if (p.bh)
        return fail(STEPS_ERR_ARG, "the server batches direct-sum simulations; run Barnes-Hut ones through steps_sim");
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
    // the scheduler fails its pending requests and frees its device copy once
    // nothing of it is in flight
    S->sims[(size_t)sim].removed = true;
    S->cv_work.notify_all();
    return STEPS_OK;
}

int64_t steps_server_sim_bytes(steps_server *S, int64_t sim) {
    if (!S) return fail(STEPS_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (sim < 0 || sim >= (int64_t)S->sims.size()) return fail(STEPS_ERR_ARG, "unknown simulation");
    return (int64_t)S->sims[(size_t)sim].bytes;
}

int steps_server_submit(steps_server *S, int64_t sim, int32_t kind, int32_t precision, const steps_step_params *sp,
                        int64_t *ticket) {
    if (!S || !ticket) return fail(STEPS_ERR_ARG, "null argument");
    if (kind != STEPS_REQ_FORCE && kind != STEPS_REQ_STEP) return fail(STEPS_ERR_ARG, "unknown request kind");
    if (precision != STEPS_PREC_FP64 && precision != STEPS_PREC_MIXED) return fail(STEPS_ERR_ARG, "unknown precision");
    if (kind == STEPS_REQ_STEP && !sp) return fail(STEPS_ERR_ARG, "step request without parameters");
    std::lock_guard<std::mutex> lk(S->mu);
    if (sim < 0 || sim >= (int64_t)S->sims.size() || S->sims[(size_t)sim].removed)
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
    S->sims[(size_t)sim].pending.push_back(r.ticket);
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
    SimEntry &e = S->sims[(size_t)sim];
    S->cv_done.wait(lk, [&] { return !e.busy && e.pending.empty(); });
    if (e.lost) return fail(STEPS_ERR_DEVICE, "simulation state lost in an eviction");
    if (e.res) {
        // copied without the lock; busy keeps the scheduler off this simulation
        e.busy = true;
        steps_sim *res = e.res;
        lk.unlock();
        const int rc = sim_download(res, x, v, F);
        lk.lock();
        e.busy = false;
        S->cv_done.notify_all();
        S->cv_work.notify_all();
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
    out->admitting = S->admitting;
    out->tickets_held = (int64_t)S->reqs.size();
    int64_t n = 0;
    for (const auto &e : S->sims) n += e.res ? 1 : 0;
    out->resident_sims = n;
    return STEPS_OK;
}

double steps_server_clock(steps_server *S) { return S ? S->now() : 0.0; }

}  // extern "C"
