// sim_internal.h -- resident-simulation machinery shared by the C ABI
// (engine.hip) and the serving loop (server.hip).  Not part of the public ABI.
#pragma once

#include "internal.h"

#include <cstdint>

#define STEPS_HIP(call, what)                                  \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return ::steps::hip_fail(e_, (what)); \
    } while (0)

namespace steps {

int cu_count(int device);
int set_device(int device);
int check_finite(const double *x, int64_t n);

// Grow-only device buffer.
struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    int reserve(size_t bytes, const char *what);
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Device copy of a host table, re-uploaded only when its pointer, size or
// content changes (a hash of every entry).
struct TableCopy {
    const double *host = nullptr;
    size_t n = 0;
    uint64_t fp = 0;
    DevBuf dev;
    int lo = 0;  // T3 tables: first x/y node of the device layout
    int sync(const double *h, size_t count, hipStream_t st, const char *what);
    // T3 / S1R2: upload the table in the pair kernels' layout
    // (t3_device_layout, s1r2_device_layout)
    int sync_t3(const double *h, int ng, hipStream_t st);
    int sync_s1r2(const double *h, int nrho, int nz, hipStream_t st);
    const double *ptr() const { return dev.as<const double>(); }
};

// Validate a descriptor and plan one evaluation (variant, splits, external term).
int make_plan(const steps_force_desc &d, int64_t N, int64_t id_min, int64_t id_max, EvalPlan &p);
// The descriptor selects a variant with a mixed-precision kernel.
inline bool desc_mixed_capable(const steps_force_desc &d) {
    if (d.bh_theta != 0.0) return false;  // the octree is FP64 only
    return d.topology == STEPS_TOPO_R3 || d.topology == STEPS_TOPO_T3 ||
           (d.topology == STEPS_TOPO_S1R2 &&
            (d.is_periodic == 1 || (d.is_periodic >= 2 && d.s1r2_lookup && d.interp_order != 0)));
}
size_t scratch_bytes(const EvalPlan &p);

}  // namespace steps

// Resident simulation: particle state, scratch, tables and per-step scalars
// all in device memory; one stream; optional captured graphs per
// (request kind, precision).
struct steps_sim {
    int device = 0;
    steps_force_desc desc{};
    int64_t N = 0;
    bool both_precisions = false;          // scratch sized for the mixed path too
    steps::DeviceParticles dp;
    steps::DevBuf jrec, jsoft, v, F, part, stage, blockmax, dparams;
    steps::TableCopy t3, s1r2, radial, radial_bh;
    steps::KdkParams *hparams = nullptr;   // pinned: copied to dparams by the step sequence
    double *herrmax = nullptr;             // pinned: errmax copied back by the step sequence
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    hipGraphExec_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [kind][precision]
    int64_t last_nrows = 0;
    bool evaluated = false;  // F holds last_nrows rows of a force evaluation of the current positions
    bool full = false;       // ... and those rows are all N particles in input order (what a KDK step kicks with)
    bool timed = false;      // ev0..ev2 bracket the last force evaluation
};

namespace steps {

// Create/destroy without touching the global error state of other sims.
int sim_create(const steps_force_desc &d, int64_t N, bool both_precisions, steps_sim **out);
void sim_destroy(steps_sim *s);
// Plan of one evaluation of this sim at the given precision.
int sim_plan(steps_sim *s, int64_t id_min, int64_t id_max, int precision, EvalPlan &p);
// Enqueue (no synchronisation): forces of all rows, or a whole KDK step whose
// scalars the caller has written to s->hparams.  Both are graph-capturable.
hipError_t sim_enqueue_forces(steps_sim *s, const EvalPlan &p, hipStream_t st, bool events);
hipError_t sim_enqueue_step(steps_sim *s, const EvalPlan &p, hipStream_t st);
// The step around its force evaluation: parameters in, first kick and drift
// (before); second kick, errmax and its copy to the pinned host block (after).
hipError_t sim_enqueue_step_before(steps_sim *s, hipStream_t st);
hipError_t sim_enqueue_step_after(steps_sim *s, hipStream_t st);
// Fill the pinned parameter block from the public step parameters.
int sim_fill_params(steps_sim *s, const steps_step_params &sp);
// Host <-> device state (synchronous on the sim's stream).
int sim_upload(steps_sim *s, const double *x, const double *v, const double *M, const double *soft,
               const double *F);
int sim_download(steps_sim *s, double *x, double *v, double *F);

}  // namespace steps
