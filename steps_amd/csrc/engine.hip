// engine.hip -- host side of the C ABI (include/steps_force.h).
//
// Owns device memory, streams and table copies.  Unlike the reference host
// wrappers (forces_cuda.cu:878-1117, which cudaMalloc/copy/free six buffers and
// the whole Ewald table on every call), device buffers are grown once and
// reused, and tables are uploaded once per content.
#include "sim_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>
#include <string>

namespace steps {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string &msg) { g_error = msg; }

int fail(int code, const std::string &msg) {
    set_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    set_error(m);
    return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? STEPS_ERR_NOMEM
                                                                        : STEPS_ERR_DEVICE;
}

int cu_count(int device) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0)
        n = 256;
    return n;
}

int DevBuf::reserve(size_t bytes, const char *what) {
    if (bytes <= cap && ptr) return STEPS_OK;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        ptr = nullptr;
        return hip_fail(e, what);
    }
    cap = want;
    return STEPS_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

namespace {
// Hash of every word of a host table (four independent multiply-rotate lanes
// in the manner of xxHash64's round, then folded): a table edited in place
// between calls, at any entry, gets a new fingerprint and is uploaded again.
// It reads the table once (about 1 ms for the 6 MB T3 table of IS_PERIODIC 2),
// which is small next to the O(N^2) evaluation that follows.
uint64_t fingerprint(const double *h, size_t n) {
    constexpr uint64_t P1 = 11400714785074694791ull, P2 = 14029467366897019727ull, P3 = 1609587929392839161ull;
    auto rotl = [](uint64_t x, int r) { return (x << r) | (x >> (64 - r)); };
    auto round = [&](uint64_t acc, uint64_t w) { return rotl(acc + w * P2, 31) * P1; };
    uint64_t a[4] = {P1 + P2, P2, 0, 0 - P1};
    const unsigned char *b = reinterpret_cast<const unsigned char *>(h);
    size_t i = 0;
    for (; i + 4 <= n; i += 4) {
        uint64_t w[4];
        std::memcpy(w, b + i * 8, sizeof w);
        for (int l = 0; l < 4; ++l) a[l] = round(a[l], w[l]);
    }
    uint64_t f = rotl(a[0], 1) + rotl(a[1], 7) + rotl(a[2], 12) + rotl(a[3], 18);
    for (; i < n; ++i) {
        uint64_t w;
        std::memcpy(&w, b + i * 8, sizeof w);
        f = rotl(f ^ round(0, w), 27) * P1 + P3;
    }
    f ^= (uint64_t)n * P3;
    f ^= f >> 33;
    f *= P2;
    f ^= f >> 29;
    return f;
}
}  // namespace

int TableCopy::sync(const double *h, size_t count, hipStream_t st, const char *what) {
    if (!h || count == 0) return STEPS_OK;
    const uint64_t f = fingerprint(h, count);
    if (h == host && count == n && f == fp && dev.ptr) return STEPS_OK;
    int rc = dev.reserve(count * sizeof(double), what);
    if (rc) return rc;
    STEPS_HIP(hipMemcpyAsync(dev.ptr, h, count * sizeof(double), hipMemcpyHostToDevice, st), what);
    STEPS_HIP(hipStreamSynchronize(st), what);
    host = h;
    n = count;
    fp = f;
    return STEPS_OK;
}

int TableCopy::sync_t3(const double *h, int ng, hipStream_t st) {
    const size_t count = (size_t)ng * ng * ng * 3;
    if (!h || ng < 1) return STEPS_OK;
    const uint64_t f = fingerprint(h, count);
    if (h == host && count == n && f == fp && dev.ptr) return STEPS_OK;
    std::vector<double> layout;
    const int l = t3_device_layout(h, ng, layout);
    int rc = dev.reserve(layout.size() * sizeof(double), "T3 table");
    if (rc) return rc;
    STEPS_HIP(hipMemcpyAsync(dev.ptr, layout.data(), layout.size() * sizeof(double), hipMemcpyHostToDevice, st),
              "T3 table");
    STEPS_HIP(hipStreamSynchronize(st), "T3 table");
    host = h;
    n = count;
    fp = f;
    lo = l;
    return STEPS_OK;
}

int TableCopy::sync_s1r2(const double *h, int nrho, int nz, hipStream_t st) {
    const size_t count = (size_t)nrho * nz * 2;
    if (!h || nrho < 1 || nz < 1) return STEPS_OK;
    const uint64_t f = fingerprint(h, count);
    if (h == host && count == n && f == fp && dev.ptr) return STEPS_OK;
    std::vector<double> layout;
    s1r2_device_layout(h, nrho, nz, layout);
    int rc = dev.reserve(layout.size() * sizeof(double), "S1R2 table");
    if (rc) return rc;
    STEPS_HIP(hipMemcpyAsync(dev.ptr, layout.data(), layout.size() * sizeof(double), hipMemcpyHostToDevice, st),
              "S1R2 table");
    STEPS_HIP(hipStreamSynchronize(st), "S1R2 table");
    host = h;
    n = count;
    fp = f;
    return STEPS_OK;
}

int make_plan(const steps_force_desc &d, int64_t N, int64_t id_min, int64_t id_max, EvalPlan &p) {
    if (N < 0) return fail(STEPS_ERR_ARG, "N < 0");
    if (N > (int64_t)INT32_MAX * 4) return fail(STEPS_ERR_ARG, "N too large");
    if (id_max < id_min - 1) return fail(STEPS_ERR_RANGE, "id_max < id_min - 1");
    if (id_max >= id_min && (id_min < 0 || id_max >= N))
        return fail(STEPS_ERR_RANGE, "row range outside [0, N)");
    if (d.precision != STEPS_PREC_FP64 && d.precision != STEPS_PREC_MIXED)
        return fail(STEPS_ERR_ARG, "unknown precision");
    p = EvalPlan{};
    p.N = N;
    p.id_min = id_min;
    p.nrows = std::max<int64_t>(0, id_max - id_min + 1);
    p.L = d.L;
    p.Rsim = d.Rsim;
    p.is_periodic = d.is_periodic;
    p.cosmology = d.cosmology;
    p.comoving = d.comoving;
    p.precision = d.precision;
    const double DE = d.H0 * d.H0 * d.Omega_lambda;  // forces_cuda.cu:887
    switch (d.topology) {
    case STEPS_TOPO_R3:
        p.variant = PV_R3;
        if (d.cosmology == 1) {
            p.ext_mode = 1;
            p.ext_coeff = (d.comoving == 1) ? d.mass_in_unit_sphere : DE;
        }
        break;
    case STEPS_TOPO_T3:
        if (d.is_periodic < 1) return fail(STEPS_ERR_ARG, "T3 needs IS_PERIODIC >= 1 (main.cc:405)");
        if (!(d.L > 0)) return fail(STEPS_ERR_ARG, "T3 needs L > 0");
        if (d.is_periodic == 1) {
            p.variant = PV_T3_NEAREST;
        } else {
            if (!d.t3_table || d.t3_ngrid < 1) return fail(STEPS_ERR_TABLE, "T3 Ewald table missing");
            if (d.t3_ngrid < 3 || d.t3_ngrid > kT3MaxGrid)
                return fail(STEPS_ERR_TABLE, "T3 table grid must have 3 to 1023 nodes per axis");
            p.variant = PV_T3_EWALD;
            p.t3_ng = d.t3_ngrid;
        }
        break;
    case STEPS_TOPO_S1R2: {
        if (d.is_periodic < 1) return fail(STEPS_ERR_ARG, "S1R2 needs IS_PERIODIC >= 1 (main.cc:536)");
        if (!(d.L > 0)) return fail(STEPS_ERR_ARG, "S1R2 needs L > 0");
        const bool radial_used = (d.is_periodic == 1 || !d.s1r2_lookup);
        if (d.is_periodic == 1) {
            p.variant = PV_S1R2_NEAREST;
        } else if (d.s1r2_lookup) {
            if (!d.s1r2_table || d.s1r2_nrho < 2 || d.s1r2_nz < 1)
                return fail(STEPS_ERR_TABLE, "S1R2 Ewald table missing (or fewer than 2 rho nodes)");
            if (s1r2_device_nodes(d.s1r2_nrho, d.s1r2_nz) >= ((size_t)1 << 28))
                return fail(STEPS_ERR_TABLE, "S1R2 table too large (2^28 nodes at most)");
            if (d.s1r2_nz + 3 >= (1 << 24) || d.s1r2_nrho + 2 >= (1 << 24))  // 24-bit row offsets in the kernels
                return fail(STEPS_ERR_TABLE, "S1R2 table dimensions must stay below 2^24 - 3 nodes");
            p.variant = d.interp_order == 0 ? PV_S1R2_NGP
                        : d.interp_order == 2 ? PV_S1R2_CIC
                                              : PV_S1R2_TSC;  // forces_cuda.cu:449-451 fallback
            p.s1r2_nrho = d.s1r2_nrho;
            p.s1r2_nz = d.s1r2_nz;
        } else {
            p.variant = PV_S1R2_IMAGES;
            p.images = d.is_periodic + 1;  // forces_cuda.cu:659-660
        }
        if (d.cosmology == 1 && d.comoving == 1) {
            p.ext_coeff = d.mass_in_unit_sphere;
            p.ext_mode = radial_used ? 3 : 2;
            if (radial_used) {
                if (!d.radial_table || d.radial_table_size < 2)
                    return fail(STEPS_ERR_TABLE, "S1R2 radial force table missing");
                if (!(d.Rsim > 0)) return fail(STEPS_ERR_ARG, "S1R2 needs Rsim > 0");
                p.radial_size = d.radial_table_size;
            }
        } else if (d.cosmology == 1) {
            p.ext_mode = 2;
            p.ext_coeff = DE;
        }
        if (p.variant >= PV_S1R2_NGP && p.variant <= PV_S1R2_TSC && !(d.Rsim > 0))
            return fail(STEPS_ERR_ARG, "S1R2 table needs Rsim > 0");
        break;
    }
    default:
        return fail(STEPS_ERR_ARG, "unknown topology");
    }
    if (d.bh_theta != 0.0) {
        // USE_BH (main.cc:302-303): the octree instead of the direct sum
        if (!(d.bh_theta > 0.0) || !std::isfinite(d.bh_theta))
            return fail(STEPS_ERR_ARG, "Barnes-Hut opening angle must be positive and finite");
        if (!(d.bh_particle_radius >= 0.0) || !std::isfinite(d.bh_particle_radius))
            return fail(STEPS_ERR_ARG, "Barnes-Hut particle radius must be non-negative and finite");
        if (d.precision != STEPS_PREC_FP64) return fail(STEPS_ERR_ARG, "Barnes-Hut forces are FP64 only");
        if (bh_item_capacity(N) > INT32_MAX) return fail(STEPS_ERR_ARG, "N too large for the octree");
        p.bh = true;
        p.bh_theta = d.bh_theta;
        p.bh_particle_radius = d.bh_particle_radius;
        // the CPU build's T3 table interpolation (ewald_space.cc:452-466): even
        // orders up to 8, anything else falls back to 2
        p.bh_order = d.interp_order;
        if (p.bh_order < 2 || p.bh_order % 2 != 0) p.bh_order = 2;
        if (p.bh_order > 8) p.bh_order = 8;
        if (d.bh_randomize) {
            // RANDOMIZE_BH: draws rand()/RAND_MAX keep the moved root cube
            // around the particles
            p.bh_randomize = true;
            for (int k = 0; k < 6; ++k) {
                if (!(d.bh_random[k] >= 0.0 && d.bh_random[k] <= 1.0))
                    return fail(STEPS_ERR_ARG, "randomised octree draws must lie in [0, 1]");
                p.bh_u[k] = d.bh_random[k];
            }
        }
        if (d.radial_bh_table) {
            // USE_RADIAL_BH_CORRECTION (main.cc:791-1091): R3 and S1R2 only,
            // applied with the comoving external term
            if (d.topology == STEPS_TOPO_T3)
                return fail(STEPS_ERR_ARG, "the radial octree correction exists for R3 and S1R2 only");
            if (d.radial_bh_table_size < 2) return fail(STEPS_ERR_TABLE, "radial octree correction table needs 2 bins");
            if (!(d.Rsim > 0)) return fail(STEPS_ERR_ARG, "radial octree correction needs Rsim > 0");
            if (d.cosmology == 1 && d.comoving == 1) p.bh_radial_size = d.radial_bh_table_size;
        }
    } else if (d.bh_randomize || d.radial_bh_table) {
        return fail(STEPS_ERR_ARG, "randomisation and the radial correction need the octree (bh_theta > 0)");
    }
    if (p.precision == STEPS_PREC_MIXED && !mixed_supported(p.variant))
        return fail(STEPS_ERR_ARG, "mixed precision is implemented for R3, T3 and S1R2 (nearest z-image, CIC or TSC correction)");

    // split the source range so that (row blocks) x (splits) fills the chip
    const int ni = rows_per_lane(p.variant, p.precision);
    const int64_t rows_per_block = (int64_t)kBlock * ni;
    const int64_t nib = std::max<int64_t>(1, (p.nrows + rows_per_block - 1) / rows_per_block);
    const int64_t target = (int64_t)cu_count(d.device) * 8;
    int64_t splits = (target + nib - 1) / nib;
    splits = std::max<int64_t>(1, std::min<int64_t>(splits, std::max<int64_t>(1, N / 256)));
    // Source chunks of at most kL2Sources (2.6 MB of records + 0.5 MB of
    // softening lengths, under the 4 MB L2 of an XCD): workgroups resident at
    // the same time work on the same chunk (row blocks vary fastest in the
    // grid), so the scalar loads of the source records hit L2.  At N = 2^21
    // this took the R3 FP64 kernel from 2.59 s to 2.40 s and its L2 misses
    // from 420 GB to 1.8 GB.  The partial sums it costs are capped at 2 GiB.
    constexpr int64_t kL2Sources = 65536;
    constexpr int64_t kMaxPartialBytes = (int64_t)2 << 30;
    splits = std::max<int64_t>(splits, (N + kL2Sources - 1) / kL2Sources);
    splits = std::min<int64_t>(splits, std::max<int64_t>(1, kMaxPartialBytes / (24 * std::max<int64_t>(p.nrows, 1))));
    int64_t jchunk = std::max<int64_t>(1, (N + splits - 1) / splits);
    jchunk = ((jchunk + kTile - 1) / kTile) * kTile;  // splits start on tile boundaries
    splits = std::max<int64_t>(1, (N + jchunk - 1) / jchunk);
    p.splits = (int)splits;
    p.jchunk = jchunk;
    if (t3_slab_plan(p)) {
        // one split per XCD; the packed (server) launches, which keep the
        // generic kernel, split the sources into eight equal chunks instead
        p.jchunk = ((N + 7) / 8 + kTile - 1) / kTile * kTile;
        p.splits = 8;
    }
    return STEPS_OK;
}

size_t scratch_bytes(const EvalPlan &p) {
    if (p.bh) return bh_scratch_bytes(p.N);
    return ((partial_bytes(p) + 255) & ~(size_t)255) + tile_extra_bytes(p);
}

int set_device(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return fail(STEPS_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(STEPS_ERR_DEVICE, "device ordinal out of range");
    STEPS_HIP(hipSetDevice(device), "hipSetDevice");
    return STEPS_OK;
}

int check_finite(const double *x, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(x[i])) return fail(STEPS_ERR_NONFINITE, "non-finite input value");
    return STEPS_OK;
}

// ---------------------------------------------------------------------------
// resident simulation
// ---------------------------------------------------------------------------

int sim_create(const steps_force_desc &d, int64_t N, bool both_precisions, steps_sim **out) {
    *out = nullptr;
    if (N <= 0) return fail(STEPS_ERR_ARG, "N must be positive");
    EvalPlan p;
    int rc = make_plan(d, N, 0, N - 1, p);
    if (rc) return rc;
    size_t scratch = scratch_bytes(p);
    if (both_precisions && desc_mixed_capable(d)) {
        // the scratch serves either precision (neither plan's need bounds the other's)
        steps_force_desc dd = d;
        dd.precision = (d.precision == STEPS_PREC_MIXED) ? STEPS_PREC_FP64 : STEPS_PREC_MIXED;
        EvalPlan q;
        if ((rc = make_plan(dd, N, 0, N - 1, q))) return rc;
        scratch = std::max(scratch, scratch_bytes(q));
    }
    if ((rc = set_device(d.device))) return rc;
    steps_sim *s = new (std::nothrow) steps_sim();
    if (!s) return fail(STEPS_ERR_NOMEM, "host allocation");
    s->device = d.device;
    s->desc = d;
    s->N = N;
    s->both_precisions = both_precisions;
    const size_t nb = (size_t)N * sizeof(double);
    auto cleanup = [&](int code) {
        sim_destroy(s);
        return code;
    };
    if ((rc = s->jrec.reserve((size_t)N * sizeof(double4), "alloc jrec")) ||
        (rc = s->jsoft.reserve(nb, "alloc soft")) || (rc = s->v.reserve(3 * nb, "alloc v")) ||
        (rc = s->F.reserve(3 * nb, "alloc F")) || (rc = s->part.reserve(scratch, "alloc partials")) ||
        (rc = s->stage.reserve(std::max(3 * nb, (size_t)1024), "alloc stage")) ||
        (rc = s->blockmax.reserve((size_t)(kdk_reduce_blocks(N) + 1) * sizeof(double), "alloc blockmax")) ||
        (rc = s->dparams.reserve(sizeof(KdkParams), "alloc step params")))
        return cleanup(rc);
    hipError_t e;
    if ((e = hipHostMalloc((void **)&s->hparams, sizeof(KdkParams) + sizeof(double), hipHostMallocDefault)) !=
        hipSuccess) {
        s->hparams = nullptr;
        return cleanup(hip_fail(e, "pinned step block"));
    }
    s->herrmax = (double *)(s->hparams + 1);
    std::memset(s->hparams, 0, sizeof(KdkParams) + sizeof(double));
    if ((e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&s->ev0)) != hipSuccess || (e = hipEventCreate(&s->ev1)) != hipSuccess ||
        (e = hipEventCreate(&s->ev2)) != hipSuccess)
        return cleanup(hip_fail(e, "stream/event create"));
    if (p.variant == PV_T3_EWALD && (rc = s->t3.sync_t3(d.t3_table, d.t3_ngrid, s->stream)))
        return cleanup(rc);
    if (p.variant >= PV_S1R2_NGP && p.variant <= PV_S1R2_TSC &&
        (rc = s->s1r2.sync_s1r2(d.s1r2_table, d.s1r2_nrho, d.s1r2_nz, s->stream)))
        return cleanup(rc);
    if (p.ext_mode == 3 &&
        (rc = s->radial.sync(d.radial_table, (size_t)d.radial_table_size, s->stream, "radial table")))
        return cleanup(rc);
    if (p.bh_radial_size > 0 &&
        (rc = s->radial_bh.sync(d.radial_bh_table, (size_t)d.radial_bh_table_size, s->stream, "radial octree table")))
        return cleanup(rc);
    // the sim owns device copies: host table pointers are not kept
    s->desc.t3_table = nullptr;
    s->desc.s1r2_table = nullptr;
    s->desc.radial_table = nullptr;
    s->desc.radial_bh_table = nullptr;
    s->dp = DeviceParticles{s->jrec.as<double4>(), s->jsoft.as<double>(), N};
    if ((e = hipMemsetAsync(s->v.ptr, 0, 3 * nb, s->stream)) != hipSuccess ||
        (e = hipMemsetAsync(s->F.ptr, 0, 3 * nb, s->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(s->stream)) != hipSuccess)
        return cleanup(hip_fail(e, "initialise sim"));
    *out = s;
    return STEPS_OK;
}

void sim_destroy(steps_sim *s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto &row : s->graph)
        for (auto &g : row)
            if (g) (void)hipGraphExecDestroy(g);
    s->jrec.release(); s->jsoft.release(); s->v.release(); s->F.release(); s->part.release();
    s->stage.release(); s->blockmax.release(); s->dparams.release();
    s->t3.dev.release(); s->s1r2.dev.release(); s->radial.dev.release(); s->radial_bh.dev.release();
    if (s->hparams) (void)hipHostFree(s->hparams);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->ev2) (void)hipEventDestroy(s->ev2);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int sim_plan(steps_sim *s, int64_t id_min, int64_t id_max, int precision, EvalPlan &p) {
    steps_force_desc d = s->desc;  // tables live on the device; the checks only need non-null
    d.t3_table = s->t3.ptr();
    d.s1r2_table = s->s1r2.ptr();
    d.radial_table = s->radial.ptr();
    d.radial_bh_table = s->radial_bh.ptr();
    d.precision = precision;
    int rc = make_plan(d, s->N, id_min, id_max, p);
    if (rc) return rc;
    if (scratch_bytes(p) > s->part.cap) return fail(STEPS_ERR_STATE, "scratch too small for this precision");
    p.t3 = s->t3.ptr();
    p.t3_lo = s->t3.lo;
    p.s1r2 = s->s1r2.ptr();
    p.radial = s->radial.ptr();
    p.bh_radial = s->radial_bh.ptr();
    return STEPS_OK;
}

hipError_t sim_enqueue_forces(steps_sim *s, const EvalPlan &p, hipStream_t st, bool events) {
    return launch_forces(p, s->dp, s->part.as<double>(), s->F.as<double>(), st, events ? s->ev0 : nullptr,
                         events ? s->ev1 : nullptr, events ? s->ev2 : nullptr);
}

hipError_t sim_enqueue_step_before(steps_sim *s, hipStream_t st) {
    // step.cc:128-181: params in, kick + drift + periodic wrap
    hipError_t e;
    if ((e = hipMemcpyAsync(s->dparams.ptr, s->hparams, sizeof(KdkParams), hipMemcpyHostToDevice, st)) != hipSuccess)
        return e;
    return launch_kick_drift(s->dp, s->v.as<double>(), s->F.as<double>(), s->dparams.as<const KdkParams>(),
                             s->desc.topology, s->desc.L, st);
}

hipError_t sim_enqueue_step_after(steps_sim *s, hipStream_t st) {
    // step.cc:250-269: second kick + errmax, errmax out
    const int nb = kdk_reduce_blocks(s->N);
    double *bm = s->blockmax.as<double>();
    hipError_t e = launch_kick_errmax(s->dp, s->v.as<double>(), s->F.as<double>(), s->dparams.as<const KdkParams>(),
                                      bm, nb, bm + nb, st, true);
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(s->herrmax, bm + nb, sizeof(double), hipMemcpyDeviceToHost, st);
}

hipError_t sim_enqueue_step(steps_sim *s, const EvalPlan &p, hipStream_t st) {
    // step.cc:100-312
    hipError_t e;
    if ((e = sim_enqueue_step_before(s, st)) != hipSuccess) return e;
    if ((e = launch_forces(p, s->dp, s->part.as<double>(), s->F.as<double>(), st, nullptr, nullptr, nullptr)) !=
        hipSuccess)
        return e;
    return sim_enqueue_step_after(s, st);
}

int sim_fill_params(steps_sim *s, const steps_step_params &sp) {
    if (!(sp.rho_part > 0)) return fail(STEPS_ERR_ARG, "rho_part must be positive");
    if (!(sp.a_kick1 > 0) || !(sp.a_kick2 > 0)) return fail(STEPS_ERR_ARG, "scale factors must be positive");
    KdkParams &k = *s->hparams;
    k.G = sp.G;
    k.h = sp.h;
    k.ainv3_1 = std::pow(sp.a_kick1, -3.0);  // step.cc:135: pow(a, -3.0)
    k.hubble1 = sp.hubble1;
    k.ainv3_2 = std::pow(sp.a_kick2, -3.0);
    k.hubble2 = sp.hubble2;
    k.const_beta = 3.0 / sp.rho_part / (4.0 * kPi);  // step.cc:250
    return STEPS_OK;
}

int sim_upload(steps_sim *s, const double *x, const double *v, const double *M, const double *soft,
               const double *F) {
    STEPS_HIP(hipSetDevice(s->device), "hipSetDevice");
    const size_t nb = (size_t)s->N * sizeof(double);
    double *xs = s->stage.as<double>();
    // M and soft are staged in the v buffer (v is written right after)
    double *ms = s->v.as<double>();
    double *hs = ms + s->N;
    STEPS_HIP(hipMemcpyAsync(xs, x, 3 * nb, hipMemcpyHostToDevice, s->stream), "copy x");
    STEPS_HIP(hipMemcpyAsync(ms, M, nb, hipMemcpyHostToDevice, s->stream), "copy M");
    STEPS_HIP(hipMemcpyAsync(hs, soft, nb, hipMemcpyHostToDevice, s->stream), "copy soft");
    STEPS_HIP(launch_pack(xs, ms, hs, s->N, s->dp, s->stream), "pack");
    if (v) STEPS_HIP(hipMemcpyAsync(s->v.ptr, v, 3 * nb, hipMemcpyHostToDevice, s->stream), "copy v");
    else STEPS_HIP(hipMemsetAsync(s->v.ptr, 0, 3 * nb, s->stream), "zero v");
    if (F) STEPS_HIP(hipMemcpyAsync(s->F.ptr, F, 3 * nb, hipMemcpyHostToDevice, s->stream), "copy F");
    else STEPS_HIP(hipMemsetAsync(s->F.ptr, 0, 3 * nb, s->stream), "zero F");
    STEPS_HIP(hipStreamSynchronize(s->stream), "upload state");
    // forces handed in with the state are the full-range forces of these positions
    s->evaluated = s->full = (F != nullptr);
    s->last_nrows = F ? s->N : 0;
    s->timed = false;
    return STEPS_OK;
}

int sim_download(steps_sim *s, double *x, double *v, double *F) {
    STEPS_HIP(hipSetDevice(s->device), "hipSetDevice");
    const size_t nb = (size_t)s->N * sizeof(double);
    if (x) {
        STEPS_HIP(launch_unpack_positions(s->dp, s->stage.as<double>(), s->stream), "unpack");
        STEPS_HIP(hipMemcpyAsync(x, s->stage.ptr, 3 * nb, hipMemcpyDeviceToHost, s->stream), "copy x");
    }
    if (v) STEPS_HIP(hipMemcpyAsync(v, s->v.ptr, 3 * nb, hipMemcpyDeviceToHost, s->stream), "copy v");
    if (F) STEPS_HIP(hipMemcpyAsync(F, s->F.ptr, 3 * nb, hipMemcpyDeviceToHost, s->stream), "copy F");
    STEPS_HIP(hipStreamSynchronize(s->stream), "download state");
    return STEPS_OK;
}

namespace {

// State of the one-shot entry point (one device at a time, under a lock).
struct OneShot {
    std::mutex mu;
    int device = -1;
    hipStream_t stream = nullptr;
    DevBuf x, M, soft, jrec, jsoft, part, F;
    TableCopy t3, s1r2, radial, radial_bh;
};
OneShot g_oneshot;

}  // namespace
}  // namespace steps

using namespace steps;

extern "C" {

void steps_force_desc_init(steps_force_desc *d) {
    if (!d) return;
    std::memset(d, 0, sizeof *d);
    d->topology = STEPS_TOPO_R3;
    d->interp_order = 2;
    d->s1r2_lookup = 1;
    d->precision = STEPS_PREC_FP64;
}

int steps_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char *steps_last_error(void) { return g_error.c_str(); }

const char *steps_version(void) { return "steps_amd 0.2 (gfx950)"; }

int steps_force_eval(const steps_force_desc *d, int64_t N, const double *x, const double *M,
                     const double *soft, double *F, int64_t id_min, int64_t id_max) {
    if (!d) return fail(STEPS_ERR_ARG, "null descriptor");
    EvalPlan p;
    int rc = make_plan(*d, N, id_min, id_max, p);
    if (rc) return rc;
    if (p.nrows == 0) return STEPS_OK;
    if (!x || !M || !soft || !F) return fail(STEPS_ERR_ARG, "null buffer");
    if ((rc = check_finite(x, 3 * N))) return rc;
    std::lock_guard<std::mutex> lock(g_oneshot.mu);
    OneShot &o = g_oneshot;
    if ((rc = set_device(d->device))) return rc;
    if (o.device != d->device) {
        // buffers belong to the previous device: drop them
        o.x.release(); o.M.release(); o.soft.release(); o.jrec.release(); o.jsoft.release();
        o.part.release(); o.F.release();
        o.t3 = TableCopy{}; o.s1r2 = TableCopy{}; o.radial = TableCopy{}; o.radial_bh = TableCopy{};
        if (o.stream) (void)hipStreamDestroy(o.stream);
        o.stream = nullptr;
        STEPS_HIP(hipStreamCreateWithFlags(&o.stream, hipStreamNonBlocking), "hipStreamCreate");
        o.device = d->device;
    }
    const size_t nb = (size_t)N * sizeof(double);
    if ((rc = o.x.reserve(3 * nb, "alloc x")) || (rc = o.M.reserve(nb, "alloc M")) ||
        (rc = o.soft.reserve(nb, "alloc soft")) || (rc = o.jrec.reserve((size_t)N * sizeof(double4), "alloc jrec")) ||
        (rc = o.jsoft.reserve(nb, "alloc jsoft")) || (rc = o.part.reserve(scratch_bytes(p), "alloc partials")) ||
        (rc = o.F.reserve(3 * (size_t)p.nrows * sizeof(double), "alloc F")))
        return rc;
    hipStream_t st = o.stream;
    if (p.variant == PV_T3_EWALD && (rc = o.t3.sync_t3(d->t3_table, d->t3_ngrid, st))) return rc;
    if (p.variant >= PV_S1R2_NGP && p.variant <= PV_S1R2_TSC &&
        (rc = o.s1r2.sync_s1r2(d->s1r2_table, d->s1r2_nrho, d->s1r2_nz, st)))
        return rc;
    if (p.ext_mode == 3 && (rc = o.radial.sync(d->radial_table, (size_t)d->radial_table_size, st, "radial table")))
        return rc;
    if (p.bh_radial_size > 0 &&
        (rc = o.radial_bh.sync(d->radial_bh_table, (size_t)d->radial_bh_table_size, st, "radial octree table")))
        return rc;
    p.t3 = o.t3.ptr();
    p.t3_lo = o.t3.lo;
    p.s1r2 = o.s1r2.ptr();
    p.radial = o.radial.ptr();
    p.bh_radial = o.radial_bh.ptr();
    STEPS_HIP(hipMemcpyAsync(o.x.ptr, x, 3 * nb, hipMemcpyHostToDevice, st), "copy x");
    STEPS_HIP(hipMemcpyAsync(o.M.ptr, M, nb, hipMemcpyHostToDevice, st), "copy M");
    STEPS_HIP(hipMemcpyAsync(o.soft.ptr, soft, nb, hipMemcpyHostToDevice, st), "copy soft");
    DeviceParticles dp{o.jrec.as<double4>(), o.jsoft.as<double>(), N};
    STEPS_HIP(launch_pack(o.x.as<double>(), o.M.as<double>(), o.soft.as<double>(), N, dp, st), "pack");
    STEPS_HIP(launch_forces(p, dp, o.part.as<double>(), o.F.as<double>(), st, nullptr, nullptr, nullptr),
              "force kernels");
    STEPS_HIP(hipMemcpyAsync(F, o.F.ptr, 3 * (size_t)p.nrows * sizeof(double), hipMemcpyDeviceToHost, st),
              "copy F");
    STEPS_HIP(hipStreamSynchronize(st), "force evaluation");
    return STEPS_OK;
}

// ---------------------------------------------------------------------------

int steps_sim_create(const steps_force_desc *d, int64_t N, steps_sim **out) {
    if (!d || !out) return fail(STEPS_ERR_ARG, "null argument");
    return sim_create(*d, N, true, out);
}

int steps_sim_destroy(steps_sim *s) {
    sim_destroy(s);
    return STEPS_OK;
}

int steps_sim_set_bh_random(steps_sim *s, const double *u) {
    if (!s || !u) return fail(STEPS_ERR_ARG, "null argument");
    if (!s->desc.bh_randomize) return fail(STEPS_ERR_STATE, "the simulation's octree is not randomised");
    for (int k = 0; k < 6; ++k)
        if (!(u[k] >= 0.0 && u[k] <= 1.0)) return fail(STEPS_ERR_ARG, "randomised octree draws must lie in [0, 1]");
    std::memcpy(s->desc.bh_random, u, sizeof s->desc.bh_random);
    return STEPS_OK;
}

int steps_sim_set_precision(steps_sim *s, int32_t precision) {
    if (!s) return fail(STEPS_ERR_ARG, "null sim");
    if (precision != STEPS_PREC_FP64 && precision != STEPS_PREC_MIXED) return fail(STEPS_ERR_ARG, "unknown precision");
    if (precision == STEPS_PREC_MIXED && !desc_mixed_capable(s->desc))
        return fail(STEPS_ERR_ARG, "mixed precision is implemented for R3, T3 and S1R2 (nearest z-image, CIC or TSC correction)");
    s->desc.precision = precision;
    return STEPS_OK;
}

int steps_sim_set_state(steps_sim *s, const double *x, const double *v, const double *M,
                        const double *soft) {
    if (!s || !x || !M || !soft) return fail(STEPS_ERR_ARG, "null argument");
    int rc = check_finite(x, 3 * s->N);
    if (rc) return rc;
    return sim_upload(s, x, v, M, soft, nullptr);
}

int steps_sim_set_positions(steps_sim *s, const double *x) {
    if (!s || !x) return fail(STEPS_ERR_ARG, "null argument");
    int rc = check_finite(x, 3 * s->N);
    if (rc) return rc;
    STEPS_HIP(hipSetDevice(s->device), "hipSetDevice");
    const size_t nb = (size_t)s->N * sizeof(double);
    STEPS_HIP(hipMemcpyAsync(s->stage.ptr, x, 3 * nb, hipMemcpyHostToDevice, s->stream), "copy x");
    STEPS_HIP(launch_pack_positions(s->stage.as<double>(), s->N, s->dp, s->stream), "pack positions");
    STEPS_HIP(hipStreamSynchronize(s->stream), "set positions");
    s->evaluated = s->full = s->timed = false;
    return STEPS_OK;
}

int steps_sim_get_state(steps_sim *s, double *x, double *v) {
    if (!s) return fail(STEPS_ERR_ARG, "null sim");
    return sim_download(s, x, v, nullptr);
}

int steps_sim_forces(steps_sim *s, int64_t id_min, int64_t id_max) {
    if (!s) return fail(STEPS_ERR_ARG, "null sim");
    EvalPlan p;
    int rc = sim_plan(s, id_min, id_max, s->desc.precision, p);
    if (rc) return rc;
    STEPS_HIP(hipSetDevice(s->device), "hipSetDevice");
    s->evaluated = s->full = s->timed = false;  // F is overwritten from here on
    STEPS_HIP(sim_enqueue_forces(s, p, s->stream, true), "force kernels");
    s->last_nrows = p.nrows;
    s->evaluated = s->timed = true;
    s->full = p.id_min == 0 && p.nrows == s->N;
    return STEPS_OK;
}

int steps_sim_get_forces(steps_sim *s, double *F) {
    if (!s || !F) return fail(STEPS_ERR_ARG, "null argument");
    if (!s->evaluated) return fail(STEPS_ERR_STATE, "no force evaluation queued");
    STEPS_HIP(hipSetDevice(s->device), "hipSetDevice");
    STEPS_HIP(hipMemcpyAsync(F, s->F.ptr, 3 * (size_t)s->last_nrows * sizeof(double), hipMemcpyDeviceToHost,
                             s->stream),
              "copy F");
    STEPS_HIP(hipStreamSynchronize(s->stream), "get forces");
    return STEPS_OK;
}

int steps_sim_synchronize(steps_sim *s) {
    if (!s) return fail(STEPS_ERR_ARG, "null sim");
    STEPS_HIP(hipSetDevice(s->device), "hipSetDevice");
    STEPS_HIP(hipStreamSynchronize(s->stream), "synchronize");
    return STEPS_OK;
}

int steps_sim_last_kernel_ms(steps_sim *s, float *pair_ms, float *total_ms) {
    if (!s) return fail(STEPS_ERR_ARG, "null sim");
    if (!s->timed) return fail(STEPS_ERR_STATE, "no timed force evaluation queued (steps carry no timing events)");
    STEPS_HIP(hipEventSynchronize(s->ev2), "event sync");
    float a = 0, b = 0;
    STEPS_HIP(hipEventElapsedTime(&a, s->ev0, s->ev1), "elapsed");
    STEPS_HIP(hipEventElapsedTime(&b, s->ev0, s->ev2), "elapsed");
    if (pair_ms) *pair_ms = a;
    if (total_ms) *total_ms = b;
    return STEPS_OK;
}

int steps_sim_step(steps_sim *s, const steps_step_params *sp, double *errmax) {
    if (!s || !sp) return fail(STEPS_ERR_ARG, "null argument");
    STEPS_HIP(hipSetDevice(s->device), "hipSetDevice");
    EvalPlan p;
    int rc = sim_plan(s, 0, s->N - 1, s->desc.precision, p);
    if (rc) return rc;
    if ((rc = sim_fill_params(s, *sp))) return rc;
    // step.cc:128-140 kicks with the forces of the current positions, all N rows
    if (!s->full)
        return fail(STEPS_ERR_STATE, "a KDK step needs the full-range forces of the current positions "
                                     "(call steps_sim_forces(s, 0, N-1) first)");
    s->evaluated = s->full = s->timed = false;
    STEPS_HIP(sim_enqueue_step(s, p, s->stream), "KDK step");
    STEPS_HIP(hipStreamSynchronize(s->stream), "KDK step");
    s->last_nrows = p.nrows;
    s->evaluated = s->full = true;  // the step's second force evaluation, all rows (no timing events)
    if (errmax) *errmax = *s->herrmax;
    return STEPS_OK;
}

int steps_sim_errmax(steps_sim *s, double a, double hubble, double G, double rho_part, double *errmax) {
    // step.cc:35-98: errmax of the stored forces, no kick
    if (!s) return fail(STEPS_ERR_ARG, "null sim");
    if (!s->full) return fail(STEPS_ERR_STATE, "errmax needs the full-range forces of the current positions");
    STEPS_HIP(hipSetDevice(s->device), "hipSetDevice");
    steps_step_params sp{0.0, a, hubble, a, hubble, G, rho_part};
    int rc = sim_fill_params(s, sp);
    if (rc) return rc;
    const int nb = kdk_reduce_blocks(s->N);
    double *bm = s->blockmax.as<double>();
    STEPS_HIP(hipMemcpyAsync(s->dparams.ptr, s->hparams, sizeof(KdkParams), hipMemcpyHostToDevice, s->stream),
              "params");
    STEPS_HIP(launch_kick_errmax(s->dp, s->v.as<double>(), s->F.as<double>(), s->dparams.as<const KdkParams>(), bm,
                                 nb, bm + nb, s->stream, false),
              "errmax");
    STEPS_HIP(hipMemcpyAsync(s->herrmax, bm + nb, sizeof(double), hipMemcpyDeviceToHost, s->stream), "errmax");
    STEPS_HIP(hipStreamSynchronize(s->stream), "errmax");
    if (errmax) *errmax = *s->herrmax;
    return STEPS_OK;
}

int steps_softening_lengths(const double *M, int64_t N, double particle_radius, double *soft,
                            double *M_min, double *rho_part) {
    // utils.cc:59-82
    if (!M || !soft || N <= 0) return fail(STEPS_ERR_ARG, "bad softening arguments");
    if (!(particle_radius > 0)) return fail(STEPS_ERR_ARG, "particle radius must be positive");
    double mmin = M[0];
    for (int64_t i = 0; i < N; ++i)
        if (mmin > M[i]) mmin = M[i];
    const double rp = mmin / (4.0 * kPi * std::pow(particle_radius, 3.0) / 3.0);
    const double cb = 3.0 / rp / (4.0 * kPi);
    for (int64_t i = 0; i < N; ++i) soft[i] = std::cbrt(M[i] * cb);
    if (M_min) *M_min = mmin;
    if (rho_part) *rho_part = rp;
    return STEPS_OK;
}

int steps_t3_table(int32_t is_periodic, double L, double *table, int32_t *ngrid) {
    int ng = 0;
    int rc = build_t3_table(is_periodic, L, table, &ng);
    if (ngrid) *ngrid = ng;
    return rc;
}

int steps_s1r2_table(int32_t is_periodic, double L, double Rsim, double *table, int32_t *nrho,
                     int32_t *nz) {
    int a = 0, b = 0;
    int rc = build_s1r2_table(is_periodic, L, Rsim, table, &a, &b);
    if (nrho) *nrho = a;
    if (nz) *nz = b;
    return rc;
}

int steps_radial_table(double R, double Lz, int32_t size, int32_t accuracy, double *table) {
    return build_radial_table(R, Lz, size, accuracy, table);
}

// utils.cc:96-156: per call, the radial forces are summed per bin in a local
// table, which is then added to the running one (as the reference adds each
// iteration's FORCE_TABLE_LOC to RADIAL_BH_FORCE_TABLE).
int steps_radial_bh_table_accumulate(int32_t topology, double Rsim, int32_t size, int64_t N, const double *x,
                                     const double *F, double *sum, int32_t *count) {
    if (topology != STEPS_TOPO_R3 && topology != STEPS_TOPO_S1R2)
        return fail(STEPS_ERR_ARG, "the radial octree correction exists for R3 and S1R2 only");
    if (size < 1 || N < 0 || !(Rsim > 0) || !sum || !count || (N > 0 && (!x || !F)))
        return fail(STEPS_ERR_ARG, "bad radial octree table arguments");
    std::vector<double> loc((size_t)size, 0.0);
    std::vector<int32_t> cnt((size_t)size, 0);
    const double w = Rsim / (double)size;
    for (int64_t i = 0; i < N; ++i) {
        const double *p = x + 3 * i, *f = F + 3 * i;
        double r, fr;
        if (topology == STEPS_TOPO_R3) {
            r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            fr = (f[0] * p[0] + f[1] * p[1] + f[2] * p[2]) / r;
        } else {
            r = std::sqrt(p[0] * p[0] + p[1] * p[1]);
            fr = (f[0] * p[0] + f[1] * p[1]) / r;
        }
        if (!std::isfinite(r)) return fail(STEPS_ERR_NONFINITE, "non-finite position in the radial octree table");
        const double bf = std::floor(r / w);
        const int b = bf >= (double)size ? size - 1 : (int)bf;
        loc[(size_t)b] += fr;
        ++cnt[(size_t)b];
    }
    for (int b = 0; b < size; ++b) {
        sum[b] += loc[(size_t)b];
        count[b] += cnt[(size_t)b];
    }
    return STEPS_OK;
}

// main.cc:1062-1075: the mean radial force per bin, 0 where a bin is empty
int steps_radial_bh_table_finish(int32_t size, double *sum, const int32_t *count) {
    if (size < 1 || !sum || !count) return fail(STEPS_ERR_ARG, "bad radial octree table arguments");
    for (int b = 0; b < size; ++b) sum[b] = count[b] > 0 ? sum[b] / (double)count[b] : 0.0;
    return STEPS_OK;
}

}  // extern "C"
