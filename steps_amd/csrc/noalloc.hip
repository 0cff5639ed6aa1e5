// noalloc.hip -- steps_force_eval_noalloc: the one-shot force evaluation of
// steps_force_eval without any device allocation.
//
// No counterpart in the reference.  It is for callers that cannot allocate
// device memory (tight serving loops, harnesses that forbid hipMalloc): every
// device buffer the evaluation needs -- the staged inputs, the packed source
// records, the forces, the pair kernels' partial sums / tiles / ordering
// scratch and the device copies of the tables -- is carved out of one
// file-scope device array, `scratch`, of kNoAllocScratchBytes per device.
// The array's address on each device is looked up once (hipGetSymbolAddress,
// with that device current) and only that address is used from the host.
// Calls are serialised by a mutex; every call on a device reuses the same
// array, so calls from different processes, or through other libraries that
// reach the same array, must not overlap.  Nothing the kernels read needs to
// start at zero: every byte they read is written earlier in the same call.
#include "sim_internal.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

namespace steps {
namespace {

constexpr size_t kNoAllocScratchBytes = (size_t)STEPS_NOALLOC_SCRATCH_MB << 20;
constexpr int kNoAllocMaxDevices = 64;

// The device memory of this entry point: held by every device the library
// runs on until the process ends.  Host code never takes this name's address
// for device use (it is a host-side stand-in); it uses g_na.base[device].
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct NoAllocState {
    std::mutex mu;
    unsigned char *base[kNoAllocMaxDevices] = {};
};
NoAllocState g_na;

// The array's address on `device` (current), looked up on the first call there.
int scratch_base(int device, unsigned char **out) {
    if (device < 0 || device >= kNoAllocMaxDevices) return fail(STEPS_ERR_DEVICE, "device ordinal out of range");
    if (!g_na.base[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch address");
        g_na.base[device] = static_cast<unsigned char *>(ptr);
    }
    *out = g_na.base[device];
    return STEPS_OK;
}

// Consecutive 256-byte-aligned pieces of the scratch array.
struct Carver {
    unsigned char *base;
    size_t used = 0;
    template <class T> T *take(size_t bytes) {
        T *p = reinterpret_cast<T *>(base + used);
        used += (bytes + 255) & ~(size_t)255;
        return p;
    }
};

// The device layouts of engine.hip's t3_device_layout / s1r2_device_layout,
// written on the device from a raw copy of the caller's table.
__global__ void t3_layout_kernel(const double *__restrict__ T, int ng, int lo, double *__restrict__ out) {
    const size_t n = (size_t)ng, nx = (size_t)(ng - lo), nz = n + 3;
    const size_t total = nx * nx * nz * 3u;
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += (size_t)gridDim.x * blockDim.x) {
        const size_t c = o % 3u, z = (o / 3u) % nz, my = (o / (3u * nz)) % nx, mx = o / (3u * nz * nx);
        const size_t iz = (z + n - 1) % n;
        out[o] = T[((((size_t)lo + mx) * n + (size_t)lo + my) * n + iz) * 3u + c];
    }
}

__global__ void s1r2_layout_kernel(const double *__restrict__ T, int nrho, int nz, double *__restrict__ out) {
    const size_t nzl = (size_t)nz + 3, total = ((size_t)nrho + 2) * nzl * 2u;
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += (size_t)gridDim.x * blockDim.x) {
        const size_t c = o & 1u, z = (o / 2u) % nzl, r = o / (2u * nzl);
        const size_t ir = r == 0 ? 0 : ((r - 1 < (size_t)nrho - 1) ? r - 1 : (size_t)nrho - 1);
        const size_t iz = (z + (size_t)nz - 1) % (size_t)nz;
        out[o] = T[(ir * (size_t)nz + iz) * 2u + c];
    }
}

unsigned grid_for(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 4096); }

}  // namespace
}  // namespace steps

using namespace steps;

extern "C" int steps_force_eval_noalloc(const steps_force_desc *d, int64_t N, const double *x, const double *M,
                                        const double *soft, double *F, int64_t id_min, int64_t id_max) {
    if (!d) return fail(STEPS_ERR_ARG, "null descriptor");
    EvalPlan p;
    int rc = make_plan(*d, N, id_min, id_max, p);
    if (rc) return rc;
    if (p.nrows == 0) return STEPS_OK;
    if (!x || !M || !soft || !F) return fail(STEPS_ERR_ARG, "null buffer");
    if ((rc = check_finite(x, 3 * N))) return rc;
    std::lock_guard<std::mutex> lock(g_na.mu);
    if ((rc = set_device(d->device))) return rc;
    unsigned char *base = nullptr;
    if ((rc = scratch_base(d->device, &base))) return rc;

    // what the evaluation needs, in the order it is carved
    const size_t nb = (size_t)N * sizeof(double);
    const bool t3 = p.variant == PV_T3_EWALD, s1r2 = is_s1r2_lookup(p.variant), radial = p.ext_mode == 3;
    int lo = 0;
    size_t t3_raw = 0, t3_dev = 0, s1_raw = 0, s1_dev = 0, rad = 0;
    if (t3) {
        lo = t3_layout_lo(d->t3_table, d->t3_ngrid);
        t3_raw = (size_t)d->t3_ngrid * d->t3_ngrid * d->t3_ngrid * 3u * sizeof(double);
        t3_dev = t3_device_entries(d->t3_ngrid, lo) * sizeof(double);
    }
    if (s1r2) {
        s1_raw = (size_t)d->s1r2_nrho * d->s1r2_nz * 2u * sizeof(double);
        s1_dev = s1r2_device_nodes(d->s1r2_nrho, d->s1r2_nz) * 2u * sizeof(double);
    }
    if (radial) rad = (size_t)d->radial_table_size * sizeof(double);
    const size_t rbh = (size_t)p.bh_radial_size * sizeof(double);
    const size_t pieces[] = {3 * nb, nb, nb, (size_t)N * sizeof(double4), nb, 3 * (size_t)p.nrows * sizeof(double),
                             scratch_bytes(p), t3_raw, t3_dev, s1_raw, s1_dev, rad, rbh};
    size_t need = 0;
    for (size_t b : pieces) need += (b + 255) & ~(size_t)255;
    if (need > kNoAllocScratchBytes)
        return fail(STEPS_ERR_ARG, "evaluation needs " + std::to_string(need) + " bytes of device scratch; " +
                                       "steps_force_eval_noalloc has " + std::to_string(kNoAllocScratchBytes));
    Carver cv{base};
    double *dx = cv.take<double>(3 * nb), *dM = cv.take<double>(nb), *dsoft = cv.take<double>(nb);
    DeviceParticles dp{cv.take<double4>((size_t)N * sizeof(double4)), cv.take<double>(nb), N};
    double *dF = cv.take<double>(3 * (size_t)p.nrows * sizeof(double));
    double *part = cv.take<double>(scratch_bytes(p));
    double *t3r = cv.take<double>(t3_raw), *t3d = cv.take<double>(t3_dev);
    double *s1r = cv.take<double>(s1_raw), *s1d = cv.take<double>(s1_dev);
    double *radd = cv.take<double>(rad);
    double *rbhd = cv.take<double>(rbh);

    const hipStream_t st = nullptr;  // the null stream: no stream is created
    if (t3) {
        STEPS_HIP(hipMemcpyAsync(t3r, d->t3_table, t3_raw, hipMemcpyHostToDevice, st), "copy T3 table");
        hipLaunchKernelGGL(t3_layout_kernel, dim3(grid_for(t3_dev / 8)), dim3(256), 0, st, t3r, d->t3_ngrid, lo, t3d);
        STEPS_HIP(hipGetLastError(), "T3 table layout");
        p.t3 = t3d;
        p.t3_lo = lo;
    }
    if (s1r2) {
        STEPS_HIP(hipMemcpyAsync(s1r, d->s1r2_table, s1_raw, hipMemcpyHostToDevice, st), "copy S1R2 table");
        hipLaunchKernelGGL(s1r2_layout_kernel, dim3(grid_for(s1_dev / 8)), dim3(256), 0, st, s1r, d->s1r2_nrho,
                           d->s1r2_nz, s1d);
        STEPS_HIP(hipGetLastError(), "S1R2 table layout");
        p.s1r2 = s1d;
    }
    if (radial) {
        STEPS_HIP(hipMemcpyAsync(radd, d->radial_table, rad, hipMemcpyHostToDevice, st), "copy radial table");
        p.radial = radd;
    }
    if (rbh) {
        STEPS_HIP(hipMemcpyAsync(rbhd, d->radial_bh_table, rbh, hipMemcpyHostToDevice, st), "copy radial octree table");
        p.bh_radial = rbhd;
    }
    STEPS_HIP(hipMemcpyAsync(dx, x, 3 * nb, hipMemcpyHostToDevice, st), "copy x");
    STEPS_HIP(hipMemcpyAsync(dM, M, nb, hipMemcpyHostToDevice, st), "copy M");
    STEPS_HIP(hipMemcpyAsync(dsoft, soft, nb, hipMemcpyHostToDevice, st), "copy soft");
    STEPS_HIP(launch_pack(dx, dM, dsoft, N, dp, st), "pack");
    STEPS_HIP(launch_forces(p, dp, part, dF, st, nullptr, nullptr, nullptr), "force kernels");
    STEPS_HIP(hipMemcpyAsync(F, dF, 3 * (size_t)p.nrows * sizeof(double), hipMemcpyDeviceToHost, st), "copy F");
    STEPS_HIP(hipStreamSynchronize(st), "force evaluation");
    return STEPS_OK;
}
