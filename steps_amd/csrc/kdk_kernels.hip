// kdk_kernels.hip -- device side of the kick-drift-kick step (step.cc:100-312).
//
// The reference runs the kicks, drift, periodic wrap and the errmax reduction
// on MPI rank 0's host (step.cc:128-181, :250-269); here they run where the
// particle state lives, so a step moves no particle data over PCIe.  errmax is
// a max-reduction (exact, so the block order does not matter).  The per-step
// scalars come from a device parameter block written before the launch, so a
// whole step can be captured once in a hipGraph and replayed.
#include "internal.h"

#include <algorithm>

namespace steps {

namespace {

constexpr int kRedBlock = 256;

__global__ __launch_bounds__(256) void kick_drift_kernel(double4 *__restrict__ jrec,
                                                         double *__restrict__ v,
                                                         const double *__restrict__ F, int64_t n,
                                                         const KdkParams *__restrict__ prm,
                                                         int topology, double L) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double G = prm->G, ainv3 = prm->ainv3_1, hubble = prm->hubble1, h = prm->h;
    double4 p = jrec[i];
    double x[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        // step.cc:133-140: A = G F a^-3 - 2 H v; v += A h/2; x += v h
        const double A = G * F[3 * i + k] * ainv3 - 2.0 * hubble * v[3 * i + k];
        const double vn = v[3 * i + k] + A * (h / 2.0);
        v[3 * i + k] = vn;
        x[k] = x[k] + vn * h;
    }
    if (topology == STEPS_TOPO_T3) {
        // step.cc:151-166
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (x[k] < 0) x[k] += L;
            else if (x[k] >= L) x[k] -= L;
        }
    } else if (topology == STEPS_TOPO_S1R2) {
        // step.cc:167-179
        if (x[2] < 0) x[2] += L;
        else if (x[2] >= L) x[2] -= L;
    }
    jrec[i] = make_double4(x[0], x[1], x[2], p.w);
}

__global__ __launch_bounds__(kRedBlock) void kick_errmax_kernel(
    const double4 *__restrict__ jrec, double *__restrict__ v, const double *__restrict__ F, int64_t n,
    const KdkParams *__restrict__ prm, int kick, double *__restrict__ block_max) {
    __shared__ double red[kRedBlock];
    const double G = prm->G, ainv3 = prm->ainv3_2, hubble = prm->hubble2, h = prm->h;
    const double const_beta = prm->const_beta;
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double A[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            A[k] = G * F[3 * i + k] * ainv3 - 2.0 * hubble * v[3 * i + k];
            if (kick) v[3 * i + k] += A[k] * (h / 2.0);  // step.cc:258-262
        }
        // step.cc:263: |A| / cbrt(M * 3/(4 pi rho_part))
        const double err = sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]) / cbrt(jrec[i].w * const_beta);
        m = fmax(m, err);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kRedBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) block_max[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kRedBlock) void max_finish_kernel(const double *__restrict__ block_max,
                                                               int nb, double *__restrict__ out) {
    __shared__ double red[kRedBlock];
    double m = 0.0;
    for (int i = threadIdx.x; i < nb; i += kRedBlock) m = fmax(m, block_max[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kRedBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

}  // namespace

int kdk_reduce_blocks(int64_t n) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + kRedBlock - 1) / kRedBlock));
}

hipError_t launch_kick_drift(DeviceParticles &dp, double *v, const double *F, const KdkParams *prm,
                             int topology, double L, hipStream_t st) {
    if (dp.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kick_drift_kernel, dim3((unsigned)((dp.n + 255) / 256)), dim3(256), 0, st, dp.jrec, v,
                       F, dp.n, prm, topology, L);
    return hipGetLastError();
}

hipError_t launch_kick_errmax(const DeviceParticles &dp, double *v, const double *F, const KdkParams *prm,
                              double *block_max, int nblocks, double *errmax_dev, hipStream_t st,
                              bool kick) {
    if (dp.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kick_errmax_kernel, dim3((unsigned)nblocks), dim3(kRedBlock), 0, st, dp.jrec, v, F, dp.n,
                       prm, kick ? 1 : 0, block_max);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(max_finish_kernel, dim3(1), dim3(kRedBlock), 0, st, block_max, nblocks, errmax_dev);
    return hipGetLastError();
}

}  // namespace steps
