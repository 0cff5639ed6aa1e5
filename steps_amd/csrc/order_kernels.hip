// order_kernels.hip -- spatial ordering of the sources (and, for T3, the rows).
//
// The FP32 pair arithmetic works on offsets from the centre of each 64-source
// tile; its rounding error grows with the distance of both particles from that
// centre, so the tiles have to be spatially compact: the mixed paths (and the
// FP64 R3 far-tile test) put the sources in Morton order of their positions
// (30-bit keys over the bounding box).  The T3 Ewald pair kernel is bound by
// its table gathers, whose cost depends on how many cache lines the 64 lanes
// of one load touch: rows and sources go into z pencils, ordered by the table
// cell column (x, y) and then by z, so that neighbouring lanes read
// neighbouring nodes of the same contiguous z runs.  Both keys are sorted with
// a stable LSD radix sort and the records gathered into sorted copies.  Every
// step is deterministic (exact min/max, integer counts, stable placement) and
// no step uses memset or host synchronisation, so the sequence can be captured
// in a graph.  There is no counterpart in the reference (its FP32 mode is
// all-FP32 on unsorted input, global_variables.h:26-32, and its kernels read
// particles in input order).
#include "internal.h"

#include <algorithm>

namespace steps {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int64_t kChunk = (int64_t)kSortThreads * kSortItems;  // elements per block
constexpr int kBboxBlocks = 1024;

size_t a256(size_t b) { return (b + 255) & ~(size_t)255; }

int sort_blocks(int64_t n) { return (int)std::max<int64_t>(1, (n + kChunk - 1) / kChunk); }

__global__ __launch_bounds__(256) void bbox_partial_kernel(const double4 *__restrict__ jrec, int64_t n,
                                                           double *__restrict__ part) {
    __shared__ double red[6][256];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double4 p = jrec[i];
        lo[0] = fmin(lo[0], p.x);
        lo[1] = fmin(lo[1], p.y);
        lo[2] = fmin(lo[2], p.z);
        hi[0] = fmax(hi[0], p.x);
        hi[1] = fmax(hi[1], p.y);
        hi[2] = fmax(hi[2], p.z);
    }
    const int t = threadIdx.x;
    for (int a = 0; a < 3; ++a) {
        red[a][t] = lo[a];
        red[3 + a][t] = hi[a];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int a = 0; a < 3; ++a) {
                red[a][t] = fmin(red[a][t], red[a][t + s]);
                red[3 + a][t] = fmax(red[3 + a][t], red[3 + a][t + s]);
            }
        __syncthreads();
    }
    if (t < 6) part[blockIdx.x * 6 + t] = red[t][0];
}

// Fold the per-block boxes (at most kBboxBlocks of them) into one.
__global__ __launch_bounds__(256) void bbox_final_kernel(double *__restrict__ part, int nparts,
                                                         double *__restrict__ box) {
    __shared__ double red[6][256];
    const int t = threadIdx.x;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = t; k < nparts; k += 256)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fmin(lo[a], part[k * 6 + a]);
            hi[a] = fmax(hi[a], part[k * 6 + 3 + a]);
        }
    for (int a = 0; a < 3; ++a) {
        red[a][t] = lo[a];
        red[3 + a][t] = hi[a];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int a = 0; a < 3; ++a) {
                red[a][t] = fmin(red[a][t], red[a][t + s]);
                red[3 + a][t] = fmax(red[3 + a][t], red[3 + a][t + s]);
            }
        __syncthreads();
    }
    if (t < 6) box[t] = red[t][0];
}

// Morton key of each source (30 bits) and the identity permutation.
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ __launch_bounds__(256) void morton_kernel(const double4 *__restrict__ jrec, int64_t n,
                                                     const double *__restrict__ box, uint32_t *__restrict__ key,
                                                     uint32_t *__restrict__ val) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 p = jrec[i];
    const double c[3] = {p.x, p.y, p.z};
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double ext = box[3 + a] - box[a];
        const double u = (ext > 0.0) ? (c[a] - box[a]) * (1024.0 / ext) : 0.0;
        q[a] = (uint32_t)min(1023, max(0, (int)u));
    }
    key[i] = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
    val[i] = (uint32_t)i;
}

// z-pencil key of each particle (30 bits): the T3 table cell column (x, y),
// 10 bits each (Ng <= 1023; y reversed in odd x columns), then z on 1024
// levels.  Coordinates are taken
// modulo L, so a box on [0, L) and one on [-L/2, L/2) give the same pencils
// up to a rotation of their order.
__device__ __forceinline__ uint32_t periodic_bin(double c, double inv_L, int nb) {
    double u = c * inv_L;
    u -= floor(u);
    return (uint32_t)min(nb - 1, max(0, (int)(u * (double)nb)));
}

__global__ __launch_bounds__(256) void pencil_kernel(const double4 *__restrict__ jrec, int64_t n, double L,
                                                     int ng, uint32_t *__restrict__ key, uint32_t *__restrict__ val) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 p = jrec[i];
    const double inv_L = 1.0 / L;
    const uint32_t bx = periodic_bin(p.x, inv_L, ng), by = periodic_bin(p.y, inv_L, ng);
    // serpentine: y runs backwards in odd x columns, so consecutive pencils
    // stay neighbours across a column change
    key[i] = (bx << 20) | (((bx & 1u) ? (uint32_t)(ng - 1) - by : by) << 10) | periodic_bin(p.z, inv_L, 1024);
    val[i] = (uint32_t)i;
}

__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const uint32_t *__restrict__ key, int64_t n,
                                                                  int shift, int nblk,
                                                                  uint32_t *__restrict__ counts) {
    __shared__ uint32_t h[256];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kChunk;
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t i = base + (int64_t)r * kSortThreads + t;
        if (i < n) atomicAdd(&h[(key[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(size_t)t * nblk + blockIdx.x] = h[t];  // digit-major: stable order of blocks per digit
}

// In-place exclusive scan of `total` counts by one workgroup.
__global__ __launch_bounds__(1024) void exclusive_scan_kernel(uint32_t *__restrict__ a, int64_t total) {
    __shared__ uint32_t s[1024];
    const int t = threadIdx.x;
    const int64_t per = (total + 1023) / 1024;
    const int64_t b = (int64_t)t * per, e = min(total, b + per);
    uint32_t sum = 0;
    for (int64_t i = b; i < e; ++i) sum += a[i];
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t v = (t >= off) ? s[t - off] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    uint32_t run = s[t] - sum;
    for (int64_t i = b; i < e; ++i) {
        const uint32_t v = a[i];
        a[i] = run;
        run += v;
    }
}

// The same scan over many workgroups, for large count arrays: each
// workgroup scans kScanChunk consecutive counts in place and writes its total
// (scan_chunk_kernel), one workgroup scans the totals (exclusive_scan_kernel),
// and every count gets its chunk's offset added (scan_add_kernel).  Integer
// sums: the result is the single-workgroup scan's exactly.
constexpr int kScanChunk = 1024;
constexpr int64_t kScanOneGroupMax = 16384;  // up to this many counts: one workgroup, one launch

__global__ __launch_bounds__(256) void scan_chunk_kernel(uint32_t *__restrict__ a, int64_t total,
                                                         uint32_t *__restrict__ bsum) {
    __shared__ uint32_t s[256];
    constexpr int kPer = kScanChunk / 256;
    const int t = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * kScanChunk + (int64_t)t * kPer;
    uint32_t c[kPer], sum = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        c[q] = (i0 + q < total) ? a[i0 + q] : 0u;
        sum += c[q];
    }
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t v = (t >= off) ? s[t - off] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    uint32_t run = s[t] - sum;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        if (i0 + q < total) a[i0 + q] = run;
        run += c[q];
    }
    if (t == 255) bsum[blockIdx.x] = s[255];
}

__global__ __launch_bounds__(256) void scan_add_kernel(uint32_t *__restrict__ a, int64_t total,
                                                       const uint32_t *__restrict__ bsum) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) a[i] += bsum[i / kScanChunk];
}

int64_t scan_chunks(int64_t total) { return (total + kScanChunk - 1) / kScanChunk; }

// Exclusive scan of total counts in place; bsum holds scan_chunks(total) words.
hipError_t launch_counts_scan(uint32_t *counts, int64_t total, uint32_t *bsum, hipStream_t st) {
    if (total <= kScanOneGroupMax) {
        hipLaunchKernelGGL(exclusive_scan_kernel, dim3(1), dim3(1024), 0, st, counts, total);
        return hipGetLastError();
    }
    const int64_t nc = scan_chunks(total);
    hipLaunchKernelGGL(scan_chunk_kernel, dim3((unsigned)nc), dim3(256), 0, st, counts, total, bsum);
    hipLaunchKernelGGL(exclusive_scan_kernel, dim3(1), dim3(1024), 0, st, bsum, nc);
    hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, counts, total, bsum);
    return hipGetLastError();
}

// Stable scatter by one 8-bit digit: within a block, elements are placed in
// index order (rounds of 256, waves in order, lanes in order); the rank of a
// lane among the lanes of its wave holding the same digit comes from eight
// ballots.
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin, int64_t n, int shift,
    const uint32_t *__restrict__ offs, int nblk, uint32_t *__restrict__ kout, uint32_t *__restrict__ vout) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kSortThreads / 64][256];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    run[t] = offs[(size_t)t * nblk + blockIdx.x];
    const uint64_t below = (l == 0) ? 0ull : (~0ull >> (64 - l));
    const int64_t base = (int64_t)blockIdx.x * kChunk;
    for (int r = 0; r < kSortItems; ++r) {
#pragma unroll
        for (int ww = 0; ww < kSortThreads / 64; ++ww) wcnt[ww][t] = 0u;
        __syncthreads();
        const int64_t i = base + (int64_t)r * kSortThreads + t;
        const bool valid = i < n;
        uint32_t k = 0, v = 0, d = 0;
        if (valid) {
            k = kin[i];
            v = vin[i];
            d = (k >> shift) & 255u;
        }
        uint64_t eq = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(valid && bit);
            eq &= bit ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(eq & below);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(eq);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (int ww = 0; ww < w; ++ww) pos += wcnt[ww][d];
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int ww = 0; ww < kSortThreads / 64; ++ww) add += wcnt[ww][t];
        run[t] += add;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void gather_kernel(const double4 *__restrict__ jrec, const double *__restrict__ soft,
                                                     const uint32_t *__restrict__ perm, int64_t n,
                                                     double4 *__restrict__ jrec_s, double *__restrict__ soft_s) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = perm[i];
    jrec_s[i] = jrec[s];
    soft_s[i] = soft[s];
}

__global__ __launch_bounds__(256) void iota_copy_kernel(const uint32_t *__restrict__ kin, int64_t n,
                                                        uint32_t *__restrict__ kout, uint32_t *__restrict__ v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    kout[i] = kin[i];
    v[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void gather_u32_kernel(const uint32_t *__restrict__ src,
                                                         const uint32_t *__restrict__ perm, int64_t n,
                                                         uint32_t *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = src[perm[i]];
}

// One stable LSD pass over 8 bits of key[cur] (moving val[cur] along).
hipError_t radix_pass(uint32_t *key[2], uint32_t *val[2], int &cur, int64_t n, int shift, uint32_t *counts,
                      uint32_t *bsum, hipStream_t st) {
    const int nblk = sort_blocks(n);
    hipLaunchKernelGGL(radix_hist_kernel, dim3(nblk), dim3(kSortThreads), 0, st, key[cur], n, shift, nblk, counts);
    hipError_t e;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = launch_counts_scan(counts, (int64_t)256 * nblk, bsum, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(radix_scatter_kernel, dim3(nblk), dim3(kSortThreads), 0, st, key[cur], val[cur], n, shift,
                       counts, nblk, key[1 - cur], val[1 - cur]);
    cur = 1 - cur;
    return hipGetLastError();
}

}  // namespace

size_t sort_u64_scratch_bytes(int64_t n) {
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    return 4 * a256(nn * 4) + a256((size_t)256 * sort_blocks(n) * 4) +
           a256((size_t)scan_chunks((int64_t)256 * sort_blocks(n)) * 4) + 256;
}

hipError_t launch_sort_words(const uint32_t *const *words, int nwords, int64_t n, void *scratch, const uint32_t **perm,
                             hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (nwords < 1) return hipErrorInvalidValue;
    const size_t nn = (size_t)n;
    char *p = (char *)scratch;
    uint32_t *key[2], *val[2];
    key[0] = (uint32_t *)p;
    key[1] = (uint32_t *)(p + a256(nn * 4));
    val[0] = (uint32_t *)(p + 2 * a256(nn * 4));
    val[1] = (uint32_t *)(p + 3 * a256(nn * 4));
    uint32_t *counts = (uint32_t *)(p + 4 * a256(nn * 4));
    uint32_t *bsum = (uint32_t *)(p + 4 * a256(nn * 4) + a256((size_t)256 * sort_blocks(n) * 4));
    const unsigned ng = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(iota_copy_kernel, dim3(ng), dim3(256), 0, st, words[nwords - 1], n, key[0], val[0]);
    hipError_t e;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    int cur = 0;
    for (int shift = 0; shift < 32; shift += 8)
        if ((e = radix_pass(key, val, cur, n, shift, counts, bsum, st)) != hipSuccess) return e;
    // each more significant word, in the order the passes so far left: a
    // stable sort by it completes the sort by the words up to it
    for (int w = nwords - 2; w >= 0; --w) {
        hipLaunchKernelGGL(gather_u32_kernel, dim3(ng), dim3(256), 0, st, words[w], val[cur], n, key[cur]);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        for (int shift = 0; shift < 32; shift += 8)
            if ((e = radix_pass(key, val, cur, n, shift, counts, bsum, st)) != hipSuccess) return e;
    }
    *perm = val[cur];
    return hipSuccess;
}

hipError_t launch_sort_u64(const uint32_t *hi, const uint32_t *lo, int64_t n, void *scratch, const uint32_t **perm,
                           hipStream_t st) {
    const uint32_t *words[2] = {hi, lo};
    return launch_sort_words(words, 2, n, scratch, perm, st);
}

size_t order_scratch_bytes(int64_t n) {
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    return 4 * a256(nn * 4) + a256((size_t)256 * sort_blocks(n) * 4) + a256((size_t)kBboxBlocks * 6 * 8) +
           a256(6 * 8) + a256((size_t)scan_chunks((int64_t)256 * sort_blocks(n)) * 4) + 256;
}

const uint32_t *spatial_order_keys(const void *scratch, int64_t n) {
    (void)n;  // the sorted keys end in key[0], at the start of the scratch
    return (const uint32_t *)scratch;
}

const uint32_t *spatial_order_perm(const void *scratch, int64_t n) {
    // the sort's four passes leave the permutation in val[0], after the two key arrays
    return (const uint32_t *)((const char *)scratch + 2 * a256((size_t)std::max<int64_t>(n, 1) * 4));
}

hipError_t launch_spatial_order(const double4 *jrec, const double *soft, int64_t n, void *scratch,
                                double4 *jrec_s, double *soft_s, hipStream_t st, double pencil_L, int pencil_ng) {
    if (n <= 0) return hipSuccess;
    const size_t nn = (size_t)n;
    char *p = (char *)scratch;
    uint32_t *key[2], *val[2];
    key[0] = (uint32_t *)p;
    p += a256(nn * 4);
    key[1] = (uint32_t *)p;
    p += a256(nn * 4);
    val[0] = (uint32_t *)p;
    p += a256(nn * 4);
    val[1] = (uint32_t *)p;
    p += a256(nn * 4);
    const int nblk = sort_blocks(n);
    uint32_t *counts = (uint32_t *)p;
    p += a256((size_t)256 * nblk * 4);
    double *part = (double *)p;
    p += a256((size_t)kBboxBlocks * 6 * 8);
    double *box = (double *)p;
    p += a256(6 * 8);
    uint32_t *bsum = (uint32_t *)p;

    hipError_t e;
    const unsigned ng = (unsigned)((n + 255) / 256);
    if (pencil_ng > 0) {
        hipLaunchKernelGGL(pencil_kernel, dim3(ng), dim3(256), 0, st, jrec, n, pencil_L, pencil_ng, key[0], val[0]);
    } else {
        const int nbb = (int)std::min<int64_t>(kBboxBlocks, (n + 255) / 256);
        hipLaunchKernelGGL(bbox_partial_kernel, dim3(nbb), dim3(256), 0, st, jrec, n, part);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(bbox_final_kernel, dim3(1), dim3(256), 0, st, part, nbb, box);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(morton_kernel, dim3(ng), dim3(256), 0, st, jrec, n, box, key[0], val[0]);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    int cur = 0;
    for (int shift = 0; shift < 32; shift += 8) {
        hipLaunchKernelGGL(radix_hist_kernel, dim3(nblk), dim3(kSortThreads), 0, st, key[cur], n, shift, nblk,
                           counts);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = launch_counts_scan(counts, (int64_t)256 * nblk, bsum, st)) != hipSuccess) return e;
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(nblk), dim3(kSortThreads), 0, st, key[cur], val[cur], n,
                           shift, counts, nblk, key[1 - cur], val[1 - cur]);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        cur = 1 - cur;
    }
    hipLaunchKernelGGL(gather_kernel, dim3(ng), dim3(256), 0, st, jrec, soft, val[cur], n, jrec_s, soft_s);
    return hipGetLastError();
}

}  // namespace steps
