// internal.h -- declarations shared by the steps_amd HIP translation units.
// Not part of the public ABI (that is include/steps_force.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "steps_force.h"

namespace steps {

constexpr double kPi = 3.14159265358979323846264338327950288419716939937510;  // global_variables.h:16
constexpr double kRadialExtent = 2.25;                                          // global_variables.h:39
constexpr int kBlock = 256;      // threads per workgroup of the pair kernels (4 wave64)
constexpr int kRowsPerLane = 1;  // i-particles held in registers per lane (FP64 path)
constexpr int kRowsPerLaneMixed = 2;

// Pair-term variants: one kernel instantiation each.
enum PairVariant : int {
    PV_R3 = 0,            // forces_cuda.cu:522-563
    PV_T3_NEAREST = 1,    // forces_cuda.cu:576-608   (IS_PERIODIC == 1)
    PV_T3_EWALD = 2,      // forces_cuda.cu:609-643   (IS_PERIODIC >= 2)
    PV_S1R2_NEAREST = 3,  // forces_cuda.cu:772-818   (IS_PERIODIC == 1)
    PV_S1R2_NGP = 4,      // forces_cuda.cu:819-863, EWALD_INTERPOLATION_ORDER 0
    PV_S1R2_CIC = 5,      //   ... order 2
    PV_S1R2_TSC = 6,      //   ... order 4
    PV_S1R2_IMAGES = 7,   // forces_cuda.cu:711-759   (PERIODIC_Z_NOLOOKUP, IS_PERIODIC >= 2)
};

constexpr bool is_s1r2_lookup(int v) { return v == PV_S1R2_NGP || v == PV_S1R2_CIC || v == PV_S1R2_TSC; }
// Variants with a mixed-precision kernel: R3, T3 (nearest image, tricubic
// Ewald correction) and S1R2 with the nearest z-image (plain, or with the CIC
// / TSC interpolated Ewald correction).  NGP is left out: its correction is
// piecewise constant, so FP32 rounding of (rho, dz) moves pairs near cell
// boundaries to the neighbouring node.  So is the S1R2 image sum.
constexpr bool mixed_supported(int v) {
    return v == PV_R3 || v == PV_T3_NEAREST || v == PV_T3_EWALD || v == PV_S1R2_NEAREST || v == PV_S1R2_CIC ||
           v == PV_S1R2_TSC;
}
// i-particles per lane of each kernel instantiation: table variants gather
// 8..192 table words per pair and keep one row per lane, and so does the
// mixed T3 nearest-image kernel (a wave of 64 rows straddles fewer periodic
// wraps than one of 128: 2145 against 1714 G pairs/s at 64^3).
constexpr int rows_per_lane(int v, int precision) {
    return (v == PV_T3_EWALD || is_s1r2_lookup(v) || (v == PV_T3_NEAREST && precision == STEPS_PREC_MIXED))
               ? 1
               : (precision == STEPS_PREC_MIXED ? kRowsPerLaneMixed : kRowsPerLane);
}

// Device-resident particle set in the layout the pair kernels read:
// jrec[i] = (x, y, z, m) for wave-uniform scalar loads of the j-particle,
// soft[i] = h.
struct DeviceParticles {
    double4 *jrec = nullptr;
    double *soft = nullptr;
    int64_t n = 0;
};

// Everything one force evaluation needs on the device.
struct EvalPlan {
    int variant = PV_R3;
    int precision = STEPS_PREC_FP64;
    int64_t N = 0;
    int64_t id_min = 0, nrows = 0;
    int splits = 1;        // j-range splits (partials reduced in fixed order)
    int64_t jchunk = 0;    // j per split
    // topology scalars
    double L = 0, Rsim = 0;
    int is_periodic = 0;
    int cosmology = 0, comoving = 0;
    double ext_coeff = 0;  // mass_in_unit_sphere (comoving) or DE (non-comoving)
    int ext_mode = 0;      // 0 none, 1 xyz, 2 xy, 3 xy * radial correction
    // tables (device)
    const double *t3 = nullptr; int t3_ng = 0;
    int t3_lo = 0;         // first x/y node of the T3 device layout (t3_device_layout)
    const double *s1r2 = nullptr; int s1r2_nrho = 0, s1r2_nz = 0;
    const double *radial = nullptr; int radial_size = 0;
    int images = 0;        // S1R2 image-sum: ewald_max
    // Barnes-Hut octree instead of the direct sum (USE_BH; forces.cc:373-506,
    // 582-772, 881-1217, CPU-only in the reference): the variant above then
    // names the pair term of each accepted node, and bh_order the T3 table
    // interpolation of the CPU build (Lagrange order 2, 6, 8; 4 Catmull-Rom)
    bool bh = false;
    double bh_theta = 0, bh_particle_radius = 0;  // opening angle, ParticleRadi
    int bh_order = 2;
    // the randomised octree (RANDOMIZE_BH) and its draws u0..u5
    bool bh_randomize = false;
    double bh_u[6] = {0, 0, 0, 0, 0, 0};
    // the radial octree correction table (device), applied with the
    // comoving external term of R3 and S1R2
    const double *bh_radial = nullptr; int bh_radial_size = 0;
};

// The T3 Ewald evaluation deals its sources out to the eight XCDs by table
// slab (pair_kernels.hip, body_t3_slab): make_plan then sets eight splits.
bool t3_slab_plan(const EvalPlan &p);

// Scratch of one evaluation: partial sums [splits][3][nrows] (FP64), followed
// (256-byte aligned) by the source-tile table and, for STEPS_PREC_MIXED, the
// FP32 tile-recentred source records and an FP32 copy of the S1R2 table.
size_t partial_bytes(const EvalPlan &p);
size_t tile_extra_bytes(const EvalPlan &p);
constexpr int kTile = 64;  // sources per tile (one wave64 prepares one tile)
constexpr int kSub = 8;    // sources per sub-tile of the mixed S1R2 table kernels

// Launch the tile pass, the pair kernel and the fixed-order reduction; F is
// device AoS [3*nrows].  Non-null events are recorded right before and after
// the pair kernel and after the reduction (on `stream`).
hipError_t launch_forces(const EvalPlan &p, const DeviceParticles &dp, double *partials,
                         double *F, hipStream_t stream, hipEvent_t ev_begin,
                         hipEvent_t ev_pair_end, hipEvent_t ev_end);
// The same in three phases: source preparation (tiles; for the mixed path the
// FP32 table copy, spatial ordering and recentred records), the pair kernel,
// and the fixed-order reduction with the external term.
hipError_t launch_forces_prepare(const EvalPlan &p, const DeviceParticles &dp, double *partials,
                                 hipStream_t stream);
hipError_t launch_forces_pairs(const EvalPlan &p, const DeviceParticles &dp, double *partials,
                               hipStream_t stream);
hipError_t launch_forces_finish(const EvalPlan &p, const DeviceParticles &dp, double *partials, double *F,
                                hipStream_t stream);

// Packed launches: the pair kernels of up to kMaxPack evaluations of one
// (variant, precision) in one grid.  pack_describe writes the descriptor of
// evaluation idx (host memory, pack_desc_bytes() each) whose workgroups start
// at blk_begin and returns its workgroup count; launch_pair_packed runs the
// descriptors after they have been copied to device memory.  Each evaluation
// runs exactly the kernel body, plan and reduction order it runs alone.
constexpr int kMaxPack = 16;
size_t pack_desc_bytes();
int64_t pack_describe(const EvalPlan &p, const DeviceParticles &dp, double *partials, void *host_descs, int idx,
                      int64_t blk_begin);
hipError_t launch_pair_packed(int variant, int precision, const void *dev_descs, int nseg, int64_t total_blocks,
                              hipStream_t stream);

// Spatial ordering of n particles into sorted copies jrec_s / soft_s: Morton
// order over the bounding box, or, with pencil_ng > 0, z pencils over the T3
// table's cell columns (box side pencil_L, pencil_ng cells per axis);
// `scratch` holds order_scratch_bytes(n) bytes (order_kernels.hip).  After the
// launch, spatial_order_perm(scratch, n)[k] is the input index of sorted
// particle k.
size_t order_scratch_bytes(int64_t n);
hipError_t launch_spatial_order(const double4 *jrec, const double *soft, int64_t n, void *scratch,
                                double4 *jrec_s, double *soft_s, hipStream_t stream, double pencil_L = 0.0,
                                int pencil_ng = 0);
const uint32_t *spatial_order_perm(const void *scratch, int64_t n);
// ... and spatial_order_keys(scratch, n)[k] its sort key.
const uint32_t *spatial_order_keys(const void *scratch, int64_t n);

// Stable sort of n 64-bit keys given as high and low 32-bit halves (device):
// *perm receives (inside scratch, sort_u64_scratch_bytes(n) bytes) the input
// index of the k-th smallest key.
size_t sort_u64_scratch_bytes(int64_t n);
hipError_t launch_sort_u64(const uint32_t *hi, const uint32_t *lo, int64_t n, void *scratch, const uint32_t **perm,
                           hipStream_t stream);
// The same for keys of nwords 32-bit words, words[0] the most significant
// (the same scratch size: one word is sorted at a time).
hipError_t launch_sort_words(const uint32_t *const *words, int nwords, int64_t n, void *scratch, const uint32_t **perm,
                             hipStream_t stream);

// Barnes-Hut octree forces (bh_kernels.hip): tree build and traversal for
// rows id_min .. id_min+nrows-1 into F (device AoS, external term included);
// `scratch` holds bh_scratch_bytes(N) bytes.  A non-null ev_tree_end is
// recorded between the tree build and the traversal.
int64_t bh_item_capacity(int64_t n);
size_t bh_scratch_bytes(int64_t n);
hipError_t launch_bh_forces(const EvalPlan &p, const DeviceParticles &dp, void *scratch, double *F,
                            hipStream_t stream, hipEvent_t ev_tree_end);

// AoS x [3n] + M [n] + soft [n] (device) -> DeviceParticles layout.
hipError_t launch_pack(const double *x, const double *M, const double *soft, int64_t n,
                       DeviceParticles &dp, hipStream_t stream);
// Replace positions only (AoS x [3n], device) keeping masses.
hipError_t launch_pack_positions(const double *x, int64_t n, DeviceParticles &dp,
                                 hipStream_t stream);
// DeviceParticles positions -> AoS x [3n] (device).
hipError_t launch_unpack_positions(const DeviceParticles &dp, double *x, hipStream_t stream);

// KDK pieces (kdk_kernels.hip).  Per-step scalars live in device memory
// (written by the host before the launch) so that captured graphs replay.
struct KdkParams {
    double G;
    double h;
    double ainv3_1, hubble1;  // before the step (first kick)
    double ainv3_2, hubble2;  // after the Friedmann update (second kick, errmax)
    double const_beta;        // 3/(4 pi rho_part)
    double pad;
};
hipError_t launch_kick_drift(DeviceParticles &dp, double *v, const double *F, const KdkParams *prm,
                             int topology, double L, hipStream_t stream);
hipError_t launch_kick_errmax(const DeviceParticles &dp, double *v, const double *F, const KdkParams *prm,
                              double *block_max, int nblocks_max, double *errmax_dev,
                              hipStream_t stream, bool kick);
int kdk_reduce_blocks(int64_t n);

// T3 table in the layout the pair kernel reads.  The reference table is
// [(ix*Ng + iy)*Ng + iz]*3 + c.  The device layout keeps x and y nodes
// lo..Ng-1 and every z node plus one periodic ghost below and two above:
// element ((mx*(Ng-lo) + my)*(Ng+3) + iz+1)*3 + c holds node (lo+mx, lo+my, iz).
// lo = Ng/2 when the table is exactly antisymmetric in x and in y (the
// kernel reads a node below lo at its mirror Ng-1-i with that component
// negated), else 0.  Returns lo.
int t3_device_layout(const double *T, int ng, std::vector<double> &out);
// The lo t3_device_layout picks for this table (Ng/2 when it folds, else 0).
int t3_layout_lo(const double *T, int ng);
inline size_t t3_device_entries(int ng, int lo) {
    return (size_t)(ng - lo) * (size_t)(ng - lo) * (size_t)(ng + 3) * 3u;
}
constexpr int kT3MaxGrid = 1023;  // element offsets of the layout stay below 2^32
// S1R2 table in the layout the pair kernels read.  The reference table is
// [(ir*Nz + iz)*2 + c].  The device layout keeps rho rows with z contiguous
// but adds ghosts: rows for rho nodes -1 .. Nrho (repeating the edge nodes,
// the reference's clamp), each with z nodes -1 .. Nz + 1 (periodic), node
// (ir, iz) at ((ir + 1)*(Nz + 3) + iz + 1)*2 + c, so no stencil needs a wrap
// or a clamp.
void s1r2_device_layout(const double *T, int nrho, int nz, std::vector<double> &out);
inline size_t s1r2_device_nodes(int nrho, int nz) { return ((size_t)nrho + 2) * ((size_t)nz + 3); }

// Table builders (table_kernels.hip); all buffers host memory.
int build_t3_table(int is_periodic, double L, double *table, int *ngrid);
int build_s1r2_table(int is_periodic, double L, double Rsim, double *table, int *nrho, int *nz);
int build_radial_table(double R, double Lz, int size, int accuracy, double *table);

// error text (engine.hip)
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

}  // namespace steps
