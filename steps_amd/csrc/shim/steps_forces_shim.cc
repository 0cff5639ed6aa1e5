// steps_forces_shim.cc -- drop-in replacement for StePS's force translation unit.
//
// The reference links exactly one of forces_cuda.o / forces.o
// (StePS/Template-LinuxGCC-Makefile:124,145); that unit defines, with C++
// linkage and the variant chosen by -DPERIODIC / -DPERIODIC_Z:
//   void forces(REAL*, REAL*, int, int)             forces_cuda.cu:73-79
//   void forces_periodic(REAL*, REAL*, int, int)    forces_cuda.cu:169-173
//   void forces_periodic_z(REAL*, REAL*, int, int)  forces_cuda.cu:457-461
//   void recalculate_softening()                    forces_cuda.cu:868-875
// This file defines the same symbols, reads the same globals (declared here
// with the types of StePS/src/global_variables.h and defined by StePS's
// main.cc), and forwards to the C ABI of include/steps_force.h.  Compile it
// with the same -DPERIODIC / -DPERIODIC_Z / -DPERIODIC_Z_NOLOOKUP /
// -DEWALD_INTERPOLATION_ORDER flags as the rest of StePS and link
// libsteps_amd.so instead of forces_cuda.o.
//
// Behaviour kept from the reference: rows ID_min..ID_max (inclusive) are
// written to F at offset 3*(i-ID_min) (forces_cuda.cu:1078-1097, overwrite);
// a failure prints to stderr and sets ForceError, which main.cc:1851 checks
// once per step.  Not kept: per-call device allocation and copies of the
// tables (buffers and tables stay resident), and the n_GPU row split (one
// device per process: STEPS_DEVICE, or rank modulo the device count).
// STEPS_FORCE_PRECISION=mixed selects the mixed-precision R^3 path.  Built
// with -DUSE_BH (the reference's Barnes-Hut build, which links forces.o) the
// same symbols evaluate the octree with the opening angle THETA and the
// node softening of ParticleRadi, globals main.cc:61-64 defines; with
// -DRANDOMIZE_BH=SEED too, each call draws its frame with rand() in the
// reference's order (main.cc:310 seeds it); once USE_RADIAL_BH_CORRECTION
// is set, R3 and S1R2 calls subtract RADIAL_BH_FORCE_TABLE's correction.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "steps_force.h"

#ifdef USE_SINGLE_PRECISION
#error "steps_forces_shim: the engine takes FP64 state (REAL=double); use STEPS_FORCE_PRECISION=mixed for the fast path"
#endif
typedef double REAL;

// ---- globals read by the reference force unit (global_variables.h) --------
extern int N;
extern REAL *M;
extern REAL *SOFT_LENGTH;
extern int rank;
extern int IS_PERIODIC;
extern int COSMOLOGY;
extern int COMOVING_INTEGRATION;
extern REAL L;
extern REAL Rsim;
extern REAL mass_in_unit_sphere;
extern double H0;
extern double Omega_lambda;
extern bool ForceError;
extern REAL beta;
extern REAL ParticleRadi;
extern REAL rho_part;
extern REAL M_min;
#ifdef USE_BH
extern REAL THETA;  // main.cc:64, set from USE_BH at main.cc:303
extern int RADIAL_BH_FORCE_TABLE_SIZE;  // main.cc:126-133
extern REAL *RADIAL_BH_FORCE_TABLE;
extern bool USE_RADIAL_BH_CORRECTION;
#endif
#if defined(PERIODIC)
extern REAL *T3_EWALD_FORCE_TABLE;
extern int N_EWALD_FORCE_GRID;
#elif defined(PERIODIC_Z)
#if !defined(PERIODIC_Z_NOLOOKUP)
extern REAL *S1R2_EWALD_FORCE_TABLE;
extern int Nz_EWALD_FORCE_GRID;
extern int Nrho_EWALD_FORCE_GRID;
#endif
extern int RADIAL_FORCE_TABLE_SIZE;
extern REAL *RADIAL_FORCE_TABLE;
#endif

#ifndef EWALD_INTERPOLATION_ORDER
#define EWALD_INTERPOLATION_ORDER 2  // Template-LinuxGCC-Makefile:35
#endif

namespace {

int pick_device() {
    const char *e = std::getenv("STEPS_DEVICE");
    if (e && *e) return std::atoi(e);
    const int n = steps_device_count();
    return n > 0 ? rank % n : 0;
}

int pick_precision() {
    const char *e = std::getenv("STEPS_FORCE_PRECISION");
    return (e && std::strcmp(e, "mixed") == 0) ? STEPS_PREC_MIXED : STEPS_PREC_FP64;
}

steps_force_desc descriptor(int topology) {
    steps_force_desc d;
    steps_force_desc_init(&d);
    d.topology = topology;
    d.is_periodic = IS_PERIODIC;
    d.cosmology = COSMOLOGY;
    d.comoving = COMOVING_INTEGRATION;
    d.L = L;
    d.Rsim = Rsim;
    d.mass_in_unit_sphere = mass_in_unit_sphere;
    d.H0 = H0;
    d.Omega_lambda = Omega_lambda;
    d.interp_order = EWALD_INTERPOLATION_ORDER;
    d.precision = pick_precision();
    d.device = pick_device();
#ifdef USE_BH
    d.bh_theta = THETA;
    d.bh_particle_radius = ParticleRadi;
#if !defined(PERIODIC)
    if (USE_RADIAL_BH_CORRECTION) {  // forces.cc:479-485, 1143-1147, 1180-1184
        d.radial_bh_table = RADIAL_BH_FORCE_TABLE;
        d.radial_bh_table_size = RADIAL_BH_FORCE_TABLE_SIZE;
    }
#endif
#ifdef RANDOMIZE_BH
    // the draws of this call, in the order the reference's force routine
    // makes them: R3 centre x, y, z, angle, axis theta, axis z (forces.cc:
    // 422-430, 362-363); T3 shift x, y, z (:664-667); S1R2 angle, centre x,
    // y, z shift (:1027-1030)
    const int draws = topology == STEPS_TOPO_R3 ? 6 : topology == STEPS_TOPO_T3 ? 3 : 4;
    d.bh_randomize = 1;
    for (int k = 0; k < draws; ++k) d.bh_random[k] = (double)rand() / (double)RAND_MAX;
#endif
#endif
#if defined(PERIODIC)
    d.t3_table = T3_EWALD_FORCE_TABLE;
    d.t3_ngrid = N_EWALD_FORCE_GRID;
#elif defined(PERIODIC_Z)
#if defined(PERIODIC_Z_NOLOOKUP)
    d.s1r2_lookup = 0;
#else
    d.s1r2_lookup = 1;
    d.s1r2_table = S1R2_EWALD_FORCE_TABLE;
    d.s1r2_nrho = Nrho_EWALD_FORCE_GRID;
    d.s1r2_nz = Nz_EWALD_FORCE_GRID;
#endif
    d.radial_table = RADIAL_FORCE_TABLE;
    d.radial_table_size = RADIAL_FORCE_TABLE_SIZE;
#endif
    return d;
}

#ifdef RANDOMIZE_BH
// The HIP runtime draws from rand() as it starts, inside the library's first
// call: library calls run on a scratch random() state, so that the process's
// sequence (srand(RANDOMIZE_BH), main.cc:310) advances only by the draws the
// reference's force routine makes.
class ScratchRandomState {
  public:
    ScratchRandomState() { prev_ = initstate(1u, buf_, sizeof buf_); }
    ~ScratchRandomState() { setstate(prev_); }
    ScratchRandomState(const ScratchRandomState &) = delete;
    ScratchRandomState &operator=(const ScratchRandomState &) = delete;

  private:
    char buf_[128];
    char *prev_;
};
#endif

void evaluate(int topology, REAL *x, REAL *F, int ID_min, int ID_max) {
    const steps_force_desc d = descriptor(topology);
#ifdef RANDOMIZE_BH
    ScratchRandomState guard;
#endif
    const int rc = steps_force_eval(&d, (int64_t)N, x, M, SOFT_LENGTH, F, ID_min, ID_max);
    if (rc != STEPS_OK) {
        std::fprintf(stderr, "MPI rank %i: steps force evaluation failed (%d): %s\n", rank, rc,
                     steps_last_error());
        ForceError = true;
    }
}

}  // namespace

#if defined(PERIODIC)
void forces_periodic(REAL *x, REAL *F, int ID_min, int ID_max) { evaluate(STEPS_TOPO_T3, x, F, ID_min, ID_max); }
#elif defined(PERIODIC_Z)
void forces_periodic_z(REAL *x, REAL *F, int ID_min, int ID_max) {
    evaluate(STEPS_TOPO_S1R2, x, F, ID_min, ID_max);
}
#else
void forces(REAL *x, REAL *F, int ID_min, int ID_max) { evaluate(STEPS_TOPO_R3, x, F, ID_min, ID_max); }
#endif

// forces_cuda.cu:868-875
void recalculate_softening() {
    beta = ParticleRadi;
    if (COSMOLOGY == 1)
        rho_part = M_min / (4.0 * 3.14159265358979323846264338327950288419716939937510 * std::pow(beta, 3.0) / 3.0);
}
