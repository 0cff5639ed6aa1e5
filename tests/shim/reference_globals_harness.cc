// Test harness for the drop-in shim: defines the globals that StePS's main.cc
// defines (global_variables.h types) and exposes C entry points so the tests
// can set them and call the shim's force routine the way step.cc:191-197 does.
#include <cstdint>

typedef double REAL;
int N = 0;
REAL *M = nullptr;
REAL *SOFT_LENGTH = nullptr;
int rank = 0;
int IS_PERIODIC = 0;
int COSMOLOGY = 0;
int COMOVING_INTEGRATION = 0;
REAL L = 0;
REAL Rsim = 0;
REAL mass_in_unit_sphere = 0;
double H0 = 0;
double Omega_lambda = 0;
bool ForceError = false;
REAL beta = 0;
REAL ParticleRadi = 0;
REAL rho_part = 0;
REAL M_min = 0;
#ifdef USE_BH
REAL THETA = 0;
int RADIAL_BH_FORCE_TABLE_SIZE = 0;
REAL *RADIAL_BH_FORCE_TABLE = nullptr;
bool USE_RADIAL_BH_CORRECTION = false;
#endif
#if defined(PERIODIC)
REAL *T3_EWALD_FORCE_TABLE = nullptr;
int N_EWALD_FORCE_GRID = 0;
#elif defined(PERIODIC_Z)
#if !defined(PERIODIC_Z_NOLOOKUP)
REAL *S1R2_EWALD_FORCE_TABLE = nullptr;
int Nz_EWALD_FORCE_GRID = 0;
int Nrho_EWALD_FORCE_GRID = 0;
#endif
int RADIAL_FORCE_TABLE_SIZE = 0;
REAL *RADIAL_FORCE_TABLE = nullptr;
#endif

void forces(REAL *x, REAL *F, int ID_min, int ID_max);
void forces_periodic(REAL *x, REAL *F, int ID_min, int ID_max);
void forces_periodic_z(REAL *x, REAL *F, int ID_min, int ID_max);
void recalculate_softening();

extern "C" {
void harness_set_particles(int n, REAL *m, REAL *soft) { N = n; M = m; SOFT_LENGTH = soft; }
void harness_set_scalars(int is_periodic, int cosmology, int comoving, double l, double rsim, double misph,
                         double h0, double omega_l) {
    IS_PERIODIC = is_periodic; COSMOLOGY = cosmology; COMOVING_INTEGRATION = comoving; L = l; Rsim = rsim;
    mass_in_unit_sphere = misph; H0 = h0; Omega_lambda = omega_l;
}
void harness_set_tables(REAL *t3, int ng, REAL *s1r2, int nrho, int nz, REAL *radial, int nradial) {
#if defined(PERIODIC)
    T3_EWALD_FORCE_TABLE = t3; N_EWALD_FORCE_GRID = ng;
#elif defined(PERIODIC_Z)
#if !defined(PERIODIC_Z_NOLOOKUP)
    S1R2_EWALD_FORCE_TABLE = s1r2; Nrho_EWALD_FORCE_GRID = nrho; Nz_EWALD_FORCE_GRID = nz;
#endif
    RADIAL_FORCE_TABLE = radial; RADIAL_FORCE_TABLE_SIZE = nradial;
#endif
    (void)t3; (void)ng; (void)s1r2; (void)nrho; (void)nz; (void)radial; (void)nradial;
}
void harness_set_softening(double particle_radi, double m_min) { ParticleRadi = particle_radi; M_min = m_min; }
void harness_set_theta(double theta) {
#ifdef USE_BH
    THETA = theta;
#endif
    (void)theta;
}
void harness_set_radial_bh(REAL *table, int size, int use) {
#ifdef USE_BH
    RADIAL_BH_FORCE_TABLE = table; RADIAL_BH_FORCE_TABLE_SIZE = size; USE_RADIAL_BH_CORRECTION = use != 0;
#endif
    (void)table; (void)size; (void)use;
}
int harness_variant() {
#if defined(PERIODIC)
    return 1;
#elif defined(PERIODIC_Z)
    return 2;
#else
    return 0;
#endif
}
int harness_call(REAL *x, REAL *F, int id_min, int id_max) {
    ForceError = false;
#if defined(PERIODIC)
    forces_periodic(x, F, id_min, id_max);
#elif defined(PERIODIC_Z)
    forces_periodic_z(x, F, id_min, id_max);
#else
    forces(x, F, id_min, id_max);
#endif
    return ForceError ? 1 : 0;
}
void harness_recalculate_softening(double *beta_out, double *rho_part_out) {
    recalculate_softening();
    *beta_out = beta;
    *rho_part_out = rho_part;
}
}
