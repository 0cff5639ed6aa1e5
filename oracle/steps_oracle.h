/*
 * steps_oracle.h -- CPU oracle for the StePS direct-summation force path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the steps_amd library,
 * its C ABI or the C++ drop-in shim) links, loads or calls this code.  Only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py use it,
 * as the checker and as the timed CPU baseline.
 *
 * Every function restates the reference algorithm from its source; the
 * comment above each definition in steps_oracle.c cites the file:line it
 * follows (paths relative to the StePS repository root).
 *
 * Pinning status (how each function is checked independently of itself):
 *   force_softening        closed form + enclosed-mass integral of the
 *                          Gadget-2 spline density (tests/test_oracle_softening.py)
 *   forces R3              long-double brute force, Newton's third law,
 *                          shell theorem, exact external term
 *   t3 table + forces      antisymmetry, O_h symmetry, independent Ewald sum
 *                          with another alpha, perfect-lattice zero force,
 *                          Catmull-Rom node reproduction
 *   s1r2 table + forces    direct real-space z-image sum; an exact-erfc
 *                          numpy 1P Ewald (itself checked against the image
 *                          sum) to 1e-11 max|D| where the A&S fit carries no
 *                          weight, within the fit's bound elsewhere; uniform
 *                          periodic cylinder (comoving term cancels the pair
 *                          sum in every mode); CIC bilinear / TSC linear data
 *                          off the nodes, NGP half-way rounding
 *   radial table           independent adaptive quadrature of the integrand,
 *                          long-cylinder limit; its correction of order 1-3
 *                          reproduces linear / quadratic / cubic data
 *   softening setup, kdk   closed forms, two-body orbit, force-free Hubble
 *                          drag against v a^2 = const and x(t) by quadrature,
 *                          periodic wraps
 *   r3 fp32 stand-in      two-body case exact to float rounding, error
 *                          against the FP64 oracle within float bounds and
 *                          growing with N (tests/test_oracle_r3.py); it is a
 *                          comparison point, not a parity oracle
 *   forces_bh(_ex)         tiny opening angle = direct sum in every topology
 *                          and mode, with and without the random frame; a far
 *                          cluster taken at its centre of mass; error falls
 *                          with theta; identity draws = plain octree bit for
 *                          bit; root cubes by the placement rules; averaging
 *                          over draws shrinks the error (tests/test_oracle_bh*.py);
 *                          with a huge opening angle (every row takes the
 *                          root) a row's force is the closed form at the
 *                          extended-precision centre of mass, for the
 *                          running averages and for com_from_children
 *                          (tests/test_oracle_bh_com.py)
 *   radial_bh correction   linear data between bin centres come back, the
 *                          origin line below bin 1's centre; the table is the
 *                          binned mean radial force; the corrected force is
 *                          the plain one less x c(r) / r, and only comoving
 *   fast_erfcx             the published A&S 7.1.26 bound (1.5e-7 on erfc)
 *                          and its reflection for x < 0; its contribution to
 *                          the S1R2 table is bounded node by node (<= ~1e-10
 *                          max|D| at the S1R2 workload's box).
 */
#ifndef STEPS_ORACLE_H
#define STEPS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_R3 = 0, ORC_T3 = 1, ORC_S1R2 = 2 };

/* Everything the reference force routines read from globals
 * (global_variables.h:42-147, main.cc:37-147). */
typedef struct oracle_params {
    int topology;              /* ORC_R3 / ORC_T3 / ORC_S1R2 (build flag PERIODIC / PERIODIC_Z) */
    int is_periodic;           /* IS_PERIODIC */
    int cosmology;             /* COSMOLOGY */
    int comoving;              /* COMOVING_INTEGRATION */
    double L;                  /* L: T3 box size, S1R2 z period */
    double Rsim;               /* Rsim */
    double mass_in_unit_sphere;
    double H0;                 /* code units (km/s/Mpc / UNIT_V) */
    double Omega_lambda;
    int s1r2_lookup;           /* 1: Ewald table (default build); 0: PERIODIC_Z_NOLOOKUP */
    int interp_order;          /* T3: 4 Catmull-Rom (GPU), 2 trilinear (CPU default);
                                  S1R2: 0 NGP, 2 CIC, 4 TSC (EWALD_INTERPOLATION_ORDER) */
    int cut_inclusive;         /* S1R2 no-lookup image cut: 1 '<=' (forces.cc:1274), 0 '<' (forces_cuda.cu:730) */
    const double *t3_table;    /* T3_EWALD_FORCE_TABLE [Ng^3*3] */
    int t3_ngrid;              /* N_EWALD_FORCE_GRID */
    const double *s1r2_table;  /* S1R2_EWALD_FORCE_TABLE [Nrho*Nz*2] */
    int s1r2_nrho, s1r2_nz;
    const double *radial_table;/* RADIAL_FORCE_TABLE */
    int radial_table_size;
} oracle_params;

/* softening law */
double oracle_force_softening(double r, double beta);
double oracle_force_softening_gpu(double r, double beta);

/* direct sums: F[3*(i-id_min)+k] overwritten for rows id_min..id_max (inclusive).
 * Deterministic: rows in parallel, j summed in order 0..N-1 per row.
 * Returns 0, or a negative code for bad arguments. */
int oracle_forces(const oracle_params *p, int N, const double *x, const double *M,
                  const double *h, double *F, int id_min, int id_max);
/* same over an explicit row list (F row k belongs to rows[k]) */
int oracle_forces_rows(const oracle_params *p, int N, const double *x, const double *M,
                       const double *h, double *F, const int *rows, int nrows);

/* the reference's single-precision (REAL = float) R3 kernel over an explicit
 * row list: all-float arithmetic and accumulation, a statistical stand-in
 * for that build (the comparison point of the mixed-precision path).
 * Returns -4 for a topology other than R3. */
int oracle_forces_r3_fp32_rows(const oracle_params *p, int N, const double *x, const double *M,
                               const double *h, double *F, const int *rows, int nrows);

/* reference-faithful OpenMP baseline (schedule(dynamic,chunk), pow, three
 * omp atomic adds per pair, -O3 -ffast-math); used only for timing. */
int oracle_forces_faithful_rows(const oracle_params *p, int N, const double *x,
                                const double *M, const double *h, double *F,
                                const int *rows, int nrows);
/* optimised CPU variant of the same mathematics (no atomics, register
 * accumulation); context row for the CPU baseline. */
int oracle_forces_fast_rows(const oracle_params *p, int N, const double *x,
                            const double *M, const double *h, double *F,
                            const int *rows, int nrows);
int oracle_max_threads(void);

/* long-double brute force, R3 only (small N) */
void oracle_forces_r3_longdouble(const oracle_params *p, int N, const double *x,
                                 const double *M, const double *h, long double *F);

/* T3 Ewald table and interpolation */
int  oracle_ewald_space_count(double R);
int  oracle_ewald_space(double R, int *idx /* [count*4] */);
void oracle_t3_ewald_correction(const double r[3], double L, double alpha,
                                double rel_cut, double rec_cut, double D[3]);
void oracle_t3_ewald_table(int Ngrid, double L, double alpha, double rel_cut,
                           double rec_cut, double *table);
void oracle_t3_params(int is_periodic, double L, double *alpha, double *rel_cut,
                      double *rec_cut, int *ngrid);
void oracle_t3_interpolate(int Ngrid, double L, const double *table, double dx,
                           double dy, double dz, double D[3], int order);

/* S1R2 Ewald table and interpolation */
double oracle_fast_erfcx(double x);
void oracle_s1r2_force_pair(double dx, double dy, double dz, double Lz, double alpha,
                            int nmax, int mmax, int rspace, double F[3]);
void oracle_s1r2_table(int Nrho, int Nz, double rho_max, double Lz, double alpha,
                       int nmax, int mmax, int rspace, double *table);
void oracle_s1r2_params(int is_periodic, double L, double Rsim, int rspace,
                        double *alpha, int *rel_cut, int *rec_cut, int *nz, int *nrho);
void oracle_s1r2_interpolate(const double *T, int Nrho, int Nz, double rho_max,
                             double Lz, double dx, double dy, double dz, int order,
                             double D[3]);

/* cylindrical radial force table and its linear interpolation */
void oracle_cylindrical_force_table(double *table, double R, double Lz, int size,
                                    int accuracy);
double oracle_cylindrical_force_correction(double r, double R, const double *table,
                                           int size);
/* the CPU twin with interpolation order 1, 2 or 3 (forces.cc:279-316) */
double oracle_cylindrical_force_correction_order(double r, double R, const double *table,
                                                 int size, int order);

/* softening set-up: h_i = cbrt(M_i*3/(4 pi rho_part)) */
void oracle_softening_lengths(const double *M, int N, double particle_radius,
                              double *h, double *M_min, double *rho_part);

/* cosmology (COSMOPARAM 0) */
typedef struct oracle_cosmo {
    double H0, Omega_m, Omega_lambda, Omega_r, Omega_k;
} oracle_cosmo;
double oracle_friedmann_step(const oracle_cosmo *c, double a0, double h);
double oracle_hubble(const oracle_cosmo *c, double a);

/* one KDK step (rank-0 view, one process): returns errmax.  On entry F holds
 * the force at x; on exit x, v, F, *a are advanced. */
typedef struct oracle_step_state {
    double a, hubble, h;          /* scale factor, H(a), step size */
    double rho_part;
    double G;                     /* +1 (GLASS_MAKING would be -1) */
} oracle_step_state;
double oracle_kdk_step(const oracle_params *p, const oracle_cosmo *c,
                       oracle_step_state *s, int N, double *x, double *v,
                       double *F, const double *M, const double *hsoft);
double oracle_initial_errmax(const oracle_params *p, oracle_step_state *s, int N,
                             double *x, const double *v, const double *F,
                             const double *M);

/* Barnes-Hut octree forces (USE_BH build; forces.cc:97-187, 373-506 R3,
 * 582-772 T3, 881-1217 S1R2): the reference's pointer octree built by
 * inserting particles 0..N-1 into the topology's root cube (R3: side
 * 2.00002 max|x_i| at the origin; T3: the box; S1R2: side L, or twice the
 * largest cylindrical radius if that is not below L, on the axis at z =
 * side/2), then per row a recursive traversal (children 0..7) that takes a
 * leaf, or a node with nodesize / dist < theta, at its centre of mass with
 * softening cbrt(mass / M_min) * particle_radius + h_i, and opens the others;
 * the own leaf and empty nodes are skipped.  Pair terms: R3 plain; T3
 * nearest image, plus for IS_PERIODIC >= 2 the table correction of the CPU
 * build (oracle_t3_interpolate, order p->interp_order); S1R2 nearest
 * z-image, plus the table correction for the lookup build; the S1R2 image
 * sum (s1r2_lookup = 0) walks once per z image with the (ewald_max - 0.4) L
 * cut.  External terms as forces.cc:474-491 (R3) and 1127-1190 (S1R2, x and
 * y only, times the CPU cylindrical correction of order 1 unless the table
 * is in use); the radial BH correction table is not restated.  Rows
 * id_min..id_max, F overwritten.  Returns 0, -3 for a missing table, or -5
 * when two particles are too close to separate within 64 levels. */
int oracle_forces_bh(const oracle_params *p, double theta, double particle_radius, int N, const double *x,
                     const double *M, const double *h, double *F, int id_min, int id_max);
/* Nodes the last oracle_forces_bh call examined (massless and own-leaf nodes
 * excluded) and took, summed over its rows. */
void oracle_bh_last_counts(long long *visited, long long *accepted);

/* The octree options of other build flags.  randomize: the RANDOMIZE_BH
 * build, with u[] the draws rand()/RAND_MAX of the call in the reference's
 * order (forces.cc:422-430 R3: centre x, y, z, angle, axis theta, axis z;
 * :662-667 T3: shift x, y, z; :1026-1034 S1R2: angle, centre x, y, z
 * shift).  radial_bh_table: USE_RADIAL_BH_CORRECTION with
 * RADIAL_BH_FORCE_TABLE (R3 and S1R2, comoving cosmology only). */
typedef struct oracle_bh_opts {
    int randomize;
    double u[6];
    const double *radial_bh_table;
    int radial_bh_table_size;
    /* 0: the reference's centres of mass, running averages updated at each
     * insertion (forces.cc:164-170); 1: after the insertions, each internal
     * node's mass and centre of mass summed from its children's, bottom-up
     * (the same tree and walk; a check that isolates the rounding of the
     * running averages, not the reference's arithmetic) */
    int com_from_children;
} oracle_bh_opts;
/* oracle_forces_bh with those options (o may be NULL): the positions are
 * moved into the random frame (R3 forces.cc:324-356, 420-434; T3 :662-685;
 * S1R2 :884-899, 1026-1073) on a copy, the octree is built and walked
 * there, the external term and the radial correction are added there, and
 * the force is turned back (forces.cc:500-503, 1207-1210). */
int oracle_forces_bh_ex(const oracle_params *p, double theta, double particle_radius, const oracle_bh_opts *o,
                        int N, const double *x, const double *M, const double *h, double *F, int id_min,
                        int id_max);
/* The root cube of the last oracle_forces_bh(_ex) call: centre x, y, z, side. */
void oracle_bh_last_root(double out[4]);
/* get_BH_radial_force_correction with RADIAL_BH_FORCE_ORDER 1 (forces.cc:
 * 182-245); r within rounding of R, whose bin index is size, is taken in the
 * last bin (the reference reads one entry past the table there). */
double oracle_radial_bh_correction(double r, double R, const double *table, int size);
/* get_radial_bh_force_correction_table (utils.cc:96-156): adds one force
 * evaluation's per-bin radial force sums to table[] and counts to count[]. */
void oracle_radial_bh_table_accumulate(int topology, double Rsim, int size, int N, const double *x,
                                       const double *F, double *table, int *count);

#ifdef __cplusplus
}
#endif
#endif
