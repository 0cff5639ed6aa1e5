/*
 * steps_force.h -- C ABI of the MI355X direct-summation gravity force engine.
 *
 * The reference selects its force engine at link time: forces_cuda.o or
 * forces.o exports forces()/forces_periodic()/forces_periodic_z() and
 * recalculate_softening() with C++ linkage, reading every other input from
 * globals (StePS/src/forces_cuda.cu:73-79, :169-173, :457-461, :868-875;
 * StePS/src/global_variables.h:42-147).  This header is the plain C boundary
 * underneath that: every input the reference reads from a global is a field
 * of steps_force_desc, all buffers are plain pointers with explicit sizes, and
 * every entry point returns a status code instead of setting ForceError or
 * calling exit().  The C++ shim in steps_amd/csrc/shim/steps_forces_shim.cc
 * maps the reference's globals onto these calls.
 *
 * Units are the reference's: Mpc, 1e11 Msun, G = 1.  Forces are accelerations
 * before the a^-3 factor of the integrator (StePS/src/step.cc:135).
 * Thread safety: one-shot calls serialise on an internal lock; a steps_sim is
 * owned by one caller thread at a time; the server is thread safe.
 */
#ifndef STEPS_FORCE_H
#define STEPS_FORCE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define STEPS_OK 0
#define STEPS_ERR_ARG (-1)       /* null pointer, negative size, bad enum       */
#define STEPS_ERR_RANGE (-2)     /* id_min/id_max outside [0, N)                 */
#define STEPS_ERR_TABLE (-3)     /* a table the topology needs is missing        */
#define STEPS_ERR_DEVICE (-4)    /* no HIP device, or a HIP runtime call failed  */
#define STEPS_ERR_NOMEM (-5)     /* device or host allocation failed             */
#define STEPS_ERR_ADMISSION (-6) /* server: request exceeds the memory budget    */
#define STEPS_ERR_STATE (-7)     /* handle used in the wrong state               */
#define STEPS_ERR_NONFINITE (-8) /* NaN/Inf in positions (input validation)      */

/* ---- configuration ------------------------------------------------------- */
enum steps_topology {
    STEPS_TOPO_R3 = 0,   /* open compactified R^3 (no -DPERIODIC*)          */
    STEPS_TOPO_T3 = 1,   /* periodic T^3 (-DPERIODIC)                       */
    STEPS_TOPO_S1R2 = 2  /* periodic z axis, open x-y (-DPERIODIC_Z)        */
};

enum steps_precision {
    STEPS_PREC_FP64 = 0, /* all pair arithmetic in FP64                      */
    STEPS_PREC_MIXED = 1 /* FP32 pair arithmetic (tile-recentred positions,
                            or for the T3 Ewald table FP64 displacements
                            rounded once), FP64 accumulation across tiles,
                            FP64 for pairs inside the softening bound; R3, T3
                            (nearest image, Ewald) and S1R2 (nearest z-image,
                            CIC, TSC)                                         */
};

/* Everything the reference force routines read from globals. */
typedef struct steps_force_desc {
    int32_t topology;            /* enum steps_topology                                  */
    int32_t is_periodic;         /* IS_PERIODIC: T3 1 nearest image, >=2 Ewald;
                                    S1R2 1 nearest z-image, >=2 Ewald / image sum          */
    int32_t cosmology;           /* COSMOLOGY                                             */
    int32_t comoving;            /* COMOVING_INTEGRATION                                  */
    double L;                    /* L: T3 box size, S1R2 z period                         */
    double Rsim;                 /* Rsim (S1R2 table extent 2.25*Rsim, radial table)      */
    double mass_in_unit_sphere;  /* external-term constant (main.cc:1288, :1313)          */
    double H0;                   /* code units; DE = H0^2*Omega_lambda (forces_cuda.cu:887) */
    double Omega_lambda;
    int32_t s1r2_lookup;         /* 1: S1R2 Ewald table (default build);
                                    0: direct z-image sum (-DPERIODIC_Z_NOLOOKUP)          */
    int32_t interp_order;        /* S1R2 EWALD_INTERPOLATION_ORDER: 0 NGP, 2 CIC, 4 TSC.
                                    T3 direct sum: always tricubic Catmull-Rom (main.cc:501);
                                    T3 Barnes-Hut: the CPU build's order (ewald_space.cc:452-
                                    523): 2 trilinear, 4 Catmull-Rom, 6 / 8 Lagrange     */
    const double *t3_table;      /* T3_EWALD_FORCE_TABLE [ng^3*3], host memory            */
    int32_t t3_ngrid;            /* N_EWALD_FORCE_GRID                                    */
    const double *s1r2_table;    /* S1R2_EWALD_FORCE_TABLE [nrho*nz*2], host memory       */
    int32_t s1r2_nrho, s1r2_nz;
    const double *radial_table;  /* RADIAL_FORCE_TABLE [size], host memory                */
    int32_t radial_table_size;
    int32_t precision;           /* enum steps_precision                                  */
    int32_t device;              /* HIP device ordinal                                    */
    double bh_theta;             /* > 0: Barnes-Hut octree forces with this opening angle
                                    (USE_BH, main.cc:303; every topology and IS_PERIODIC
                                    mode of forces.cc:373-1217; FP64 only); 0: direct sum */
    double bh_particle_radius;   /* ParticleRadi: a node's softening length is
                                    cbrt(mass / M_min) * ParticleRadi + h_i (forces.cc:386) */
    int32_t bh_randomize;        /* nonzero: the randomised octree of the RANDOMIZE_BH build
                                    (forces.cc:420-434, 662-685, 1026-1058, 1194-1211): the
                                    octree is built in a randomly moved frame and the forces
                                    are turned back.  R3: rotation by angle u3*pi about the
                                    unit vector of (theta = u4*2pi, z = u5*2-1), root cube of
                                    side 4.00004 max|x| centred at (u0..u2 - 0.5)*2 max|x|.
                                    T3: shift by (u0..u2 - 0.5)*L, wrapped into [0, L).
                                    S1R2: rotation about z by u0*2pi, root centred at
                                    ((u1, u2) - 0.5)*2 max r_xy, z shifted by (u3 - 0.5)*L
                                    and wrapped, root side L unless 4 max r_xy >= L (then
                                    4 max r_xy).  Positions are not modified.              */
    double bh_random[6];         /* u0..u5: the draws rand()/RAND_MAX of one evaluation, in
                                    the order the reference draws them (R3 six, T3 three,
                                    S1R2 four); steps_sim_set_bh_random changes them      */
    const double *radial_bh_table; /* RADIAL_BH_FORCE_TABLE [size] (host memory) or NULL:
                                    with the octree, R3 or S1R2, comoving cosmology, each
                                    row's force loses x * c(r) / r, c linearly interpolated
                                    between bin centres (forces.cc:189-267, 479-485,
                                    1143-1147, 1180-1184; r = |x| in R3, the cylindrical
                                    radius and x, y only in S1R2); see
                                    steps_radial_bh_table_accumulate                       */
    int32_t radial_bh_table_size;  /* RADIAL_BH_FORCE_TABLE_SIZE, at least 2               */
} steps_force_desc;

/* Fill *d with the defaults (R3, FP64, no cosmology, device 0). */
void steps_force_desc_init(steps_force_desc *d);

/* ---- library -------------------------------------------------------------- */
int steps_device_count(void);
const char *steps_last_error(void); /* thread-local text of the last failure */
const char *steps_version(void);

/* ---- one-shot force evaluation (the reference's force routine) ------------
 * Replaces forces_cuda() (forces_cuda.cu:878-1117), forces_periodic_cuda()
 * (:1121-1374) and forces_periodic_z_cuda() (:1378-1704) for one process and
 * one device.  x is host AoS [3*N]; M and soft are host [N]; F is host AoS
 * [3*(id_max-id_min+1)] and rows id_min..id_max (inclusive) are overwritten
 * at offset 3*(i-id_min), as the GPU reference does.  Device buffers and
 * tables are cached across calls instead of allocated per call. */
int steps_force_eval(const steps_force_desc *d, int64_t N, const double *x, const double *M,
                     const double *soft, double *F, int64_t id_min, int64_t id_max);

/* steps_force_eval without any device allocation.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, results
 * (bitwise) and error behaviour as steps_force_eval, but every device buffer
 * the evaluation needs (staged inputs, packed records, forces, the pair
 * kernels' scratch, the tables' device copies) comes from one static device
 * array of STEPS_NOALLOC_SCRATCH_MB MiB per device, held until the process
 * ends; an evaluation that needs more (R3 FP64: up to N of about 2^18)
 * fails with STEPS_ERR_ARG and says how many bytes it needed.  Calls are
 * serialised within the process; all calls on a device share the array,
 * so calls that overlap on one device (other processes) race.  Work runs on
 * the null stream. */
#define STEPS_NOALLOC_SCRATCH_MB 256
int steps_force_eval_noalloc(const steps_force_desc *d, int64_t N, const double *x, const double *M,
                             const double *soft, double *F, int64_t id_min, int64_t id_max);

/* ---- resident simulation (state kept in HBM between calls) ---------------- */
typedef struct steps_sim steps_sim;

int steps_sim_create(const steps_force_desc *d, int64_t N, steps_sim **out);
int steps_sim_destroy(steps_sim *s);
/* Switch the precision of later evaluations (STEPS_PREC_FP64 / _MIXED).  No
 * counterpart in the reference, which fixes the precision at compile time
 * (USE_SINGLE_PRECISION, global_variables.h:26-32); the scratch of a sim
 * whose variant has a mixed path is sized for both, so switching allocates
 * nothing. */
int steps_sim_set_precision(steps_sim *s, int32_t precision);
/* The draws of the randomised octree for later evaluations (u[6], the
 * descriptor's bh_random): the reference draws them afresh in every force
 * routine call (forces.cc:422-430, 662-667, 1026-1034). */
int steps_sim_set_bh_random(steps_sim *s, const double *u);
/* x, v AoS [3N] (v may be NULL: zero), M and soft [N]; host memory. */
int steps_sim_set_state(steps_sim *s, const double *x, const double *v, const double *M,
                        const double *soft);
int steps_sim_set_positions(steps_sim *s, const double *x);
int steps_sim_get_state(steps_sim *s, double *x, double *v);
/* Queue one force evaluation for rows id_min..id_max on the sim's stream. */
int steps_sim_forces(steps_sim *s, int64_t id_min, int64_t id_max);
/* Copy the last evaluation's forces (AoS [3*(id_max-id_min+1)]) to host. */
int steps_sim_get_forces(steps_sim *s, double *F);
int steps_sim_synchronize(steps_sim *s);
/* Duration of the last force evaluation's kernels, measured with HIP events
 * on the sim's stream: pair kernel and whole evaluation (pair + reduction). */
int steps_sim_last_kernel_ms(steps_sim *s, float *pair_ms, float *total_ms);

/* One kick-drift-kick step with Hubble drag (StePS/src/step.cc:100-312),
 * all on the device: A = G F a^-3 - 2 H v; v += A h/2, x += v h, periodic
 * wrap (T3: x,y,z; S1R2: z), forces, A' with the updated a and H, v += A' h/2,
 * errmax = max_i |A'_i| / cbrt(M_i 3/(4 pi rho_part)).  a_kick1/hubble1 are the
 * scale factor and H(a) before the step, a_kick2/hubble2 after the host's
 * Friedmann update (step.cc:233-243); G is +1.  The stored F must hold the
 * forces of the current positions for all N rows: after steps_sim_forces(s,
 * 0, N-1) or a previous step; otherwise (no evaluation yet, a sub-range
 * evaluation, new positions) STEPS_ERR_STATE.  After a step F holds the
 * step's own full-range forces, so steps_sim_get_forces works without
 * another evaluation; steps_sim_last_kernel_ms times only steps_sim_forces. */
typedef struct steps_step_params {
    double h;
    double a_kick1, hubble1;
    double a_kick2, hubble2;
    double G;
    double rho_part;
} steps_step_params;
int steps_sim_step(steps_sim *s, const steps_step_params *p, double *errmax);
/* errmax of the stored forces without stepping (step.cc:35-98); needs the
 * full-range forces as steps_sim_step does. */
int steps_sim_errmax(steps_sim *s, double a, double hubble, double G, double rho_part,
                     double *errmax);

/* ---- tables (built on the device, returned in host memory) ----------------
 * T3 Ewald correction, ewald_space.cc:288-383 with main.cc:422-448 params;
 * call with table==NULL to get *ngrid, then with a [ngrid^3*3] buffer.
 * The three builders cache their results by parameters (up to 2 GiB). */
int steps_t3_table(int32_t is_periodic, double L, double *table, int32_t *ngrid);
/* S1R2 Tornberg 1P correction, ewald_space.cc:754-798 with main.cc:569-602;
 * table==NULL queries (*nrho, *nz). */
int steps_s1r2_table(int32_t is_periodic, double L, double Rsim, double *table,
                     int32_t *nrho, int32_t *nz);
/* finite-cylinder radial force table, utils.cc:162-228 */
int steps_radial_table(double R, double Lz, int32_t size, int32_t accuracy, double *table);
/* The radial octree correction table (main.cc:950-1088, utils.cc:96-156):
 * one call per octree force evaluation of the glass adds each particle's
 * radial force (F . x / |x| in R3; the x-y parts over the cylindrical
 * radius in S1R2) to sum[b] and one to count[b], b = floor(r / (Rsim /
 * size)) capped at size - 1; host memory, N >= 0 rows of x and F.  After the
 * last call steps_radial_bh_table_finish turns the sums into the table:
 * sum[b] / count[b], or 0 for an empty bin. */
int steps_radial_bh_table_accumulate(int32_t topology, double Rsim, int32_t size, int64_t N, const double *x,
                                     const double *F, double *sum, int32_t *count);
int steps_radial_bh_table_finish(int32_t size, double *sum, const int32_t *count);

/* ---- serving loop ----------------------------------------------------------
 * No counterpart in the reference (one simulation per MPI job there).  A
 * server keeps an ensemble of independent simulations, takes force and step
 * requests from any thread, and runs them on one device with continuous
 * batching: simulations are admitted into HBM by memory budget (idle ones are
 * evicted LRU to host memory and re-admitted with their state), up to
 * `streams` requests run concurrently (one per simulation at a time, in
 * submission order per simulation), each simulation's force/step sequence is
 * replayed from a captured hipGraph, and the pair kernels of ready requests
 * of small simulations (N <= pack_max_n, same variant and precision) are
 * packed into one launch.  Results do not depend on the mix. */
typedef struct steps_server steps_server;

typedef struct steps_server_config {
    int32_t device;
    int32_t streams;           /* concurrent requests on the device, 1..16 (default 4) */
    int32_t use_graphs;        /* 1: replay captured per-simulation graphs (default)    */
    int64_t hbm_budget_bytes;  /* resident simulations' state + scratch (default 200 GiB) */
    int64_t pack_max_n;        /* pack simulations with N <= this (default 65536; 0: never) */
} steps_server_config;

enum steps_request_kind { STEPS_REQ_FORCE = 0, STEPS_REQ_STEP = 1 };

typedef struct steps_request_result {
    int32_t status;            /* STEPS_OK or an error code                        */
    int32_t kind;
    int64_t sim;
    int64_t ticket;
    double errmax;             /* STEP requests: max_i |A_i| / h_i after the step  */
    double t_submit, t_start, t_end;  /* seconds on steps_server_clock()           */
} steps_request_result;

typedef struct steps_server_stats {
    int64_t submitted, completed, failed;
    int64_t admissions, evictions, graph_captures;
    int64_t resident_sims, bytes_resident, max_backlog;
    int64_t packed_launches, packed_requests;  /* launches holding >= 2 requests, and those requests */
    int64_t tickets_held;      /* requests whose result has not been collected yet */
    int64_t admitting;         /* admissions in progress (device copies being made) */
} steps_server_stats;

void steps_server_config_init(steps_server_config *c);
int steps_server_create(const steps_server_config *c, steps_server **out);
/* Abandons pending requests (they finish with STEPS_ERR_STATE), waits for the
 * ones on the device, frees everything. */
int steps_server_destroy(steps_server *s);
/* Registers a simulation (host arrays are copied; v may be NULL).  It becomes
 * resident when it first has work.  Fails with STEPS_ERR_ADMISSION if it
 * could never fit the budget.  Table pointers of *d may be NULL: the server
 * then builds the tables the reference would for that descriptor (T3 and
 * S1R2 Ewald tables, main.cc:422-448 and :569-602; the radial table with
 * RADIAL_FORCE_TABLE_SIZE = RADIAL_FORCE_ACCURACY = 1000 unless
 * radial_table_size is set, main.cc:1266-1302); built tables are cached by
 * their parameters, so an ensemble sharing a box builds each table once. */
int steps_server_add_sim(steps_server *s, const steps_force_desc *d, int64_t N, const double *x,
                         const double *v, const double *M, const double *soft, int64_t *sim_id);
int steps_server_remove_sim(steps_server *s, int64_t sim_id);
/* Device bytes the simulation takes while resident (state + scratch + tables). */
int64_t steps_server_sim_bytes(steps_server *s, int64_t sim_id);
/* kind: STEPS_REQ_FORCE (all rows) or STEPS_REQ_STEP (p required); precision
 * per request (enum steps_precision).  A step kicks with the forces of the
 * current positions, so a simulation's first request must be a force
 * evaluation: a step before it finishes with STEPS_ERR_STATE. */
int steps_server_submit(steps_server *s, int64_t sim_id, int32_t kind, int32_t precision,
                        const steps_step_params *p, int64_t *ticket);
/* Blocks until the request is done and returns its result once: the ticket
 * is then collected, and the server forgets collected tickets in order, so a
 * long-running loop keeps only the window of tickets not yet collected. */
int steps_server_wait(steps_server *s, int64_t ticket, steps_request_result *out);
/* Blocks until every submitted request is done (results stay collectable). */
int steps_server_drain(steps_server *s);
/* After the simulation's requests finished: x, v, F (AoS [3N], any may be NULL). */
int steps_server_get_state(steps_server *s, int64_t sim_id, double *x, double *v, double *F);
int steps_server_get_stats(steps_server *s, steps_server_stats *out);
double steps_server_clock(steps_server *s);

/* ---- softening set-up (utils.cc:59-82) ------------------------------------ */
int steps_softening_lengths(const double *M, int64_t N, double particle_radius, double *soft,
                            double *M_min, double *rho_part);

#ifdef __cplusplus
}
#endif
#endif
