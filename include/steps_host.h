/*
 * steps_host.h -- the host side of a StePS run around the force engine:
 * parameter files, initial conditions and snapshots, and a driver
 * (steps_run) that runs a StePS parameter file on the device through
 * steps_force.h.
 *
 * These restate the reference's host code that the force routine's callers
 * depend on, in the formats a StePS user already has on disk:
 *   read_param()                     StePS/src/read_paramfile.cc:83-735
 *   measure_N_part_from_ascii_snapshot, read_ascii_ic
 *                                    StePS/src/inputoutput.cc:53-62, 100-179
 *   load_snapshot + reordering + gadget_format_conversion (IC_FORMAT 1)
 *                                    StePS/src/read_gadget_ic.cc:80-413
 *   read_OUT_LST                     StePS/src/inputoutput.cc:181-312
 *   write_ascii_snapshot             StePS/src/inputoutput.cc:826-909
 * HDF5 input and output (IC_FORMAT / OUTPUT_FORMAT 2) need the HDF5 library,
 * which this build does not link: they are refused, as the reference refuses
 * them when built without HAVE_HDF5 (inputoutput.cc:2446-2451,
 * read_paramfile.cc:196-201).
 *
 * Functions return 0 (STEPS_OK) or a negative steps_force.h error code; the
 * message is in steps_host_last_error().
 */
#ifndef STEPS_HOST_H
#define STEPS_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STEPS_UNIT_T 47.14829951063323  /* Gy per internal time unit (global_variables.h:17) */
#define STEPS_UNIT_V 20.738652969925447 /* km/s per internal velocity unit (global_variables.h:18) */
#define STEPS_PATH_MAX 1024

/* Every parameter read_param() reads, with its global's name.  Unread keys
 * keep the reference's start values: 0 (globals are zero-initialised),
 * RADIAL_FORCE_ACCURACY = RADIAL_FORCE_TABLE_SIZE = 1000 for the S1xR2 build
 * (read_paramfile.cc:116-117); the octree build's keys (USE_BH,
 * read_paramfile.cc:130-141, 348-379) start at RADIAL_BH_FORCE_CORRECTION
 * 0, GLASS_FILE_FOR_BH_FORCE_CORRECTION "None", RADIAL_BH_FORCE_TABLE_SIZE
 * 64, RADIAL_BH_FORCE_TABLE_ITERATION 4 and are clamped as read (a
 * correction flag other than 0 or 1 reads as 0; at least 5 bins and one
 * iteration).  After the file is read, as there: with
 * COSMOLOGY = 1, STEP_MIN/STEP_MAX are converted from Gy to internal units,
 * and with H0_INDEPENDENT_UNITS = 1, L_BOX and PARTICLE_RADII from Mpc/h to
 * Mpc; with COSMOLOGY != 1, H0_INDEPENDENT_UNITS is forced to 0. */
typedef struct steps_run_params {
    double Omega_b, Omega_lambda, Omega_m, Omega_r;
    double H0;                      /* HubbleConstant / UNIT_V (internal units) */
    char IC_FILE[STEPS_PATH_MAX];
    double L;                       /* L_BOX */
    int32_t IS_PERIODIC;
    int32_t OUTPUT_FORMAT;
    char OUT_DIR[STEPS_PATH_MAX];   /* always ends with '/' when set */
    double a_max;
    uint32_t N_snapshot;            /* SNAPSHOT_START_NUMBER */
    double ACC_PARAM;
    double h_min, h_max;            /* STEP_MIN, STEP_MAX */
    int32_t COSMOLOGY;
    double a_start;
    int32_t COMOVING_INTEGRATION;   /* clamped to 0..1 */
    double TIME_LIMIT_IN_MINS;      /* TIME_LIMIT_IN_MIN */
    double FIRST_T_OUT;
    int32_t IC_FORMAT;
    int32_t OUTPUT_TIME_VARIABLE;
    char OUT_LST[STEPS_PATH_MAX];
    double H_OUT;
    double ParticleRadi;            /* PARTICLE_RADII */
    double MIN_REDSHIFT;
    int32_t REDSHIFT_CONE;
    int32_t H0_INDEPENDENT_UNITS;   /* H_INDEPENDENT_UNITS */
    double Rsim;                    /* R_SIM */
    int32_t RADIAL_FORCE_ACCURACY;  /* S1xR2 build only; at least 10 */
    int32_t RADIAL_FORCE_TABLE_SIZE;/* S1xR2 build only; at least 5 */
    int32_t RADIAL_BH_FORCE_CORRECTION;
    char GLASS_FILE_FOR_BH_FORCE_CORRECTION[STEPS_PATH_MAX];
    int32_t RADIAL_BH_FORCE_TABLE_SIZE;
    int32_t RADIAL_BH_FORCE_TABLE_ITERATION;
    double w0, wa;                  /* COSMOPARAM 1 and 2 builds */
    char EXPANSION_FILE[STEPS_PATH_MAX];  /* COSMOPARAM -1 build */
    int32_t INTERPOLATION_ORDER;    /* COSMOPARAM -1: 1..3, anything else reads as 1 */
} steps_run_params;

const char *steps_host_last_error(void);

/* Parse a StePS parameter file as read_param() does: each line is tested
 * against every key with strstr (so a key matches anywhere in a line, and
 * a line can set several parameters), and the value is the line's second
 * whitespace-separated token.  topology (enum steps_topology) plays the part
 * of the build flags: IS_PERIODIC > 4 is clamped to 4 except for S1xR2, and
 * the RADIAL_FORCE_* keys exist only for S1xR2. */
int steps_read_params(const char *path, int32_t topology, steps_run_params *p);

/* ASCII initial conditions / snapshots: one particle per line,
 * x y z vx vy vz M.  steps_ascii_count_particles counts newline characters,
 * as measure_N_part_from_ascii_snapshot does. */
int64_t steps_ascii_count_particles(const char *path);
int steps_read_ascii_ic(const char *path, int64_t N, double *x, double *v, double *M);

/* GADGET-2 format-1 initial conditions (one file): header, positions,
 * velocities, IDs, variable masses.  The type-1 particles, brought into ID
 * order, with positions converted from kpc/h to Mpc and velocities kept in
 * km/s; *L is BoxSize converted the same way; M receives the masses the file
 * carries (1e10 Msun/h), which StePS itself replaces (main.cc:1305-1317).
 * With x == NULL only *N (the type-1 count), *L and *a are filled. */
int steps_read_gadget_ic(const char *path, int64_t *N, double *x, double *v, double *M, double *L,
                         double *a);

/* The output list (OUT_LST): one value per line, times in Gy or redshifts by
 * OUTPUT_TIME_VARIABLE; values are returned in file order.  With out == NULL
 * only the count is returned (>= 0). */
int64_t steps_read_output_list(const char *path, double *out, int64_t cap);

/* The expansion history (friedmann_solver.cc) of the COSMOPARAM builds:
 * 0 LCDM, 1 wCDM (w0), 2 w0-wa (w = w0 + wa (1 - a)), -1 a tabulated history
 * `table` of `rows` rows (t, a, H) in internal units, interpolated with
 * `order` 1, 2 or 3 and a cursor `index` the calls move (expansion_index).
 * Omega_k = 1 - Omega_m - Omega_lambda - Omega_r (main.cc:1261). */
typedef struct steps_expansion {
    int32_t cosmoparam;
    double H0, Omega_m, Omega_r, Omega_lambda, Omega_k, w0, wa;
    const double *table;
    int64_t rows;
    int32_t order;
    int64_t index;
} steps_expansion;
/* CALCULATE_Hubble_param (friedmann_solver.cc:161-164 LCDM, 237-240 wCDM,
 * 322-325 w0-wa, 381-415 tabulated) */
double steps_hubble(const steps_expansion *e, double a);
/* friedmann_solver_step (friedmann_solver.cc:100-159, 176-235, 255-320;
 * tabulated 343-379): a after one RK4 step of length h from a0 (the
 * tabulated history: a at the time T at the end of the step) */
double steps_friedmann_step(steps_expansion *e, double a0, double h, double T);
/* friedmann_solver_start (friedmann_solver.cc:33-95; tabulated 438-467):
 * the cosmic time of a_start (RK4 step h backwards to the big bang, halved
 * until converged) */
double steps_friedmann_start(steps_expansion *e, double a_start, double h);
/* read_expansion_history (inputoutput.cc:64-97): rows (t Gy, a, H km/s/Mpc)
 * into table [3 rows] in internal units; table == NULL returns the row
 * count */
int64_t steps_read_expansion_history(const char *path, double *table, int64_t cap);

/* save_function_to_ascii_table (inputoutput.cc:1038-1061): the header line,
 * then "%.6f\t%.15f" rows of (x_min + k deltax, values[k]) while the
 * argument does not exceed x_max. */
int steps_write_ascii_table(const char *path, double x_min, double x_max, double deltax, const double *values,
                            int32_t n, const char *header);

/* write_ascii_snapshot: x and M multiplied by h0_dimless (1, or H0/100 with
 * H0-independent units), velocities by sqrt(a) * UNIT_V (km/s sqrt(a));
 * "%.16f\t" for each of the seven columns, then a newline. */
int steps_write_ascii_snapshot(const char *path, int64_t N, const double *x, const double *v, const double *M,
                               double a, double h0_dimless);

#ifdef __cplusplus
}
#endif
#endif
