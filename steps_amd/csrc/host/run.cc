// run.cc -- steps_run: a StePS simulation driven from a StePS parameter file,
// with the force evaluations and the kick-drift-kick steps on the device.
//
//   steps_run <parameter file> [--topology r3|t3|s1r2] [--precision fp64|mixed]
//             [--interp-order 0|2|4|6|8] [--no-lookup] [--bh THETA]
//             [--bh-seed SEED] [--cosmo-param 0|1|2|-1] [--max-steps K] [--device D]
//
// The topology, the interpolation order, the table / no-table mode, the
// Barnes-Hut octree with its opening angle and its randomisation are build
// flags in the reference (PERIODIC, PERIODIC_Z, EWALD_INTERPOLATION_ORDER,
// PERIODIC_Z_NOLOOKUP, USE_BH=THETA, RANDOMIZE_BH=SEED); here they are
// options, and so is the expansion history (COSMOPARAM: 0 LCDM, 1 wCDM with
// w0, 2 w0-wa, -1 the tabulated EXPANSION_FILE with INTERPOLATION_ORDER;
// cosmo.cc).  With the octree, RADIAL_BH_FORCE_CORRECTION = 1 in R3 or S1R2
// first builds the radial correction table from the octree forces of the
// glass file (or the IC file) as main.cc:791-1091 does, averaged over
// RADIAL_BH_FORCE_TABLE_ITERATION randomised evaluations, writes it to
// OUT_DIR/radial_bh_force_table.dat and subtracts it in every evaluation.
// With --bh-seed, rand() is seeded as main.cc:310 does and every octree
// evaluation draws its frame from it in the reference's order.
// The run follows main.cc:1100-1864 and step.cc:100-312:
//   * initial conditions (ASCII or GADGET), H0-independent units, velocity
//     rescaling (main.cc:1144-1181); equal GADGET masses for T3, the mass
//     consistency check of the periodic modes (main.cc:1304-1364);
//   * softening lengths (utils.cc:59-82), the tables of the mode
//     (main.cc:404-726, 1256-1302), the initial time from the Friedmann
//     equation, the output schedule (main.cc:1422-1565);
//   * the initial time from the expansion history (friedmann_solver.cc:32-95);
//   * initial force, initial step length (step.cc:35-98), then steps until
//     a (or T) reaches a_max: a device KDK step, the scale factor advanced by
//     RK4 in between its two kicks, snapshots at output times or redshifts,
//     the next step length sqrt(2 ACC_PARAM / errmax) clamped to
//     [STEP_MIN, STEP_MAX] and cut to land on the next output time.
// Outputs are ASCII snapshots (OUTPUT_FORMAT 0) and a log of the steps.
// Not provided: HDF5 input/output, redshift-cone output, glass making.
#include "host_common.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

struct Opts {
    std::string param;
    int topology = STEPS_TOPO_R3;
    int precision = STEPS_PREC_FP64;
    int interp_order = 2;  // EWALD_INTERPOLATION_ORDER of the reference's Makefile
    int lookup = 1;
    double bh_theta = 0.0;
    long bh_seed = -1;  // RANDOMIZE_BH
    int cosmoparam = 0;  // COSMOPARAM (Template-LinuxGCC-Makefile:44)
    long max_steps = -1;
    int device = 0;
};

int die(const std::string &msg, int code = 1) {
    std::fprintf(stderr, "steps_run: %s\n", msg.c_str());
    return code;
}

bool parse_opts(int argc, char **argv, Opts &o) {
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&]() -> std::string { return (i + 1 < argc) ? argv[++i] : std::string(); };
        if (a == "--topology") {
            const std::string v = val();
            o.topology = v == "r3" ? STEPS_TOPO_R3 : v == "t3" ? STEPS_TOPO_T3 : v == "s1r2" ? STEPS_TOPO_S1R2 : -1;
            if (o.topology < 0) return false;
        } else if (a == "--precision") {
            const std::string v = val();
            o.precision = v == "mixed" ? STEPS_PREC_MIXED : STEPS_PREC_FP64;
        } else if (a == "--interp-order") {
            o.interp_order = std::atoi(val().c_str());
        } else if (a == "--no-lookup") {
            o.lookup = 0;
        } else if (a == "--bh") {
            o.bh_theta = std::atof(val().c_str());
            if (!(o.bh_theta > 0.0)) return false;
        } else if (a == "--bh-seed") {
            const std::string v = val();
            if (v.empty()) return false;
            o.bh_seed = std::atol(v.c_str());
            if (o.bh_seed < 0) return false;
        } else if (a == "--cosmo-param") {
            const std::string v = val();
            if (v != "0" && v != "1" && v != "2" && v != "-1") return false;
            o.cosmoparam = std::atoi(v.c_str());
        } else if (a == "--max-steps") {
            o.max_steps = std::atol(val().c_str());
        } else if (a == "--device") {
            o.device = std::atoi(val().c_str());
        } else if (o.param.empty() && a.size() && a[0] != '-') {
            o.param = a;
        } else {
            return false;
        }
    }
    return !o.param.empty() && (o.bh_seed < 0 || o.bh_theta > 0.0);
}

// rand()'s state for the randomised octree: the state srand(seed) gives
// (glibc's 128-byte default kind), kept apart from the process's own, which
// the HIP runtime draws from as it starts.
char g_draw_state[128];

void seed_draws(unsigned seed) {  // main.cc:310
    char *prev = initstate(seed, g_draw_state, sizeof g_draw_state);
    setstate(prev);
}

// The draws of one randomised octree evaluation, rand()/RAND_MAX in the
// order the reference's force routine makes them (forces.cc:422-430 and
// 362-363 R3, 664-667 T3, 1027-1030 S1R2).
void draw_frame(int topology, double *u) {
    const int n = topology == STEPS_TOPO_R3 ? 6 : topology == STEPS_TOPO_T3 ? 3 : 4;
    char *prev = setstate(g_draw_state);
    for (int k = 0; k < 6; ++k) u[k] = k < n ? (double)std::rand() / (double)RAND_MAX : 0.0;
    setstate(prev);
}

// write_ascii_snapshot's file name: t<100 t Gy>.dat or z<1000 z>.dat
// (inputoutput.cc:836-872), t_next being the output time or redshift due.
std::string snapshot_name(const steps_run_params &P, double t_next) {
    char A[64];
    if (P.COSMOLOGY == 1)
        std::snprintf(A, sizeof A, "%d", (int)std::round(P.OUTPUT_TIME_VARIABLE == 0 ? 100 * t_next * STEPS_UNIT_T
                                                                                    : 1000 * t_next));
    else
        std::snprintf(A, sizeof A, "%d", (int)std::round(100 * t_next));
    return std::string(P.OUT_DIR) + (P.OUTPUT_TIME_VARIABLE == 0 ? "t" : "z") + A + ".dat";
}

}  // namespace

int main(int argc, char **argv) {
    Opts o;
    if (!parse_opts(argc, argv, o))
        return die("usage: steps_run <parameter file> [--topology r3|t3|s1r2] [--precision fp64|mixed] "
                   "[--interp-order 0|2|4|6|8] [--no-lookup] [--bh THETA] [--bh-seed SEED] "
                   "[--cosmo-param 0|1|2|-1] [--max-steps K] [--device D]", 2);
    if (o.bh_seed >= 0) seed_draws((unsigned)o.bh_seed);
    steps_run_params P;
    if (steps_read_params(o.param.c_str(), o.topology, &P)) return die(steps_host_last_error());
    if (P.IC_FORMAT == 2) return die("IC_FORMAT 2 (HDF5) needs the HDF5 library, which this build does not link");
    if (P.IC_FORMAT != 0 && P.IC_FORMAT != 1) return die("bad IC format");
    if (P.REDSHIFT_CONE == 1) return die("REDSHIFT_CONE output is not provided by this driver");
    const bool cosmo = P.COSMOLOGY == 1, comoving = cosmo && P.COMOVING_INTEGRATION == 1;
    if (cosmo && !comoving && P.IS_PERIODIC > 0)
        return die("COSMOLOGY = 1, IS_PERIODIC > 0 and COMOVING_INTEGRATION = 0 cannot be run (main.cc:1369-1374)");
    if (o.topology != STEPS_TOPO_R3 && P.IS_PERIODIC < 1) return die("periodic topologies need IS_PERIODIC >= 1");

    // ---- initial conditions (inputoutput.cc:2440-2501)
    int64_t N = 0;
    std::vector<double> x, v, M;
    if (P.IC_FORMAT == 0) {
        N = steps_ascii_count_particles(P.IC_FILE);
        if (N <= 0) return die(N < 0 ? steps_host_last_error() : std::string("empty IC file ") + P.IC_FILE);
        x.resize(3 * N), v.resize(3 * N), M.resize(N);
        if (steps_read_ascii_ic(P.IC_FILE, N, x.data(), v.data(), M.data())) return die(steps_host_last_error());
    } else {
        double Lg = 0, ag = 0;
        if (steps_read_gadget_ic(P.IC_FILE, &N, nullptr, nullptr, nullptr, &Lg, &ag)) return die(steps_host_last_error());
        x.resize(3 * N), v.resize(3 * N), M.resize(N);
        if (steps_read_gadget_ic(P.IC_FILE, &N, x.data(), v.data(), M.data(), &Lg, &ag))
            return die(steps_host_last_error());
        P.L = Lg;  // "The IC determines the box size" (inputoutput.cc:2466)
        if (!(o.topology == STEPS_TOPO_T3 && cosmo))
            return die("GADGET initial conditions carry usable masses only for periodic cosmological runs "
                       "(main.cc:1305-1317)");
    }
    std::printf("steps_run: %lld particles from %s\n", (long long)N, P.IC_FILE);
    const double h0d = P.H0 * STEPS_UNIT_V / 100.0;
    if (P.H0_INDEPENDENT_UNITS != 0 && cosmo) {  // main.cc:1131-1145
        if (P.H0 == 0.0) return die("H0 is zero while H0-independent units are used");
        for (auto &c : x) c /= h0d;
        for (auto &m : M) m /= h0d;
    }
    if (comoving)  // main.cc:1147-1163: GADGET's sqrt(a) convention
        for (auto &c : v) c = c / std::sqrt(P.a_start) / STEPS_UNIT_V;
    else if (cosmo)
        for (auto &c : v) c /= STEPS_UNIT_V;

    // ---- cosmology, masses (main.cc:1256-1369)
    steps_expansion E{};
    E.cosmoparam = o.cosmoparam;
    E.H0 = P.H0;
    E.Omega_m = P.Omega_m;
    E.Omega_r = P.Omega_r;
    E.Omega_lambda = P.Omega_lambda;
    E.Omega_k = 1. - P.Omega_m - P.Omega_lambda - P.Omega_r;  // main.cc:1261
    E.w0 = P.w0;
    E.wa = P.wa;
    E.order = P.INTERPOLATION_ORDER;
    std::vector<double> expansion;
    if (cosmo && o.cosmoparam == -1) {  // main.cc:1436-1446
        const int64_t rows = steps_read_expansion_history(P.EXPANSION_FILE, nullptr, 0);
        if (rows < 0) return die(std::string("the expansion history file ") + P.EXPANSION_FILE + " does not exist");
        if (rows < 4) return die("the expansion history needs at least 4 rows");
        expansion.resize(3 * (size_t)rows);
        if (steps_read_expansion_history(P.EXPANSION_FILE, expansion.data(), rows) != rows)
            return die(steps_host_last_error());
        E.table = expansion.data();
        E.rows = rows;
    }
    const double rho_crit = 3.0 * P.H0 * P.H0 / (8.0 * kPi);
    double misph = 0.0;
    if (comoving) {
        misph = (o.topology == STEPS_TOPO_S1R2) ? 2.0 * kPi * rho_crit * P.Omega_m
                                                : 4.0 * kPi * rho_crit * P.Omega_m / 3.0;
        if (P.IC_FORMAT == 1 && o.topology == STEPS_TOPO_T3)
            for (auto &m : M) m = P.Omega_m * rho_crit * std::pow(P.L, 3.0) / (double)N;
        if (o.topology != STEPS_TOPO_R3) {
            double sum = 0.0, comp = 0.0;  // Kahan, as main.cc:1321-1330
            for (double m : M) {
                const double y = m - comp, t = sum + y;
                comp = (t - sum) - y;
                sum = t;
            }
            const double vol = (o.topology == STEPS_TOPO_T3) ? std::pow(P.L, 3.0) : kPi * P.Rsim * P.Rsim * P.L;
            const double ratio = sum / vol / (rho_crit * P.Omega_m);
            if (std::fabs(ratio - 1) > 1e-2)
                return die("the particle masses are inconsistent with the cosmological parameters (rho_part/rho_cosm = " +
                           std::to_string(ratio) + ")");
            if (std::fabs(ratio - 1) > 1e-5) {
                std::printf("steps_run: rescaling the particle masses by 1/%.6f\n", ratio);
                for (auto &m : M) m /= ratio;
            }
        }
    }
    // positions outside a periodic box are moved in (step.cc:40-66), here
    // before the initial force rather than after it
    if (o.topology != STEPS_TOPO_R3)
        for (int64_t i = 0; i < N; ++i)
            for (int c = (o.topology == STEPS_TOPO_T3 ? 0 : 2); c < 3; ++c) {
                double &q = x[3 * i + c];
                if (q < 0) q += P.L;
                if (q >= P.L) q -= P.L;
            }

    // ---- softening (utils.cc:59-82) and the force descriptor with its tables
    std::vector<double> soft(N);
    double M_min = 0, rho_part = 0;
    if (steps_softening_lengths(M.data(), N, P.ParticleRadi, soft.data(), &M_min, &rho_part))
        return die(steps_last_error());
    steps_force_desc d;
    steps_force_desc_init(&d);
    d.topology = o.topology;
    d.is_periodic = P.IS_PERIODIC;
    d.cosmology = P.COSMOLOGY;
    d.comoving = P.COMOVING_INTEGRATION;
    d.L = P.L;
    d.Rsim = P.Rsim;
    d.mass_in_unit_sphere = misph;
    d.H0 = P.H0;
    d.Omega_lambda = P.Omega_lambda;
    d.s1r2_lookup = o.lookup;
    d.interp_order = o.interp_order;
    d.precision = o.precision;
    d.device = o.device;
    d.bh_theta = o.bh_theta;             // USE_BH (main.cc:302-303)
    d.bh_particle_radius = P.ParticleRadi;
    std::vector<double> t3, s1r2, radial;
    if (o.topology == STEPS_TOPO_T3 && P.IS_PERIODIC >= 2) {
        int32_t ng = 0;
        if (steps_t3_table(P.IS_PERIODIC, P.L, nullptr, &ng)) return die(steps_last_error());
        t3.resize((size_t)ng * ng * ng * 3);
        if (steps_t3_table(P.IS_PERIODIC, P.L, t3.data(), &ng)) return die(steps_last_error());
        d.t3_table = t3.data();
        d.t3_ngrid = ng;
    }
    if (o.topology == STEPS_TOPO_S1R2) {
        if (P.IS_PERIODIC >= 2 && o.lookup) {
            int32_t nr = 0, nz = 0;
            if (steps_s1r2_table(P.IS_PERIODIC, P.L, P.Rsim, nullptr, &nr, &nz)) return die(steps_last_error());
            s1r2.resize((size_t)nr * nz * 2);
            if (steps_s1r2_table(P.IS_PERIODIC, P.L, P.Rsim, s1r2.data(), &nr, &nz)) return die(steps_last_error());
            d.s1r2_table = s1r2.data();
            d.s1r2_nrho = nr;
            d.s1r2_nz = nz;
        }
        if (comoving && (P.IS_PERIODIC == 1 || !o.lookup)) {  // main.cc:1266-1302
            const double Lz = (P.IS_PERIODIC == 1) ? 0.5 * P.L : P.L * ((double)(P.IS_PERIODIC + 1) - 0.4);
            radial.resize(P.RADIAL_FORCE_TABLE_SIZE);
            if (steps_radial_table(P.Rsim, Lz, P.RADIAL_FORCE_TABLE_SIZE, P.RADIAL_FORCE_ACCURACY, radial.data()))
                return die(steps_last_error());
            d.radial_table = radial.data();
            d.radial_table_size = P.RADIAL_FORCE_TABLE_SIZE;
        }
    }

    d.bh_randomize = o.bh_seed >= 0 ? 1 : 0;  // RANDOMIZE_BH

    // ---- the radial octree correction table (main.cc:791-1091), before the
    // run's own force evaluations, as there
    std::vector<double> radial_bh;
    if (o.bh_theta > 0.0 && P.RADIAL_BH_FORCE_CORRECTION == 1 && o.topology != STEPS_TOPO_T3) {
        std::string glass = P.GLASS_FILE_FOR_BH_FORCE_CORRECTION;
        if (glass != "None") {
            if (std::FILE *g = std::fopen(glass.c_str(), "r")) {
                std::fclose(g);
            } else {
                std::fprintf(stderr, "steps_run: the glass file %s does not exist; using the IC file\n", glass.c_str());
                glass = "None";
            }
        }
        if (glass == "None") glass = P.IC_FILE;
        int64_t Ng = 0;
        std::vector<double> xg, vg, Mg;
        if (P.IC_FORMAT == 0) {
            Ng = steps_ascii_count_particles(glass.c_str());
            if (Ng <= 0) return die(Ng < 0 ? steps_host_last_error() : "empty glass file " + glass);
            xg.resize(3 * Ng), vg.resize(3 * Ng), Mg.resize(Ng);
            if (steps_read_ascii_ic(glass.c_str(), Ng, xg.data(), vg.data(), Mg.data()))
                return die(steps_host_last_error());
        } else {
            double Lg = 0, ag = 0;
            if (steps_read_gadget_ic(glass.c_str(), &Ng, nullptr, nullptr, nullptr, &Lg, &ag))
                return die(steps_host_last_error());
            xg.resize(3 * Ng), vg.resize(3 * Ng), Mg.resize(Ng);
            if (steps_read_gadget_ic(glass.c_str(), &Ng, xg.data(), vg.data(), Mg.data(), &Lg, &ag))
                return die(steps_host_last_error());
        }
        if (Ng != N)  // main.cc:1110-1116
            return die("the radial octree correction's particle count (" + std::to_string(Ng) +
                       ") differs from the initial conditions' (" + std::to_string(N) + ")");
        if (P.H0_INDEPENDENT_UNITS != 0 && cosmo) {  // main.cc:817-830
            for (auto &c : xg) c /= h0d;
            for (auto &m : Mg) m /= h0d;
        }
        std::vector<double> softg(Ng), Fg(3 * Ng);
        double mmin_g = 0, rho_g = 0;
        if (steps_softening_lengths(Mg.data(), Ng, P.ParticleRadi, softg.data(), &mmin_g, &rho_g))
            return die(steps_last_error());
        const int iters = o.bh_seed >= 0 ? P.RADIAL_BH_FORCE_TABLE_ITERATION : 1;  // main.cc:979-983
        const int size = P.RADIAL_BH_FORCE_TABLE_SIZE;
        radial_bh.assign((size_t)size, 0.0);
        std::vector<int32_t> count((size_t)size, 0);
        for (int it = 0; it < iters; ++it) {
            steps_force_desc dg = d;
            if (o.bh_seed >= 0) draw_frame(o.topology, dg.bh_random);
            if (steps_force_eval(&dg, Ng, xg.data(), Mg.data(), softg.data(), Fg.data(), 0, Ng - 1))
                return die(std::string("radial octree correction: ") + steps_last_error());
            if (steps_radial_bh_table_accumulate(o.topology, P.Rsim, size, Ng, xg.data(), Fg.data(), radial_bh.data(),
                                                 count.data()))
                return die(steps_last_error());
        }
        steps_radial_bh_table_finish(size, radial_bh.data(), count.data());
        for (int b = 0; b < size; ++b)
            if (count[(size_t)b] == 0) std::printf("steps_run: no particles in shell %d, its correction is zero\n", b);
        const std::string tname = std::string(P.OUT_DIR) + "radial_bh_force_table.dat";
        if (steps_write_ascii_table(tname.c_str(), 0.5 * P.Rsim / size, P.Rsim, P.Rsim / size, radial_bh.data(), size,
                                    "# Radial BH force correction table\n# r[Mpc]    F_correction[internal units]"))
            return die(steps_host_last_error());
        std::printf("steps_run: radial octree correction table (%d bins, %d evaluations) written to %s\n", size, iters,
                    tname.c_str());
        d.radial_bh_table = radial_bh.data();
        d.radial_bh_table_size = size;
    }

    // ---- time and the output schedule (main.cc:1422-1565)
    double a = cosmo ? P.a_start : 1.0, T = 0.0, H = 0.0, a_max = P.a_max, t_next = 0.0, dT_out = 0.0;
    if (cosmo) T = steps_friedmann_start(&E, P.a_start, P.h_min * 0.05);  // main.cc:1454
    if (comoving) H = steps_hubble(&E, a);
    if (cosmo && !comoving) {
        a = 1.0;
        a_max = P.a_max / STEPS_UNIT_T;  // a_max is the final time in Gy here (main.cc:1551)
    }
    std::vector<double> out_list;
    size_t out_idx = 0;
    const bool have_list = P.OUT_LST[0] != '\0';
    if (have_list) {
        const int64_t n = steps_read_output_list(P.OUT_LST, nullptr, 0);
        if (n < 0) return die(steps_host_last_error());
        out_list.resize(n);
        steps_read_output_list(P.OUT_LST, out_list.data(), n);
        if (P.OUTPUT_TIME_VARIABLE == 1) {  // redshifts, descending, closed by z(a_max) (inputoutput.cc:266-271)
            std::sort(out_list.begin(), out_list.end(), [](double p, double q) { return p > q; });
            out_list.push_back(1.0 / P.a_max - 1.0);
        } else {  // times in Gy -> internal units, ascending
            std::sort(out_list.begin(), out_list.end());
            for (auto &t : out_list) t /= STEPS_UNIT_T;
        }
        if (out_list.empty()) return die("empty output list");
        const double cur = (P.OUTPUT_TIME_VARIABLE == 1) ? 1.0 / a - 1.0 : T;
        while (out_idx < out_list.size() &&
               ((P.OUTPUT_TIME_VARIABLE == 1) ? out_list[out_idx] >= cur : out_list[out_idx] <= cur))
            ++out_idx;
        if (out_idx >= out_list.size()) return die("no valid output time or redshift in the output list");
        t_next = out_list[out_idx];
    } else if (cosmo) {
        if (P.OUTPUT_TIME_VARIABLE == 0) {
            dT_out = P.H_OUT / STEPS_UNIT_T;
            t_next = (P.FIRST_T_OUT >= T * STEPS_UNIT_T) ? P.FIRST_T_OUT / STEPS_UNIT_T : T + dT_out;
        } else {
            dT_out = P.H_OUT;
            t_next = (P.FIRST_T_OUT >= T) ? P.FIRST_T_OUT : P.FIRST_T_OUT - dT_out;
        }
    } else {
        dT_out = P.H_OUT;
        t_next = T + dT_out;
    }

    // ---- the resident simulation, initial force and step length
    steps_sim *sim = nullptr;
    if (steps_sim_create(&d, N, &sim)) return die(steps_last_error());
    auto fail_sim = [&](const char *what) {
        const std::string m = std::string(what) + ": " + steps_last_error();
        steps_sim_destroy(sim);
        return die(m);
    };
    if (steps_sim_set_state(sim, x.data(), v.data(), M.data(), soft.data())) return fail_sim("state");
    auto next_frame = [&]() -> int {  // a randomised octree draws afresh per evaluation
        if (o.bh_seed < 0) return 0;
        double u[6];
        draw_frame(o.topology, u);
        return steps_sim_set_bh_random(sim, u);
    };
    if (next_frame() || steps_sim_forces(sim, 0, N - 1)) return fail_sim("initial force");
    double errmax = 0.0;
    if (steps_sim_errmax(sim, a, H, 1.0, rho_part, &errmax)) return fail_sim("initial step length");
    double h = std::sqrt(2.0 * P.ACC_PARAM / errmax);
    h = std::min(std::max(h, P.h_min), P.h_max);  // main.cc:1652-1667
    const std::string logname = std::string(P.OUT_DIR) + "Logfile.dat";
    std::FILE *log = std::fopen(logname.c_str(), "w");
    if (!log) return fail_sim(("cannot write " + logname).c_str());
    std::fprintf(log, "#step\tt[internal]\ta\tz\tH[internal]\th\terrmax\n");
    auto snapshot = [&](double due) -> int {
        if (steps_sim_get_state(sim, x.data(), v.data())) return -1;
        const std::string name = snapshot_name(P, due);
        std::printf("steps_run: snapshot %s (t = %.8f, a = %.8f)\n", name.c_str(), T, a);
        return steps_write_ascii_snapshot(name.c_str(), N, x.data(), v.data(), M.data(), comoving ? a : 1.0,
                                          (P.H0_INDEPENDENT_UNITS != 0) ? h0d : 1.0);
    };
    const auto t_wall0 = std::chrono::steady_clock::now();
    double a_tmp = cosmo && comoving ? a : 0.0;
    long step = 0;
    for (; a_tmp < a_max; ++step) {
        if (o.max_steps >= 0 && step >= o.max_steps) break;
        T += h;
        // step.cc:128-269: kick (a, H), drift, force, a <- RK4, kick (a', H')
        const double a1 = a, H1 = H;
        if (comoving) {  // step.cc:233-240
            a = steps_friedmann_step(&E, a, h, T);
            H = steps_hubble(&E, a);
        }
        steps_step_params sp{h, a1, H1, a, H, 1.0, rho_part};
        if (next_frame() || steps_sim_step(sim, &sp, &errmax)) {
            std::fclose(log);
            return fail_sim("KDK step");
        }
        a_tmp = comoving ? a : T;
        std::fprintf(log, "%ld\t%.15e\t%.15e\t%.15e\t%.15e\t%.15e\t%.15e\n", step, T, a, 1.0 / a - 1.0, H, h, errmax);
        // outputs (main.cc:1720-1806)
        if (!have_list) {
            if (P.OUTPUT_TIME_VARIABLE == 0) {
                if (T > t_next) {
                    if (snapshot(t_next)) return fail_sim("snapshot");
                    t_next += dT_out;
                }
            } else if (1.0 / a - 1.0 < t_next) {
                if (snapshot(t_next)) return fail_sim("snapshot");
                t_next -= dT_out;
            }
        } else if (P.OUTPUT_TIME_VARIABLE == 1) {
            if (1.0 / a - 1.0 < t_next) {
                if (snapshot(t_next)) return fail_sim("snapshot");
                if (out_idx + 1 < out_list.size()) t_next = out_list[++out_idx];
                else t_next = -1.0;  // past the list
            }
        } else if (T >= t_next) {
            if (snapshot(t_next)) return fail_sim("snapshot");
            if (out_idx + 1 < out_list.size()) t_next = out_list[++out_idx];
            else t_next = 1e300;
        }
        h = std::sqrt(2.0 * P.ACC_PARAM / errmax);  // main.cc:1807-1819
        h = std::min(std::max(h, P.h_min), P.h_max);
        if (h + T > t_next && P.OUTPUT_TIME_VARIABLE == 0) h = t_next - T + 1e-9 * P.h_min;
        const double mins = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_wall0).count() / 60.0;
        if (P.TIME_LIMIT_IN_MINS != 0 && mins >= P.TIME_LIMIT_IN_MINS) {
            std::printf("steps_run: wall-clock limit reached (%.1f min)\n", mins);
            break;
        }
    }
    int rc = 0;
    if (P.OUTPUT_TIME_VARIABLE == 0) rc = snapshot(t_next);  // main.cc:1839-1845
    std::fclose(log);
    std::printf("steps_run: %ld steps, t = %.10f, a = %.10f, z = %.10f\n", step, T, a, 1.0 / a - 1.0);
    steps_sim_destroy(sim);
    return rc ? die(steps_host_last_error()) : 0;
}
