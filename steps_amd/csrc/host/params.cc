// params.cc -- StePS parameter files (read_paramfile.cc:83-735).
//
// read_param() tests every line against each key with strstr and, on a
// match, reads the line's second whitespace-separated token with sscanf
// ("%*s\t%lf" and the like).  The keys and their targets are kept here as a
// table; a line is offered to every entry in the reference's order, so a line
// that contains two keys sets both, as there.  The post-read conversions and
// clamps follow read_paramfile.cc:186-193, 240-247, 316-333, 613-727.
#include "host_common.h"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace {

enum Kind { KDBL, KINT, KUINT, KSTR };

struct Key {
    const char *name;
    Kind kind;
    size_t off;
    bool s1r2_only;  // RADIAL_FORCE_* exist only in the PERIODIC_Z build
};
// The USE_BH build's keys are read whatever the build flags here: they act
// only when steps_run evaluates the octree.

#define F(field) offsetof(steps_run_params, field)
// read_paramfile.cc:91-123, in the order the reference tests them
const Key kKeys[] = {
    {"Omega_b", KDBL, F(Omega_b), false},
    {"Omega_lambda", KDBL, F(Omega_lambda), false},
    {"Omega_m", KDBL, F(Omega_m), false},
    {"Omega_r", KDBL, F(Omega_r), false},
    {"HubbleConstant", KDBL, F(H0), false},
    {"IC_FILE", KSTR, F(IC_FILE), false},
    {"L_BOX", KDBL, F(L), false},
    {"IS_PERIODIC", KINT, F(IS_PERIODIC), false},
    {"OUTPUT_FORMAT", KINT, F(OUTPUT_FORMAT), false},
    {"OUT_DIR", KSTR, F(OUT_DIR), false},
    {"a_max", KDBL, F(a_max), false},
    {"SNAPSHOT_START_NUMBER", KUINT, F(N_snapshot), false},
    {"ACC_PARAM", KDBL, F(ACC_PARAM), false},
    {"STEP_MIN", KDBL, F(h_min), false},
    {"COSMOLOGY", KINT, F(COSMOLOGY), false},
    {"a_start", KDBL, F(a_start), false},
    {"STEP_MAX", KDBL, F(h_max), false},
    {"COMOVING_INTEGRATION", KINT, F(COMOVING_INTEGRATION), false},
    {"TIME_LIMIT_IN_MIN", KDBL, F(TIME_LIMIT_IN_MINS), false},
    {"FIRST_T_OUT", KDBL, F(FIRST_T_OUT), false},
    {"IC_FORMAT", KINT, F(IC_FORMAT), false},
    {"OUTPUT_TIME_VARIABLE", KINT, F(OUTPUT_TIME_VARIABLE), false},
    {"OUT_LST", KSTR, F(OUT_LST), false},
    {"H_OUT", KDBL, F(H_OUT), false},
    {"PARTICLE_RADII", KDBL, F(ParticleRadi), false},
    {"MIN_REDSHIFT", KDBL, F(MIN_REDSHIFT), false},
    {"REDSHIFT_CONE", KINT, F(REDSHIFT_CONE), false},
    {"H_INDEPENDENT_UNITS", KINT, F(H0_INDEPENDENT_UNITS), false},
    {"R_SIM", KDBL, F(Rsim), false},
    {"RADIAL_FORCE_ACCURACY", KINT, F(RADIAL_FORCE_ACCURACY), true},
    {"RADIAL_FORCE_TABLE_SIZE", KINT, F(RADIAL_FORCE_TABLE_SIZE), true},
    {"RADIAL_BH_FORCE_CORRECTION", KINT, F(RADIAL_BH_FORCE_CORRECTION), false},
    {"GLASS_FILE_FOR_BH_FORCE_CORRECTION", KSTR, F(GLASS_FILE_FOR_BH_FORCE_CORRECTION), false},
    {"RADIAL_BH_FORCE_TABLE_SIZE", KINT, F(RADIAL_BH_FORCE_TABLE_SIZE), false},
    {"RADIAL_BH_FORCE_TABLE_ITERATION", KINT, F(RADIAL_BH_FORCE_TABLE_ITERATION), false},
    // the COSMOPARAM builds' keys (read_paramfile.cc:115-121), read here whatever
    // the expansion history steps_run is given
    {"w0", KDBL, F(w0), false},
    {"wa", KDBL, F(wa), false},
    {"EXPANSION_FILE", KSTR, F(EXPANSION_FILE), false},
    {"INTERPOLATION_ORDER", KINT, F(INTERPOLATION_ORDER), false},
};
#undef F

// The value of a matching line: its second whitespace-separated token.
// Returns false when there is none (sscanf would leave the target alone).
bool second_token(const char *line, char *tok, size_t cap) {
    char first[STEPS_PATH_MAX], second[STEPS_PATH_MAX];
    if (std::sscanf(line, "%1023s %1023s", first, second) != 2) return false;
    std::snprintf(tok, cap, "%s", second);
    return true;
}

// One key's assignment from a matching line, with its clamps.
void assign(const Key &k, const char *line, steps_run_params *p) {
    char tok[STEPS_PATH_MAX];
    if (!second_token(line, tok, sizeof tok)) return;
    char *base = reinterpret_cast<char *>(p) + k.off;
    switch (k.kind) {
    case KDBL: {
        double v;
        if (std::sscanf(tok, "%lf", &v) == 1) std::memcpy(base, &v, sizeof v);
        break;
    }
    case KINT: {
        int v;
        if (std::sscanf(tok, "%i", &v) == 1) {
            int32_t w = v;
            std::memcpy(base, &w, sizeof w);
        }
        break;
    }
    case KUINT: {
        unsigned v;
        if (std::sscanf(tok, "%u", &v) == 1) {
            uint32_t w = v;
            std::memcpy(base, &w, sizeof w);
        }
        break;
    }
    case KSTR:
        std::snprintf(base, STEPS_PATH_MAX, "%s", tok);
        break;
    }
}

}  // namespace

extern "C" int steps_read_params(const char *path, int32_t topology, steps_run_params *p) {
    if (!path || !p) return host_fail(STEPS_ERR_ARG, "null argument");
    if (topology != STEPS_TOPO_R3 && topology != STEPS_TOPO_T3 && topology != STEPS_TOPO_S1R2)
        return host_fail(STEPS_ERR_ARG, "unknown topology");
    std::FILE *f = std::fopen(path, "r");
    if (!f) return host_fail(STEPS_ERR_ARG, std::string("cannot open parameter file ") + path);
    std::memset(p, 0, sizeof *p);
    const bool s1r2 = topology == STEPS_TOPO_S1R2;
    if (s1r2) p->RADIAL_FORCE_ACCURACY = p->RADIAL_FORCE_TABLE_SIZE = 1000;  // read_paramfile.cc:116-117
    // read_paramfile.cc:136-141
    std::snprintf(p->GLASS_FILE_FOR_BH_FORCE_CORRECTION, sizeof p->GLASS_FILE_FOR_BH_FORCE_CORRECTION, "None");
    p->RADIAL_BH_FORCE_TABLE_SIZE = 64;
    p->RADIAL_BH_FORCE_TABLE_ITERATION = 4;
    char line[1024];
    while (std::fgets(line, sizeof line, f)) {
        for (const Key &k : kKeys) {
            if (k.s1r2_only && !s1r2) continue;
            if (!std::strstr(line, k.name)) continue;
            assign(k, line, p);
            // clamps applied as each key is read (read_paramfile.cc:186-193, 240-247, 316-333)
            if (!std::strcmp(k.name, "HubbleConstant")) p->H0 = p->H0 / STEPS_UNIT_V;
            if (!std::strcmp(k.name, "IS_PERIODIC") && !s1r2 && p->IS_PERIODIC > 4) p->IS_PERIODIC = 4;
            if (!std::strcmp(k.name, "COMOVING_INTEGRATION"))
                p->COMOVING_INTEGRATION = p->COMOVING_INTEGRATION > 1 ? 1 : (p->COMOVING_INTEGRATION < 0 ? 0 : p->COMOVING_INTEGRATION);
            if (!std::strcmp(k.name, "RADIAL_FORCE_ACCURACY") && p->RADIAL_FORCE_ACCURACY < 10) p->RADIAL_FORCE_ACCURACY = 10;
            if (!std::strcmp(k.name, "RADIAL_FORCE_TABLE_SIZE") && p->RADIAL_FORCE_TABLE_SIZE < 5)
                p->RADIAL_FORCE_TABLE_SIZE = 5;
            // read_paramfile.cc:351-378
            if (!std::strcmp(k.name, "RADIAL_BH_FORCE_CORRECTION") && p->RADIAL_BH_FORCE_CORRECTION != 0 &&
                p->RADIAL_BH_FORCE_CORRECTION != 1)
                p->RADIAL_BH_FORCE_CORRECTION = 0;
            if (!std::strcmp(k.name, "RADIAL_BH_FORCE_TABLE_SIZE") && p->RADIAL_BH_FORCE_TABLE_SIZE < 5)
                p->RADIAL_BH_FORCE_TABLE_SIZE = 5;
            if (!std::strcmp(k.name, "RADIAL_BH_FORCE_TABLE_ITERATION") && p->RADIAL_BH_FORCE_TABLE_ITERATION < 1)
                p->RADIAL_BH_FORCE_TABLE_ITERATION = 1;
        }
    }
    std::fclose(f);
    // read_paramfile.cc:627-631
    if (p->INTERPOLATION_ORDER < 1 || p->INTERPOLATION_ORDER > 3) p->INTERPOLATION_ORDER = 1;
    // OUT_DIR ends with a '/' (read_paramfile.cc:205-210 appends one when the
    // value lacks it)
    const size_t n = std::strlen(p->OUT_DIR);
    if (n > 0 && p->OUT_DIR[n - 1] != '/' && n + 1 < sizeof p->OUT_DIR) {
        p->OUT_DIR[n] = '/';
        p->OUT_DIR[n + 1] = '\0';
    }
    // OUTPUT_FORMAT other than ASCII needs HDF5 (read_paramfile.cc:196-201)
    if (p->OUTPUT_FORMAT != 0) p->OUTPUT_FORMAT = 0;
    if (p->COSMOLOGY == 1) {
        p->h_min /= STEPS_UNIT_T;  // Gy -> internal units (read_paramfile.cc:634-635)
        p->h_max /= STEPS_UNIT_T;
        if (p->H0_INDEPENDENT_UNITS == 1) {  // Mpc/h -> Mpc (:719-723)
            const double hd = p->H0 * STEPS_UNIT_V / 100.0;
            p->L /= hd;
            p->ParticleRadi /= hd;
        }
    } else {
        p->H0_INDEPENDENT_UNITS = 0;
    }
    return STEPS_OK;
}
