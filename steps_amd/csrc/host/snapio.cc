// snapio.cc -- initial conditions and snapshots in the formats StePS reads
// and writes without HDF5: ASCII (inputoutput.cc:53-62, 100-179, 826-909),
// GADGET-2 format 1 (read_gadget_ic.cc:80-413) and the output list
// (inputoutput.cc:181-312).
#include "host_common.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

namespace {
thread_local std::string g_host_error;

struct File {
    std::FILE *f;
    explicit File(std::FILE *ff) : f(ff) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

// GADGET-2 snapshot header: 256 bytes (the layout of GADGET's io_header).
#pragma pack(push, 1)
struct GadgetHeader {
    int32_t npart[6];
    double mass[6];
    double time;
    double redshift;
    int32_t flag_sfr, flag_feedback;
    int32_t npartTotal[6];
    int32_t flag_cooling, num_files;
    double BoxSize, Omega0, OmegaLambda, HubbleParam;
    char fill[256 - 6 * 4 - 6 * 8 - 2 * 8 - 2 * 4 - 6 * 4 - 2 * 4 - 4 * 8];
};
#pragma pack(pop)
static_assert(sizeof(GadgetHeader) == 256, "GADGET header is 256 bytes");

// One Fortran-style record: 4-byte length, payload, 4-byte length.
bool read_record(std::FILE *f, std::vector<char> &buf) {
    uint32_t n0 = 0, n1 = 0;
    if (std::fread(&n0, 4, 1, f) != 1) return false;
    buf.resize(n0);
    if (n0 && std::fread(buf.data(), 1, n0, f) != n0) return false;
    if (std::fread(&n1, 4, 1, f) != 1) return false;
    return n0 == n1;
}
}  // namespace

int host_fail(int code, const std::string &msg) {
    g_host_error = msg;
    return code;
}

extern "C" {

const char *steps_host_last_error(void) { return g_host_error.c_str(); }

int64_t steps_ascii_count_particles(const char *path) {
    if (!path) return host_fail(STEPS_ERR_ARG, "null path");
    File f(std::fopen(path, "rb"));
    if (!f.f) return host_fail(STEPS_ERR_ARG, std::string("cannot open ") + path);
    int64_t n = 0;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f.f)) > 0) n += std::count(buf, buf + got, '\n');
    return n;
}

int steps_read_ascii_ic(const char *path, int64_t N, double *x, double *v, double *M) {
    if (!path || N < 0 || (N > 0 && (!x || !v || !M))) return host_fail(STEPS_ERR_ARG, "bad arguments");
    File f(std::fopen(path, "r"));
    if (!f.f) return host_fail(STEPS_ERR_ARG, std::string("cannot open ") + path);
    for (int64_t i = 0; i < N; ++i) {
        double r[7];
        for (int c = 0; c < 7; ++c)
            if (std::fscanf(f.f, "%lf", &r[c]) != 1)
                return host_fail(STEPS_ERR_ARG, std::string("IC file ended or malformed at particle ") +
                                                    std::to_string(i) + " in " + path);
        for (int c = 0; c < 3; ++c) {
            x[3 * i + c] = r[c];
            v[3 * i + c] = r[3 + c];
        }
        M[i] = r[6];
    }
    return STEPS_OK;
}

int steps_read_gadget_ic(const char *path, int64_t *N, double *x, double *v, double *M, double *L, double *a) {
    if (!path || !N) return host_fail(STEPS_ERR_ARG, "bad arguments");
    File f(std::fopen(path, "rb"));
    if (!f.f) return host_fail(STEPS_ERR_ARG, std::string("cannot open ") + path);
    std::vector<char> rec;
    if (!read_record(f.f, rec) || rec.size() != sizeof(GadgetHeader))
        return host_fail(STEPS_ERR_ARG, "not a GADGET-2 format-1 file (header record)");
    GadgetHeader h;
    std::memcpy(&h, rec.data(), sizeof h);
    int64_t ntot = 0, nmass = 0, n1 = h.npart[1], first1 = 0;
    for (int k = 0; k < 6; ++k) {
        if (h.npart[k] < 0) return host_fail(STEPS_ERR_ARG, "negative particle count in GADGET header");
        if (k < 1) first1 += h.npart[k];
        ntot += h.npart[k];
        if (h.mass[k] == 0) nmass += h.npart[k];
    }
    if (!(h.HubbleParam > 0)) return host_fail(STEPS_ERR_ARG, "GADGET header: HubbleParam must be positive");
    *N = n1;
    // read_gadget_ic.cc:88: BoxSize in kpc/h -> Mpc
    if (L) *L = h.BoxSize / 1000.0 / h.HubbleParam;
    if (a) *a = h.time;
    if (!x) return STEPS_OK;
    if (!v || !M) return host_fail(STEPS_ERR_ARG, "bad arguments");
    std::vector<char> pos, vel, ids, mass;
    if (!read_record(f.f, pos) || pos.size() != (size_t)ntot * 12 || !read_record(f.f, vel) ||
        vel.size() != (size_t)ntot * 12 || !read_record(f.f, ids) || ids.size() != (size_t)ntot * 4)
        return host_fail(STEPS_ERR_ARG, "GADGET file: position, velocity or ID block has the wrong size");
    if (nmass > 0 && (!read_record(f.f, mass) || mass.size() != (size_t)nmass * 4))
        return host_fail(STEPS_ERR_ARG, "GADGET file: mass block has the wrong size");
    const float *P = reinterpret_cast<const float *>(pos.data());
    const float *V = reinterpret_cast<const float *>(vel.data());
    const int32_t *I = reinterpret_cast<const int32_t *>(ids.data());
    const float *Mv = reinterpret_cast<const float *>(mass.data());
    // the variable-mass block holds the particles of the types with mass 0 in type order
    int64_t moff = 0;
    for (int k = 0; k < 1; ++k)
        if (h.mass[k] == 0) moff += h.npart[k];
    // type-1 particles in ID order (read_gadget_ic.cc reorders by ID, then keeps type 1)
    std::vector<int64_t> ord(n1);
    std::iota(ord.begin(), ord.end(), first1);
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t p, int64_t q) { return I[p] < I[q]; });
    for (int64_t i = 0; i < n1; ++i) {
        const int64_t s = ord[i];
        for (int c = 0; c < 3; ++c) {
            x[3 * i + c] = (double)P[3 * s + c] / 1000.0 / h.HubbleParam;  // kpc/h -> Mpc
            v[3 * i + c] = (double)V[3 * s + c];                            // km/s
        }
        M[i] = (h.mass[1] == 0) ? (double)Mv[moff + (s - first1)] : h.mass[1];
    }
    return STEPS_OK;
}

int64_t steps_read_output_list(const char *path, double *out, int64_t cap) {
    if (!path) return host_fail(STEPS_ERR_ARG, "null path");
    File f(std::fopen(path, "r"));
    if (!f.f) return host_fail(STEPS_ERR_ARG, std::string("the output list ") + path + " does not exist");
    int64_t n = 0;
    double val;
    while (std::fscanf(f.f, "%lf", &val) == 1) {
        if (out && n < cap) out[n] = val;
        ++n;
    }
    return n;
}

int steps_write_ascii_snapshot(const char *path, int64_t N, const double *x, const double *v, const double *M,
                               double a, double h0_dimless) {
    if (!path || N < 0 || (N > 0 && (!x || !v || !M))) return host_fail(STEPS_ERR_ARG, "bad arguments");
    File f(std::fopen(path, "w"));
    if (!f.f) return host_fail(STEPS_ERR_ARG, std::string("cannot write ") + path);
    const double vs = std::sqrt(a) * STEPS_UNIT_V;  // km/s * sqrt(a) (inputoutput.cc:901)
    for (int64_t i = 0; i < N; ++i) {
        for (int c = 0; c < 3; ++c) std::fprintf(f.f, "%.16f\t", x[3 * i + c] * h0_dimless);
        for (int c = 0; c < 3; ++c) std::fprintf(f.f, "%.16f\t", v[3 * i + c] * vs);
        std::fprintf(f.f, "%.16f\t\n", M[i] * h0_dimless);
    }
    if (std::ferror(f.f)) return host_fail(STEPS_ERR_ARG, std::string("write error on ") + path);
    return STEPS_OK;
}

// inputoutput.cc:1038-1061
int steps_write_ascii_table(const char *path, double x_min, double x_max, double deltax, const double *values,
                            int32_t n, const char *header) {
    if (!path || n < 0 || (n > 0 && !values) || !header) return host_fail(STEPS_ERR_ARG, "bad arguments");
    File f(std::fopen(path, "w"));
    if (!f.f) return host_fail(STEPS_ERR_ARG, std::string("cannot write ") + path);
    std::fprintf(f.f, "%s\n", header);
    double arg = x_min;
    for (int32_t i = 0; i < n; ++i) {
        if (arg > x_max) break;
        std::fprintf(f.f, "%.6f\t%.15f\n", arg, values[i]);
        arg += deltax;
    }
    if (std::ferror(f.f)) return host_fail(STEPS_ERR_ARG, std::string("write error on ") + path);
    return STEPS_OK;
}

}  // extern "C"
