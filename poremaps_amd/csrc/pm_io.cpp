// pm_io.cpp -- the reference's file formats: the ordered .inp parameter file
// (src/read_problem_params.cc:50-156), interior raw dumps
// (src/output.cc:53-190) and the appended log (src/output.cc:193-230).
#include "../../include/poremaps_amd.h"
#include "../../include/poremaps_amd_io.h"

#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_io_err;

int io_fail(int code, const char *fmt, ...) {
    char buf[600];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_io_err = buf;
    return code;
}

// Whitespace-separated tokens of the whole file (fscanf's "%d"/"%s"/"%lf"
// conversions and the literal keys all skip leading whitespace).
bool tokenize(const char *path, std::vector<std::string> &tok) {
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    std::string cur;
    int ch;
    while ((ch = std::fgetc(f)) != EOF) {
        if (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\v' || ch == '\f') {
            if (!cur.empty()) { tok.push_back(cur); cur.clear(); }
        } else {
            cur.push_back((char)ch);
        }
    }
    if (!cur.empty()) tok.push_back(cur);
    std::fclose(f);
    return true;
}

bool to_int(const std::string &s, int *v) {
    char *end = nullptr;
    errno = 0;
    long x = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || *end != '\0' || errno || x < INT_MIN || x > INT_MAX) return false;
    *v = (int)x;
    return true;
}

bool to_double(const std::string &s, double *v) {
    char *end = nullptr;
    errno = 0;
    double x = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end != '\0') return false;
    *v = x;
    return true;
}

}  // namespace

extern "C" const char *pm_io_last_error(void) { return g_io_err.c_str(); }

// The 14 ordered lines of src/read_problem_params.cc:75-90.  Deviation from
// fscanf: a missing or misspelt key is an error instead of leaving the
// remaining parameters uninitialised.
extern "C" int pm_read_inp(const char *path, pm_params *p, pm_inp_extra *x) {
    if (!path || !p || !x) return io_fail(PM_E_ARG, "null argument");
    std::vector<std::string> t;
    if (!tokenize(path, t)) return io_fail(PM_E_ARG, "Input file doesnt exist: %s", path);
    pm_params_default(p);
    std::memset(x, 0, sizeof *x);
    size_t pos = 0;
    auto key = [&](const char *k) -> bool {
        if (pos >= t.size() || t[pos] != k) return false;
        pos++;
        return true;
    };
    auto ints = [&](int *v, int n) -> bool {
        for (int q = 0; q < n; q++)
            if (pos >= t.size() || !to_int(t[pos++], v + q)) return false;
        return true;
    };
    auto dbl = [&](double *v) -> bool { return pos < t.size() && to_double(t[pos++], v); };
    auto str = [&](char *s, size_t cap) -> bool {
        if (pos >= t.size() || t[pos].size() + 1 > cap) return false;
        std::memcpy(s, t[pos].c_str(), t[pos].size() + 1);
        pos++;
        return true;
    };
#define NEED(cond, what)                                                                   \
    do {                                                                                   \
        if (!(cond)) return io_fail(PM_E_ARG, "%s: cannot read '%s' (token %zu)", path, what, pos); \
    } while (0)
    NEED(key("dom_decomposition") && ints(x->dom_decomposition, 3), "dom_decomposition");
    NEED(key("boundary_method") && ints(&p->boundary_method, 1), "boundary_method");
    NEED(key("geometry_file_name") && str(x->geometry_file, sizeof x->geometry_file), "geometry_file_name");
    int sz[3];
    NEED(key("size_x_y_z") && ints(sz, 3), "size_x_y_z");
    p->nx = sz[0]; p->ny = sz[1]; p->nz = sz[2];
    NEED(key("voxel_size") && dbl(&p->voxel_size), "voxel_size");
    NEED(key("max_iter") && ints(&p->max_iter, 1), "max_iter");
    NEED(key("it_eval") && ints(&p->it_eval, 1), "it_eval");
    NEED(key("it_write") && ints(&p->it_write, 1), "it_write");
    NEED(key("log_file_name") && str(x->log_file, sizeof x->log_file), "log_file_name");
    NEED(key("solving_algorithm") && ints(&p->solving_algorithm, 1), "solving_algorithm");
    NEED(key("eps") && dbl(&p->eps), "eps");
    NEED(key("porosity") && dbl(&p->porosity), "porosity");
    NEED(key("dom_interest") && ints(p->dom_interest, 6), "dom_interest");
    NEED(key("write_output") && ints(x->write_output, 4), "write_output");
#undef NEED
    // src/read_problem_params.cc:96-104, :144-146: warnings only
    if (p->eps <= 0) std::fprintf(stderr, "Setting EPS to a value smaller equal than zero.");
    if (p->it_eval > 0 && p->it_write % p->it_eval)
        std::fprintf(stderr, "Set writing and evaluating freq to proper values (it_write mod it_eval != 0).");
    if (p->porosity > 1.0 || (p->porosity < -0.0 && p->porosity != -1.0))
        std::fprintf(stderr, "Porosity has to be 0.0 < porosity <= 1.0 or -1.0 to be evaluated.");
    const int *d = x->dom_decomposition;
    if (!(d[0] == 0 && d[1] == 0 && d[2] == 0) && !(d[0] != 0 && d[1] != 0 && d[2] != 0))
        std::fprintf(stderr, "Check domain decomposition.");
    return PM_OK;
}

// src/output.cc:53-104 / :106-153: interior values, x fastest, native byte
// order, written from offset 0 of a file opened read-write and created if
// missing (MPI_MODE_RDWR | MPI_MODE_CREATE: an existing longer file keeps
// its tail).  Velocities are scaled by the voxel size by the caller.
static int write_at_zero(const char *path, const void *data, size_t bytes) {
    int fd = ::open(path, O_RDWR | O_CREAT, 0644);
    if (fd < 0) return io_fail(PM_E_ARG, "cannot open %s: %s", path, std::strerror(errno));
    const char *b = (const char *)data;
    size_t done = 0;
    while (done < bytes) {
        ssize_t w = ::pwrite(fd, b + done, bytes - done, (off_t)done);
        if (w <= 0) {
            ::close(fd);
            return io_fail(PM_E_ARG, "write to %s failed: %s", path, std::strerror(errno));
        }
        done += (size_t)w;
    }
    ::close(fd);
    return PM_OK;
}

extern "C" int pm_write_raw_f64(const char *path, const double *v, int64_t n, double scale, int rescale) {
    if (!path || (!v && n) || n < 0) return io_fail(PM_E_ARG, "null argument or negative size");
    if (!rescale) return write_at_zero(path, v, sizeof(double) * (size_t)n);
    std::vector<double> t((size_t)n);
    for (int64_t q = 0; q < n; q++) t[(size_t)q] = v[q] * scale;
    return write_at_zero(path, t.data(), sizeof(double) * (size_t)n);
}

extern "C" int pm_write_raw_i32(const char *path, const int32_t *v, int64_t n) {
    if (!path || (!v && n) || n < 0) return io_fail(PM_E_ARG, "null argument or negative size");
    return write_at_zero(path, v, sizeof(int32_t) * (size_t)n);
}

// src/output.cc:193-230
extern "C" int pm_write_log_row(const char *path, const pm_log_row *r, int firstline) {
    if (!path || !r) return io_fail(PM_E_ARG, "null argument");
    FILE *f = std::fopen(path, "a");
    if (!f) return io_fail(PM_E_ARG, "cannot open log %s: %s", path, std::strerror(errno));
    if (firstline)
        std::fprintf(f, "# iteration, conv, TPS [1/s], wmax_velz [m/s], wmean_velz [m/s], k13 [m^2], "
                        "k23 [m^2], k33 [m^2], wk13 [m^2], wk23 [m^2], wk33 [m^2]\n");
    std::fprintf(f, "%i, %e, %e, %e, %e, %e, %e, %e, %e, %e, %e\n", r->iteration, r->conv, r->tps,
                 r->wmax_velz, r->wmean_velz, r->k13, r->k23, r->k33, r->wk13, r->wk23, r->wk33);
    std::fclose(f);
    return PM_OK;
}

extern "C" int pm_read_raw_u8(const char *path, uint8_t *out, int64_t n) {
    if (!path || !out || n < 0) return io_fail(PM_E_ARG, "null argument or negative size");
    FILE *f = std::fopen(path, "rb");
    if (!f) return io_fail(PM_E_ARG, "cannot open geometry %s: %s", path, std::strerror(errno));
    size_t got = std::fread(out, 1, (size_t)n, f);
    std::fclose(f);
    if ((int64_t)got != n)
        return io_fail(PM_E_GEOMETRY, "geometry %s holds %zu bytes, size_x_y_z needs %lld", path, got, (long long)n);
    return PM_OK;
}

// src/output.cc:156-190 (write_domain_decomposition) with the blocks of
// src/geometry.cc:129-171 (get_dom_limits: along each axis the first
// size % dims blocks hold one cell more) and the rank of each block in the
// Cartesian communicator (src/main_stokes_solver.cc:182-193; MPICH numbers
// the blocks with z fastest, then y, then x).
extern "C" int pm_rank_map(int nx, int ny, int nz, const int dims_in[3], int32_t *out) {
    if (!dims_in || !out || nx < 1 || ny < 1 || nz < 1) return io_fail(PM_E_ARG, "pm_rank_map: bad argument");
    const int size[3] = {nx, ny, nz};
    int dims[3];
    for (int a = 0; a < 3; a++) {
        dims[a] = dims_in[a] < 1 ? 1 : dims_in[a];
        if (dims[a] > size[a])
            return io_fail(PM_E_ARG, "pm_rank_map: %d blocks along an axis of %d cells", dims[a], size[a]);
    }
    // block coordinate of every cell along each axis
    std::vector<int> coord[3];
    for (int a = 0; a < 3; a++) {
        const int base = size[a] / dims[a], extra = size[a] - dims[a] * base;
        coord[a].resize((size_t)size[a]);
        for (int c = 0, x = 0; c < dims[a]; c++)
            for (int m = 0; m < base + (c < extra ? 1 : 0); m++) coord[a][(size_t)x++] = c;
    }
    for (int i = 0; i < nz; i++)
        for (int j = 0; j < ny; j++) {
            int32_t *row = out + ((int64_t)i * ny + j) * nx;
            const int yz = coord[1][(size_t)j] * dims[2] + coord[2][(size_t)i];
            for (int k = 0; k < nx; k++) row[k] = (int32_t)(coord[0][(size_t)k] * dims[1] * dims[2] + yz);
        }
    return PM_OK;
}
