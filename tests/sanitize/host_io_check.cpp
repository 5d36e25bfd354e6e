// Host-side checks of the file-format layer (poremaps_amd/csrc/pm_io.cpp),
// built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_host_sanitizers.py: the ordered .inp parser on the reference's
// own input files and on malformed variants (truncated, misspelt keys,
// overlong names, non-numbers, out-of-range integers, empty), the raw writers
// and reader, the log writer and the rank map.  Exit status 0 when every check
// holds; a sanitizer report ends the process with its own nonzero status.
//
//   host_io_check <golden dir> <scratch dir>
#include "../../include/poremaps_amd.h"
#include "../../include/poremaps_amd_io.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

// The parser starts from the solver library's defaults; the GPU library is
// not linked here, so its fields start from zero (the parser's own output is
// what is checked).
extern "C" void pm_params_default(pm_params *p) { std::memset(p, 0, sizeof *p); }

static int failures = 0;
#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            failures++;                                                      \
        }                                                                    \
    } while (0)

static std::string slurp(const std::string &path) {
    std::string s;
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return s;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
    std::fclose(f);
    return s;
}

static void spit(const std::string &path, const std::string &s) {
    FILE *f = std::fopen(path.c_str(), "wb");
    std::fwrite(s.data(), 1, s.size(), f);
    std::fclose(f);
}

static std::string replace(std::string s, const std::string &a, const std::string &b) {
    const size_t at = s.find(a);
    if (at != std::string::npos) s.replace(at, a.size(), b);
    return s;
}

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    const std::string gold = argv[1], tmp = argv[2];
    pm_params p;
    pm_inp_extra x;

    // the reference's own input files
    CHECK(pm_read_inp((gold + "/input_rect_channel_24_14_50_vs_5e-05.inp").c_str(), &p, &x) == PM_OK);
    CHECK(p.nx == 24 && p.ny == 14 && p.nz == 50 && p.solving_algorithm == 2 && p.eps == 1e-7);
    CHECK(std::strcmp(x.geometry_file, "rect_channel_24_14_50_vs_5e-05.raw") == 0);
    CHECK(pm_read_inp((gold + "/input_template.inp").c_str(), &p, &x) == PM_OK);
    CHECK(x.dom_decomposition[0] == 1 && x.dom_decomposition[1] == 2 && x.dom_decomposition[2] == 4);

    // malformed variants: every one is an error with a message, none reads
    // past its tokens
    const std::string good = slurp(gold + "/input_rect_channel_24_14_50_vs_5e-05.inp");
    CHECK(!good.empty());
    std::vector<std::string> bad = {
        "",
        good.substr(0, good.size() / 2),
        good.substr(0, good.find("write_output") + 16),
        replace(good, "voxel_size", "voxel_sizes"),
        replace(good, "it_eval    100", "it_eval    1e2"),
        replace(good, "size_x_y_z  24 14 50", "size_x_y_z  24 14"),
        replace(good, "it_write   100", "it_write   99999999999999999999"),
        replace(good, "rect_channel_24_14_50_vs_5e-05.raw", std::string(5000, 'g')),
        replace(good, "eps 1e-07", "eps x"),
        replace(good, "max_iter    100000001", "max_iter    3000000000"),
    };
    for (size_t i = 0; i < bad.size(); i++) {
        const std::string f = tmp + "/bad" + std::to_string(i) + ".inp";
        spit(f, bad[i]);
        const int rc = pm_read_inp(f.c_str(), &p, &x);
        CHECK(rc == PM_E_ARG);
        CHECK(pm_io_last_error()[0] != '\0');
    }
    // no trailing newline after the last token is still a whole file
    {
        std::string s = good;
        while (!s.empty() && (s.back() == '\n' || s.back() == ' ')) s.pop_back();
        spit(tmp + "/nonl.inp", s);
        CHECK(pm_read_inp((tmp + "/nonl.inp").c_str(), &p, &x) == PM_OK && x.write_output[3] == 1);
    }
    CHECK(pm_read_inp((tmp + "/missing.inp").c_str(), &p, &x) == PM_E_ARG);
    CHECK(pm_read_inp(nullptr, &p, &x) == PM_E_ARG);

    // raw files: written at offset 0, an existing longer file keeps its tail
    {
        std::vector<double> v(37);
        for (size_t i = 0; i < v.size(); i++) v[i] = 0.5 * (double)i;
        const std::string f = tmp + "/velz.raw";
        spit(f, std::string(400, '\x7f'));
        CHECK(pm_write_raw_f64(f.c_str(), v.data(), (int64_t)v.size(), 2.0, 1) == PM_OK);
        const std::string s = slurp(f);
        CHECK(s.size() == 400);
        double back[37];
        std::memcpy(back, s.data(), sizeof back);
        for (int i = 0; i < 37; i++) CHECK(back[i] == (double)i);
        CHECK(s[399] == '\x7f');
        CHECK(pm_write_raw_f64(f.c_str(), v.data(), -1, 1.0, 1) == PM_E_ARG);
        std::vector<int32_t> r(11, 7);
        CHECK(pm_write_raw_i32((tmp + "/r.raw").c_str(), r.data(), (int64_t)r.size()) == PM_OK);
        CHECK(pm_write_raw_i32((tmp + "/r.raw").c_str(), r.data(), -3) == PM_E_ARG);
        uint8_t g[64];
        spit(tmp + "/g.raw", std::string(40, '\1'));
        CHECK(pm_read_raw_u8((tmp + "/g.raw").c_str(), g, 40) == PM_OK && g[39] == 1);
        CHECK(pm_read_raw_u8((tmp + "/g.raw").c_str(), g, 64) == PM_E_GEOMETRY);
        CHECK(pm_read_raw_u8((tmp + "/g.raw").c_str(), g, -1) == PM_E_ARG);
    }

    // log rows
    {
        pm_log_row row;
        std::memset(&row, 0, sizeof row);
        row.iteration = 1100;
        row.wk33 = 8.368958e-09;
        const std::string f = tmp + "/perm.log";
        CHECK(pm_write_log_row(f.c_str(), &row, 1) == PM_OK);
        CHECK(pm_write_log_row(f.c_str(), &row, 0) == PM_OK);
        const std::string s = slurp(f);
        CHECK(s.rfind("# iteration", 0) == 0 && s.find("1100, 0.000000e+00") != std::string::npos);
    }

    // rank map against a direct count, uneven and degenerate splits
    {
        const int shapes[][3] = {{10, 8, 18}, {5, 3, 2}, {1, 1, 1}, {7, 9, 4}};
        const int dims[][3] = {{3, 2, 2}, {1, 1, 9}, {0, 0, 0}, {2, 3, 4}, {5, 1, 1}};
        for (const auto &sh : shapes)
            for (const auto &d : dims) {
                const int nx = sh[0], ny = sh[1], nz = sh[2];
                std::vector<int32_t> m((size_t)nx * ny * nz, -1);
                const int rc = pm_rank_map(nx, ny, nz, d, m.data());
                const bool fits = d[0] <= nx && d[1] <= ny && d[2] <= nz;
                CHECK(rc == (fits ? PM_OK : PM_E_ARG));
                if (rc != PM_OK) continue;
                int D[3], sz[3] = {nx, ny, nz};
                for (int a = 0; a < 3; a++) D[a] = d[a] < 1 ? 1 : d[a];
                auto blk = [&](int a, int c) {
                    const int base = sz[a] / D[a], extra = sz[a] % D[a];
                    for (int b = 0, lo = 0; b < D[a]; b++) {
                        const int n = base + (b < extra ? 1 : 0);
                        if (c < lo + n) return b;
                        lo += n;
                    }
                    return -1;
                };
                for (int i = 0; i < nz; i++)
                    for (int j = 0; j < ny; j++)
                        for (int k = 0; k < nx; k++) {
                            const int want = (blk(0, k) * D[1] + blk(1, j)) * D[2] + blk(2, i);
                            CHECK(m[((size_t)i * ny + j) * nx + k] == want);
                        }
            }
        CHECK(pm_rank_map(0, 1, 1, dims[0], nullptr) == PM_E_ARG);
    }

    if (failures) std::fprintf(stderr, "%d check(s) failed\n", failures);
    else std::printf("host io checks: ok\n");
    return failures ? 1 : 0;
}
