// poremaps_main.cpp -- the `POREMAPS [input.inp]` command line, a drop-in for
// src/main_stokes_solver.cc:57-416 on one GPU (no mpirun).
//
// Reference behaviour kept: input file argv[1] or "input_param.inp"; the
// ordered .inp; output files "velx_"/"vely_"/"velz_"/"press_"/
// "voxel_neighborhood_"/"domain_decomp_" + geometry_file_name verbatim in the
// working directory, written in the order nhood, press, vels, decomposition;
// velocities scaled by the voxel size; the log appended with a header; a
// progress line every 100 iterations.
//
// Build knobs (defaults = reference behaviour), as extra arguments or
// environment variables:
//   --mode=reference|jacobi|sor3        POREMAPS_MODE
//   --precision=f64|f32dc               POREMAPS_PRECISION
//   --omega=<w>                         POREMAPS_OMEGA (0, the fast modes' default: the pair by porosity)
//   --omega-p=<w>                       POREMAPS_OMEGA_P (pressure relaxation; default omega)
//   --device=<n>                        POREMAPS_DEVICE
//   --walls=reference|noslip|slip       POREMAPS_WALLS (noslip / slip: the non-periodic x / y
//                                       faces are no-slip / slip walls; the reference leaves them
//                                       open; slip: fast modes only)
// Deviation: hard errors (missing file, bad .inp, bad geometry, HIP failure)
// exit non-zero instead of continuing.
#include "../../include/poremaps_amd.h"
#include "../../include/poremaps_amd_io.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

namespace {

struct LogCtx {
    const char *log_file;
    bool firstline;
};

int on_log(const pm_log_row *row, void *user) {
    LogCtx *lc = (LogCtx *)user;
    if (pm_write_log_row(lc->log_file, row, lc->firstline) != PM_OK)
        std::fprintf(stderr, "%s\n", pm_io_last_error());
    lc->firstline = false;
    return 0;
}

int on_progress(int it, double conv, double seconds, void *) {
    // src/calc_flow.cc:159-163
    std::printf("Iteration: %i\t Convergence: %e\t", it, conv);
    std::printf("Time: %f\t TPS: %f\n", seconds, 1.0 / (seconds > 0 ? seconds : 1e-12));
    std::fflush(stdout);
    return 0;
}

bool knob(const std::vector<std::string> &args, const char *flag, const char *env, std::string *out) {
    const std::string pre = std::string(flag) + "=";
    for (const auto &a : args)
        if (a.compare(0, pre.size(), pre) == 0) { *out = a.substr(pre.size()); return true; }
    const char *e = std::getenv(env);
    if (e && *e) { *out = e; return true; }
    return false;
}

}  // namespace

int main(int argc, char **argv) {
    std::vector<std::string> flags;
    std::string input_fn;
    for (int a = 1; a < argc; a++) {
        if (std::strncmp(argv[a], "--", 2) == 0) flags.push_back(argv[a]);
        else if (input_fn.empty()) input_fn = argv[a];
        else { std::fprintf(stderr, "unexpected argument %s\n", argv[a]); return 2; }
    }
    std::time_t start = std::time(nullptr);
    std::printf("\nStarting Stokes Solver (POREMAPS, MI355X build) at %s", std::asctime(std::localtime(&start)));
    if (input_fn.empty()) {
        input_fn = "input_param.inp";
        std::printf("No input file specified. Using input_param.inp\n");
    } else {
        std::printf("Input file: %s\n", input_fn.c_str());
    }

    pm_params prm;
    pm_inp_extra extra;
    if (pm_read_inp(input_fn.c_str(), &prm, &extra) != PM_OK) {
        std::fprintf(stderr, "%s\n", pm_io_last_error());
        return 2;
    }
    std::string v;
    if (knob(flags, "--mode", "POREMAPS_MODE", &v)) {
        if (v == "reference") prm.mode = PM_MODE_REFERENCE;
        else if (v == "jacobi") prm.mode = PM_MODE_JACOBI;
        else if (v == "sor3") prm.mode = PM_MODE_SOR3;
        else { std::fprintf(stderr, "unknown mode %s\n", v.c_str()); return 2; }
    }
    if (knob(flags, "--precision", "POREMAPS_PRECISION", &v)) {
        if (v == "f64") prm.precision = PM_PREC_F64;
        else if (v == "f32dc") prm.precision = PM_PREC_F32DC;
        else { std::fprintf(stderr, "unknown precision %s\n", v.c_str()); return 2; }
    }
    // fast modes without --omega: the relaxation pair chosen from the porosity
    if (prm.mode != PM_MODE_REFERENCE) prm.omega = 0.0;
    if (knob(flags, "--omega", "POREMAPS_OMEGA", &v)) prm.omega = std::atof(v.c_str());
    if (knob(flags, "--omega-p", "POREMAPS_OMEGA_P", &v)) prm.omega_p = std::atof(v.c_str());
    if (knob(flags, "--device", "POREMAPS_DEVICE", &v)) prm.device = std::atoi(v.c_str());
    if (knob(flags, "--walls", "POREMAPS_WALLS", &v)) {
        if (v == "reference") prm.walls = 0;
        else if (v == "noslip") prm.walls = 1;
        else if (v == "slip") prm.walls = 2;
        else { std::fprintf(stderr, "unknown walls %s\n", v.c_str()); return 2; }
    }

    // dom_decomposition P Q R in the reference mode: the iterates of
    // `mpirun -np P*Q*R`, each block swept against its halo copies of the
    // other blocks as the sweep started; 0 0 0 (MPI_Dims_create over one
    // rank) is one domain.  The fast modes ignore it.
    const int *dd = extra.dom_decomposition;
    const bool blocks = prm.mode == PM_MODE_REFERENCE && dd[0] > 0 && dd[1] > 0 && dd[2] > 0 &&
                        dd[0] * dd[1] * dd[2] > 1;
    if (blocks)
        for (int a = 0; a < 3; a++) prm.dom_decomposition[a] = dd[a];
    std::printf("Given domain decomposition nx = %i, ny = %i, nz = %i (%s)\n", dd[0], dd[1], dd[2],
                blocks ? "its blocks reproduced on one GPU" : "one domain on one GPU");
    std::printf("Domain size: domx = %i, domy = %i, domz = %i; boundary_method %i; solver %i\n", prm.nx, prm.ny,
                prm.nz, prm.boundary_method, prm.solving_algorithm);

    const int64_t n = (int64_t)prm.nx * prm.ny * prm.nz;
    std::vector<uint8_t> raw((size_t)n);
    std::printf("Reading input geometry\n");
    if (pm_read_raw_u8(extra.geometry_file, raw.data(), n) != PM_OK) {
        std::fprintf(stderr, "%s\n", pm_io_last_error());
        return 2;
    }
    pm_ctx *ctx = nullptr;
    if (pm_ctx_create(&prm, &ctx) != PM_OK) {
        std::fprintf(stderr, "%s\n", pm_last_error());
        return 2;
    }
    std::printf("Evaluate fluid voxel neighborhood\n");
    if (pm_load_geometry(ctx, raw.data()) != PM_OK) {
        std::fprintf(stderr, "%s\n", pm_last_error());
        pm_destroy(ctx);
        return 2;
    }
    std::printf("START MAIN COMPUTATION\n");
    std::fflush(stdout);
    LogCtx lc{extra.log_file, true};
    pm_result res;
    int rc = pm_solve(ctx, on_log, on_progress, &lc, &res);
    int exit_code = 0;
    if (rc != PM_OK) {
        std::fprintf(stderr, "%s\n", pm_last_error());
        exit_code = rc == PM_E_DIVERGED ? 3 : 2;
        if (rc != PM_E_DIVERGED) { pm_destroy(ctx); return exit_code; }
    }

    const std::string geom = extra.geometry_file;
    std::printf("Writing output to file\n");
    const int *wo = extra.write_output;
    std::vector<double> f((size_t)n);
    bool ok = true;
    if (wo[2] == 1) {
        std::printf("Writing neighborhood to file\n");
        std::vector<int32_t> code((size_t)n);
        ok = ok && pm_get_field(ctx, PM_FIELD_NHOOD, code.data()) == PM_OK &&
             pm_write_raw_i32(("voxel_neighborhood_" + geom).c_str(), code.data(), n) == PM_OK;
    }
    if (ok && wo[1] == 1) {
        std::printf("Writing press to file\n");
        ok = pm_get_field(ctx, PM_FIELD_PRESS, f.data()) == PM_OK &&
             pm_write_raw_f64(("press_" + geom).c_str(), f.data(), n, prm.voxel_size, 0) == PM_OK;
    }
    if (ok && wo[0] == 1) {
        std::printf("Writing vels to file\n");
        const char *pre[3] = {"velx_", "vely_", "velz_"};
        for (int q = 0; q < 3 && ok; q++)
            ok = pm_get_field(ctx, PM_FIELD_VELX + q, f.data()) == PM_OK &&
                 pm_write_raw_f64((pre[q] + geom).c_str(), f.data(), n, prm.voxel_size, 1) == PM_OK;
    }
    if (ok && wo[3] == 1) {
        std::printf("Writing domain decomposition to file\n");
        // the rank of every voxel, as the run with these blocks writes it
        std::vector<int32_t> rank((size_t)n, 0);
        const int ones[3] = {1, 1, 1};
        ok = pm_rank_map(prm.nx, prm.ny, prm.nz, blocks ? dd : ones, rank.data()) == PM_OK &&
             pm_write_raw_i32(("domain_decomp_" + geom).c_str(), rank.data(), n) == PM_OK;
    }
    if (!ok) {
        std::fprintf(stderr, "%s %s\n", pm_last_error(), pm_io_last_error());
        exit_code = 2;
    }
    std::printf("Solve: %d sweeps, %.3f s, %.3f GLUP/s, k33(whole) %e\n", res.sweeps, res.seconds, res.glups,
                res.last.wk33);
    pm_destroy(ctx);
    return exit_code;
}
