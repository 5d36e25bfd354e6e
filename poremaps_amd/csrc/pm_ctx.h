// pm_ctx.h -- the solver context (one geometry on one GPU), shared by the
// single-solve C ABI (pm_solver.hip) and the batch engine (pm_batch.hip).
#pragma once

#include "../../include/poremaps_amd.h"
#include "pm_kernels.h"

#include <vector>

// Event pairs around kernel groups (pm_time_kernels); inactive when null.
struct Timer {
    struct Rec { int cat; hipEvent_t a, b; };
    std::vector<Rec> recs;
    int open_cat = -1;
    hipEvent_t open_ev = nullptr;
};

enum TimerCat { TC_SWEEP = 0, TC_GHOST, TC_FOLD, TC_RESIDUAL, TC_REF, TC_FIRST0, TC_N };
// JI_RECORDS: the interleaved state's lagged pass reading per-voxel records
// (codes of the cell and its recompute targets, gathered residuals).
// JI_MARCH: the interleaved state's lagged pass as a z-march of tiles through
// LDS (pm_march.hip).
// JI_COMPACT: the lagged pass on the compact state (fluid cells only, linked
// entries; FluidLists::compact_jac).
enum JacobiImpl { JI_PAIR = 0, JI_LIST = 1, JI_STREAM = 2, JI_RECORDS = 3, JI_MARCH = 4, JI_COMPACT = 5 };
static const char *kTimerNames[TC_N] = {"sweep", "ghost_fill", "dc_fold", "dc_residual", "reference_sweep",
                                        "dc_first_pass"};

// One Jacobi kernel configuration.  Every configuration produces the same
// iterates bit for bit, so the choice only changes the speed.
struct SweepVariant {
    int impl;           // JacobiImpl
    bool select_loads;  // fluid-list kernels: all neighbours + select, or case-predicated loads
    int zs;             // dense kernel / z-march: planes per z segment
    int ty = 4;         // z-march: tile rows (4 or 8)
    int occ = 0;        // FP32+DC record pass: occupancy target (0, 6, 8 waves per SIMD)
    bool skip2 = false; // FP32+DC record pass: +-2 loads only where the case reads them
    bool skip1 = false; // with skip2: every other stencil load too, only where used
};

// Autotuner: the first sweeps of a solve are run with each candidate in turn
// (windows of `window` sweeps timed with events on the solver's stream), then
// the fastest is kept for the rest of the context's life.
struct Tuner {
    static constexpr int kMax = 12;
    bool active = false, done = false;
    int n = 0, idx = 0, in_window = 0, window = 6, first_scored = 0;
    SweepVariant cand[kMax];
    hipEvent_t ev0[kMax] = {}, ev1[kMax] = {};
    float ms[kMax] = {};
    int best = -1;
};

struct pm_ctx {
    Timer *timer = nullptr;
    pm_params prm;
    pm::Grid g;
    hipStream_t stream = nullptr;
    bool loaded = false;
    // geometry
    uint8_t *G = nullptr, *pc = nullptr;
    int32_t *code = nullptr;
    // FP64 state (reference mode, FP64 fast modes, defect-correction master)
    double *X[4] = {nullptr, nullptr, nullptr, nullptr};   // p, u, v, w
    double *X2[4] = {nullptr, nullptr, nullptr, nullptr};  // Jacobi second buffer
    double *qX[2] = {nullptr, nullptr};
    // FP32 correction state of the defect-correction modes, interleaved
    // ({p, u, v, w} per cell): two sets for Jacobi's out-of-place sweeps, one
    // for SOR's in-place sweeps; its outlet buffers and the FP32 residual
    float4 *S4[2] = {nullptr, nullptr};
    float *qD[2] = {nullptr, nullptr};
    pm::Res4 R4 = {nullptr, nullptr, nullptr};
    // reductions
    pm::FluidLists lists;   // fluid cells of the interior (fast modes)
    void *red_scratch = nullptr;
    pm::RedSums *red_dev = nullptr;
    pm::RedSums *red_host = nullptr;
    // iteration state
    int t = 0;            // sweeps done
    double kprev = 0.0;   // permeability of the previous convergence test
    double phi = 0.0;
    int64_t fluid = 0;
    // exact-order sweep: LDS-tiled wavefront (default) or the single-workgroup
    // / per-hyperplane forms (POREMAPS_REF=tiles|onewg|planes)
    int ref_impl = 0;
    // fluid-list column (rows x cells) for the next pm_load_geometry (0: whole
    // extent).  On the 512^3 field 64 x 64 is best: interleaved FP32 lagged
    // pass 0.96 ms against 1.05 ms with whole rows and 1.43 ms in plain
    // (z, y, x) order (flat from 16 x 32 to 128 x 128); FP32 record pass
    // 0.860 ms (128 x 64: 0.862, 32 x 128: 0.879, 64 x whole rows: 0.929);
    // FP64 record pass 1.60 ms against 1.64 ms with whole rows.  FP32+DC SOR
    // uses whole rows (pm_ctx_create)
    // pm_params.omega was 0: the relaxation pair is chosen at each geometry
    // load from the porosity
    bool auto_relax = false;
    // reference mode with a domain decomposition: the blocks (device table),
    // the most tiles of any block per axis, the state as each sweep starts
    // and the outlet plane as the last sweep left it before its pins
    pm::RefBox *ref_tab = nullptr;
    int ref_nblk = 0;
    int ref_ntile[3] = {0, 0, 0};
    double *ref_snap[4] = {nullptr, nullptr, nullptr, nullptr};
    double *ref_outlet = nullptr;
    // the FP32 correction state is zero everywhere (after a fold or at the
    // start): the next first pass needs only the records and the residual
    bool dzero = false;
    int prm_tile_y = 64;
    int prm_tile_x = 64;
    // Jacobi kernels: velocity + pressure pair over the fluid list, one lagged
    // pass over the fluid list, or one lagged dense 2.5-D pass (z segments of
    // stream_zs planes)
    int jac_impl = JI_PAIR;
    int stream_zs = 64;
    // z-march form: tile rows (4 or 8) and planes per z segment (0: all)
    int march_ty = 4, march_zseg = 128;
    int sor_chunk = 0;          // FluidLists::sor_chunk (POREMAPS_SOR_CHUNK)
    int rec_occ = 0;            // FluidLists::rec_occ (POREMAPS_REC_OCC)
    bool rec_skip2 = false;     // FluidLists::rec_skip2 (POREMAPS_REC_SKIP2=1)
    bool rec_skip1 = false;     // FluidLists::rec_skip1 (POREMAPS_REC_SKIP2=2: both)
    bool sor_skip2 = true;      // FluidLists::sor_skip2 (POREMAPS_SOR_SKIP2=0: off)
    bool sor_fused = true;      // FluidLists::sor_fused (POREMAPS_SOR_STAGE=seq: false)
    bool sor_compact = true;    // FluidLists::compact (POREMAPS_SOR=split / natural: off)
    bool sor_packed = true;     // FluidLists::packed (POREMAPS_SOR_LINKS=full: off)
    bool jac_packed = true;     // FluidLists::packed for the compact Jacobi (POREMAPS_JAC_LINKS=full / aos: off)
    int pack_span = 4095;       // launch_pack_links' span (POREMAPS_SOR_PACK_SPAN)
    int pack_far = -1;          // entries left on the full links (-1: links not packed)
    bool jac_compact = true;    // FluidLists::compact_jac (POREMAPS_JACOBI naming a dense form: off)
    float layout_ms[2] = {0.f, 0.f};  // FP32+DC Jacobi: ms per sweep, compact / dense layout (0: not timed)
    bool jac_links_aos = false; // FluidLists::links_aos (POREMAPS_JAC_LINKS=aos)
    Tuner tuner;
    // the state holds (u^n, p^{n-1}) of a lagged Jacobi sequence
    bool lagged = false;
    // restarts after a divergence (pm_restart_slower)
    int restarts = 0;

    bool fast() const { return prm.mode != PM_MODE_REFERENCE; }
    bool dc() const { return fast() && prm.precision == PM_PREC_F32DC; }
    bool jacobi() const { return prm.mode == PM_MODE_JACOBI; }
};


// Context internals used by the batch engine (defined in pm_solver.hip).
int pm_fail(int code, const char *fmt, ...);
int pm_one_sweep(pm_ctx *c);
int pm_settle(pm_ctx *c);
int pm_restart_slower(pm_ctx *c);
int pm_reduce_state(pm_ctx *c, bool full, pm::RedSums *out);
double pm_convergence_from(const pm::RedSums &r, double *kprev);
void pm_permeability_from(const pm::RedSums &r, const int *dom, double dx, double phi, double *out);
double pm_used_porosity(const pm_ctx *c);
