// pm_solver.hip -- host side of the solver: context, device memory, the
// iteration loop of src/calc_flow.cc:53-169 and the C ABI of poremaps_amd.h.
#include "../../include/poremaps_amd.h"
#include "pm_ctx.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(PM_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),      \
                        __FILE__, __LINE__);                                                  \
    } while (0)

using clk = std::chrono::steady_clock;

__global__ void k_fill_f32(float *a, int64_t n, float v) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

}  // namespace

extern "C" const char *pm_last_error(void) { return g_err.c_str(); }

extern "C" void pm_params_default(pm_params *p) {
    std::memset(p, 0, sizeof *p);
    p->boundary_method = 0;
    p->voxel_size = 1.0;
    p->max_iter = 100000000;
    p->it_eval = 100;
    p->it_write = 100;
    p->solving_algorithm = 2;
    p->eps = 1e-6;
    p->porosity = -1.0;
    p->mode = PM_MODE_REFERENCE;
    p->precision = PM_PREC_F64;
    p->omega = 1.0;
    p->device = 0;
    p->omega_p = 0.0;
}

// The decomposition of the reference mode (x, y, z blocks; 1 1 1 = one
// domain): dom_decomposition, or 1 1 z_slabs.
static void ref_dims(const pm_params *p, int d[3]) {
    for (int a = 0; a < 3; a++) d[a] = std::max(1, p->dom_decomposition[a]);
    if (d[0] == 1 && d[1] == 1 && d[2] == 1 && p->z_slabs > 1) d[2] = p->z_slabs;
}
static bool ref_decomposed(const pm_params *p) {
    int d[3];
    ref_dims(p, d);
    return p->mode == PM_MODE_REFERENCE && d[0] * d[1] * d[2] > 1;
}

// porosity below which the pair by porosity (pm_params.omega = 0) takes the
// large pressure relaxation
constexpr double kAutoTightPhi = 0.20;

static int check_params(const pm_params *p) {
    if (!p) return fail(PM_E_ARG, "null parameters");
    if (p->nx < 2 || p->ny < 2 || p->nz < 2)
        return fail(PM_E_ARG, "size_x_y_z must be >= 2 on every axis (got %d %d %d)", p->nx, p->ny, p->nz);
    if ((int64_t)p->nx * p->ny * p->nz > (int64_t)1 << 31)
        return fail(PM_E_ARG, "domain of %d x %d x %d exceeds 2^31 voxels", p->nx, p->ny, p->nz);
    if (pm::make_grid(p->nx, p->ny, p->nz, 0).total >= ((int64_t)1 << 31))
        return fail(PM_E_ARG, "padded domain of %d x %d x %d exceeds 32-bit cell offsets", p->nx, p->ny, p->nz);
    if (p->boundary_method < 0 || p->boundary_method > 4)
        return fail(PM_E_ARG, "boundary_method must be 0..4 (got %d)", p->boundary_method);
    if (p->it_eval <= 0 || p->it_write <= 0)
        return fail(PM_E_ARG, "it_eval and it_write must be positive");
    if (p->mode < PM_MODE_REFERENCE || p->mode > PM_MODE_SOR3)
        return fail(PM_E_ARG, "unknown mode %d", p->mode);
    if (p->precision != PM_PREC_F64 && p->precision != PM_PREC_F32DC)
        return fail(PM_E_ARG, "unknown precision %d", p->precision);
    if (p->mode == PM_MODE_REFERENCE && p->precision != PM_PREC_F64)
        return fail(PM_E_ARG, "the reference ordering runs in FP64 only");
    if (p->mode != PM_MODE_REFERENCE && !(p->omega >= 0.0 && p->omega < 100.0))
        return fail(PM_E_ARG, "omega must be in (0, 100), or 0 for the pair by porosity (got %g)", p->omega);
    if (!(p->omega_p >= 0.0 && p->omega_p < 100.0))
        return fail(PM_E_ARG, "omega_p must be in [0, 100) (got %g)", p->omega_p);
    const int *dd = p->dom_decomposition;
    if (p->z_slabs < 0 || p->z_slabs > p->nz)
        return fail(PM_E_ARG, "z_slabs must be 0 .. nz (got %d with nz %d)", p->z_slabs, p->nz);
    if (dd[0] < 0 || dd[1] < 0 || dd[2] < 0 || dd[0] > p->nx || dd[1] > p->ny || dd[2] > p->nz)
        return fail(PM_E_ARG, "dom_decomposition %d %d %d must lie in 0 .. size_x_y_z (%d %d %d)", dd[0], dd[1],
                    dd[2], p->nx, p->ny, p->nz);
    if (p->z_slabs > 1 && (dd[0] > 1 || dd[1] > 1 || dd[2] > 1))
        return fail(PM_E_ARG, "give either z_slabs or dom_decomposition");
    if (p->walls < 0 || p->walls > 2) return fail(PM_E_ARG, "walls must be 0, 1 or 2 (got %d)", p->walls);
    if (p->walls == 2 && p->mode == PM_MODE_REFERENCE)
        return fail(PM_E_ARG, "slip walls (walls 2) run in the fast modes (mode jacobi or sor3)");
    int d[3];
    ref_dims(p, d);
    if ((int64_t)d[0] * d[1] * d[2] > PM_MAX_BLOCKS)
        return fail(PM_E_ARG, "dom_decomposition %d %d %d has more than %d blocks", d[0], d[1], d[2],
                    PM_MAX_BLOCKS);
    return PM_OK;
}

extern "C" int64_t pm_device_bytes(const pm_params *p) {
    if (check_params(p) != PM_OK) return -1;
    pm::Grid g = pm::make_grid(p->nx, p->ny, p->nz, p->boundary_method);
    int64_t per = 1 + 1 + 4 + 32;  // G, pc, code, X
    const bool fast = p->mode != PM_MODE_REFERENCE;
    const bool dc = fast && p->precision == PM_PREC_F32DC;
    const bool jac = p->mode == PM_MODE_JACOBI;
    if (fast && !dc && jac) per += 32;
    // interleaved correction sets (16 B each: two for Jacobi, one for SOR),
    // continuity residual (4 B) and momentum residual (16 B per fluid cell);
    // Jacobi: the gathered residuals (16 B per fluid cell)
    if (dc) per += (jac ? 32 : 16) + 4 + 16 + (jac ? 16 : 0);
    // Jacobi and FP32+DC SOR: per-voxel records (8 B per fluid cell)
    if (fast && (jac || dc)) per += 8;
    // FP32+DC on the compact state: the links (24 B per entry) and the outlet
    // buffer(s) by entry (4 B each; two for Jacobi); the packed links (SOR:
    // 12 B per entry and 72 B per 64; Jacobi: 24 B and 144 B per 64)
    if (dc) per += jac ? 32 + 27 : 28 + 14;
    // FP32+DC Jacobi: the z-march form's row-segment starts (4 B per 64 cells)
    if (dc && jac) per += 1;
    // fluid lists (upper bound: every cell fluid); SOR on per-field arrays
    // keeps separate colour lists
    if (fast) per += (jac || dc) ? 4 : 8;
    // reference mode in blocks: the state as a sweep starts, the outlet plane
    // and the block table
    const int64_t blocks = ref_decomposed(p) ? 32 * g.total + 8 * g.qtotal() + (int64_t)sizeof(pm::RefBox) *
                                                   PM_MAX_BLOCKS : 0;
    return per * g.total + blocks + 4 * g.qtotal() * 8 + (int64_t)pm::reduce_scratch_bytes() + (1 << 20);
}

static void free_all(pm_ctx *c) {
    auto f = [](void *p) { if (p) (void)hipFree(p); };
    f(c->G); f(c->pc); f(c->code);
    for (int q = 0; q < 4; q++) { f(c->X[q]); f(c->X2[q]); }
    for (int q = 0; q < 2; q++) { f(c->qX[q]); f(c->qD[q]); f(c->S4[q]); c->S4[q] = nullptr; }
    f(c->R4.p); f(c->R4.m); f(c->R4.q); f(c->R4.pc); c->R4 = pm::Res4{nullptr, nullptr, nullptr, nullptr};
    f(c->red_scratch); f(c->red_dev);
    f(c->lists.all);
    f(c->lists.rec);
    f(c->lists.rowstart);
    f(c->lists.clink); f(c->lists.chalo); f(c->lists.cout); f(c->lists.nat); f(c->lists.natpos);
    f(c->lists.cpack); f(c->lists.cbase);
    for (int q = 0; q < 4; q++) { f(c->ref_snap[q]); c->ref_snap[q] = nullptr; }
    f(c->ref_outlet); f(c->ref_tab);
    c->ref_outlet = nullptr; c->ref_tab = nullptr; c->ref_nblk = 0;
    if (!c->lists.col_in_all)
        for (int q = 0; q < 3; q++) f(c->lists.col[q]);
    c->lists = pm::FluidLists();
    c->pack_far = -1;
    if (c->red_host) (void)hipHostFree(c->red_host);
    c->G = c->pc = nullptr; c->code = nullptr;
    for (int q = 0; q < 4; q++) c->X[q] = c->X2[q] = nullptr;
    c->qX[0] = c->qX[1] = nullptr; c->qD[0] = c->qD[1] = nullptr;
    c->red_scratch = nullptr; c->red_dev = nullptr; c->red_host = nullptr;
}

extern "C" int pm_ctx_create(const pm_params *p, pm_ctx **out) {
    if (!out) return fail(PM_E_ARG, "null output pointer");
    *out = nullptr;
    int rc = check_params(p);
    if (rc != PM_OK) return rc;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (p->device < 0 || p->device >= ndev)
        return fail(PM_E_ARG, "device %d not present (%d devices)", p->device, ndev);
    HIPCHK(hipSetDevice(p->device));
    pm_ctx *c = new (std::nothrow) pm_ctx;
    if (!c) return fail(PM_E_MEMORY, "host allocation failed");
    c->prm = *p;
    c->auto_relax = p->mode != PM_MODE_REFERENCE && p->omega == 0.0;
    c->g = pm::make_grid(p->nx, p->ny, p->nz, p->boundary_method);
    c->g.slip = p->walls == 2 ? 1 : 0;
    // FP32+DC 3-colour SOR on colour-split rows: columns of whole rows, so
    // that a wavefront's runs in each block are as long as the row allows.
    // 512^3 field, ms per sweep: 64 rows x whole rows 1.456, 128 x whole 1.455,
    // 256 x whole 1.467, plain (z, y, x) order 1.496, 64 x 192 1.507, 64 x 64
    // 1.558
    // The compact state (the default; POREMAPS_SOR=split or natural keeps
    // every cell) takes its entries in 64 x 64 columns: 0.745 ms per sweep
    // against 0.762 ms in columns of 64 whole rows.
    if (const char *sc = std::getenv("POREMAPS_SOR_CHUNK")) c->sor_chunk = std::max(0, std::atoi(sc));
    if (const char *so = std::getenv("POREMAPS_SOR"))
        c->sor_compact = std::strcmp(so, "split") != 0 && std::strcmp(so, "natural") != 0;
    if (const char *sl = std::getenv("POREMAPS_SOR_LINKS")) c->sor_packed = std::strcmp(sl, "full") != 0;
    if (const char *sp = std::getenv("POREMAPS_PACK_SPAN")) c->pack_span = std::max(0, std::atoi(sp));
    // slip walls and the z-chunked passes: colour-split rows
    if (p->walls == 2 || c->sor_chunk > 0) c->sor_compact = false;
    if (p->mode == PM_MODE_SOR3 && p->precision == PM_PREC_F32DC && !c->sor_compact) c->prm_tile_x = 0;
    if (const char *ss = std::getenv("POREMAPS_SOR_STAGE")) c->sor_fused = std::strcmp(ss, "seq") != 0;
    // the z-chunked SOR walks the colour lists plane by plane
    if (p->mode == PM_MODE_SOR3 && p->precision == PM_PREC_F32DC && c->sor_chunk > 0) c->prm_tile_y = 0;
    if (const char *ty = std::getenv("POREMAPS_LIST_TILE")) c->prm_tile_y = std::max(0, std::atoi(ty));
    if (const char *tx = std::getenv("POREMAPS_LIST_TILE_X")) c->prm_tile_x = std::max(0, std::atoi(tx));
    if (const char *ri = std::getenv("POREMAPS_REF")) {
        if (std::strcmp(ri, "tiles") == 0) c->ref_impl = pm::REF_TILES;
        else if (std::strcmp(ri, "onewg") == 0) c->ref_impl = pm::REF_ONE_WG;
        else if (std::strcmp(ri, "planes") == 0) c->ref_impl = pm::REF_PLANES;
    }
    if (const char *jm = std::getenv("POREMAPS_JACOBI")) {
        if (std::strcmp(jm, "pair") == 0) c->jac_impl = JI_PAIR;
        else if (std::strcmp(jm, "list") == 0) c->jac_impl = JI_LIST;
        else if (std::strcmp(jm, "stream") == 0) c->jac_impl = JI_STREAM;
        else if (std::strcmp(jm, "records") == 0) c->jac_impl = JI_RECORDS;
        else if (std::strcmp(jm, "march") == 0) c->jac_impl = JI_MARCH;
        // a named dense form keeps the dense layout; "compact" fixes the
        // compact pass (no tuning)
        if (*jm && std::strcmp(jm, "compact") != 0) c->jac_compact = false;
    }
    // FP32+DC Jacobi on the compact state: not with slip walls (their
    // mirrored halos are not copies of images)
    if (p->walls == 2) c->jac_compact = false;
    if (const char *jl = std::getenv("POREMAPS_JAC_LINKS")) {
        c->jac_links_aos = std::strcmp(jl, "aos") == 0;
        c->jac_packed = std::strcmp(jl, "full") != 0 && !c->jac_links_aos;
    }
    if (const char *zs = std::getenv("POREMAPS_STREAM_ZS")) c->stream_zs = std::max(1, std::atoi(zs));
    if (c->g.slip && (c->jac_impl == JI_STREAM || c->jac_impl == JI_MARCH)) c->jac_impl = JI_RECORDS;
    // z-march form (POREMAPS_JACOBI=march): tile rows 4 or 8, planes per z
    // segment (0: all)
    if (const char *mt = std::getenv("POREMAPS_MARCH_TY")) c->march_ty = std::atoi(mt) == 8 ? 8 : 4;
    if (const char *rk = std::getenv("POREMAPS_REC_SKIP2")) {
        c->rec_skip2 = std::atoi(rk) >= 1;
        c->rec_skip1 = std::atoi(rk) == 2 && !c->g.slip;
    }
    if (const char *sk = std::getenv("POREMAPS_SOR_SKIP2")) c->sor_skip2 = std::atoi(sk) != 0;
    if (const char *ro = std::getenv("POREMAPS_REC_OCC")) {
        const int o = std::atoi(ro);
        c->rec_occ = o == 6 || o == 8 ? o : 0;
    }
    if (const char *mz = std::getenv("POREMAPS_MARCH_ZSEG")) c->march_zseg = std::max(0, std::atoi(mz));
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(PM_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    *out = c;
    return PM_OK;
}

extern "C" void pm_destroy(pm_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->prm.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

template <typename T>
static int dalloc(T **p, int64_t n) {
    hipError_t e = hipMalloc((void **)p, sizeof(T) * (size_t)n);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(PM_E_MEMORY, "hipMalloc of %lld bytes: %s", (long long)(sizeof(T) * n),
                    hipGetErrorString(e));
    }
    return PM_OK;
}

#define RC(expr)                      \
    do {                              \
        int rc_ = (expr);             \
        if (rc_ != PM_OK) return rc_; \
    } while (0)

static void tuner_setup(pm_ctx *c);

static int fill_f32(pm_ctx *c, float *a, int64_t n, float v) {
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a, n, v);
    HIPCHK(hipGetLastError());
    return PM_OK;
}

static void tb(pm_ctx *c, int cat) {
    if (!c->timer) return;
    hipEvent_t e;
    (void)hipEventCreate(&e);
    (void)hipEventRecord(e, c->stream);
    c->timer->open_cat = cat;
    c->timer->open_ev = e;
}
static void te(pm_ctx *c) {
    if (!c->timer || c->timer->open_cat < 0) return;
    hipEvent_t e;
    (void)hipEventCreate(&e);
    (void)hipEventRecord(e, c->stream);
    c->timer->recs.push_back({c->timer->open_cat, c->timer->open_ev, e});
    c->timer->open_cat = -1;
}

// Outlet-plane buffer, residual: the start of every defect-correction cycle.
static int dc_prepare(pm_ctx *c) {
    const pm::Grid &g = c->g;
    tb(c, TC_RESIDUAL);
    pm::launch_outlet_plane(g, c->pc, c->X[0], c->X[1], c->X[2], c->X[3], 3.0, c->qX[0], c->stream);
    pm::launch_residual4(g, c->lists, c->pc, c->X[0], c->X[1], c->X[2], c->X[3], c->qX[0], c->R4, c->stream);
    pm::launch_gather_rq(g, c->lists, c->R4, c->stream);
    te(c);
    // the compact SOR state's outlet buffer holds one value per entry
    const int64_t nq = std::max<int64_t>(g.qtotal(), c->lists.n_centries);
    RC(fill_f32(c, c->qD[0], c->lists.compact || c->lists.compact_jac ? nq : g.qtotal(), 0.f));
    RC(fill_f32(c, c->qD[1], c->lists.compact_jac ? nq : g.qtotal(), 0.f));
    HIPCHK(hipGetLastError());
    // every caller has just zeroed the correction state (fold or start)
    c->dzero = true;
    return PM_OK;
}

// The initial state of src/initialize.cc:69-121 (linear pressure, zero
// velocities, pins) and the mode's derived buffers; the geometry is loaded.
static int init_state(pm_ctx *c) {
    const pm::Grid &g = c->g;
    for (int q = 0; q < 4; q++) HIPCHK(hipMemsetAsync(c->X[q], 0, sizeof(double) * (size_t)g.total, c->stream));
    pm::launch_init_fields(g, c->G, c->X[0], c->X[1], c->X[2], c->X[3], c->stream);
    // src/calc_flow.cc:105-108: the pins are applied before sweep 0
    pm::launch_pins<double>(g, c->G, c->X[0], (double)(g.nz + 2), 2.0, 3.0, c->stream);
    if (c->ref_outlet)
        // the pinned initial outlet plane is also its value before the pins
        HIPCHK(hipMemcpyAsync(c->ref_outlet + g.base, c->X[0] + g.base + (int64_t)(g.nz + 1) * g.sy,
                              sizeof(double) * (size_t)g.sy, hipMemcpyDeviceToDevice, c->stream));
    if (c->fast()) {
        pm::launch_outlet_plane(g, c->pc, c->X[0], c->X[1], c->X[2], c->X[3], 3.0, c->qX[0], c->stream);
        HIPCHK(hipMemcpyAsync(c->qX[1], c->qX[0], sizeof(double) * (size_t)g.qtotal(),
                              hipMemcpyDeviceToDevice, c->stream));
    }
    if (c->X2[0])
        for (int q = 0; q < 4; q++)
            HIPCHK(hipMemcpyAsync(c->X2[q], c->X[q], sizeof(double) * (size_t)g.total, hipMemcpyDeviceToDevice,
                                  c->stream));
    if (c->dc()) {
        for (int q = 0; q < 2; q++)
            if (c->S4[q]) RC(fill_f32(c, (float *)c->S4[q], 4 * g.total, 0.f));
        RC(fill_f32(c, c->R4.p, g.total, 0.f));
        RC(fill_f32(c, (float *)c->R4.m, 4 * std::max<int64_t>(1, c->lists.n_all), 0.f));
        RC(dc_prepare(c));
    }
    c->t = 0;
    c->kprev = 0.0;
    c->lagged = false;
    HIPCHK(hipGetLastError());
    return PM_OK;
}

// Start a fast-mode solve again from the initial state with a smaller
// pressure relaxation, after it diverged: the relaxation pair crossed the
// stability bound of the symplectic sweep (omega * omega_p near 19 on every
// medium probed).  Returns PM_E_DIVERGED when the pair is already small, or
// when the medium has no fluid (the convergence value is then 0/0, as in the
// reference, and no relaxation changes that).
int pm_restart_slower(pm_ctx *c) {
    const double wp = c->prm.omega_p > 0.0 ? c->prm.omega_p : c->prm.omega;
    if (!c->fast() || c->fluid == 0 || c->prm.omega * wp <= 8.0) return PM_E_DIVERGED;
    c->prm.omega_p = 0.75 * wp;
    c->restarts++;
    RC(init_state(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PM_OK;
}

// The compact SOR state's tables (FluidLists::compact) from the interior
// geometry raw (x fastest, 0 fluid) and the fluid list `all` (natural
// offsets, colour lists back to back): entry t is list position t; then one
// entry per fluid cell of a periodic face halo, with the image it copies
// (the dense ghost fill's sources and pins); then the zero entry.  Links by
// coordinates, so that nothing beyond the halo box is ever an entry.  The
// records' part of the interior entries is filled on the device from the
// colour passes' records (launch_lag_records).  Falls back to the
// colour-split rows where the medium does not fit the compact form.
static int build_compact_sor(pm_ctx *c, const uint8_t *raw, const std::vector<int32_t> &all) {
    const pm::Grid &g = c->g;
    const int64_t n_all = (int64_t)all.size();
    std::vector<int32_t> cidx((size_t)g.total, -1);
    for (int64_t t = 0; t < n_all; t++) cidx[(size_t)all[(size_t)t]] = (int32_t)t;
    auto fluid_at = [&](int i, int j, int k) {  // halo coordinates of an interior cell
        return raw[((int64_t)(i - 2) * g.ny + (j - 2)) * g.nx + (k - 2)] == 0;
    };
    // the halo cells that get entries: (image entry, natural offset, zero_p,
    // with_q); zero_p: the copy's p is 0; with_q: its outlet buffer value is
    // copied too
    struct HaloCell {
        int32_t img;
        int64_t off;
        bool zero_p, with_q;
    };
    std::vector<HaloCell> cand;
    auto add = [&](int i, int j, int k, int si, int sj, int sk, bool zero_p, bool with_q) {
        if (!fluid_at(si, sj, sk)) return;
        cand.push_back({cidx[(size_t)g.at(si, sj, sk)], g.at(i, j, k), zero_p, with_q});
    };
    const int iout = g.nz + 1;
    if (g.px)
        for (int i = 2; i < g.nz + 2; i++)
            for (int j = 2; j < g.ny + 2; j++) {
                const int src[4] = {g.nx, g.nx + 1, 2, 3}, dst[4] = {0, 1, g.nx + 2, g.nx + 3};
                for (int e = 0; e < 4; e++) add(i, j, dst[e], i, j, src[e], i == iout, i == iout);
            }
    if (g.py)
        for (int i = 2; i < g.nz + 2; i++)
            for (int k = 2; k < g.nx + 2; k++) {
                const int src[4] = {g.ny, g.ny + 1, 2, 3}, dst[4] = {0, 1, g.ny + 2, g.ny + 3};
                for (int e = 0; e < 4; e++) add(i, dst[e], k, i, src[e], k, i == iout, i == iout);
            }
    if (g.pz)
        for (int j = 2; j < g.ny + 2; j++)
            for (int k = 2; k < g.nx + 2; k++) {
                // the copies of the outlet plane and the pinned planes 1 and
                // nz + 2 carry p = 0
                const int src[4] = {g.nz, g.nz + 1, 2, 3}, dst[4] = {0, 1, g.nz + 2, g.nz + 3};
                for (int e = 0; e < 4; e++)
                    add(dst[e], j, k, src[e], j, k, dst[e] == 1 || dst[e] == g.nz + 2, false);
            }
    // halo entries in the order of their images, so that the ghost fill reads
    // and writes both in ascending order
    std::stable_sort(cand.begin(), cand.end(), [](const HaloCell &a, const HaloCell &b) { return a.img < b.img; });
    std::vector<int64_t> hoff;   // natural offsets of the halo entries
    std::vector<int2> halo;      // (entry | bit 31: with_q, image entry | bit 31: p = 0)
    int64_t next = n_all;
    for (const HaloCell &h : cand) {
        cidx[(size_t)h.off] = (int32_t)next;
        hoff.push_back(h.off);
        halo.push_back(make_int2(h.with_q ? (int)((uint32_t)next | 0x80000000u) : (int)next,
                                 h.zero_p ? (int)((uint32_t)h.img | 0x80000000u) : (int)h.img));
        next++;
    }
    const int64_t zero = next, n_ent = next + 1;
    if (!pm::compact_supported(g, n_ent)) return PM_OK;
    // links: (-x, +x), (-y, +y), (-z, +z) arrays of n_ent entries (24 B per entry)
    std::vector<uint2> lk((size_t)(3 * n_ent));
    std::vector<int32_t> outl;
    auto link = [&](int64_t off, int64_t e) {
        const int64_t r = off - g.base;
        const int i = (int)(r / g.sy), j = (int)((r % g.sy) / g.sx), k = (int)(r % g.sx);
        auto at = [&](int ii, int jj, int kk) -> uint32_t {
            if (ii < 0 || ii >= g.NZ || jj < 0 || jj >= g.NY || kk < 0 || kk >= g.NX) return (uint32_t)zero;
            const int32_t m = cidx[(size_t)g.at(ii, jj, kk)];
            return (uint32_t)(m < 0 ? zero : m);
        };
        lk[(size_t)e] = make_uint2(at(i, j, k - 1), at(i, j, k + 1));
        lk[(size_t)(n_ent + e)] = make_uint2(at(i, j - 1, k), at(i, j + 1, k));
        lk[(size_t)(2 * n_ent + e)] = make_uint2(at(i - 1, j, k), at(i + 1, j, k));
        return i;
    };
    for (int64_t t = 0; t < n_all; t++)
        if (link(all[(size_t)t], t) == iout) outl.push_back((int32_t)t);
    for (size_t h = 0; h < hoff.size(); h++) link(hoff[h], n_all + (int64_t)h);
    for (int a = 0; a < 3; a++) lk[(size_t)(a * n_ent + zero)] = make_uint2((uint32_t)zero, (uint32_t)zero);
    RC(dalloc(&c->lists.clink, 3 * n_ent));
    HIPCHK(hipMemcpy(c->lists.clink, lk.data(), sizeof(uint2) * lk.size(), hipMemcpyHostToDevice));
    RC(dalloc(&c->lists.chalo, std::max<int64_t>(1, (int64_t)halo.size())));
    if (!halo.empty())
        HIPCHK(hipMemcpy(c->lists.chalo, halo.data(), sizeof(int2) * halo.size(), hipMemcpyHostToDevice));
    RC(dalloc(&c->lists.cout, std::max<int64_t>(1, (int64_t)outl.size())));
    if (!outl.empty())
        HIPCHK(hipMemcpy(c->lists.cout, outl.data(), sizeof(int32_t) * outl.size(), hipMemcpyHostToDevice));
    // the outlet buffer by entry
    if (n_ent > g.qtotal()) {
        (void)hipFree(c->qD[0]);
        c->qD[0] = nullptr;
        RC(dalloc(&c->qD[0], n_ent));
    }
    c->lists.n_centries = (int)n_ent;
    c->lists.n_chalo = (int)halo.size();
    c->lists.n_cout = (int)outl.size();
    c->lists.compact = true;
    c->lists.links_aos = false;
    return PM_OK;
}

// Use the packed records unless more than a quarter of the entries fit no
// base: such an entry reads its record and then the link arrays, so a layout
// where many do (a medium whose fluid cells of one colour are far apart in the
// neighbouring colours' lists) runs faster on the link arrays alone.  With the
// test setting POREMAPS_PACK_KEEP=1 the records are kept whatever the count.
static int keep_packed(pm_ctx *c, int far) {
    pm::FluidLists &L = c->lists;
    c->pack_far = far;
    const char *keep = std::getenv("POREMAPS_PACK_KEEP");
    const bool always = keep && std::strcmp(keep, "0") != 0;
    if (!always && (int64_t)far * 4 > L.n_all) {
        (void)hipFree(L.cpack);
        (void)hipFree(L.cbase);
        L.cpack = nullptr;
        L.cbase = nullptr;
        L.packed = false;
        return PM_OK;
    }
    L.packed = true;
    return PM_OK;
}

// The compact SOR's links packed to 12 B per entry (launch_pack_links), with
// 18 bases per wave of 64 entries; the 24-byte links stay for the entries
// whose links fit no base (none on the 512^3 Gaussian random field).
static int pack_sor_links(pm_ctx *c) {
    pm::FluidLists &L = c->lists;
    int64_t waves = 0;
    for (int col = 0; col < 3; col++) {
        L.wave_off[col] = (int)waves;
        waves += (L.n_col[col] + 63) / 64;
    }
    RC(dalloc(&L.cpack, 3 * std::max<int64_t>(1, L.n_all)));
    RC(dalloc(&L.cbase, 18 * std::max<int64_t>(1, waves)));
    int *d_cnt = nullptr;  // entries on the full links, records that do not decode
    RC(dalloc(&d_cnt, 2));
    int cnt[2] = {0, 0};
    hipError_t e = hipMemsetAsync(d_cnt, 0, 2 * sizeof(int), c->stream);
    if (e == hipSuccess) {
        pm::launch_pack_links(L, c->pack_span, d_cnt, d_cnt + 1, c->stream);
        e = hipMemcpyAsync(cnt, d_cnt, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_cnt);
    if (e != hipSuccess) return fail(PM_E_HIP, "packed links: %s", hipGetErrorString(e));
    HIPCHK(hipGetLastError());
    if (cnt[1] != 0) return fail(PM_E_HIP, "packed links: %d records do not decode to their links", cnt[1]);
    return keep_packed(c, cnt[0]);
}

// The compact Jacobi's packed records (launch_pack_jac_links): 24 B per entry
// holding the lagged pass's twelve links, 36 bases per wave of 64 entries;
// the link arrays stay for the other passes and the entries that fit no base.
static int pack_jac_links(pm_ctx *c) {
    pm::FluidLists &L = c->lists;
    const int64_t waves = (L.n_all + 63) / 64;
    RC(dalloc(&L.cpack, 6 * std::max<int64_t>(1, L.n_all)));
    RC(dalloc(&L.cbase, 36 * std::max<int64_t>(1, waves)));
    int *d_cnt = nullptr;  // entries on the full links, records that do not decode
    RC(dalloc(&d_cnt, 2));
    int cnt[2] = {0, 0};
    hipError_t e = hipMemsetAsync(d_cnt, 0, 2 * sizeof(int), c->stream);
    if (e == hipSuccess) {
        pm::launch_pack_jac_links(L, std::min(c->pack_span, 1023), d_cnt, d_cnt + 1, c->stream);
        e = hipMemcpyAsync(cnt, d_cnt, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_cnt);
    if (e != hipSuccess) return fail(PM_E_HIP, "packed links: %s", hipGetErrorString(e));
    HIPCHK(hipGetLastError());
    if (cnt[1] != 0) return fail(PM_E_HIP, "packed links: %d records do not decode to their links", cnt[1]);
    return keep_packed(c, cnt[0]);
}

// The compact FP32+DC Jacobi state (FluidLists::compact_jac): entry t is list
// position t, then the zero entry; no halo entries, a link across a periodic
// face goes to the image cell itself.  Links built on the device from the
// uploaded fluid list (launch_torus_links), in three arrays or
// (POREMAPS_JAC_LINKS=aos) one record per entry; the records' flags follow
// in launch_lag_records.
static int build_compact_jacobi(pm_ctx *c) {
    const pm::Grid &g = c->g;
    const int64_t n_all = c->lists.n_all, n_ent = n_all + 1;
    if (!pm::compact_supported(g, n_ent)) return PM_OK;
    const bool aos = c->jac_links_aos;
    RC(dalloc(&c->lists.clink, 3 * n_ent));
    int32_t *cidx = nullptr;
    RC(dalloc(&cidx, g.total));
    hipError_t e = pm::launch_torus_links(g, c->lists.all, (int)n_all, cidx, c->lists.clink, (int)n_ent, aos,
                                          c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(cidx);
    if (e != hipSuccess) return fail(PM_E_HIP, "compact links: %s", hipGetErrorString(e));
    HIPCHK(hipGetLastError());
    // the outlet buffers by entry
    if (n_ent > g.qtotal())
        for (int q = 0; q < 2; q++) {
            (void)hipFree(c->qD[q]);
            c->qD[q] = nullptr;
            RC(dalloc(&c->qD[q], n_ent));
        }
    // the neighbours' continuity residuals come from the entries' residual
    // array through the links, not gathered per cell
    RC(dalloc(&c->R4.pc, n_ent));
    RC(fill_f32(c, c->R4.pc, n_ent, 0.f));
    c->lists.n_centries = (int)n_ent;
    c->lists.compact_jac = true;
    c->lists.links_aos = aos;
    return PM_OK;
}

// The layout of the FP32+DC Jacobi correction state is chosen like a kernel
// form: the compact pass (each load only where used) and the dense record
// pass (its unused +-2 loads skipped where the state allows) each run a few
// sweeps from the zero correction state on the solve's own residual, timed
// with HIP events on the solver's stream (after a warm-up of 3 sweeps); then
// both correction sets and outlet buffers are zeroed again -- the FP64 state is
// never touched -- so the solve starts from exactly the state it would have
// started from, and the faster layout is kept (the other's buffers are
// freed).  On the 512^3 Gaussian field the compact pass is the faster, on
// denser media the dense one can be.  POREMAPS_JACOBI (a fixed form) or
// POREMAPS_TUNE=0 keeps the compact layout untimed.
static int choose_jacobi_layout(pm_ctx *c) {
    if (!c->dc() || !c->jacobi() || !c->lists.compact_jac || c->lists.n_all == 0) return PM_OK;
    const char *fixed = std::getenv("POREMAPS_JACOBI");
    const char *off = std::getenv("POREMAPS_TUNE");
    if ((fixed && *fixed) || (off && std::strcmp(off, "0") == 0)) return PM_OK;
    const pm::Grid &g = c->g;
    hipStream_t s = c->stream;
    const double om = c->prm.omega, wp = c->prm.omega_p > 0.0 ? c->prm.omega_p : om;
    const float fcu = (float)(om * pm::kDt), fcp = (float)(wp * pm::kC2 * pm::kDt);
    const int64_t nq = std::max<int64_t>(g.qtotal(), c->lists.n_centries);
    constexpr int kWarm = 3, kWin = 6;
    float ms[2] = {0.f, 0.f};
    for (int lay = 0; lay < 2; lay++) {
        pm::FluidLists &L = c->lists;
        L.compact_jac = lay == 0;
        L.use_rec = lay == 1;
        L.select_loads = true;
        L.rec_occ = 0;
        L.rec_skip2 = lay == 1 && pm::rec_skip2_supported(g);
        L.rec_skip1 = lay == 0;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        float4 *A = c->S4[0], *B = c->S4[1];
        float *qa = c->qD[0], *qb = c->qD[1];
        for (int it = 0; it < kWarm + kWin; it++) {
            if (it == kWarm) HIPCHK(hipEventRecord(e0, s));
            pm::launch_jacobi4(it == 0 ? pm::JAC_FIRST_ZERO : pm::JAC_LAG, g, L, c->pc, A, qa, B, qb, c->R4, fcu, fcp,
                               s);
            std::swap(A, B);
            std::swap(qa, qb);
        }
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(&ms[lay], e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        for (int q = 0; q < 2; q++) {
            RC(fill_f32(c, (float *)c->S4[q], 4 * g.total, 0.f));
            RC(fill_f32(c, c->qD[q], nq, 0.f));
        }
    }
    c->layout_ms[0] = ms[0] / kWin;
    c->layout_ms[1] = ms[1] / kWin;
    const bool compact = ms[0] <= ms[1];
    pm::FluidLists &L = c->lists;
    L.compact_jac = compact;
    if (compact) {
        (void)hipFree(c->R4.q);
        c->R4.q = nullptr;
        c->jac_impl = JI_COMPACT;
    } else {
        (void)hipFree(L.clink);
        (void)hipFree(c->R4.pc);
        (void)hipFree(L.cpack);
        (void)hipFree(L.cbase);
        L.clink = nullptr;
        c->R4.pc = nullptr;
        L.cpack = nullptr;
        L.cbase = nullptr;
        L.packed = false;
        L.n_centries = 0;
        c->jac_impl = JI_RECORDS;
    }
    L.use_rec = c->jac_impl == JI_RECORDS;
    L.rec_occ = c->rec_occ;
    L.rec_skip2 = c->rec_skip2 && pm::rec_skip2_supported(g);
    L.rec_skip1 = c->rec_skip1;
    HIPCHK(hipStreamSynchronize(s));
    return PM_OK;
}

extern "C" int pm_load_geometry(pm_ctx *c, const uint8_t *raw) {
    if (!c || !raw) return fail(PM_E_ARG, "null argument");
    HIPCHK(hipSetDevice(c->prm.device));
    const pm::Grid &g = c->g;
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    int64_t fluid = 0;
    for (int64_t q = 0; q < n; q++) {
        if (raw[q] > 1)
            return fail(PM_E_GEOMETRY, "geometry byte %lld is %u; only 0 (fluid) and 1 (solid) are valid",
                        (long long)q, (unsigned)raw[q]);
        fluid += raw[q] == 0;
    }
    free_all(c);
    c->loaded = false;
    RC(dalloc(&c->G, g.total));
    RC(dalloc(&c->pc, g.total));
    RC(dalloc(&c->code, g.total));
    for (int q = 0; q < 4; q++) RC(dalloc(&c->X[q], g.total));
    for (int q = 0; q < 2; q++) RC(dalloc(&c->qX[q], g.qtotal()));
    if (ref_decomposed(&c->prm)) {
        int d[3];
        ref_dims(&c->prm, d);
        std::vector<pm::RefBox> blocks;
        pm::ref_blocks(g, d, &blocks);
        c->ref_nblk = (int)blocks.size();
        for (const pm::RefBox &b : blocks)
            for (int a = 0; a < 3; a++) c->ref_ntile[a] = std::max(c->ref_ntile[a], (b.n[a] + 7) / 8);
        for (int q = 0; q < 4; q++) RC(dalloc(&c->ref_snap[q], g.total));
        RC(dalloc(&c->ref_outlet, g.qtotal()));
        RC(dalloc(&c->ref_tab, c->ref_nblk));
        HIPCHK(hipMemcpy(c->ref_tab, blocks.data(), sizeof(pm::RefBox) * blocks.size(), hipMemcpyHostToDevice));
    }
    if (c->fast() && !c->dc() && c->jacobi())
        for (int q = 0; q < 4; q++) RC(dalloc(&c->X2[q], g.total));
    if (c->dc()) {
        // interleaved correction state: Jacobi ping-pongs two sets, SOR
        // updates one in place
        for (int q = 0; q < (c->jacobi() ? 2 : 1); q++) RC(dalloc(&c->S4[q], g.total));
        RC(dalloc(&c->R4.p, g.total));  // the momentum part follows the fluid list
        for (int q = 0; q < 2; q++) RC(dalloc(&c->qD[q], g.qtotal()));
    }
    RC(dalloc((uint8_t **)&c->red_scratch, (int64_t)pm::reduce_scratch_bytes()));
    RC(dalloc(&c->red_dev, 1));
    HIPCHK(hipHostMalloc((void **)&c->red_host, sizeof(pm::RedSums), hipHostMallocDefault));
    if (c->fast()) {
        // fluid lists, all and per colour (i+j+k) mod 3 of the halo-including
        // indices, in columns of tile_y rows x tile_x cells marching along z:
        // (column, z, y, x) order.  Every list kernel treats its entries
        // independently, so the order only changes cache reuse: the +-2-plane
        // neighbours of a cell are touched again within a few column-planes
        // (L2) instead of a few whole planes, and a workgroup's 256 cells
        // cover a compact patch of a plane whose +-y rows it mostly loads
        // itself (L1).  On the 512^3 field the lagged pass takes 1.43 ms in
        // plain (z, y, x) order, 1.05 ms in 64-row slabs of whole rows and
        // 0.96 ms in 64 x 64 columns.  0: whole extent.
        const int ty = c->prm_tile_y > 0 ? std::min(c->prm_tile_y, g.ny) : g.ny;
        const int tx = c->prm_tile_x > 0 ? std::min(c->prm_tile_x, g.nx) : g.nx;
        std::vector<int32_t> all;
        std::vector<int32_t> col[3];
        std::vector<int32_t> nat, natpos;
        // FP32+DC Jacobi with columns 64 cells wide: the list position of
        // every row segment's first fluid cell (the z-march form's index)
        // (one column as wide as the domain is the same for nx <= 64)
        const bool rows = c->dc() && c->jacobi() && (tx == 64 || (tx == g.nx && g.nx <= 64));
        const int ntx = (g.nx + 63) / 64;
        std::vector<int32_t> rowstart(rows ? (size_t)g.nz * g.ny * ntx : 0);
        all.reserve((size_t)fluid);
        for (int q = 0; q < 3; q++) col[q].reserve((size_t)fluid / 3 + 1);
        for (int j0 = 0; j0 < g.ny; j0 += ty)
            for (int k0 = 0; k0 < g.nx; k0 += tx)
            for (int i = 0; i < g.nz; i++)
                for (int j = j0; j < std::min(j0 + ty, g.ny); j++) {
                    const uint8_t *row = raw + ((int64_t)i * g.ny + j) * g.nx;
                    if (rows) rowstart[((size_t)i * g.ny + j) * ntx + k0 / 64] = (int32_t)all.size();
                    for (int k = k0; k < std::min(k0 + tx, g.nx); k++) {
                        if (row[k] != 0) continue;
                        const int32_t e = (int32_t)g.at(i + 2, j + 2, k + 2);
                        all.push_back(e);
                        col[(i + j + k + 6) % 3].push_back(e);
                    }
                }
        c->lists.tile_y = c->prm_tile_y > 0 ? ty : 0;
        c->lists.tile_x = c->prm_tile_x > 0 ? tx : 0;
        const char *hf = std::getenv("POREMAPS_HALOS");
        c->lists.fill_halos = !(hf && std::strcmp(hf, "separate") == 0);
        const char *ld = std::getenv("POREMAPS_SWEEP_LOADS");
        c->lists.select_loads = !(ld && std::strcmp(ld, "branch") == 0);
        c->lists.use_rec = c->jac_impl == JI_RECORDS;
        if (rows) {
            RC(dalloc(&c->lists.rowstart, (int64_t)rowstart.size()));
            HIPCHK(hipMemcpy(c->lists.rowstart, rowstart.data(), sizeof(int32_t) * rowstart.size(),
                             hipMemcpyHostToDevice));
        }
        c->lists.march_ty = c->march_ty;
        c->lists.march_zseg = c->march_zseg;
        c->lists.rec_occ = c->rec_occ;
        c->lists.rec_skip2 = c->rec_skip2 && pm::rec_skip2_supported(g);
        c->lists.rec_skip1 = c->rec_skip1;
        c->lists.sor_skip2 = c->sor_skip2;
        const bool sor = c->prm.mode == PM_MODE_SOR3;
        if (sor && c->dc()) {
            // the colour lists back to back: the residual's momentum part,
            // indexed by list position, then serves every colour pass.  The
            // correction state lies in colour-split rows unless
            // POREMAPS_SOR=natural.
            const char *so = std::getenv("POREMAPS_SOR");
            c->lists.split = !(so && std::strcmp(so, "natural") == 0);
            // the list in column order and each cell's position in the
            // colour lists: the compact state's residual and fold walk it
            nat = all;
            natpos.resize(all.size());
            {
                int64_t cnt[3] = {0, 0, 0};
                const int64_t off[3] = {0, (int64_t)col[0].size(), (int64_t)(col[0].size() + col[1].size())};
                for (size_t u = 0; u < all.size(); u++) {
                    const int64_t r = all[u] - g.base;
                    const int i = (int)(r / g.sy), j = (int)((r % g.sy) / g.sx), k = (int)(r % g.sx);
                    const int q = (i + j + k + 6) % 3;
                    natpos[u] = (int32_t)(off[q] + cnt[q]++);
                }
            }
            all.clear();
            for (int q = 0; q < 3; q++) {
                c->lists.off_col[q] = (int)all.size();
                c->lists.n_col[q] = (int)col[q].size();
                all.insert(all.end(), col[q].begin(), col[q].end());
            }
            if (c->lists.split && c->sor_chunk > 0 && ty == g.ny && tx == g.nx) {
                // plane offsets of the colour lists, which are in (z, y, x)
                // order here
                const int np = g.nz + 1;
                c->lists.plane_off.assign((size_t)3 * np, 0);
                for (int q = 0; q < 3; q++) {
                    int32_t *po = c->lists.plane_off.data() + (size_t)q * np;
                    size_t t = 0;
                    for (int i = 0; i < g.nz; i++) {
                        po[i] = (int32_t)t;
                        while (t < col[q].size() && (col[q][t] - g.base) / g.sy == i + 2) t++;
                    }
                    po[g.nz] = (int32_t)t;
                    if (t != col[q].size()) c->lists.plane_off.clear();  // not in plane order
                    if (c->lists.plane_off.empty()) break;
                }
                c->lists.sor_chunk = c->lists.plane_off.empty() ? 0 : c->sor_chunk;
                c->lists.sor_fused = c->sor_fused;
            }
        }
        c->lists.n_all = (int)all.size();
        RC(dalloc(&c->lists.all, std::max<int64_t>(1, c->lists.n_all)));
        if (c->dc()) RC(dalloc(&c->R4.m, std::max<int64_t>(1, c->lists.n_all)));
        if (c->jacobi() || c->lists.split) RC(dalloc(&c->lists.rec, std::max<int64_t>(1, c->lists.n_all)));
        c->lists.use_march = c->jac_impl == JI_MARCH && pm::march_supported(g, c->lists);
        if (c->dc() && c->jacobi()) RC(dalloc(&c->R4.q, std::max<int64_t>(1, c->lists.n_all)));
        HIPCHK(hipMemcpy(c->lists.all, all.data(), sizeof(int32_t) * all.size(), hipMemcpyHostToDevice));
        if (sor && c->dc()) {
            c->lists.col_in_all = true;
            for (int q = 0; q < 3; q++) c->lists.col[q] = c->lists.all + c->lists.off_col[q];
            if (c->lists.split && c->sor_compact && c->lists.sor_chunk == 0) RC(build_compact_sor(c, raw, all));
            if (c->lists.compact) {
                RC(dalloc(&c->lists.nat, std::max<int64_t>(1, (int64_t)nat.size())));
                RC(dalloc(&c->lists.natpos, std::max<int64_t>(1, (int64_t)natpos.size())));
                HIPCHK(hipMemcpy(c->lists.nat, nat.data(), sizeof(int32_t) * nat.size(), hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(c->lists.natpos, natpos.data(), sizeof(int32_t) * natpos.size(),
                                 hipMemcpyHostToDevice));
            }
        } else if (c->dc() && c->jacobi() && c->jac_compact) {
            RC(build_compact_jacobi(c));
            if (c->lists.compact_jac) c->jac_impl = JI_COMPACT;
        } else if (sor)
            for (int q = 0; q < 3; q++) {
                c->lists.n_col[q] = (int)col[q].size();
                RC(dalloc(&c->lists.col[q], std::max<int64_t>(1, c->lists.n_col[q])));
                HIPCHK(hipMemcpy(c->lists.col[q], col[q].data(), sizeof(int32_t) * col[q].size(),
                                 hipMemcpyHostToDevice));
            }
    }

    uint8_t *raw_dev = nullptr;
    RC(dalloc(&raw_dev, n));
    hipError_t e = hipMemcpyAsync(raw_dev, raw, (size_t)n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->G, 0, (size_t)g.total, c->stream);
    if (e != hipSuccess) {
        (void)hipFree(raw_dev);
        return fail(PM_E_HIP, "geometry upload: %s", hipGetErrorString(e));
    }
    pm::launch_geom_fill(g, raw_dev, c->G, c->prm.walls, c->stream);
    pm::launch_exchange_full<uint8_t>(g, c->G, nullptr, nullptr, nullptr, c->stream);
    HIPCHK(hipMemsetAsync(c->pc, 0xff, (size_t)g.total, c->stream));
    HIPCHK(hipMemsetAsync(c->code, 0, sizeof(int32_t) * (size_t)g.total, c->stream));
    pm::launch_eval_geometry(g, c->G, c->code, c->pc, c->stream);
    pm::launch_lag_records(g, c->lists, c->pc, c->stream);
    if (c->lists.compact || c->lists.compact_jac) {
        // the +-2 links into the solid sides' slots (out of place)
        uint2 *L1 = nullptr;
        RC(dalloc(&L1, 3 * (int64_t)c->lists.n_centries));
        pm::launch_link_subst(c->lists, c->lists.clink, L1, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        (void)hipFree(c->lists.clink);
        c->lists.clink = L1;
        if (c->lists.compact && c->sor_packed) RC(pack_sor_links(c));
        if (c->lists.compact_jac && c->jac_packed && !c->lists.links_aos) RC(pack_jac_links(c));
    }
    RC(init_state(c));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    (void)hipFree(raw_dev);
    c->fluid = fluid;
    // src/geometry.cc:335-364: fluid_count / vox_count in double
    c->phi = (double)fluid / (double)(g.nx * g.ny * g.nz);
    if (c->auto_relax) {
        // The relaxation pair by porosity.  The sweep is stable while
        // omega * omega_p stays below about 19; within that, omega speeds up
        // the viscous relaxation of the velocities and omega_p the Darcy-like
        // diffusion of the pressure.  Tight media (porosity below 0.20) are
        // limited by the pressure mode: (0.6, 30) needs 2-4x fewer sweeps
        // there than (2.5, 7), which needs 3-4x fewer on open media (the 64
        // media of the batch workload: 1,255 -> 868 G voxel updates in all).
        const bool tight = c->phi < kAutoTightPhi;
        c->prm.omega = tight ? 0.6 : 2.5;
        c->prm.omega_p = tight ? 30.0 : 7.0;
    }
    c->t = 0;
    c->kprev = 0.0;
    c->loaded = true;
    c->lagged = false;
    RC(choose_jacobi_layout(c));
    tuner_setup(c);
    return PM_OK;
}

// Complete a lagged Jacobi state (u^n, p^{n-1}): the pressure and outlet
// buffer of the last sweep, then the ghost fill.  Afterwards the state is the
// one the velocity + pressure kernel pair would have left.
static int settle(pm_ctx *c) {
    if (!c->lagged) return PM_OK;
    c->lagged = false;
    const pm::Grid &g = c->g;
    hipStream_t s = c->stream;
    const double om = c->prm.omega, wp = c->prm.omega_p > 0.0 ? c->prm.omega_p : om;
    const double cu = om * pm::kDt, cp = wp * pm::kC2 * pm::kDt;
    if (c->dc()) {
        pm::launch_jacobi4(pm::JAC_SETTLE, g, c->lists, c->pc, c->S4[0], c->qD[0], c->S4[0], c->qD[0], c->R4,
                           (float)cu, (float)cp, s);
        pm::launch_ghost4_lists(g, c->lists, c->G, c->S4[0], c->qD[0], s);
    } else {
        pm::launch_jacobi_stage(pm::JAC_SETTLE, g, c->lists, c->pc, c->X[0], c->X[1], c->X[2], c->X[3], c->qX[0],
                                c->X[0], c->X[1], c->X[2], c->X[3], c->qX[0], cu, cp, 3.0, s);
        pm::launch_ghost_fast(g, c->G, c->X[0], c->X[1], c->X[2], c->X[3], c->qX[0], (double)(g.nz + 2),
                                      2.0, 3.0, s);
    }
    HIPCHK(hipGetLastError());
    return PM_OK;
}

// One FP64 Jacobi sweep from set A (p, u, v, w; outlet buffer qa) into set B
// with the context's kernels: the velocity + pressure pair, or one lagged
// pass (dense 2.5-D or fluid-list form).
static void jacobi_pass(pm_ctx *c, double *const *A, double *qa, double *const *B, double *qb, double cu,
                        double cp, double pin_out) {
    const pm::Grid &g = c->g;
    if (c->jac_impl == JI_PAIR) {
        (void)settle(c);  // the pair reads a complete state
        pm::launch_jacobi(g, c->lists, c->pc, A[0], A[1], A[2], A[3], qa, B[0], B[1], B[2], B[3], qb, cu, cp,
                          pin_out, c->stream);
        return;
    }
    const int st = c->lagged ? pm::JAC_LAG : pm::JAC_FIRST;
    if (c->jac_impl == JI_STREAM)
        pm::launch_jacobi_stream(st, c->stream_zs, g, c->pc, A[0], A[1], A[2], A[3], qa, B[0], B[1], B[2], B[3],
                                 cu, cp, pin_out, c->stream);
    else
        pm::launch_jacobi_stage(st, g, c->lists, c->pc, A[0], A[1], A[2], A[3], qa, B[0], B[1], B[2], B[3], qb,
                                cu, cp, pin_out, c->stream);
    c->lagged = true;
}

// X += d on the fluid interior, d = 0 (defect correction); the ghost fill
// then refills d's periodic halo copies with zeros.
static void fold_now(pm_ctx *c) {
    const pm::Grid &g = c->g;
    pm::launch_fold4(c->lists, c->X[0], c->X[1], c->X[2], c->X[3], c->S4[0], c->stream);
    pm::launch_ghost4_lists(g, c->lists, c->G, c->S4[0], c->qD[0], c->stream);
}

static void apply_variant(pm_ctx *c, const SweepVariant &v) {
    if (v.impl == JI_COMPACT) {
        // the compact pass: every load, or each only where it is used (also
        // at 6 and 8 waves per SIMD)
        c->jac_impl = JI_COMPACT;
        c->lists.rec_skip1 = c->rec_skip1 = v.skip1;
        c->lists.rec_occ = c->rec_occ = v.skip1 ? v.occ : 0;
        return;
    }
    c->jac_impl = v.impl;
    c->lists.select_loads = v.select_loads;
    c->stream_zs = v.zs;
    c->lists.use_rec = v.impl == JI_RECORDS;
    c->lists.use_march = v.impl == JI_MARCH && pm::march_supported(c->g, c->lists);
    if (v.impl == JI_MARCH) {
        c->lists.march_ty = c->march_ty = v.ty;
        c->lists.march_zseg = c->march_zseg = v.zs;
    }
    c->lists.rec_occ = c->rec_occ = v.impl == JI_RECORDS ? v.occ : 0;
    c->rec_skip2 = v.impl == JI_RECORDS && v.skip2;
    c->rec_skip1 = c->rec_skip2 && v.skip1;
    c->lists.rec_skip2 = c->rec_skip2 && pm::rec_skip2_supported(c->g);
    c->lists.rec_skip1 = c->rec_skip1;
}

static const char *variant_name(const pm_ctx *c, const SweepVariant &v, char *buf, size_t len) {
    if (v.impl == JI_COMPACT && v.skip1 && v.occ > 0)
        std::snprintf(buf, len, "interleaved-compact%s-skipall-occ%d", c->lists.packed ? "-packed" : "", v.occ);
    else if (v.impl == JI_COMPACT)
        std::snprintf(buf, len, "interleaved-compact%s%s", c->lists.packed ? "-packed" : "", v.skip1 ? "-skipall" : "");
    else if (c->dc() && v.impl == JI_PAIR) std::snprintf(buf, len, "interleaved-pair");
    else if (c->dc() && v.impl == JI_MARCH) std::snprintf(buf, len, "interleaved-zmarch-ty%d-zs%d", v.ty, v.zs);
    else if (c->dc() && v.impl == JI_RECORDS && v.skip2)
        std::snprintf(buf, len, "interleaved-lagged-records-%s", v.skip1 ? "skipall" : "skip2");
    else if (c->dc() && v.impl == JI_RECORDS && v.occ > 0)
        std::snprintf(buf, len, "interleaved-lagged-records-occ%d", v.occ);
    else if (c->dc())
        std::snprintf(buf, len, "interleaved-lagged-%s",
                      v.impl == JI_RECORDS ? "records" : v.select_loads ? "eager" : "branch");
    else if (v.impl == JI_STREAM) std::snprintf(buf, len, "dense-2.5d-zs%d", v.zs);
    else if (v.impl == JI_RECORDS) std::snprintf(buf, len, "list-lagged-records%s", v.skip2 ? "-skip2" : "");
    else std::snprintf(buf, len, "%s-%s", v.impl == JI_PAIR ? "list-pair" : "list-lagged",
                       v.select_loads ? "select" : "branch");
    return buf;
}

// Candidates of the Jacobi autotuner (fast Jacobi mode), unless the kernels
// are fixed by POREMAPS_JACOBI or tuning is off (POREMAPS_TUNE=0).
static void tuner_setup(pm_ctx *c) {
    Tuner &tn = c->tuner;
    tn = Tuner();
    const char *fixed = std::getenv("POREMAPS_JACOBI");
    const char *off = std::getenv("POREMAPS_TUNE");
    if (!c->fast() || !c->jacobi() || c->lists.n_all == 0 || (fixed && *fixed) || (off && std::strcmp(off, "0") == 0))
        return;
    // the first window (caches and clocks still cold after the load) is run
    // with the last candidate and not scored
    const SweepVariant soa[] = {{JI_STREAM, true, 16}, {JI_PAIR, true, 64}, {JI_PAIR, false, 64},
                                {JI_LIST, true, 64}, {JI_RECORDS, true, 64}, {JI_RECORDS, true, 64, 4, 0, true},
                                {JI_STREAM, true, 64}, {JI_STREAM, true, 16}};
    // the record pass also at 6 and 8 waves per SIMD (its registers capped at
    // 80 and 64, a few spilled), and with its +-2 stencil loads made only by
    // the lanes whose case reads them
    const SweepVariant aos[] = {{JI_RECORDS, true, 64}, {JI_PAIR, true, 64}, {JI_LIST, false, 64},
                                {JI_LIST, true, 64}, {JI_RECORDS, true, 64, 4, 6}, {JI_RECORDS, true, 64, 4, 8},
                                {JI_RECORDS, true, 64, 4, 0, true}, {JI_RECORDS, true, 64, 4, 0, true, true},
                                {JI_RECORDS, true, 64}};
    // the z-march form: 64 x 4 tiles, z segments of 128 planes (64 x 8 tiles
    // and 64-plane segments measured slower on the 512^3 field: 4.09 and 2.40
    // against 2.37 ms per sweep)
    const SweepVariant march[] = {{JI_MARCH, true, 128, 4}};
    // slip walls are held and mirrored by the list forms only (not by the
    // dense 2.5-D form or the z-march, nor by the pass that skips every unused
    // load, which reads the -1 neighbours by the code's own solid flags)
    const auto usable = [&](const SweepVariant &v) {
        if (v.skip2 && !pm::rec_skip2_supported(c->g)) return false;
        return !(c->g.slip && (v.impl == JI_STREAM || v.impl == JI_MARCH || v.skip1));
    };
    if (c->lists.compact_jac) {
        // the compact state: its pass with every load and with each load only
        // where it is used (the dense forms read another layout)
        // (8 waves per SIMD spills and measured 1.54 against 0.70 ms per sweep
        // on the 512^3 field; it is not a candidate)
        const SweepVariant cmp[] = {{JI_COMPACT, true, 64}, {JI_COMPACT, true, 64, 4, 0, true, true},
                                    {JI_COMPACT, true, 64, 4, 6, true, true}, {JI_COMPACT, true, 64}};
        for (const auto &v : cmp) tn.cand[tn.n++] = v;
    } else if (c->dc()) {
        for (const auto &v : aos)
            if (usable(v)) tn.cand[tn.n++] = v;
        if (pm::march_supported(c->g, c->lists))
            for (const auto &v : march)
                if (usable(v)) tn.cand[tn.n++] = v;
    }
    else {
        for (const auto &v : soa)
            if (usable(v)) tn.cand[tn.n++] = v;
        // the unscored first window's form, scored again at the end (the
        // dense form is not there under slip)
        if (c->g.slip) tn.cand[tn.n++] = tn.cand[0];
    }
    tn.first_scored = 1;
    tn.active = true;
}

// Before a Jacobi sweep: the candidate of the current window.
static int tuner_begin(pm_ctx *c) {
    Tuner &tn = c->tuner;
    if (!tn.active || tn.done) return PM_OK;
    if (tn.in_window == 0) {
        apply_variant(c, tn.cand[tn.idx]);
        if (c->jac_impl == JI_PAIR) RC(settle(c));
        HIPCHK(hipEventCreate(&tn.ev0[tn.idx]));
        HIPCHK(hipEventRecord(tn.ev0[tn.idx], c->stream));
    }
    return PM_OK;
}

// After a Jacobi sweep (ghost fill included): close the window; after the
// last one keep the fastest candidate.
static int tuner_end(pm_ctx *c) {
    Tuner &tn = c->tuner;
    if (!tn.active || tn.done) return PM_OK;
    if (++tn.in_window < tn.window) return PM_OK;
    HIPCHK(hipEventCreate(&tn.ev1[tn.idx]));
    HIPCHK(hipEventRecord(tn.ev1[tn.idx], c->stream));
    tn.in_window = 0;
    if (++tn.idx < tn.n) return PM_OK;
    HIPCHK(hipEventSynchronize(tn.ev1[tn.n - 1]));
    tn.best = tn.first_scored;
    for (int q = 0; q < tn.n; q++) {
        float e = 0.f;
        (void)hipEventElapsedTime(&e, tn.ev0[q], tn.ev1[q]);
        tn.ms[q] = e / tn.window;
        (void)hipEventDestroy(tn.ev0[q]);
        (void)hipEventDestroy(tn.ev1[q]);
        tn.ev0[q] = tn.ev1[q] = nullptr;
        if (q >= tn.first_scored && tn.ms[q] < tn.ms[tn.best]) tn.best = q;
    }
    tn.done = true;
    apply_variant(c, tn.cand[tn.best]);
    return PM_OK;
}

// One sweep of the context's ordering, ghost fill and pins included; for the
// defect-correction mode also the fold when the sweep index hits it_eval.
// Jacobi sweeps run in lagged form (one pass each); settle() completes the
// state before anything reads the pressure.
static int one_sweep(pm_ctx *c) {
    const pm::Grid &g = c->g;
    hipStream_t s = c->stream;
    const double om = c->prm.omega, wp = c->prm.omega_p > 0.0 ? c->prm.omega_p : om;
    const double cu = om * pm::kDt, cp = wp * pm::kC2 * pm::kDt;
    if (!c->fast()) {
        tb(c, TC_REF);
        if (c->ref_nblk > 1)
            pm::launch_ref_block_sweep(g, c->G, c->code, c->X, c->ref_snap, c->ref_outlet, c->ref_tab, c->ref_nblk,
                                       c->ref_ntile, c->prm.solving_algorithm, s);
        else
            pm::launch_ref_sweeps(g, c->G, c->code, c->X[0], c->X[1], c->X[2], c->X[3], c->prm.solving_algorithm,
                                  1, c->ref_impl, s);
        te(c);
    } else if (!c->dc()) {
        const double pin_in = (double)(g.nz + 2);
        if (c->jacobi()) {
            RC(tuner_begin(c));
            tb(c, TC_SWEEP);
            jacobi_pass(c, c->X, c->qX[0], c->X2, c->qX[1], cu, cp, 3.0);
            te(c);
            tb(c, TC_GHOST);
            pm::launch_ghost_fast(g, c->G, c->X2[0], c->X2[1], c->X2[2], c->X2[3], c->qX[1], pin_in, 2.0,
                                          3.0, s);
            te(c);
            for (int q = 0; q < 4; q++) std::swap(c->X[q], c->X2[q]);
            std::swap(c->qX[0], c->qX[1]);
            RC(tuner_end(c));
        } else {
            tb(c, TC_SWEEP);
            pm::launch_colour(g, c->lists, c->pc, c->X[0], c->X[1], c->X[2], c->X[3], c->qX[0], cu, cp, 3.0, s);
            te(c);
            tb(c, TC_GHOST);
            pm::launch_ghost_fast(g, c->G, c->X[0], c->X[1], c->X[2], c->X[3], c->qX[0], pin_in, 2.0,
                                          3.0, s);
            te(c);
        }
    } else {
        const float fcu = (float)cu, fcp = (float)cp;
        if (c->jacobi()) {
            RC(tuner_begin(c));
            // the first pass from the zero state is timed apart: it is not a
            // full sweep
            const bool first0 = !c->lagged && c->dzero && c->jac_impl != JI_PAIR &&
                                pm::jacobi4_fills_halos(pm::JAC_FIRST_ZERO, c->lists);
            tb(c, first0 ? TC_FIRST0 : TC_SWEEP);
            int stage = pm::JAC_PAIR;
            if (c->jac_impl == JI_PAIR) {
                RC(settle(c));
                pm::launch_jacobi4(pm::JAC_PAIR, g, c->lists, c->pc, c->S4[0], c->qD[0], c->S4[1], c->qD[1], c->R4,
                                   fcu, fcp, s);
            } else {
                stage = c->lagged ? pm::JAC_LAG : c->dzero ? pm::JAC_FIRST_ZERO : pm::JAC_FIRST;
                pm::launch_jacobi4(stage, g, c->lists, c->pc, c->S4[0], c->qD[0], c->S4[1], c->qD[1], c->R4, fcu,
                                   fcp, s);
                c->lagged = true;
            }
            te(c);
            if (!pm::jacobi4_fills_halos(stage, c->lists)) {
                tb(c, TC_GHOST);
                pm::launch_ghost4(g, c->G, c->S4[1], c->qD[1], false, s);
                te(c);
            }
            std::swap(c->S4[0], c->S4[1]);
            std::swap(c->qD[0], c->qD[1]);
            c->dzero = false;
            RC(tuner_end(c));
        } else {
            tb(c, TC_SWEEP);
            pm::launch_colour4(g, c->lists, c->pc, c->S4[0], c->qD[0], c->R4, fcu, fcp, s);
            c->dzero = false;
            te(c);
            tb(c, TC_GHOST);
            pm::launch_ghost4_lists(g, c->lists, c->G, c->S4[0], c->qD[0], s);
            te(c);
        }
        if (c->t % c->prm.it_eval == 0) {
            tb(c, TC_FOLD);
            RC(settle(c));
            fold_now(c);
            pm::launch_ghost_fast(g, c->G, c->X[0], c->X[1], c->X[2], c->X[3], nullptr,
                                          (double)(g.nz + 2), 2.0, 3.0, s);
            te(c);
            RC(dc_prepare(c));
        }
    }
    HIPCHK(hipGetLastError());
    c->t++;
    return PM_OK;
}

// Fold the outstanding correction into the FP64 state (defect-correction
// mode), so that X is the solution state.
static int dc_flush(pm_ctx *c) {
    RC(settle(c));
    if (!c->dc()) return PM_OK;
    const pm::Grid &g = c->g;
    fold_now(c);
    pm::launch_ghost_fast(g, c->G, c->X[0], c->X[1], c->X[2], c->X[3], nullptr, (double)(g.nz + 2), 2.0,
                                  3.0, c->stream);
    RC(dc_prepare(c));
    HIPCHK(hipGetLastError());
    return PM_OK;
}

static int reduce(pm_ctx *c, bool full, pm::RedSums *out) {
    RC(settle(c));
    pm::DomBox box;
    const int *d = c->prm.dom_interest;
    box.on = !(d[0] == 0 && d[1] == 0 && d[2] == 0 && d[3] == 0 && d[4] == 0 && d[5] == 0);
    for (int q = 0; q < 6; q++) box.d[q] = d[q];
    pm::launch_reduce(c->g, c->G, c->X[0], c->X[1], c->X[2], c->X[3], box, full, c->red_scratch, c->red_dev,
                      c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->red_host, c->red_dev, sizeof(pm::RedSums), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *out = *c->red_host;
    return PM_OK;
}

// src/evaluation.cc:216-262 (one rank).
static double convergence_from(const pm::RedSums &r, double *kprev) {
    double velz_sum = r.sum[2];
    int flcount = (int)r.count[0];
    if (flcount == 0) velz_sum = 0.0;
    const double mean_velz = velz_sum / flcount;
    const double kcur = mean_velz / pm::kRe;
    const double conv = std::fabs(kcur - *kprev) / kcur;
    *kprev = kcur;
    return conv;
}

// src/evaluation.cc:54-211 (one rank); out = k13 k23 k33 wk13 wk23 wk33 wmax wmean.
static void permeability_from(const pm::RedSums &r, const int *dom, double dx, double phi, double *out) {
    const bool dofi = !(dom[0] == 0 && dom[1] == 0 && dom[2] == 0 && dom[3] == 0 && dom[4] == 0 && dom[5] == 0);
    double wsx = r.sum[0], wsy = r.sum[1], wsz = r.sum[2], wmax = r.wmax;
    int wfl = (int)r.count[0];
    double sx = r.sum[3], sy = r.sum[4], sz = r.sum[5], p0 = r.sum[6], pend = r.sum[7];
    int fl = dofi ? (int)r.count[1] : 0, p0n = (int)r.count[2], pendn = (int)r.count[3];
    if (wfl == 0) { wsx = 0.0; wsy = 0.0; wsz = 0.0; wmax = 0.0; }
    if (fl == 0) { sx = 0.0; sy = 0.0; sz = 0.0; }
    if (p0n == 0) p0 = 0.0;
    if (pendn == 0) pend = 0.0;
    const double dx2 = dx * dx;
    if (dofi) {
        const double pgrad = ((pend / pendn) - (p0 / p0n)) / (dom[5] - dom[4]);
        out[0] = (sx / fl) * (1.0 / pm::kRe) * phi * (1.0 / pgrad) * dx2;
        out[1] = (sy / fl) * (1.0 / pm::kRe) * phi * (1.0 / pgrad) * dx2;
        out[2] = (sz / fl) * (1.0 / pm::kRe) * phi * (1.0 / pgrad) * dx2;
    } else {
        out[0] = out[1] = out[2] = 0.0;
    }
    out[3] = (wsx / wfl) * (1.0 / pm::kRe) * phi * dx2;
    out[4] = (wsy / wfl) * (1.0 / pm::kRe) * phi * dx2;
    out[5] = (wsz / wfl) * (1.0 / pm::kRe) * phi * dx2;
    out[6] = wmax * dx;
    out[7] = wsz / wfl * dx;
}

static double used_porosity(const pm_ctx *c) {
    // src/calc_flow.cc:92-97
    return c->prm.porosity == -1.0 ? c->phi : c->prm.porosity;
}

static void fill_row(pm_log_row *row, int t, double conv, double tps, const double *k) {
    row->iteration = t;
    row->conv = conv;
    row->tps = tps;
    row->wmax_velz = k[6];
    row->wmean_velz = k[7];
    row->k13 = k[0]; row->k23 = k[1]; row->k33 = k[2];
    row->wk13 = k[3]; row->wk23 = k[4]; row->wk33 = k[5];
}

extern "C" int pm_solve(pm_ctx *c, pm_log_fn log_fn, pm_progress_fn progress_fn, void *user, pm_result *res) {
    if (!c) return fail(PM_E_ARG, "null context");
    if (!c->loaded) return fail(PM_E_STATE, "pm_solve before pm_load_geometry");
    HIPCHK(hipSetDevice(c->prm.device));
    const pm_params &P = c->prm;
    const double phi = used_porosity(c);
    double conv = 1.0 + P.eps;
    int nrows = 0;
    pm_log_row last;
    std::memset(&last, 0, sizeof last);
    int status = PM_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    const auto t_start = clk::now();
    auto t_mark = t_start;
    int t_mark_it = c->t;
    // src/calc_flow.cc:101: while (iteration < max_iterations && perm_conv >= eps)
    while (c->t < P.max_iter && conv >= P.eps) {
        // run sweeps up to the next iteration that needs the host
        int stop = c->t;
        while (stop + 1 < P.max_iter) {
            const bool eval = (stop % P.it_eval == 0 && stop > 1000);
            const bool prog = progress_fn && stop % 100 == 0;
            const bool fold = c->dc() && stop % P.it_eval == 0;
            if (eval || prog || fold) break;
            stop++;
        }
        while (c->t <= stop) RC(one_sweep(c));
        const int it = stop;  // the iteration index whose end we are at
        if (it % P.it_eval == 0 && it > 1000) {
            pm::RedSums r;
            RC(reduce(c, false, &r));
            conv = convergence_from(r, &c->kprev);
            if (it % P.it_write == 0) {
                RC(reduce(c, true, &r));
                double k[8];
                permeability_from(r, P.dom_interest, P.voxel_size, phi, k);
                const auto now = clk::now();
                const double dt = std::chrono::duration<double>(now - t_mark).count();
                const int nit = std::max(1, c->t - t_mark_it);
                fill_row(&last, it, conv, nit / std::max(dt, 1e-12), k);
                nrows++;
                if (log_fn && log_fn(&last, user) != 0) { status = fail(PM_E_ABORTED, "log callback stopped the solve"); break; }
            }
        }
        if (progress_fn && it % 100 == 0) {
            HIPCHK(hipStreamSynchronize(c->stream));
            const auto now = clk::now();
            const double dt = std::chrono::duration<double>(now - t_mark).count();
            const int nit = std::max(1, c->t - t_mark_it);
            t_mark = now;
            t_mark_it = c->t;
            if (progress_fn(it, conv, dt / nit, user) != 0) { status = fail(PM_E_ABORTED, "progress callback stopped the solve"); break; }
        }
        if (!std::isfinite(conv)) {
            // a fast-mode relaxation beyond the stability bound: start over slower
            if (pm_restart_slower(c) != PM_OK) break;
            conv = 1.0 + P.eps;
            t_mark = clk::now();
            t_mark_it = c->t;
        }
    }
    RC(dc_flush(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double secs = std::chrono::duration<double>(clk::now() - t_start).count();
    if (res) {
        pm::RedSums r;
        RC(reduce(c, true, &r));
        double k[8];
        permeability_from(r, P.dom_interest, P.voxel_size, phi, k);
        std::memset(res, 0, sizeof *res);
        res->sweeps = c->t;
        res->conv = conv;
        res->porosity = phi;
        res->seconds = secs;
        res->glups = (double)c->g.nx * c->g.ny * c->g.nz * c->t / std::max(secs, 1e-12) / 1e9;
        res->n_log_rows = nrows;
        res->restarts = c->restarts;
        fill_row(&res->last, c->t - 1, conv, 0.0, k);
    }
    if (status != PM_OK) return status;
    if (!std::isfinite(conv)) return fail(PM_E_DIVERGED, "convergence value is %g at iteration %d", conv, c->t - 1);
    return PM_OK;
}

extern "C" int pm_step(pm_ctx *c, int n) {
    if (!c) return fail(PM_E_ARG, "null context");
    if (!c->loaded) return fail(PM_E_STATE, "pm_step before pm_load_geometry");
    if (n < 0) return fail(PM_E_ARG, "negative sweep count");
    HIPCHK(hipSetDevice(c->prm.device));
    for (int q = 0; q < n; q++) RC(one_sweep(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PM_OK;
}

static int copy_field(pm_ctx *c, int field, std::vector<double> &host) {
    RC(settle(c));
    const pm::Grid &g = c->g;
    host.resize((size_t)g.total);
    double *src = nullptr;
    if (field >= PM_FIELD_VELX && field <= PM_FIELD_VELZ) src = c->X[1 + field - PM_FIELD_VELX];
    else if (field == PM_FIELD_PRESS) src = c->X[0];
    else return fail(PM_E_ARG, "unknown field %d", field);
    HIPCHK(hipMemcpyAsync(host.data(), src, sizeof(double) * (size_t)g.total, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PM_OK;
}

extern "C" int pm_get_field(pm_ctx *c, int field, void *out) {
    if (!c || !out) return fail(PM_E_ARG, "null argument");
    if (!c->loaded) return fail(PM_E_STATE, "no geometry loaded");
    HIPCHK(hipSetDevice(c->prm.device));
    const pm::Grid &g = c->g;
    if (field == PM_FIELD_NHOOD) {
        std::vector<int32_t> h((size_t)g.total);
        HIPCHK(hipMemcpyAsync(h.data(), c->code, sizeof(int32_t) * (size_t)g.total, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        int32_t *o = (int32_t *)out;
        for (int i = 0; i < g.nz; i++)
            for (int j = 0; j < g.ny; j++)
                std::memcpy(o + ((int64_t)i * g.ny + j) * g.nx, h.data() + g.at(i + 2, j + 2, 2),
                            sizeof(int32_t) * g.nx);
        return PM_OK;
    }
    std::vector<double> h;
    RC(copy_field(c, field, h));
    double *o = (double *)out;
    for (int i = 0; i < g.nz; i++)
        for (int j = 0; j < g.ny; j++)
            std::memcpy(o + ((int64_t)i * g.ny + j) * g.nx, h.data() + g.at(i + 2, j + 2, 2), sizeof(double) * g.nx);
    return PM_OK;
}

extern "C" int pm_get_field_halo(pm_ctx *c, int field, double *out) {
    if (!c || !out) return fail(PM_E_ARG, "null argument");
    if (!c->loaded) return fail(PM_E_STATE, "no geometry loaded");
    HIPCHK(hipSetDevice(c->prm.device));
    const pm::Grid &g = c->g;
    std::vector<double> h;
    RC(copy_field(c, field, h));
    for (int i = 0; i < g.NZ; i++)
        for (int j = 0; j < g.NY; j++)
            std::memcpy(out + ((int64_t)i * g.NY + j) * g.NX, h.data() + g.at(i, j, 0), sizeof(double) * g.NX);
    return PM_OK;
}

extern "C" int pm_porosity(pm_ctx *c, double *phi) {
    if (!c || !phi) return fail(PM_E_ARG, "null argument");
    if (!c->loaded) return fail(PM_E_STATE, "no geometry loaded");
    *phi = c->phi;
    return PM_OK;
}

extern "C" int pm_permeability(pm_ctx *c, double out[8]) {
    if (!c || !out) return fail(PM_E_ARG, "null argument");
    if (!c->loaded) return fail(PM_E_STATE, "no geometry loaded");
    HIPCHK(hipSetDevice(c->prm.device));
    pm::RedSums r;
    RC(reduce(c, true, &r));
    permeability_from(r, c->prm.dom_interest, c->prm.voxel_size, used_porosity(c), out);
    return PM_OK;
}

// Algorithmic HBM bytes of one launch group per voxel of the interior
// (S = bytes per value of the swept state): every operand read once, every
// result written once.
static double algo_bytes(const pm_ctx *c, int cat) {
    const double n = (double)c->g.nx * c->g.ny * c->g.nz;
    const double S = c->dc() ? 4.0 : 8.0;
    const double R = c->dc() ? 16.0 : 0.0;
    switch (cat) {
    case TC_SWEEP:  // read p,u,v,w + code (+ R), write p,u,v,w
        return n * (8.0 * S + 1.0 + R);
    case TC_REF:    // same minimum for the exact-order sweep (FP64, int32 code + bool geometry)
        return n * (8.0 * 8.0 + 5.0);
    case TC_GHOST: {
        const pm::Grid &g = c->g;
        const double faces = 2.0 * ((double)g.ny * g.nz + (double)g.nx * g.nz + (double)g.nx * g.ny);
        return faces * 2.0 * 4.0 * S * 2.0;  // 2 layers, 4 fields, read + write
    }
    case TC_FOLD:      // read X (32) + d (16) + code, write X (32) + d (16)
        return n * 97.0;
    case TC_RESIDUAL:  // read X (32) + code, write R (16)
        return n * 49.0;
    case TC_FIRST0:    // first pass from the zero state: read code + R (16), write the state (16)
        return n * 33.0;
    default:
        return 0.0;
    }
}

extern "C" int pm_relaxation(pm_ctx *c, double *omega, double *omega_p) {
    if (!c || !omega || !omega_p) return fail(PM_E_ARG, "null argument");
    *omega = c->prm.omega;
    *omega_p = c->prm.omega_p > 0.0 ? c->prm.omega_p : c->prm.omega;
    return PM_OK;
}

extern "C" int pm_kernel_choice(pm_ctx *c, char *name, int len, char *tried, int tried_len) {
    if (!c || !name || len <= 0) return fail(PM_E_ARG, "null argument");
    char buf[64];
    const SweepVariant cur = {c->jac_impl, c->lists.select_loads,
                              c->jac_impl == JI_MARCH ? c->march_zseg : c->stream_zs, c->march_ty,
                              c->jac_impl == JI_RECORDS || c->jac_impl == JI_COMPACT ? c->lists.rec_occ : 0,
                              c->jac_impl == JI_RECORDS && c->lists.rec_skip2,
                              (c->jac_impl == JI_RECORDS && c->lists.rec_skip2 && c->lists.rec_skip1) ||
                                  (c->jac_impl == JI_COMPACT && c->lists.rec_skip1)};
    const char *sor = !c->fast() ? "n/a" : !c->dc() ? "list-colour" : c->lists.split ? "interleaved-colour-split"
                                                                                    : "interleaved-colour";
    if (c->fast() && c->dc() && c->lists.split) {
        char z[32] = "";
        if (c->lists.sor_chunk > 0)
            std::snprintf(z, sizeof z, "-z%d%s", c->lists.sor_chunk, c->lists.sor_fused ? "" : "-seq");
        std::snprintf(buf, sizeof buf, "interleaved-colour-split%s%s", z,
                      c->lists.sor_skip2 && pm::rec_skip2_supported(c->g) ? "-skip2" : "");
        sor = buf;
        if (c->lists.compact) sor = c->lists.packed ? "interleaved-colour-compact-packed" : "interleaved-colour-compact";
    }
    std::snprintf(name, (size_t)len, "%s", c->jacobi() ? variant_name(c, cur, buf, sizeof buf) : sor);
    if (tried && tried_len > 0) {
        std::string t;
        if (c->layout_ms[0] > 0.f) {
            // the FP32+DC Jacobi layout choice (choose_jacobi_layout)
            char e[96];
            std::snprintf(e, sizeof e, "layout-compact=%.4f;layout-dense=%.4f", c->layout_ms[0], c->layout_ms[1]);
            t += e;
        }
        if (c->pack_far >= 0) {
            // the entries whose links fit no packed record (the records are
            // kept unless more than a quarter of the entries are such)
            char e[48];
            std::snprintf(e, sizeof e, "%spacked-far=%d", t.empty() ? "" : ";", c->pack_far);
            t += e;
        }
        const Tuner &tn = c->tuner;
        if (tn.done)
            for (int q = tn.first_scored; q < tn.n; q++) {
                char e[96];
                std::snprintf(e, sizeof e, "%s%s=%.4f", q > tn.first_scored || !t.empty() ? ";" : "",
                              variant_name(c, tn.cand[q], buf, sizeof buf), tn.ms[q]);
                t += e;
            }
        std::snprintf(tried, (size_t)tried_len, "%s", t.c_str());
    }
    return PM_OK;
}

extern "C" int pm_time_kernels(pm_ctx *c, int n, pm_kernel_timing *out, int max_out, int *n_out) {
    if (!c || !out || !n_out) return fail(PM_E_ARG, "null argument");
    if (!c->loaded) return fail(PM_E_STATE, "no geometry loaded");
    HIPCHK(hipSetDevice(c->prm.device));
    Timer tm;
    c->timer = &tm;
    int rc = PM_OK;
    for (int q = 0; q < n && rc == PM_OK; q++) rc = one_sweep(c);
    c->timer = nullptr;
    HIPCHK(hipStreamSynchronize(c->stream));
    double ms[TC_N] = {0};
    int cnt[TC_N] = {0};
    for (auto &r : tm.recs) {
        float e = 0.f;
        (void)hipEventElapsedTime(&e, r.a, r.b);
        ms[r.cat] += e;
        cnt[r.cat]++;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    if (rc != PM_OK) return rc;
    int k = 0;
    for (int cat = 0; cat < TC_N && k < max_out; cat++) {
        if (!cnt[cat]) continue;
        std::snprintf(out[k].name, sizeof out[k].name, "%s", kTimerNames[cat]);
        out[k].launches = cnt[cat];
        out[k].avg_ms = ms[cat] / cnt[cat];
        out[k].bytes_per_launch = algo_bytes(c, cat);
        k++;
    }
    *n_out = k;
    return PM_OK;
}

// Context internals for the batch engine (declared in pm_ctx.h).
int pm_fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return fail(code, "%s", buf);
}
int pm_one_sweep(pm_ctx *c) { return one_sweep(c); }
int pm_settle(pm_ctx *c) { return settle(c); }
int pm_reduce_state(pm_ctx *c, bool full, pm::RedSums *out) { return reduce(c, full, out); }
double pm_convergence_from(const pm::RedSums &r, double *kprev) { return convergence_from(r, kprev); }
void pm_permeability_from(const pm::RedSums &r, const int *dom, double dx, double phi, double *out) {
    permeability_from(r, dom, dx, phi, out);
}
double pm_used_porosity(const pm_ctx *c) { return used_porosity(c); }
