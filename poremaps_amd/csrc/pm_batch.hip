// pm_batch.hip -- many geometries solved together on one GPU (no reference
// counterpart; the reference runs one geometry per process).
//
// Every admitted geometry is a full solver context (pm_ctx: its own arrays,
// fluid list, convergence state).  One interval = it_eval sweeps of every
// running sample + the defect-correction fold/residual + the per-sample
// reductions, launched as batched kernels: each phase is one launch over a
// table of (sample, first fluid voxel) chunks of 256, so 64 geometries cost
// the launches of one.  The interval is captured once as a HIP graph and
// replayed while the set of running samples is unchanged; finished samples
// drop out of the tables (eviction) and new ones join between intervals.
#include "../../include/poremaps_amd_batch.h"
#include "pm_compact.h"
#include "pm_ctx.h"
#include "pm_reduce.h"
#include "pm_voxel.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#pragma clang fp contract(off)

namespace pm {

constexpr int kB = 256;  // fluid voxels / halo lines / plane cells per chunk

// Device descriptor of one running sample.  A/B: the swept FP64 state of the
// FP64 mode (A4/B4 in the defect-correction mode), A the current one at the
// start of an interval.
struct BSample {
    Grid g;
    const int32_t *list;
    int n;
    const uint8_t *G, *pc;
    void *A[4], *B[4];
    void *qA, *qB;
    double *X[4];
    double *qX;
    // defect correction: interleaved correction sets and residual, and the
    // lagged pass's per-voxel records
    float4 *A4, *B4;
    Res4 R4;
    const uint2 *rec;
    // relaxation of this sample: omega*dt, omega_p*c2*dt
    double cu, cp;
    // dom_interest sub-domain of the permeability (interior indices, inclusive)
    DomBox box;
    int chunk_begin, chunk_end;
    // the lagged pass reads its +-2 stencil neighbours only where the case
    // uses them (rec_skip2_supported)
    int skip2;
    // the correction state is the compact one (pm_compact.h; FluidLists::
    // compact_jac): entry t = list position t, links clink of nE entries,
    // outlet buffers and R4.pc by entry, no halo copies
    int compact;
    const uint2 *clink;
    int nE;
    // the compact lagged pass from packed records (jac_packed_links; chunks
    // start at multiples of 64, so a wave's entries share their bases)
    int packed;
    const uint32_t *cpack;
    const int32_t *cbase;
};

struct BChunk {
    int s, start;
};

template <typename T>
__device__ __forceinline__ T *fld(void *const *set, int q) { return (T *)set[q]; }

template <typename T>
__global__ __launch_bounds__(kB) void k_b_vel(const BSample *__restrict__ S, const BChunk *__restrict__ C,
                                              int parity) {
    const BChunk ch = C[blockIdx.x];
    const BSample &s = S[ch.s];
    const int t = ch.start + threadIdx.x;
    if (t >= s.n) return;
    const int64_t c = s.list[t];
    void *const *src = parity ? s.B : s.A;
    void *const *dst = parity ? s.A : s.B;
    const T *q0 = (const T *)(parity ? s.qB : s.qA);
    vel_voxel<T, true>(s.g, c, s.pc[c], fld<T>(src, 0), fld<T>(src, 1), fld<T>(src, 2), fld<T>(src, 3), q0,
                       fld<T>(dst, 1), fld<T>(dst, 2), fld<T>(dst, 3), (T)s.cu);
}

template <typename T>
__global__ __launch_bounds__(kB) void k_b_pres(const BSample *__restrict__ S, const BChunk *__restrict__ C,
                                               int parity) {
    const BChunk ch = C[blockIdx.x];
    const BSample &s = S[ch.s];
    const int t = ch.start + threadIdx.x;
    if (t >= s.n) return;
    const int64_t c = s.list[t];
    void *const *src = parity ? s.B : s.A;
    void *const *dst = parity ? s.A : s.B;
    T *q1 = (T *)(parity ? s.qA : s.qB);
    pres_voxel<T>(s.g, c, s.pc[c], fld<T>(src, 0), fld<T>(dst, 1), fld<T>(dst, 2), fld<T>(dst, 3), fld<T>(dst, 0),
                  q1, (T)s.cp, (T)3.0);
}

// Interleaved correction state (defect correction): the passes of an
// interval, ghost fill, fold (over the fluid list: X += d, d = 0; the halo
// copies of d are refilled with zeros by the following ghost fill) and
// residual.
// The first pass of an interval: every interval starts from a zero
// correction state (the previous interval ended with the fold, the ghost fill
// of the zeroed state and the cleared outlet buffers; a new sample starts
// from zeros), so first_voxel4_zero computes it from the records and the
// residual, halo copies included.
__global__ __launch_bounds__(kB) void k_b_first0(const BSample *__restrict__ S, const BChunk *__restrict__ C,
                                                 int parity) {
    const BChunk ch = C[blockIdx.x];
    const BSample &s = S[ch.s];
    const int t = ch.start + threadIdx.x;
    if (t >= s.n) return;
    if (s.compact) first_zero_compact<false>(t, s.clink, s.nE, parity ? s.A4 : s.B4, s.R4, (float)s.cu);
    else first_voxel4_zero(s.g, t, s.rec, parity ? s.A4 : s.B4, s.R4, (float)s.cu);
}

// Lagged single pass (the sweeps after the first of an interval; from the
// per-voxel records, every load issued up front) and the settle pass that
// completes the pressure at the interval's end, in place.
// One lagged pass at list position t of sample s: the compact pass (each load
// only where used) or the dense record pass.
__device__ __forceinline__ void b_lag_cell(const BSample &s, int t, int parity) {
    if (s.compact && s.packed) {
        const float4 *S0 = parity ? s.B4 : s.A4;
        lag_voxel4_compact<false, false, true>(t, s.clink, s.nE, links_rsrc(s.clink, s.nE), state_rsrc(S0, s.nE), S0,
                                               parity ? s.A4 : s.B4, s.R4, (float)s.cu, (float)s.cp, 0.f, s.cpack,
                                               jac_wave_bases(s.cbase, t), t & 63);
    } else if (s.compact) {
        const float4 *S0 = parity ? s.B4 : s.A4;
        lag_voxel4_compact<false, false>(t, s.clink, s.nE, links_rsrc(s.clink, s.nE), state_rsrc(S0, s.nE), S0,
                                         parity ? s.A4 : s.B4, s.R4, (float)s.cu, (float)s.cp, 0.f);
    } else if (s.skip2)
        lag_voxel4_rec<true, true>(s.g, t, s.rec, parity ? s.B4 : s.A4, (const float *)(parity ? s.qB : s.qA),
                                   parity ? s.A4 : s.B4, s.R4, (float)s.cu, (float)s.cp, 0.f);
    else
        lag_voxel4_rec<true>(s.g, t, s.rec, parity ? s.B4 : s.A4, (const float *)(parity ? s.qB : s.qA),
                             parity ? s.A4 : s.B4, s.R4, (float)s.cu, (float)s.cp, 0.f);
}

__global__ __launch_bounds__(kB) void k_b_lag4(const BSample *__restrict__ S, const BChunk *__restrict__ C,
                                               int parity) {
    const BChunk ch = C[blockIdx.x];
    const BSample &s = S[ch.s];
    const int t = ch.start + threadIdx.x;
    if (t >= s.n) return;
    b_lag_cell(s, t, parity);
}

// The lagged pass of the samples on packed compact records (the run order
// puts them first), without the dense branch, so that it runs at 6 waves per
// SIMD as the single-solve pass does.
__global__ __launch_bounds__(kB, 6) void k_b_lag4p(const BSample *__restrict__ S, const BChunk *__restrict__ C,
                                                  int parity) {
    const BChunk ch = C[blockIdx.x];
    const BSample &s = S[ch.s];
    const int t = ch.start + threadIdx.x;
    if (t >= s.n) return;
    const float4 *S0 = parity ? s.B4 : s.A4;
    lag_voxel4_compact<false, false, true>(t, s.clink, s.nE, links_rsrc(s.clink, s.nE), state_rsrc(S0, s.nE), S0,
                                           parity ? s.A4 : s.B4, s.R4, (float)s.cu, (float)s.cp, 0.f, s.cpack,
                                           jac_wave_bases(s.cbase, t), t & 63);
}

// The same two passes as persistent kernels over tile queues: the samples'
// (sample, chunk) tiles back to back, largest sample first (the table
// order), cut into nq contiguous queues (one per XCD by default: workgroup b
// serves queue b mod nq, and workgroups are dispatched to the XCDs round-robin
// by id, so a queue's counter and its tiles' lines stay with one XCD).  A
// fixed grid of resident workgroups takes tickets from its queue's counter
// (one per queue and pass, zeroed at the start of the interval by
// k_b_zero_counters, a kernel node rather than a memset); a ticket stands for
// `group` consecutive tiles, and the ticket of the next group is fetched
// while the current one is worked on.  A ticket past the queue's end ends the
// workgroup, so every workgroup exits once its queue is drained, whatever the
// others do (no workgroup waits on another), and every queue has workgroups
// (nq <= grid).  A tile's cells get the same work as in the grid-wide form,
// so the results are the same bits.
template <typename F>
__device__ __forceinline__ void queue_tiles(int *ctr, int nC, int group, int nq, F &&work) {
    const int q = (int)(blockIdx.x % (unsigned)nq);
    const int lo = (int)((int64_t)nC * q / nq), hi = (int)((int64_t)nC * (q + 1) / nq);
    __shared__ int tk[2];
    if (threadIdx.x == 0) tk[0] = atomicAdd(ctr + q, 1);
    __syncthreads();
    for (int k = 0;; k++) {
        const int64_t first = lo + (int64_t)tk[k & 1] * group;
        if (first >= hi) return;  // the same value in every thread of the block
        if (threadIdx.x == 0) tk[(k + 1) & 1] = atomicAdd(ctr + q, 1);
        const int last = (int)std::min<int64_t>(first + group, hi);
        for (int b = (int)first; b < last; b++) work(b);
        __syncthreads();
    }
}

__global__ __launch_bounds__(kB) void k_b_zero_counters(int *ctr, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) ctr[i] = 0;
}

__global__ __launch_bounds__(kB) void k_b_first0_q(const BSample *__restrict__ S, const BChunk *__restrict__ C,
                                                   int nC, int *ctr, int group, int nq, int parity) {
    queue_tiles(ctr, nC, group, nq, [&](int b) {
        const BChunk ch = C[b];
        const BSample &s = S[ch.s];
        const int t = ch.start + threadIdx.x;
        if (t >= s.n) return;
        if (s.compact) first_zero_compact<false>(t, s.clink, s.nE, parity ? s.A4 : s.B4, s.R4, (float)s.cu);
        else first_voxel4_zero(s.g, t, s.rec, parity ? s.A4 : s.B4, s.R4, (float)s.cu);
    });
}

__global__ __launch_bounds__(kB) void k_b_lag4_q(const BSample *__restrict__ S, const BChunk *__restrict__ C, int nC,
                                                 int *ctr, int group, int nq, int parity) {
    queue_tiles(ctr, nC, group, nq, [&](int b) {
        const BChunk ch = C[b];
        const BSample &s = S[ch.s];
        const int t = ch.start + threadIdx.x;
        if (t >= s.n) return;
        b_lag_cell(s, t, parity);
    });
}

__global__ __launch_bounds__(kB) void k_b_settle4(const BSample *__restrict__ S, const BChunk *__restrict__ C,
                                                  int which) {
    const BChunk ch = C[blockIdx.x];
    const BSample &s = S[ch.s];
    const int t = ch.start + threadIdx.x;
    if (t >= s.n) return;
    float4 *set = which ? s.B4 : s.A4;
    if (s.compact) {
        settle_compact<false>(t, s.clink, s.nE, set, (float *)(which ? s.qB : s.qA), s.R4, (float)s.cp);
        return;
    }
    const int64_t c = s.list[t];
    pres_voxel4(s.g, c, s.pc[c], set, set, (float *)(which ? s.qB : s.qA), s.R4, (float)s.cp, 0.f);
}

__global__ __launch_bounds__(kB) void k_b_ghost4(const BSample *__restrict__ S, const BChunk *__restrict__ C,
                                                 int which) {
    const BChunk ch = C[blockIdx.x];
    const BSample &s = S[ch.s];
    const int64_t t = (int64_t)ch.start + threadIdx.x;
    // the compact state has no halo copies; its settle pass pins the outlet
    if (s.compact || t >= ghost_lines(s.g)) return;
    ghost_line4(s.g, t, s.G, which ? s.B4 : s.A4, (float *)(which ? s.qB : s.qA));
}

__global__ __launch_bounds__(kB) void k_b_fold4(const BSample *__restrict__ S, const BChunk *__restrict__ C,
                                                int which) {
    const BChunk ch = C[blockIdx.x];
    const BSample &s = S[ch.s];
    const int t = ch.start + threadIdx.x;
    if (t >= s.n) return;
    if (s.compact) fold_compact(t, s.list[t], s.X[0], s.X[1], s.X[2], s.X[3], which ? s.B4 : s.A4);
    else fold_voxel4(s.list[t], s.X[0], s.X[1], s.X[2], s.X[3], which ? s.B4 : s.A4);
}

__global__ __launch_bounds__(kB) void k_b_residual4(const BSample *__restrict__ S, const BChunk *__restrict__ C) {
    const BChunk ch = C[blockIdx.x];
    const BSample &s = S[ch.s];
    const int t = ch.start + threadIdx.x;
    if (t >= s.n) return;
    const int64_t c = s.list[t];
    residual_voxel4(s.g, c, t, s.pc[c], s.X[0], s.X[1], s.X[2], s.X[3], s.qX, s.R4);
}

// The gathered continuity residuals of the lagged pass (k_gather_rq).
__global__ __launch_bounds__(kB) void k_b_gather_rq(const BSample *__restrict__ S, const BChunk *__restrict__ C) {
    const BChunk ch = C[blockIdx.x];
    const BSample &s = S[ch.s];
    const int t = ch.start + threadIdx.x;
    if (t >= s.n) return;
    const int64_t c = s.list[t];
    if (s.compact) {
        // the compact pass reads the continuity residual per entry
        s.R4.pc[t] = s.R4.p[c];
        return;
    }
    int i, j, k;
    cell_ijk(s.g, c, i, j, k);
    int64_t nxn, nyn, nzn;
    lag_targets(s.g, c, i, j, k, nxn, nyn, nzn);
    s.R4.q[t] = make_float4(s.R4.p[c], s.R4.p[nxn], s.R4.p[nyn], s.R4.p[nzn]);
}

// which: 0 = set A, 1 = set B (the swept state, correction pins 0 in the
// defect-correction mode), 2 = the FP64 solution X (solution pins).
template <typename T, bool DC>
__global__ __launch_bounds__(kB) void k_b_ghost(const BSample *__restrict__ S, const BChunk *__restrict__ C,
                                                int which) {
    const BChunk ch = C[blockIdx.x];
    const BSample &s = S[ch.s];
    const int64_t t = (int64_t)ch.start + threadIdx.x;
    if (t >= ghost_lines(s.g)) return;
    const bool solution = which == 2 || !DC;
    const T pin_in = solution ? (T)(s.g.nz + 2) : (T)0, pin_hi = solution ? (T)2.0 : (T)0,
            pin_out = solution ? (T)3.0 : (T)0;
    if (which == 2) {
        ghost_line<T>(s.g, t, s.G, (T *)s.X[0], (T *)s.X[1], (T *)s.X[2], (T *)s.X[3], nullptr, pin_in, pin_hi,
                      pin_out);
    } else {
        void *const *set = which ? s.B : s.A;
        ghost_line<T>(s.g, t, s.G, fld<T>(set, 0), fld<T>(set, 1), fld<T>(set, 2), fld<T>(set, 3),
                      (T *)(which ? s.qB : s.qA), pin_in, pin_hi, pin_out);
    }
}

__global__ __launch_bounds__(kB) void k_b_outlet(const BSample *__restrict__ S, const BChunk *__restrict__ C) {
    const BChunk ch = C[blockIdx.x];
    const BSample &s = S[ch.s];
    const int t = ch.start + threadIdx.x;
    if (t >= s.g.NY * s.g.NX) return;
    outlet_cell(s.g, t / s.g.NX, t % s.g.NX, s.pc, s.X[0], s.X[1], s.X[2], s.X[3], 3.0, s.qX);
}

__global__ __launch_bounds__(kB) void k_b_zero_q(const BSample *__restrict__ S, const BChunk *__restrict__ C) {
    const BChunk ch = C[blockIdx.x];
    const BSample &s = S[ch.s];
    const int t = ch.start + threadIdx.x;
    // the compact state's outlet buffers (by entry) are written by its settle
    // pass and never read by the batch's passes
    if (s.compact || t >= s.g.NY * s.g.NX) return;
    const int64_t e = s.g.qat(t / s.g.NX, t % s.g.NX);
    ((float *)s.qA)[e] = 0.f;
    ((float *)s.qB)[e] = 0.f;
}


// Per-chunk partial sums of the solution state (whole-domain part of
// src/evaluation.cc:54-262).  which: 0/1 = set A/B (FP64 mode), 2 = X.
__global__ __launch_bounds__(kB) void k_b_reduce(const BSample *__restrict__ S, const BChunk *__restrict__ C, int which,
                                                 RedAcc *__restrict__ part) {
    const BChunk ch = C[blockIdx.x];
    const BSample &s = S[ch.s];
    const int t = ch.start + threadIdx.x;
    RedAcc a;
    acc_zero(a);
    if (t < s.n) {
        const int64_t c = s.list[t];
        const double *u = which == 2 ? s.X[1] : (const double *)(which ? s.B[1] : s.A[1]);
        const double *v = which == 2 ? s.X[2] : (const double *)(which ? s.B[2] : s.A[2]);
        const double *w = which == 2 ? s.X[3] : (const double *)(which ? s.B[3] : s.A[3]);
        const double wc = w[c];
        a.n[0] = 1;
        nadd(a.s[0], u[c]);
        nadd(a.s[1], v[c]);
        nadd(a.s[2], wc);
        a.wmax = wc > 0.0 ? wc : 0.0;
        if (s.box.on) {
            // the sub-domain sums of k_reduce_state<true>
            int i, j, k;
            cell_ijk(s.g, c, i, j, k);
            i -= 2; j -= 2; k -= 2;
            const DomBox &bx = s.box;
            if (k >= bx.d[0] && k <= bx.d[1] && j >= bx.d[2] && j <= bx.d[3] && i >= bx.d[4] && i <= bx.d[5]) {
                const double *p = which == 2 ? s.X[0] : (const double *)(which ? s.B[0] : s.A[0]);
                a.n[1] = 1;
                nadd(a.s[3], u[c]);
                nadd(a.s[4], v[c]);
                nadd(a.s[5], wc);
                if (i == bx.d[4]) { a.n[2] = 1; nadd(a.s[6], p[c]); }
                if (i == bx.d[5]) { a.n[3] = 1; nadd(a.s[7], p[c]); }
            }
        }
    }
    block_reduce_store(a, part + blockIdx.x);
}

// One block per sample: merge its chunks' partials in chunk order.
__global__ __launch_bounds__(kB) void k_b_reduce_final(const BSample *__restrict__ S, const RedAcc *__restrict__ part,
                                                       RedSums *__restrict__ out) {
    const BSample &s = S[blockIdx.x];
    RedAcc a;
    acc_zero(a);
    for (int b = s.chunk_begin + threadIdx.x; b < s.chunk_end; b += kB) acc_merge(a, part[b]);
    __shared__ RedAcc res;
    block_reduce_store(a, &res);
    __syncthreads();
    if (threadIdx.x == 0) {
        RedSums &o = out[blockIdx.x];
        for (int q = 0; q < 8; q++) o.sum[q] = res.s[q].s + res.s[q].c;
        for (int q = 0; q < 4; q++) o.count[q] = res.n[q];
        o.wmax = res.wmax;
    }
}

}  // namespace pm

using pm::BChunk;
using pm::BSample;

struct pm_batch {
    pm_batch_params prm;
    hipStream_t stream = nullptr;
    std::vector<pm_ctx *> ctx;       // slot -> sample (nullptr: free)
    std::vector<int> state;          // PM_SAMPLE_*
    std::vector<double> conv;
    std::vector<int64_t> bytes;
    std::vector<pm_log_row> last;
    std::vector<int> n_rows;
    int64_t used = 0, budget = 0;
    // tables of the running samples
    std::vector<int> run;            // table index -> slot
    BSample *dS = nullptr;
    BChunk *dC = nullptr, *dG = nullptr, *dP = nullptr;
    pm::RedAcc *dPart = nullptr;
    pm::RedSums *dOut = nullptr, *hOut = nullptr;
    int capS = 0, capC = 0, capG = 0, capP = 0;
    int capPart = 0, capOut = 0;     // reduction partials / outputs (grow-only)
    int nS = 0, nC = 0, nG = 0, nP = 0;
    int nCp = 0;                     // the first nCp chunks: samples on packed compact records
    bool shape_dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool in_flight = false;          // launched, not yet finished
    bool skip2 = true;               // BSample::skip2 where supported (POREMAPS_BATCH_SKIP2=0: off)
    bool launched_empty = false;     // the launched step had no running sample
    // the interval's passes as persistent kernels over the tile queues
    // (POREMAPS_BATCH_QUEUE=1; the default is one workgroup per tile, which
    // measured faster: 33,365 against 26,977 solves/h on the 64-media batch
    // workload); qgrid resident workgroups, dQ one counter per queue and pass
    bool queue = false;
    int qgrid = 0;
    int qgroup = 4;                  // tiles per ticket (POREMAPS_BATCH_QGROUP)
    int nq = 8;                      // queues (POREMAPS_BATCH_QUEUES)
    int *dQ = nullptr;
    int capQ = 0;
    bool dc() const { return prm.precision == PM_PREC_F32DC; }
};

extern "C" void pm_batch_params_default(pm_batch_params *p) {
    std::memset(p, 0, sizeof *p);
    p->precision = PM_PREC_F32DC;
    p->omega = 0.0;  // the relaxation pair by each sample's porosity
    p->omega_p = 0.0;
    p->it_eval = 100;
    p->it_write = 100;
    p->max_iter = 100000000;
    p->eps = 1e-6;
    p->device = 0;
    p->use_graph = 1;
    p->budget_bytes = 0;
}

static pm_params sample_params(const pm_batch_params *p, int nx, int ny, int nz, int bc, double dx, double phi) {
    pm_params q;
    pm_params_default(&q);
    q.nx = nx; q.ny = ny; q.nz = nz;
    q.boundary_method = bc;
    q.voxel_size = dx;
    q.porosity = phi;
    q.max_iter = p->max_iter;
    q.it_eval = p->it_eval;
    q.it_write = p->it_write;
    q.eps = p->eps;
    q.mode = PM_MODE_JACOBI;
    q.precision = p->precision;
    q.omega = p->omega;
    q.omega_p = p->omega_p;
    q.device = p->device;
    return q;
}

#define BCHK(expr)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return pm_fail(PM_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

extern "C" int pm_batch_create(const pm_batch_params *p, pm_batch **out) {
    if (!p || !out) return pm_fail(PM_E_ARG, "null argument");
    *out = nullptr;
    if (p->precision != PM_PREC_F64 && p->precision != PM_PREC_F32DC)
        return pm_fail(PM_E_ARG, "unknown precision %d", p->precision);
    if (p->it_eval <= 0 || p->it_write <= 0) return pm_fail(PM_E_ARG, "it_eval and it_write must be positive");
    if (!(p->omega >= 0.0 && p->omega < 100.0))
        return pm_fail(PM_E_ARG, "omega must be in (0, 100), or 0 for the pair by porosity");
    if (!(p->omega_p >= 0.0 && p->omega_p < 100.0)) return pm_fail(PM_E_ARG, "omega_p must be in [0, 100)");
    BCHK(hipSetDevice(p->device));
    pm_batch *b = new (std::nothrow) pm_batch;
    if (!b) return pm_fail(PM_E_MEMORY, "host allocation failed");
    b->prm = *p;
    if (const char *sk = std::getenv("POREMAPS_BATCH_SKIP2")) b->skip2 = std::atoi(sk) != 0;
    if (const char *bq = std::getenv("POREMAPS_BATCH_QUEUE")) b->queue = std::atoi(bq) != 0;
    if (const char *bg = std::getenv("POREMAPS_BATCH_QGROUP")) b->qgroup = std::max(1, std::atoi(bg));
    if (const char *bn = std::getenv("POREMAPS_BATCH_QUEUES")) b->nq = std::max(1, std::min(64, std::atoi(bn)));
    if (b->queue) {
        // the queue kernels' grid: the workgroups one wave of the lagged pass
        // keeps resident on every CU
        int per_cu = 0, cus = 0;
        hipError_t oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pm::k_b_lag4_q, pm::kB, 0);
        if (oe == hipSuccess) oe = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device);
        if (oe != hipSuccess || per_cu <= 0 || cus <= 0) b->queue = false;
        else b->qgrid = per_cu * cus;
    }
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return pm_fail(PM_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    if (p->budget_bytes > 0) {
        b->budget = p->budget_bytes;
    } else {
        size_t fr = 0, tot = 0;
        (void)hipMemGetInfo(&fr, &tot);
        b->budget = (int64_t)(0.9 * (double)fr);
    }
    *out = b;
    return PM_OK;
}

extern "C" int64_t pm_batch_sample_bytes(const pm_batch_params *p, int nx, int ny, int nz) {
    if (!p) return -1;
    pm_params q = sample_params(p, nx, ny, nz, 0, 1.0, -1.0);
    return pm_device_bytes(&q);
}

// First sweep of a sample on its own stream (and, in the defect-correction
// mode, the fold after sweep 0), so that every batch interval afterwards ends
// on an iteration index that is a multiple of it_eval.  The batch runs its own
// kernels, so the context's autotuner is off and the sweep uses the pair.
static int prologue(pm_ctx *c) {
    c->tuner = Tuner();
    c->jac_impl = c->lists.compact_jac ? JI_COMPACT : JI_PAIR;
    int rc = pm_one_sweep(c);
    if (rc == PM_OK) rc = pm_settle(c);
    if (rc == PM_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = pm_fail(PM_E_HIP, "first sweep failed");
    return rc;
}

extern "C" int pm_batch_add(pm_batch *b, const uint8_t *raw, int nx, int ny, int nz, int bc, double dx, double phi,
                            int *slot) {
    if (!b || !raw || !slot) return pm_fail(PM_E_ARG, "null argument");
    if (b->in_flight) return pm_fail(PM_E_STATE, "batch interval launched and not finished");
    pm_params q = sample_params(&b->prm, nx, ny, nz, bc, dx, phi);
    const int64_t need = pm_device_bytes(&q);
    if (need < 0) return PM_E_ARG;
    if (b->used + need > b->budget)
        return pm_fail(PM_E_MEMORY, "admission refused: %lld + %lld bytes exceed the budget of %lld",
                       (long long)b->used, (long long)need, (long long)b->budget);
    pm_ctx *c = nullptr;
    int rc = pm_ctx_create(&q, &c);
    if (rc != PM_OK) return rc;
    // each sample's layout is timed at load like a solo solve's
    // (choose_jacobi_layout); the batch's kernels read either
    rc = pm_load_geometry(c, raw);
    if (rc == PM_OK) rc = prologue(c);
    if (rc != PM_OK) {
        pm_destroy(c);
        return rc;
    }
    int s = -1;
    for (size_t i = 0; i < b->ctx.size(); i++)
        if (!b->ctx[i]) { s = (int)i; break; }
    if (s < 0) {
        s = (int)b->ctx.size();
        b->ctx.push_back(nullptr);
        b->state.push_back(0);
        b->conv.push_back(0.0);
        b->bytes.push_back(0);
        b->last.push_back(pm_log_row());
        b->n_rows.push_back(0);
    }
    b->ctx[s] = c;
    // a geometry without fluid voxels has nothing to iterate: its fields stay
    // at their initial values, so it is finished as soon as it is admitted
    b->state[s] = c->lists.n_all > 0 ? PM_SAMPLE_RUNNING : PM_SAMPLE_MAX_ITER;
    b->conv[s] = 1.0 + b->prm.eps;
    b->bytes[s] = need;
    std::memset(&b->last[s], 0, sizeof(pm_log_row));
    b->n_rows[s] = 0;
    b->used += need;
    b->shape_dirty = true;
    *slot = s;
    return PM_OK;
}

template <typename T>
static int grow(T **p, int *cap, int n) {
    if (n <= *cap) return PM_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    const int c = std::max(n, *cap * 2);
    if (hipMalloc((void **)p, sizeof(T) * (size_t)c) != hipSuccess)
        return pm_fail(PM_E_MEMORY, "batch table allocation of %d entries failed", c);
    *cap = c;
    return PM_OK;
}

// Descriptors of the running samples from their contexts' current pointers.
static void fill_desc(const pm_batch *b, std::vector<BSample> &S, std::vector<BChunk> &C, std::vector<BChunk> &G,
                      std::vector<BChunk> &P) {
    for (int slot : b->run) {
        pm_ctx *c = b->ctx[slot];
        BSample d;
        std::memset(&d, 0, sizeof d);
        d.g = c->g;
        d.list = c->lists.all;
        d.n = c->lists.n_all;
        d.G = c->G;
        d.pc = c->pc;
        for (int q = 0; q < 4; q++) {
            // the FP64 mode's two sets (the defect-correction mode uses A4 / B4)
            d.A[q] = c->X[q];
            d.B[q] = c->X2[q];
            d.X[q] = c->X[q];
        }
        d.A4 = c->S4[0];
        d.B4 = c->S4[1];
        d.R4 = c->R4;
        d.rec = c->lists.rec;
        const int *dm = c->prm.dom_interest;
        d.box.on = !(dm[0] == 0 && dm[1] == 0 && dm[2] == 0 && dm[3] == 0 && dm[4] == 0 && dm[5] == 0);
        for (int q = 0; q < 6; q++) d.box.d[q] = dm[q];
        const double om = c->prm.omega, wp = c->prm.omega_p > 0.0 ? c->prm.omega_p : om;
        d.cu = om * pm::kDt;
        d.cp = wp * pm::kC2 * pm::kDt;
        d.qA = b->dc() ? (void *)c->qD[0] : (void *)c->qX[0];
        d.qB = b->dc() ? (void *)c->qD[1] : (void *)c->qX[1];
        d.qX = c->qX[0];
        d.skip2 = b->skip2 && pm::rec_skip2_supported(d.g) ? 1 : 0;
        d.compact = c->lists.compact_jac ? 1 : 0;
        d.clink = c->lists.clink;
        d.nE = c->lists.n_centries;
        d.packed = c->lists.compact_jac && c->lists.packed ? 1 : 0;
        d.cpack = c->lists.cpack;
        d.cbase = c->lists.cbase;
        const int s = (int)S.size();
        d.chunk_begin = (int)C.size();
        for (int t = 0; t < d.n; t += pm::kB) C.push_back({s, t});
        d.chunk_end = (int)C.size();
        const int64_t lines = pm::ghost_lines(d.g);
        for (int64_t t = 0; t < lines; t += pm::kB) G.push_back({s, (int)t});
        const int plane = d.g.NY * d.g.NX;
        for (int t = 0; t < plane; t += pm::kB) P.push_back({s, t});
        S.push_back(d);
    }
}

static void launch_interval(pm_batch *b) {
    using namespace pm;
    hipStream_t st = b->stream;
    const int n = b->prm.it_eval;
    const dim3 blk(kB);
    // the relaxation factors are per sample (sample descriptors)
    if (b->dc()) {
        // symplectic Jacobi in lagged form: the first sweep from the zero
        // state, then one lagged pass per sweep (both write their cells' halo
        // copies themselves), then the settle pass
        if (b->queue) {
            // no more workgroups than there are tickets, at least one per queue
            const int nq = std::max(1, std::min(b->nq, b->nC));
            const int tickets = (b->nC + b->qgroup - 1) / b->qgroup;
            const dim3 grid((unsigned)std::max(nq, std::min(tickets, b->qgrid)));
            hipLaunchKernelGGL(k_b_zero_counters, dim3(1), blk, 0, st, b->dQ, n * nq);
            for (int s = 0; s < n; s++) {
                const int par = s & 1;
                if (s == 0)
                    hipLaunchKernelGGL(k_b_first0_q, grid, blk, 0, st, b->dS, b->dC, b->nC, b->dQ + s * nq,
                                       b->qgroup, nq, par);
                else
                    hipLaunchKernelGGL(k_b_lag4_q, grid, blk, 0, st, b->dS, b->dC, b->nC, b->dQ + s * nq,
                                       b->qgroup, nq, par);
            }
        } else {
            for (int s = 0; s < n; s++) {
                const int par = s & 1;
                if (s == 0) {
                    hipLaunchKernelGGL(k_b_first0, dim3(b->nC), blk, 0, st, b->dS, b->dC, par);
                } else {
                    if (b->nCp > 0) hipLaunchKernelGGL(k_b_lag4p, dim3(b->nCp), blk, 0, st, b->dS, b->dC, par);
                    if (b->nC > b->nCp)
                        hipLaunchKernelGGL(k_b_lag4, dim3(b->nC - b->nCp), blk, 0, st, b->dS, b->dC + b->nCp, par);
                }
            }
        }
        const int cur = n & 1;
        hipLaunchKernelGGL(k_b_settle4, dim3(b->nC), blk, 0, st, b->dS, b->dC, cur);
        hipLaunchKernelGGL(k_b_ghost4, dim3(b->nG), blk, 0, st, b->dS, b->dG, cur);
        hipLaunchKernelGGL(k_b_fold4, dim3(b->nC), blk, 0, st, b->dS, b->dC, cur);
        hipLaunchKernelGGL(k_b_ghost4, dim3(b->nG), blk, 0, st, b->dS, b->dG, cur);
        hipLaunchKernelGGL((k_b_ghost<double, true>), dim3(b->nG), blk, 0, st, b->dS, b->dG, 2);
        hipLaunchKernelGGL(k_b_outlet, dim3(b->nP), blk, 0, st, b->dS, b->dP);
        hipLaunchKernelGGL(k_b_residual4, dim3(b->nC), blk, 0, st, b->dS, b->dC);
        hipLaunchKernelGGL(k_b_gather_rq, dim3(b->nC), blk, 0, st, b->dS, b->dC);
        hipLaunchKernelGGL(k_b_zero_q, dim3(b->nP), blk, 0, st, b->dS, b->dP);
        hipLaunchKernelGGL(k_b_reduce, dim3(b->nC), blk, 0, st, b->dS, b->dC, 2, b->dPart);
    } else {
        for (int s = 0; s < n; s++) {
            const int par = s & 1;
            hipLaunchKernelGGL((k_b_vel<double>), dim3(b->nC), blk, 0, st, b->dS, b->dC, par);
            hipLaunchKernelGGL((k_b_pres<double>), dim3(b->nC), blk, 0, st, b->dS, b->dC, par);
            hipLaunchKernelGGL((k_b_ghost<double, false>), dim3(b->nG), blk, 0, st, b->dS, b->dG, 1 - par);
        }
        hipLaunchKernelGGL(k_b_reduce, dim3(b->nC), blk, 0, st, b->dS, b->dC, n & 1, b->dPart);
    }
    hipLaunchKernelGGL(k_b_reduce_final, dim3(b->nS), blk, 0, st, b->dS, b->dPart, b->dOut);
}

static void drop_graph(pm_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// After an interval with an odd number of sweeps the other buffer is current:
// swap the context's pointers as pm_one_sweep does.
static void swap_current(pm_batch *b, pm_ctx *c) {
    if (b->dc()) {
        std::swap(c->S4[0], c->S4[1]);
        std::swap(c->qD[0], c->qD[1]);
    } else {
        for (int q = 0; q < 4; q++) std::swap(c->X[q], c->X2[q]);
        std::swap(c->qX[0], c->qX[1]);
    }
}

#define NOT_IN_FLIGHT(b)                                                                         \
    do {                                                                                         \
        if ((b)->in_flight) return pm_fail(PM_E_STATE, "batch interval launched and not finished"); \
    } while (0)

extern "C" int pm_batch_step(pm_batch *b, int *n_running) {
    int rc = pm_batch_launch(b);
    return rc == PM_OK ? pm_batch_finish(b, n_running) : rc;
}

extern "C" int pm_batch_ready(pm_batch *b) {
    if (!b || !b->in_flight || b->launched_empty) return 1;
    return hipStreamQuery(b->stream) == hipErrorNotReady ? 0 : 1;
}

extern "C" int pm_batch_launch(pm_batch *b) {
    if (!b) return pm_fail(PM_E_ARG, "null batch");
    NOT_IN_FLIGHT(b);
    BCHK(hipSetDevice(b->prm.device));
    if (b->shape_dirty) {
        // the captured interval's nodes refer to the tables, the reduction
        // buffers and every running sample's arrays: it is destroyed before
        // any of them is freed, reallocated or rewritten
        drop_graph(b);
        b->run.clear();
        for (size_t s = 0; s < b->ctx.size(); s++)
            if (b->ctx[s] && b->state[s] == PM_SAMPLE_RUNNING && b->ctx[s]->lists.n_all > 0)
                b->run.push_back((int)s);
        // the samples on packed compact records first (their chunks get the
        // lagged kernel without the dense branch), then largest sample first:
        // the tile queue hands out its tiles first, the small samples' tiles
        // fill the end of each pass
        const auto packed = [b](int x) { return b->ctx[x]->lists.compact_jac && b->ctx[x]->lists.packed; };
        std::stable_sort(b->run.begin(), b->run.end(), [b, packed](int x, int y) {
            if (packed(x) != packed(y)) return packed(x);
            return b->ctx[x]->lists.n_all > b->ctx[y]->lists.n_all;
        });
    }
    if (b->run.empty()) {
        b->in_flight = b->launched_empty = true;
        return PM_OK;
    }
    std::vector<BSample> S;
    std::vector<BChunk> C, G, P;
    fill_desc(b, S, C, G, P);
    int rc;
    if ((rc = grow(&b->dS, &b->capS, (int)S.size())) || (rc = grow(&b->dC, &b->capC, (int)C.size())) ||
        (rc = grow(&b->dG, &b->capG, (int)G.size())) || (rc = grow(&b->dP, &b->capP, (int)P.size())))
        return rc;
    // reduction partials (one per chunk) and outputs (one per sample):
    // grow-only, reallocated only while no graph refers to them (sizes change
    // only with the membership, which dropped the graph above)
    if (b->capPart < b->capC) {
        if (b->dPart) (void)hipFree(b->dPart);
        b->dPart = nullptr;
        b->capPart = 0;
        BCHK(hipMalloc((void **)&b->dPart, sizeof(pm::RedAcc) * (size_t)b->capC));
        b->capPart = b->capC;
    }
    if (b->capOut < b->capS) {
        if (b->dOut) (void)hipFree(b->dOut);
        if (b->hOut) (void)hipHostFree(b->hOut);
        b->dOut = nullptr;
        b->hOut = nullptr;
        b->capOut = 0;
        BCHK(hipMalloc((void **)&b->dOut, sizeof(pm::RedSums) * (size_t)b->capS));
        BCHK(hipHostMalloc((void **)&b->hOut, sizeof(pm::RedSums) * (size_t)b->capS, hipHostMallocDefault));
        b->capOut = b->capS;
    }
    if (b->dc() && b->queue && b->capQ < b->prm.it_eval * b->nq) {
        if (b->dQ) (void)hipFree(b->dQ);
        b->dQ = nullptr;
        b->capQ = 0;
        BCHK(hipMalloc((void **)&b->dQ, sizeof(int) * (size_t)b->prm.it_eval * (size_t)b->nq));
        b->capQ = b->prm.it_eval * b->nq;
    }
    BCHK(hipMemcpyAsync(b->dS, S.data(), sizeof(BSample) * S.size(), hipMemcpyHostToDevice, b->stream));
    if (b->shape_dirty) {
        BCHK(hipMemcpyAsync(b->dC, C.data(), sizeof(BChunk) * C.size(), hipMemcpyHostToDevice, b->stream));
        BCHK(hipMemcpyAsync(b->dG, G.data(), sizeof(BChunk) * G.size(), hipMemcpyHostToDevice, b->stream));
        BCHK(hipMemcpyAsync(b->dP, P.data(), sizeof(BChunk) * P.size(), hipMemcpyHostToDevice, b->stream));
        b->nS = (int)S.size();
        b->nC = (int)C.size();
        b->nCp = 0;
        for (const BSample &d : S)
            if (d.compact && d.packed) b->nCp = d.chunk_end;
        b->nG = (int)G.size();
        b->nP = (int)P.size();
    }
    if (b->prm.use_graph) {
        if (!b->exec) {
            BCHK(hipStreamSynchronize(b->stream));
            BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            launch_interval(b);
            BCHK(hipStreamEndCapture(b->stream, &b->graph));
            BCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        BCHK(hipGraphLaunch(b->exec, b->stream));
    } else {
        launch_interval(b);
    }
    BCHK(hipGetLastError());
    BCHK(hipMemcpyAsync(b->hOut, b->dOut, sizeof(pm::RedSums) * (size_t)b->nS, hipMemcpyDeviceToHost, b->stream));
    b->shape_dirty = false;
    b->in_flight = true;
    b->launched_empty = false;
    return PM_OK;
}

extern "C" int pm_batch_finish(pm_batch *b, int *n_running) {
    if (!b) return pm_fail(PM_E_ARG, "null batch");
    if (!b->in_flight) return pm_fail(PM_E_STATE, "pm_batch_finish without pm_batch_launch");
    b->in_flight = false;
    if (b->launched_empty) {
        if (n_running) *n_running = 0;
        return PM_OK;
    }
    BCHK(hipSetDevice(b->prm.device));
    BCHK(hipStreamSynchronize(b->stream));
    int running = 0;
    for (size_t i = 0; i < b->run.size(); i++) {
        const int slot = b->run[i];
        pm_ctx *c = b->ctx[slot];
        c->t += b->prm.it_eval;
        if (b->prm.it_eval & 1) swap_current(b, c);
        const int it = c->t - 1;  // iteration index of the last sweep
        const pm::RedSums &r = b->hOut[i];
        double k[8];
        pm_permeability_from(r, c->prm.dom_interest, c->prm.voxel_size, pm_used_porosity(c), k);
        pm_log_row &row = b->last[slot];
        row.iteration = it;
        row.wmax_velz = k[6]; row.wmean_velz = k[7];
        row.k13 = k[0]; row.k23 = k[1]; row.k33 = k[2];
        row.wk13 = k[3]; row.wk23 = k[4]; row.wk33 = k[5];
        if (it % b->prm.it_eval == 0 && it > 1000) {
            b->conv[slot] = pm_convergence_from(r, &c->kprev);
            if (it % b->prm.it_write == 0) b->n_rows[slot]++;
        }
        row.conv = b->conv[slot];
        int st = PM_SAMPLE_RUNNING;
        if (!std::isfinite(b->conv[slot])) {
            // relaxation beyond the stability bound: this sample starts over
            // with a smaller pressure relaxation (the others are unaffected)
            if (pm_restart_slower(c) == PM_OK && prologue(c) == PM_OK) {
                b->conv[slot] = 1.0 + c->prm.eps;
                b->shape_dirty = true;
                running++;
                continue;
            }
            st = PM_SAMPLE_DIVERGED;
        } else if (b->conv[slot] < c->prm.eps) st = PM_SAMPLE_CONVERGED;
        else if (c->t >= b->prm.max_iter) st = PM_SAMPLE_MAX_ITER;
        b->state[slot] = st;
        if (st != PM_SAMPLE_RUNNING) b->shape_dirty = true;
        else running++;
    }
    if (n_running) *n_running = running;
    return PM_OK;
}

extern "C" int pm_batch_set_eps(pm_batch *b, int slot, double eps) {
    if (!b || slot < 0 || slot >= (int)b->ctx.size() || !b->ctx[slot]) return pm_fail(PM_E_ARG, "no sample in slot %d", slot);
    NOT_IN_FLIGHT(b);
    if (!(eps > 0.0)) return pm_fail(PM_E_ARG, "eps must be positive");
    b->ctx[slot]->prm.eps = eps;
    return PM_OK;
}

extern "C" int pm_batch_set_dom_interest(pm_batch *b, int slot, const int box[6]) {
    if (!b || slot < 0 || slot >= (int)b->ctx.size() || !b->ctx[slot]) return pm_fail(PM_E_ARG, "no sample in slot %d", slot);
    NOT_IN_FLIGHT(b);
    if (!box) return pm_fail(PM_E_ARG, "null box");
    pm_ctx *c = b->ctx[slot];
    const int n[3] = {c->prm.nx, c->prm.ny, c->prm.nz};
    const bool off = box[0] == 0 && box[1] == 0 && box[2] == 0 && box[3] == 0 && box[4] == 0 && box[5] == 0;
    for (int a = 0; a < 3 && !off; a++)
        if (box[2 * a] < 0 || box[2 * a] > box[2 * a + 1] || box[2 * a + 1] >= n[a])
            return pm_fail(PM_E_ARG, "dom_interest %d %d outside axis %d of %d cells", box[2 * a], box[2 * a + 1], a,
                           n[a]);
    for (int q = 0; q < 6; q++) c->prm.dom_interest[q] = box[q];
    b->shape_dirty = true;  // the sample descriptors carry the box
    return PM_OK;
}

extern "C" int pm_batch_status(pm_batch *b, int slot, int *state, pm_result *res) {
    if (!b || slot < 0 || slot >= (int)b->ctx.size() || !b->ctx[slot]) return pm_fail(PM_E_ARG, "no sample in slot %d", slot);
    NOT_IN_FLIGHT(b);
    pm_ctx *c = b->ctx[slot];
    if (state) *state = b->state[slot];
    if (res) {
        std::memset(res, 0, sizeof *res);
        res->sweeps = c->t;
        res->conv = b->conv[slot];
        res->porosity = pm_used_porosity(c);
        res->n_log_rows = b->n_rows[slot];
        res->last = b->last[slot];
        res->restarts = c->restarts;
    }
    return PM_OK;
}

extern "C" int pm_batch_get_field(pm_batch *b, int slot, int field, void *out) {
    if (!b || slot < 0 || slot >= (int)b->ctx.size() || !b->ctx[slot]) return pm_fail(PM_E_ARG, "no sample in slot %d", slot);
    NOT_IN_FLIGHT(b);
    BCHK(hipStreamSynchronize(b->stream));
    return pm_get_field(b->ctx[slot], field, out);
}

extern "C" int pm_batch_remove(pm_batch *b, int slot) {
    if (!b || slot < 0 || slot >= (int)b->ctx.size() || !b->ctx[slot]) return pm_fail(PM_E_ARG, "no sample in slot %d", slot);
    NOT_IN_FLIGHT(b);
    BCHK(hipStreamSynchronize(b->stream));
    // the captured interval refers to this sample's arrays
    drop_graph(b);
    pm_destroy(b->ctx[slot]);
    b->ctx[slot] = nullptr;
    b->used -= b->bytes[slot];
    b->bytes[slot] = 0;
    b->shape_dirty = true;
    return PM_OK;
}

extern "C" int pm_batch_memory(pm_batch *b, int64_t *used, int64_t *budget) {
    if (!b) return pm_fail(PM_E_ARG, "null batch");
    if (used) *used = b->used;
    if (budget) *budget = b->budget;
    return PM_OK;
}

extern "C" void pm_batch_destroy(pm_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->prm.device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (pm_ctx *c : b->ctx)
        if (c) pm_destroy(c);
    if (b->dS) (void)hipFree(b->dS);
    if (b->dC) (void)hipFree(b->dC);
    if (b->dG) (void)hipFree(b->dG);
    if (b->dP) (void)hipFree(b->dP);
    if (b->dPart) (void)hipFree(b->dPart);
    if (b->dOut) (void)hipFree(b->dOut);
    if (b->hOut) (void)hipHostFree(b->hOut);
    if (b->dQ) (void)hipFree(b->dQ);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}
