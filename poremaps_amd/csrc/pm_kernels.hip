// pm_kernels.hip -- HIP kernels of the POREMAPS Stokes solver for gfx950.
//
// Reference mode: the in-place lexicographic sweep of src/solver.cc:112-223,
// reproduced exactly by updating the hyperplanes i+j+k = s in increasing order
// (every voxel reads new values only from its -1/-2 neighbours, whose i+j+k is
// smaller, and old values only from +1/+2 neighbours; halos are frozen during
// a sweep, src/calc_flow.cc:109-117).  FP64, no contraction, divisions kept.
//
// Fast modes: symplectic Jacobi (out of place) and 3-colour SOR (in place) in
// FP64 or FP32, with a one-pass ghost fill, device reductions and the FP64
// defect-correction passes.
#include "pm_kernels.h"
#include "pm_reduce.h"

#include <algorithm>

#pragma clang fp contract(off)

namespace pm {

// ---------------------------------------------------------------------------
// geometry and initial state
// ---------------------------------------------------------------------------

// src/geometry.cc:98-115: interior bytes shifted by the halo, everything else 0
// (fluid).  The build's opt-in walls on the non-periodic x / y faces (periodic
// halos are copied by the exchange afterwards): walls = 1, no-slip, makes
// those halo layers solid; walls = 2, slip (symmetry planes), fills them with
// the mirror image of the first two interior layers (k = 1, 0 from 2, 3;
// nx + 2, nx + 3 from nx + 1, nx; rows likewise).
__global__ void k_geom_fill(Grid g, const uint8_t *__restrict__ raw, uint8_t *__restrict__ G, int walls) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int i = blockIdx.z;
    if (k >= g.NX) return;
    const bool hx = k < 2 || k >= g.nx + 2, hy = j < 2 || j >= g.ny + 2;
    int ks = k, js = j;
    if (walls == 2) {
        if (!g.px && hx) ks = k < 2 ? 3 - k : 2 * g.nx + 3 - k;
        if (!g.py && hy) js = j < 2 ? 3 - j : 2 * g.ny + 3 - j;
    }
    const bool in = i >= 2 && i < g.nz + 2 && js >= 2 && js < g.ny + 2 && ks >= 2 && ks < g.nx + 2;
    const uint8_t halo = (walls == 1 && ((hx && !g.px) || (hy && !g.py))) ? 1 : 0;
    G[g.at(i, j, k)] = in ? raw[((int64_t)(i - 2) * g.ny + (js - 2)) * g.nx + (ks - 2)] : halo;
}

// Two-layer periodic copy along one axis over the full extent of the other
// two (src/parallelization.cc:67-190, :193-318 at one rank).  Up to four
// fields of one type; null pointers are skipped.
template <typename T>
__global__ void k_exchange_axis(Grid g, int axis, T *f0, T *f1, T *f2, T *f3) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    T *fs[4] = {f0, f1, f2, f3};
    if (axis == 0) {
        if (t >= (int64_t)g.NZ * g.NY) return;
        const int i = t / g.NY, j = t % g.NY;
        const int64_t r = g.at(i, j, 0);
        for (int q = 0; q < 4; q++) {
            T *f = fs[q];
            if (!f) continue;
            const T a0 = f[r + 2], a1 = f[r + 3], b0 = f[r + g.NX - 4], b1 = f[r + g.NX - 3];
            f[r + g.NX - 2] = a0; f[r + g.NX - 1] = a1; f[r] = b0; f[r + 1] = b1;
        }
    } else if (axis == 1) {
        if (t >= (int64_t)g.NZ * g.NX) return;
        const int i = t / g.NX, k = t % g.NX;
        for (int q = 0; q < 4; q++) {
            T *f = fs[q];
            if (!f) continue;
            const T a0 = f[g.at(i, 2, k)], a1 = f[g.at(i, 3, k)];
            const T b0 = f[g.at(i, g.NY - 4, k)], b1 = f[g.at(i, g.NY - 3, k)];
            f[g.at(i, g.NY - 2, k)] = a0; f[g.at(i, g.NY - 1, k)] = a1;
            f[g.at(i, 0, k)] = b0; f[g.at(i, 1, k)] = b1;
        }
    } else {
        if (t >= (int64_t)g.NY * g.NX) return;
        const int j = t / g.NX, k = t % g.NX;
        for (int q = 0; q < 4; q++) {
            T *f = fs[q];
            if (!f) continue;
            const T a0 = f[g.at(2, j, k)], a1 = f[g.at(3, j, k)];
            const T b0 = f[g.at(g.NZ - 4, j, k)], b1 = f[g.at(g.NZ - 3, j, k)];
            f[g.at(g.NZ - 2, j, k)] = a0; f[g.at(g.NZ - 1, j, k)] = a1;
            f[g.at(0, j, k)] = b0; f[g.at(1, j, k)] = b1;
        }
    }
}

// src/geometry.cc:174-302: per-axis wall case of every interior fluid cell,
// code = 100*cx + 10*cy + cz; solid -1, other cells 0.  Also the packed
// uint8 code the fast kernels read ((cx-1)*36 + (cy-1)*6 + (cz-1), 255 solid).
__device__ __forceinline__ int axis_case(const uint8_t *G, int64_t c, int64_t s) {
    const int l = G[c - s] == 0 ? 0 : 1;
    const int r = G[c + s] == 0 ? 0 : 1;
    int ll = -1, rr = -1;
    if (l == 0) ll = G[c - 2 * s] == 0 ? 0 : 1;
    if (r == 0) rr = G[c + 2 * s] == 0 ? 0 : 1;
    if (l == 1 && r == 1) return 1;
    if (l == 1 && r == 0 && rr == 1) return 2;
    if (ll == 1 && l == 0 && r == 1) return 3;
    if (l == 1 && r == 0 && rr == 0) return 4;
    if (ll == 0 && l == 0 && r == 1) return 5;
    if (l == 0 && r == 0) return 6;
    return 0;
}

__global__ void k_eval_geometry(Grid g, const uint8_t *__restrict__ G, int32_t *__restrict__ code,
                                uint8_t *__restrict__ pc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int i = blockIdx.z;
    if (k >= g.NX) return;
    const int64_t c = g.at(i, j, k);
    const bool in = i >= 2 && i < g.nz + 2 && j >= 2 && j < g.ny + 2 && k >= 2 && k < g.nx + 2;
    if (G[c] != 0) { code[c] = -1; pc[c] = kSolid; return; }
    if (!in) { code[c] = 0; pc[c] = kSolid; return; }
    const int cx = axis_case(G, c, 1), cy = axis_case(G, c, g.sx), cz = axis_case(G, c, g.sy);
    code[c] = 100 * cx + 10 * cy + cz;
    pc[c] = (uint8_t)((cx - 1) * 36 + (cy - 1) * 6 + (cz - 1));
}

// src/initialize.cc:69-112: u = v = w = 0, p = -1; interior fluid p is linear
// in z ((nz - i) + 4 at one rank), interior solid p = 0.
__global__ void k_init_fields(Grid g, const uint8_t *__restrict__ G, double *p, double *u,
                              double *v, double *w) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int i = blockIdx.z;
    if (k >= g.NX) return;
    const int64_t c = g.at(i, j, k);
    const bool in = i >= 2 && i < g.nz + 2 && j >= 2 && j < g.ny + 2 && k >= 2 && k < g.nx + 2;
    u[c] = 0.0; v[c] = 0.0; w[c] = 0.0;
    p[c] = in ? (G[c] == 0 ? (double)((g.nz - 0 - i) + 2 + 2) : 0.0) : -1.0;
}

// src/boundary_conditions.cc:46-78 at one rank.
template <typename T>
__global__ void k_pins(Grid g, const uint8_t *__restrict__ G, T *p, T pin_in, T pin_hi,
                       T pin_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int j = blockIdx.y + 2;
    if (k >= g.nx + 2) return;
    const int64_t c1 = g.at(1, j, k), chi = g.at(g.NZ - 2, j, k), cout = g.at(g.NZ - 3, j, k);
    if (G[c1] == 0) p[c1] = pin_in;
    if (G[chi] == 0) p[chi] = pin_hi;
    if (G[cout] == 0) p[cout] = pin_out;
}

// ---------------------------------------------------------------------------
// reference mode (exact lexicographic order)
// ---------------------------------------------------------------------------

// src/solver.cc:136-219 for one voxel, expression for expression, at element
// c of arrays with row pitch sj and plane pitch si (the padded global field,
// or a tile copied into LDS).
template <typename I>
__device__ __forceinline__ void ref_update(const uint8_t *G, const int32_t *code, double *p, double *u, double *v,
                                           double *w, int solver, I c, I sj, I si) {
    if (G[c] != 0) return;
    double dp_dz, dp_dy, dp_dx;
    if (G[c - si] == 0) dp_dz = -p[c - si] + p[c]; else dp_dz = 0.;
    if (G[c - sj] == 0) dp_dy = -p[c - sj] + p[c]; else dp_dy = 0.;
    if (G[c - 1] == 0) dp_dx = -p[c - 1] + p[c]; else dp_dx = 0.;
    int num = code[c];
    const int cz = num % 10; num /= 10;
    const int cy = num % 10; num /= 10;
    const int cx = num % 10;
    const double d2vxdx2 = (u[c - 1] - 2 * u[c] + u[c + 1]);
    const double d2vydy2 = (v[c - sj] - 2 * v[c] + v[c + sj]);
    const double d2vzdz2 = (w[c - si] - 2 * w[c] + w[c + si]);
    const double d2vydx2 = d2<double>(cx, v[c - 2], v[c - 1], v[c], v[c + 1], v[c + 2]);
    const double d2vzdx2 = d2<double>(cx, w[c - 2], w[c - 1], w[c], w[c + 1], w[c + 2]);
    const double d2vxdy2 = d2<double>(cy, u[c - 2 * sj], u[c - sj], u[c], u[c + sj], u[c + 2 * sj]);
    const double d2vzdy2 = d2<double>(cy, w[c - 2 * sj], w[c - sj], w[c], w[c + sj], w[c + 2 * sj]);
    const double d2vxdz2 = d2<double>(cz, u[c - 2 * si], u[c - si], u[c], u[c + si], u[c + 2 * si]);
    const double d2vydz2 = d2<double>(cz, v[c - 2 * si], v[c - si], v[c], v[c + si], v[c + 2 * si]);
    const double velx_sum = d2vxdx2 + d2vxdy2 + d2vxdz2;
    const double vely_sum = d2vydy2 + d2vydz2 + d2vydx2;
    const double velz_sum = d2vzdz2 + d2vzdx2 + d2vzdy2;
    double v_sum = 0.0;
    if (solver == 1)
        v_sum = (w[c + si] - w[c] + v[c + sj] - v[c] + u[c + 1] - u[c]);
    if (solver == 1 || solver == 2) {
        if (G[c - 1] == 1) u[c] = 0.0;
        else u[c] = u[c] + (-dp_dx + velx_sum / kRe) * kDt;
        if (G[c - sj] == 1) v[c] = 0.0;
        else v[c] = v[c] + (-dp_dy + vely_sum / kRe) * kDt;
        if (G[c - si] == 1) w[c] = 0.0;
        else w[c] = w[c] + (-dp_dz + velz_sum / kRe) * kDt;
    } else if (solver == 3) {
        u[c] = (1.0 - kOmegaRef) * u[c] + kOmegaRef * (-dp_dx + velx_sum / kRe) * kDt;
        v[c] = (1.0 - kOmegaRef) * v[c] + kOmegaRef * (-dp_dy + vely_sum / kRe) * kDt;
        w[c] = (1.0 - kOmegaRef) * w[c] + kOmegaRef * (-dp_dz + velz_sum / kRe) * kDt;
    }
    if (solver == 2 || solver == 3)
        v_sum = (w[c + si] - w[c] + v[c + sj] - v[c] + u[c + 1] - u[c]);
    p[c] = p[c] - kC2 * v_sum * kDt;
}

// ref_update split by field for the tiled wavefront: COMP 0/1/2 updates u/v/w
// of cell c (each reads only its own field, the pressure and the geometry,
// so the three run side by side on different waves), ref_pressure then p from
// the new velocities (solvers 2, 3) or from the divergence vs_old taken
// before any velocity of the cell changed (solver 1).  The same expressions
// in the same order as ref_update, the case switch of get_2nd_derivation
// replaced by its select form (same arithmetic), so the same values.
template <int COMP>
__device__ __forceinline__ void ref_velocity(const uint8_t *G, const int32_t *code, const double *p, double *f,
                                             int solver, int c, int sj, int si) {
    const int s = COMP == 0 ? 1 : COMP == 1 ? sj : si;  // the component's own axis
    double dp;
    if (G[c - s] == 0) dp = -p[c - s] + p[c]; else dp = 0.;
    int num = code[c];
    const int cz = num % 10; num /= 10;
    const int cy = num % 10; num /= 10;
    const int cx = num % 10;
    // own-axis Laplacian term, then the other axes' get_2nd_derivation in the
    // reference's summation order: u: x, y, z; v: y, z, x; w: z, x, y
    const double own = (f[c - s] - 2 * f[c] + f[c + s]);
    const int ca = COMP == 0 ? cy : COMP == 1 ? cz : cx;   // first other axis
    const int cb = COMP == 0 ? cz : COMP == 1 ? cx : cy;   // second other axis
    const int sa = COMP == 0 ? sj : COMP == 1 ? si : 1;
    const int sb = COMP == 0 ? si : COMP == 1 ? 1 : sj;
    const double da = d2_pick<double>(ca, f[c - 2 * sa], f[c - sa], f[c], f[c + sa], f[c + 2 * sa]);
    const double db = d2_pick<double>(cb, f[c - 2 * sb], f[c - sb], f[c], f[c + sb], f[c + 2 * sb]);
    const double sum = own + da + db;
    if (solver == 1 || solver == 2) {
        if (G[c - s] == 1) f[c] = 0.0;
        else f[c] = f[c] + (-dp + sum / kRe) * kDt;
    } else if (solver == 3) {
        f[c] = (1.0 - kOmegaRef) * f[c] + kOmegaRef * (-dp + sum / kRe) * kDt;
    }
}

__device__ __forceinline__ double ref_divergence(const double *u, const double *v, const double *w, int c, int sj,
                                                 int si) {
    return (w[c + si] - w[c] + v[c + sj] - v[c] + u[c + 1] - u[c]);
}

__device__ __forceinline__ void ref_voxel(const Grid &g, const uint8_t *__restrict__ G,
                                          const int32_t *__restrict__ code, double *p, double *u,
                                          double *v, double *w, int solver, int i, int j, int k) {
    ref_update<int64_t>(G, code, p, u, v, w, solver, g.at(i, j, k), g.sx, g.sy);
}

// Hyperplane s of the interior (0-based coordinates a=z, b=y, c=x, a+b+c=s).
__device__ __forceinline__ void hyperplane_range(const Grid &g, int s, int &a_lo, int &a_hi) {
    a_lo = max(0, s - (g.ny - 1) - (g.nx - 1));
    a_hi = min(g.nz - 1, s);
}

// One hyperplane per launch (large domains).
__global__ void k_ref_hyperplane(Grid g, const uint8_t *__restrict__ G,
                                 const int32_t *__restrict__ code, double *p, double *u,
                                 double *v, double *w, int solver, int s) {
    int a_lo, a_hi;
    hyperplane_range(g, s, a_lo, a_hi);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)(a_hi - a_lo + 1) * g.ny) return;
    const int a = a_lo + (int)(t / g.ny), b = (int)(t % g.ny), c = s - a - b;
    if (c < 0 || c >= g.nx) return;
    ref_voxel(g, G, code, p, u, v, w, solver, a + 2, b + 2, c + 2);
}

// Tiled wavefront of the exact-order sweep.  The interior is cut into kRT^3
// tiles; tile (A, B, C) lies on tile-diagonal D = A + B + C.  The stencil
// reads only axis-aligned neighbours (+-1, +-2 on one axis), so a tile touches
// only the tiles next to it on each axis, which lie on diagonals D - 1 and
// D + 1: the tiles of one diagonal are independent, and running the diagonals
// in order gives every cell new values from its -1/-2 neighbours and old ones
// from its +1/+2 neighbours -- the reference's lexicographic sweep exactly.
// Within a tile the cells go by local hyperplanes.  The tile and its 2-cell
// rim are copied into LDS first: the rim belongs to tiles of diagonal D - 1
// (finished), of diagonal D + 1 (not started) or to the halo (frozen during a
// sweep), so it does not change while the tile runs.  One workgroup of four
// waves per tile: all copy; per local hyperplane three waves update u, v, w
// side by side (one (z, y) line of the tile per lane), then the first one the
// pressure.  54 x 54 x 50 tube: 0.40 ms per sweep (one wave computing every
// field: 0.62 ms; one workgroup for the whole domain: 2.03 ms; one launch per
// hyperplane: 0.76 ms); 128^3 sphere pack 1.15 ms, 256^3 4.05 ms.
constexpr int kRT = 8;                   // tile edge
constexpr int kRB = kRT + 4;             // tile + rim
constexpr int kRN = kRB * kRB * kRB;     // 1728 cells, 63.9 KB of LDS

constexpr int kRThreads = 4 * kRT * kRT;  // 4 waves copy, the first one computes

// Tile (A, B, C) of the box lo + [0, n) (interior coordinates, x, y, z) of
// the grid g (whole workgroup; the caller checked that the tile exists).
// Rim cells inside the box come from the live fields f, those outside from
// fo: for one domain (the box is the interior, fo = f) they are the halo,
// frozen during a sweep; for a block of a decomposition they are the other
// blocks' cells and the halo as they were when the sweep started (the
// block's own halo copies in the reference's multi-rank run).
__device__ __forceinline__ void ref_tile(const Grid &g, const uint8_t *__restrict__ G,
                                         const int32_t *__restrict__ code, const RefFields &f, const RefFields &fo,
                                         const RefBox &box, int solver, int A, int B, int C) {
    __shared__ double Lp[kRN], Lu[kRN], Lv[kRN], Lw[kRN];
    __shared__ int32_t Lc[kRN];
    __shared__ uint8_t LG[kRN];
    // tile origin in interior coordinates = origin of its rim box in padded ones
    const int i0 = box.lo[2] + A * kRT, j0 = box.lo[1] + B * kRT, k0 = box.lo[0] + C * kRT;
    const int tz = min(kRT, box.n[2] - A * kRT), ty = min(kRT, box.n[1] - B * kRT),
              tx = min(kRT, box.n[0] - C * kRT);
    const int by = ty + 4, bx = tx + 4, nbox = (tz + 4) * by * bx;
    for (int t = threadIdx.x; t < nbox; t += blockDim.x) {
        const int a = t / (by * bx), r = t - a * by * bx, b = r / bx, x = r - b * bx;
        const int64_t gc = g.at(i0 + a, j0 + b, k0 + x);
        const int l = (a * kRB + b) * kRB + x;
        // interior coordinates of the cell: i0 + a - 2, ...
        const bool in = (unsigned)(i0 + a - 2 - box.lo[2]) < (unsigned)box.n[2] &&
                        (unsigned)(j0 + b - 2 - box.lo[1]) < (unsigned)box.n[1] &&
                        (unsigned)(k0 + x - 2 - box.lo[0]) < (unsigned)box.n[0];
        const RefFields &src = in ? f : fo;
        Lp[l] = src.f[0][gc]; Lu[l] = src.f[1][gc]; Lv[l] = src.f[2][gc]; Lw[l] = src.f[3][gc];
        Lc[l] = code[gc]; LG[l] = G[gc];
    }
    __syncthreads();
    // each of the first three waves takes one velocity field of the tile's
    // current local hyperplane, one (z, y) line of the tile per lane; then
    // the first wave updates the pressure from the new velocities.  The
    // fourth wave only helps with the copies.
    const int q = (int)threadIdx.x / (kRT * kRT), l = (int)threadIdx.x % (kRT * kRT);
    const int a = l / kRT, b = l % kRT;
    const int smax = (tz - 1) + (ty - 1) + (tx - 1);
    constexpr int sj = kRB, si = kRB * kRB;
    for (int sh = 0; sh <= smax; sh++) {
        const int x = sh - a - b;
        const int c = ((a + 2) * kRB + (b + 2)) * kRB + (x + 2);
        const bool fl = a < tz && b < ty && x >= 0 && x < tx && LG[c] == 0;
        double vs = 0.0;
        if (solver == 1) {
            // src/solver.cc: the divergence before the cell's velocities change
            if (q == 0 && fl) vs = ref_divergence(Lu, Lv, Lw, c, sj, si);
            __syncthreads();
        }
        if (fl) {
            if (q == 0) ref_velocity<0>(LG, Lc, Lp, Lu, solver, c, sj, si);
            else if (q == 1) ref_velocity<1>(LG, Lc, Lp, Lv, solver, c, sj, si);
            else if (q == 2) ref_velocity<2>(LG, Lc, Lp, Lw, solver, c, sj, si);
        }
        __syncthreads();
        if (q == 0 && fl) {
            if (solver == 2 || solver == 3) vs = ref_divergence(Lu, Lv, Lw, c, sj, si);
            Lp[c] = Lp[c] - kC2 * vs * kDt;
        }
        __syncthreads();
    }
    const int ntile = tz * ty * tx;
    for (int t = threadIdx.x; t < ntile; t += blockDim.x) {
        const int a2 = t / (ty * tx), r = t - a2 * ty * tx, b2 = r / tx, x2 = r - b2 * tx;
        const int64_t gc = g.at(i0 + 2 + a2, j0 + 2 + b2, k0 + 2 + x2);
        const int l = ((a2 + 2) * kRB + (b2 + 2)) * kRB + (x2 + 2);
        f.f[0][gc] = Lp[l]; f.f[1][gc] = Lu[l]; f.f[2][gc] = Lv[l]; f.f[3][gc] = Lw[l];
    }
}

__global__ __launch_bounds__(kRThreads) void k_ref_tile(Grid g, const uint8_t *__restrict__ G,
                                                       const int32_t *__restrict__ code, RefFields f, int solver,
                                                       int D, int a0, int nty, int ntx) {
    const int A = a0 + (int)blockIdx.x / nty, B = (int)blockIdx.x % nty, C = D - A - B;
    if (C < 0 || C >= ntx) return;  // the same for the whole workgroup
    const RefBox box = {{0, 0, 0}, {g.nx, g.ny, g.nz}};
    ref_tile(g, G, code, f, f, box, solver, A, B, C);
}

// The tiles of tile-diagonal D of every block of a decomposition at once (the
// blocks are independent within a sweep, each swept against the state fo
// taken when the sweep started): workgroup -> (block, A, B), C = D - A - B,
// tile counts per axis up to ntz, nty of the largest block.
__global__ __launch_bounds__(kRThreads) void k_ref_tile_blocks(Grid g, const uint8_t *__restrict__ G,
                                                              const int32_t *__restrict__ code, RefFields f,
                                                              RefFields fo, const RefBox *__restrict__ tab,
                                                              int solver, int D, int ntz, int nty) {
    const int per = ntz * nty;
    const int k = (int)blockIdx.x / per, A = ((int)blockIdx.x / nty) % ntz, B = (int)blockIdx.x % nty;
    const RefBox box = tab[k];
    const int C = D - A - B;
    if (A >= (box.n[2] + kRT - 1) / kRT || B >= (box.n[1] + kRT - 1) / kRT || C < 0 ||
        C >= (box.n[0] + kRT - 1) / kRT)
        return;  // the same for the whole workgroup
    ref_tile(g, G, code, f, fo, box, solver, A, B, C);
}

// Whole sweeps in one workgroup (small domains): hyperplanes, then the
// reference's exchange X, Y, Z of p, u, v, w, then the pins -- the exact
// order of src/calc_flow.cc:109-117.
__global__ __launch_bounds__(1024) void k_ref_sweeps_1wg(Grid g, const uint8_t *__restrict__ G,
                                                         const int32_t *__restrict__ code,
                                                         double *p, double *u, double *v,
                                                         double *w, int solver, int nsweeps) {
    const int smax = (g.nx - 1) + (g.ny - 1) + (g.nz - 1);
    double *fs[4] = {p, u, v, w};
    for (int it = 0; it < nsweeps; it++) {
        for (int s = 0; s <= smax; s++) {
            int a_lo, a_hi;
            hyperplane_range(g, s, a_lo, a_hi);
            const int n = (a_hi - a_lo + 1) * g.ny;
            for (int t = threadIdx.x; t < n; t += blockDim.x) {
                const int a = a_lo + t / g.ny, b = t % g.ny, c = s - a - b;
                if (c >= 0 && c < g.nx) ref_voxel(g, G, code, p, u, v, w, solver, a + 2, b + 2, c + 2);
            }
            __syncthreads();
        }
        if (g.px) {
            for (int t = threadIdx.x; t < g.NZ * g.NY; t += blockDim.x) {
                const int64_t r = g.at(t / g.NY, t % g.NY, 0);
                for (int q = 0; q < 4; q++) {
                    double *f = fs[q];
                    const double a0 = f[r + 2], a1 = f[r + 3], b0 = f[r + g.NX - 4], b1 = f[r + g.NX - 3];
                    f[r + g.NX - 2] = a0; f[r + g.NX - 1] = a1; f[r] = b0; f[r + 1] = b1;
                }
            }
            __syncthreads();
        }
        if (g.py) {
            for (int t = threadIdx.x; t < g.NZ * g.NX; t += blockDim.x) {
                const int i = t / g.NX, k = t % g.NX;
                for (int q = 0; q < 4; q++) {
                    double *f = fs[q];
                    const double a0 = f[g.at(i, 2, k)], a1 = f[g.at(i, 3, k)];
                    const double b0 = f[g.at(i, g.NY - 4, k)], b1 = f[g.at(i, g.NY - 3, k)];
                    f[g.at(i, g.NY - 2, k)] = a0; f[g.at(i, g.NY - 1, k)] = a1;
                    f[g.at(i, 0, k)] = b0; f[g.at(i, 1, k)] = b1;
                }
            }
            __syncthreads();
        }
        if (g.pz) {
            for (int t = threadIdx.x; t < g.NY * g.NX; t += blockDim.x) {
                const int j = t / g.NX, k = t % g.NX;
                for (int q = 0; q < 4; q++) {
                    double *f = fs[q];
                    const double a0 = f[g.at(2, j, k)], a1 = f[g.at(3, j, k)];
                    const double b0 = f[g.at(g.NZ - 4, j, k)], b1 = f[g.at(g.NZ - 3, j, k)];
                    f[g.at(g.NZ - 2, j, k)] = a0; f[g.at(g.NZ - 1, j, k)] = a1;
                    f[g.at(0, j, k)] = b0; f[g.at(1, j, k)] = b1;
                }
            }
            __syncthreads();
        }
        const double pin_in = (double)(g.nz + 2);
        for (int t = threadIdx.x; t < g.ny * g.nx; t += blockDim.x) {
            const int j = t / g.nx + 2, k = t % g.nx + 2;
            const int64_t c1 = g.at(1, j, k);
            if (G[c1] == 0) p[c1] = pin_in;
        }
        for (int t = threadIdx.x; t < g.ny * g.nx; t += blockDim.x) {
            const int j = t / g.nx + 2, k = t % g.nx + 2;
            const int64_t chi = g.at(g.NZ - 2, j, k), cout = g.at(g.NZ - 3, j, k);
            if (G[chi] == 0) p[chi] = 2.0;
            if (G[cout] == 0) p[cout] = 3.0;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// fast modes
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// reductions (deterministic, compensated FP64)
// ---------------------------------------------------------------------------


// FULL = false: only the sum of w and the fluid count (compute_convergence).
template <bool FULL>
__global__ __launch_bounds__(kRedThreads) void k_reduce_state(
    Grid g, const uint8_t *__restrict__ G, const double *__restrict__ p,
    const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ w,
    DomBox box, RedAcc *__restrict__ partial) {
    RedAcc a;
    acc_zero(a);
    // 64 lanes along x, 4 rows per block step: coalesced, no per-cell
    // division.  The operands are loaded whatever the cell and a solid cell
    // adds 0.0, which leaves a compensated sum exactly as it was: the loads of
    // a row do not wait for its geometry bytes.
    const int lane = threadIdx.x & 63, sub = threadIdx.x >> 6;
    const int64_t rows = (int64_t)g.ny * g.nz;
    for (int64_t r = (int64_t)blockIdx.x * 4 + sub; r < rows; r += (int64_t)gridDim.x * 4) {
        const int j = (int)(r % g.ny), i = (int)(r / g.ny);
        const int64_t row = g.at(i + 2, j + 2, 2);
#pragma unroll 4
    for (int k = lane; k < g.nx; k += 64) {
        const int64_t c = row + k;
        const bool fl = G[c] == 0;
        const double wc = w[c];
        const double uc = FULL ? u[c] : 0.0, vc = FULL ? v[c] : 0.0;
        a.n[0] += fl ? 1 : 0;
        nadd(a.s[2], fl ? wc : 0.0);
        if (FULL) {
            nadd(a.s[0], fl ? uc : 0.0);
            nadd(a.s[1], fl ? vc : 0.0);
            if (fl && wc > a.wmax) a.wmax = wc;
            if (fl && box.on && k >= box.d[0] && k <= box.d[1] && j >= box.d[2] && j <= box.d[3] &&
                i >= box.d[4] && i <= box.d[5]) {
                a.n[1] += 1;
                nadd(a.s[3], uc);
                nadd(a.s[4], vc);
                nadd(a.s[5], wc);
                if (i == box.d[4]) { a.n[2] += 1; nadd(a.s[6], p[c]); }
                if (i == box.d[5]) { a.n[3] += 1; nadd(a.s[7], p[c]); }
            }
        }
    }
    }
    block_reduce_store(a, partial + blockIdx.x);
}

__global__ __launch_bounds__(kRedThreads) void k_reduce_final(const RedAcc *__restrict__ partial,
                                                              int nparts, RedSums *__restrict__ out) {
    RedAcc a;
    acc_zero(a);
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) acc_merge(a, partial[b]);
    __shared__ RedAcc res;
    block_reduce_store(a, &res);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 8; q++) out->sum[q] = res.s[q].s + res.s[q].c;
        for (int q = 0; q < 4; q++) out->count[q] = res.n[q];
        out->wmax = res.wmax;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline dim3 row_grid(const Grid &g) {
    return dim3((unsigned)((g.NX + 63) / 64), (unsigned)g.NY, (unsigned)g.NZ);
}
static inline unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_geom_fill(const Grid &g, const uint8_t *raw, uint8_t *G, int walls, hipStream_t s) {
    hipLaunchKernelGGL(k_geom_fill, row_grid(g), dim3(64), 0, s, g, raw, G, walls);
}

template <typename T>
void launch_exchange_full(const Grid &g, T *f0, T *f1, T *f2, T *f3, hipStream_t s) {
    if (g.px)
        hipLaunchKernelGGL(k_exchange_axis<T>, dim3(nblocks((int64_t)g.NZ * g.NY, 256)), dim3(256), 0, s,
                           g, 0, f0, f1, f2, f3);
    if (g.py)
        hipLaunchKernelGGL(k_exchange_axis<T>, dim3(nblocks((int64_t)g.NZ * g.NX, 256)), dim3(256), 0, s,
                           g, 1, f0, f1, f2, f3);
    if (g.pz)
        hipLaunchKernelGGL(k_exchange_axis<T>, dim3(nblocks((int64_t)g.NY * g.NX, 256)), dim3(256), 0, s,
                           g, 2, f0, f1, f2, f3);
}
template void launch_exchange_full<uint8_t>(const Grid &, uint8_t *, uint8_t *, uint8_t *, uint8_t *, hipStream_t);
template void launch_exchange_full<double>(const Grid &, double *, double *, double *, double *, hipStream_t);
template void launch_exchange_full<float>(const Grid &, float *, float *, float *, float *, hipStream_t);

void launch_eval_geometry(const Grid &g, const uint8_t *G, int32_t *code, uint8_t *pc, hipStream_t s) {
    hipLaunchKernelGGL(k_eval_geometry, row_grid(g), dim3(64), 0, s, g, G, code, pc);
}

void launch_init_fields(const Grid &g, const uint8_t *G, double *p, double *u, double *v, double *w,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_init_fields, row_grid(g), dim3(64), 0, s, g, G, p, u, v, w);
    launch_exchange_full<double>(g, p, nullptr, nullptr, nullptr, s);
}

template <typename T>
void launch_pins(const Grid &g, const uint8_t *G, T *p, T pin_in, T pin_hi, T pin_out, hipStream_t s) {
    hipLaunchKernelGGL(k_pins<T>, dim3(nblocks(g.nx, 64), (unsigned)g.ny), dim3(64), 0, s, g, G, p, pin_in,
                       pin_hi, pin_out);
}
template void launch_pins<double>(const Grid &, const uint8_t *, double *, double, double, double, hipStream_t);
template void launch_pins<float>(const Grid &, const uint8_t *, float *, float, float, float, hipStream_t);

void launch_ref_sweeps(const Grid &g, const uint8_t *G, const int32_t *code, double *p, double *u,
                       double *v, double *w, int solver, int nsweeps, int impl, hipStream_t s) {
    if (nsweeps <= 0) return;
    if (impl == REF_ONE_WG) {
        hipLaunchKernelGGL(k_ref_sweeps_1wg, dim3(1), dim3(1024), 0, s, g, G, code, p, u, v, w, solver,
                           nsweeps);
        return;
    }
    if (impl == REF_TILES) {
        const int ntz = (g.nz + kRT - 1) / kRT, nty = (g.ny + kRT - 1) / kRT, ntx = (g.nx + kRT - 1) / kRT;
        for (int it = 0; it < nsweeps; it++) {
            for (int D = 0; D <= (ntz - 1) + (nty - 1) + (ntx - 1); D++) {
                const int a_lo = std::max(0, D - (nty - 1) - (ntx - 1)), a_hi = std::min(ntz - 1, D);
                hipLaunchKernelGGL(k_ref_tile, dim3((unsigned)((a_hi - a_lo + 1) * nty)), dim3(kRThreads), 0, s,
                                   g, G, code, RefFields{{p, u, v, w}}, solver, D, a_lo, nty, ntx);
            }
            launch_exchange_full<double>(g, p, u, v, w, s);
            launch_pins<double>(g, G, p, (double)(g.nz + 2), 2.0, 3.0, s);
        }
        return;
    }
    const int smax = (g.nx - 1) + (g.ny - 1) + (g.nz - 1);
    for (int it = 0; it < nsweeps; it++) {
        for (int sp = 0; sp <= smax; sp++) {
            const int a_lo = std::max(0, sp - (g.ny - 1) - (g.nx - 1));
            const int a_hi = std::min(g.nz - 1, sp);
            const int64_t n = (int64_t)(a_hi - a_lo + 1) * g.ny;
            hipLaunchKernelGGL(k_ref_hyperplane, dim3(nblocks(n, 256)), dim3(256), 0, s, g, G, code, p, u, v,
                               w, solver, sp);
        }
        launch_exchange_full<double>(g, p, u, v, w, s);
        launch_pins<double>(g, G, p, (double)(g.nz + 2), 2.0, 3.0, s);
    }
}

void ref_blocks(const Grid &g, const int dims_in[3], std::vector<RefBox> *out) {
    const int size[3] = {g.nx, g.ny, g.nz};
    int dims[3];
    for (int a = 0; a < 3; a++) dims[a] = std::max(1, std::min(dims_in[a], size[a]));
    out->clear();
    // src/geometry.cc:129-171: along each axis the first size % dims blocks
    // hold one cell more; blocks in rank order (z fastest)
    for (int cx = 0; cx < dims[0]; cx++)
        for (int cy = 0; cy < dims[1]; cy++)
            for (int cz = 0; cz < dims[2]; cz++) {
                const int c[3] = {cx, cy, cz};
                RefBox b;
                for (int a = 0; a < 3; a++) {
                    const int base = size[a] / dims[a], extra = size[a] - dims[a] * base;
                    b.lo[a] = c[a] * base + std::min(c[a], extra);
                    b.n[a] = base + (c[a] < extra ? 1 : 0);
                }
                out->push_back(b);
            }
}

void launch_ref_block_sweep(const Grid &g, const uint8_t *G, const int32_t *code, double *const f[4],
                            double *const snap[4], double *outlet, const RefBox *tab, int nblk, const int ntile[3],
                            int solver, hipStream_t s) {
    // the state as the sweep starts, with the outlet plane as it was swept
    // (before its pins): what the blocks' halo copies hold in the reference
    for (int q = 0; q < 4; q++)
        (void)hipMemcpyAsync(snap[q], f[q], sizeof(double) * (size_t)g.total, hipMemcpyDeviceToDevice, s);
    const int64_t plane = g.base + (int64_t)(g.nz + 1) * g.sy;
    (void)hipMemcpyAsync(snap[0] + plane, outlet + g.base, sizeof(double) * (size_t)g.sy, hipMemcpyDeviceToDevice,
                         s);
    const RefFields live{{f[0], f[1], f[2], f[3]}}, old{{snap[0], snap[1], snap[2], snap[3]}};
    for (int D = 0; D <= (ntile[0] - 1) + (ntile[1] - 1) + (ntile[2] - 1); D++)
        hipLaunchKernelGGL(k_ref_tile_blocks, dim3((unsigned)(nblk * ntile[2] * ntile[1])), dim3(kRThreads), 0, s,
                           g, G, code, live, old, tab, solver, D, ntile[2], ntile[1]);
    launch_exchange_full<double>(g, f[0], f[1], f[2], f[3], s);
    (void)hipMemcpyAsync(outlet + g.base, f[0] + plane, sizeof(double) * (size_t)g.sy, hipMemcpyDeviceToDevice, s);
    launch_pins<double>(g, G, f[0], (double)(g.nz + 2), 2.0, 3.0, s);
}

void launch_reduce(const Grid &g, const uint8_t *G, const double *p, const double *u, const double *v,
                   const double *w, const DomBox &box, bool full, void *scratch, RedSums *out,
                   hipStream_t s) {
    RedAcc *part = (RedAcc *)scratch;
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t rows = n / g.nx;
    const int nb = (int)std::min<int64_t>(kRedBlocksMax, std::max<int64_t>(1, (rows + 3) / 4));
    if (full)
        hipLaunchKernelGGL(k_reduce_state<true>, dim3(nb), dim3(kRedThreads), 0, s, g, G, p, u, v, w, box, part);
    else
        hipLaunchKernelGGL(k_reduce_state<false>, dim3(nb), dim3(kRedThreads), 0, s, g, G, p, u, v, w, box, part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedThreads), 0, s, part, nb, out);
}

size_t reduce_scratch_bytes() { return sizeof(RedAcc) * kRedBlocksMax; }

}  // namespace pm
