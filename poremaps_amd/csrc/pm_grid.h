// pm_grid.h -- device layout and the per-voxel stencil shared by all kernels.
//
// Every field lives in HBM in the reference's halo-2 shape (nz+4, ny+4, nx+4),
// x fastest, with the row pitch (3 ceil((nx+4)/3) >= nx+4 cells, room for a
// colour-split row) rounded up to 32 elements and a base offset of 30 so that
// the first interior cell of every row (k = 2) starts a 128-byte (FP32) /
// 256-byte (FP64) segment: the interior rows a wavefront streams are aligned,
// the halo columns spill into the neighbouring segments.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace pm {

constexpr double kRe = 0.01;     // src/constants.h:57
constexpr double kC2 = 1.5e6;    // src/constants.h:60
constexpr double kDt = 1.0e-4;   // src/constants.h:64
constexpr double kOmegaRef = 1.2;  // src/constants.h:51-52 (reference SOR)
constexpr uint8_t kSolid = 255;  // packed code of a solid cell

struct Grid {
    int nx, ny, nz;   // interior
    int NX, NY, NZ;   // with the 2-cell halo
    int64_t sx, sy;   // row and plane pitch in elements
    int64_t base;     // element offset of cell (0,0,0)
    int64_t total;    // elements to allocate per field
    int px, py, pz;   // periodic axes
    int cb;           // block length of a colour-split row, ceil(NX / 3)
    int slip;         // slip walls on the non-periodic x / y faces (pm_params.walls = 2)

    __host__ __device__ int64_t at(int i, int j, int k) const {
        return base + (int64_t)i * sy + (int64_t)j * sx + k;
    }
    // outlet-plane buffer: one pitched plane
    __host__ __device__ int64_t qat(int j, int k) const { return base + (int64_t)j * sx + k; }
    int64_t qtotal() const { return base + sy; }

    // Colour-split rows (the 3-colour SOR correction state): a row's cells
    // are stored as three blocks by k mod 3, cell k at b * cb + k / 3 with
    // b = k mod 3.  The colour (i + j + k) mod 3 of a row's cells is one
    // block, so the cells one colour pass updates lie contiguous in every
    // row, and so do their -+y / -+z neighbours (same block of the
    // neighbouring rows) and their -+x neighbours (the other two blocks).
    // Same pitches, base and size as the natural layout (3 * cb <= sx).
    __host__ __device__ int64_t ks(int k) const { return (int64_t)(k % 3) * cb + k / 3; }
    __host__ __device__ int64_t at_s(int i, int j, int k) const {
        return base + (int64_t)i * sy + (int64_t)j * sx + ks(k);
    }
    __host__ __device__ int64_t qat_s(int j, int k) const { return base + (int64_t)j * sx + ks(k); }
};

// FP32 residual of the interleaved defect-correction state: the continuity
// part Rp per cell (it is read at the cell and at its -1 neighbours) and the
// momentum part (Ru, Rv, Rw) per fluid voxel in fluid-list order, with Rp
// again in its fourth lane (read only by the voxel's own thread; the SOR's
// colour passes take Rp from there), which keeps the residual traffic of a
// sweep at fluid size.  q (Jacobi): per fluid voxel, the continuity part at the cell
// and at the three cells whose pressure its lagged pass recomputes (-x, -y
// neighbour or their periodic images, -z neighbour), gathered once per
// residual so that a sweep reads them with one streamed load.
// pc (FP32+DC Jacobi on the compact state): the continuity part per entry of
// the compact state (entry order, 0 at the zero entry), read at the cell's
// -x, -y, -z neighbours through their links.
struct Res4 {
    float *p;
    float4 *m;
    float4 *q;
    float *pc = nullptr;
};

inline Grid make_grid(int nx, int ny, int nz, int bc) {
    Grid g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.NX = nx + 4; g.NY = ny + 4; g.NZ = nz + 4;
    g.cb = (g.NX + 2) / 3;
    g.sx = ((int64_t)3 * g.cb + 31) / 32 * 32;
    g.sy = g.sx * g.NY;
    g.base = 30;
    g.total = g.base + g.sy * g.NZ;
    // src/main_stokes_solver.cc:163-165
    g.px = (bc == 0);
    g.py = (bc == 0);
    g.pz = (bc == 0 || bc == 1 || bc == 2);
    g.slip = 0;
    return g;
}

// src/solver.cc:57-91 (get_2nd_derivation); the constant quotients are folded
// in the working precision exactly as the reference's compiler folds them.
template <typename T>
__device__ __forceinline__ T d2(int cs, T ll, T l, T c, T r, T rr) {
    const T K83 = (T)(8.0 / 3.0), K163 = (T)(16.0 / 3.0), K15 = (T)(1.0 / 5.0);
    switch (cs) {
    case 1: return (T)-8.0 * c;
    case 2: return K83 * r - K163 * c;
    case 3: return K83 * l - K163 * c;
    case 4: return ((T)2.0 * r - (T)5.0 * c) - K15 * rr;
    case 5: return ((T)2.0 * l - K15 * ll) - (T)5.0 * c;
    default: return (l - (T)2.0 * c) + r;
    }
}

// A case in {1, 2, 4} means the -1 neighbour on that axis is solid
// (src/geometry.cc:246-289), which is exactly the test src/solver.cc applies
// to geom[i-1], geom[j-1], geom[k-1].
__device__ __forceinline__ bool back_solid(int cs) { return cs == 1 || cs == 2 || cs == 4; }

struct Cases {
    int cx, cy, cz;
    bool bx, by, bz;
};

__device__ __forceinline__ Cases decode(uint8_t pc) {
    Cases c;
    c.cx = pc / 36 + 1;
    c.cy = (pc / 6) % 6 + 1;
    c.cz = pc % 6 + 1;
    c.bx = back_solid(c.cx);
    c.by = back_solid(c.cy);
    c.bz = back_solid(c.cz);
    return c;
}

// A case in {1, 3, 5} means the +1 neighbour on that axis is solid.
__device__ __forceinline__ bool fwd_solid(int cs) { return cs == 1 || cs == 3 || cs == 5; }

// get_2nd_derivation reading only the neighbours the case uses.  Velocities
// of solid cells are exactly +0 (never written, copied as 0 into halos), so
// the arithmetic is the same as loading every operand.
template <typename T>
__device__ __forceinline__ T d2_load(int cs, const T *__restrict__ f, int64_t c, int64_t s, T vc) {
    const T K83 = (T)(8.0 / 3.0), K163 = (T)(16.0 / 3.0), K15 = (T)(1.0 / 5.0);
    switch (cs) {
    case 1: return (T)-8.0 * vc;
    case 2: return K83 * f[c + s] - K163 * vc;
    case 3: return K83 * f[c - s] - K163 * vc;
    case 4: return ((T)2.0 * f[c + s] - (T)5.0 * vc) - K15 * f[c + 2 * s];
    case 5: return ((T)2.0 * f[c - s] - K15 * f[c - 2 * s]) - (T)5.0 * vc;
    default: return (f[c - s] - (T)2.0 * vc) + f[c + s];
    }
}

// Own-axis second difference (l - 2c) + r with solid neighbours read as 0.
template <typename T>
__device__ __forceinline__ T lap_load(int cs, const T *__restrict__ f, int64_t c, int64_t s, T vc) {
    const T l = back_solid(cs) ? (T)0 : f[c - s];
    const T r = fwd_solid(cs) ? (T)0 : f[c + s];
    return (l - (T)2.0 * vc) + r;
}

// d2 with every case's expression evaluated and the case's value picked by
// select (no branches; the same arithmetic as d2).
template <typename T>
__device__ __forceinline__ T d2_pick(int cs, T ll, T l, T vc, T r, T rr) {
    const T K83 = (T)(8.0 / 3.0), K163 = (T)(16.0 / 3.0), K15 = (T)(1.0 / 5.0);
    const T r1 = (T)-8.0 * vc;
    const T r2 = K83 * r - K163 * vc;
    const T r3 = K83 * l - K163 * vc;
    const T r4 = ((T)2.0 * r - (T)5.0 * vc) - K15 * rr;
    const T r5 = ((T)2.0 * l - K15 * ll) - (T)5.0 * vc;
    const T r6 = (l - (T)2.0 * vc) + r;
    T o = r6;
    o = cs == 1 ? r1 : o;
    o = cs == 2 ? r2 : o;
    o = cs == 3 ? r3 : o;
    o = cs == 4 ? r4 : o;
    o = cs == 5 ? r5 : o;
    return o;
}

// get_2nd_derivation with all four neighbours loaded up front and the case's
// expression picked by select: no divergent branches around the loads, the
// same arithmetic as d2_load.
template <typename T>
__device__ __forceinline__ T d2_sel(int cs, const T *__restrict__ f, int64_t c, int64_t s, T vc) {
    return d2_pick<T>(cs, f[c - 2 * s], f[c - s], vc, f[c + s], f[c + 2 * s]);
}

// Momentum forces, every neighbour loaded (branch-free), same arithmetic as
// fast_force_p.
template <typename T>
__device__ __forceinline__ void fast_force_s(const Cases &cs, int64_t c, int64_t sj, int64_t si,
                                             const T *__restrict__ p, const T *__restrict__ u,
                                             const T *__restrict__ v, const T *__restrict__ w,
                                             const T *__restrict__ q, int64_t qc,
                                             T &fu, T &fv, T &fw) {
    const T INV_RE = (T)(1.0 / kRe);
    const T pc = p[c];
    const T pxm = q ? q[qc - 1] : p[c - 1];
    const T pym = q ? q[qc - sj] : p[c - sj];
    const T pzm = p[c - si];
    const T dpx = cs.bx ? (T)0 : pc - pxm;
    const T dpy = cs.by ? (T)0 : pc - pym;
    const T dpz = cs.bz ? (T)0 : pc - pzm;
    const T uc = u[c], vc = v[c], wc = w[c];
    const T ul = u[c - 1], ur = u[c + 1], vl = v[c - sj], vr = v[c + sj], wl = w[c - si], wr = w[c + si];
    const T lxx = ((cs.bx ? (T)0 : ul) - (T)2.0 * uc) + (fwd_solid(cs.cx) ? (T)0 : ur);
    const T lyy = ((cs.by ? (T)0 : vl) - (T)2.0 * vc) + (fwd_solid(cs.cy) ? (T)0 : vr);
    const T lzz = ((cs.bz ? (T)0 : wl) - (T)2.0 * wc) + (fwd_solid(cs.cz) ? (T)0 : wr);
    const T dyx = d2_sel<T>(cs.cx, v, c, 1, vc);
    const T dzx = d2_sel<T>(cs.cx, w, c, 1, wc);
    const T dxy = d2_sel<T>(cs.cy, u, c, sj, uc);
    const T dzy = d2_sel<T>(cs.cy, w, c, sj, wc);
    const T dxz = d2_sel<T>(cs.cz, u, c, si, uc);
    const T dyz = d2_sel<T>(cs.cz, v, c, si, vc);
    const T su = (lxx + dxy) + dxz;
    const T sv = (lyy + dyz) + dyx;
    const T sw = (lzz + dzx) + dzy;
    fu = cs.bx ? (T)0 : su * INV_RE - dpx;
    fv = cs.by ? (T)0 : sv * INV_RE - dpy;
    fw = cs.bz ? (T)0 : sw * INV_RE - dpz;
}

// Momentum forces with case-predicated neighbour loads; same operation order
// as fast_force below and the CPU oracle's stencil_force.
template <typename T>
__device__ __forceinline__ void fast_force_p(const Cases &cs, int64_t c, int64_t sj, int64_t si,
                                             const T *__restrict__ p, const T *__restrict__ u,
                                             const T *__restrict__ v, const T *__restrict__ w,
                                             const T *__restrict__ q, int64_t qc,
                                             T &fu, T &fv, T &fw) {
    const T INV_RE = (T)(1.0 / kRe);
    const T pc = p[c];
    const T dpx = cs.bx ? (T)0 : pc - (q ? q[qc - 1] : p[c - 1]);
    const T dpy = cs.by ? (T)0 : pc - (q ? q[qc - sj] : p[c - sj]);
    const T dpz = cs.bz ? (T)0 : pc - p[c - si];
    const T uc = u[c], vc = v[c], wc = w[c];
    const T lxx = lap_load<T>(cs.cx, u, c, 1, uc);
    const T lyy = lap_load<T>(cs.cy, v, c, sj, vc);
    const T lzz = lap_load<T>(cs.cz, w, c, si, wc);
    const T dyx = d2_load<T>(cs.cx, v, c, 1, vc);
    const T dzx = d2_load<T>(cs.cx, w, c, 1, wc);
    const T dxy = d2_load<T>(cs.cy, u, c, sj, uc);
    const T dzy = d2_load<T>(cs.cy, w, c, sj, wc);
    const T dxz = d2_load<T>(cs.cz, u, c, si, uc);
    const T dyz = d2_load<T>(cs.cz, v, c, si, vc);
    const T su = (lxx + dxy) + dxz;
    const T sv = (lyy + dyz) + dyx;
    const T sw = (lzz + dzx) + dzy;
    fu = cs.bx ? (T)0 : su * INV_RE - dpx;
    fv = cs.by ? (T)0 : sv * INV_RE - dpy;
    fw = cs.bz ? (T)0 : sw * INV_RE - dpz;
}

// Momentum forces of the fast modes at cell c (element index), reading the
// state through raw pointers with the given strides; q (outlet plane only)
// replaces the -x/-y pressure neighbours.  Same operation order as the CPU
// oracle's stencil_force.
template <typename T>
__device__ __forceinline__ void fast_force(const Cases &cs, int64_t c, int64_t sj, int64_t si,
                                           const T *__restrict__ p, const T *__restrict__ u,
                                           const T *__restrict__ v, const T *__restrict__ w,
                                           const T *__restrict__ q, int64_t qc,
                                           T &fu, T &fv, T &fw) {
    const T INV_RE = (T)(1.0 / kRe);
    const T pc = p[c];
    const T pxm = q ? q[qc - 1] : p[c - 1];
    const T pym = q ? q[qc - sj] : p[c - sj];
    const T dpx = cs.bx ? (T)0 : pc - pxm;
    const T dpy = cs.by ? (T)0 : pc - pym;
    const T dpz = cs.bz ? (T)0 : pc - p[c - si];
    const T uc = u[c], vc = v[c], wc = w[c];
    const T lxx = (u[c - 1] - (T)2.0 * uc) + u[c + 1];
    const T lyy = (v[c - sj] - (T)2.0 * vc) + v[c + sj];
    const T lzz = (w[c - si] - (T)2.0 * wc) + w[c + si];
    const T dyx = d2<T>(cs.cx, v[c - 2], v[c - 1], vc, v[c + 1], v[c + 2]);
    const T dzx = d2<T>(cs.cx, w[c - 2], w[c - 1], wc, w[c + 1], w[c + 2]);
    const T dxy = d2<T>(cs.cy, u[c - 2 * sj], u[c - sj], uc, u[c + sj], u[c + 2 * sj]);
    const T dzy = d2<T>(cs.cy, w[c - 2 * sj], w[c - sj], wc, w[c + sj], w[c + 2 * sj]);
    const T dxz = d2<T>(cs.cz, u[c - 2 * si], u[c - si], uc, u[c + si], u[c + 2 * si]);
    const T dyz = d2<T>(cs.cz, v[c - 2 * si], v[c - si], vc, v[c + si], v[c + 2 * si]);
    const T su = (lxx + dxy) + dxz;
    const T sv = (lyy + dyz) + dyx;
    const T sw = (lzz + dzx) + dzy;
    fu = cs.bx ? (T)0 : su * INV_RE - dpx;
    fv = cs.by ? (T)0 : sv * INV_RE - dpy;
    fw = cs.bz ? (T)0 : sw * INV_RE - dpz;
}

}  // namespace pm
