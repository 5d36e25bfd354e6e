// pm_sweep.hip -- fast-mode sweep kernels over fluid-voxel lists.
//
// Porous media are mostly solid (porosity 0.15-0.45 in the target cases) and
// a solid cell's state never changes, so the sweeps launch one thread per
// fluid voxel: a sorted list of the fluid cells' element offsets (z, y, x
// order, so consecutive lanes take consecutive fluid voxels of a row) built
// once per geometry.  Neighbour loads are predicated on the voxel's case
// digits: the velocity of a solid cell is exactly 0 and is not read.
#include "pm_kernels.h"

#pragma clang fp contract(off)

namespace pm {

constexpr int kListBlock = 256;
constexpr int kXcds = 8;   // MI355X: 8 accelerator dies, each with its own L2

// Workgroups are dispatched to the XCDs round-robin by id; renumber them so
// that XCD x walks the contiguous x-th eighth of the (z-sorted) fluid list and
// the z-neighbour planes its waves share stay in that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int xcd = b % kXcds, local = b / kXcds;
    const int q = nb / kXcds, r = nb % kXcds;
    return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

__device__ __forceinline__ void cell_jk(const Grid &g, int64_t c, int &i, int &j, int &k) {
    const uint32_t e = (uint32_t)(c - g.base);
    const uint32_t sy = (uint32_t)g.sy, sx = (uint32_t)g.sx;
    i = (int)(e / sy);
    const uint32_t r = e - (uint32_t)i * sy;
    j = (int)(r / sx);
    k = (int)(r - (uint32_t)j * sx);
}

// Symplectic Jacobi, velocity pass over the fluid list.
template <typename T, bool HAS_R, bool SEL>
__global__ __launch_bounds__(kListBlock) void k_jac_vel(Grid g, const int32_t *__restrict__ list, int n,
                                                        const uint8_t *__restrict__ pc,
                                                        const T *__restrict__ p0, const T *__restrict__ u0,
                                                        const T *__restrict__ v0, const T *__restrict__ w0,
                                                        const T *__restrict__ q0, T *__restrict__ u1,
                                                        T *__restrict__ v1, T *__restrict__ w1,
                                                        const float *__restrict__ Ru,
                                                        const float *__restrict__ Rv,
                                                        const float *__restrict__ Rw, T cu) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t c = list[t];
    const Cases cs = decode(pc[c]);
    const int64_t o0 = g.at(g.nz + 1, 0, 0);
    const bool out = c >= o0 && c < o0 + g.sy;
    T fu, fv, fw;
    if (SEL)
        fast_force_s<T>(cs, c, g.sx, g.sy, p0, u0, v0, w0, out ? q0 : nullptr, c - (int64_t)(g.nz + 1) * g.sy,
                        fu, fv, fw);
    else
        fast_force_p<T>(cs, c, g.sx, g.sy, p0, u0, v0, w0, out ? q0 : nullptr, c - (int64_t)(g.nz + 1) * g.sy,
                        fu, fv, fw);
    if (HAS_R) { fu = fu + (T)Ru[c]; fv = fv + (T)Rv[c]; fw = fw + (T)Rw[c]; }
    u1[c] = cs.bx ? (T)0 : u0[c] + fu * cu;
    v1[c] = cs.by ? (T)0 : v0[c] + fv * cu;
    w1[c] = cs.bz ? (T)0 : w0[c] + fw * cu;
}

// Symplectic Jacobi, pressure pass: divergence of the new velocities (+1
// neighbours wrapped on periodic axes, 0 where solid) and the old pressure.
template <typename T, bool HAS_R>
__global__ __launch_bounds__(kListBlock) void k_jac_pres(Grid g, const int32_t *__restrict__ list, int n,
                                                         const uint8_t *__restrict__ pc,
                                                         const T *__restrict__ p0,
                                                         const T *__restrict__ u1,
                                                         const T *__restrict__ v1,
                                                         const T *__restrict__ w1, T *__restrict__ p1,
                                                         T *__restrict__ q1, const float *__restrict__ Rp,
                                                         T cp, T pin_out) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t c = list[t];
    const Cases cs = decode(pc[c]);
    int i, j, k;
    cell_jk(g, c, i, j, k);
    const int64_t cxp = (k == g.nx + 1 && g.px) ? c + 1 - g.nx : c + 1;
    const int64_t cyp = (j == g.ny + 1 && g.py) ? c + g.sx - (int64_t)g.ny * g.sx : c + g.sx;
    const int64_t czp = (i == g.nz + 1 && g.pz) ? c + g.sy - (int64_t)g.nz * g.sy : c + g.sy;
    const T uc = u1[c], vc = v1[c], wc = w1[c];
    const T wz = fwd_solid(cs.cz) ? (T)0 : w1[czp];
    const T vy = fwd_solid(cs.cy) ? (T)0 : v1[cyp];
    const T ux = fwd_solid(cs.cx) ? (T)0 : u1[cxp];
    const T div0 = ((((wz - wc) + vy) - vc) + ux) - uc;
    T div = div0;
    if (HAS_R) div = div + (T)Rp[c];
    p1[c] = p0[c] - div * cp;
    if (i == g.nz + 1) q1[g.qat(j, k)] = pin_out - div0 * (T)(kC2 * kDt);
}

// 3-colour SOR: one colour's list, in place.
template <typename T, bool HAS_R, bool SEL>
__global__ __launch_bounds__(kListBlock) void k_colour(Grid g, const int32_t *__restrict__ list, int n,
                                                       const uint8_t *__restrict__ pc, T *p, T *u, T *v,
                                                       T *w, T *q, T cu, T cp, T pin_out,
                                                       const float *__restrict__ Rp,
                                                       const float *__restrict__ Ru,
                                                       const float *__restrict__ Rv,
                                                       const float *__restrict__ Rw) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t c = list[t];
    const Cases cs = decode(pc[c]);
    const int64_t o0 = g.at(g.nz + 1, 0, 0);
    const bool out = c >= o0 && c < o0 + g.sy;
    const int64_t qc = c - (int64_t)(g.nz + 1) * g.sy;
    T fu, fv, fw;
    if (SEL)
        fast_force_s<T>(cs, c, g.sx, g.sy, p, u, v, w, out ? q : nullptr, qc, fu, fv, fw);
    else
        fast_force_p<T>(cs, c, g.sx, g.sy, p, u, v, w, out ? q : nullptr, qc, fu, fv, fw);
    if (HAS_R) { fu = fu + (T)Ru[c]; fv = fv + (T)Rv[c]; fw = fw + (T)Rw[c]; }
    const T un = cs.bx ? (T)0 : u[c] + fu * cu;
    const T vn = cs.by ? (T)0 : v[c] + fv * cu;
    const T wn = cs.bz ? (T)0 : w[c] + fw * cu;
    u[c] = un; v[c] = vn; w[c] = wn;
    const T wz = fwd_solid(cs.cz) ? (T)0 : w[c + g.sy];
    const T vy = fwd_solid(cs.cy) ? (T)0 : v[c + g.sx];
    const T ux = fwd_solid(cs.cx) ? (T)0 : u[c + 1];
    const T div0 = ((((wz - wn) + vy) - vn) + ux) - un;
    T div = div0;
    if (HAS_R) div = div + (T)Rp[c];
    p[c] = p[c] - div * cp;
    if (out) q[qc] = pin_out - div0 * (T)(kC2 * kDt);
}

// Momentum and continuity residual of a FP64 state (defect correction).
__global__ __launch_bounds__(kListBlock) void k_residual(Grid g, const int32_t *__restrict__ list, int n,
                                                         const uint8_t *__restrict__ pc,
                                                         const double *__restrict__ p,
                                                         const double *__restrict__ u,
                                                         const double *__restrict__ v,
                                                         const double *__restrict__ w,
                                                         const double *__restrict__ q,
                                                         float *__restrict__ Rp, float *__restrict__ Ru,
                                                         float *__restrict__ Rv, float *__restrict__ Rw) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t c = list[t];
    const Cases cs = decode(pc[c]);
    const int64_t o0 = g.at(g.nz + 1, 0, 0);
    const bool out = c >= o0 && c < o0 + g.sy;
    double fu, fv, fw;
    fast_force_p<double>(cs, c, g.sx, g.sy, p, u, v, w, out ? q : nullptr, c - (int64_t)(g.nz + 1) * g.sy,
                         fu, fv, fw);
    Ru[c] = (float)fu; Rv[c] = (float)fv; Rw[c] = (float)fw;
    const double wz = fwd_solid(cs.cz) ? 0.0 : w[c + g.sy];
    const double vy = fwd_solid(cs.cy) ? 0.0 : v[c + g.sx];
    const double ux = fwd_solid(cs.cx) ? 0.0 : u[c + 1];
    const double div = ((((wz - w[c]) + vy) - v[c]) + ux) - u[c];
    Rp[c] = (float)div;
}

static inline unsigned lblocks(int n) { return (unsigned)((n + kListBlock - 1) / kListBlock); }

template <typename T, bool HAS_R, bool SEL>
static void jacobi_pair(const Grid &g, const FluidLists &L, const uint8_t *pc, const T *p0, const T *u0,
                        const T *v0, const T *w0, const T *q0, T *p1, T *u1, T *v1, T *w1, T *q1, T cu, T cp,
                        T pin_out, const float *Rp, const float *Ru, const float *Rv, const float *Rw,
                        hipStream_t s) {
    hipLaunchKernelGGL((k_jac_vel<T, HAS_R, SEL>), dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, g, L.all,
                       L.n_all, pc, p0, u0, v0, w0, q0, u1, v1, w1, Ru, Rv, Rw, cu);
    hipLaunchKernelGGL((k_jac_pres<T, HAS_R>), dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, g, L.all, L.n_all,
                       pc, p0, u1, v1, w1, p1, q1, Rp, cp, pin_out);
}

template <typename T>
void launch_jacobi(const Grid &g, const FluidLists &L, const uint8_t *pc, const T *p0, const T *u0,
                   const T *v0, const T *w0, const T *q0, T *p1, T *u1, T *v1, T *w1, T *q1, T cu, T cp,
                   T pin_out, const float *Rp, const float *Ru, const float *Rv, const float *Rw,
                   hipStream_t s) {
    if (L.n_all == 0) return;
    if (Rp && L.select_loads)
        jacobi_pair<T, true, true>(g, L, pc, p0, u0, v0, w0, q0, p1, u1, v1, w1, q1, cu, cp, pin_out, Rp, Ru, Rv,
                                   Rw, s);
    else if (Rp)
        jacobi_pair<T, true, false>(g, L, pc, p0, u0, v0, w0, q0, p1, u1, v1, w1, q1, cu, cp, pin_out, Rp, Ru, Rv,
                                    Rw, s);
    else if (L.select_loads)
        jacobi_pair<T, false, true>(g, L, pc, p0, u0, v0, w0, q0, p1, u1, v1, w1, q1, cu, cp, pin_out, Rp, Ru,
                                    Rv, Rw, s);
    else
        jacobi_pair<T, false, false>(g, L, pc, p0, u0, v0, w0, q0, p1, u1, v1, w1, q1, cu, cp, pin_out, Rp, Ru,
                                     Rv, Rw, s);
}
template void launch_jacobi<double>(const Grid &, const FluidLists &, const uint8_t *, const double *,
                                    const double *, const double *, const double *, const double *, double *,
                                    double *, double *, double *, double *, double, double, double,
                                    const float *, const float *, const float *, const float *, hipStream_t);
template void launch_jacobi<float>(const Grid &, const FluidLists &, const uint8_t *, const float *,
                                   const float *, const float *, const float *, const float *, float *, float *,
                                   float *, float *, float *, float, float, float, const float *, const float *,
                                   const float *, const float *, hipStream_t);

template <typename T, bool HAS_R, bool SEL>
static void colour_stage(const Grid &g, const int32_t *list, int n, const uint8_t *pc, T *p, T *u, T *v, T *w,
                         T *q, T cu, T cp, T pin_out, const float *Rp, const float *Ru, const float *Rv,
                         const float *Rw, hipStream_t s) {
    hipLaunchKernelGGL((k_colour<T, HAS_R, SEL>), dim3(lblocks(n)), dim3(kListBlock), 0, s, g, list, n, pc, p, u,
                       v, w, q, cu, cp, pin_out, Rp, Ru, Rv, Rw);
}

template <typename T>
void launch_colour(const Grid &g, const FluidLists &L, const uint8_t *pc, T *p, T *u, T *v, T *w, T *q, T cu,
                   T cp, T pin_out, const float *Rp, const float *Ru, const float *Rv, const float *Rw,
                   hipStream_t s) {
    for (int col = 0; col < 3; col++) {
        const int n = L.n_col[col];
        if (n == 0) continue;
        if (Rp && L.select_loads)
            colour_stage<T, true, true>(g, L.col[col], n, pc, p, u, v, w, q, cu, cp, pin_out, Rp, Ru, Rv, Rw, s);
        else if (Rp)
            colour_stage<T, true, false>(g, L.col[col], n, pc, p, u, v, w, q, cu, cp, pin_out, Rp, Ru, Rv, Rw, s);
        else if (L.select_loads)
            colour_stage<T, false, true>(g, L.col[col], n, pc, p, u, v, w, q, cu, cp, pin_out, Rp, Ru, Rv, Rw, s);
        else
            colour_stage<T, false, false>(g, L.col[col], n, pc, p, u, v, w, q, cu, cp, pin_out, Rp, Ru, Rv, Rw, s);
    }
}
template void launch_colour<double>(const Grid &, const FluidLists &, const uint8_t *, double *, double *,
                                    double *, double *, double *, double, double, double, const float *,
                                    const float *, const float *, const float *, hipStream_t);
template void launch_colour<float>(const Grid &, const FluidLists &, const uint8_t *, float *, float *, float *,
                                   float *, float *, float, float, float, const float *, const float *,
                                   const float *, const float *, hipStream_t);

void launch_residual(const Grid &g, const FluidLists &L, const uint8_t *pc, const double *p, const double *u,
                     const double *v, const double *w, const double *q, float *Rp, float *Ru, float *Rv,
                     float *Rw, hipStream_t s) {
    if (L.n_all == 0) return;
    hipLaunchKernelGGL(k_residual, dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, g, L.all, L.n_all, pc, p, u, v,
                       w, q, Rp, Ru, Rv, Rw);
}

}  // namespace pm
