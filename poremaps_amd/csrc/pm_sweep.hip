// pm_sweep.hip -- fast-mode kernels of one geometry.
//
// Porous media are mostly solid (porosity 0.15-0.45 in the target cases) and
// a solid cell's state never changes, so the sweeps launch one thread per
// fluid voxel: a sorted list of the fluid cells' element offsets (z, y, x
// order, so consecutive lanes take consecutive fluid voxels of a row) built
// once per geometry.  The per-cell work lives in pm_voxel.h.
#include "pm_kernels.h"
#include "pm_voxel.h"

#pragma clang fp contract(off)

namespace pm {

constexpr int kListBlock = 256;
constexpr int kXcds = 8;  // MI355X: 8 accelerator dies, each with its own L2

// Workgroups are dispatched to the XCDs round-robin by id; renumber them so
// that XCD x walks the contiguous x-th eighth of the (z-sorted) fluid list and
// the z-neighbour planes its waves share stay in that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int xcd = b % kXcds, local = b / kXcds;
    const int q = nb / kXcds, r = nb % kXcds;
    return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

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
    vel_voxel<T, HAS_R, SEL>(g, c, pc[c], p0, u0, v0, w0, q0, u1, v1, w1, Ru, Rv, Rw, cu);
}

template <typename T, bool HAS_R>
__global__ __launch_bounds__(kListBlock) void k_jac_pres(Grid g, const int32_t *__restrict__ list, int n,
                                                         const uint8_t *__restrict__ pc,
                                                         const T *__restrict__ p0, const T *__restrict__ u1,
                                                         const T *__restrict__ v1, const T *__restrict__ w1,
                                                         T *__restrict__ p1, T *__restrict__ q1,
                                                         const float *__restrict__ Rp, T cp, T pin_out) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t c = list[t];
    pres_voxel<T, HAS_R>(g, c, pc[c], p0, u1, v1, w1, p1, q1, Rp, cp, pin_out);
}

template <typename T, bool HAS_R, bool SEL>
__global__ __launch_bounds__(kListBlock) void k_colour(Grid g, const int32_t *__restrict__ list, int n,
                                                       const uint8_t *__restrict__ pc, T *p, T *u, T *v, T *w,
                                                       T *q, T cu, T cp, T pin_out,
                                                       const float *__restrict__ Rp,
                                                       const float *__restrict__ Ru,
                                                       const float *__restrict__ Rv,
                                                       const float *__restrict__ Rw) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t c = list[t];
    colour_voxel<T, HAS_R, SEL>(g, c, pc[c], p, u, v, w, q, cu, cp, pin_out, Rp, Ru, Rv, Rw);
}

__global__ __launch_bounds__(kListBlock) void k_residual(Grid g, const int32_t *__restrict__ list, int n,
                                                         const uint8_t *__restrict__ pc,
                                                         const double *__restrict__ p,
                                                         const double *__restrict__ u,
                                                         const double *__restrict__ v,
                                                         const double *__restrict__ w,
                                                         const double *__restrict__ q, float *__restrict__ Rp,
                                                         float *__restrict__ Ru, float *__restrict__ Rv,
                                                         float *__restrict__ Rw) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t c = list[t];
    residual_voxel(g, c, pc[c], p, u, v, w, q, Rp, Ru, Rv, Rw);
}

template <typename T>
__global__ void k_ghost_fast(Grid g, const uint8_t *__restrict__ G, T *p, T *u, T *v, T *w, T *q, T pin_in,
                             T pin_hi, T pin_out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    ghost_line<T>(g, t, G, p, u, v, w, q, pin_in, pin_hi, pin_out);
}

__global__ void k_outlet_plane(Grid g, const uint8_t *__restrict__ pc, const double *__restrict__ p,
                               const double *__restrict__ u, const double *__restrict__ v,
                               const double *__restrict__ w, double pin_out, double *__restrict__ q) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= g.NX) return;
    outlet_cell(g, j, k, pc, p, u, v, w, pin_out, q);
}

__global__ void k_fold(Grid g, const uint8_t *__restrict__ pc, double *Xp, double *Xu, double *Xv, double *Xw,
                       float *dp, float *du, float *dv, float *dw) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.total) return;
    fold_cell(g, c, pc, Xp, Xu, Xv, Xw, dp, du, dv, dw);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline unsigned lblocks(int n) { return (unsigned)((n + kListBlock - 1) / kListBlock); }
static inline unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

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

template <typename T>
void launch_ghost_fast(const Grid &g, const uint8_t *G, T *p, T *u, T *v, T *w, T *q, T pin_in, T pin_hi,
                       T pin_out, hipStream_t s) {
    hipLaunchKernelGGL(k_ghost_fast<T>, dim3(nblocks(ghost_lines(g), 256)), dim3(256), 0, s, g, G, p, u, v, w, q,
                       pin_in, pin_hi, pin_out);
}
template void launch_ghost_fast<double>(const Grid &, const uint8_t *, double *, double *, double *, double *,
                                        double *, double, double, double, hipStream_t);
template void launch_ghost_fast<float>(const Grid &, const uint8_t *, float *, float *, float *, float *,
                                       float *, float, float, float, hipStream_t);

void launch_outlet_plane(const Grid &g, const uint8_t *pc, const double *p, const double *u, const double *v,
                         const double *w, double pin_out, double *q, hipStream_t s) {
    hipLaunchKernelGGL(k_outlet_plane, dim3(nblocks(g.NX, 64), (unsigned)g.NY), dim3(64), 0, s, g, pc, p, u, v,
                       w, pin_out, q);
}

void launch_fold(const Grid &g, const uint8_t *pc, double *Xp, double *Xu, double *Xv, double *Xw, float *dp,
                 float *du, float *dv, float *dw, hipStream_t s) {
    hipLaunchKernelGGL(k_fold, dim3(nblocks(g.total, 256)), dim3(256), 0, s, g, pc, Xp, Xu, Xv, Xw, dp, du, dv,
                       dw);
}

}  // namespace pm
