// pm_voxel.h -- per-cell work of the fast modes, shared by the single-solve
// kernels (pm_sweep.hip, pm_kernels.hip) and the batched kernels
// (pm_batch.hip).  Each function handles one fluid cell (element offset c) or
// one halo line of one geometry described by a Grid.
#pragma once

#include "pm_grid.h"

#pragma clang fp contract(off)

namespace pm {

__device__ __forceinline__ void cell_ijk(const Grid &g, int64_t c, int &i, int &j, int &k) {
    const uint32_t e = (uint32_t)(c - g.base);
    const uint32_t sy = (uint32_t)g.sy, sx = (uint32_t)g.sx;
    i = (int)(e / sy);
    const uint32_t r = e - (uint32_t)i * sy;
    j = (int)(r / sx);
    k = (int)(r - (uint32_t)j * sx);
}

// Slip walls (Grid::slip): a cell of the first interior column (row) next to
// a slip x (y) face has the wall on its -x (-y) face, where u (v) is held at
// 0 -- the same as a solid -1 neighbour does to that component and its
// pressure difference.  Its second derivatives keep their own cases: they see
// the mirrored halos.  j, k: the cell's indices with the halo (k == 2: the
// first interior column).
__device__ __forceinline__ void slip_hold(const Grid &g, int j, int k, Cases &cs) {
    if (!g.slip) return;
    if (!g.px && k == 2) cs.bx = true;
    if (!g.py && j == 2) cs.by = true;
}
__device__ __forceinline__ void slip_hold_at(const Grid &g, int64_t c, Cases &cs) {
    if (!g.slip) return;
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    slip_hold(g, j, k, cs);
}

// The cell's mirror images in the slip-wall halos (Grid::slip), written by
// the passes that write their cells' halo copies themselves: columns k = 2, 3
// mirror into k = 1, 0 and k = nx + 1, nx into nx + 2, nx + 3 with u = 0, rows
// likewise with v = 0; p = 0 (the halo pressure of a correction state).
typedef float slip_v4f __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void slip_images4(const Grid &g, int64_t c, int j, int k, slip_v4f o, float4 *S1) {
    if (!g.slip) return;
    if (!g.px) {
        slip_v4f m = o;
        m.x = 0.f;
        m.y = 0.f;
        if (k == 2) __builtin_nontemporal_store(m, reinterpret_cast<slip_v4f *>(&S1[c - 1]));
        if (k == 3) __builtin_nontemporal_store(m, reinterpret_cast<slip_v4f *>(&S1[c - 3]));
        if (k == g.nx + 1) __builtin_nontemporal_store(m, reinterpret_cast<slip_v4f *>(&S1[c + 1]));
        if (k == g.nx) __builtin_nontemporal_store(m, reinterpret_cast<slip_v4f *>(&S1[c + 3]));
    }
    if (!g.py) {
        slip_v4f m = o;
        m.x = 0.f;
        m.z = 0.f;
        const int64_t sx = g.sx;
        if (j == 2) __builtin_nontemporal_store(m, reinterpret_cast<slip_v4f *>(&S1[c - sx]));
        if (j == 3) __builtin_nontemporal_store(m, reinterpret_cast<slip_v4f *>(&S1[c - 3 * sx]));
        if (j == g.ny + 1) __builtin_nontemporal_store(m, reinterpret_cast<slip_v4f *>(&S1[c + sx]));
        if (j == g.ny) __builtin_nontemporal_store(m, reinterpret_cast<slip_v4f *>(&S1[c + 3 * sx]));
    }
}

__device__ __forceinline__ bool on_outlet(const Grid &g, int64_t c) {
    const int64_t o0 = g.at(g.nz + 1, 0, 0);
    return c >= o0 && c < o0 + g.sy;
}

// Symplectic Jacobi, velocity part: (p0,u0,v0,w0) -> (u1,v1,w1) at c.
template <typename T, bool SEL>
__device__ __forceinline__ void vel_voxel(const Grid &g, int64_t c, uint8_t pcv, const T *__restrict__ p0,
                                          const T *__restrict__ u0, const T *__restrict__ v0,
                                          const T *__restrict__ w0, const T *__restrict__ q0,
                                          T *__restrict__ u1, T *__restrict__ v1, T *__restrict__ w1, T cu) {
    Cases cs = decode(pcv);
    slip_hold_at(g, c, cs);
    const bool out = on_outlet(g, c);
    const int64_t qc = c - (int64_t)(g.nz + 1) * g.sy;
    T fu, fv, fw;
    if (SEL) fast_force_s<T>(cs, c, g.sx, g.sy, p0, u0, v0, w0, out ? q0 : nullptr, qc, fu, fv, fw);
    else fast_force_p<T>(cs, c, g.sx, g.sy, p0, u0, v0, w0, out ? q0 : nullptr, qc, fu, fv, fw);
    u1[c] = cs.bx ? (T)0 : u0[c] + fu * cu;
    v1[c] = cs.by ? (T)0 : v0[c] + fv * cu;
    w1[c] = cs.bz ? (T)0 : w0[c] + fw * cu;
}

// Symplectic Jacobi, pressure part: divergence of the new velocities (+1
// neighbours wrapped on periodic axes, 0 where solid) and the old pressure;
// on the outlet plane also q1 = pin_out - c2*dt*div (residual excluded).
template <typename T>
__device__ __forceinline__ void pres_voxel(const Grid &g, int64_t c, uint8_t pcv, const T *__restrict__ p0,
                                           const T *__restrict__ u1, const T *__restrict__ v1,
                                           const T *__restrict__ w1, T *__restrict__ p1, T *__restrict__ q1,
                                           T cp, T pin_out) {
    const Cases cs = decode(pcv);
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    const int64_t cxp = (k == g.nx + 1 && g.px) ? c + 1 - g.nx : c + 1;
    const int64_t cyp = (j == g.ny + 1 && g.py) ? c + g.sx - (int64_t)g.ny * g.sx : c + g.sx;
    const int64_t czp = (i == g.nz + 1 && g.pz) ? c + g.sy - (int64_t)g.nz * g.sy : c + g.sy;
    const T uc = u1[c], vc = v1[c], wc = w1[c];
    const T wz = fwd_solid(cs.cz) ? (T)0 : w1[czp];
    const T vy = fwd_solid(cs.cy) ? (T)0 : v1[cyp];
    const T ux = fwd_solid(cs.cx) ? (T)0 : u1[cxp];
    const T div0 = ((((wz - wc) + vy) - vc) + ux) - uc;
    p1[c] = p0[c] - div0 * cp;
    if (i == g.nz + 1) q1[g.qat(j, k)] = pin_out - div0 * (T)(kC2 * kDt);
}

// ---------------------------------------------------------------------------
// Symplectic Jacobi in lagged form.  A sweep n maps (u^n, p^n) to
// (u^{n+1}, p^{n+1}) with u^{n+1} = u^n + cu*F(u^n, p^n) and
// p^{n+1} = p^n - cp*(div u^{n+1} + Rp).  The lagged pass maps
// (u^n, p^{n-1}) to (u^{n+1}, p^n): it recomputes p^n = p^{n-1} -
// cp*(div u^n + Rp) at the cell and at its -x/-y/-z neighbours from the
// input state (halos filled) instead of reading a pressure written by a
// second kernel, so one pass over the fluid list does the work of the
// velocity and pressure kernels.  Every value is produced by the same
// expressions as vel_voxel / pres_voxel, so the lagged sequence is bit-equal
// to the two-kernel one; a final pres_voxel pass ("settle") supplies p of the
// last sweep.
// ---------------------------------------------------------------------------

// Viscous sums (su, sv, sw) of the momentum forces at c: the stencil of
// fast_force_p (SEL false: case-predicated loads) or fast_force_s (SEL true).
template <typename T, bool SEL>
__device__ __forceinline__ void visc_sums(const Cases &cs, int64_t c, int64_t sj, int64_t si,
                                          const T *__restrict__ u, const T *__restrict__ v,
                                          const T *__restrict__ w, T uc, T vc, T wc, T &su, T &sv, T &sw) {
    T lxx, lyy, lzz, dyx, dzx, dxy, dzy, dxz, dyz;
    if (SEL) {
        const T ul = u[c - 1], ur = u[c + 1], vl = v[c - sj], vr = v[c + sj], wl = w[c - si], wr = w[c + si];
        lxx = ((cs.bx ? (T)0 : ul) - (T)2.0 * uc) + (fwd_solid(cs.cx) ? (T)0 : ur);
        lyy = ((cs.by ? (T)0 : vl) - (T)2.0 * vc) + (fwd_solid(cs.cy) ? (T)0 : vr);
        lzz = ((cs.bz ? (T)0 : wl) - (T)2.0 * wc) + (fwd_solid(cs.cz) ? (T)0 : wr);
        dyx = d2_sel<T>(cs.cx, v, c, 1, vc);
        dzx = d2_sel<T>(cs.cx, w, c, 1, wc);
        dxy = d2_sel<T>(cs.cy, u, c, sj, uc);
        dzy = d2_sel<T>(cs.cy, w, c, sj, wc);
        dxz = d2_sel<T>(cs.cz, u, c, si, uc);
        dyz = d2_sel<T>(cs.cz, v, c, si, vc);
    } else {
        lxx = lap_load<T>(cs.cx, u, c, 1, uc);
        lyy = lap_load<T>(cs.cy, v, c, sj, vc);
        lzz = lap_load<T>(cs.cz, w, c, si, wc);
        dyx = d2_load<T>(cs.cx, v, c, 1, vc);
        dzx = d2_load<T>(cs.cx, w, c, 1, wc);
        dxy = d2_load<T>(cs.cy, u, c, sj, uc);
        dzy = d2_load<T>(cs.cy, w, c, sj, wc);
        dxz = d2_load<T>(cs.cz, u, c, si, uc);
        dyz = d2_load<T>(cs.cz, v, c, si, vc);
    }
    su = (lxx + dxy) + dxz;
    sv = (lyy + dyz) + dyx;
    sw = (lzz + dzx) + dzy;
}

// Divergence of the velocities at fluid cell n as pres_voxel forms it (the +1
// neighbours read through the filled halos, 0 where solid).
template <typename T>
__device__ __forceinline__ T lag_div0(const Grid &g, int64_t n, uint8_t pcn, const T *__restrict__ u,
                                      const T *__restrict__ v, const T *__restrict__ w) {
    const Cases cs = decode(pcn);
    const T wz = fwd_solid(cs.cz) ? (T)0 : w[n + g.sy];
    const T vy = fwd_solid(cs.cy) ? (T)0 : v[n + g.sx];
    const T ux = fwd_solid(cs.cx) ? (T)0 : u[n + 1];
    return ((((wz - w[n]) + vy) - v[n]) + ux) - u[n];
}

// p^n at a fluid interior cell n off the outlet plane.
template <typename T>
__device__ __forceinline__ T lag_p(const Grid &g, int64_t n, uint8_t pcn, const T *__restrict__ p0,
                                   const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w, T cp) {
    return p0[n] - lag_div0<T>(g, n, pcn, u, v, w) * cp;
}

// Outlet-plane buffer q^n at a fluid interior cell n of z = nz+1.
template <typename T>
__device__ __forceinline__ T lag_q(const Grid &g, int64_t n, uint8_t pcn, const T *__restrict__ u,
                                   const T *__restrict__ v, const T *__restrict__ w, T pin_out) {
    return pin_out - lag_div0<T>(g, n, pcn, u, v, w) * (T)(kC2 * kDt);
}

// Pressure seen by fluid cell c from its fluid -1 neighbour along an in-plane
// axis (stride s, index idx of c on that axis, n_ax interior cells, periodic
// flag per): recomputed at the neighbour or at its periodic image; a wall
// halo keeps its stored value (p, or q on the outlet plane).
template <typename T>
__device__ __forceinline__ T lag_back(const Grid &g, int64_t c, int64_t s, int idx, int n_ax, bool per, bool out,
                                      int64_t qn, const uint8_t *__restrict__ pc, const T *__restrict__ p0,
                                      const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
                                      const T *__restrict__ q0, T cp, T pin_out) {
    int64_t n = c - s;
    if (idx == 2) {
        if (!per) return out ? q0[qn] : p0[n];
        n += (int64_t)n_ax * s;
    }
    return out ? lag_q<T>(g, n, pc[n], u, v, w, pin_out) : lag_p<T>(g, n, pc[n], p0, u, v, w, cp);
}

// One lagged pass at fluid cell c: reads (p^{n-1}, u^n, v^n, w^n) = set 0,
// writes (p^n, u^{n+1}, v^{n+1}, w^{n+1}) into set 1.
template <typename T, bool SEL>
__device__ __forceinline__ void lag_voxel(const Grid &g, int64_t c, uint8_t pcv, const uint8_t *__restrict__ pc,
                                          const T *__restrict__ p0, const T *__restrict__ u0,
                                          const T *__restrict__ v0, const T *__restrict__ w0,
                                          const T *__restrict__ q0, T *__restrict__ p1, T *__restrict__ u1,
                                          T *__restrict__ v1, T *__restrict__ w1, T cu, T cp, T pin_out) {
    Cases cs = decode(pcv);
    slip_hold_at(g, c, cs);
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    const bool out = i == g.nz + 1;
    const T Pc = out ? pin_out : lag_p<T>(g, c, pcv, p0, u0, v0, w0, cp);
    T dpx = (T)0, dpy = (T)0, dpz = (T)0;
    if (!cs.bx)
        dpx = Pc - lag_back<T>(g, c, 1, k, g.nx, g.px, out, g.qat(j, k - 1), pc, p0, u0, v0, w0, q0, cp, pin_out);
    if (!cs.by)
        dpy = Pc - lag_back<T>(g, c, g.sx, j, g.ny, g.py, out, g.qat(j - 1, k), pc, p0, u0, v0, w0, q0, cp,
                               pin_out);
    if (!cs.bz) {
        const int64_t n = c - g.sy;
        dpz = Pc - (i == 2 ? p0[n] : lag_p<T>(g, n, pc[n], p0, u0, v0, w0, cp));
    }
    const T uc = u0[c], vc = v0[c], wc = w0[c];
    T su, sv, sw;
    visc_sums<T, SEL>(cs, c, g.sx, g.sy, u0, v0, w0, uc, vc, wc, su, sv, sw);
    const T INV_RE = (T)(1.0 / kRe);
    T fu = cs.bx ? (T)0 : su * INV_RE - dpx;
    T fv = cs.by ? (T)0 : sv * INV_RE - dpy;
    T fw = cs.bz ? (T)0 : sw * INV_RE - dpz;
    u1[c] = cs.bx ? (T)0 : uc + fu * cu;
    v1[c] = cs.by ? (T)0 : vc + fv * cu;
    w1[c] = cs.bz ? (T)0 : wc + fw * cu;
    p1[c] = Pc;
}

// ---------------------------------------------------------------------------
// FP32 correction state of the defect-correction Jacobi mode, interleaved:
// one float4 {p, u, v, w} per cell (and the FP32 residual likewise), so that a
// neighbour costs one 16-byte load instead of up to four 4-byte gathers.  Same
// expressions, in the same order, as the per-field functions above.
// ---------------------------------------------------------------------------

// The 12 stencil neighbours of a cell (+-1, +-2 on each axis).
struct Star4 {
    float4 XM, XP, XM2, XP2, YM, YP, YM2, YP2, ZM, ZP, ZM2, ZP2;
};

__device__ __forceinline__ Star4 load_star4(const float4 *__restrict__ S, int64_t c, int64_t sj, int64_t si) {
    Star4 s;
    s.XM = S[c - 1]; s.XP = S[c + 1]; s.XM2 = S[c - 2]; s.XP2 = S[c + 2];
    s.YM = S[c - sj]; s.YP = S[c + sj]; s.YM2 = S[c - 2 * sj]; s.YP2 = S[c + 2 * sj];
    s.ZM = S[c - si]; s.ZP = S[c + si]; s.ZM2 = S[c - 2 * si]; s.ZP2 = S[c + 2 * si];
    return s;
}

// The +-2 neighbours are read only by the one-sided cases 4 (+2) and 5 (-2)
// of the second derivatives (about 9 % of fluid voxels per axis at porosity
// 0.2).  load_star4_masked reads them with buffer loads whose offset is out of
// range for the lanes whose case does not use them: those lanes make no
// memory request and get 0, which d2 never reads, so every result is the
// same.  The buffer rs covers the state array (byte offsets below 2^32).
__device__ __forceinline__ float4 buf_load4(__amdgpu_buffer_rsrc_t rs, int64_t e, bool need) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    const uint32_t off = need ? (uint32_t)(e * 16) : 0xFFFFFFF0u;
    const v4f v = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
    return make_float4(v.x, v.y, v.z, v.w);
}

// One float of the state (component q of element e) the same way.
__device__ __forceinline__ float buf_load1(__amdgpu_buffer_rsrc_t rs, int64_t e, int q, bool need) {
    const uint32_t off = need ? (uint32_t)(e * 16 + 4 * q) : 0xFFFFFFF0u;
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
}

// ALL1: the +-1 neighbours likewise, each read unless its side is solid
// (a -1 neighbour is read only when the case is not backward-solid on that
// axis, a +1 neighbour only when it is not forward-solid: every use of it,
// the second derivatives, the divergence and the pressure difference, drops
// out otherwise).
template <bool ALL1 = false>
__device__ __forceinline__ Star4 load_star4_masked(const float4 *__restrict__ S, __amdgpu_buffer_rsrc_t rs,
                                                   int64_t c, int64_t sj, int64_t si, const Cases &cs) {
    Star4 s;
    if (ALL1) {
        s.XM = buf_load4(rs, c - 1, !cs.bx); s.XP = buf_load4(rs, c + 1, !fwd_solid(cs.cx));
        s.YM = buf_load4(rs, c - sj, !cs.by); s.YP = buf_load4(rs, c + sj, !fwd_solid(cs.cy));
        s.ZM = buf_load4(rs, c - si, !cs.bz); s.ZP = buf_load4(rs, c + si, !fwd_solid(cs.cz));
    } else {
        s.XM = S[c - 1]; s.XP = S[c + 1];
        s.YM = S[c - sj]; s.YP = S[c + sj];
        s.ZM = S[c - si]; s.ZP = S[c + si];
    }
    s.XM2 = buf_load4(rs, c - 2, cs.cx == 5); s.XP2 = buf_load4(rs, c + 2, cs.cx == 4);
    s.YM2 = buf_load4(rs, c - 2 * sj, cs.cy == 5); s.YP2 = buf_load4(rs, c + 2 * sj, cs.cy == 4);
    s.ZM2 = buf_load4(rs, c - 2 * si, cs.cz == 5); s.ZP2 = buf_load4(rs, c + 2 * si, cs.cz == 4);
    return s;
}

// Viscous sums at a cell from its value C and its stencil (the stencil of
// fast_force_s).
__device__ __forceinline__ void visc_from4(const Cases &cs, const Star4 &s, const float4 &C, float &su, float &sv,
                                           float &sw) {
    const float4 &XM = s.XM, &XP = s.XP, &XM2 = s.XM2, &XP2 = s.XP2;
    const float4 &YM = s.YM, &YP = s.YP, &YM2 = s.YM2, &YP2 = s.YP2;
    const float4 &ZM = s.ZM, &ZP = s.ZP, &ZM2 = s.ZM2, &ZP2 = s.ZP2;
    const float lxx = ((cs.bx ? 0.f : XM.y) - 2.0f * C.y) + (fwd_solid(cs.cx) ? 0.f : XP.y);
    const float lyy = ((cs.by ? 0.f : YM.z) - 2.0f * C.z) + (fwd_solid(cs.cy) ? 0.f : YP.z);
    const float lzz = ((cs.bz ? 0.f : ZM.w) - 2.0f * C.w) + (fwd_solid(cs.cz) ? 0.f : ZP.w);
    const float dyx = d2<float>(cs.cx, XM2.z, XM.z, C.z, XP.z, XP2.z);
    const float dzx = d2<float>(cs.cx, XM2.w, XM.w, C.w, XP.w, XP2.w);
    const float dxy = d2<float>(cs.cy, YM2.y, YM.y, C.y, YP.y, YP2.y);
    const float dzy = d2<float>(cs.cy, YM2.w, YM.w, C.w, YP.w, YP2.w);
    const float dxz = d2<float>(cs.cz, ZM2.y, ZM.y, C.y, ZP.y, ZP2.y);
    const float dyz = d2<float>(cs.cz, ZM2.z, ZM.z, C.z, ZP.z, ZP2.z);
    su = (lxx + dxy) + dxz;
    sv = (lyy + dyz) + dyx;
    sw = (lzz + dzx) + dzy;
}

// Viscous sums at c from the interleaved state.
__device__ __forceinline__ void visc_sums4(const Cases &cs, int64_t c, int64_t sj, int64_t si,
                                           const float4 *__restrict__ S, const float4 &C, float &su, float &sv,
                                           float &sw) {
    visc_from4(cs, load_star4(S, c, sj, si), C, su, sv, sw);
}

// Velocities from the forces (vel_voxel's tail); returns the new cell.
__device__ __forceinline__ float4 vel_from4(const Cases &cs, const float4 &C, float su, float sv, float sw,
                                            float dpx, float dpy, float dpz, const float4 &Rm, float p_out,
                                            float cu) {
    const float INV_RE = (float)(1.0 / kRe);
    float fu = cs.bx ? 0.f : su * INV_RE - dpx;
    float fv = cs.by ? 0.f : sv * INV_RE - dpy;
    float fw = cs.bz ? 0.f : sw * INV_RE - dpz;
    fu = fu + Rm.x; fv = fv + Rm.y; fw = fw + Rm.z;
    float4 o;
    o.x = p_out;
    o.y = cs.bx ? 0.f : C.y + fu * cu;
    o.z = cs.by ? 0.f : C.z + fv * cu;
    o.w = cs.bz ? 0.f : C.w + fw * cu;
    return o;
}

// Velocity part of the pair (vel_voxel): writes u, v, w of S1[c] and carries
// p over when first is set (the first pass of a lagged sequence); the pair's
// pressure pass overwrites p.
__device__ __forceinline__ void vel_voxel4(const Grid &g, int64_t c, int t, uint8_t pcv,
                                           const float4 *__restrict__ S0, const float *__restrict__ q0,
                                           float4 *__restrict__ S1, Res4 R, float cu) {
    Cases cs = decode(pcv);
    slip_hold_at(g, c, cs);
    const bool out = on_outlet(g, c);
    const int64_t qc = c - (int64_t)(g.nz + 1) * g.sy;
    const float4 C = S0[c];
    const float pxm = out ? q0[qc - 1] : S0[c - 1].x;
    const float pym = out ? q0[qc - g.sx] : S0[c - g.sx].x;
    const float pzm = S0[c - g.sy].x;
    const float dpx = cs.bx ? 0.f : C.x - pxm;
    const float dpy = cs.by ? 0.f : C.x - pym;
    const float dpz = cs.bz ? 0.f : C.x - pzm;
    float su, sv, sw;
    visc_sums4(cs, c, g.sx, g.sy, S0, C, su, sv, sw);
    S1[c] = vel_from4(cs, C, su, sv, sw, dpx, dpy, dpz, R.m[t], C.x, cu);
}

// Divergence of the interleaved velocities at fluid cell n (pres_voxel /
// lag_div0 order), +1 neighbours through the filled halos.
__device__ __forceinline__ float div4(const Grid &g, int64_t n, uint8_t pcn, const float4 *__restrict__ S,
                                      const float4 &N) {
    const Cases cs = decode(pcn);
    const float wz = fwd_solid(cs.cz) ? 0.f : S[n + g.sy].w;
    const float vy = fwd_solid(cs.cy) ? 0.f : S[n + g.sx].z;
    const float ux = fwd_solid(cs.cx) ? 0.f : S[n + 1].y;
    return ((((wz - N.w) + vy) - N.z) + ux) - N.y;
}

// Pressure part of the pair (pres_voxel), in S1 (velocities already new), or
// in place for the settle pass: p of S1[c] and, on the outlet plane, q1.
__device__ __forceinline__ void pres_voxel4(const Grid &g, int64_t c, uint8_t pcv, const float4 *__restrict__ S0,
                                            float4 *S1, float *__restrict__ q1, Res4 R, float cp, float pin_out) {
    const Cases cs = decode(pcv);
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    const int64_t cxp = (k == g.nx + 1 && g.px) ? c + 1 - g.nx : c + 1;
    const int64_t cyp = (j == g.ny + 1 && g.py) ? c + g.sx - (int64_t)g.ny * g.sx : c + g.sx;
    const int64_t czp = (i == g.nz + 1 && g.pz) ? c + g.sy - (int64_t)g.nz * g.sy : c + g.sy;
    const float4 N = S1[c];
    const float wz = fwd_solid(cs.cz) ? 0.f : S1[czp].w;
    const float vy = fwd_solid(cs.cy) ? 0.f : S1[cyp].z;
    const float ux = fwd_solid(cs.cx) ? 0.f : S1[cxp].y;
    const float div0 = ((((wz - N.w) + vy) - N.z) + ux) - N.y;
    const float div = div0 + R.p[c];
    S1[c].x = S0[c].x - div * cp;
    if (i == g.nz + 1) q1[g.qat(j, k)] = pin_out - div0 * (float)(kC2 * kDt);
}

// Lagged pressure / outlet buffer at a neighbour n (lag_back on the
// interleaved state).
__device__ __forceinline__ float lag_back4(const Grid &g, int64_t c, int64_t s, int idx, int n_ax, bool per, bool out,
                                           int64_t qn, const uint8_t *__restrict__ pc, const float4 *__restrict__ S0,
                                           const float *__restrict__ q0, Res4 R, float cp, float pin_out) {
    int64_t n = c - s;
    if (idx == 2) {
        if (!per) return out ? q0[qn] : S0[n].x;
        n += (int64_t)n_ax * s;
    }
    const float4 N = S0[n];
    const float d0 = div4(g, n, pc[n], S0, N);
    if (out) return pin_out - d0 * (float)(kC2 * kDt);
    return N.x - (d0 + R.p[n]) * cp;
}

// One lagged pass at fluid cell c on the interleaved state (lag_voxel).
__device__ __forceinline__ void lag_voxel4(const Grid &g, int64_t c, int t, uint8_t pcv,
                                           const uint8_t *__restrict__ pc, const float4 *__restrict__ S0,
                                           const float *__restrict__ q0, float4 *__restrict__ S1, Res4 R, float cu,
                                           float cp, float pin_out) {
    Cases cs = decode(pcv);
    slip_hold_at(g, c, cs);
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    const bool out = i == g.nz + 1;
    const float4 C = S0[c];
    const float Pc = out ? pin_out : C.x - (div4(g, c, pcv, S0, C) + R.p[c]) * cp;
    float dpx = 0.f, dpy = 0.f, dpz = 0.f;
    if (!cs.bx)
        dpx = Pc - lag_back4(g, c, 1, k, g.nx, g.px, out, g.qat(j, k - 1), pc, S0, q0, R, cp, pin_out);
    if (!cs.by)
        dpy = Pc - lag_back4(g, c, g.sx, j, g.ny, g.py, out, g.qat(j - 1, k), pc, S0, q0, R, cp, pin_out);
    if (!cs.bz) {
        const int64_t n = c - g.sy;
        if (i == 2) {
            dpz = Pc - S0[n].x;
        } else {
            const float4 N = S0[n];
            dpz = Pc - (N.x - (div4(g, n, pc[n], S0, N) + R.p[n]) * cp);
        }
    }
    float su, sv, sw;
    visc_sums4(cs, c, g.sx, g.sy, S0, C, su, sv, sw);
    S1[c] = vel_from4(cs, C, su, sv, sw, dpx, dpy, dpz, R.m[t], Pc, cu);
}

// Divergence at a cell from its value N, the +1 neighbours' components and
// its code (div4 with the loads done by the caller).
__device__ __forceinline__ float div_from4(uint8_t pcn, const float4 &N, float ux, float vy, float wz) {
    const Cases cs = decode(pcn);
    const float a = fwd_solid(cs.cz) ? 0.f : wz;
    const float b = fwd_solid(cs.cy) ? 0.f : vy;
    const float d = fwd_solid(cs.cx) ? 0.f : ux;
    return ((((a - N.w) + b) - N.z) + d) - N.y;
}

// lag_voxel4 with every load of the voxel issued before any condition is
// evaluated (on a 512^3 field: 1.07 ms per pass against 1.29-1.38 ms): the pressures of the -x, -y and -z neighbours are recomputed
// whether or not the face is open and selected afterwards, so the loads wait
// on the list entry only, not on the cell's code.  Same expressions in the
// same order, so the same values.  Every address stays inside the 2-cell halo
// (c is an interior cell; a periodic image is an interior cell too).
__device__ __forceinline__ void lag_voxel4_eager(const Grid &g, int64_t c, int t, const uint8_t *__restrict__ pc,
                                                 const float4 *__restrict__ S0, const float *__restrict__ q0,
                                                 float4 *__restrict__ S1, Res4 R, float cu, float cp,
                                                 float pin_out) {
    const int64_t sx = g.sx, sy = g.sy;
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    const bool out = i == g.nz + 1;
    // the -x / -y neighbour whose pressure is recomputed: the cell before, or
    // its periodic image at the first interior index
    const bool wx = k == 2 && g.px, wy = j == 2 && g.py;
    const int64_t nxn = wx ? c - 1 + g.nx : c - 1;
    const int64_t nyn = wy ? c - sx + (int64_t)g.ny * sx : c - sx;
    const int64_t nzn = c - sy;
    const uint8_t pcv = pc[c], pcx = pc[nxn], pcy = pc[nyn], pcz = pc[nzn];
    const float rpc = R.p[c], rpx = R.p[nxn], rpy = R.p[nyn], rpz = R.p[nzn];
    const float4 C = S0[c];
    const Star4 s = load_star4(S0, c, sx, sy);
    const float vyx = S0[nxn + sx].z, wzx = S0[nxn + sy].w;
    const float uxy = S0[nyn + 1].y, wzy = S0[nyn + sy].w;
    const float uxz = S0[nzn + 1].y, vyz = S0[nzn + sx].z;
    // the residual entry and the new cell are touched once per pass: streamed
    // (non-temporal), so that the stencil's lines stay in cache
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f r = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(&R.m[t]));
    const float4 Rm = make_float4(r.x, r.y, r.z, r.w);
    // without wrap the neighbour and its +1 cell are already loaded
    float4 NX = s.XM, NY = s.YM;
    float uxx = C.y, vyy = C.z;
    if (wx) { NX = S0[nxn]; uxx = S0[nxn + 1].y; }
    if (wy) { NY = S0[nyn]; vyy = S0[nyn + sx].z; }

    Cases cs = decode(pcv);
    slip_hold_at(g, c, cs);
    const float Pc = out ? pin_out : C.x - (div_from4(pcv, C, s.XP.y, s.YP.z, s.ZP.w) + rpc) * cp;
    const float c2dt = (float)(kC2 * kDt);
    const float dx0 = div_from4(pcx, NX, uxx, vyx, wzx);
    const float dy0 = div_from4(pcy, NY, uxy, vyy, wzy);
    const float dz0 = div_from4(pcz, s.ZM, uxz, vyz, C.w);
    float px = out ? pin_out - dx0 * c2dt : NX.x - (dx0 + rpx) * cp;
    float py = out ? pin_out - dy0 * c2dt : NY.x - (dy0 + rpy) * cp;
    const float pz = i == 2 ? s.ZM.x : s.ZM.x - (dz0 + rpz) * cp;
    // a wall halo keeps its stored value (q on the outlet plane)
    if (k == 2 && !g.px) px = out ? q0[g.qat(j, k - 1)] : s.XM.x;
    if (j == 2 && !g.py) py = out ? q0[g.qat(j - 1, k)] : s.YM.x;
    const float dpx = cs.bx ? 0.f : Pc - px;
    const float dpy = cs.by ? 0.f : Pc - py;
    const float dpz = cs.bz ? 0.f : Pc - pz;
    float su, sv, sw;
    visc_from4(cs, s, C, su, sv, sw);
    const float4 o = vel_from4(cs, C, su, sv, sw, dpx, dpy, dpz, Rm, Pc, cu);
    const v4f ov = {o.x, o.y, o.z, o.w};
    __builtin_nontemporal_store(ov, reinterpret_cast<v4f *>(&S1[c]));
}

// First pass (vel_voxel4 with p carried) from a correction state that is zero
// everywhere, halos and outlet buffer included: the state just after a fold
// or at the start.  Every viscous sum and pressure difference is then +0, so
// vel_from4 gives u = 0 + (0 + Ru) * cu (0 on a closed face) and p = 0: the same
// values from the voxel's record and momentum residual alone.  Also writes
// the cell's periodic halo copies (the ghost fill of the new state; the
// pressure it pins is 0 already).
__device__ __forceinline__ void first_voxel4_zero(const Grid &g, int t, const uint2 *__restrict__ rec,
                                                  float4 *__restrict__ S1, Res4 R, float cu) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    typedef unsigned int v2u __attribute__((ext_vector_type(2)));
    const v2u rc = __builtin_nontemporal_load(reinterpret_cast<const v2u *>(&rec[t]));
    const int64_t c = (int64_t)rc.x;
    Cases cs = decode((uint8_t)(rc.y & 0xff));
    slip_hold_at(g, c, cs);
    const v4f r = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(&R.m[t]));
    const float z = 0.f;
    v4f o;
    o.x = z;
    o.y = cs.bx ? z : z + (z + r.x) * cu;
    o.z = cs.by ? z : z + (z + r.y) * cu;
    o.w = cs.bz ? z : z + (z + r.z) * cu;
    __builtin_nontemporal_store(o, reinterpret_cast<v4f *>(&S1[c]));
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    const int64_t sx = g.sx, sy = g.sy;
    if (g.pz) {
        const int64_t zs = (int64_t)g.nz * sy;
        if (i >= g.nz) __builtin_nontemporal_store(o, reinterpret_cast<v4f *>(&S1[c - zs]));
        if (i <= 3) __builtin_nontemporal_store(o, reinterpret_cast<v4f *>(&S1[c + zs]));
    }
    if (g.px) {
        if (k >= g.nx) __builtin_nontemporal_store(o, reinterpret_cast<v4f *>(&S1[c - g.nx]));
        if (k <= 3) __builtin_nontemporal_store(o, reinterpret_cast<v4f *>(&S1[c + g.nx]));
    }
    if (g.py) {
        const int64_t ys = (int64_t)g.ny * sx;
        if (j >= g.ny) __builtin_nontemporal_store(o, reinterpret_cast<v4f *>(&S1[c - ys]));
        if (j <= 3) __builtin_nontemporal_store(o, reinterpret_cast<v4f *>(&S1[c + ys]));
    }
    slip_images4(g, c, j, k, o, S1);
}

// The cells whose pressure the lagged pass at c recomputes: the -x and -y
// neighbours, or their periodic images when c is on the first interior
// index of a periodic axis, and the -z neighbour.
__device__ __forceinline__ void lag_targets(const Grid &g, int64_t c, int i, int j, int k, int64_t &nxn,
                                            int64_t &nyn, int64_t &nzn) {
    (void)i;
    nxn = (k == 2 && g.px) ? c - 1 + g.nx : c - 1;
    nyn = (j == 2 && g.py) ? c - g.sx + (int64_t)g.ny * g.sx : c - g.sx;
    nzn = c - g.sy;
}

// Per-voxel record of the lagged pass: the cell's element offset and the
// packed codes of the cell and of its three recompute targets (8 bits each,
// cell lowest).
__device__ __forceinline__ uint2 lag_record(const Grid &g, int64_t c, const uint8_t *__restrict__ pc) {
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    int64_t nxn, nyn, nzn;
    lag_targets(g, c, i, j, k, nxn, nyn, nzn);
    const uint32_t codes = (uint32_t)pc[c] | ((uint32_t)pc[nxn] << 8) | ((uint32_t)pc[nyn] << 16) |
                           ((uint32_t)pc[nzn] << 24);
    return make_uint2((uint32_t)c, codes);
}

// lag_voxel4_eager with the codes from the voxel's record and the four
// continuity residuals from R.q: two streamed loads instead of eight gathers,
// the same values.
// SKIP2: the +-2 neighbours through load_star4_masked (the state below 2^32
// bytes); SKIP1 (with SKIP2): the +-1 neighbours and the six values of the
// recompute targets too, each read only where it is used (the targets of a
// solid side are never used; a target's forward velocity only where that
// face of the target is open).
template <bool IMAGES = false, bool SKIP2 = false, bool SKIP1 = false>
__device__ __forceinline__ void lag_voxel4_rec(const Grid &g, int t, const uint2 *__restrict__ rec,
                                               const float4 *__restrict__ S0, const float *__restrict__ q0,
                                               float4 *__restrict__ S1, Res4 R, float cu, float cp, float pin_out) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    typedef unsigned int v2u __attribute__((ext_vector_type(2)));
    const v2u rc = __builtin_nontemporal_load(reinterpret_cast<const v2u *>(&rec[t]));
    const int64_t c = (int64_t)rc.x;
    const uint8_t pcv = (uint8_t)(rc.y & 0xff), pcx = (uint8_t)((rc.y >> 8) & 0xff),
                  pcy = (uint8_t)((rc.y >> 16) & 0xff), pcz = (uint8_t)(rc.y >> 24);
    const int64_t sx = g.sx, sy = g.sy;
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    const bool out = i == g.nz + 1;
    const bool wx = k == 2 && g.px, wy = j == 2 && g.py;
    int64_t nxn, nyn, nzn;
    lag_targets(g, c, i, j, k, nxn, nyn, nzn);
    const v4f rq = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(&R.q[t]));
    const float rpc = rq.x, rpx = rq.y, rpy = rq.z, rpz = rq.w;
    const float4 C = S0[c];
    Star4 s;
    float vyx, wzx, uxy, wzy, uxz, vyz;
    if (SKIP2) {
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void *)S0, (short)0, (int)(uint32_t)(g.total * 16), 0x00020000);
        const Cases cc = decode(pcv);
        s = load_star4_masked<SKIP1>(S0, rs, c, sx, sy, cc);
        if (SKIP1) {
            const Cases tx = decode(pcx), ty = decode(pcy), tz = decode(pcz);
            vyx = buf_load1(rs, nxn + sx, 2, !cc.bx && !fwd_solid(tx.cy));
            wzx = buf_load1(rs, nxn + sy, 3, !cc.bx && !fwd_solid(tx.cz));
            uxy = buf_load1(rs, nyn + 1, 1, !cc.by && !fwd_solid(ty.cx));
            wzy = buf_load1(rs, nyn + sy, 3, !cc.by && !fwd_solid(ty.cz));
            uxz = buf_load1(rs, nzn + 1, 1, !cc.bz && !fwd_solid(tz.cx));
            vyz = buf_load1(rs, nzn + sx, 2, !cc.bz && !fwd_solid(tz.cy));
        } else {
            vyx = S0[nxn + sx].z; wzx = S0[nxn + sy].w;
            uxy = S0[nyn + 1].y; wzy = S0[nyn + sy].w;
            uxz = S0[nzn + 1].y; vyz = S0[nzn + sx].z;
        }
    } else {
        s = load_star4(S0, c, sx, sy);
        vyx = S0[nxn + sx].z; wzx = S0[nxn + sy].w;
        uxy = S0[nyn + 1].y; wzy = S0[nyn + sy].w;
        uxz = S0[nzn + 1].y; vyz = S0[nzn + sx].z;
    }
    const v4f r = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(&R.m[t]));
    const float4 Rm = make_float4(r.x, r.y, r.z, r.w);
    float4 NX = s.XM, NY = s.YM;
    float uxx = C.y, vyy = C.z;
    if (wx) { NX = S0[nxn]; uxx = S0[nxn + 1].y; }
    if (wy) { NY = S0[nyn]; vyy = S0[nyn + sx].z; }

    Cases cs = decode(pcv);
    slip_hold(g, j, k, cs);
    const float Pc = out ? pin_out : C.x - (div_from4(pcv, C, s.XP.y, s.YP.z, s.ZP.w) + rpc) * cp;
    const float c2dt = (float)(kC2 * kDt);
    const float dx0 = div_from4(pcx, NX, uxx, vyx, wzx);
    const float dy0 = div_from4(pcy, NY, uxy, vyy, wzy);
    const float dz0 = div_from4(pcz, s.ZM, uxz, vyz, C.w);
    float px = out ? pin_out - dx0 * c2dt : NX.x - (dx0 + rpx) * cp;
    float py = out ? pin_out - dy0 * c2dt : NY.x - (dy0 + rpy) * cp;
    const float pz = i == 2 ? s.ZM.x : s.ZM.x - (dz0 + rpz) * cp;
    if (k == 2 && !g.px) px = out ? q0[g.qat(j, k - 1)] : s.XM.x;
    if (j == 2 && !g.py) py = out ? q0[g.qat(j - 1, k)] : s.YM.x;
    const float dpx = cs.bx ? 0.f : Pc - px;
    const float dpy = cs.by ? 0.f : Pc - py;
    const float dpz = cs.bz ? 0.f : Pc - pz;
    float su, sv, sw;
    visc_from4(cs, s, C, su, sv, sw);
    const float4 o = vel_from4(cs, C, su, sv, sw, dpx, dpy, dpz, Rm, Pc, cu);
    const v4f ov = {o.x, o.y, o.z, o.w};
    __builtin_nontemporal_store(ov, reinterpret_cast<v4f *>(&S1[c]));
    if (IMAGES) {
        // The ghost fill of the new state (ghost_line4), cell by cell: the
        // cell's copies in the periodic halos, with the pressure pinned to 0
        // on halo planes 1 and nz+2 (the copies of planes nz+1 and 2; the
        // geometry's halos are periodic too, so the copy is fluid).  An outlet
        // cell's pressure is already pin_out = 0.  Solid cells' copies stay 0,
        // wall halos are never written, and the lagged pass leaves the outlet
        // buffer alone, so nothing else of the fill changes the state.
        if (g.pz) {
            const int64_t zs = (int64_t)g.nz * sy;
            if (i >= g.nz) {
                v4f w = ov;
                if (i == g.nz + 1) w.x = 0.f;
                __builtin_nontemporal_store(w, reinterpret_cast<v4f *>(&S1[c - zs]));
            }
            if (i <= 3) {
                v4f w = ov;
                if (i == 2) w.x = 0.f;
                __builtin_nontemporal_store(w, reinterpret_cast<v4f *>(&S1[c + zs]));
            }
        }
        if (g.px) {
            if (k >= g.nx) __builtin_nontemporal_store(ov, reinterpret_cast<v4f *>(&S1[c - g.nx]));
            if (k <= 3) __builtin_nontemporal_store(ov, reinterpret_cast<v4f *>(&S1[c + g.nx]));
        }
        if (g.py) {
            const int64_t ys = (int64_t)g.ny * sx;
            if (j >= g.ny) __builtin_nontemporal_store(ov, reinterpret_cast<v4f *>(&S1[c - ys]));
            if (j <= 3) __builtin_nontemporal_store(ov, reinterpret_cast<v4f *>(&S1[c + ys]));
        }
        slip_images4(g, c, j, k, ov, S1);
    }
}

// lag_voxel (per-field state, SEL form) from the voxel's record, every load
// issued before any condition: the state of the cell, its 12 stencil
// neighbours per field, the pressures of the three recompute targets and the
// six diagonal velocities their divergences need.  The same expressions in the
// same order as lag_voxel, so the same values.
// An element of a per-field array the same way as buf_load4 (element bytes
// sizeof(T); the array below 2^32 bytes).
template <typename T>
__device__ __forceinline__ T buf_load_t(__amdgpu_buffer_rsrc_t rs, int64_t e, bool need) {
    const uint32_t off = need ? (uint32_t)(e * (int64_t)sizeof(T)) : 0xFFFFFFF0u;
    if constexpr (sizeof(T) == 8) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
    }
}

// SKIP2: the twelve +-2 loads only for the lanes whose case reads them
// (buf_load_t; each field below 2^32 bytes).
template <typename T, bool SKIP2 = false>
__device__ __forceinline__ void lag_voxel_rec(const Grid &g, int t, const uint2 *__restrict__ rec,
                                              const T *__restrict__ p0, const T *__restrict__ u0,
                                              const T *__restrict__ v0, const T *__restrict__ w0,
                                              const T *__restrict__ q0, T *__restrict__ p1, T *__restrict__ u1,
                                              T *__restrict__ v1, T *__restrict__ w1, T cu, T cp, T pin_out) {
    typedef unsigned int v2u __attribute__((ext_vector_type(2)));
    const v2u rc = __builtin_nontemporal_load(reinterpret_cast<const v2u *>(&rec[t]));
    const int64_t c = (int64_t)rc.x;
    const uint8_t pcv = (uint8_t)(rc.y & 0xff), pcx = (uint8_t)((rc.y >> 8) & 0xff),
                  pcy = (uint8_t)((rc.y >> 16) & 0xff), pcz = (uint8_t)(rc.y >> 24);
    const int64_t sx = g.sx, sy = g.sy;
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    const bool out = i == g.nz + 1;
    const bool wx = k == 2 && g.px, wy = j == 2 && g.py;
    int64_t nxn, nyn, nzn;
    lag_targets(g, c, i, j, k, nxn, nyn, nzn);
    // the cell and its stencil
    const T pc_ = p0[c], uc = u0[c], vc = v0[c], wc = w0[c];
    const T uxm = u0[c - 1], uxp = u0[c + 1];
    const T uym = u0[c - sx], uyp = u0[c + sx];
    const T uzm = u0[c - sy], uzp = u0[c + sy];
    const T vym = v0[c - sx], vyp = v0[c + sx];
    const T vxm = v0[c - 1], vxp = v0[c + 1];
    const T vzm = v0[c - sy], vzp = v0[c + sy];
    const T wzm = w0[c - sy], wzp = w0[c + sy];
    const T wxm = w0[c - 1], wxp = w0[c + 1];
    const T wym = w0[c - sx], wyp = w0[c + sx];
    T uym2, uyp2, uzm2, uzp2, vxm2, vxp2, vzm2, vzp2, wxm2, wxp2, wym2, wyp2;
    if (SKIP2) {
        const int nb = (int)(uint32_t)(g.total * (int64_t)sizeof(T));
        const __amdgpu_buffer_rsrc_t ru = __builtin_amdgcn_make_buffer_rsrc((void *)u0, (short)0, nb, 0x00020000);
        const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc((void *)v0, (short)0, nb, 0x00020000);
        const __amdgpu_buffer_rsrc_t rw = __builtin_amdgcn_make_buffer_rsrc((void *)w0, (short)0, nb, 0x00020000);
        const Cases c2 = decode(pcv);
        uym2 = buf_load_t<T>(ru, c - 2 * sx, c2.cy == 5); uyp2 = buf_load_t<T>(ru, c + 2 * sx, c2.cy == 4);
        uzm2 = buf_load_t<T>(ru, c - 2 * sy, c2.cz == 5); uzp2 = buf_load_t<T>(ru, c + 2 * sy, c2.cz == 4);
        vxm2 = buf_load_t<T>(rv, c - 2, c2.cx == 5); vxp2 = buf_load_t<T>(rv, c + 2, c2.cx == 4);
        vzm2 = buf_load_t<T>(rv, c - 2 * sy, c2.cz == 5); vzp2 = buf_load_t<T>(rv, c + 2 * sy, c2.cz == 4);
        wxm2 = buf_load_t<T>(rw, c - 2, c2.cx == 5); wxp2 = buf_load_t<T>(rw, c + 2, c2.cx == 4);
        wym2 = buf_load_t<T>(rw, c - 2 * sx, c2.cy == 5); wyp2 = buf_load_t<T>(rw, c + 2 * sx, c2.cy == 4);
    } else {
        uym2 = u0[c - 2 * sx]; uyp2 = u0[c + 2 * sx]; uzm2 = u0[c - 2 * sy]; uzp2 = u0[c + 2 * sy];
        vxm2 = v0[c - 2]; vxp2 = v0[c + 2]; vzm2 = v0[c - 2 * sy]; vzp2 = v0[c + 2 * sy];
        wxm2 = w0[c - 2]; wxp2 = w0[c + 2]; wym2 = w0[c - 2 * sx]; wyp2 = w0[c + 2 * sx];
    }
    // recompute targets: pressures and the +1 velocities their divergences need
    const T pxt = p0[nxn], pyt = p0[nyn], pzt = p0[nzn];
    const T vx_d = v0[nxn + sx], wx_d = w0[nxn + sy];
    const T uy_d = u0[nyn + 1], wy_d = w0[nyn + sy];
    const T uz_d = u0[nzn + 1], vz_d = v0[nzn + sx];
    // the -x / -y targets' own velocities and +1 neighbour: already loaded
    // unless the target is a periodic image
    T uX = uxm, vX = vxm, wX = wxm, uXp = uc;
    T uY = uym, vY = vym, wY = wym, vYp = vc;
    if (wx) { uX = u0[nxn]; vX = v0[nxn]; wX = w0[nxn]; uXp = u0[nxn + 1]; }
    if (wy) { uY = u0[nyn]; vY = v0[nyn]; wY = w0[nyn]; vYp = v0[nyn + sx]; }

    Cases cs = decode(pcv);
    slip_hold(g, j, k, cs);
    // lag_div0 at the cell and at the three targets
    const auto div0 = [](uint8_t code, T u, T v, T w, T uxp1, T vyp1, T wzp1) {
        const Cases d = decode(code);
        const T wz = fwd_solid(d.cz) ? (T)0 : wzp1;
        const T vy = fwd_solid(d.cy) ? (T)0 : vyp1;
        const T ux = fwd_solid(d.cx) ? (T)0 : uxp1;
        return ((((wz - w) + vy) - v) + ux) - u;
    };
    const T C2DT = (T)(kC2 * kDt);
    const T Pc = out ? pin_out : pc_ - div0(pcv, uc, vc, wc, uxp, vyp, wzp) * cp;
    T dpx = (T)0, dpy = (T)0, dpz = (T)0;
    if (!cs.bx) {
        T bx;
        if (k == 2 && !g.px) bx = out ? q0[g.qat(j, k - 1)] : p0[c - 1];
        else {
            const T d = div0(pcx, uX, vX, wX, uXp, vx_d, wx_d);
            bx = out ? pin_out - d * C2DT : pxt - d * cp;
        }
        dpx = Pc - bx;
    }
    if (!cs.by) {
        T by;
        if (j == 2 && !g.py) by = out ? q0[g.qat(j - 1, k)] : p0[c - sx];
        else {
            const T d = div0(pcy, uY, vY, wY, uy_d, vYp, wy_d);
            by = out ? pin_out - d * C2DT : pyt - d * cp;
        }
        dpy = Pc - by;
    }
    if (!cs.bz) dpz = Pc - (i == 2 ? pzt : pzt - div0(pcz, uzm, vzm, wzm, uz_d, vz_d, wc) * cp);
    // visc_sums, SEL form
    const T lxx = ((cs.bx ? (T)0 : uxm) - (T)2.0 * uc) + (fwd_solid(cs.cx) ? (T)0 : uxp);
    const T lyy = ((cs.by ? (T)0 : vym) - (T)2.0 * vc) + (fwd_solid(cs.cy) ? (T)0 : vyp);
    const T lzz = ((cs.bz ? (T)0 : wzm) - (T)2.0 * wc) + (fwd_solid(cs.cz) ? (T)0 : wzp);
    const T dyx = d2_pick<T>(cs.cx, vxm2, vxm, vc, vxp, vxp2);
    const T dzx = d2_pick<T>(cs.cx, wxm2, wxm, wc, wxp, wxp2);
    const T dxy = d2_pick<T>(cs.cy, uym2, uym, uc, uyp, uyp2);
    const T dzy = d2_pick<T>(cs.cy, wym2, wym, wc, wyp, wyp2);
    const T dxz = d2_pick<T>(cs.cz, uzm2, uzm, uc, uzp, uzp2);
    const T dyz = d2_pick<T>(cs.cz, vzm2, vzm, vc, vzp, vzp2);
    const T su = (lxx + dxy) + dxz;
    const T sv = (lyy + dyz) + dyx;
    const T sw = (lzz + dzx) + dzy;
    const T INV_RE = (T)(1.0 / kRe);
    const T fu = cs.bx ? (T)0 : su * INV_RE - dpx;
    const T fv = cs.by ? (T)0 : sv * INV_RE - dpy;
    const T fw = cs.bz ? (T)0 : sw * INV_RE - dpz;
    u1[c] = cs.bx ? (T)0 : uc + fu * cu;
    v1[c] = cs.by ? (T)0 : vc + fv * cu;
    w1[c] = cs.bz ? (T)0 : wc + fw * cu;
    p1[c] = Pc;
}

// One 3-colour SOR update at fluid cell c on the interleaved state, in place
// (colour_voxel): velocities from the current state, then the pressure from
// the new velocities at c and the current ones of its +1 neighbours.  Every
// stencil neighbour has another colour, so nothing read here is written by the
// same launch.  t is the cell's position in the fluid list the residual's
// momentum part follows.  All loads are issued before any condition.
__device__ __forceinline__ void colour_voxel4(const Grid &g, int64_t c, int t, const uint8_t *__restrict__ pc,
                                              float4 *S, float *q, Res4 R, float cu, float cp, float pin_out) {
    const uint8_t pcv = pc[c];
    const float4 C = S[c];
    const Star4 s = load_star4(S, c, g.sx, g.sy);
    const float rp = R.p[c];
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f r = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(&R.m[t]));
    const float4 Rm = make_float4(r.x, r.y, r.z, r.w);
    const bool out = on_outlet(g, c);
    const int64_t qc = c - (int64_t)(g.nz + 1) * g.sy;
    float pxm = s.XM.x, pym = s.YM.x;
    if (out) { pxm = q[qc - 1]; pym = q[qc - g.sx]; }
    Cases cs = decode(pcv);
    slip_hold_at(g, c, cs);
    const float dpx = cs.bx ? 0.f : C.x - pxm;
    const float dpy = cs.by ? 0.f : C.x - pym;
    const float dpz = cs.bz ? 0.f : C.x - s.ZM.x;
    float su, sv, sw;
    visc_from4(cs, s, C, su, sv, sw);
    float4 N = vel_from4(cs, C, su, sv, sw, dpx, dpy, dpz, Rm, C.x, cu);
    const float wz = fwd_solid(cs.cz) ? 0.f : s.ZP.w;
    const float vy = fwd_solid(cs.cy) ? 0.f : s.YP.z;
    const float ux = fwd_solid(cs.cx) ? 0.f : s.XP.y;
    const float div0 = ((((wz - N.w) + vy) - N.z) + ux) - N.y;
    const float div = div0 + rp;
    N.x = C.x - div * cp;
    S[c] = N;
    if (out) q[qc] = pin_out - div0 * (float)(kC2 * kDt);
}

// Record of a fluid voxel for the colour-split SOR pass: its colour-split
// offset and its packed code, block k mod 3 (bits 8-9) and outlet flag (bit
// 10).
__device__ __forceinline__ uint2 colour_record(const Grid &g, int64_t c, const uint8_t *__restrict__ pc) {
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    const uint32_t f = (uint32_t)pc[c] | ((uint32_t)(k % 3) << 8) | ((uint32_t)(i == g.nz + 1) << 10);
    return make_uint2((uint32_t)g.at_s(i, j, k), f);
}

// colour_voxel4 on the colour-split state from the voxel's record: the same
// values in the same order, from the same logical cells.  The x neighbours
// lie in the other blocks of the row: k-1 is block b-1 (slot m-1 from block
// 0), k+1 block b+1 (slot m+1 from block 2), k-2 and k+2 likewise.  The
// continuity residual is read from the w lane of the momentum residual.
// SKIP2: the +-2 neighbours only for the lanes whose case reads them
// (buf_load4; the state below 2^32 bytes).
template <bool SKIP2 = false>
__device__ __forceinline__ void colour_voxel4_split(const Grid &g, int t, const uint2 *__restrict__ rec,
                                                    float4 *S, float *q, Res4 R, float cu, float cp,
                                                    float pin_out) {
    typedef unsigned int v2u __attribute__((ext_vector_type(2)));
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v2u rr = __builtin_nontemporal_load(reinterpret_cast<const v2u *>(&rec[t]));
    const int64_t c = rr.x;
    const uint32_t f = rr.y;
    const int b = (int)((f >> 8) & 3u);
    const bool out = (f >> 10) & 1u;
    const int64_t cb = g.cb, sj = g.sx, si = g.sy;
    const int64_t oM1 = b == 0 ? 2 * cb - 1 : -cb, oP1 = b == 2 ? 1 - 2 * cb : cb;
    const int64_t oM2 = b == 2 ? -2 * cb : cb - 1, oP2 = b == 0 ? 2 * cb : 1 - cb;
    const float4 C = S[c];
    Star4 s;
    s.XM = S[c + oM1]; s.XP = S[c + oP1];
    s.YM = S[c - sj]; s.YP = S[c + sj];
    s.ZM = S[c - si]; s.ZP = S[c + si];
    if (SKIP2) {
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void *)S, (short)0, (int)(uint32_t)(g.total * 16), 0x00020000);
        const Cases c2 = decode((uint8_t)(f & 0xffu));
        s.XM2 = buf_load4(rs, c + oM2, c2.cx == 5); s.XP2 = buf_load4(rs, c + oP2, c2.cx == 4);
        s.YM2 = buf_load4(rs, c - 2 * sj, c2.cy == 5); s.YP2 = buf_load4(rs, c + 2 * sj, c2.cy == 4);
        s.ZM2 = buf_load4(rs, c - 2 * si, c2.cz == 5); s.ZP2 = buf_load4(rs, c + 2 * si, c2.cz == 4);
    } else {
        s.XM2 = S[c + oM2]; s.XP2 = S[c + oP2];
        s.YM2 = S[c - 2 * sj]; s.YP2 = S[c + 2 * sj];
        s.ZM2 = S[c - 2 * si]; s.ZP2 = S[c + 2 * si];
    }
    const v4f r = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(&R.m[t]));
    const float4 Rm = make_float4(r.x, r.y, r.z, 0.f);
    const float rp = r.w;
    const int64_t qc = c - (int64_t)(g.nz + 1) * si;
    float pxm = s.XM.x, pym = s.YM.x;
    if (out) { pxm = q[qc + oM1]; pym = q[qc - sj]; }
    Cases cs = decode((uint8_t)(f & 0xffu));
    if (g.slip) {
        // c is a colour-split offset: the first interior column (k = 2) is
        // position 0 of block 2 of its row
        int i2, j2, r2;
        cell_ijk(g, c, i2, j2, r2);
        slip_hold(g, j2, r2 == 2 * g.cb ? 2 : -1, cs);
    }
    const float dpx = cs.bx ? 0.f : C.x - pxm;
    const float dpy = cs.by ? 0.f : C.x - pym;
    const float dpz = cs.bz ? 0.f : C.x - s.ZM.x;
    float su, sv, sw;
    visc_from4(cs, s, C, su, sv, sw);
    float4 N = vel_from4(cs, C, su, sv, sw, dpx, dpy, dpz, Rm, C.x, cu);
    const float wz = fwd_solid(cs.cz) ? 0.f : s.ZP.w;
    const float vy = fwd_solid(cs.cy) ? 0.f : s.YP.z;
    const float ux = fwd_solid(cs.cx) ? 0.f : s.XP.y;
    const float div0 = ((((wz - N.w) + vy) - N.z) + ux) - N.y;
    const float div = div0 + rp;
    N.x = C.x - div * cp;
    S[c] = N;
    if (out) q[qc] = pin_out - div0 * (float)(kC2 * kDt);
}

// Residual of the FP64 state at c into the interleaved FP32 residual.
__device__ __forceinline__ void residual_voxel4(const Grid &g, int64_t c, int t, uint8_t pcv,
                                                const double *__restrict__ p, const double *__restrict__ u,
                                                const double *__restrict__ v, const double *__restrict__ w,
                                                const double *__restrict__ q, Res4 R) {
    const Cases cs = decode(pcv);
    const bool out = on_outlet(g, c);
    double fu, fv, fw;
    fast_force_p<double>(cs, c, g.sx, g.sy, p, u, v, w, out ? q : nullptr, c - (int64_t)(g.nz + 1) * g.sy, fu,
                         fv, fw);
    const double wz = fwd_solid(cs.cz) ? 0.0 : w[c + g.sy];
    const double vy = fwd_solid(cs.cy) ? 0.0 : v[c + g.sx];
    const double ux = fwd_solid(cs.cx) ? 0.0 : u[c + 1];
    const double div = ((((wz - w[c]) + vy) - v[c]) + ux) - u[c];
    R.p[c] = (float)div;
    R.m[t] = make_float4((float)fu, (float)fv, (float)fw, (float)div);
}

// Ghost fill of the interleaved correction state (ghost_line with float4
// copies; the pins of a correction state are 0).  SPLIT: the state and its
// outlet buffer in colour-split rows (the geometry G stays natural).
template <bool SPLIT = false>
__device__ __forceinline__ void ghost_line4(const Grid &g, int64_t t, const uint8_t *__restrict__ G, float4 *S,
                                            float *q) {
    const int64_t nX = (int64_t)g.nz * g.ny, nY = (int64_t)g.nz * g.nx, nZ = (int64_t)g.ny * g.nx;
    const int iout = g.nz + 1;
    auto sat = [&](int i, int j, int k) { return SPLIT ? g.at_s(i, j, k) : g.at(i, j, k); };
    auto sqat = [&](int j, int k) { return SPLIT ? g.qat_s(j, k) : g.qat(j, k); };
    if (t < nX) {
        const int i = (int)(t / g.ny) + 2, j = (int)(t % g.ny) + 2;
        if (!g.px) {
            // slip walls: v, w mirrored, u and p 0
            if (g.slip)
                for (int e = 0; e < 2; e++) {
                    float4 m = S[sat(i, j, 2 + e)], h = S[sat(i, j, g.nx + 1 - e)];
                    m.x = 0.f; m.y = 0.f; h.x = 0.f; h.y = 0.f;
                    S[sat(i, j, 1 - e)] = m;
                    S[sat(i, j, g.nx + 2 + e)] = h;
                }
            return;
        }
        const int src[4] = {g.nx, g.nx + 1, 2, 3};
        const int dst[4] = {0, 1, g.nx + 2, g.nx + 3};
        for (int e = 0; e < 4; e++) {
            float4 v = S[sat(i, j, src[e])];
            if (i == iout && G[g.at(i, j, src[e])] == 0) v.x = 0.f;
            S[sat(i, j, dst[e])] = v;
            if (q && i == iout) q[sqat(j, dst[e])] = q[sqat(j, src[e])];
        }
    } else if (t < nX + nY) {
        const int64_t tt = t - nX;
        const int i = (int)(tt / g.nx) + 2, k = (int)(tt % g.nx) + 2;
        if (!g.py) {
            // slip walls: u, w mirrored, v and p 0
            if (g.slip)
                for (int e = 0; e < 2; e++) {
                    float4 m = S[sat(i, 2 + e, k)], h = S[sat(i, g.ny + 1 - e, k)];
                    m.x = 0.f; m.z = 0.f; h.x = 0.f; h.z = 0.f;
                    S[sat(i, 1 - e, k)] = m;
                    S[sat(i, g.ny + 2 + e, k)] = h;
                }
            return;
        }
        const int src[4] = {g.ny, g.ny + 1, 2, 3};
        const int dst[4] = {0, 1, g.ny + 2, g.ny + 3};
        for (int e = 0; e < 4; e++) {
            float4 v = S[sat(i, src[e], k)];
            if (i == iout && G[g.at(i, src[e], k)] == 0) v.x = 0.f;
            S[sat(i, dst[e], k)] = v;
            if (q && i == iout) q[sqat(dst[e], k)] = q[sqat(src[e], k)];
        }
    } else if (t < nX + nY + nZ) {
        const int64_t tt = t - nX - nY;
        const int j = (int)(tt / g.nx) + 2, k = (int)(tt % g.nx) + 2;
        const bool out_fluid = G[g.at(iout, j, k)] == 0;
        if (g.pz) {
            const int src[4] = {g.nz, g.nz + 1, 2, 3};
            const int dst[4] = {0, 1, g.nz + 2, g.nz + 3};
            for (int e = 0; e < 4; e++) {
                float4 v = S[sat(src[e], j, k)];
                if (src[e] == iout && out_fluid) v.x = 0.f;
                S[sat(dst[e], j, k)] = v;
            }
        }
        if (G[g.at(1, j, k)] == 0) S[sat(1, j, k)].x = 0.f;
        if (G[g.at(g.nz + 2, j, k)] == 0) S[sat(g.nz + 2, j, k)].x = 0.f;
        if (out_fluid) S[sat(iout, j, k)].x = 0.f;
    }
}

// Fold of the interleaved correction into the FP64 state at fluid interior
// cell c: X += d, d = 0.  Only fluid cells carry a correction; the caller
// refills the halo copies of d (with zeros) by a ghost fill afterwards.
__device__ __forceinline__ void fold_voxel4(int64_t c, double *Xp, double *Xu, double *Xv, double *Xw, float4 *d) {
    const float4 v = d[c];
    Xp[c] = Xp[c] + (double)v.x;
    Xu[c] = Xu[c] + (double)v.y;
    Xv[c] = Xv[c] + (double)v.z;
    Xw[c] = Xw[c] + (double)v.w;
    d[c] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// fold_voxel4 for a colour-split correction state: c natural, cs its
// colour-split offset.
__device__ __forceinline__ void fold_voxel4_split(int64_t c, int64_t cs, double *Xp, double *Xu, double *Xv,
                                                  double *Xw, float4 *d) {
    const float4 v = d[cs];
    Xp[c] = Xp[c] + (double)v.x;
    Xu[c] = Xu[c] + (double)v.y;
    Xv[c] = Xv[c] + (double)v.z;
    Xw[c] = Xw[c] + (double)v.w;
    d[cs] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// 3-colour SOR: one cell of the current colour, in place.
template <typename T, bool SEL>
__device__ __forceinline__ void colour_voxel(const Grid &g, int64_t c, uint8_t pcv, T *p, T *u, T *v, T *w,
                                             T *q, T cu, T cp, T pin_out) {
    Cases cs = decode(pcv);
    slip_hold_at(g, c, cs);
    const bool out = on_outlet(g, c);
    const int64_t qc = c - (int64_t)(g.nz + 1) * g.sy;
    T fu, fv, fw;
    if (SEL) fast_force_s<T>(cs, c, g.sx, g.sy, p, u, v, w, out ? q : nullptr, qc, fu, fv, fw);
    else fast_force_p<T>(cs, c, g.sx, g.sy, p, u, v, w, out ? q : nullptr, qc, fu, fv, fw);
    const T un = cs.bx ? (T)0 : u[c] + fu * cu;
    const T vn = cs.by ? (T)0 : v[c] + fv * cu;
    const T wn = cs.bz ? (T)0 : w[c] + fw * cu;
    u[c] = un; v[c] = vn; w[c] = wn;
    const T wz = fwd_solid(cs.cz) ? (T)0 : w[c + g.sy];
    const T vy = fwd_solid(cs.cy) ? (T)0 : v[c + g.sx];
    const T ux = fwd_solid(cs.cx) ? (T)0 : u[c + 1];
    const T div0 = ((((wz - wn) + vy) - vn) + ux) - un;
    p[c] = p[c] - div0 * cp;
    if (out) q[qc] = pin_out - div0 * (T)(kC2 * kDt);
}

// Number of halo lines the fast-mode ghost fill walks (x-face, y-face and
// z-face lines of the interior).
__host__ __device__ inline int64_t ghost_lines(const Grid &g) {
    return (int64_t)g.nz * g.ny + (int64_t)g.nz * g.nx + (int64_t)g.ny * g.nx;
}

// Fast-mode ghost fill, line t: the outlet plane's interior is pinned,
// periodic faces are copied (the outlet plane's copies taken from the pinned
// values), then the inlet/top halo planes are pinned.  Corner halo cells are
// never read by the axis-aligned stencil and are left alone.
template <typename T>
__device__ __forceinline__ void ghost_line(const Grid &g, int64_t t, const uint8_t *__restrict__ G, T *p, T *u,
                                           T *v, T *w, T *q, T pin_in, T pin_hi, T pin_out) {
    const int64_t nX = (int64_t)g.nz * g.ny, nY = (int64_t)g.nz * g.nx, nZ = (int64_t)g.ny * g.nx;
    const int iout = g.nz + 1;
    if (t < nX) {
        const int i = (int)(t / g.ny) + 2, j = (int)(t % g.ny) + 2;
        const int64_t r = g.at(i, j, 0);
        if (!g.px) {
            // slip walls: v, w mirrored, u 0, p left as it is
            if (g.slip)
                for (int e = 0; e < 2; e++) {
                    const int64_t lo = r + 1 - e, lm = r + 2 + e, hi = r + g.nx + 2 + e, hm = r + g.nx + 1 - e;
                    u[lo] = (T)0; v[lo] = v[lm]; w[lo] = w[lm];
                    u[hi] = (T)0; v[hi] = v[hm]; w[hi] = w[hm];
                }
            return;
        }
        const int src[4] = {g.nx, g.nx + 1, 2, 3};
        const int dst[4] = {0, 1, g.nx + 2, g.nx + 3};
        for (int e = 0; e < 4; e++) {
            const int64_t s = r + src[e], d = r + dst[e];
            u[d] = u[s]; v[d] = v[s]; w[d] = w[s];
            p[d] = (i == iout && G[s] == 0) ? pin_out : p[s];
            if (q && i == iout) q[g.qat(j, dst[e])] = q[g.qat(j, src[e])];
        }
    } else if (t < nX + nY) {
        const int64_t tt = t - nX;
        const int i = (int)(tt / g.nx) + 2, k = (int)(tt % g.nx) + 2;
        if (!g.py) {
            // slip walls: u, w mirrored, v 0, p left as it is
            if (g.slip)
                for (int e = 0; e < 2; e++) {
                    const int64_t lo = g.at(i, 1 - e, k), lm = g.at(i, 2 + e, k);
                    const int64_t hi = g.at(i, g.ny + 2 + e, k), hm = g.at(i, g.ny + 1 - e, k);
                    v[lo] = (T)0; u[lo] = u[lm]; w[lo] = w[lm];
                    v[hi] = (T)0; u[hi] = u[hm]; w[hi] = w[hm];
                }
            return;
        }
        const int src[4] = {g.ny, g.ny + 1, 2, 3};
        const int dst[4] = {0, 1, g.ny + 2, g.ny + 3};
        for (int e = 0; e < 4; e++) {
            const int64_t s = g.at(i, src[e], k), d = g.at(i, dst[e], k);
            u[d] = u[s]; v[d] = v[s]; w[d] = w[s];
            p[d] = (i == iout && G[s] == 0) ? pin_out : p[s];
            if (q && i == iout) q[g.qat(dst[e], k)] = q[g.qat(src[e], k)];
        }
    } else if (t < nX + nY + nZ) {
        const int64_t tt = t - nX - nY;
        const int j = (int)(tt / g.nx) + 2, k = (int)(tt % g.nx) + 2;
        const int64_t cout = g.at(iout, j, k);
        const bool out_fluid = G[cout] == 0;
        if (g.pz) {
            const int src[4] = {g.nz, g.nz + 1, 2, 3};
            const int dst[4] = {0, 1, g.nz + 2, g.nz + 3};
            for (int e = 0; e < 4; e++) {
                const int64_t s = g.at(src[e], j, k), d = g.at(dst[e], j, k);
                u[d] = u[s]; v[d] = v[s]; w[d] = w[s];
                p[d] = (src[e] == iout && out_fluid) ? pin_out : p[s];
            }
        }
        const int64_t c1 = g.at(1, j, k), chi = g.at(g.nz + 2, j, k);
        if (G[c1] == 0) p[c1] = pin_in;
        if (G[chi] == 0) p[chi] = pin_hi;
        if (out_fluid) p[cout] = pin_out;
    }
}

// Outlet-plane buffer of a FP64 state at plane position (j, k) (any j < NY,
// k < NX): pin_out - c2*dt*div at interior fluid cells of z = nz+1, p at the
// (periodic image of the) cell elsewhere.
__device__ __forceinline__ void outlet_cell(const Grid &g, int j, int k, const uint8_t *__restrict__ pc,
                                            const double *__restrict__ p, const double *__restrict__ u,
                                            const double *__restrict__ v, const double *__restrict__ w,
                                            double pin_out, double *__restrict__ q) {
    const int i = g.nz + 1;
    int jj = j, kk = k;
    if (g.px) { if (kk < 2) kk += g.nx; else if (kk >= g.nx + 2) kk -= g.nx; }
    if (g.py) { if (jj < 2) jj += g.ny; else if (jj >= g.ny + 2) jj -= g.ny; }
    const bool in = jj >= 2 && jj < g.ny + 2 && kk >= 2 && kk < g.nx + 2;
    const int64_t c = g.at(i, jj, kk);
    double val;
    if (in && pc[c] != kSolid) {
        const double div = ((((w[c + g.sy] - w[c]) + v[c + g.sx]) - v[c]) + u[c + 1]) - u[c];
        val = pin_out - div * (kC2 * kDt);
    } else {
        val = p[c];
    }
    q[g.qat(j, k)] = val;
}

}  // namespace pm
