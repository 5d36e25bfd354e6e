// pm_compact.h -- per-cell work of the FP32+DC Jacobi on the compact state,
// shared by the single-solve kernels (pm_sweep.hip) and the batched kernels
// (pm_batch.hip).
#pragma once

#include "pm_kernels.h"
#include "pm_voxel.h"

#pragma clang fp contract(off)

namespace pm {

// ---------------------------------------------------------------------------
// FP32+DC Jacobi on the compact state (FluidLists::compact_jac): one entry per
// fluid cell in list order and the zero entry, links as for the SOR, but a
// link across a periodic face goes straight to the image cell.  The record
// pass and the passes before it write their cells' periodic halo copies
// themselves, so in the dense layout a halo cell always holds its image's
// values, except the pressure of the z-halo planes 1 and nz + 2, which the
// fill pins to 0: plane 1 copies the outlet plane, whose correction pressure
// is 0 too, and no pass reads the pressure of a +z neighbour.  Reading the
// image therefore gives the same values, and the compact state needs no
// halo entries and no ghost fill.  The lagged pass recomputes the pressure at
// its -x, -y, -z neighbours exactly as lag_voxel4_rec does (the wrapped
// neighbour is the same periodic image); their +1 neighbours are second hops
// through the neighbours' links.  Flags (36 bits over the links' top bits):
// the packed codes of the cell and of its -x, -y, -z neighbours (8 bits
// each), the outlet flag (bit 32) and the first-plane flag (bit 33).
// ---------------------------------------------------------------------------
struct JLinks {
    int32_t xm, xp, ym, yp, zm, zp;
    uint64_t f;
};

// Where link word w (0 -x, 1 +x, 2 -y, 3 +y, 4 -z, 5 +z) of entry e lies,
// in 32-bit words: three arrays of (-, +) pairs, or (AOS) one record of six
// words per entry ordered +x, +y, +z, -x, -y, -z, so that the +1 links a
// second hop needs from one neighbour are adjacent.
template <bool AOS>
__device__ __forceinline__ uint32_t link_word(int32_t e, int w, int nE) {
    if (AOS) return (uint32_t)e * 6u + (uint32_t)((w & 1) ? (w >> 1) : 3 + (w >> 1));
    return ((uint32_t)(w >> 1) * (uint32_t)nE + (uint32_t)e) * 2u + (uint32_t)(w & 1);
}

template <bool AOS>
__device__ __forceinline__ int32_t buf_load_lw(__amdgpu_buffer_rsrc_t rs, int32_t e, int w, int nE, bool need) {
    const uint32_t off = need ? link_word<AOS>(e, w, nE) * 4u : 0xFFFFFFF0u;
    return (int32_t)((uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0) & kLinkMask);
}

template <bool AOS>
__device__ __forceinline__ JLinks jac_links(const uint2 *__restrict__ L, int nE, int t) {
    uint32_t w[6];
    if (AOS) {
        typedef unsigned int v3u __attribute__((ext_vector_type(3)));
        const uint32_t *W = reinterpret_cast<const uint32_t *>(L) + (size_t)t * 6;
        const v3u p = __builtin_nontemporal_load(reinterpret_cast<const v3u *>(W));
        const v3u m = __builtin_nontemporal_load(reinterpret_cast<const v3u *>(W + 3));
        w[0] = m.x; w[1] = p.x; w[2] = m.y; w[3] = p.y; w[4] = m.z; w[5] = p.z;
    } else {
        typedef unsigned int v2u __attribute__((ext_vector_type(2)));
        const v2u lx = __builtin_nontemporal_load(reinterpret_cast<const v2u *>(&L[t]));
        const v2u ly = __builtin_nontemporal_load(reinterpret_cast<const v2u *>(&L[nE + t]));
        const v2u lz = __builtin_nontemporal_load(reinterpret_cast<const v2u *>(&L[2 * nE + t]));
        w[0] = lx.x; w[1] = lx.y; w[2] = ly.x; w[3] = ly.y; w[4] = lz.x; w[5] = lz.y;
    }
    JLinks r;
    r.xm = (int32_t)(w[0] & kLinkMask); r.xp = (int32_t)(w[1] & kLinkMask);
    r.ym = (int32_t)(w[2] & kLinkMask); r.yp = (int32_t)(w[3] & kLinkMask);
    r.zm = (int32_t)(w[4] & kLinkMask); r.zp = (int32_t)(w[5] & kLinkMask);
    r.f = 0;
    for (int q = 0; q < 6; q++) r.f |= (uint64_t)(w[q] >> kLinkBits) << (6 * q);
    return r;
}

// The cell's stencil: the +-1 neighbours (each only where its side is open,
// with ALL1) and the +-2 neighbours as second hops, only where the case reads
// them.
template <bool ALL1, bool AOS>
__device__ __forceinline__ Star4 star_compact(const JLinks &l, const Cases &cc, const float4 *__restrict__ S,
                                              __amdgpu_buffer_rsrc_t rsL, __amdgpu_buffer_rsrc_t rsS, int nE) {
    Star4 s;
    if (ALL1) {
        // a slot is read where its side is open or where it holds the +-2 link
        s.XM = buf_load4(rsS, l.xm, !cc.bx || cc.cx == 4); s.XP = buf_load4(rsS, l.xp, !fwd_solid(cc.cx) || cc.cx == 5);
        s.YM = buf_load4(rsS, l.ym, !cc.by || cc.cy == 4); s.YP = buf_load4(rsS, l.yp, !fwd_solid(cc.cy) || cc.cy == 5);
        s.ZM = buf_load4(rsS, l.zm, !cc.bz || cc.cz == 4); s.ZP = buf_load4(rsS, l.zp, !fwd_solid(cc.cz) || cc.cz == 5);
    } else {
        s.XM = S[l.xm]; s.XP = S[l.xp];
        s.YM = S[l.ym]; s.YP = S[l.yp];
        s.ZM = S[l.zm]; s.ZP = S[l.zp];
    }
    // the +-2 neighbour comes through the slot of the solid side (case 4:
    // the -1 slot holds the +2 link, case 5: the +1 slot the -2 link)
    s.XP2 = s.XM; s.XM2 = s.XP;
    s.YP2 = s.YM; s.YM2 = s.YP;
    s.ZP2 = s.ZM; s.ZM2 = s.ZP;
    return s;
}

// The +1 links of the recompute targets a lagged pass hops through, with one
// load per target in the record layout: (+y, +z) of the -x target, (+x, +z)
// of the -y target, (+x, +y) of the -z target; each pair only where one of
// its two values is used.
template <bool AOS>
__device__ __forceinline__ void target_links(__amdgpu_buffer_rsrc_t rsL, const JLinks &l, int nE, const bool need[6],
                                             int32_t out[6]) {
    if (AOS) {
        typedef unsigned int v2u __attribute__((ext_vector_type(2)));
        typedef unsigned int v3u __attribute__((ext_vector_type(3)));
        const bool nx = need[0] || need[1], ny = need[2] || need[3], nz = need[4] || need[5];
        const v2u a = __builtin_bit_cast(v2u, __builtin_amdgcn_raw_buffer_load_b64(
                                                   rsL, nx ? ((uint32_t)l.xm * 6u + 1u) * 4u : 0xFFFFFFF0u, 0, 0));
        const v3u b = __builtin_bit_cast(v3u, __builtin_amdgcn_raw_buffer_load_b96(
                                                   rsL, ny ? (uint32_t)l.ym * 24u : 0xFFFFFFF0u, 0, 0));
        const v2u c = __builtin_bit_cast(v2u, __builtin_amdgcn_raw_buffer_load_b64(
                                                   rsL, nz ? (uint32_t)l.zm * 24u : 0xFFFFFFF0u, 0, 0));
        out[0] = (int32_t)(a.x & kLinkMask); out[1] = (int32_t)(a.y & kLinkMask);
        out[2] = (int32_t)(b.x & kLinkMask); out[3] = (int32_t)(b.z & kLinkMask);
        out[4] = (int32_t)(c.x & kLinkMask); out[5] = (int32_t)(c.y & kLinkMask);
    } else {
        out[0] = buf_load_lw<false>(rsL, l.xm, 3, nE, need[0]); out[1] = buf_load_lw<false>(rsL, l.xm, 5, nE, need[1]);
        out[2] = buf_load_lw<false>(rsL, l.ym, 1, nE, need[2]); out[3] = buf_load_lw<false>(rsL, l.ym, 5, nE, need[3]);
        out[4] = buf_load_lw<false>(rsL, l.zm, 1, nE, need[4]); out[5] = buf_load_lw<false>(rsL, l.zm, 3, nE, need[5]);
    }
}

// lag_voxel4_rec on the compact state.  ALL: every load, else each +-1
// neighbour and each recompute-target value only where it is used.  AOS: the
// Packed links of the compact Jacobi (FluidLists::packed with compact_jac),
// 24 B per entry: twelve links -- slots 0-5 the +-1 neighbours as in the link
// arrays (after the +-2 substitution), slots 6-11 the recompute targets'
// links target_links reads (the -x neighbour's +y and +z links, the -y
// neighbour's +x and +z, the -z neighbour's +x and +y) -- each in 12 bits: 2
// bits naming one of three bases of the entry's wave of 64 and the slot (3:
// the zero entry) and a 10-bit offset, link = base + lane + offset.  Word w
// holds slots 2w, 2w + 1 in bits 0-23 and flag bits 6w .. 6w + 5 in bits
// 24-29; bit 31 of word 5 marks an entry whose links fit no base (read from
// the link arrays).  36 bases per wave, [base][slot].
constexpr int kJPackSpan = 1023;
constexpr uint32_t kJPackZero = 0xFFFu;

// One float of a buffer of floats (element e), only for lanes with need set.
__device__ __forceinline__ float buf_loadf(__amdgpu_buffer_rsrc_t rs, int32_t e, bool need) {
    const uint32_t off = need ? (uint32_t)e * 4u : 0xFFFFFFF0u;
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
}

// The bases are read through the constant address space: nothing writes
// them during a pass, and a wave-uniform address there is read with scalar
// loads even where the pointer comes from memory the kernel also writes
// through (the batch's sample table), which would otherwise make every lane
// load all 36.
typedef const int32_t __attribute__((address_space(4))) kconst_i32;

__device__ __forceinline__ void jac_packed_links(int t, int lane, const kconst_i32 *bw,
                                                 const uint32_t *__restrict__ P, const uint2 *__restrict__ L, int nE,
                                                 __amdgpu_buffer_rsrc_t rsL, JLinks &l, int32_t (&dl)[6]) {
    typedef unsigned int v2u __attribute__((ext_vector_type(2)));
    const uint32_t *rec = P + (size_t)t * 6;
    const v2u w01 = __builtin_nontemporal_load(reinterpret_cast<const v2u *>(rec));
    const v2u w23 = __builtin_nontemporal_load(reinterpret_cast<const v2u *>(rec + 2));
    const v2u w45 = __builtin_nontemporal_load(reinterpret_cast<const v2u *>(rec + 4));
    const uint32_t W[6] = {w01.x, w01.y, w23.x, w23.y, w45.x, w45.y};
    if (W[5] >> 31) {
        l = jac_links<false>(L, nE, t);
        const bool all[6] = {true, true, true, true, true, true};
        target_links<false>(rsL, l, nE, all, dl);
        return;
    }
    const int32_t zero = nE - 1;
    int32_t lk[12];
#pragma unroll
    for (int sl = 0; sl < 12; sl++) {
        const uint32_t c = (W[sl >> 1] >> (12 * (sl & 1))) & 0xFFFu;
        const uint32_t sel = c >> 10;
        // the base by masks, not by choosing between array elements (which
        // the compiler turns into a per-lane indexed read)
        const int32_t m1 = -(int32_t)(sel == 1), m2 = -(int32_t)(sel == 2), m0 = ~(m1 | m2);
        const int32_t base = (bw[sl] & m0) | (bw[12 + sl] & m1) | (bw[24 + sl] & m2);
        lk[sl] = c == kJPackZero ? zero : base + lane + (int32_t)(c & 0x3FFu);
    }
    l.xm = lk[0]; l.xp = lk[1]; l.ym = lk[2]; l.yp = lk[3]; l.zm = lk[4]; l.zp = lk[5];
    l.f = 0;
#pragma unroll
    for (int w = 0; w < 6; w++) l.f |= (uint64_t)((W[w] >> 24) & 63u) << (6 * w);
#pragma unroll
    for (int q = 0; q < 6; q++) dl[q] = lk[6 + q];
}

// The 36 bases of the wave that holds list position r (B and r the same on
// every lane of the wave).
__device__ __forceinline__ const kconst_i32 *jac_wave_bases(const int32_t *B, int r) {
    const uint64_t a = (uint64_t)(B + (size_t)(r >> 6) * 36);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return (const kconst_i32 *)((uint64_t)hi << 32 | lo);
}

// links as one record per entry.
template <bool ALL, bool AOS, bool PACKED = false>
__device__ __forceinline__ void lag_voxel4_compact(int t, const uint2 *__restrict__ L, int nE,
                                                   __amdgpu_buffer_rsrc_t rsL, __amdgpu_buffer_rsrc_t rsS,
                                                   const float4 *__restrict__ S0, float4 *__restrict__ S1, Res4 R,
                                                   float cu, float cp, float pin_out,
                                                   const uint32_t *__restrict__ P = nullptr,
                                                   const kconst_i32 *bw = nullptr, int lane = 0) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    JLinks l;
    int32_t dlp[6];
    if (PACKED) jac_packed_links(t, lane, bw, P, L, nE, rsL, l, dlp);
    else l = jac_links<AOS>(L, nE, t);
    const uint8_t pcv = (uint8_t)(l.f & 0xff), pcx = (uint8_t)((l.f >> 8) & 0xff),
                  pcy = (uint8_t)((l.f >> 16) & 0xff), pcz = (uint8_t)((l.f >> 24) & 0xff);
    const bool out = (l.f >> 32) & 1u, first = (l.f >> 33) & 1u;
    const int32_t zero = nE - 1;
    const Cases cc = decode(pcv);
    // the -x, -y, -z neighbours' continuity residuals (their recomputed
    // pressures); without ALL only where used: an open side, off the outlet
    // plane (x, y) and off the first plane (z)
    float rpx, rpy, rpz;
    if (ALL) {
        rpx = R.pc[l.xm]; rpy = R.pc[l.ym]; rpz = R.pc[l.zm];
    } else {
        const __amdgpu_buffer_rsrc_t rsP =
            __builtin_amdgcn_make_buffer_rsrc((void *)R.pc, (short)0, (int)((uint32_t)nE * 4u), 0x00020000);
        rpx = buf_loadf(rsP, l.xm, !cc.bx && !out);
        rpy = buf_loadf(rsP, l.ym, !cc.by && !out);
        rpz = buf_loadf(rsP, l.zm, !cc.bz && !first);
    }
    const float4 C = S0[t];
    const Star4 s = star_compact<ALL, AOS>(l, cc, S0, rsL, rsS, nE);
    // the recompute targets' +1 neighbours: v of (x-1, y+1), w of (x-1, z+1),
    // u of (x+1, y-1), w of (y-1, z+1), u of (x+1, z-1), v of (y+1, z-1)
    bool need[6];
    if (ALL) {
        for (int q = 0; q < 6; q++) need[q] = true;
    } else {
        const Cases tx = decode(pcx), ty = decode(pcy), tz = decode(pcz);
        need[0] = !cc.bx && !fwd_solid(tx.cy); need[1] = !cc.bx && !fwd_solid(tx.cz);
        need[2] = !cc.by && !fwd_solid(ty.cx); need[3] = !cc.by && !fwd_solid(ty.cz);
        need[4] = !cc.bz && !fwd_solid(tz.cx); need[5] = !cc.bz && !fwd_solid(tz.cy);
    }
    int32_t dl[6];
    if (PACKED) {
        for (int q = 0; q < 6; q++) dl[q] = dlp[q];
    } else {
        target_links<AOS>(rsL, l, nE, need, dl);
    }
    const float vyx = buf_load1(rsS, dl[0], 2, need[0]), wzx = buf_load1(rsS, dl[1], 3, need[1]);
    const float uxy = buf_load1(rsS, dl[2], 1, need[2]), wzy = buf_load1(rsS, dl[3], 3, need[3]);
    const float uxz = buf_load1(rsS, dl[4], 1, need[4]), vyz = buf_load1(rsS, dl[5], 2, need[5]);
    const v4f r = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(&R.m[t]));
    const float4 Rm = make_float4(r.x, r.y, r.z, r.w);
    const float rpc = r.w;  // the cell's continuity residual
    // the -x / -y neighbour is the periodic image where the dense pass wraps,
    // and the image's +x / +y neighbour is this cell
    const float4 NX = s.XM, NY = s.YM;
    const float uxx = C.y, vyy = C.z;
    const Cases cs = cc;
    const float Pc = out ? pin_out : C.x - (div_from4(pcv, C, s.XP.y, s.YP.z, s.ZP.w) + rpc) * cp;
    const float c2dt = (float)(kC2 * kDt);
    const float dx0 = div_from4(pcx, NX, uxx, vyx, wzx);
    const float dy0 = div_from4(pcy, NY, uxy, vyy, wzy);
    const float dz0 = div_from4(pcz, s.ZM, uxz, vyz, C.w);
    float px = out ? pin_out - dx0 * c2dt : NX.x - (dx0 + rpx) * cp;
    float py = out ? pin_out - dy0 * c2dt : NY.x - (dy0 + rpy) * cp;
    const float pz = first ? s.ZM.x : s.ZM.x - (dz0 + rpz) * cp;
    // a wall halo keeps its stored value, 0 (as does its outlet buffer
    // value); a solid neighbour's value is never used
    if (l.xm == zero) px = s.XM.x;
    if (l.ym == zero) py = s.YM.x;
    const float dpx = cs.bx ? 0.f : Pc - px;
    const float dpy = cs.by ? 0.f : Pc - py;
    const float dpz = cs.bz ? 0.f : Pc - pz;
    float su, sv, sw;
    visc_from4(cs, s, C, su, sv, sw);
    const float4 o = vel_from4(cs, C, su, sv, sw, dpx, dpy, dpz, Rm, Pc, cu);
    const v4f ov = {o.x, o.y, o.z, o.w};
    __builtin_nontemporal_store(ov, reinterpret_cast<v4f *>(&S1[t]));
}

// first_voxel4_zero on the compact state: the first pass of an interval from
// a correction state that is zero everywhere, from the cell's code and
// momentum residual alone.
template <bool AOS>
__device__ __forceinline__ void first_zero_compact(int t, const uint2 *__restrict__ L, int nE,
                                                   float4 *__restrict__ S1, Res4 R, float cu) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    const uint32_t *W = reinterpret_cast<const uint32_t *>(L);
    const uint32_t wxm = W[link_word<AOS>(t, 0, nE)], wxp = W[link_word<AOS>(t, 1, nE)];
    const uint32_t code = (wxm >> kLinkBits) | (((wxp >> kLinkBits) & 3u) << 6);
    const Cases cs = decode((uint8_t)code);
    const v4f r = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(&R.m[t]));
    const float z = 0.f;
    v4f o;
    o.x = z;
    o.y = cs.bx ? z : z + (z + r.x) * cu;
    o.z = cs.by ? z : z + (z + r.y) * cu;
    o.w = cs.bz ? z : z + (z + r.z) * cu;
    __builtin_nontemporal_store(o, reinterpret_cast<v4f *>(&S1[t]));
}

// pres_voxel4 in place (the settle pass) on the compact state, with the pin
// the dense ghost fill then applies: an outlet cell's correction pressure is
// 0 and its outlet buffer value pin_out - c2 dt div.  The continuity residual
// of the cell is the fourth lane of its momentum residual.
template <bool AOS>
__device__ __forceinline__ void settle_compact(int t, const uint2 *__restrict__ L, int nE, float4 *S,
                                               float *__restrict__ q1, Res4 R, float cp) {
    const JLinks l = jac_links<AOS>(L, nE, t);
    const Cases cs = decode((uint8_t)(l.f & 0xff));
    const bool out = (l.f >> 32) & 1u;
    const float4 N = S[t];
    const float wz = fwd_solid(cs.cz) ? 0.f : S[l.zp].w;
    const float vy = fwd_solid(cs.cy) ? 0.f : S[l.yp].z;
    const float ux = fwd_solid(cs.cx) ? 0.f : S[l.xp].y;
    const float div0 = ((((wz - N.w) + vy) - N.z) + ux) - N.y;
    const float div = div0 + R.m[t].w;
    float p = N.x - div * cp;
    if (out) {
        q1[t] = 0.f - div0 * (float)(kC2 * kDt);
        p = 0.f;
    }
    S[t].x = p;
}

// X += d, d = 0 at the fluid cell of entry t (natural offset c).
__device__ __forceinline__ void fold_compact(int t, int64_t c, double *Xp, double *Xu, double *Xv, double *Xw,
                                             float4 *d) {
    const float4 v = d[t];
    Xp[c] = Xp[c] + (double)v.x;
    Xu[c] = Xu[c] + (double)v.y;
    Xv[c] = Xv[c] + (double)v.z;
    Xw[c] = Xw[c] + (double)v.w;
    d[t] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// The buffer resources of a compact pass: the links (three arrays or one
// record per entry, 24 B per entry) and the state (16 B per entry).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t links_rsrc(const uint2 *L, int nE) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)L, (short)0, (int)((uint32_t)nE * 24u), 0x00020000);
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t state_rsrc(const float4 *S, int nE) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)S, (short)0, (int)((uint32_t)nE * 16u), 0x00020000);
}

}  // namespace pm
