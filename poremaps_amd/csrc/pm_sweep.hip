// pm_sweep.hip -- fast-mode kernels of one geometry.
//
// Porous media are mostly solid (porosity 0.15-0.45 in the target cases) and
// a solid cell's state never changes, so the sweeps launch one thread per
// fluid voxel: a sorted list of the fluid cells' element offsets (z, y, x
// order, so consecutive lanes take consecutive fluid voxels of a row) built
// once per geometry.  The per-cell work lives in pm_voxel.h.
#include "pm_compact.h"
#include "pm_kernels.h"
#include "pm_voxel.h"

#include <algorithm>

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

template <typename T, bool SEL>
__global__ __launch_bounds__(kListBlock) void k_jac_vel(Grid g, const int32_t *__restrict__ list, int n,
                                                        const uint8_t *__restrict__ pc,
                                                        const T *__restrict__ p0, const T *__restrict__ u0,
                                                        const T *__restrict__ v0, const T *__restrict__ w0,
                                                        const T *__restrict__ q0, T *__restrict__ u1,
                                                        T *__restrict__ v1, T *__restrict__ w1, T cu) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t c = list[t];
    vel_voxel<T, SEL>(g, c, pc[c], p0, u0, v0, w0, q0, u1, v1, w1, cu);
}

template <typename T>
__global__ __launch_bounds__(kListBlock) void k_jac_pres(Grid g, const int32_t *__restrict__ list, int n,
                                                         const uint8_t *__restrict__ pc,
                                                         const T *__restrict__ p0, const T *__restrict__ u1,
                                                         const T *__restrict__ v1, const T *__restrict__ w1,
                                                         T *__restrict__ p1, T *__restrict__ q1, T cp, T pin_out) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t c = list[t];
    pres_voxel<T>(g, c, pc[c], p0, u1, v1, w1, p1, q1, cp, pin_out);
}

// First sweep from a settled state (p^n given): velocities into set 1 and
// p^n carried over, so that set 1 holds (u^{n+1}, p^n) for the lagged passes.
template <typename T, bool SEL>
__global__ __launch_bounds__(kListBlock) void k_jac_first(Grid g, const int32_t *__restrict__ list, int n,
                                                          const uint8_t *__restrict__ pc,
                                                          const T *__restrict__ p0, const T *__restrict__ u0,
                                                          const T *__restrict__ v0, const T *__restrict__ w0,
                                                          const T *__restrict__ q0, T *__restrict__ p1,
                                                          T *__restrict__ u1, T *__restrict__ v1,
                                                          T *__restrict__ w1, T cu) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t c = list[t];
    vel_voxel<T, SEL>(g, c, pc[c], p0, u0, v0, w0, q0, u1, v1, w1, cu);
    p1[c] = p0[c];
}

template <typename T, bool SEL>
__global__ __launch_bounds__(kListBlock) void k_jac_lag(Grid g, const int32_t *__restrict__ list, int n,
                                                        const uint8_t *__restrict__ pc,
                                                        const T *__restrict__ p0, const T *__restrict__ u0,
                                                        const T *__restrict__ v0, const T *__restrict__ w0,
                                                        const T *__restrict__ q0, T *__restrict__ p1,
                                                        T *__restrict__ u1, T *__restrict__ v1,
                                                        T *__restrict__ w1, T cu, T cp, T pin_out) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t c = list[t];
    lag_voxel<T, SEL>(g, c, pc[c], pc, p0, u0, v0, w0, q0, p1, u1, v1, w1, cu, cp, pin_out);
}

// Interleaved FP32 correction state (defect-correction Jacobi): stage 0 =
// first pass (velocities, p carried), 1 = lagged pass, 2 = pair velocity
// pass, 3 = pair pressure pass / settle, 4 = lagged pass with every load
// issued up front.
template <int STAGE>
__global__ __launch_bounds__(kListBlock) void k_jac4(Grid g, const int32_t *__restrict__ list, int n,
                                                     const uint8_t *__restrict__ pc, const float4 *__restrict__ S0,
                                                     const float *__restrict__ q0, float4 *S1, float *q1,
                                                     Res4 R, float cu, float cp) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t c = list[t];
    if (STAGE == 0 || STAGE == 2) vel_voxel4(g, c, t, pc[c], S0, q0, S1, R, cu);
    else if (STAGE == 1) lag_voxel4(g, c, t, pc[c], pc, S0, q0, S1, R, cu, cp, 0.f);
    else if (STAGE == 4) lag_voxel4_eager(g, c, t, pc, S0, q0, S1, R, cu, cp, 0.f);
    else pres_voxel4(g, c, pc[c], S0, S1, q1, R, cp, 0.f);
}

// 3-colour SOR on the interleaved state: one colour's cells, positions t0.. of
// the fluid list (the colour lists lie back to back in it).
__global__ __launch_bounds__(kListBlock) void k_colour4(Grid g, const int32_t *__restrict__ list, int n, int t0,
                                                        const uint8_t *__restrict__ pc, float4 *S, float *q,
                                                        Res4 R, float cu, float cp) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    colour_voxel4(g, list[t0 + t], t0 + t, pc, S, q, R, cu, cp, 0.f);
}

// The same on the colour-split state, from the voxels' records
// (colour_voxel4_split).
template <bool SKIP2>
__global__ __launch_bounds__(kListBlock) void k_colour4_split(Grid g, const uint2 *__restrict__ rec, int n, int t0,
                                                              float4 *S, float *q, Res4 R, float cu, float cp) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    colour_voxel4_split<SKIP2>(g, t0 + t, rec, S, q, R, cu, cp, 0.f);
}

// Up to three colours' list ranges of one stage of the z-chunked SOR in one
// launch: blocks [b[q], b[q + 1]) take colour q's positions t0[q] ..
// t0[q] + n[q] - 1.  The ranges lie on planes far enough apart that no cell
// of one is read or written by another (launch_colour4).
struct ColourStage {
    int t0[3], n[3], b[3];
};

template <bool SKIP2>
__global__ __launch_bounds__(kListBlock) void k_colour4_stage(Grid g, const uint2 *__restrict__ rec, ColourStage st,
                                                              float4 *S, float *q, Res4 R, float cu, float cp) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int col = b >= st.b[2] ? 2 : (b >= st.b[1] ? 1 : 0);
    const int t = (b - st.b[col]) * kListBlock + threadIdx.x;
    if (t >= st.n[col]) return;
    colour_voxel4_split<SKIP2>(g, st.t0[col] + t, rec, S, q, R, cu, cp, 0.f);
}

// ---------------------------------------------------------------------------
// 3-colour SOR on the compact correction state (FluidLists::compact).  The
// colour-split rows keep every cell of the domain, so at porosity 0.2 a wave's
// 64 fluid cells and their neighbours lie on about 40 lines of 128 B per load
// and a colour pass reads nearly every line of the state (1.6 GB of a 512^3
// field, three times per sweep).  Here the state holds the fluid cells only,
// in list order, so the cells of one colour pass and their neighbours (the
// other two colours' cells of the same rows, in the same column order) lie on
// about a third of those lines; each entry carries the entries of its six
// +-1 neighbours, and a +-2 neighbour is the +-1 neighbour's own link on the
// same axis (read, like the value behind it, only by the lanes whose case
// uses it).  Periodic halo cells have entries of their own that the ghost
// fill refreshes after each sweep, as the dense halo copies are, so every
// value read is the one colour_voxel4_split reads, and the same expressions
// in the same order give the same bits.
// ---------------------------------------------------------------------------

// Link word w (0 -x, 1 +x, 2 -y, 3 +y, 4 -z, 5 +z) of entry e, only for lanes
// with need set (buffer load over the three link arrays of nE entries; an
// out-of-range offset makes no request).
__device__ __forceinline__ int32_t buf_load_link(__amdgpu_buffer_rsrc_t rs, int32_t e, int w, int nE, bool need) {
    const uint32_t off = need ? ((uint32_t)(w >> 1) * (uint32_t)nE + (uint32_t)e) * 8u + 4u * (uint32_t)(w & 1)
                              : 0xFFFFFFF0u;
    return (int32_t)((uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0) & kLinkMask);
}

// The colour pass of entry t from its six links (0 -x, 1 +x, 2 -y, 3 +y, 4 -z,
// 5 +z) and its record flags (packed code, block, outlet flag: bits 0-10).
__device__ __forceinline__ void colour_body4(int t, const int32_t (&lk)[6], uint32_t f, float4 *S, float *q, Res4 R,
                                             float cu, float cp, float pin_out) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    const int32_t xm = lk[0], xp = lk[1], ym = lk[2], yp = lk[3], zm = lk[4], zp = lk[5];
    const bool out = (f >> 10) & 1u;
    const Cases c2 = decode((uint8_t)(f & 0xffu));
    const float4 C = S[t];
    Star4 s;
    s.XM = S[xm]; s.XP = S[xp];
    s.YM = S[ym]; s.YP = S[yp];
    s.ZM = S[zm]; s.ZP = S[zp];
    // the +-2 neighbour comes through the slot of the solid side (case 4:
    // the -1 slot holds the +2 link, case 5: the +1 slot the -2 link)
    s.XP2 = s.XM; s.XM2 = s.XP;
    s.YP2 = s.YM; s.YM2 = s.YP;
    s.ZP2 = s.ZM; s.ZM2 = s.ZP;
    const v4f r = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(&R.m[t]));
    const float4 Rm = make_float4(r.x, r.y, r.z, 0.f);
    const float rp = r.w;
    // the outlet buffer is indexed by entry: the -x / -y neighbours' values
    // are those of the neighbours' entries (the zero entry's is 0, as a wall
    // halo's outlet value is)
    float pxm = s.XM.x, pym = s.YM.x;
    if (out) { pxm = q[xm]; pym = q[ym]; }
    const Cases cs = c2;
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
    S[t] = N;
    if (out) q[t] = pin_out - div0 * (float)(kC2 * kDt);
}

// The links and flags of entry t from the three link arrays.
__device__ __forceinline__ uint32_t full_links(int t, const uint2 *__restrict__ L, int nE, int32_t (&lk)[6]) {
    typedef unsigned int v2u __attribute__((ext_vector_type(2)));
    const v2u lx = __builtin_nontemporal_load(reinterpret_cast<const v2u *>(&L[t]));
    const v2u ly = __builtin_nontemporal_load(reinterpret_cast<const v2u *>(&L[nE + t]));
    const v2u lz = __builtin_nontemporal_load(reinterpret_cast<const v2u *>(&L[2 * nE + t]));
    lk[0] = (int32_t)(lx.x & kLinkMask); lk[1] = (int32_t)(lx.y & kLinkMask);
    lk[2] = (int32_t)ly.x; lk[3] = (int32_t)ly.y; lk[4] = (int32_t)lz.x; lk[5] = (int32_t)lz.y;
    return (lx.x >> kLinkBits) | ((lx.y >> kLinkBits) << 6);
}

__device__ __forceinline__ void colour_voxel4_compact(int t, const uint2 *__restrict__ L, int nE, float4 *S, float *q,
                                                      Res4 R, float cu, float cp, float pin_out) {
    int32_t lk[6];
    const uint32_t f = full_links(t, L, nE, lk);
    colour_body4(t, lk, f, S, q, R, cu, cp, pin_out);
}

// Packed links.  The interior entries of one colour go by waves of 64 (list
// positions t0 + 64 k ...), as the colour pass runs them.  Within a wave the
// links of one slot are nearly consecutive: the neighbours of consecutive
// cells of one colour are consecutive cells of the neighbouring colour, so
// link - lane is the same up to the fluid cells the rows skip, and takes a
// few values apart where the wave's cells reach across a column edge (or a
// +-2 link in a solid side's slot reaches across the other one).  Each slot
// therefore has up to three bases per wave, and a link is
// base[sel][slot] + lane + d with d < 4096, coded in 14 bits (sel 3: the
// zero entry).  Record of an entry, three words: codes of slots 0-1, 2-3,
// 4-5 in bits 0-27 of each, the 11 flag bits in bits 28-31 of the first two
// and 28-30 of the third, bit 31 of the third set when a link fits no base
// (then all six are read from the link arrays).
constexpr int kPackSpan = 4095;
constexpr uint32_t kPackZero = 0x3FFFu;

// One thread per entry of colour list positions t0 .. t0 + n - 1, whole waves
// (lanes past n take part in the wave's choices without an entry).  The
// bases of a slot: three times the value of link - lane that starts the
// window [base, base + span] holding most of the wave's links not yet placed
// (ties to the lowest lane); span <= kPackSpan (smaller in tests, to send
// entries to the full links).
__global__ __launch_bounds__(256) void k_pack_links(const uint2 *__restrict__ L, int nE, int t0, int n, int wb,
                                                    int span, uint32_t *__restrict__ P, int32_t *__restrict__ B,
                                                    int *__restrict__ nfar) {
    const int r = (int)blockIdx.x * 256 + (int)threadIdx.x;
    const int lane = (int)threadIdx.x & 63;
    const bool ok = r < n;
    const int t = t0 + r;
    const int32_t zero = nE - 1;
    int32_t lk[6] = {zero, zero, zero, zero, zero, zero};
    uint32_t f = 0;
    if (ok) f = full_links(t, L, nE, lk);
    int32_t *bw = B + (size_t)(wb + (r >> 6)) * 18;
    uint32_t code[6];
    bool far = false;
    for (int sl = 0; sl < 6; sl++) {
        const bool live = ok && lk[sl] != zero;
        const int32_t v = lk[sl] - lane;
        bool placed = !live;
        uint32_t c = kPackZero;
        for (int k = 0; k < 3; k++) {
            int cnt = 0;
            for (int j = 0; j < 64; j++) {
                const int32_t vj = __shfl(v, j, 64);
                const int pj = __shfl((int)placed, j, 64);
                cnt += (!pj && vj >= v && vj - v <= span) ? 1 : 0;
            }
            int key = placed ? -1 : ((cnt << 6) | (63 - lane));
            for (int o = 32; o > 0; o >>= 1) key = max(key, __shfl_xor(key, o, 64));
            const int32_t base = key < 0 ? 0 : __shfl(v, 63 - (key & 63), 64);
            if (key >= 0 && !placed && v >= base && v - base <= span) {
                placed = true;
                c = ((uint32_t)k << 12) | (uint32_t)(v - base);
            }
            if (lane == 0 && r < n) bw[k * 6 + sl] = base;
        }
        if (!placed) far = true;
        code[sl] = c;
    }
    if (!ok) return;
    uint32_t *rec = P + (size_t)t * 3;
    rec[0] = code[0] | (code[1] << 14) | ((f & 15u) << 28);
    rec[1] = code[2] | (code[3] << 14) | (((f >> 4) & 15u) << 28);
    rec[2] = code[4] | (code[5] << 14) | (((f >> 8) & 7u) << 28) | (far ? 0x80000000u : 0u);
    if (far) atomicAdd(nfar, 1);
}

// The links and flags of entry t from its packed record (b: the 18 bases of
// its wave; lane: its place in the wave).
__device__ __forceinline__ uint32_t packed_links(int t, int lane, const int32_t (&b)[18],
                                                 const uint32_t *__restrict__ P, const uint2 *__restrict__ L, int nE,
                                                 int32_t (&lk)[6]) {
    typedef unsigned int v3u __attribute__((ext_vector_type(3)));
    const v3u w = __builtin_nontemporal_load(reinterpret_cast<const v3u *>(P + (size_t)t * 3));
    if (w.z >> 31) return full_links(t, L, nE, lk);
    const uint32_t cw[3] = {w.x, w.y, w.z};
    const int32_t zero = nE - 1;
#pragma unroll
    for (int sl = 0; sl < 6; sl++) {
        const uint32_t c = (cw[sl >> 1] >> (14 * (sl & 1))) & 0x3FFFu;
        const uint32_t sel = c >> 12;
        // the base by masks, not by choosing between array elements (which
        // the compiler turns into a per-lane indexed read)
        const int32_t m1 = -(int32_t)(sel == 1), m2 = -(int32_t)(sel == 2), m0 = ~(m1 | m2);
        const int32_t base = (b[sl] & m0) | (b[6 + sl] & m1) | (b[12 + sl] & m2);
        lk[sl] = c == kPackZero ? zero : base + lane + (int32_t)(c & 0xFFFu);
    }
    return (w.x >> 28) | ((w.y >> 28) << 4) | (((w.z >> 28) & 7u) << 8);
}

// The 18 bases of the wave of colour position r (wave-uniform).
__device__ __forceinline__ void wave_bases(const int32_t *__restrict__ B, int wb, int r, int32_t (&b)[18]) {
    const int wave = __builtin_amdgcn_readfirstlane(r >> 6);
    const int32_t *bw = B + (size_t)(wb + wave) * 18;
#pragma unroll
    for (int i = 0; i < 18; i++) b[i] = bw[i];
}

// Load-time check of the packed records of colour positions t0 .. t0 + n - 1
// against the link arrays: counts the entries whose decoded links or flags
// differ (none, or the packed pass is not used).
__global__ __launch_bounds__(kListBlock) void k_check_packed(const uint32_t *__restrict__ P,
                                                             const int32_t *__restrict__ B, int wb,
                                                             const uint2 *__restrict__ L, int nE, int n, int t0,
                                                             int *__restrict__ bad) {
    const int r = (int)blockIdx.x * kListBlock + (int)threadIdx.x;
    if (r >= n) return;
    int32_t b[18];
    wave_bases(B, wb, r, b);
    int32_t a[6], c[6];
    const uint32_t fa = full_links(t0 + r, L, nE, a);
    const uint32_t fc = packed_links(t0 + r, (int)threadIdx.x & 63, b, P, L, nE, c);
    bool same = (fa & 0x7FFu) == (fc & 0x7FFu);
    for (int i = 0; i < 6; i++) same = same && a[i] == c[i];
    if (!same) atomicAdd(bad, 1);
}

// One colour's entries t0 .. t0 + n - 1 (its list positions).  nE: entries
// of the state (the buffers' extents).
__global__ __launch_bounds__(kListBlock) void k_colour4_compact(const uint2 *__restrict__ L, int nE, int n, int t0,
                                                                float4 *S, float *q, Res4 R, float cu, float cp) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    colour_voxel4_compact(t0 + t, L, nE, S, q, R, cu, cp, 0.f);
}

// The same from the packed links; wb: the colour's first wave in B.
__global__ __launch_bounds__(kListBlock) void k_colour4_packed(const uint32_t *__restrict__ P,
                                                               const int32_t *__restrict__ B, int wb,
                                                               const uint2 *__restrict__ L, int nE, int n, int t0,
                                                               float4 *S, float *q, Res4 R, float cu, float cp) {
    const int r = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (r >= n) return;
    int32_t b[18];
    wave_bases(B, wb, r, b);
    int32_t lk[6];
    const uint32_t f = packed_links(t0 + r, (int)threadIdx.x & 63, b, P, L, nE, lk);
    colour_body4(t0 + r, lk, f, S, q, R, cu, cp, 0.f);
}

// FP32+DC Jacobi on the compact state: pm_compact.h.

// OCC: waves per SIMD the registers must allow (0: the compiler's choice)
template <bool ALL, bool AOS, int OCC = 0>
__global__ __launch_bounds__(kListBlock, OCC > 0 ? OCC : 1) void k_jac4c_lag(const uint2 *__restrict__ L, int nE,
                                                                             int n, const float4 *__restrict__ S0,
                                                                             float4 *__restrict__ S1, Res4 R, float cu,
                                                                             float cp) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const __amdgpu_buffer_rsrc_t rsL =
        __builtin_amdgcn_make_buffer_rsrc((void *)L, (short)0, (int)((uint32_t)nE * 24u), 0x00020000);
    const __amdgpu_buffer_rsrc_t rsS =
        __builtin_amdgcn_make_buffer_rsrc((void *)S0, (short)0, (int)((uint32_t)nE * 16u), 0x00020000);
    lag_voxel4_compact<ALL, AOS>(t, L, nE, rsL, rsS, S0, S1, R, cu, cp, 0.f);
}

// The lagged pass from the packed records (jac_packed_links): the same cell
// work, with the recompute targets' links in the entry's own record.
template <bool ALL, int OCC = 0>
__global__ __launch_bounds__(kListBlock, OCC > 0 ? OCC : 1) void k_jac4c_lagp(
    const uint32_t *__restrict__ P, const int32_t *__restrict__ B, const uint2 *__restrict__ L, int nE, int n,
    const float4 *__restrict__ S0, float4 *__restrict__ S1, Res4 R, float cu, float cp) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const __amdgpu_buffer_rsrc_t rsL =
        __builtin_amdgcn_make_buffer_rsrc((void *)L, (short)0, (int)((uint32_t)nE * 24u), 0x00020000);
    const __amdgpu_buffer_rsrc_t rsS =
        __builtin_amdgcn_make_buffer_rsrc((void *)S0, (short)0, (int)((uint32_t)nE * 16u), 0x00020000);
    lag_voxel4_compact<ALL, false, true>(t, L, nE, rsL, rsS, S0, S1, R, cu, cp, 0.f, P, jac_wave_bases(B, t),
                                         (int)threadIdx.x & 63);
}

// The twelve links a packed Jacobi record holds, from the link arrays.
__device__ __forceinline__ uint64_t jac_twelve(int t, const uint2 *__restrict__ L, int nE, int32_t (&lk)[12]) {
    const JLinks l = jac_links<false>(L, nE, t);
    const uint32_t *W = reinterpret_cast<const uint32_t *>(L);
    lk[0] = l.xm; lk[1] = l.xp; lk[2] = l.ym; lk[3] = l.yp; lk[4] = l.zm; lk[5] = l.zp;
    lk[6] = (int32_t)(W[link_word<false>(l.xm, 3, nE)] & kLinkMask);
    lk[7] = (int32_t)(W[link_word<false>(l.xm, 5, nE)] & kLinkMask);
    lk[8] = (int32_t)(W[link_word<false>(l.ym, 1, nE)] & kLinkMask);
    lk[9] = (int32_t)(W[link_word<false>(l.ym, 5, nE)] & kLinkMask);
    lk[10] = (int32_t)(W[link_word<false>(l.zm, 1, nE)] & kLinkMask);
    lk[11] = (int32_t)(W[link_word<false>(l.zm, 3, nE)] & kLinkMask);
    return l.f;
}

// Packed Jacobi records of list positions 0 .. n - 1, one thread per entry,
// whole waves (lanes past n take part in the choices without an entry); the
// bases as k_pack_links chooses them, with windows [base, base + span],
// span <= kJPackSpan.
__global__ __launch_bounds__(256) void k_pack_jac(const uint2 *__restrict__ L, int nE, int n, int span,
                                                  uint32_t *__restrict__ P, int32_t *__restrict__ B,
                                                  int *__restrict__ nfar) {
    const int r = (int)blockIdx.x * 256 + (int)threadIdx.x;
    const int lane = (int)threadIdx.x & 63;
    const bool ok = r < n;
    const int32_t zero = nE - 1;
    int32_t lk[12];
    for (int q = 0; q < 12; q++) lk[q] = zero;
    uint64_t f = 0;
    if (ok) f = jac_twelve(r, L, nE, lk);
    int32_t *bw = B + (size_t)(r >> 6) * 36;
    uint32_t code[12];
    bool far = false;
    for (int sl = 0; sl < 12; sl++) {
        const bool live = ok && lk[sl] != zero;
        const int32_t v = lk[sl] - lane;
        bool placed = !live;
        uint32_t c = kJPackZero;
        for (int k = 0; k < 3; k++) {
            int cnt = 0;
            for (int j = 0; j < 64; j++) {
                const int32_t vj = __shfl(v, j, 64);
                const int pj = __shfl((int)placed, j, 64);
                cnt += (!pj && vj >= v && vj - v <= span) ? 1 : 0;
            }
            int key = placed ? -1 : ((cnt << 6) | (63 - lane));
            for (int o = 32; o > 0; o >>= 1) key = max(key, __shfl_xor(key, o, 64));
            const int32_t base = key < 0 ? 0 : __shfl(v, 63 - (key & 63), 64);
            if (key >= 0 && !placed && v >= base && v - base <= span) {
                placed = true;
                c = ((uint32_t)k << 10) | (uint32_t)(v - base);
            }
            if (lane == 0 && r < n) bw[k * 12 + sl] = base;
        }
        if (!placed) far = true;
        code[sl] = c;
    }
    if (!ok) return;
    uint32_t *rec = P + (size_t)r * 6;
    for (int w = 0; w < 6; w++)
        rec[w] = code[2 * w] | (code[2 * w + 1] << 12) | ((uint32_t)((f >> (6 * w)) & 63u) << 24) |
                 (w == 5 && far ? 0x80000000u : 0u);
    if (far) atomicAdd(nfar, 1);
}

// Load-time check of the packed Jacobi records against the link arrays:
// counts the entries whose twelve decoded links or flags differ.
__global__ __launch_bounds__(kListBlock) void k_check_jac(const uint32_t *__restrict__ P,
                                                          const int32_t *__restrict__ B,
                                                          const uint2 *__restrict__ L, int nE, int n,
                                                          int *__restrict__ bad) {
    const int r = (int)blockIdx.x * kListBlock + (int)threadIdx.x;
    if (r >= n) return;
    const __amdgpu_buffer_rsrc_t rsL =
        __builtin_amdgcn_make_buffer_rsrc((void *)L, (short)0, (int)((uint32_t)nE * 24u), 0x00020000);
    int32_t a[12];
    const uint64_t fa = jac_twelve(r, L, nE, a);
    JLinks l;
    int32_t dl[6];
    jac_packed_links(r, (int)threadIdx.x & 63, jac_wave_bases(B, r), P, L, nE, rsL, l, dl);
    const int32_t c[12] = {l.xm, l.xp, l.ym, l.yp, l.zm, l.zp, dl[0], dl[1], dl[2], dl[3], dl[4], dl[5]};
    bool same = (fa & 0xFFFFFFFFFull) == (l.f & 0xFFFFFFFFFull);
    for (int q = 0; q < 12; q++) same = same && a[q] == c[q];
    if (!same) atomicAdd(bad, 1);
}

void launch_pack_jac_links(const FluidLists &L, int span, int *d_far, int *d_bad, hipStream_t s) {
    const int n = L.n_all;
    if (n == 0) return;
    span = std::max(0, std::min(span, kJPackSpan));
    hipLaunchKernelGGL(k_pack_jac, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, L.clink, L.n_centries, n,
                       span, L.cpack, L.cbase, d_far);
    hipLaunchKernelGGL(k_check_jac, dim3((unsigned)((n + kListBlock - 1) / kListBlock)), dim3(kListBlock), 0, s,
                       L.cpack, L.cbase, L.clink, L.n_centries, n, d_bad);
}

// first_voxel4_zero on the compact state (no halo copies to write).
template <bool AOS>
__global__ __launch_bounds__(kListBlock) void k_jac4c_first_zero(const uint2 *__restrict__ L, int nE, int n,
                                                                 float4 *__restrict__ S1, Res4 R, float cu) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    first_zero_compact<AOS>(t, L, nE, S1, R, cu);
}

// vel_voxel4 (the first pass from a settled state, p carried) on the compact
// state; the outlet buffer by entry.
template <bool AOS>
__global__ __launch_bounds__(kListBlock) void k_jac4c_first(const uint2 *__restrict__ L, int nE, int n,
                                                            const float4 *__restrict__ S0,
                                                            const float *__restrict__ q0, float4 *__restrict__ S1,
                                                            Res4 R, float cu) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const __amdgpu_buffer_rsrc_t rsL =
        __builtin_amdgcn_make_buffer_rsrc((void *)L, (short)0, (int)((uint32_t)nE * 24u), 0x00020000);
    const __amdgpu_buffer_rsrc_t rsS =
        __builtin_amdgcn_make_buffer_rsrc((void *)S0, (short)0, (int)((uint32_t)nE * 16u), 0x00020000);
    const JLinks l = jac_links<AOS>(L, nE, t);
    const Cases cs = decode((uint8_t)(l.f & 0xff));
    const bool out = (l.f >> 32) & 1u;
    const float4 C = S0[t];
    const Star4 s = star_compact<false, AOS>(l, cs, S0, rsL, rsS, nE);
    const float pxm = out ? q0[l.xm] : s.XM.x;
    const float pym = out ? q0[l.ym] : s.YM.x;
    const float pzm = s.ZM.x;
    const float dpx = cs.bx ? 0.f : C.x - pxm;
    const float dpy = cs.by ? 0.f : C.x - pym;
    const float dpz = cs.bz ? 0.f : C.x - pzm;
    float su, sv, sw;
    visc_from4(cs, s, C, su, sv, sw);
    S1[t] = vel_from4(cs, C, su, sv, sw, dpx, dpy, dpz, R.m[t], C.x, cu);
}

// The settle pass on the compact state (settle_compact).
template <bool AOS>
__global__ __launch_bounds__(kListBlock) void k_jac4c_settle(const uint2 *__restrict__ L, int nE, int n, float4 *S,
                                                             float *__restrict__ q1, Res4 R, float cp) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    settle_compact<AOS>(t, L, nE, S, q1, R, cp);
}

// The compact Jacobi flags into the links' top bits, from the lagged pass's
// per-voxel records (natural offset; codes of the cell and of its three
// recompute targets) and the cell's plane.
template <bool AOS>
__global__ __launch_bounds__(kListBlock) void k_compact_jac_flags(Grid g, const uint2 *__restrict__ rec, int n,
                                                                  int nE, uint2 *L) {
    const int t = blockIdx.x * kListBlock + threadIdx.x;
    if (t >= n) return;
    const uint2 r = rec[t];
    int i, j, k;
    cell_ijk(g, (int64_t)r.x, i, j, k);
    const uint64_t f = (uint64_t)r.y | ((uint64_t)(i == g.nz + 1) << 32) | ((uint64_t)(i == 2) << 33);
    uint32_t *W = reinterpret_cast<uint32_t *>(L);
    for (int q = 0; q < 6; q++) {
        uint32_t &w = W[link_word<AOS>(t, q, nE)];
        w = (w & kLinkMask) | ((uint32_t)((f >> (6 * q)) & 63u) << kLinkBits);
    }
}

// The compact Jacobi state's links on the device: cidx (Grid::total entries)
// maps a cell to its entry, -1 elsewhere; entry t's six neighbours one step
// away, across a periodic face at the image cell, are looked up there (the
// zero entry where there is none).
__global__ __launch_bounds__(kListBlock) void k_cidx_scatter(const int32_t *__restrict__ list, int n, int32_t *cidx) {
    const int t = blockIdx.x * kListBlock + threadIdx.x;
    if (t < n) cidx[list[t]] = t;
}

template <bool AOS>
__global__ __launch_bounds__(kListBlock) void k_torus_links(Grid g, const int32_t *__restrict__ list, int n,
                                                            const int32_t *__restrict__ cidx, int nE, uint2 *L) {
    const int t = blockIdx.x * kListBlock + threadIdx.x;
    if (t >= nE) return;
    const int32_t zero = nE - 1;
    uint32_t w[6];
    if (t >= n) {
        for (int q = 0; q < 6; q++) w[q] = (uint32_t)zero;
    } else {
        int i, j, k;
        cell_ijk(g, list[t], i, j, k);
        auto at = [&](int ii, int jj, int kk) -> uint32_t {
            if (g.px) kk = kk < 2 ? kk + g.nx : kk >= g.nx + 2 ? kk - g.nx : kk;
            if (g.py) jj = jj < 2 ? jj + g.ny : jj >= g.ny + 2 ? jj - g.ny : jj;
            if (g.pz) ii = ii < 2 ? ii + g.nz : ii >= g.nz + 2 ? ii - g.nz : ii;
            if (ii < 0 || ii >= g.NZ || jj < 0 || jj >= g.NY || kk < 0 || kk >= g.NX) return (uint32_t)zero;
            const int32_t m = cidx[g.at(ii, jj, kk)];
            return (uint32_t)(m < 0 ? zero : m);
        };
        w[0] = at(i, j, k - 1); w[1] = at(i, j, k + 1);
        w[2] = at(i, j - 1, k); w[3] = at(i, j + 1, k);
        w[4] = at(i - 1, j, k); w[5] = at(i + 1, j, k);
    }
    uint32_t *W = reinterpret_cast<uint32_t *>(L);
    for (int q = 0; q < 6; q++) W[link_word<AOS>(t, q, nE)] = w[q];
}

hipError_t launch_torus_links(const Grid &g, const int32_t *list, int n, int32_t *cidx, uint2 *L, int nE, bool aos,
                              hipStream_t s) {
    const hipError_t e = hipMemsetAsync(cidx, 0xff, sizeof(int32_t) * (size_t)g.total, s);
    if (e != hipSuccess) return e;
    if (n > 0)
        hipLaunchKernelGGL(k_cidx_scatter, dim3((unsigned)((n + kListBlock - 1) / kListBlock)), dim3(kListBlock), 0, s,
                           list, n, cidx);
    const dim3 grid((unsigned)((nE + kListBlock - 1) / kListBlock)), blk(kListBlock);
    if (aos) hipLaunchKernelGGL(k_torus_links<true>, grid, blk, 0, s, g, list, n, cidx, nE, L);
    else hipLaunchKernelGGL(k_torus_links<false>, grid, blk, 0, s, g, list, n, cidx, nE, L);
    return hipSuccess;
}

// The +-2 links folded into the +-1 slots (both compact states).  A +-2
// neighbour is read only by the one-sided cases 4 (+2: the -1 neighbour is
// solid) and 5 (-2: the +1 neighbour is solid), where the slot of the solid
// side links to the zero entry and its value is never used: that slot takes
// the +-2 link instead (the +1 neighbour's own +1 link, the -1 neighbour's
// own -1 link), so a pass reads the +-2 neighbour with the load it makes for
// that slot anyway and needs no second hop.  Out of place (L0 read, L1
// written), interior entries t < n; the flags in the top bits are kept.
template <bool AOS>
__global__ __launch_bounds__(kListBlock) void k_link_subst(const uint2 *__restrict__ L0, uint2 *__restrict__ L1, int n,
                                                           int nE) {
    const int64_t t = (int64_t)blockIdx.x * kListBlock + threadIdx.x;
    const uint32_t *W0 = reinterpret_cast<const uint32_t *>(L0);
    uint32_t *W1 = reinterpret_cast<uint32_t *>(L1);
    if (t >= nE) return;
    uint32_t w[6];
    for (int q = 0; q < 6; q++) w[q] = W0[link_word<AOS>((int32_t)t, q, nE)];
    if (t < n) {
        const uint32_t code = (w[0] >> kLinkBits) | (((w[1] >> kLinkBits) & 3u) << 6);
        const Cases cs = decode((uint8_t)code);
        const int cc[3] = {cs.cx, cs.cy, cs.cz};
        for (int a = 0; a < 3; a++) {
            const int m = 2 * a, pl = 2 * a + 1;
            if (cc[a] == 4) {
                const int32_t e1 = (int32_t)(w[pl] & kLinkMask);
                const uint32_t e2 = W0[link_word<AOS>(e1, pl, nE)] & kLinkMask;
                w[m] = (w[m] & ~kLinkMask) | e2;
            } else if (cc[a] == 5) {
                const int32_t e1 = (int32_t)(w[m] & kLinkMask);
                const uint32_t e2 = W0[link_word<AOS>(e1, m, nE)] & kLinkMask;
                w[pl] = (w[pl] & ~kLinkMask) | e2;
            }
        }
    }
    for (int q = 0; q < 6; q++) W1[link_word<AOS>((int32_t)t, q, nE)] = w[q];
}

void launch_pack_links(const FluidLists &L, int span, int *d_far, int *d_bad, hipStream_t s) {
    span = std::max(0, std::min(span, kPackSpan));
    for (int col = 0; col < 3; col++) {
        const int n = L.n_col[col];
        if (n == 0) continue;
        hipLaunchKernelGGL(k_pack_links, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, L.clink, L.n_centries,
                           L.off_col[col], n, L.wave_off[col], span, L.cpack, L.cbase, d_far);
    }
    for (int col = 0; col < 3; col++) {
        const int n = L.n_col[col];
        if (n == 0) continue;
        hipLaunchKernelGGL(k_check_packed, dim3((unsigned)((n + kListBlock - 1) / kListBlock)), dim3(kListBlock), 0, s,
                           L.cpack, L.cbase, L.wave_off[col], L.clink, L.n_centries, n, L.off_col[col], d_bad);
    }
}

void launch_link_subst(const FluidLists &L, const uint2 *L0, uint2 *L1, hipStream_t s) {
    const int nE = L.n_centries;
    const dim3 grid((unsigned)((nE + kListBlock - 1) / kListBlock)), blk(kListBlock);
    if (L.links_aos) hipLaunchKernelGGL(k_link_subst<true>, grid, blk, 0, s, L0, L1, L.n_all, nE);
    else hipLaunchKernelGGL(k_link_subst<false>, grid, blk, 0, s, L0, L1, L.n_all, nE);
}

// The compact state's ghost fill (ghost_line4 on the entries): halo entries
// copied from their images, with p 0 where the dense fill pins it (bit 31 of
// the image: copies of the outlet plane and the z-halo planes 1 and nz + 2),
// the x / y halo copies of the outlet plane with their outlet buffer value
// (bit 31 of the halo entry), then the outlet plane's p pinned to 0.  A copy
// whose image is an outlet entry zeroes p itself, so it does not depend on
// the order of the two writes.
// Four items per thread (item k at t + k * gridDim.x * blockDim.x), all
// loads issued before any store: a few hundred thousand independent 16-byte
// copies are latency-bound, so each thread keeps four in flight.  Reading
// every item before writing any is the same as the one-at-a-time order: the
// images are interior entries (no copy writes what another reads), and a
// copy of an outlet entry zeroes its p itself.
constexpr int kGhostItems = 4;
__global__ __launch_bounds__(256) void k_ghost4_compact(const int2 *__restrict__ halo, int nh,
                                                        const int32_t *__restrict__ outl, int no, float4 *S,
                                                        float *q) {
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int kind[kGhostItems];  // 0: nothing, 1: halo copy, 2: outlet pin
    int32_t dst[kGhostItems];
    int2 h[kGhostItems];
    float4 v[kGhostItems];
    float qv[kGhostItems];
#pragma unroll
    for (int k = 0; k < kGhostItems; k++) {
        const int64_t t = t0 + k * stride;
        kind[k] = 0;
        dst[k] = 0;
        qv[k] = 0.f;
        v[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        h[k] = make_int2(0, 0);
        if (t < nh) {
            h[k] = halo[t];
            const int32_t src = h[k].y & 0x7fffffff;
            dst[k] = h[k].x & 0x7fffffff;
            v[k] = S[src];
            if (h[k].x < 0) qv[k] = q[src];
            kind[k] = 1;
        } else if (t - nh < no) {
            dst[k] = outl[t - nh];
            kind[k] = 2;
        }
    }
#pragma unroll
    for (int k = 0; k < kGhostItems; k++) {
        if (kind[k] == 1) {
            float4 w = v[k];
            if (h[k].y < 0) w.x = 0.f;
            if (h[k].x < 0) q[dst[k]] = qv[k];
            S[dst[k]] = w;
        } else if (kind[k] == 2) {
            S[dst[k]].x = 0.f;
        }
    }
}

// The colour passes' record flags (packed code, block, outlet flag: bits
// 0-10 of rec.y) into the top bits of the interior entries' -x / +x links,
// list positions 0 .. n - 1.
__global__ __launch_bounds__(kListBlock) void k_compact_records(const uint2 *__restrict__ rec, int n, uint2 *L) {
    const int t = blockIdx.x * kListBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t f = rec[t].y;
    uint2 l = L[t];
    l.x = (l.x & kLinkMask) | ((f & 63u) << kLinkBits);
    l.y = (l.y & kLinkMask) | (((f >> 6) & 31u) << kLinkBits);
    L[t] = l;
}

__global__ __launch_bounds__(kListBlock) void k_colour_records(Grid g, const int32_t *__restrict__ list, int n,
                                                               const uint8_t *__restrict__ pc, uint2 *rec) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    rec[t] = colour_record(g, list[t], pc);
}

// FP64 per-field lagged pass from the per-voxel records (lag_voxel_rec);
// SKIP2: its +-2 loads only where the case reads them.
template <typename T, bool SKIP2 = false>
__global__ __launch_bounds__(kListBlock) void k_jac_rec(Grid g, const uint2 *__restrict__ rec, int n,
                                                        const T *__restrict__ p0, const T *__restrict__ u0,
                                                        const T *__restrict__ v0, const T *__restrict__ w0,
                                                        const T *__restrict__ q0, T *__restrict__ p1,
                                                        T *__restrict__ u1, T *__restrict__ v1, T *__restrict__ w1,
                                                        T cu, T cp, T pin_out) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    lag_voxel_rec<T, SKIP2>(g, t, rec, p0, u0, v0, w0, q0, p1, u1, v1, w1, cu, cp, pin_out);
}

// Lagged pass from the per-voxel records (lag_voxel4_rec); IMAGES: the pass
// also writes the ghost fill's periodic copies and halo pins of its cells.
// OCC: waves per SIMD the register allocation must allow (0: the compiler's
// choice, 81 VGPRs = 5 waves); the same code at higher occupancy, for the
// autotuner.
// SKIP2: the +-2 stencil loads only for the lanes whose case reads them
// (load_star4_masked; the state below 2^32 bytes).
template <bool IMAGES, int OCC, bool SKIP2 = false, bool SKIP1 = false>
__global__ __launch_bounds__(kListBlock, OCC > 0 ? OCC : 1) void k_jac4_rec(
    Grid g, const uint2 *__restrict__ rec, int n, const float4 *__restrict__ S0, const float *__restrict__ q0,
    float4 *S1, Res4 R, float cu, float cp) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    lag_voxel4_rec<IMAGES, SKIP2, SKIP1>(g, t, rec, S0, q0, S1, R, cu, cp, 0.f);
}

// The masked stencil loads address the state with 32-bit byte offsets.
bool rec_skip2_supported(const Grid &g) { return g.total * 16 <= (int64_t)0xF0000000LL; }

// First pass from a zero correction state (first_voxel4_zero).
__global__ __launch_bounds__(kListBlock) void k_jac4_first_zero(Grid g, const uint2 *__restrict__ rec, int n,
                                                                float4 *S1, Res4 R, float cu) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    first_voxel4_zero(g, t, rec, S1, R, cu);
}

__global__ __launch_bounds__(kListBlock) void k_lag_records(Grid g, const int32_t *__restrict__ list, int n,
                                                            const uint8_t *__restrict__ pc, uint2 *rec) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    rec[t] = lag_record(g, list[t], pc);
}

// R.q from R.p: the continuity residual at each voxel and its recompute
// targets.
__global__ __launch_bounds__(kListBlock) void k_gather_rq(Grid g, const int32_t *__restrict__ list, int n, Res4 R) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t c = list[t];
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    int64_t nxn, nyn, nzn;
    lag_targets(g, c, i, j, k, nxn, nyn, nzn);
    R.q[t] = make_float4(R.p[c], R.p[nxn], R.p[nyn], R.p[nzn]);
}

// Thread -> list position for the once-per-interval kernels.  Plain lists:
// the identity.  Colour lists back to back (SOR): the i-th voxels of the three
// colour lists go to three consecutive threads, so that a wavefront walks the
// domain in space order instead of one colour at a time (the entries keep
// their list positions, which index the momentum residual).
struct ListWalk {
    int off[3], n[3];
    bool colours;
};

__device__ __forceinline__ int walk_pos(const ListWalk &W, int T, int n) {
    if (!W.colours) return T < n ? T : -1;
    const int q = T % 3, i = T / 3;
    return i < W.n[q] ? W.off[q] + i : -1;
}

static ListWalk list_walk(const FluidLists &L, int &threads) {
    ListWalk W;
    W.colours = L.col_in_all;
    int most = 0;
    for (int q = 0; q < 3; q++) {
        W.off[q] = L.off_col[q];
        W.n[q] = L.n_col[q];
        most = std::max(most, L.n_col[q]);
    }
    threads = W.colours ? 3 * most : L.n_all;
    return W;
}

__global__ __launch_bounds__(kListBlock) void k_residual4(Grid g, const int32_t *__restrict__ list, int n,
                                                          ListWalk W, const uint8_t *__restrict__ pc,
                                                          const double *__restrict__ p,
                                                          const double *__restrict__ u,
                                                          const double *__restrict__ v,
                                                          const double *__restrict__ w,
                                                          const double *__restrict__ q, Res4 R) {
    const int t = walk_pos(W, xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x, n);
    if (t < 0) return;
    const int64_t c = list[t];
    residual_voxel4(g, c, t, pc[c], p, u, v, w, q, R);
}

template <bool SPLIT>
__global__ void k_ghost4(Grid g, const uint8_t *__restrict__ G, float4 *S, float *q) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    ghost_line4<SPLIT>(g, t, G, S, q);
}

__global__ __launch_bounds__(kListBlock) void k_fold4(const int32_t *__restrict__ list, int n, double *Xp,
                                                      double *Xu, double *Xv, double *Xw, float4 *d) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    fold_voxel4(list[t], Xp, Xu, Xv, Xw, d);
}

// X += d, d = 0 at the fluid cells of the compact state (entry t = list
// position t, natural offset list[t]), the colour lists walked in space order
// as k_fold4_split walks them; the halo entries are refilled by the ghost fill
// after.
__global__ __launch_bounds__(kListBlock) void k_fold4_compact(const int32_t *__restrict__ list, int n, ListWalk W,
                                                              double *Xp, double *Xu, double *Xv, double *Xw,
                                                              float4 *d) {
    const int t = walk_pos(W, xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x, n);
    if (t < 0) return;
    fold_compact(t, list[t], Xp, Xu, Xv, Xw, d);
}

__global__ __launch_bounds__(kListBlock) void k_fold4_split(const int32_t *__restrict__ list,
                                                            const uint2 *__restrict__ rec, int n, ListWalk W,
                                                            double *Xp, double *Xu, double *Xv, double *Xw,
                                                            float4 *d) {
    const int t = walk_pos(W, xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x, n);
    if (t < 0) return;
    fold_voxel4_split(list[t], rec[t].x, Xp, Xu, Xv, Xw, d);
}

// ---------------------------------------------------------------------------
// Dense 2.5-D sweep (first or lagged Jacobi pass, pm_voxel.h semantics).
//
// A workgroup owns a tile of kSX x kSY cells in (x, y) and marches along z
// over a segment of planes.  Each thread keeps its own column's window of the
// state in registers (u, v at z-2..z+2, w at z-1..z+1: the z-derivatives), the
// tile's current plane of u, v, w with a 2-cell ring goes through LDS (the x-
// and y-derivatives), and the plane of lagged pressures p^n of the tile and
// its -1 ring is computed once into LDS, so a cell costs about ten coalesced
// global loads instead of the ~40 gathers of the list kernels.  Interior solid
// cells are skipped (their velocities are exactly +0 and no fluid cell reads
// their pressure), so fully solid lines are never fetched.  The arithmetic is
// pm_voxel.h's, value for value: the results are bit-equal to the list kernels.
// ---------------------------------------------------------------------------
constexpr int kSX = 64, kSY = 8, kSN = kSX * kSY;   // tile, threads per workgroup
constexpr int kLX = kSX + 4, kLY = kSY + 4;          // LDS plane with a 2-cell ring
constexpr int kPX = kSX + 1, kPY = kSY + 1;          // pressure plane with the -1 ring
constexpr int kSHalo = 4 * kLX + 4 * kSY;            // ring cells of the LDS plane
constexpr int kSRing = kSX + 1 + kSY;                // ring cells of the pressure plane

__device__ __forceinline__ bool interior(const Grid &g, int i, int j, int k) {
    return i >= 2 && i < g.nz + 2 && j >= 2 && j < g.ny + 2 && k >= 2 && k < g.nx + 2;
}

__device__ __forceinline__ uint8_t pc_at(const Grid &g, const uint8_t *__restrict__ pc, int i, int j, int k) {
    return (i < g.NZ && j < g.NY && k < g.NX) ? pc[g.at(i, j, k)] : kSolid;
}

// State value at (i, j, k): 0 outside the padded field and at interior solid
// cells (exact: their velocities are +0), else loaded.
template <typename T>
__device__ __forceinline__ T st_ld(const Grid &g, const T *__restrict__ f, int i, int j, int k, uint8_t pcv) {
    if (j >= g.NY || k >= g.NX) return (T)0;
    if (pcv == kSolid && interior(g, i, j, k)) return (T)0;
    return f[g.at(i, j, k)];
}

template <typename T, int MODE>
__global__ __launch_bounds__(kSN, 4) void k_jac_stream(Grid g, int ntx, int nty, int zs,
                                                     const uint8_t *__restrict__ pc, const T *__restrict__ p0,
                                                     const T *__restrict__ u0, const T *__restrict__ v0,
                                                     const T *__restrict__ w0, const T *__restrict__ q0,
                                                     T *__restrict__ p1, T *__restrict__ u1,
                                                     T *__restrict__ v1, T *__restrict__ w1,
                                                     T cu, T cp, T pin_out) {
    __shared__ T Lu[2][kLY][kLX], Lv[2][kLY][kLX], Lw[2][kLY][kLX], LP[2][kPY][kPX];
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int txt = b % ntx, rest = b / ntx, tyt = rest % nty, seg = rest / nty;
    const int k0 = 2 + txt * kSX, j0 = 2 + tyt * kSY;
    const int i0 = 2 + seg * zs, i1 = min(i0 + zs, g.nz + 2);
    const int tid = threadIdx.x, tx = tid % kSX, ty = tid / kSX;
    const int j = j0 + ty, k = k0 + tx;
    const bool own_in = j < g.ny + 2 && k < g.nx + 2;
    const T INV_RE = (T)(1.0 / kRe), C2DT = (T)(kC2 * kDt);

    // LDS-ring cell loaded by this thread
    int hly = -1, hlx = -1;
    if (tid < kSHalo) {
        if (tid < 2 * kLX) { hly = tid / kLX; hlx = tid % kLX; }
        else if (tid < 4 * kLX) { const int h = tid - 2 * kLX; hly = kSY + 2 + h / kLX; hlx = h % kLX; }
        else { const int h = tid - 4 * kLX; hly = 2 + h / 4; const int s = h % 4; hlx = s < 2 ? s : kSX + s; }
    }
    const int hj = j0 - 2 + hly, hk = k0 - 2 + hlx;
    const bool h_on = hly >= 0 && hj < g.NY && hk < g.NX;
    // pressure-ring cell computed by this thread (corner (j0-1, k0-1) unused)
    int rly = -1, rlx = -1;
    if (tid < kSRing) {
        if (tid < kSX + 1) { rly = 1; rlx = 1 + tid; }
        else { rly = 2 + (tid - kSX - 1); rlx = 1; }
    }
    const int rj = j0 - 2 + rly, rk = k0 - 2 + rlx;
    const bool r_on = rly >= 0 && rj <= g.ny + 1 && rk <= g.nx + 1 && !(rj < 2 && rk < 2);
    const bool r_int = r_on && rj >= 2 && rk >= 2;

    // Software pipeline: everything step i reads from global memory was
    // requested during step i-1 (or the prologue), with the solid-cell
    // predicates taken from codes fetched one step earlier still.
    uint8_t c0 = pc_at(g, pc, i0, j, k), c1 = pc_at(g, pc, i0 + 1, j, k), c2 = pc_at(g, pc, i0 + 2, j, k),
            c3 = pc_at(g, pc, i0 + 3, j, k);
    const uint8_t cm2 = pc_at(g, pc, i0 - 2, j, k), cm1 = pc_at(g, pc, i0 - 1, j, k);
    T um2 = st_ld(g, u0, i0 - 2, j, k, cm2), vm2 = st_ld(g, v0, i0 - 2, j, k, cm2);
    T um1 = st_ld(g, u0, i0 - 1, j, k, cm1), vm1 = st_ld(g, v0, i0 - 1, j, k, cm1);
    T wm1 = st_ld(g, w0, i0 - 1, j, k, cm1);
    T uc = st_ld(g, u0, i0, j, k, c0), vc = st_ld(g, v0, i0, j, k, c0), wc = st_ld(g, w0, i0, j, k, c0);
    T up1 = st_ld(g, u0, i0 + 1, j, k, c1), vp1 = st_ld(g, v0, i0 + 1, j, k, c1);
    T wp1 = st_ld(g, w0, i0 + 1, j, k, c1);
    T up2 = st_ld(g, u0, i0 + 2, j, k, c2), vp2 = st_ld(g, v0, i0 + 2, j, k, c2);
    T pcur = (T)0;
    if (own_in && c0 != kSolid) pcur = p0[g.at(i0, j, k)];
    T hu = (T)0, hv = (T)0, hw = (T)0;
    uint8_t hc1 = kSolid;
    if (h_on) {
        const uint8_t hc0 = pc[g.at(i0, hj, hk)];
        hu = st_ld(g, u0, i0, hj, hk, hc0);
        hv = st_ld(g, v0, i0, hj, hk, hc0);
        hw = st_ld(g, w0, i0, hj, hk, hc0);
        hc1 = pc[g.at(i0 + 1, hj, hk)];
    }
    uint8_t rc0 = kSolid, rc1 = kSolid;
    T rp0 = (T)0, rw1 = (T)0;
    if (r_int) {
        const int64_t r = g.at(i0, rj, rk);
        rc0 = pc[r];
        if (rc0 != kSolid) {
            rp0 = p0[r];
            rw1 = w0[r + g.sy];
        }
        rc1 = pc[r + g.sy];
    }
    // pressure of the own column one plane below the segment
    T Pprev = (T)0;
    if (own_in) {
        if (i0 == 2 || MODE == JAC_FIRST) Pprev = p0[g.at(i0 - 1, j, k)];
        else if (cm1 != kSolid) Pprev = lag_p<T>(g, g.at(i0 - 1, j, k), cm1, p0, u0, v0, w0, cp);
    }

    for (int i = i0; i < i1; i++) {
        const int buf = i & 1;
        const bool out = i == g.nz + 1;
        const bool more = i + 1 < i1;

        // requests for step i+1
        uint8_t c4 = kSolid, hc2 = kSolid, rc2 = kSolid;
        T u3 = (T)0, v3 = (T)0, w2 = (T)0, pn = (T)0;
        T hun = (T)0, hvn = (T)0, hwn = (T)0, rp0n = (T)0, rw1n = (T)0;
        if (more) {
            c4 = pc_at(g, pc, i + 4, j, k);
            u3 = st_ld(g, u0, i + 3, j, k, c3);
            v3 = st_ld(g, v0, i + 3, j, k, c3);
            w2 = st_ld(g, w0, i + 2, j, k, c2);
            if (own_in && c1 != kSolid) pn = p0[g.at(i + 1, j, k)];
            if (h_on) {
                hun = st_ld(g, u0, i + 1, hj, hk, hc1);
                hvn = st_ld(g, v0, i + 1, hj, hk, hc1);
                hwn = st_ld(g, w0, i + 1, hj, hk, hc1);
                hc2 = pc[g.at(i + 2, hj, hk)];
            }
            if (r_int) {
                const int64_t r = g.at(i + 1, rj, rk);
                if (rc1 != kSolid) {
                    rp0n = p0[r];
                    rw1n = w0[r + g.sy];
                }
                rc2 = pc[r + g.sy];
            }
        }

        Lu[buf][ty + 2][tx + 2] = uc;
        Lv[buf][ty + 2][tx + 2] = vc;
        Lw[buf][ty + 2][tx + 2] = wc;
        if (hly >= 0) {
            Lu[buf][hly][hlx] = hu;
            Lv[buf][hly][hlx] = hv;
            Lw[buf][hly][hlx] = hw;
        }
        __syncthreads();

        // lagged pressure (first pass: the given one) of the own cell and the
        // -1 ring; on the outlet plane the ring/neighbour value is the buffer q
        const bool own_fluid = own_in && c0 != kSolid;
        T Pc = (T)0, lp = (T)0;
        if (own_fluid) {
            if (MODE == JAC_FIRST) {
                Pc = pcur;
                lp = out ? q0[g.qat(j, k)] : Pc;
            } else {
                const Cases cs = decode(c0);
                const T wz = fwd_solid(cs.cz) ? (T)0 : wp1;
                const T vy = fwd_solid(cs.cy) ? (T)0 : Lv[buf][ty + 3][tx + 2];
                const T ux = fwd_solid(cs.cx) ? (T)0 : Lu[buf][ty + 2][tx + 3];
                const T d0 = ((((wz - wc) + vy) - vc) + ux) - uc;
                if (out) {
                    Pc = pin_out;
                    lp = pin_out - d0 * C2DT;
                } else {
                    Pc = pcur - d0 * cp;
                    lp = Pc;
                }
            }
        }
        LP[buf][ty + 1][tx + 1] = lp;
        if (rly >= 0) {
            T val = (T)0;
            if (r_int) {
                if (rc0 != kSolid) {
                    if (MODE == JAC_FIRST) {
                        val = out ? q0[g.qat(rj, rk)] : rp0;
                    } else {
                        const Cases rs = decode(rc0);
                        const T wz = fwd_solid(rs.cz) ? (T)0 : rw1;
                        const T vy = fwd_solid(rs.cy) ? (T)0 : Lv[buf][rly + 1][rlx];
                        const T ux = fwd_solid(rs.cx) ? (T)0 : Lu[buf][rly][rlx + 1];
                        const T d0 = ((((wz - Lw[buf][rly][rlx]) + vy) - Lv[buf][rly][rlx]) + ux) -
                                     Lu[buf][rly][rlx];
                        if (out) {
                            val = pin_out - d0 * C2DT;
                        } else {
                            val = rp0 - d0 * cp;
                        }
                    }
                }
            } else if (r_on) {
                // ring cell in the domain's x or y halo
                const bool xh = rk < 2;
                const bool per = xh ? g.px : g.py;
                if (MODE == JAC_FIRST || !per) {
                    val = out ? q0[g.qat(rj, rk)] : p0[g.at(i, rj, rk)];
                } else {
                    const int wj = xh ? rj : rj + g.ny, wk = xh ? rk + g.nx : rk;
                    const int64_t n = g.at(i, wj, wk);
                    const uint8_t nc = pc[n];
                    if (nc != kSolid)
                        val = out ? lag_q<T>(g, n, nc, u0, v0, w0, pin_out)
                                  : lag_p<T>(g, n, nc, p0, u0, v0, w0, cp);
                }
            }
            LP[buf][rly - 1][rlx - 1] = val;
        }
        __syncthreads();

        if (own_fluid) {
            const int64_t c = g.at(i, j, k);
            const Cases cs = decode(c0);
            const T dpx = cs.bx ? (T)0 : Pc - LP[buf][ty + 1][tx];
            const T dpy = cs.by ? (T)0 : Pc - LP[buf][ty][tx + 1];
            const T dpz = cs.bz ? (T)0 : Pc - Pprev;
            const int ly = ty + 2, lx = tx + 2;
            const T lxx = ((cs.bx ? (T)0 : Lu[buf][ly][lx - 1]) - (T)2.0 * uc) +
                          (fwd_solid(cs.cx) ? (T)0 : Lu[buf][ly][lx + 1]);
            const T lyy = ((cs.by ? (T)0 : Lv[buf][ly - 1][lx]) - (T)2.0 * vc) +
                          (fwd_solid(cs.cy) ? (T)0 : Lv[buf][ly + 1][lx]);
            const T lzz = ((cs.bz ? (T)0 : wm1) - (T)2.0 * wc) + (fwd_solid(cs.cz) ? (T)0 : wp1);
            const T dyx = d2<T>(cs.cx, Lv[buf][ly][lx - 2], Lv[buf][ly][lx - 1], vc, Lv[buf][ly][lx + 1],
                                Lv[buf][ly][lx + 2]);
            const T dzx = d2<T>(cs.cx, Lw[buf][ly][lx - 2], Lw[buf][ly][lx - 1], wc, Lw[buf][ly][lx + 1],
                                Lw[buf][ly][lx + 2]);
            const T dxy = d2<T>(cs.cy, Lu[buf][ly - 2][lx], Lu[buf][ly - 1][lx], uc, Lu[buf][ly + 1][lx],
                                Lu[buf][ly + 2][lx]);
            const T dzy = d2<T>(cs.cy, Lw[buf][ly - 2][lx], Lw[buf][ly - 1][lx], wc, Lw[buf][ly + 1][lx],
                                Lw[buf][ly + 2][lx]);
            const T dxz = d2<T>(cs.cz, um2, um1, uc, up1, up2);
            const T dyz = d2<T>(cs.cz, vm2, vm1, vc, vp1, vp2);
            const T su = (lxx + dxy) + dxz;
            const T sv = (lyy + dyz) + dyx;
            const T sw = (lzz + dzx) + dzy;
            T fu = cs.bx ? (T)0 : su * INV_RE - dpx;
            T fv = cs.by ? (T)0 : sv * INV_RE - dpy;
            T fw = cs.bz ? (T)0 : sw * INV_RE - dpz;
            u1[c] = cs.bx ? (T)0 : uc + fu * cu;
            v1[c] = cs.by ? (T)0 : vc + fv * cu;
            w1[c] = cs.bz ? (T)0 : wc + fw * cu;
            p1[c] = Pc;
        }

        // advance the pipeline one plane
        Pprev = Pc;
        um2 = um1; um1 = uc; uc = up1; up1 = up2; up2 = u3;
        vm2 = vm1; vm1 = vc; vc = vp1; vp1 = vp2; vp2 = v3;
        wm1 = wc; wc = wp1; wp1 = w2;
        c0 = c1; c1 = c2; c2 = c3; c3 = c4;
        pcur = pn;
        hu = hun; hv = hvn; hw = hwn; hc1 = hc2;
        rc0 = rc1; rc1 = rc2; rp0 = rp0n; rw1 = rw1n;
    }
}

template <typename T, bool SEL>
__global__ __launch_bounds__(kListBlock) void k_colour(Grid g, const int32_t *__restrict__ list, int n,
                                                       const uint8_t *__restrict__ pc, T *p, T *u, T *v, T *w,
                                                       T *q, T cu, T cp, T pin_out) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t c = list[t];
    colour_voxel<T, SEL>(g, c, pc[c], p, u, v, w, q, cu, cp, pin_out);
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

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline unsigned lblocks(int n) { return (unsigned)((n + kListBlock - 1) / kListBlock); }
static inline unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// The FP64 fast modes' launchers (per-field state).
template <bool SEL>
static void jacobi_pair(const Grid &g, const FluidLists &L, const uint8_t *pc, const double *p0, const double *u0,
                        const double *v0, const double *w0, const double *q0, double *p1, double *u1, double *v1,
                        double *w1, double *q1, double cu, double cp, double pin_out, hipStream_t s) {
    hipLaunchKernelGGL((k_jac_vel<double, SEL>), dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, g, L.all, L.n_all,
                       pc, p0, u0, v0, w0, q0, u1, v1, w1, cu);
    hipLaunchKernelGGL(k_jac_pres<double>, dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, g, L.all, L.n_all, pc, p0,
                       u1, v1, w1, p1, q1, cp, pin_out);
}

void launch_jacobi(const Grid &g, const FluidLists &L, const uint8_t *pc, const double *p0, const double *u0,
                   const double *v0, const double *w0, const double *q0, double *p1, double *u1, double *v1,
                   double *w1, double *q1, double cu, double cp, double pin_out, hipStream_t s) {
    if (L.n_all == 0) return;
    if (L.select_loads)
        jacobi_pair<true>(g, L, pc, p0, u0, v0, w0, q0, p1, u1, v1, w1, q1, cu, cp, pin_out, s);
    else
        jacobi_pair<false>(g, L, pc, p0, u0, v0, w0, q0, p1, u1, v1, w1, q1, cu, cp, pin_out, s);
}

template <bool SEL>
static void jacobi_stage_t(int stage, const Grid &g, const FluidLists &L, const uint8_t *pc, const double *p0,
                           const double *u0, const double *v0, const double *w0, const double *q0, double *p1,
                           double *u1, double *v1, double *w1, double *q1, double cu, double cp, double pin_out,
                           hipStream_t s) {
    const dim3 grid(lblocks(L.n_all)), blk(kListBlock);
    if (stage == JAC_FIRST)
        hipLaunchKernelGGL((k_jac_first<double, SEL>), grid, blk, 0, s, g, L.all, L.n_all, pc, p0, u0, v0, w0, q0,
                           p1, u1, v1, w1, cu);
    else if (stage == JAC_LAG && L.rec && L.use_rec && L.rec_skip2 && rec_skip2_supported(g))
        hipLaunchKernelGGL((k_jac_rec<double, true>), grid, blk, 0, s, g, L.rec, L.n_all, p0, u0, v0, w0, q0, p1, u1,
                           v1, w1, cu, cp, pin_out);
    else if (stage == JAC_LAG && L.rec && L.use_rec)
        hipLaunchKernelGGL(k_jac_rec<double>, grid, blk, 0, s, g, L.rec, L.n_all, p0, u0, v0, w0, q0, p1, u1, v1, w1,
                           cu, cp, pin_out);
    else if (stage == JAC_LAG)
        hipLaunchKernelGGL((k_jac_lag<double, SEL>), grid, blk, 0, s, g, L.all, L.n_all, pc, p0, u0, v0, w0, q0, p1,
                           u1, v1, w1, cu, cp, pin_out);
    else
        hipLaunchKernelGGL(k_jac_pres<double>, grid, blk, 0, s, g, L.all, L.n_all, pc, p0, u0, v0, w0, p1, q1, cp,
                           pin_out);
}

void launch_jacobi_stage(int stage, const Grid &g, const FluidLists &L, const uint8_t *pc, const double *p0,
                         const double *u0, const double *v0, const double *w0, const double *q0, double *p1,
                         double *u1, double *v1, double *w1, double *q1, double cu, double cp, double pin_out,
                         hipStream_t s) {
    if (L.n_all == 0) return;
    if (L.select_loads)
        jacobi_stage_t<true>(stage, g, L, pc, p0, u0, v0, w0, q0, p1, u1, v1, w1, q1, cu, cp, pin_out, s);
    else
        jacobi_stage_t<false>(stage, g, L, pc, p0, u0, v0, w0, q0, p1, u1, v1, w1, q1, cu, cp, pin_out, s);
}

bool jacobi4_fills_halos(int stage, const FluidLists &L) {
    if (L.compact_jac) return true;  // no halo entries: nothing to fill
    if (stage == JAC_FIRST_ZERO) return L.rec && L.fill_halos;
    if (stage == JAC_LAG && L.use_march && L.rowstart) return true;
    return stage == JAC_LAG && L.select_loads && L.rec && L.use_rec && L.fill_halos;
}

template <bool AOS>
static void launch_jacobi4_compact(int stage, const FluidLists &L, const float4 *S0, const float *q0, float4 *S1,
                                   float *q1, Res4 R, float cu, float cp, hipStream_t s) {
    const dim3 grid(lblocks(L.n_all)), blk(kListBlock);
    const int nE = L.n_centries;
    if (stage == JAC_FIRST_ZERO)
        hipLaunchKernelGGL(k_jac4c_first_zero<AOS>, grid, blk, 0, s, L.clink, nE, L.n_all, S1, R, cu);
    else if (stage == JAC_FIRST)
        hipLaunchKernelGGL(k_jac4c_first<AOS>, grid, blk, 0, s, L.clink, nE, L.n_all, S0, q0, S1, R, cu);
    else if (stage == JAC_LAG && L.packed && L.rec_skip1 && L.rec_occ == 6)
        hipLaunchKernelGGL((k_jac4c_lagp<false, 6>), grid, blk, 0, s, L.cpack, L.cbase, L.clink, nE, L.n_all, S0,
                           S1, R, cu, cp);
    else if (stage == JAC_LAG && L.packed && L.rec_skip1)
        hipLaunchKernelGGL((k_jac4c_lagp<false>), grid, blk, 0, s, L.cpack, L.cbase, L.clink, nE, L.n_all, S0, S1,
                           R, cu, cp);
    else if (stage == JAC_LAG && L.packed)
        hipLaunchKernelGGL((k_jac4c_lagp<true>), grid, blk, 0, s, L.cpack, L.cbase, L.clink, nE, L.n_all, S0, S1, R,
                           cu, cp);
    else if (stage == JAC_LAG && L.rec_skip1 && L.rec_occ == 6)
        hipLaunchKernelGGL((k_jac4c_lag<false, AOS, 6>), grid, blk, 0, s, L.clink, nE, L.n_all, S0, S1, R, cu, cp);
    else if (stage == JAC_LAG && L.rec_skip1)
        hipLaunchKernelGGL((k_jac4c_lag<false, AOS>), grid, blk, 0, s, L.clink, nE, L.n_all, S0, S1, R, cu, cp);
    else if (stage == JAC_LAG)
        hipLaunchKernelGGL((k_jac4c_lag<true, AOS>), grid, blk, 0, s, L.clink, nE, L.n_all, S0, S1, R, cu, cp);
    else  // JAC_SETTLE (in place, S1 == S0); the pair has no compact form
        hipLaunchKernelGGL(k_jac4c_settle<AOS>, grid, blk, 0, s, L.clink, nE, L.n_all, S1, q1, R, cp);
}

void launch_jacobi4(int stage, const Grid &g, const FluidLists &L, const uint8_t *pc, const float4 *S0,
                    const float *q0, float4 *S1, float *q1, Res4 R, float cu, float cp, hipStream_t s) {
    if (L.n_all == 0) return;
    const dim3 grid(lblocks(L.n_all)), blk(kListBlock);
    if (L.compact_jac) {
        if (L.links_aos) launch_jacobi4_compact<true>(stage, L, S0, q0, S1, q1, R, cu, cp, s);
        else launch_jacobi4_compact<false>(stage, L, S0, q0, S1, q1, R, cu, cp, s);
        return;
    }
    if (stage == JAC_FIRST_ZERO && !jacobi4_fills_halos(stage, L)) stage = JAC_FIRST;  // the state is zero
    if (stage == JAC_FIRST_ZERO) {
        hipLaunchKernelGGL(k_jac4_first_zero, grid, blk, 0, s, g, L.rec, L.n_all, S1, R, cu);
    } else if (stage == JAC_LAG && L.use_march && L.rowstart) {
        launch_jacobi4_march(g, L, pc, S0, q0, S1, R, cu, cp, L.march_ty, L.march_zseg, s);
    } else if (stage == JAC_FIRST) {
        hipLaunchKernelGGL(k_jac4<0>, grid, blk, 0, s, g, L.all, L.n_all, pc, S0, q0, S1, q1, R, cu, cp);
    } else if (jacobi4_fills_halos(stage, L)) {
        if (L.rec_skip2 && L.rec_skip1 && rec_skip2_supported(g))
            hipLaunchKernelGGL((k_jac4_rec<true, 5, true, true>), grid, blk, 0, s, g, L.rec, L.n_all, S0, q0, S1, R,
                               cu, cp);
        else if (L.rec_skip2 && rec_skip2_supported(g))
            hipLaunchKernelGGL((k_jac4_rec<true, 0, true>), grid, blk, 0, s, g, L.rec, L.n_all, S0, q0, S1, R, cu,
                               cp);
        else if (L.rec_occ == 8)
            hipLaunchKernelGGL((k_jac4_rec<true, 8>), grid, blk, 0, s, g, L.rec, L.n_all, S0, q0, S1, R, cu, cp);
        else if (L.rec_occ == 6)
            hipLaunchKernelGGL((k_jac4_rec<true, 6>), grid, blk, 0, s, g, L.rec, L.n_all, S0, q0, S1, R, cu, cp);
        else
            hipLaunchKernelGGL((k_jac4_rec<true, 0>), grid, blk, 0, s, g, L.rec, L.n_all, S0, q0, S1, R, cu, cp);
    } else if (stage == JAC_LAG && L.select_loads && L.rec && L.use_rec) {
        hipLaunchKernelGGL((k_jac4_rec<false, 0>), grid, blk, 0, s, g, L.rec, L.n_all, S0, q0, S1, R, cu, cp);
    } else if (stage == JAC_LAG && L.select_loads) {
        hipLaunchKernelGGL(k_jac4<4>, grid, blk, 0, s, g, L.all, L.n_all, pc, S0, q0, S1, q1, R, cu, cp);
    } else if (stage == JAC_LAG) {
        hipLaunchKernelGGL(k_jac4<1>, grid, blk, 0, s, g, L.all, L.n_all, pc, S0, q0, S1, q1, R, cu, cp);
    } else if (stage == JAC_PAIR) {
        hipLaunchKernelGGL(k_jac4<2>, grid, blk, 0, s, g, L.all, L.n_all, pc, S0, q0, S1, q1, R, cu, cp);
        hipLaunchKernelGGL(k_jac4<3>, grid, blk, 0, s, g, L.all, L.n_all, pc, S0, q0, S1, q1, R, cu, cp);
    } else {
        hipLaunchKernelGGL(k_jac4<3>, grid, blk, 0, s, g, L.all, L.n_all, pc, S0, q0, S1, q1, R, cu, cp);
    }
}

// Colour q's list range on planes [z0, z1) (clipped to the interior).
static void colour_planes(const Grid &g, const FluidLists &L, int q, int z0, int z1, int &t0, int &n) {
    z0 = std::max(z0, 0);
    z1 = std::min(z1, g.nz);
    t0 = L.off_col[q];
    n = 0;
    if (z1 <= z0) return;
    const int32_t *po = L.plane_off.data() + (size_t)q * (g.nz + 1);
    t0 += po[z0];
    n = po[z1] - po[z0];
}

// The three colour passes of one SOR sweep in z-chunks of Z planes.  A cell
// of colour q reads the cells within 2 planes of its own, of the two other
// colours, as the three whole passes leave them: colour 1 on plane i needs
// colour 0 finished on planes <= i + 2, and colour 0 on plane i + 2 must
// still read colour 1's old values on plane i.  Colour 1 therefore trails
// colour 0 by 3 planes and colour 2 trails colour 1 by 3: stage k updates
// colour 0 on [kZ, kZ + Z), colour 1 on [(k-1)Z - 3, kZ - 3) and colour 2 on
// [(k-2)Z - 6, (k-1)Z - 6) (fused: one launch; their reads and writes are
// at least one plane apart), or, in sequence, colour q on [kZ - 3q,
// kZ + Z - 3q) for q = 0, 1, 2.  Both give the three whole passes' values
// bit for bit; halos are frozen during a sweep as there.
static void launch_colour4_chunked(const Grid &g, const FluidLists &L, float4 *S, float *q, Res4 R, float cu,
                                   float cp, hipStream_t s) {
    const int Z = L.sor_chunk;
    const bool skip2 = L.sor_skip2 && rec_skip2_supported(g);
    for (int k = 0;; k++) {
        ColourStage st;
        int lo[3];
        for (int col = 0; col < 3; col++) lo[col] = L.sor_fused ? (k - col) * Z - 3 * col : k * Z - 3 * col;
        if (lo[2] >= g.nz) break;
        if (L.sor_fused) {
            int nb = 0;
            for (int col = 0; col < 3; col++) {
                colour_planes(g, L, col, lo[col], lo[col] + Z, st.t0[col], st.n[col]);
                st.b[col] = nb;
                nb += (int)lblocks(st.n[col]);
            }
            if (nb > 0 && skip2)
                hipLaunchKernelGGL(k_colour4_stage<true>, dim3((unsigned)nb), dim3(kListBlock), 0, s, g, L.rec, st, S,
                                   q, R, cu, cp);
            else if (nb > 0)
                hipLaunchKernelGGL(k_colour4_stage<false>, dim3((unsigned)nb), dim3(kListBlock), 0, s, g, L.rec, st,
                                   S, q, R, cu, cp);
        } else {
            for (int col = 0; col < 3; col++) {
                int t0, n;
                colour_planes(g, L, col, lo[col], lo[col] + Z, t0, n);
                if (n > 0 && skip2)
                    hipLaunchKernelGGL(k_colour4_split<true>, dim3(lblocks(n)), dim3(kListBlock), 0, s, g, L.rec, n,
                                       t0, S, q, R, cu, cp);
                else if (n > 0)
                    hipLaunchKernelGGL(k_colour4_split<false>, dim3(lblocks(n)), dim3(kListBlock), 0, s, g, L.rec, n,
                                       t0, S, q, R, cu, cp);
            }
        }
    }
}

bool compact_supported(const Grid &g, int64_t n_entries) {
    return !g.slip && n_entries > 0 && n_entries <= (int64_t)kLinkMask;
}

void launch_colour4(const Grid &g, const FluidLists &L, const uint8_t *pc, float4 *S, float *q, Res4 R, float cu,
                    float cp, hipStream_t s) {
    if (L.compact) {
        for (int col = 0; col < 3; col++) {
            const int n = L.n_col[col];
            if (n == 0) continue;
            if (L.packed)
                hipLaunchKernelGGL(k_colour4_packed, dim3(lblocks(n)), dim3(kListBlock), 0, s, L.cpack, L.cbase,
                                   L.wave_off[col], L.clink, L.n_centries, n, L.off_col[col], S, q, R, cu, cp);
            else
                hipLaunchKernelGGL(k_colour4_compact, dim3(lblocks(n)), dim3(kListBlock), 0, s, L.clink,
                                   L.n_centries, n, L.off_col[col], S, q, R, cu, cp);
        }
        return;
    }
    if (L.split && L.sor_chunk > 0 && !L.plane_off.empty()) {
        launch_colour4_chunked(g, L, S, q, R, cu, cp, s);
        return;
    }
    for (int col = 0; col < 3; col++) {
        const int n = L.n_col[col];
        if (n == 0) continue;
        if (L.split && L.sor_skip2 && rec_skip2_supported(g))
            hipLaunchKernelGGL(k_colour4_split<true>, dim3(lblocks(n)), dim3(kListBlock), 0, s, g, L.rec, n,
                               L.off_col[col], S, q, R, cu, cp);
        else if (L.split)
            hipLaunchKernelGGL(k_colour4_split<false>, dim3(lblocks(n)), dim3(kListBlock), 0, s, g, L.rec, n,
                               L.off_col[col], S, q, R, cu, cp);
        else
            hipLaunchKernelGGL(k_colour4, dim3(lblocks(n)), dim3(kListBlock), 0, s, g, L.all, n, L.off_col[col], pc,
                               S, q, R, cu, cp);
    }
}

void launch_lag_records(const Grid &g, const FluidLists &L, const uint8_t *pc, hipStream_t s) {
    if (L.n_all == 0 || !L.rec) return;
    if (L.split) {
        hipLaunchKernelGGL(k_colour_records, dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, g, L.all, L.n_all, pc,
                           L.rec);
        if (L.compact)
            hipLaunchKernelGGL(k_compact_records, dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, L.rec, L.n_all,
                               L.clink);
    }
    else {
        hipLaunchKernelGGL(k_lag_records, dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, g, L.all, L.n_all, pc,
                           L.rec);
        if (L.compact_jac)
        {
            if (L.links_aos)
                hipLaunchKernelGGL(k_compact_jac_flags<true>, dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, g, L.rec,
                                   L.n_all, L.n_centries, L.clink);
            else
                hipLaunchKernelGGL(k_compact_jac_flags<false>, dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, g,
                                   L.rec, L.n_all, L.n_centries, L.clink);
        }
    }
}

// The compact state's continuity residual per entry (k_gather_rq's
// counterpart): entry t is list position t.
__global__ __launch_bounds__(kListBlock) void k_gather_rp_compact(const int32_t *__restrict__ list, int n, Res4 R) {
    const int t = blockIdx.x * kListBlock + threadIdx.x;
    if (t >= n) return;
    R.pc[t] = R.p[list[t]];
}

void launch_gather_rq(const Grid &g, const FluidLists &L, Res4 R, hipStream_t s) {
    if (L.n_all == 0) return;
    // the compact state's residual per entry (also while its layout is being
    // timed against the dense one, which needs the gathered one)
    if (R.pc)
        hipLaunchKernelGGL(k_gather_rp_compact, dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, L.all, L.n_all, R);
    if (!R.q) return;
    hipLaunchKernelGGL(k_gather_rq, dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, g, L.all, L.n_all, R);
}

// The residual and the fold of the compact SOR state walked in column order
// (nat) rather than colour by colour: the FP64 state is read in space order,
// the residual's momentum part and the correction written at the cells'
// colour-list positions (natpos).
__global__ __launch_bounds__(kListBlock) void k_residual4_nat(Grid g, const int32_t *__restrict__ nat,
                                                              const int32_t *__restrict__ natpos, int n,
                                                              const uint8_t *__restrict__ pc,
                                                              const double *__restrict__ p,
                                                              const double *__restrict__ u,
                                                              const double *__restrict__ v,
                                                              const double *__restrict__ w,
                                                              const double *__restrict__ q, Res4 R) {
    const int x = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (x >= n) return;
    const int64_t c = nat[x];
    residual_voxel4(g, c, natpos[x], pc[c], p, u, v, w, q, R);
}

__global__ __launch_bounds__(kListBlock) void k_fold4_nat(const int32_t *__restrict__ nat,
                                                          const int32_t *__restrict__ natpos, int n, double *Xp,
                                                          double *Xu, double *Xv, double *Xw, float4 *d) {
    const int x = xcd_block(blockIdx.x, gridDim.x) * kListBlock + threadIdx.x;
    if (x >= n) return;
    fold_compact(natpos[x], nat[x], Xp, Xu, Xv, Xw, d);
}

void launch_residual4(const Grid &g, const FluidLists &L, const uint8_t *pc, const double *p, const double *u,
                      const double *v, const double *w, const double *q, Res4 R, hipStream_t s) {
    if (L.n_all == 0) return;
    if (L.compact && L.nat) {
        hipLaunchKernelGGL(k_residual4_nat, dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, g, L.nat, L.natpos,
                           L.n_all, pc, p, u, v, w, q, R);
        return;
    }
    int threads = 0;
    const ListWalk W = list_walk(L, threads);
    hipLaunchKernelGGL(k_residual4, dim3(lblocks(threads)), dim3(kListBlock), 0, s, g, L.all, L.n_all, W, pc, p, u,
                       v, w, q, R);
}

void launch_ghost4(const Grid &g, const uint8_t *G, float4 *S, float *q, bool split, hipStream_t s) {
    if (split)
        hipLaunchKernelGGL(k_ghost4<true>, dim3(nblocks(ghost_lines(g), 256)), dim3(256), 0, s, g, G, S, q);
    else
        hipLaunchKernelGGL(k_ghost4<false>, dim3(nblocks(ghost_lines(g), 256)), dim3(256), 0, s, g, G, S, q);
}

void launch_ghost4_lists(const Grid &g, const FluidLists &L, const uint8_t *G, float4 *S, float *q, hipStream_t s) {
    if (L.compact_jac) return;  // no halo entries; the outlet pins are the settle pass's
    if (!L.compact) {
        launch_ghost4(g, G, S, q, L.split, s);
        return;
    }
    const int64_t n = (int64_t)L.n_chalo + L.n_cout;
    if (n > 0)
        hipLaunchKernelGGL(k_ghost4_compact, dim3(nblocks(n, 256 * kGhostItems)), dim3(256), 0, s, L.chalo,
                           L.n_chalo, L.cout, L.n_cout, S, q);
}

void launch_fold4(const FluidLists &L, double *Xp, double *Xu, double *Xv, double *Xw, float4 *d, hipStream_t s) {
    if (L.compact && L.nat) {
        if (L.n_all > 0)
            hipLaunchKernelGGL(k_fold4_nat, dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, L.nat, L.natpos, L.n_all,
                               Xp, Xu, Xv, Xw, d);
        return;
    }
    if (L.compact || L.compact_jac) {
        if (L.n_all == 0) return;
        int threads = 0;
        const ListWalk W = list_walk(L, threads);
        hipLaunchKernelGGL(k_fold4_compact, dim3(lblocks(threads)), dim3(kListBlock), 0, s, L.all, L.n_all, W, Xp,
                           Xu, Xv, Xw, d);
        return;
    }
    if (L.n_all == 0) return;
    if (L.split) {
        int threads = 0;
        const ListWalk W = list_walk(L, threads);
        hipLaunchKernelGGL(k_fold4_split, dim3(lblocks(threads)), dim3(kListBlock), 0, s, L.all, L.rec, L.n_all, W,
                           Xp, Xu, Xv, Xw, d);
    }
    else
        hipLaunchKernelGGL(k_fold4, dim3(lblocks(L.n_all)), dim3(kListBlock), 0, s, L.all, L.n_all, Xp, Xu, Xv, Xw,
                           d);
}

void launch_jacobi_stream(int stage, int zs, const Grid &g, const uint8_t *pc, const double *p0, const double *u0,
                          const double *v0, const double *w0, const double *q0, double *p1, double *u1, double *v1,
                          double *w1, double cu, double cp, double pin_out, hipStream_t s) {
    const int ntx = (g.nx + kSX - 1) / kSX, nty = (g.ny + kSY - 1) / kSY;
    zs = std::max(1, zs);
    const int nseg = (g.nz + zs - 1) / zs;
    const dim3 grid((unsigned)(ntx * nty * nseg)), blk(kSN);
    if (stage == JAC_FIRST)
        hipLaunchKernelGGL((k_jac_stream<double, JAC_FIRST>), grid, blk, 0, s, g, ntx, nty, zs, pc, p0, u0, v0, w0,
                           q0, p1, u1, v1, w1, cu, cp, pin_out);
    else
        hipLaunchKernelGGL((k_jac_stream<double, JAC_LAG>), grid, blk, 0, s, g, ntx, nty, zs, pc, p0, u0, v0, w0, q0,
                           p1, u1, v1, w1, cu, cp, pin_out);
}

template <bool SEL>
static void colour_stage(const Grid &g, const int32_t *list, int n, const uint8_t *pc, double *p, double *u,
                         double *v, double *w, double *q, double cu, double cp, double pin_out, hipStream_t s) {
    hipLaunchKernelGGL((k_colour<double, SEL>), dim3(lblocks(n)), dim3(kListBlock), 0, s, g, list, n, pc, p, u, v, w,
                       q, cu, cp, pin_out);
}

void launch_colour(const Grid &g, const FluidLists &L, const uint8_t *pc, double *p, double *u, double *v, double *w,
                   double *q, double cu, double cp, double pin_out, hipStream_t s) {
    for (int col = 0; col < 3; col++) {
        const int n = L.n_col[col];
        if (n == 0) continue;
        if (L.select_loads)
            colour_stage<true>(g, L.col[col], n, pc, p, u, v, w, q, cu, cp, pin_out, s);
        else
            colour_stage<false>(g, L.col[col], n, pc, p, u, v, w, q, cu, cp, pin_out, s);
    }
}

void launch_ghost_fast(const Grid &g, const uint8_t *G, double *p, double *u, double *v, double *w, double *q,
                       double pin_in, double pin_hi, double pin_out, hipStream_t s) {
    hipLaunchKernelGGL(k_ghost_fast<double>, dim3(nblocks(ghost_lines(g), 256)), dim3(256), 0, s, g, G, p, u, v, w,
                       q, pin_in, pin_hi, pin_out);
}

void launch_outlet_plane(const Grid &g, const uint8_t *pc, const double *p, const double *u, const double *v,
                         const double *w, double pin_out, double *q, hipStream_t s) {
    hipLaunchKernelGGL(k_outlet_plane, dim3(nblocks(g.NX, 64), (unsigned)g.NY), dim3(64), 0, s, g, pc, p, u, v,
                       w, pin_out, q);
}

}  // namespace pm
