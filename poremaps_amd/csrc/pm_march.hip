// pm_march.hip -- the lagged Jacobi pass on the interleaved FP32 correction
// state as a z-march through LDS (the FP32+FP64-refined Jacobi sweep).
//
// The list form (lag_voxel4_rec) gathers ~22 float4 neighbours per fluid
// voxel straight from HBM/L2: at porosity 0.2 the 64 fluid voxels of a wave
// lie on ~5 rows, so each stencil load touches ~39 cache lines and the pass is
// bound by the L1's line-access path, not by HBM.  Here a workgroup owns a
// tile of 64 x TY cells (x, y) and a segment of z planes.  It keeps the five
// planes z-2 .. z+2 of the tile and its 2-cell rim in LDS (a ring of planes,
// each loaded once with fully coalesced 16-byte loads, two planes ahead), and
// every fluid cell of plane z reads its stencil, the velocities its lagged
// pressure recomputations need, and the recompute targets from LDS.  The
// packed codes and the continuity residual ride along in LDS rings of their
// own; the momentum residual comes from the fluid list's per-voxel array: a
// cell's list position is the start of its row segment (rowstart, the list is
// ordered in columns of 64 cells along x) plus the number of fluid cells
// before it in the row (one wave per row, a ballot), read one plane ahead.
// Every global load is issued ahead of its use: the planes two steps ahead,
// the residual entries one step ahead.
//
// Same expressions, in the same order, as lag_voxel4_rec<true>, so the same
// values: rim cells of a periodic x or y axis are loaded from their periodic
// images, which the state's halo copies equal (every pass writes them from
// the same value); z-halo planes are read as they are (their pressure is the
// pinned copy, as the list form reads it).
#include "pm_kernels.h"
#include "pm_voxel.h"

#pragma clang fp contract(off)

namespace pm {

constexpr int kMarchThreads = 256;
constexpr int kMarchTX = 64;        // tile width along x = one wave
constexpr int kMarchXcds = 8;

__device__ __forceinline__ int march_xcd_block(int b, int nb) {
    const int xcd = b % kMarchXcds, local = b / kMarchXcds;
    const int q = nb / kMarchXcds, r = nb % kMarchXcds;
    return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

template <int TY>
struct MarchTile {
    static constexpr int RX = kMarchTX + 4, RY = TY + 4, RN = RX * RY;
    static constexpr int LOADS = (RN + kMarchThreads - 1) / kMarchThreads;
    static constexpr int ROWS = TY / 4;  // rows per wave (four waves)
};

// Element offset within a plane of rim-box cell f of the tile whose interior
// origin is (xi0, yi0), or -1 when the cell lies beyond the halo.  x / y cells
// outside the interior of a periodic axis come from their periodic image.
template <int TY>
__device__ __forceinline__ int64_t march_offset(const Grid &g, int f, int xi0, int yi0) {
    using T = MarchTile<TY>;
    if (f >= T::RN) return -1;
    const int ry = f / T::RX, rx = f - ry * T::RX;
    int k = xi0 + rx, j = yi0 + ry;  // padded indices (the rim starts 2 before the tile)
    if (k > g.nx + 3 || j > g.ny + 3) return -1;
    if (g.px) { if (k < 2) k += g.nx; else if (k >= g.nx + 2) k -= g.nx; }
    if (g.py) { if (j < 2) j += g.ny; else if (j >= g.ny + 2) j -= g.ny; }
    return (int64_t)j * g.sx + k;
}

// Ring slots per plane kind.  Step i stores plane i+3 and then (after one
// barrier) reads planes i-2 .. i+2; a wave still on step i-1 reads planes
// i-3 .. i+1 (everyone has passed the barrier of step i-1, so nobody is on
// step i-2).  With seven slots plane i+3's slot is none of i-3 .. i+2.
constexpr int kMarchSlots = 7;

// The lagged pass at the fluid cells of plane i of a tile (its planes i-2 ..
// i+2 in the LDS rings): lag_voxel4_rec's expressions with every operand
// from LDS (or the outlet buffer on a wall) and the momentum residual Rm of
// each own row from registers.
template <int TY>
__device__ __forceinline__ void march_plane(const Grid &g, int i, const float4 (*L)[MarchTile<TY>::RN],
                                            const uint8_t (*Lc)[MarchTile<TY>::RN],
                                            const float (*Lr)[MarchTile<TY>::RN], const float *qxw, float qyw,
                                            float4 *S1, const float4 *Rms, float cu, float cp, int xi0, int yi0,
                                            int wv, int lane, int k, bool kin) {
    using T = MarchTile<TY>;
    constexpr int RX = T::RX, NS = kMarchSlots;
    typedef float v4f __attribute__((ext_vector_type(4)));
    const int64_t sx = g.sx, sy = g.sy;
    const float4 *Lm2 = L[(i + NS - 2) % NS], *Lm1 = L[(i + NS - 1) % NS], *L0 = L[i % NS],
                 *Lp1 = L[(i + 1) % NS], *Lp2 = L[(i + 2) % NS];
    const uint8_t *C0 = Lc[i % NS], *Cm1 = Lc[(i + NS - 1) % NS];
    const float *R0 = Lr[i % NS], *Rm1 = Lr[(i + NS - 1) % NS];
    const bool out = i == g.nz + 1;
#pragma unroll
    for (int rr = 0; rr < T::ROWS; rr++) {
        const int yrow = wv + 4 * rr;                 // tile row
        const int j = yi0 + yrow + 2;
        const int o = (yrow + 2) * RX + lane + 2;     // the cell in the rim box
        const bool in = kin && yi0 + yrow < g.ny;
        const uint8_t pcv = in ? C0[o] : kSolid;
        if (pcv == kSolid) continue;
        const int64_t c = g.at(i, j, k);
        // codes and continuity residuals of the cell and of its recompute
        // targets (the -x / -y rim cell, from the periodic image where the
        // axis wraps, and the -z cell)
        const uint8_t pcx = C0[o - 1], pcy = C0[o - RX], pcz = Cm1[o];
        const float rpc = R0[o], rpx = R0[o - 1], rpy = R0[o - RX], rpz = Rm1[o];
        const float4 Rm = Rms[rr];
        const float4 C = L0[o];
        Star4 s;
        s.XM = L0[o - 1]; s.XP = L0[o + 1]; s.XM2 = L0[o - 2]; s.XP2 = L0[o + 2];
        s.YM = L0[o - RX]; s.YP = L0[o + RX]; s.YM2 = L0[o - 2 * RX]; s.YP2 = L0[o + 2 * RX];
        s.ZM = Lm1[o]; s.ZP = Lp1[o]; s.ZM2 = Lm2[o]; s.ZP2 = Lp2[o];
        // the recompute targets' neighbourhoods (the -x / -y target's +1
        // cell is then the cell's own halo copy, equal to C)
        const float vyx = L0[o - 1 + RX].z, wzx = Lp1[o - 1].w;
        const float uxy = L0[o - RX + 1].y, wzy = Lp1[o - RX].w;
        const float uxz = Lm1[o + 1].y, vyz = Lm1[o + RX].z;
        const float4 NX = s.XM, NY = s.YM;
        const float uxx = C.y, vyy = C.z;

        const Cases cs = decode(pcv);
        const float pin_out = 0.f;
        const float Pc = out ? pin_out : C.x - (div_from4(pcv, C, s.XP.y, s.YP.z, s.ZP.w) + rpc) * cp;
        const float c2dt = (float)(kC2 * kDt);
        const float dx0 = div_from4(pcx, NX, uxx, vyx, wzx);
        const float dy0 = div_from4(pcy, NY, uxy, vyy, wzy);
        const float dz0 = div_from4(pcz, s.ZM, uxz, vyz, C.w);
        float px = out ? pin_out - dx0 * c2dt : NX.x - (dx0 + rpx) * cp;
        float py = out ? pin_out - dy0 * c2dt : NY.x - (dy0 + rpy) * cp;
        const float pz = i == 2 ? s.ZM.x : s.ZM.x - (dz0 + rpz) * cp;
        if (k == 2 && !g.px) px = out ? qxw[rr] : s.XM.x;      // q0[qat(j, k - 1)]
        if (j == 2 && !g.py) py = out ? qyw : s.YM.x;          // q0[qat(j - 1, k)]
        const float dpx = cs.bx ? 0.f : Pc - px;
        const float dpy = cs.by ? 0.f : Pc - py;
        const float dpz = cs.bz ? 0.f : Pc - pz;
        float su, sv, sw;
        visc_from4(cs, s, C, su, sv, sw);
        const float4 ov4 = vel_from4(cs, C, su, sv, sw, dpx, dpy, dpz, Rm, Pc, cu);
        const v4f ov = {ov4.x, ov4.y, ov4.z, ov4.w};
        __builtin_nontemporal_store(ov, reinterpret_cast<v4f *>(&S1[c]));
        // the ghost fill of the new state, cell by cell (lag_voxel4_rec)
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
    }
}

template <int TY>
__global__ __launch_bounds__(kMarchThreads) void k_jac4_march(Grid g, const uint8_t *__restrict__ pc,
                                                              const float4 *__restrict__ S0,
                                                              const float *__restrict__ q0, float4 *S1, Res4 R,
                                                              const int32_t *__restrict__ rowstart, int ntx, int nty,
                                                              int zseg, float cu, float cp) {
    using T = MarchTile<TY>;
    constexpr int RX = T::RX, RN = T::RN, ROWS = T::ROWS, LOADS = T::LOADS;
    // rings of the planes: state, packed codes, continuity residual
    // (kMarchSlots slots each, one barrier per plane)
    constexpr int NS = kMarchSlots;
    __shared__ float4 L[NS][RN];
    __shared__ uint8_t Lc[NS][RN];
    __shared__ float Lr[NS][RN];
    typedef float v3f __attribute__((ext_vector_type(3)));

    const int r = march_xcd_block(blockIdx.x, gridDim.x);
    const int txi = r % ntx, tyi = (r / ntx) % nty, seg = r / (ntx * nty);
    const int xi0 = txi * kMarchTX, yi0 = tyi * TY;
    const int z0 = 2 + seg * zseg, z1 = min(g.nz + 2, z0 + zseg);  // padded planes [z0, z1)
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t sy = g.sy;

    // element offsets within a plane of this thread's rim-box cells (-1: none)
    int32_t off[LOADS];
#pragma unroll
    for (int m = 0; m < LOADS; m++) off[m] = (int32_t)march_offset<TY>(g, tid + m * kMarchThreads, xi0, yi0);
    // a cell beyond the halo loads the plane's first element (a valid
    // address) and keeps zero / solid: value selects, not pointer selects
#define PM_MARCH_LOAD(i, S, Cc, Rr)                                                                           \
    _Pragma("unroll") for (int m = 0; m < LOADS; m++) {                                                       \
        const int64_t e_ = g.base + (int64_t)min((i), g.nz + 3) * sy + max(off[m], 0);                        \
        S[m] = S0[e_];                                                                                        \
        Cc[m] = pc[e_];                                                                                       \
        Rr[m] = R.p[e_];                                                                                      \
    }
#define PM_MARCH_STORE(i, S, Cc, Rr)                                                                          \
    _Pragma("unroll") for (int m = 0; m < LOADS; m++) {                                                       \
        const int f_ = tid + m * kMarchThreads;                                                               \
        const bool ok_ = off[m] >= 0;                                                                         \
        if (f_ < RN) {                                                                                        \
            float4 v_;                                                                                        \
            v_.x = ok_ ? S[m].x : 0.f; v_.y = ok_ ? S[m].y : 0.f;                                             \
            v_.z = ok_ ? S[m].z : 0.f; v_.w = ok_ ? S[m].w : 0.f;                                             \
            L[(i) % NS][f_] = v_;                                                                             \
            Lc[(i) % NS][f_] = ok_ ? Cc[m] : kSolid;                                                          \
            Lr[(i) % NS][f_] = ok_ ? Rr[m] : 0.f;                                                             \
        }                                                                                                     \
    }
    // the row-segment starts of plane i (one per own row)
#define PM_MARCH_ROWSTART(i, RS)                                                                              \
    _Pragma("unroll") for (int rr = 0; rr < ROWS; rr++) {                                                     \
        const int y_ = min(yi0 + wv + 4 * rr, g.ny - 1);                                                      \
        RS[rr] = rowstart[((int64_t)(min((i), z1 - 1) - 2) * g.ny + y_) * ntx + txi];                         \
    }
    // the momentum residual of the own cells of plane i (its codes in LDS):
    // list position = row-segment start + fluid cells before the cell
#define PM_MARCH_RESID(i, RS, RM)                                                                             \
    _Pragma("unroll") for (int rr = 0; rr < ROWS; rr++) {                                                     \
        const int yrow_ = wv + 4 * rr;                                                                        \
        const bool in_ = kin && yi0 + yrow_ < g.ny && (i) < z1;                                               \
        const uint8_t pcv_ = in_ ? Lc[(i) % NS][(yrow_ + 2) * RX + lane + 2] : kSolid;                        \
        const bool fl_ = pcv_ != kSolid;                                                                      \
        const uint64_t mask_ = __ballot(fl_);                                                                 \
        const int t_ = fl_ ? RS[rr] + __popcll(mask_ & ((1ull << lane) - 1ull)) : 0;                          \
        /* (Ru, Rv, Rw) only: a dead fourth lane would let its register be  */                              \
        /* reused at once, and that reuse waits for the load */                                               \
        const v3f rm_ = __builtin_nontemporal_load(reinterpret_cast<const v3f *>(&R.m[t_]));                  \
        RM[rr] = make_float4(rm_.x, rm_.y, rm_.z, 0.f);                                                       \
    }

    // own cells: wave wv takes tile rows wv, wv + 4, ... (lane = x)
    const int k = xi0 + lane + 2;
    const bool kin = xi0 + lane < g.nx;

    // the outlet buffer at the -x / -y wall halo of the own cells (read on
    // the outlet plane where the axis is a wall; the pass never writes it)
    float qxw[ROWS], qyw = 0.f;
#pragma unroll
    for (int rr = 0; rr < ROWS; rr++)
        qxw[rr] = (k == 2 && !g.px && yi0 + wv + 4 * rr < g.ny) ? q0[g.qat(yi0 + wv + 4 * rr + 2, 1)] : 0.f;
    if (yi0 == 0 && !g.py && kin) qyw = q0[g.qat(1, k)];

    float4 SA[LOADS], SB[LOADS];
    uint8_t CA[LOADS], CB[LOADS];
    float RA[LOADS], RB[LOADS];
    int RSA[ROWS], RSB[ROWS];
    float4 RMA[ROWS], RMB[ROWS];
    for (int i = z0 - 2; i <= z0 + 2; i++) {
        PM_MARCH_LOAD(i, SA, CA, RA)
        PM_MARCH_STORE(i, SA, CA, RA)
    }
    PM_MARCH_ROWSTART(z0, RSA)
    PM_MARCH_LOAD(z0 + 3, SB, CB, RB)
    PM_MARCH_ROWSTART(z0 + 1, RSB)
    __syncthreads();
    PM_MARCH_RESID(z0, RSA, RMA)
    PM_MARCH_LOAD(z0 + 4, SA, CA, RA)

    // step i: plane i+3 (loaded two steps ago) into its slot, then plane i
    // computed; the loads of plane i+5, of the residuals of plane i+1 and of
    // the row starts of plane i+2 are in flight meanwhile.  Two register sets
    // alternate (even / odd steps), so no value is copied between steps.
#define PM_MARCH_STEP(i, SCUR, CCUR, RCUR, RSCUR, RSNXT, RMCUR, RMNXT)                                        \
    {                                                                                                         \
        if ((i) + 3 <= z1 + 1) { PM_MARCH_STORE((i) + 3, SCUR, CCUR, RCUR) }                                  \
        __syncthreads();                                                                                      \
        PM_MARCH_LOAD((i) + 5, SCUR, CCUR, RCUR)                                                              \
        PM_MARCH_RESID((i) + 1, RSNXT, RMNXT)                                                                 \
        PM_MARCH_ROWSTART((i) + 2, RSCUR)                                                                     \
        march_plane<TY>(g, i, L, Lc, Lr, qxw, qyw, S1, RMCUR, cu, cp, xi0, yi0, wv, lane, k, kin);            \
    }
    int i = z0;
    for (; i + 1 < z1; i += 2) {
        PM_MARCH_STEP(i, SB, CB, RB, RSA, RSB, RMA, RMB)
        PM_MARCH_STEP(i + 1, SA, CA, RA, RSB, RSA, RMB, RMA)
    }
    if (i < z1) PM_MARCH_STEP(i, SB, CB, RB, RSA, RSB, RMA, RMB)
#undef PM_MARCH_STEP
#undef PM_MARCH_ROWSTART
#undef PM_MARCH_RESID
#undef PM_MARCH_LOAD
#undef PM_MARCH_STORE
}

bool march_supported(const Grid &g, const FluidLists &L) {
    // the list's row segments must be the tiles' rows: columns 64 cells wide,
    // or one column as wide as a domain of at most 64 cells
    return L.rowstart != nullptr &&
           (L.tile_x == kMarchTX || (g.nx <= kMarchTX && (L.tile_x == 0 || L.tile_x >= g.nx)));
}

void launch_jacobi4_march(const Grid &g, const FluidLists &L, const uint8_t *pc, const float4 *S0, const float *q0,
                          float4 *S1, Res4 R, float cu, float cp, int ty, int zseg, hipStream_t s) {
    if (L.n_all == 0) return;
    const int ntx = (g.nx + kMarchTX - 1) / kMarchTX;
    const int zs = zseg > 0 ? zseg : g.nz;
    const int nseg = (g.nz + zs - 1) / zs;
    if (ty == 8) {
        const int nty = (g.ny + 7) / 8;
        hipLaunchKernelGGL(k_jac4_march<8>, dim3((unsigned)(ntx * nty * nseg)), dim3(kMarchThreads), 0, s, g, pc,
                           S0, q0, S1, R, L.rowstart, ntx, nty, zs, cu, cp);
    } else {
        const int nty = (g.ny + 3) / 4;
        hipLaunchKernelGGL(k_jac4_march<4>, dim3((unsigned)(ntx * nty * nseg)), dim3(kMarchThreads), 0, s, g, pc,
                           S0, q0, S1, R, L.rowstart, ntx, nty, zs, cu, cp);
    }
}

}  // namespace pm
