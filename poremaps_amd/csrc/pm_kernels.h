// pm_kernels.h -- launchers of the HIP kernels (pm_kernels.hip).
#pragma once

#include "pm_grid.h"

#include <vector>

namespace pm {

// dom_interest box in 0-based interior coordinates (x min/max, y, z).
struct DomBox {
    int on;
    int d[6];
};

constexpr int kRedThreads = 256;
// Most workgroups of a state reduction (one partial sum each): 32 waves per
// CU, enough loads in flight.
constexpr int kRedBlocksMax = 8192;

// Results of a state reduction: sum[0..2] fluid u, v, w; sum[3..5] box fluid
// u, v, w; sum[6], sum[7] p on the box's zmin / zmax planes; count[0..3] the
// matching fluid counts; wmax = max(0, max fluid w).
struct RedSums {
    double sum[8];
    long long count[4];
    double wmax;
};

void launch_geom_fill(const Grid &g, const uint8_t *raw, uint8_t *G, int walls, hipStream_t s);
template <typename T>
void launch_exchange_full(const Grid &g, T *f0, T *f1, T *f2, T *f3, hipStream_t s);
void launch_eval_geometry(const Grid &g, const uint8_t *G, int32_t *code, uint8_t *pc, hipStream_t s);
void launch_init_fields(const Grid &g, const uint8_t *G, double *p, double *u, double *v, double *w,
                        hipStream_t s);
template <typename T>
void launch_pins(const Grid &g, const uint8_t *G, T *p, T pin_in, T pin_hi, T pin_out, hipStream_t s);

// Reference order: `nsweeps` complete sweeps (cells in lexicographic order,
// exchange, pins), as a wavefront of LDS tiles (REF_TILES), in one workgroup
// (REF_ONE_WG) or one launch per hyperplane (REF_PLANES); the same bits.
enum { REF_TILES = 0, REF_ONE_WG = 1, REF_PLANES = 2 };
// The four FP64 fields p, u, v, w (kernel argument).
struct RefFields {
    double *f[4];
};
// A block of a domain decomposition in interior coordinates: first cell and
// extent along x, y, z.
struct RefBox {
    int lo[3], n[3];
};
// The blocks of dom_decomposition dims (x, y, z) in the reference's rank
// order (src/geometry.cc:129-171; z fastest).
void ref_blocks(const Grid &g, const int dims[3], std::vector<RefBox> *out);
// One exact-order sweep the way `mpirun -np P*Q*R` runs it with
// dom_decomposition P Q R: every block swept (tiled wavefront) against the
// state as the sweep started (snap, 4 arrays of Grid::total, overwritten),
// whose outlet plane is taken from `outlet` (the plane z = nz+1 as the last
// sweep left it before its pins, Grid::qtotal elements, updated); then the
// exchange and the pins.  tab: nblk blocks in device memory; ntile: the most
// tiles of any block along x, y, z.
void launch_ref_block_sweep(const Grid &g, const uint8_t *G, const int32_t *code, double *const f[4],
                            double *const snap[4], double *outlet, const RefBox *tab, int nblk, const int ntile[3],
                            int solver, hipStream_t s);
void launch_ref_sweeps(const Grid &g, const uint8_t *G, const int32_t *code, double *p, double *u,
                       double *v, double *w, int solver, int nsweeps, int impl, hipStream_t s);

// Links of the compact SOR state (FluidLists::compact): three arrays of
// n_centries uint2 back to back, (-x, +x), (-y, +y), (-z, +z) entry by
// entry: the entries of the six +-1 neighbours (the zero entry where the
// neighbour is solid, a wall halo or beyond the halo) in the low 26 bits.
// The top 6 bits of the -x and +x links carry the interior entries' colour
// pass record: packed code (8 bits), block k mod 3 (2 bits), outlet flag.
constexpr uint32_t kLinkMask = 0x03FFFFFFu;
constexpr int kLinkBits = 26;

// Device lists of the interior fluid cells' element offsets in column order
// (columns of tile_y rows x tile_x cells, each marching along z; within a
// plane of a column ascending y, x): all of them, and split by colour
// (i+j+k) mod 3.  For 3-colour SOR on the interleaved state the colour lists
// lie back to back in `all` (colour q from off_col[q]) and col[q] point
// into it (col_in_all).
struct FluidLists {
    int32_t *all = nullptr;
    int n_all = 0;
    int32_t *col[3] = {nullptr, nullptr, nullptr};
    int n_col[3] = {0, 0, 0};
    int off_col[3] = {0, 0, 0};
    bool col_in_all = false;
    // stencil loads: every neighbour up front + select (true) or only the
    // neighbours the case needs, behind branches (false)
    bool select_loads = true;
    // column of the list order, rows x cells (0: whole extent)
    int tile_y = 0, tile_x = 0;
    // per-voxel records of the lagged pass (element offset + packed codes of
    // the cell and its recompute targets), FP32+DC Jacobi; use_rec selects
    // the record form of the lagged pass
    uint2 *rec = nullptr;
    bool use_rec = false;
    // the FP32+DC record pass writes the periodic halo copies and halo pins of
    // its cells itself (no ghost fill after it)
    bool fill_halos = true;
    // FP32+DC 3-colour SOR: the correction state in colour-split rows
    // (Grid::at_s) and rec holding the colour passes' records (colour-split
    // offset + packed code, block, outlet flag)
    bool split = false;
    // FP32+DC Jacobi: the list position of the first fluid cell of every row
    // segment (z, y, x column of tile_x cells), nz * ny * ceil(nx / tile_x)
    // entries, for the z-march form of the lagged pass (tile_x 64 only);
    // use_march selects that form, with tiles of 64 x march_ty cells and z
    // segments of march_zseg planes (0: all)
    int32_t *rowstart = nullptr;
    bool use_march = false;
    int march_ty = 4, march_zseg = 128;
    // FP32+DC 3-colour SOR in z-chunks of sor_chunk planes (0: three whole
    // passes): colour 1 trails colour 0 by 3 planes and colour 2 trails it by
    // 6, so a chunk is re-read by the next colour while it is still cached.
    // The colour lists are then in (z, y, x) order and plane_off[q * (nz + 1)
    // + i] is the position, within colour list q, of its first voxel on plane
    // i.  sor_fused: the three colours of a stage (chunk k, k - 1, k - 2) in
    // one launch; else one launch per colour and chunk.
    int sor_chunk = 0;
    bool sor_fused = true;
    // FP32+DC record pass: waves per SIMD its registers must allow (0, 6, 8;
    // 0: the compiler's allocation)
    int rec_occ = 0;
    // FP32+DC record pass: the +-2 stencil loads only where the cell's case
    // reads them (rec_skip2_supported: the state below 2^32 bytes)
    bool rec_skip2 = false;
    // with rec_skip2: the +-1 loads and the recompute targets' loads too,
    // each only where it is used
    bool rec_skip1 = false;
    // FP32+DC SOR (colour-split rows): the same for the colour passes
    bool sor_skip2 = true;
    std::vector<int32_t> plane_off;
    // FP32+DC 3-colour SOR on the compact state: the correction state holds
    // one float4 per fluid cell in list order (entry t = list position t),
    // then one per fluid cell of a periodic face halo (a copy of its image),
    // then one zero entry that stands for every solid, wall-halo or otherwise
    // unused neighbour; the outlet buffer likewise holds one float per entry.
    // clink: the entries' links and records (kLinkMask); chalo: (halo entry,
    // image entry | bit 31 when the copy's p is 0 and its outlet buffer value
    // is copied too); cout: the entries of the outlet plane (p pinned to 0
    // after a sweep).
    bool compact = false;
    // FP32+DC Jacobi on the compact state: one entry per fluid cell in list
    // order and the zero entry, links across periodic faces straight to the
    // image (no halo entries; the flags hold the codes of the cell and of its
    // recompute targets, the outlet and first-plane flags)
    bool compact_jac = false;
    // compact Jacobi: the links as one record of six words per entry (+x, +y,
    // +z, -x, -y, -z) instead of three arrays of (-, +) pairs
    bool links_aos = false;
    uint2 *clink = nullptr;
    int n_centries = 0;
    int2 *chalo = nullptr;
    int n_chalo = 0;
    int32_t *cout = nullptr;
    int n_cout = 0;
    // compact SOR: the fluid list in column order (natural offsets) and each
    // cell's position in the colour lists, for the residual and the fold
    int32_t *nat = nullptr, *natpos = nullptr;
    // compact SOR, packed links (launch_pack_links): 12 B per interior entry
    // (compact Jacobi: launch_pack_jac_links, 24 B per entry, 36 bases per
    // wave, cpack / cbase / packed used the same way)
    // instead of 24 -- per link 2 bits naming one of three bases of the
    // entry's wave and slot (3: the zero entry) and a 12-bit offset, the
    // record flags and a flag for entries whose links fit no base (read from
    // clink); cbase: 18 bases (3 x the six slots) per wave of 64 entries of a
    // colour, wave_off[col]: the colour's first wave
    bool packed = false;
    uint32_t *cpack = nullptr;
    int32_t *cbase = nullptr;
    int wave_off[3] = {0, 0, 0};
};

// FP64 fast modes (per-field state p, u, v, w): the velocity + pressure
// kernel pair of one symplectic Jacobi sweep from set 0 into set 1.
void launch_jacobi(const Grid &g, const FluidLists &L, const uint8_t *pc, const double *p0, const double *u0,
                   const double *v0, const double *w0, const double *q0, double *p1, double *u1, double *v1,
                   double *w1, double *q1, double cu, double cp, double pin_out, hipStream_t s);
// Symplectic Jacobi in lagged form (pm_voxel.h): JAC_FIRST maps a settled
// set 0 (u^n, p^n) to set 1 = (u^{n+1}, p^n); JAC_LAG maps set 0 =
// (u^n, p^{n-1}) to set 1 = (u^{n+1}, p^n); JAC_SETTLE completes set 0 in
// place: p^n and the outlet buffer from set 0's velocities (set 1 pointers:
// p1 = set 0's p, q1 = set 0's q).
enum { JAC_FIRST = 0, JAC_LAG = 1, JAC_SETTLE = 2, JAC_PAIR = 3, JAC_FIRST_ZERO = 4 };

// Interleaved FP32 correction state (one float4 {p, u, v, w} per cell) of
// the defect-correction modes.  Jacobi: stage JAC_FIRST / JAC_LAG / JAC_PAIR
// (JAC_FIRST_ZERO: JAC_FIRST from a state that is zero everywhere, halos and
// outlet buffer included, computed from the records and the residual alone)
// (velocity + pressure kernels, S1 complete) from S0 into S1; JAC_SETTLE
// completes S0 in place (pass S1 = S0, q1 = its outlet buffer).  3-colour
// SOR: the three colour launches, in place.
void launch_jacobi4(int stage, const Grid &g, const FluidLists &L, const uint8_t *pc, const float4 *S0,
                    const float *q0, float4 *S1, float *q1, Res4 R, float cu, float cp, hipStream_t s);
// Whether launch_jacobi4 of this stage leaves S1's halos filled (the record
// pass with L.fill_halos): the caller then skips launch_ghost4.
bool jacobi4_fills_halos(int stage, const FluidLists &L);
bool rec_skip2_supported(const Grid &g);
void launch_colour4(const Grid &g, const FluidLists &L, const uint8_t *pc, float4 *S, float *q, Res4 R, float cu,
                    float cp, hipStream_t s);
// The lagged pass (JAC_LAG, halo copies included) as a z-march of 64 x ty
// tiles through LDS (pm_march.hip); needs L.rowstart, L.rec and R.q.
bool march_supported(const Grid &g, const FluidLists &L);
void launch_jacobi4_march(const Grid &g, const FluidLists &L, const uint8_t *pc, const float4 *S0, const float *q0,
                          float4 *S1, Res4 R, float cu, float cp, int ty, int zseg, hipStream_t s);
// The per-voxel records L.rec of the list: the lagged pass's, or the colour
// passes' when L.split.
void launch_lag_records(const Grid &g, const FluidLists &L, const uint8_t *pc, hipStream_t s);
void launch_gather_rq(const Grid &g, const FluidLists &L, Res4 R, hipStream_t s);
void launch_residual4(const Grid &g, const FluidLists &L, const uint8_t *pc, const double *p, const double *u,
                      const double *v, const double *w, const double *q, Res4 R, hipStream_t s);
void launch_ghost4(const Grid &g, const uint8_t *G, float4 *S, float *q, bool split, hipStream_t s);
// The ghost fill of the correction state in the lists' layout: the compact
// state's halo copies and outlet pins (and the outlet buffer's periodic
// copies), else launch_ghost4 (natural or colour-split rows).
void launch_ghost4_lists(const Grid &g, const FluidLists &L, const uint8_t *G, float4 *S, float *q, hipStream_t s);
// The compact Jacobi state's links (no halo entries; links across periodic
// faces to the image) built on the device from the fluid list; cidx is
// scratch of Grid::total entries.  Returns the error of clearing cidx.
hipError_t launch_torus_links(const Grid &g, const int32_t *list, int n, int32_t *cidx, uint2 *L, int nE, bool aos,
                              hipStream_t s);
// The compact states' +-2 links folded into the slots of the solid sides
// (cases 4 and 5), from L0 into L1 (both of L.n_centries entries), after the
// records' flags are in the links.
// Packed links of the compact SOR's interior entries from clink (final links
// and flags) into L.cpack (3 n_all words) and L.cbase (18 per wave), both
// allocated and L.wave_off set by the caller; adds the number of entries left
// on the full links to the device counter *d_far, then decodes every record
// and adds the entries whose links or flags differ from clink's to *d_bad.
// span: the widest offset from a base (4095, the most the record holds;
// smaller only in tests).
void launch_pack_links(const FluidLists &L, int span, int *d_far, int *d_bad, hipStream_t s);
// Packed records of the compact Jacobi's entries (24 B each, the twelve
// links of the lagged pass and the flags) from clink into L.cpack (6 n_all
// words) and L.cbase (36 per wave of 64), both allocated by the caller;
// counters as launch_pack_links.  span <= 1023.
void launch_pack_jac_links(const FluidLists &L, int span, int *d_far, int *d_bad, hipStream_t s);
void launch_link_subst(const FluidLists &L, const uint2 *L0, uint2 *L1, hipStream_t s);
// Whether the compact SOR state can hold this medium: no slip walls, and
// every entry's byte offset below 2^32 (its masked loads are buffer loads).
bool compact_supported(const Grid &g, int64_t n_entries);
// X += d, d = 0 over the fluid list; d's halo copies need a ghost fill after.
void launch_fold4(const FluidLists &L, double *Xp, double *Xu, double *Xv, double *Xw, float4 *d, hipStream_t s);
void launch_jacobi_stage(int stage, const Grid &g, const FluidLists &L, const uint8_t *pc, const double *p0,
                         const double *u0, const double *v0, const double *w0, const double *q0, double *p1,
                         double *u1, double *v1, double *w1, double *q1, double cu, double cp, double pin_out,
                         hipStream_t s);
// The JAC_FIRST / JAC_LAG passes as the dense 2.5-D kernel: 64 x 8 tiles
// marching along z in segments of zs planes (pm_sweep.hip).
void launch_jacobi_stream(int stage, int zs, const Grid &g, const uint8_t *pc, const double *p0, const double *u0,
                          const double *v0, const double *w0, const double *q0, double *p1, double *u1, double *v1,
                          double *w1, double cu, double cp, double pin_out, hipStream_t s);
// 3-colour SOR, FP64 per-field state, in place: the three colour launches.
void launch_colour(const Grid &g, const FluidLists &L, const uint8_t *pc, double *p, double *u, double *v, double *w,
                   double *q, double cu, double cp, double pin_out, hipStream_t s);
void launch_ghost_fast(const Grid &g, const uint8_t *G, double *p, double *u, double *v, double *w, double *q,
                       double pin_in, double pin_hi, double pin_out, hipStream_t s);

void launch_reduce(const Grid &g, const uint8_t *G, const double *p, const double *u, const double *v,
                   const double *w, const DomBox &box, bool full, void *scratch, RedSums *out,
                   hipStream_t s);
// scratch: kRedBlocksMax block partials (reduce_scratch_bytes()).
size_t reduce_scratch_bytes();

void launch_outlet_plane(const Grid &g, const uint8_t *pc, const double *p, const double *u,
                         const double *v, const double *w, double pin_out, double *q, hipStream_t s);

}  // namespace pm
