/*
 * pm_oracle.c -- CPU restatement of the POREMAPS Stokes sweep (checker only).
 *
 * TEST INFRASTRUCTURE.  Nothing in poremaps_amd/ links or calls this file;
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it as
 * the reference answer for the HIP path.
 *
 * Compile with -ffp-contract=off and without -march so that every double
 * operation is one IEEE operation in the order written, like the reference's
 * g++ -O3 x86-64 build (SSE2, no FMA).
 *
 * Reference-faithful functions cite the file:line of the reference they
 * restate.  The fast-mode functions (Jacobi, 3-colour SOR, defect correction)
 * have no reference code: they define the semantics the GPU fast kernels must
 * reproduce bit for bit.
 */
#include "pm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/constants.h:51-64 */
#define ORC_RE 0.01
#define ORC_C2 1.5e6
#define ORC_DT 1.0e-4
#define ORC_OMEGA_SOR 1.2

#define HX(nx) ((nx) + 4)
#define HY(ny) ((ny) + 4)
#define HZ(nz) ((nz) + 4)

static inline size_t ix(int NX, int NY, int i, int j, int k)
{
    return ((size_t)i * NY + j) * NX + k;
}

static void periodic_axes(int bc, int *px, int *py, int *pz)
{
    /* src/main_stokes_solver.cc:163-165: 0 all periodic, 1-2 z only, 3-4 none */
    *px = (bc == 0);
    *py = (bc == 0);
    *pz = (bc == 0 || bc == 1 || bc == 2);
}

/* Two-layer periodic halo copy, X then Y then Z, every axis over the full
 * extent of the other two (src/parallelization.cc:67-190 and :193-318 at one
 * rank: MPI_Sendrecv to itself; nothing on axes without a neighbour). */
#define DEFINE_EXCHANGE(NAME, T)                                                  \
    static void NAME(T *f, int nx, int ny, int nz, int px, int py, int pz)      \
    {                                                                            \
        const int NX = HX(nx), NY = HY(ny), NZ = HZ(nz);                         \
        if (px)                                                                  \
            for (int i = 0; i < NZ; i++)                                         \
                for (int j = 0; j < NY; j++) {                                   \
                    T *r = f + ix(NX, NY, i, j, 0);                              \
                    T a0 = r[2], a1 = r[3], b0 = r[NX - 4], b1 = r[NX - 3];      \
                    r[NX - 2] = a0; r[NX - 1] = a1; r[0] = b0; r[1] = b1;        \
                }                                                                \
        if (py)                                                                  \
            for (int i = 0; i < NZ; i++)                                         \
                for (int k = 0; k < NX; k++) {                                   \
                    T a0 = f[ix(NX, NY, i, 2, k)], a1 = f[ix(NX, NY, i, 3, k)];  \
                    T b0 = f[ix(NX, NY, i, NY - 4, k)];                          \
                    T b1 = f[ix(NX, NY, i, NY - 3, k)];                          \
                    f[ix(NX, NY, i, NY - 2, k)] = a0;                            \
                    f[ix(NX, NY, i, NY - 1, k)] = a1;                            \
                    f[ix(NX, NY, i, 0, k)] = b0;                                 \
                    f[ix(NX, NY, i, 1, k)] = b1;                                 \
                }                                                                \
        if (pz)                                                                  \
            for (int j = 0; j < NY; j++)                                         \
                for (int k = 0; k < NX; k++) {                                   \
                    T a0 = f[ix(NX, NY, 2, j, k)], a1 = f[ix(NX, NY, 3, j, k)];  \
                    T b0 = f[ix(NX, NY, NZ - 4, j, k)];                          \
                    T b1 = f[ix(NX, NY, NZ - 3, j, k)];                          \
                    f[ix(NX, NY, NZ - 2, j, k)] = a0;                            \
                    f[ix(NX, NY, NZ - 1, j, k)] = a1;                            \
                    f[ix(NX, NY, 0, j, k)] = b0;                                 \
                    f[ix(NX, NY, 1, j, k)] = b1;                                 \
                }                                                                \
    }

DEFINE_EXCHANGE(exchange_u8, uint8_t)
DEFINE_EXCHANGE(exchange_f64, double)
DEFINE_EXCHANGE(exchange_f32, float)

/* src/geometry.cc:58-126: interior copied in with halo offset 2, every other
 * cell 0 (fluid), then the geometry halo exchange. */
void orc_geom_halo(const uint8_t *raw, int nx, int ny, int nz, int bc, uint8_t *G)
{
    orc_geom_halo_walls(raw, nx, ny, nz, bc, 0, G);
}

/* The build's opt-in physical walls (no reference code): walls = 1 makes the
 * halo layers of the non-periodic x and y faces solid, so the domain faces
 * are no-slip walls handled by the same one-sided stencils as interior solid
 * voxels (the reference leaves them fluid, src/geometry.cc:98-106).  walls = 0
 * is orc_geom_halo; walls = 2 (slip walls, orc_set_slip: symmetry planes)
 * fills those halo layers with the mirror image of the first two interior
 * layers (x first, over every row and plane, then y). */
void orc_geom_halo_walls(const uint8_t *raw, int nx, int ny, int nz, int bc, int walls, uint8_t *G)
{
    const int NX = HX(nx), NY = HY(ny), NZ = HZ(nz);
    int px, py, pz;
    periodic_axes(bc, &px, &py, &pz);
    memset(G, 0, (size_t)NX * NY * NZ);
    if (walls == 1)
        for (int i = 0; i < NZ; i++)
            for (int j = 0; j < NY; j++)
                for (int k = 0; k < NX; k++) {
                    const int hx = k < 2 || k >= NX - 2, hy = j < 2 || j >= NY - 2;
                    if ((hx && !px) || (hy && !py)) G[ix(NX, NY, i, j, k)] = 1;
                }
    size_t c = 0;
    for (int i = 0; i < nz; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nx; k++)
                G[ix(NX, NY, i + 2, j + 2, k + 2)] = raw[c++];
    if (walls == 2) {
        if (!px)
            for (int i = 0; i < NZ; i++)
                for (int j = 0; j < NY; j++)
                    for (int e = 0; e < 2; e++) {
                        G[ix(NX, NY, i, j, 1 - e)] = G[ix(NX, NY, i, j, 2 + e)];
                        G[ix(NX, NY, i, j, NX - 2 + e)] = G[ix(NX, NY, i, j, NX - 3 - e)];
                    }
        if (!py)
            for (int i = 0; i < NZ; i++)
                for (int k = 0; k < NX; k++)
                    for (int e = 0; e < 2; e++) {
                        G[ix(NX, NY, i, 1 - e, k)] = G[ix(NX, NY, i, 2 + e, k)];
                        G[ix(NX, NY, i, NY - 2 + e, k)] = G[ix(NX, NY, i, NY - 3 - e, k)];
                    }
    }
    exchange_u8(G, nx, ny, nz, px, py, pz);
}

/* src/geometry.cc:174-302 (eval_geometry). */
void orc_eval_geometry(const uint8_t *G, int nx, int ny, int nz, int32_t *code)
{
    const int NX = HX(nx), NY = HY(ny), NZ = HZ(nz);
    const size_t n = (size_t)NX * NY * NZ;
    for (size_t c = 0; c < n; c++) code[c] = G[c] == 0 ? 0 : -1;
    for (int dir = 0; dir < 3; dir++) {
        const int di = dir == 2, dj = dir == 1, dk = dir == 0;
        const int shift = dir == 0 ? 100 : (dir == 1 ? 10 : 1);
        for (int i = 2; i < NZ - 2; i++)
            for (int j = 2; j < NY - 2; j++)
                for (int k = 2; k < NX - 2; k++) {
                    if (G[ix(NX, NY, i, j, k)] != 0) continue;
                    int l = G[ix(NX, NY, i - di, j - dj, k - dk)] == 0 ? 0 : 1;
                    int r = G[ix(NX, NY, i + di, j + dj, k + dk)] == 0 ? 0 : 1;
                    int ll = -1, rr = -1;
                    if (l == 0) ll = G[ix(NX, NY, i - 2 * di, j - 2 * dj, k - 2 * dk)] == 0 ? 0 : 1;
                    if (r == 0) rr = G[ix(NX, NY, i + 2 * di, j + 2 * dj, k + 2 * dk)] == 0 ? 0 : 1;
                    int cs = 0;
                    if (l == 1 && r == 1) cs = 1;
                    else if (l == 1 && r == 0 && rr == 1) cs = 2;
                    else if (ll == 1 && l == 0 && r == 1) cs = 3;
                    else if (l == 1 && r == 0 && rr == 0) cs = 4;
                    else if (ll == 0 && l == 0 && r == 1) cs = 5;
                    else if (l == 0 && r == 0) cs = 6;
                    code[ix(NX, NY, i, j, k)] += cs * shift;
                }
    }
}

/* src/geometry.cc:335-364 (get_porosity at one rank). */
double orc_porosity(const uint8_t *G, int nx, int ny, int nz)
{
    const int NX = HX(nx), NY = HY(ny);
    double fluid = 0.0;
    double vox = (double)(nx * ny * nz);
    for (int i = 2; i < nz + 2; i++)
        for (int j = 2; j < ny + 2; j++)
            for (int k = 2; k < nx + 2; k++)
                if (G[ix(NX, NY, i, j, k)] == 0) fluid += 1.0;
    return fluid / vox;
}

/* src/initialize.cc:69-121 (allocate_initial_field at one rank). */
void orc_init_fields(const uint8_t *G, int nx, int ny, int nz, int bc,
                     double *p, double *u, double *v, double *w)
{
    const int NX = HX(nx), NY = HY(ny), NZ = HZ(nz);
    const size_t n = (size_t)NX * NY * NZ;
    int px, py, pz;
    for (size_t c = 0; c < n; c++) { u[c] = 0.0; v[c] = 0.0; w[c] = 0.0; p[c] = -1.0; }
    for (int i = 2; i < NZ - 2; i++)
        for (int j = 2; j < NY - 2; j++)
            for (int k = 2; k < NX - 2; k++) {
                size_t c = ix(NX, NY, i, j, k);
                p[c] = G[c] == 0 ? (double)((nz - 0 - i) + 2 + 2) : 0.0;
            }
    periodic_axes(bc, &px, &py, &pz);
    exchange_f64(p, nx, ny, nz, px, py, pz);
}

void orc_exchange_f64(double *f, int nx, int ny, int nz, int bc)
{
    int px, py, pz;
    periodic_axes(bc, &px, &py, &pz);
    exchange_f64(f, nx, ny, nz, px, py, pz);
}

/* src/boundary_conditions.cc:46-78 at one rank (inlet and outlet rank). */
#define DEFINE_PINS(NAME, T)                                                     \
    static void NAME(const uint8_t *G, int nx, int ny, int nz, T *p,            \
                     T pin_in, T pin_hi, T pin_out)                              \
    {                                                                            \
        const int NX = HX(nx), NY = HY(ny), NZ = HZ(nz);                         \
        for (int j = 2; j < NY - 2; j++)                                         \
            for (int k = 2; k < NX - 2; k++)                                     \
                if (G[ix(NX, NY, 1, j, k)] == 0) p[ix(NX, NY, 1, j, k)] = pin_in; \
        for (int j = 2; j < NY - 2; j++)                                         \
            for (int k = 2; k < NX - 2; k++) {                                   \
                if (G[ix(NX, NY, NZ - 2, j, k)] == 0) p[ix(NX, NY, NZ - 2, j, k)] = pin_hi; \
                if (G[ix(NX, NY, NZ - 3, j, k)] == 0) p[ix(NX, NY, NZ - 3, j, k)] = pin_out; \
            }                                                                    \
    }
DEFINE_PINS(pins_f64, double)
DEFINE_PINS(pins_f32, float)

void orc_pins(const uint8_t *G, int nx, int ny, int nz, double *p)
{
    /* size[2] + 2 is an int converted to double (src/boundary_conditions.cc:61) */
    pins_f64(G, nx, ny, nz, p, (double)(nz + 2), 2.0, 3.0);
}

/* src/solver.cc:57-91 (get_2nd_derivation).  A case outside 1..6 leaves the
 * reference's output unassigned; fluid interior cells never have one. */
static inline double d2_ref(int cs, double ll, double l, double c, double r, double rr)
{
    switch (cs) {
    case 1: return -8.0 * c;
    case 2: return 8.0 / 3.0 * r - 16.0 / 3.0 * c;
    case 3: return 8.0 / 3.0 * l - 16.0 / 3.0 * c;
    case 4: return 2.0 * r - 5.0 * c - 1.0 / 5.0 * rr;
    case 5: return 2.0 * l - 1.0 / 5.0 * ll - 5.0 * c;
    case 6: return l - 2.0 * c + r;
    default: return 0.0;
    }
}

/* src/solver.cc:112-223 (central_diff5): in place, lexicographic z, y, x. */
void orc_sweep_ref(int solver, const uint8_t *G, const int32_t *code,
                   int nx, int ny, int nz,
                   double *p, double *u, double *v, double *w)
{
    const int NX = HX(nx), NY = HY(ny), NZ = HZ(nz);
    const size_t sj = NX, si = (size_t)NX * NY;
    double vs = 0.0;
    for (int i = 2; i < NZ - 2; i++)
        for (int j = 2; j < NY - 2; j++)
            for (int k = 2; k < NX - 2; k++) {
                const size_t c = ix(NX, NY, i, j, k);
                if (G[c] != 0) continue;
                double dpz = G[c - si] == 0 ? -p[c - si] + p[c] : 0.;
                double dpy = G[c - sj] == 0 ? -p[c - sj] + p[c] : 0.;
                double dpx = G[c - 1] == 0 ? -p[c - 1] + p[c] : 0.;
                int num = code[c];
                int cz = num % 10; num /= 10;
                int cy = num % 10; num /= 10;
                int cx = num % 10;
                double d2vxdx2 = (u[c - 1] - 2 * u[c] + u[c + 1]);
                double d2vydy2 = (v[c - sj] - 2 * v[c] + v[c + sj]);
                double d2vzdz2 = (w[c - si] - 2 * w[c] + w[c + si]);
                double d2vydx2 = d2_ref(cx, v[c - 2], v[c - 1], v[c], v[c + 1], v[c + 2]);
                double d2vzdx2 = d2_ref(cx, w[c - 2], w[c - 1], w[c], w[c + 1], w[c + 2]);
                double d2vxdy2 = d2_ref(cy, u[c - 2 * sj], u[c - sj], u[c], u[c + sj], u[c + 2 * sj]);
                double d2vzdy2 = d2_ref(cy, w[c - 2 * sj], w[c - sj], w[c], w[c + sj], w[c + 2 * sj]);
                double d2vxdz2 = d2_ref(cz, u[c - 2 * si], u[c - si], u[c], u[c + si], u[c + 2 * si]);
                double d2vydz2 = d2_ref(cz, v[c - 2 * si], v[c - si], v[c], v[c + si], v[c + 2 * si]);
                double sx = d2vxdx2 + d2vxdy2 + d2vxdz2;
                double sy = d2vydy2 + d2vydz2 + d2vydx2;
                double sz = d2vzdz2 + d2vzdx2 + d2vzdy2;
                if (solver == 1)
                    vs = (w[c + si] - w[c] + v[c + sj] - v[c] + u[c + 1] - u[c]);
                if (solver == 1 || solver == 2) {
                    if (G[c - 1] == 1) u[c] = 0.0;
                    else u[c] = u[c] + (-dpx + sx / ORC_RE) * ORC_DT;
                    if (G[c - sj] == 1) v[c] = 0.0;
                    else v[c] = v[c] + (-dpy + sy / ORC_RE) * ORC_DT;
                    if (G[c - si] == 1) w[c] = 0.0;
                    else w[c] = w[c] + (-dpz + sz / ORC_RE) * ORC_DT;
                } else if (solver == 3) {
                    u[c] = (1.0 - ORC_OMEGA_SOR) * u[c] + ORC_OMEGA_SOR * (-dpx + sx / ORC_RE) * ORC_DT;
                    v[c] = (1.0 - ORC_OMEGA_SOR) * v[c] + ORC_OMEGA_SOR * (-dpy + sy / ORC_RE) * ORC_DT;
                    w[c] = (1.0 - ORC_OMEGA_SOR) * w[c] + ORC_OMEGA_SOR * (-dpz + sz / ORC_RE) * ORC_DT;
                }
                if (solver == 2 || solver == 3)
                    vs = (w[c + si] - w[c] + v[c + sj] - v[c] + u[c + 1] - u[c]);
                p[c] = p[c] - ORC_C2 * vs * ORC_DT;
            }
}

/* src/evaluation.cc:216-262 (compute_convergence at one rank). */
double orc_convergence(const uint8_t *G, const double *w, int nx, int ny, int nz,
                       double *kprev)
{
    const int NX = HX(nx), NY = HY(ny);
    double sum = 0.0;
    int cnt = 0;
    for (int i = 2; i < nz + 2; i++)
        for (int j = 2; j < ny + 2; j++)
            for (int k = 2; k < nx + 2; k++) {
                size_t c = ix(NX, NY, i, j, k);
                if (G[c] == 0) { cnt += 1; sum += w[c]; }
            }
    if (cnt == 0) sum = 0.0;
    double cbuf0 = sum, cbuf1 = (double)cnt;
    cnt = (int)cbuf1;
    double mean = cbuf0 / cnt;
    double kcur = mean / ORC_RE;
    double conv = fabs(kcur - *kprev) / kcur;
    *kprev = kcur;
    return conv;
}

/* src/evaluation.cc:54-211 (compute_permeability at one rank, starts = 0). */
void orc_permeability(const uint8_t *G, const double *u, const double *v,
                      const double *w, const double *p, int nx, int ny, int nz,
                      const int *dom, double dx, double phi, double *out)
{
    const int NX = HX(nx), NY = HY(ny);
    double sx = 0, sy = 0, sz = 0, wsx = 0, wsy = 0, wsz = 0, wmax = 0;
    double p0 = 0, pend = 0, pgrad = 1.0;
    int p0n = 0, pendn = 0, fl = 0, wfl = 0;
    int dofi = 1;
    for (int i = 2; i < nz + 2; i++)
        for (int j = 2; j < ny + 2; j++)
            for (int k = 2; k < nx + 2; k++) {
                size_t c = ix(NX, NY, i, j, k);
                if (G[c] == 0) {
                    wfl += 1; wsx += u[c]; wsy += v[c]; wsz += w[c];
                    if (w[c] > wmax) wmax = w[c];
                }
            }
    if (dom[0] == 0 && dom[1] == 0 && dom[2] == 0 && dom[3] == 0 && dom[4] == 0 && dom[5] == 0) {
        fl = 0;
        dofi = 0;
    } else {
        for (int i = 2; i < nz + 2; i++)
            for (int j = 2; j < ny + 2; j++)
                for (int k = 2; k < nx + 2; k++) {
                    int gi = i - 2, gj = j - 2, gk = k - 2;
                    size_t c = ix(NX, NY, i, j, k);
                    if (gk < dom[0] || gk > dom[1] || gj < dom[2] || gj > dom[3] ||
                        gi < dom[4] || gi > dom[5])
                        continue;
                    if (G[c] == 0) { fl += 1; sx += u[c]; sy += v[c]; sz += w[c]; }
                    if (gi == dom[4] && G[c] == 0) { p0n += 1; p0 += p[c]; }
                    if (gi == dom[5] && G[c] == 0) { pendn += 1; pend += p[c]; }
                }
    }
    if (wfl == 0) { wsx = 0; wsy = 0; wsz = 0; wmax = 0; }
    if (fl == 0) { sx = 0; sy = 0; sz = 0; }
    if (p0n == 0) p0 = 0;
    if (pendn == 0) pend = 0;
    const double dx2 = dx * dx; /* pow(voxelsize, 2) */
    if (dofi) {
        pgrad = ((pend / pendn) - (p0 / p0n)) / (dom[5] - dom[4]);
        out[0] = (sx / fl) * (1.0 / ORC_RE) * phi * (1.0 / pgrad) * dx2;
        out[1] = (sy / fl) * (1.0 / ORC_RE) * phi * (1.0 / pgrad) * dx2;
        out[2] = (sz / fl) * (1.0 / ORC_RE) * phi * (1.0 / pgrad) * dx2;
    } else {
        out[0] = 0.0; out[1] = 0.0; out[2] = 0.0;
    }
    out[3] = (wsx / wfl) * (1.0 / ORC_RE) * phi * dx2;
    out[4] = (wsy / wfl) * (1.0 / ORC_RE) * phi * dx2;
    out[5] = (wsz / wfl) * (1.0 / ORC_RE) * phi * dx2;
    out[7] = wsz / wfl * dx;
    out[6] = wmax * dx;
}

/* The reference's sweep under `mpirun -np P*Q*R` with dom_decomposition
 * P Q R, one OpenMP thread per block (several blocks per thread when the
 * runtime grants fewer threads).
 *
 * Blocks: src/geometry.cc:129-171 (get_dom_limits): along each axis the first
 * size % dims blocks hold one cell more.  Rank numbering (MPI_Cart_create
 * under MPICH, src/main_stokes_solver.cc:182-193): z fastest, then y, then x.
 * Neighbours: MPI_Cart_shift with the periods of
 * src/main_stokes_solver.cc:163-165; a block is its own neighbour along a
 * periodic axis with one block.
 *
 * Per sweep (src/calc_flow.cc:105-117): every block sweeps its interior in
 * lexicographic order against its own halo copies (orc_sweep_ref on the
 * block's halo-2 array); then communicate_halos (src/parallelization.cc:
 * 193-318) axis by axis: every block packs its two edge layers of the full
 * face (halo rows included), then unpacks its neighbours' layers into its
 * halo (nothing where there is no neighbour); then the blocks holding
 * global z-plane 1 / nz+1, nz+2 apply the pins
 * (src/boundary_conditions.cc:46-78).  A block's halo copy of another
 * block's cells is therefore taken before the pins: on the outlet plane
 * z = nz+1 it holds the swept value, not 3.0.
 *
 * Geometry and codes: the windows of the global halo-2 arrays (the geometry
 * halo exchange of src/parallelization.cc:67-190 gives every block exactly
 * these cells; codes depend only on them).  The state starts from the
 * windows of the global arrays, except that a block's halo cells on the
 * outlet plane inside the global interior and outside the block take
 * outlet[] (the swept values of that plane before its last pins, (ny+4) x
 * (nx+4) like a plane of the global arrays; NULL: the plane of p, right for
 * a state that has not been swept since its pins).  On return the global
 * arrays hold every block's interior and the global halo cells next to it
 * (faces, edges, corners) from its halo; outlet[] (if given) the swept
 * outlet plane.
 * 0, or -1 (allocation failure). */
typedef struct {
    int lo[3], n[3], N[3];   /* first interior cell (x, y, z), extent, extent with halo */
    int nb[6];               /* neighbour block along -x, +x, -y, +y, -z, +z; -1 none */
    uint8_t *G;
    int32_t *code;
    double *f[4];
    double *snd[2];          /* the packed low / high edge layers of the four fields */
} orc_block;

static inline size_t bix(const orc_block *b, int i, int j, int k)
{
    return ((size_t)i * b->N[1] + j) * b->N[0] + k;
}

/* Edge layers of axis a (0 x, 1 y, 2 z): layer l of the face lies at index l
 * along a; the face is every cell of the other two axes (halos included). */
static size_t face_cells(const orc_block *b, int a)
{
    return (size_t)b->N[0] * b->N[1] * b->N[2] / (size_t)b->N[a];
}

static void face_copy(orc_block *b, int a, int l, double *buf, int pack)
{
    /* buf: the four fields' layer l, face order (slow, fast) of the other two
     * axes, field by field */
    const size_t fc = face_cells(b, a);
    for (int q = 0; q < 4; q++) {
        double *f = b->f[q];
        double *o = buf + fc * (size_t)q;
        size_t t = 0;
        if (a == 0) {
            for (int i = 0; i < b->N[2]; i++)
                for (int j = 0; j < b->N[1]; j++, t++) {
                    if (pack) o[t] = f[bix(b, i, j, l)]; else f[bix(b, i, j, l)] = o[t];
                }
        } else if (a == 1) {
            for (int i = 0; i < b->N[2]; i++)
                for (int k = 0; k < b->N[0]; k++, t++) {
                    if (pack) o[t] = f[bix(b, i, l, k)]; else f[bix(b, i, l, k)] = o[t];
                }
        } else {
            for (int j = 0; j < b->N[1]; j++)
                for (int k = 0; k < b->N[0]; k++, t++) {
                    if (pack) o[t] = f[bix(b, l, j, k)]; else f[bix(b, l, j, k)] = o[t];
                }
        }
    }
}

int orc_sweeps_ref_blocks(int solver, const uint8_t *G, const int32_t *code, int nx, int ny, int nz, int bc,
                          double *p, double *u, double *v, double *w, double *outlet, int nsweeps,
                          const int dims_in[3])
{
    const int NXg = HX(nx), NYg = HY(ny);
    const int size[3] = {nx, ny, nz};
    int per[3];
    periodic_axes(bc, &per[0], &per[1], &per[2]);
    int dims[3];
    for (int a = 0; a < 3; a++) {
        dims[a] = dims_in[a] < 1 ? 1 : dims_in[a];
        if (dims[a] > size[a]) dims[a] = size[a];
    }
    const int nblk = dims[0] * dims[1] * dims[2];
    orc_block *B = calloc((size_t)nblk, sizeof(orc_block));
    if (!B) return -1;
    int fail = 0;
    for (int cx = 0; cx < dims[0]; cx++)
        for (int cy = 0; cy < dims[1]; cy++)
            for (int cz = 0; cz < dims[2]; cz++) {
                orc_block *b = B + ((size_t)cx * dims[1] + cy) * dims[2] + cz;
                const int c[3] = {cx, cy, cz};
                for (int a = 0; a < 3; a++) {
                    const int base = size[a] / dims[a], extra = size[a] - dims[a] * base;
                    b->lo[a] = c[a] * base + (c[a] < extra ? c[a] : extra);
                    b->n[a] = base + (c[a] < extra ? 1 : 0);
                    b->N[a] = b->n[a] + 4;
                    int cl = c[a] - 1, cr = c[a] + 1;
                    if (cl < 0) cl = per[a] ? dims[a] - 1 : -1;
                    if (cr >= dims[a]) cr = per[a] ? 0 : -1;
                    int cc[3] = {cx, cy, cz};
                    cc[a] = cl;
                    b->nb[2 * a] = cl < 0 ? -1 : (cc[0] * dims[1] + cc[1]) * dims[2] + cc[2];
                    cc[a] = cr;
                    b->nb[2 * a + 1] = cr < 0 ? -1 : (cc[0] * dims[1] + cc[1]) * dims[2] + cc[2];
                }
                const size_t cells = (size_t)b->N[0] * b->N[1] * b->N[2];
                size_t fmax = 0;
                for (int a = 0; a < 3; a++)
                    if (face_cells(b, a) > fmax) fmax = face_cells(b, a);
                b->G = malloc(cells);
                b->code = malloc(cells * sizeof(int32_t));
                for (int q = 0; q < 4; q++) b->f[q] = malloc(cells * sizeof(double));
                for (int h = 0; h < 2; h++) b->snd[h] = malloc(fmax * 4 * 2 * sizeof(double));
                if (!b->G || !b->code || !b->f[0] || !b->f[1] || !b->f[2] || !b->f[3] || !b->snd[0] || !b->snd[1])
                    fail = 1;
            }
    if (!fail) {
        double *glob[4] = {p, u, v, w};
#pragma omp parallel num_threads(nblk)
        {
#ifdef _OPENMP
            const int tid = omp_get_thread_num(), nth = omp_get_num_threads();
#else
            const int tid = 0, nth = 1;
#endif
            for (int bi = tid; bi < nblk; bi += nth) {
                orc_block *b = B + bi;
                for (int i = 0; i < b->N[2]; i++)
                    for (int j = 0; j < b->N[1]; j++)
                        for (int k = 0; k < b->N[0]; k++) {
                            const int gi = b->lo[2] + i, gj = b->lo[1] + j, gk = b->lo[0] + k;
                            const size_t g = ix(NXg, NYg, gi, gj, gk), l = bix(b, i, j, k);
                            b->G[l] = G[g];
                            b->code[l] = code[g];
                            for (int q = 0; q < 4; q++) b->f[q][l] = glob[q][g];
                            const int inside = i >= 2 && i < b->N[2] - 2 && j >= 2 && j < b->N[1] - 2 && k >= 2 &&
                                               k < b->N[0] - 2;
                            if (outlet && gi == nz + 1 && !inside && gj >= 2 && gj < ny + 2 && gk >= 2 &&
                                gk < nx + 2)
                                b->f[0][l] = outlet[(size_t)gj * NXg + gk];
                        }
            }
#pragma omp barrier
            for (int it = 0; it < nsweeps; it++) {
                for (int bi = tid; bi < nblk; bi += nth) {
                    orc_block *b = B + bi;
                    orc_sweep_ref(solver, b->G, b->code, b->n[0], b->n[1], b->n[2], b->f[0], b->f[1], b->f[2],
                                  b->f[3]);
                }
                for (int a = 0; a < 3; a++) {
#pragma omp barrier
                    for (int bi = tid; bi < nblk; bi += nth) {
                        orc_block *b = B + bi;
                        const size_t fc4 = face_cells(b, a) * 4;
                        /* low edge layers 2, 3; high edge layers N-4, N-3 */
                        face_copy(b, a, 2, b->snd[0], 1);
                        face_copy(b, a, 3, b->snd[0] + fc4, 1);
                        face_copy(b, a, b->N[a] - 4, b->snd[1], 1);
                        face_copy(b, a, b->N[a] - 3, b->snd[1] + fc4, 1);
                    }
#pragma omp barrier
                    for (int bi = tid; bi < nblk; bi += nth) {
                        orc_block *b = B + bi;
                        const size_t fc4 = face_cells(b, a) * 4;
                        const int r = b->nb[2 * a + 1], l = b->nb[2 * a];
                        /* from the neighbour above its low layers, from the one
                         * below its high layers */
                        if (r >= 0) {
                            face_copy(b, a, b->N[a] - 2, B[r].snd[0], 0);
                            face_copy(b, a, b->N[a] - 1, B[r].snd[0] + fc4, 0);
                        }
                        if (l >= 0) {
                            face_copy(b, a, 0, B[l].snd[1], 0);
                            face_copy(b, a, 1, B[l].snd[1] + fc4, 0);
                        }
                    }
                }
#pragma omp barrier
                for (int bi = tid; bi < nblk; bi += nth) {
                    orc_block *b = B + bi;
                    const int cz = (bi % dims[2]);
                    double *pb = b->f[0];
                    for (int j = 2; j < b->N[1] - 2; j++)
                        for (int k = 2; k < b->N[0] - 2; k++) {
                            if (cz == 0 && b->G[bix(b, 1, j, k)] == 0) pb[bix(b, 1, j, k)] = (double)(nz + 2);
                            if (cz == dims[2] - 1) {
                                const size_t hi = bix(b, b->N[2] - 2, j, k), out = bix(b, b->N[2] - 3, j, k);
                                if (outlet) outlet[(size_t)(b->lo[1] + j) * NXg + (b->lo[0] + k)] = pb[out];
                                if (b->G[hi] == 0) pb[hi] = 2.0;
                                if (b->G[out] == 0) pb[out] = 3.0;
                            }
                        }
                }
            }
#pragma omp barrier
            /* back: interiors, and the global halo cells of the boundary blocks */
            for (int bi = tid; bi < nblk; bi += nth) {
                orc_block *b = B + bi;
                for (int i = 0; i < b->N[2]; i++)
                    for (int j = 0; j < b->N[1]; j++)
                        for (int k = 0; k < b->N[0]; k++) {
                            const int gi = b->lo[2] + i, gj = b->lo[1] + j, gk = b->lo[0] + k;
                            /* on every axis: inside the block, or in the
                             * global halo (the block is on that face) */
                            const int kz = (i >= 2 && i < b->N[2] - 2) || gi < 2 || gi >= nz + 2;
                            const int ky = (j >= 2 && j < b->N[1] - 2) || gj < 2 || gj >= ny + 2;
                            const int kx = (k >= 2 && k < b->N[0] - 2) || gk < 2 || gk >= nx + 2;
                            const int keep = kz && ky && kx;
                            if (!keep) continue;
                            const size_t g = ix(NXg, NYg, gi, gj, gk), l = bix(b, i, j, k);
                            for (int q = 0; q < 4; q++) glob[q][g] = b->f[q][l];
                        }
            }
        }
    }
    for (int bi = 0; bi < nblk; bi++) {
        free(B[bi].G); free(B[bi].code);
        for (int q = 0; q < 4; q++) free(B[bi].f[q]);
        free(B[bi].snd[0]); free(B[bi].snd[1]);
    }
    free(B);
    return fail ? -1 : 0;
}

/* dom_decomposition 1 1 P: P z-slabs (orc_sweeps_ref_blocks).  Used to time
 * the CPU baseline on all host cores; P = 1 is orc_sweep_ref + exchange +
 * pins exactly. */
int orc_sweeps_ref_slabs(int solver, const uint8_t *G, const int32_t *code, int nx, int ny, int nz, int bc,
                         double *p, double *u, double *v, double *w, int nsweeps, int P)
{
    const int dims[3] = {1, 1, P};
    return orc_sweeps_ref_blocks(solver, G, code, nx, ny, nz, bc, p, u, v, w, NULL, nsweeps, dims);
}

/* src/calc_flow.cc:82-165 at one rank. */
int orc_calc_flow(const uint8_t *G, const int32_t *code, int nx, int ny, int nz,
                  int bc, const orc_params *prm,
                  double *p, double *u, double *v, double *w,
                  double *rows, int max_rows, int *n_rows)
{
    int t = 0;
    double conv = 1.0 + prm->eps;
    double kperm = 0.0;
    double phi = prm->porosity == -1.0 ? orc_porosity(G, nx, ny, nz) : prm->porosity;
    int nr = 0;
    while (t < prm->max_iter && conv >= prm->eps) {
        if (t == 0) orc_pins(G, nx, ny, nz, p);
        orc_sweep_ref(prm->solver, G, code, nx, ny, nz, p, u, v, w);
        orc_exchange_f64(p, nx, ny, nz, bc);
        orc_exchange_f64(u, nx, ny, nz, bc);
        orc_exchange_f64(v, nx, ny, nz, bc);
        orc_exchange_f64(w, nx, ny, nz, bc);
        orc_pins(G, nx, ny, nz, p);
        if (t % prm->it_eval == 0 && t > 1000) {
            conv = orc_convergence(G, w, nx, ny, nz, &kperm);
            if (t % prm->it_write == 0) {
                double k[8];
                orc_permeability(G, u, v, w, p, nx, ny, nz, prm->dom, prm->dx, phi, k);
                if (rows && nr < max_rows) {
                    double *r = rows + (size_t)nr * 11;
                    r[0] = t; r[1] = conv; r[2] = 0.0; r[3] = k[6]; r[4] = k[7];
                    for (int q = 0; q < 6; q++) r[5 + q] = k[q];
                }
                nr++;
            }
        }
        t++;
    }
    if (n_rows) *n_rows = nr;
    return t;
}

/* ------------------------------------------------------------------------ */
/* Build-side fast modes (no reference code).                                 */
/* ------------------------------------------------------------------------ */

void orc_pack_code(const int32_t *code, int nx, int ny, int nz, uint8_t *pc)
{
    const int NX = HX(nx), NY = HY(ny), NZ = HZ(nz);
    const size_t n = (size_t)NX * NY * NZ;
    for (size_t c = 0; c < n; c++) pc[c] = 255;
    for (int i = 2; i < NZ - 2; i++)
        for (int j = 2; j < NY - 2; j++)
            for (int k = 2; k < NX - 2; k++) {
                size_t c = ix(NX, NY, i, j, k);
                int num = code[c];
                if (num <= 0) continue;
                int cz = num % 10, cy = (num / 10) % 10, cx = (num / 100) % 10;
                pc[c] = (uint8_t)((cx - 1) * 36 + (cy - 1) * 6 + (cz - 1));
            }
}

/* The build's opt-in slip walls (walls = 2; no reference code): with slip on
 * a non-periodic x (y) axis the cells of the first interior column k = 2
 * (row j = 2) keep u (v) at 0 -- their -x (-y) face is the wall -- and after
 * every sweep the halo layers of those faces hold the mirror images of the
 * tangential velocities (v, w on x faces; u, w on y faces) and 0 for the
 * normal one; pressure halos are left as they are.  The halos stay fluid,
 * so the second derivatives see the mirror values: zero shear at the wall.
 * Set for the fast-mode sweeps and ghost fills that follow (test setting,
 * not thread-local). */
static int g_slip_x = 0, g_slip_y = 0;

void orc_set_slip(int slip, int bc)
{
    int px, py, pz;
    periodic_axes(bc, &px, &py, &pz);
    g_slip_x = slip && !px;
    g_slip_y = slip && !py;
}

#define REAL double
#define SUF f64
#include "pm_oracle_fast.inc"
#undef REAL
#undef SUF

#define REAL float
#define SUF f32
#include "pm_oracle_fast.inc"
#undef REAL
#undef SUF

/* Outlet-plane buffer of a FP64 state: pin_out - c2*dt*div at interior fluid
 * cells of z = nz+1, p elsewhere on that plane (wall halos keep p's frozen
 * value), then the periodic x/y copies. */
void orc_outlet_plane(const uint8_t *pc, int nx, int ny, int nz, int bc,
                      const double *p, const double *u, const double *v, const double *w,
                      double pin_out, double *q)
{
    const int NX = HX(nx), NY = HY(ny), NZ = HZ(nz);
    const size_t sj = NX, si = (size_t)NX * NY;
    int px, py, pz;
    const int i = NZ - 3;
    for (int j = 0; j < NY; j++)
        for (int k = 0; k < NX; k++) {
            const size_t c = ix(NX, NY, i, j, k);
            const int interior = j >= 2 && j < NY - 2 && k >= 2 && k < NX - 2;
            if (interior && pc[c] != 255) {
                double div = ((((w[c + si] - w[c]) + v[c + sj]) - v[c]) + u[c + 1]) - u[c];
                q[(size_t)j * NX + k] = pin_out - div * (ORC_C2 * ORC_DT);
            } else {
                q[(size_t)j * NX + k] = p[c];
            }
        }
    periodic_axes(bc, &px, &py, &pz);
    exchange_f64(q, nx, ny, -3, px, py, 0);
}

/* Momentum and continuity residual of a FP64 state (the increment of one
 * sweep per unit relaxation), rounded to FP32; q is the state's outlet-plane
 * buffer.  Solid-face velocity components and solid cells get 0. */
void orc_residual(const uint8_t *pc, int nx, int ny, int nz,
                  const double *p, const double *u, const double *v, const double *w,
                  const double *q, float *Rp, float *Ru, float *Rv, float *Rw)
{
    const int NX = HX(nx), NY = HY(ny), NZ = HZ(nz);
    const size_t n = (size_t)NX * NY * NZ;
    for (size_t c = 0; c < n; c++) { Rp[c] = 0.f; Ru[c] = 0.f; Rv[c] = 0.f; Rw[c] = 0.f; }
    for (int i = 2; i < NZ - 2; i++)
        for (int j = 2; j < NY - 2; j++)
            for (int k = 2; k < NX - 2; k++) {
                size_t c = ix(NX, NY, i, j, k);
                if (pc[c] == 255) continue;
                double f[4];
                const int out = (i == NZ - 3);
                stencil_force_f64(pc[c], c, (size_t)NX, (size_t)NX * NY, p, u, v, w,
                                  out ? q : NULL, (size_t)j * NX + k, f);
                Ru[c] = (float)f[1]; Rv[c] = (float)f[2]; Rw[c] = (float)f[3];
                double div = (((((w[c + (size_t)NX * NY] - w[c]) + v[c + NX]) - v[c]) + u[c + 1]) - u[c]);
                Rp[c] = (float)div;
            }
}

/* Defect-correction fold: X += d on the fluid interior, d = 0 everywhere,
 * then the solution ghost fill on X. */
void orc_fold(const uint8_t *G, int nx, int ny, int nz, int bc,
              double *Xp, double *Xu, double *Xv, double *Xw,
              float *dp, float *du, float *dv, float *dw)
{
    const int NX = HX(nx), NY = HY(ny), NZ = HZ(nz);
    for (int i = 0; i < NZ; i++)
        for (int j = 0; j < NY; j++)
            for (int k = 0; k < NX; k++) {
                const size_t c = ix(NX, NY, i, j, k);
                const int interior = i >= 2 && i < NZ - 2 && j >= 2 && j < NY - 2 &&
                                     k >= 2 && k < NX - 2;
                if (interior && G[c] == 0) {
                    Xp[c] = Xp[c] + (double)dp[c];
                    Xu[c] = Xu[c] + (double)du[c];
                    Xv[c] = Xv[c] + (double)dv[c];
                    Xw[c] = Xw[c] + (double)dw[c];
                }
                dp[c] = 0.f; du[c] = 0.f; dv[c] = 0.f; dw[c] = 0.f;
            }
    orc_ghost_f64(G, nx, ny, nz, bc, Xp, Xu, Xv, Xw, NULL, (double)(nz + 2), 2.0, 3.0);
}

/* Exported for the known-answer tests of get_2nd_derivation (src/solver.cc:57-91). */
double orc_d2_ref(int cs, double ll, double l, double c, double r, double rr)
{
    return d2_ref(cs, ll, l, c, r, rr);
}
double orc_d2_f64(int cs, double ll, double l, double c, double r, double rr)
{
    return d2_f64(cs, ll, l, c, r, rr);
}
float orc_d2_f32(int cs, float ll, float l, float c, float r, float rr)
{
    return d2_f32(cs, ll, l, c, r, rr);
}

/* A whole fast-mode solve (the build's Jacobi / 3-colour SOR with the
 * relaxation pair (omega, omega_p), FP64 or FP32 correction sweeps with FP64
 * defect correction) from the initial state in p, u, v, w (after the pins) to
 * the reference's stop rule (src/calc_flow.cc:101,120-128): the convergence
 * value of the FP64 state every it_eval sweeps past sweep 1000.  In the
 * defect-correction form the FP32 correction d obeys the same sweep with the
 * FP32 residual R of the FP64 state added to its forces, pins and outlet
 * buffer 0; every it_eval sweeps (sweep index t % it_eval == 0, before the
 * convergence test of that sweep) d is folded into the state and R and the
 * outlet buffer recomputed.  Returns the sweeps done (the final state in
 * p, u, v, w; *kprev the last convergence test's permeability in voxel
 * units), or -1 on allocation failure. */
int orc_solve_fast(int mode, int dc, const uint8_t *G, const uint8_t *pc, int nx, int ny, int nz, int bc,
                   double omega, double omega_p, double eps, int it_eval, int max_iter,
                   double *p, double *u, double *v, double *w, double *kprev_out)
{
    const size_t n = (size_t)HX(nx) * HY(ny) * HZ(nz), nq = (size_t)HX(nx) * HY(ny);
    const double pin_in = (double)(nz + 2);
    double *X[4] = {p, u, v, w};
    double *B[4] = {NULL, NULL, NULL, NULL}, *q = calloc(nq, sizeof(double)), *qb = NULL;
    float *d[2][4] = {{NULL}}, *qd[2] = {NULL, NULL}, *R[4] = {NULL, NULL, NULL, NULL};
    int fail = !q, t = 0;
    double kprev = 0.0, conv = 1.0 + eps;
    if (!dc && mode == 1) {
        qb = calloc(nq, sizeof(double));
        fail |= !qb;
        for (int f = 0; f < 4; f++) {
            B[f] = malloc(n * sizeof(double));
            fail |= !B[f];
            if (B[f]) memcpy(B[f], X[f], n * sizeof(double));
        }
    }
    if (dc) {
        for (int f = 0; f < 4; f++) {
            for (int h = 0; h < 2; h++) { d[h][f] = calloc(n, sizeof(float)); fail |= !d[h][f]; }
            R[f] = calloc(n, sizeof(float));
            fail |= !R[f];
        }
        for (int h = 0; h < 2; h++) { qd[h] = calloc(nq, sizeof(float)); fail |= !qd[h]; }
    }
    if (!fail) {
        orc_outlet_plane(pc, nx, ny, nz, bc, p, u, v, w, 3.0, q);
        if (qb) memcpy(qb, q, nq * sizeof(double));
        if (dc) orc_residual(pc, nx, ny, nz, p, u, v, w, q, R[0], R[1], R[2], R[3]);
        double **A = X, **Bn = B, *qa = q, *qn = qb;
        while (t < max_iter && conv >= eps) {
            if (!dc && mode == 1) {
                orc_sweep_jacobi_f64(pc, nx, ny, nz, bc, omega, omega_p, A[0], A[1], A[2], A[3], Bn[0], Bn[1],
                                     Bn[2], Bn[3], qa, qn, 3.0, NULL, NULL, NULL, NULL);
                orc_ghost_f64(G, nx, ny, nz, bc, Bn[0], Bn[1], Bn[2], Bn[3], qn, pin_in, 2.0, 3.0);
                double **tmp = A; A = Bn; Bn = tmp;
                double *tq = qa; qa = qn; qn = tq;
            } else if (!dc) {
                orc_sweep_colour_f64(pc, nx, ny, nz, omega, omega_p, A[0], A[1], A[2], A[3], qa, 3.0, NULL, NULL,
                                     NULL, NULL);
                orc_ghost_f64(G, nx, ny, nz, bc, A[0], A[1], A[2], A[3], qa, pin_in, 2.0, 3.0);
            } else {
                if (mode == 1) {
                    orc_sweep_jacobi_f32(pc, nx, ny, nz, bc, omega, omega_p, d[0][0], d[0][1], d[0][2], d[0][3],
                                         d[1][0], d[1][1], d[1][2], d[1][3], qd[0], qd[1], 0.f, R[0], R[1], R[2],
                                         R[3]);
                    orc_ghost_f32(G, nx, ny, nz, bc, d[1][0], d[1][1], d[1][2], d[1][3], qd[1], 0.f, 0.f, 0.f);
                    for (int f = 0; f < 4; f++) { float *tmp = d[0][f]; d[0][f] = d[1][f]; d[1][f] = tmp; }
                    float *tq = qd[0]; qd[0] = qd[1]; qd[1] = tq;
                } else {
                    orc_sweep_colour_f32(pc, nx, ny, nz, omega, omega_p, d[0][0], d[0][1], d[0][2], d[0][3], qd[0],
                                         0.f, R[0], R[1], R[2], R[3]);
                    orc_ghost_f32(G, nx, ny, nz, bc, d[0][0], d[0][1], d[0][2], d[0][3], qd[0], 0.f, 0.f, 0.f);
                }
                if (t % it_eval == 0) {
                    orc_fold(G, nx, ny, nz, bc, p, u, v, w, d[0][0], d[0][1], d[0][2], d[0][3]);
                    orc_outlet_plane(pc, nx, ny, nz, bc, p, u, v, w, 3.0, q);
                    orc_residual(pc, nx, ny, nz, p, u, v, w, q, R[0], R[1], R[2], R[3]);
                    memset(qd[0], 0, nq * sizeof(float));
                    memset(qd[1], 0, nq * sizeof(float));
                }
            }
            if (t % it_eval == 0 && t > 1000) conv = orc_convergence(G, dc ? w : A[3], nx, ny, nz, &kprev);
            t++;
        }
        if (!dc && A != X)
            for (int f = 0; f < 4; f++) memcpy(X[f], A[f], n * sizeof(double));
    }
    for (int f = 0; f < 4; f++) {
        free(B[f]); free(R[f]);
        for (int h = 0; h < 2; h++) free(d[h][f]);
    }
    free(q); free(qb); free(qd[0]); free(qd[1]);
    if (kprev_out) *kprev_out = kprev;
    return fail ? -1 : t;
}
