/*
 * pm_oracle.h -- CPU restatement of the POREMAPS pseudo-time Stokes sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in poremaps_amd/: only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Layout (same as the reference's jagged A[z][y][x], flattened): every field
 * has a halo of 2 on each side, shape (nz+4, ny+4, nx+4), x fastest;
 * index(i,j,k) = (i*(ny+4) + j)*(nx+4) + k with i=z, j=y, k=x.
 * Geometry: 0 = fluid, 1 = solid (uint8).  Neighbourhood codes: int32,
 * 100*cx + 10*cy + cz for interior fluid cells, -1 solid, 0 unclassified.
 *
 * Periodicity by boundary_method (reference src/main_stokes_solver.cc:163-165):
 *   0: x, y, z periodic;  1, 2: z periodic only;  3, 4: nothing periodic.
 */
#ifndef PM_ORACLE_H
#define PM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* --- reference-faithful pieces (FP64) ---------------------------------- */
void orc_geom_halo(const uint8_t *raw, int nx, int ny, int nz, int bc, uint8_t *G);
/* walls = 1: the halo layers of the non-periodic x / y faces solid (no-slip
 * domain walls; build-side option, the reference leaves them fluid). */
void orc_geom_halo_walls(const uint8_t *raw, int nx, int ny, int nz, int bc, int walls, uint8_t *G);
void orc_eval_geometry(const uint8_t *G, int nx, int ny, int nz, int32_t *code);
double orc_porosity(const uint8_t *G, int nx, int ny, int nz);
void orc_init_fields(const uint8_t *G, int nx, int ny, int nz, int bc,
                     double *p, double *u, double *v, double *w);
void orc_exchange_f64(double *f, int nx, int ny, int nz, int bc);
void orc_pins(const uint8_t *G, int nx, int ny, int nz, double *p);
void orc_sweep_ref(int solver, const uint8_t *G, const int32_t *code,
                   int nx, int ny, int nz,
                   double *p, double *u, double *v, double *w);
/* nsweeps reference sweeps (+ halo exchange + pins) as `mpirun -np P*Q*R`
 * runs them with dom_decomposition P Q R (dims = {P, Q, R}): blocks swept
 * against frozen halo copies, one OpenMP thread per block.  outlet: the
 * outlet plane (z = nz+1, (ny+4)*(nx+4)) before its last pins, carried
 * between calls (NULL: the plane of p).  0 or -1 (allocation failure). */
int orc_sweeps_ref_blocks(int solver, const uint8_t *G, const int32_t *code, int nx, int ny, int nz, int bc,
                          double *p, double *u, double *v, double *w, double *outlet, int nsweeps,
                          const int dims[3]);
/* The same with dom_decomposition 1 1 P (z-slabs); P = 1: the serial sweep. */
int orc_sweeps_ref_slabs(int solver, const uint8_t *G, const int32_t *code, int nx, int ny, int nz, int bc,
                         double *p, double *u, double *v, double *w, int nsweeps, int P);
double orc_convergence(const uint8_t *G, const double *w, int nx, int ny, int nz,
                       double *kprev);
/* out[0..7] = k13, k23, k33, wk13, wk23, wk33, wmax_velz, wmean_velz */
void orc_permeability(const uint8_t *G, const double *u, const double *v,
                      const double *w, const double *p, int nx, int ny, int nz,
                      const int *dom, double dx, double phi, double *out);

typedef struct {
    int max_iter, it_eval, it_write;
    int solver;        /* 1, 2, 3 reference; see orc_run_fast for the rest */
    double eps;
    double porosity;   /* -1.0: compute from the geometry */
    int dom[6];
    double dx;
} orc_params;

/* Reference solver loop (src/calc_flow.cc:53-169) on an initialised state.
 * Returns the number of sweeps; writes up to max_rows log rows of 11 doubles
 * (iteration, conv, tps(=0), wmax, wmean, k13, k23, k33, wk13, wk23, wk33). */
int orc_calc_flow(const uint8_t *G, const int32_t *code, int nx, int ny, int nz,
                  int bc, const orc_params *prm,
                  double *p, double *u, double *v, double *w,
                  double *rows, int max_rows, int *n_rows);

/* --- build-side fast-mode semantics ------------------------------------ */
/* Packed per-voxel code: (cx-1)*36 + (cy-1)*6 + (cz-1) for fluid, 255 solid. */
void orc_pack_code(const int32_t *code, int nx, int ny, int nz, uint8_t *pc);

/* Outlet-plane buffer q (one plane, (ny+4)*(nx+4)): on the outlet plane
 * z = nz+1 the fast modes read the -x/-y neighbour pressure through q, which
 * holds pin_out - c2*dt*div (the unpinned pressure the reference's
 * lexicographic sweep exposes there between update and re-pin).  This makes
 * every ordering and relaxation factor converge to the reference's fixed
 * point.  A solution state starts with q = its p on that plane; a correction
 * state with q = 0. */

/* One symplectic Jacobi sweep, out of place: reads (p0,u0,v0,w0,q0), writes
 * the fluid interior of (p1,u1,v1,w1,q1) -- velocities from the old state,
 * their periodic copies, then pressure from the new velocities.  R* (all NULL
 * or all set): FP32 residual added (defect-correction correction sweep).
 * omega relaxes the velocities, omega_p the pressure. */
void orc_sweep_jacobi_f64(const uint8_t *pc, int nx, int ny, int nz, int bc, double omega, double omega_p,
                          const double *p0, const double *u0, const double *v0, const double *w0,
                          double *p1, double *u1, double *v1, double *w1,
                          const double *q0, double *q1, double pin_out,
                          const float *Rp, const float *Ru, const float *Rv, const float *Rw);
void orc_sweep_jacobi_f32(const uint8_t *pc, int nx, int ny, int nz, int bc, double omega, double omega_p,
                          const float *p0, const float *u0, const float *v0, const float *w0,
                          float *p1, float *u1, float *v1, float *w1,
                          const float *q0, float *q1, float pin_out,
                          const float *Rp, const float *Ru, const float *Rv, const float *Rw);
/* One 3-colour SOR sweep in place (colour (i+j+k) mod 3, order 0, 1, 2). */
void orc_sweep_colour_f64(const uint8_t *pc, int nx, int ny, int nz, double omega, double omega_p,
                          double *p, double *u, double *v, double *w, double *q, double pin_out,
                          const float *Rp, const float *Ru, const float *Rv, const float *Rw);
void orc_sweep_colour_f32(const uint8_t *pc, int nx, int ny, int nz, double omega, double omega_p,
                          float *p, float *u, float *v, float *w, float *q, float pin_out,
                          const float *Rp, const float *Ru, const float *Rv, const float *Rw);
/* After a sweep: periodic copies of the four fields and of q (x/y), then the
 * pressure pins (z=1: pin_in, z=nz+2: pin_hi, z=nz+1: pin_out where fluid).
 * A solution state uses (nz+2, 2, 3); a correction state (0, 0, 0). */
void orc_ghost_f64(const uint8_t *G, int nx, int ny, int nz, int bc,
                   double *p, double *u, double *v, double *w, double *q,
                   double pin_in, double pin_hi, double pin_out);
void orc_ghost_f32(const uint8_t *G, int nx, int ny, int nz, int bc,
                   float *p, float *u, float *v, float *w, float *q,
                   float pin_in, float pin_hi, float pin_out);
/* q of a FP64 state from its current fields. */
void orc_outlet_plane(const uint8_t *pc, int nx, int ny, int nz, int bc,
                      const double *p, const double *u, const double *v, const double *w,
                      double pin_out, double *q);
/* Momentum forces and divergence of a FP64 state (with its q), in FP32. */
void orc_residual(const uint8_t *pc, int nx, int ny, int nz,
                  const double *p, const double *u, const double *v, const double *w,
                  const double *q, float *Rp, float *Ru, float *Rv, float *Rw);
/* Defect-correction fold: X += d on the fluid interior, d = 0, X ghost fill. */
void orc_fold(const uint8_t *G, int nx, int ny, int nz, int bc,
              double *Xp, double *Xu, double *Xv, double *Xw,
              float *dp, float *du, float *dv, float *dw);

/* A whole fast-mode solve to the reference's stop rule: mode 1 Jacobi, 2
 * 3-colour SOR; dc 0 FP64, 1 FP32 correction sweeps + FP64 defect correction
 * (fold every it_eval sweeps).  From the initial state in p, u, v, w (pins
 * applied); returns the sweeps (final state in p, u, v, w, *kprev the last
 * convergence test's mean(w)/Re) or -1. */
/* The build's slip walls for the fast-mode sweeps and ghost fills that
 * follow: slip = 1 on the non-periodic x / y faces of boundary method bc, 0
 * off (see pm_oracle.c). */
void orc_set_slip(int slip, int bc);
int orc_solve_fast(int mode, int dc, const uint8_t *G, const uint8_t *pc, int nx, int ny, int nz, int bc,
                   double omega, double omega_p, double eps, int it_eval, int max_iter,
                   double *p, double *u, double *v, double *w, double *kprev);

/* get_2nd_derivation (reference form) and the fast-mode forms. */
double orc_d2_ref(int cs, double ll, double l, double c, double r, double rr);
double orc_d2_f64(int cs, double ll, double l, double c, double r, double rr);
float orc_d2_f32(int cs, float ll, float l, float c, float r, float rr);

#ifdef __cplusplus
}
#endif
#endif
