/*
 * poremaps_amd.h -- C ABI of the MI355X-native POREMAPS Stokes solver.
 *
 * One context = one geometry solved on one GPU.  The library owns all device
 * memory; callers pass plain host pointers and sizes.  Every entry point
 * returns 0 on success or a negative PM_E* code; pm_last_error() gives the
 * message of the calling thread's last failure.  One context must not be used
 * from two threads at once.
 *
 * Reference seam replaced: calc_flow() and its setup sequence in
 *   src/main_stokes_solver.cc:195-303 and src/calc_flow.cc:53-169
 *   (calc_flow(bool*** geom, int* size, ..., double*** press, vel_x, vel_y,
 *    vel_z, int it_write, int it_eval, unsigned solver, double eps,
 *    double porosity, ..., int* dom_interest, char* log)).
 */
#ifndef POREMAPS_AMD_H
#define POREMAPS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM_OK 0
#define PM_E_ARG (-1)      /* invalid argument or parameter */
#define PM_E_HIP (-2)      /* HIP runtime failure */
#define PM_E_STATE (-3)    /* call out of order (e.g. solve before geometry) */
#define PM_E_MEMORY (-4)   /* device or host allocation failed / over budget */
#define PM_E_GEOMETRY (-5) /* geometry bytes other than 0 and 1 */
#define PM_E_ABORTED (-6)  /* a callback asked to stop */
#define PM_E_DIVERGED (-7) /* non-finite convergence value */

/* Iteration ordering. */
#define PM_MODE_REFERENCE 0 /* the reference's in-place lexicographic sweep,
                               solving_algorithm 1/2/3, reproduced exactly */
#define PM_MODE_JACOBI 1    /* symplectic Jacobi, out of place, relaxation omega */
#define PM_MODE_SOR3 2      /* 3-colour SOR ((i+j+k) mod 3), in place, omega */

/* Arithmetic of the bulk sweeps (PM_MODE_JACOBI / PM_MODE_SOR3 only). */
#define PM_PREC_F64 0    /* FP64 state and sweeps */
#define PM_PREC_F32DC 1  /* FP32 correction sweeps + FP64 defect correction */

/* Field selectors for pm_get_field. */
#define PM_FIELD_VELX 0  /* double, interior, x fastest (not scaled by dx) */
#define PM_FIELD_VELY 1
#define PM_FIELD_VELZ 2
#define PM_FIELD_PRESS 3
#define PM_FIELD_NHOOD 4 /* int32 neighbourhood codes, interior */

typedef struct pm_ctx pm_ctx;

/* The ordered .inp parameters (src/read_problem_params.cc:75-90) plus the
 * build's own knobs.  pm_params_default() fills reference behaviour. */
typedef struct {
    int nx, ny, nz;          /* size_x_y_z */
    int boundary_method;     /* 0..4 */
    double voxel_size;       /* m */
    int max_iter, it_eval, it_write;
    int solving_algorithm;   /* 1, 2, 3 (reference mode) */
    double eps;
    double porosity;         /* -1.0: computed from the geometry */
    int dom_interest[6];     /* xmin xmax ymin ymax zmin zmax (inclusive) */
    /* build knobs */
    int mode;                /* PM_MODE_* */
    int precision;           /* PM_PREC_* */
    double omega;            /* relaxation of PM_MODE_JACOBI / PM_MODE_SOR3 (velocities); 0:
                                the pair (omega, omega_p) chosen at pm_load_geometry from the
                                geometry's porosity: (0.6, 30) below 0.20, where the pressure
                                mode limits convergence, else (2.5, 7) (pm_relaxation) */
    int device;              /* HIP device ordinal */
    double omega_p;          /* pressure relaxation of the fast modes; 0: omega.  The fixed
                                point does not depend on it; the Jacobi sweep is stable while
                                omega * omega_p stays below about 19 (a divergent solve
                                restarts with omega_p cut by a quarter: pm_result.restarts) */
    int z_slabs;             /* reference mode: shorthand for dom_decomposition 1 1 P
                                (0 or 1: not used) */
    int dom_decomposition[3];/* reference mode: sweep as `mpirun -np P*Q*R` does with
                                dom_decomposition P Q R (src/main_stokes_solver.cc:161-193,
                                src/geometry.cc:129-171): every block swept against its halo
                                copies of the other blocks as the sweep started; 0 0 0 or
                                1 1 1: one domain.  1 <= P <= nx, Q <= ny, R <= nz, at most
                                PM_MAX_BLOCKS blocks.  Fast modes ignore it. */
    int walls;               /* 0: the reference's domain faces -- the halo of a non-periodic
                                face is fluid, with u = v = w = 0 and p = -1 frozen
                                (src/geometry.cc:98-106, src/parallelization.cc:193-318);
                                1: no-slip walls on the non-periodic x and y faces
                                (boundary_method 1-4): their halo layers are solid, so the
                                first cells use the one-sided wall stencils of
                                src/solver.cc:57-91.  Every mode.
                                2: slip (symmetry) walls on those faces: the halo mirrors
                                the geometry and the tangential velocities, the normal
                                velocity is held at 0 on the wall face.  Fast modes only
                                (PM_E_ARG in reference mode), not in a batch.
                                1 and 2 have no reference counterpart (opt-in physical
                                models). */
} pm_params;

#define PM_MAX_BLOCKS 4096

/* One log row (src/output.cc:193-230 column order). */
typedef struct {
    int iteration;
    double conv, tps, wmax_velz, wmean_velz;
    double k13, k23, k33, wk13, wk23, wk33;
} pm_log_row;

typedef struct {
    int sweeps;              /* sweeps performed */
    double conv;             /* last convergence value */
    double porosity;         /* porosity used for k */
    double seconds;          /* device-resident solve time (no file I/O) */
    double glups;            /* nx*ny*nz*sweeps / seconds / 1e9 */
    int n_log_rows;
    pm_log_row last;         /* permeability at the last sweep */
    int restarts;            /* fast modes: restarts from the initial state with omega_p
                                reduced by a quarter after a divergence (sweeps counts the
                                last attempt only) */
} pm_result;

/* Called at every log event (t > 1000, t % it_eval == 0, t % it_write == 0);
 * non-zero return stops the solve with PM_E_ABORTED. */
typedef int (*pm_log_fn)(const pm_log_row *row, void *user);
/* Called every 100 sweeps (src/calc_flow.cc:159-163); seconds = wall time of
 * the last sweep (mean over the sweeps since the previous call). */
typedef int (*pm_progress_fn)(int iteration, double conv, double seconds, void *user);

const char *pm_last_error(void);
void pm_params_default(pm_params *p);

/* Replaces the setup of src/main_stokes_solver.cc:195-275 (parameters,
 * allocation).  Device memory is reserved on the first pm_load_geometry. */
int pm_ctx_create(const pm_params *p, pm_ctx **out);

/* Replaces read_geometry + eval_geometry + allocate_initial_field
 * (src/geometry.cc:58-126, :174-302; src/initialize.cc:69-121): raw holds
 * nx*ny*nz bytes, x fastest, 0 = fluid, 1 = solid. */
int pm_load_geometry(pm_ctx *ctx, const uint8_t *raw);

/* Replaces calc_flow (src/calc_flow.cc:53-169): iterate until conv < eps or
 * max_iter.  Callbacks may be NULL. */
int pm_solve(pm_ctx *ctx, pm_log_fn log_fn, pm_progress_fn progress_fn, void *user,
             pm_result *res);

/* Copy a field's interior to host memory (nx*ny*nz elements, x fastest):
 * double for velocities and pressure, int32 for PM_FIELD_NHOOD.  Replaces the
 * gather in write_output_raw (src/output.cc:53-153). */
int pm_get_field(pm_ctx *ctx, int field, void *out);

/* Porosity of the loaded geometry (src/geometry.cc:335-364). */
int pm_porosity(pm_ctx *ctx, double *phi);

/* Whole-domain and dom_interest permeability of the current state
 * (src/evaluation.cc:54-211); out = k13 k23 k33 wk13 wk23 wk33 wmax wmean. */
int pm_permeability(pm_ctx *ctx, double out[8]);

/* Advance exactly n sweeps with the context's ordering (ghost fill and pins
 * included, no convergence test); for step-by-step parity checks. */
int pm_step(pm_ctx *ctx, int n);

/* Copy a FP64 field including its 2-cell halo ((nz+4)*(ny+4)*(nx+4), x
 * fastest) -- the layout of the reference's jagged arrays. */
int pm_get_field_halo(pm_ctx *ctx, int field, double *out);

/* Bytes of device memory this context holds (or will hold once loaded). */
int64_t pm_device_bytes(const pm_params *p);

/* Kernel timing of the context's sweep: run n sweeps (advancing the state
 * like pm_step) with HIP events around every kernel on the library's stream.
 * Per kernel: average duration of one launch and its algorithmic HBM bytes
 * (the minimum traffic of one launch: every field it must read once, every
 * field it writes once).  No reference counterpart (measurement only). */
typedef struct {
    char name[48];
    int launches;
    double avg_ms;
    double bytes_per_launch;
} pm_kernel_timing;
int pm_time_kernels(pm_ctx *ctx, int n, pm_kernel_timing *out, int max_out, int *n_out);

/* Jacobi sweep kernels in use and the autotuner's measurements.  The fast
 * Jacobi mode has several kernel forms that produce the same iterates bit
 * for bit (velocity + pressure kernel pair over the fluid list with select or
 * case-predicated neighbour loads; one lagged pass over the fluid list with
 * case-predicated loads, with every load issued up front, or reading
 * per-voxel records; one lagged dense 2.5-D pass); the first sweeps of a
 * solve run each in turn and the fastest is kept (environment:
 * POREMAPS_TUNE=0 disables, POREMAPS_JACOBI=pair|list|records|stream fixes the
 * form).  name receives the form in use ("n/a" in the reference mode; for
 * 3-colour SOR "list-colour" (FP64), "interleaved-colour-split" or
 * "interleaved-colour" (FP32+DC state in colour-split or natural rows));
 * tried (optional) "form=ms/sweep;..." once tuning has finished.  No
 * reference counterpart. */
int pm_kernel_choice(pm_ctx *ctx, char *name, int len, char *tried, int tried_len);

/* The relaxation pair in use (after the choice by porosity when
 * pm_params.omega was 0, and after any restart that cut omega_p).  No
 * reference counterpart. */
int pm_relaxation(pm_ctx *ctx, double *omega, double *omega_p);

void pm_destroy(pm_ctx *ctx);

/* No counterpart in the reference: pm_permeability (the same reduction
 * kernels, the same results bit for bit) for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  The caller passes its own device arrays,
 * interior only, x fastest (nx*ny*nz elements): geom 0 = fluid / 1 = solid,
 * p, u, v, w in FP64; out receives k13 k23 k33 wk13 wk23 wk33 wmax wmean in
 * host memory; porosity -1.0 computes it from geom; stream: a hipStream_t
 * (NULL: the null stream), synchronized before return.  The partial sums
 * live in a fixed 1.38 MB array of the library on each GPU, kept until the
 * process ends: the call allocates nothing, and calls that overlap on one
 * GPU race for that array.  At most PM_NOALLOC_MAX_VOXELS voxels. */
#define PM_NOALLOC_MAX_VOXELS 2147483647LL
int pm_permeability_noalloc(const uint8_t *geom, const double *p, const double *u, const double *v,
                            const double *w, int nx, int ny, int nz, const int dom_interest[6],
                            double voxel_size, double porosity, double out[8], void *stream);

#ifdef __cplusplus
}
#endif
#endif
