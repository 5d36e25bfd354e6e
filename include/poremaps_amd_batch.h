/*
 * poremaps_amd_batch.h -- many geometries solved together on one GPU.
 *
 * No reference counterpart: the reference solves one geometry per mpirun.
 * Callers: the serving loop and batch drivers that would otherwise run one
 * `POREMAPS input.inp` after another.
 *
 * Every sample is an independent solve with the reference's stop rule
 * (src/calc_flow.cc:101,120-128); all active samples advance together, one
 * it_eval interval per pm_batch_step, through batched kernels that cover every
 * sample's fluid voxels in one launch per phase (chunks of 256 fluid voxels
 * tagged with their sample), optionally captured once as a HIP graph and
 * replayed per interval.  A sample's trajectory is bit-identical to the same
 * solve run alone through pm_solve.  Fast modes only (symplectic Jacobi, FP64
 * or FP32 + FP64 defect correction).
 */
#ifndef POREMAPS_AMD_BATCH_H
#define POREMAPS_AMD_BATCH_H

#include <stdint.h>

#include "poremaps_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pm_batch pm_batch;

typedef struct {
    int precision;           /* PM_PREC_F64 or PM_PREC_F32DC */
    double omega;            /* velocity relaxation; 0 (default): the pair chosen per sample
                                from its porosity, as pm_params.omega = 0 */
    int it_eval, it_write, max_iter;
    double eps;
    int device;
    int use_graph;           /* capture each interval as a HIP graph (1) or launch directly (0) */
    int64_t budget_bytes;    /* admission limit for the samples' device memory; 0: 90 % of free HBM */
    double omega_p;          /* pressure relaxation; 0: omega (as pm_params.omega_p); ignored
                                when omega is 0 */
} pm_batch_params;

/* Sample states for pm_batch_status. */
#define PM_SAMPLE_RUNNING 0
#define PM_SAMPLE_CONVERGED 1
#define PM_SAMPLE_MAX_ITER 2
#define PM_SAMPLE_DIVERGED 3
#define PM_SAMPLE_REJECTED 4   /* pm_serve responses only: the request was not admitted (its three
                                  solves exceed the whole memory budget, or pm_batch_add refused
                                  its geometry or parameters); its k is zero */

void pm_batch_params_default(pm_batch_params *p);
int pm_batch_create(const pm_batch_params *p, pm_batch **out);

/* Device bytes one sample of this size needs (admission footprint). */
int64_t pm_batch_sample_bytes(const pm_batch_params *p, int nx, int ny, int nz);

/* Admit a geometry (nx*ny*nz bytes, x fastest, 0 fluid / 1 solid); runs its
 * first sweep and returns its slot.  PM_E_MEMORY when the budget would be
 * exceeded (nothing is allocated then). */
int pm_batch_add(pm_batch *b, const uint8_t *raw, int nx, int ny, int nz, int boundary_method,
                 double voxel_size, double porosity, int *slot);

/* Stopping tolerance of one sample (default: the batch's eps), the eps of
 * the reference's stop rule (src/main.cc:181-185). */
int pm_batch_set_eps(pm_batch *b, int slot, double eps);
/* The sample's dom_interest box (x0 x1 y0 y1 z0 z1, interior indices,
 * inclusive; all 0: none): k13, k23, k33 of its log rows and result then
 * follow src/evaluation.cc:54-211 for that sub-domain. */
int pm_batch_set_dom_interest(pm_batch *b, int slot, const int box[6]);

/* One it_eval interval for every running sample, then their convergence
 * tests; *n_running = samples still running afterwards. */
int pm_batch_step(pm_batch *b, int *n_running);

/* pm_batch_step in two halves, for callers that overlap host work or another
 * batch (on its own stream) with the interval: pm_batch_launch enqueues the
 * interval and the read-back of its sums on the batch's stream and returns;
 * pm_batch_ready gives 1 once that work has finished on the GPU, else 0;
 * pm_batch_finish waits for it and runs the convergence tests.  Between
 * launch and finish the batch takes no other call (PM_E_STATE).  No
 * counterpart in the reference. */
int pm_batch_launch(pm_batch *b);
int pm_batch_ready(pm_batch *b);
int pm_batch_finish(pm_batch *b, int *n_running);

/* State and result of a sample (result.last: permeability at the last
 * interval end). */
int pm_batch_status(pm_batch *b, int slot, int *state, pm_result *res);

/* Interior field of a sample (as pm_get_field). */
int pm_batch_get_field(pm_batch *b, int slot, int field, void *out);

/* Release a sample's device memory (evict); the slot becomes free. */
int pm_batch_remove(pm_batch *b, int slot);

/* Device bytes currently held by admitted samples, and the budget. */
int pm_batch_memory(pm_batch *b, int64_t *used, int64_t *budget);

void pm_batch_destroy(pm_batch *b);

#ifdef __cplusplus
}
#endif
#endif
