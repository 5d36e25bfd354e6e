/*
 * poremaps_amd_serve.h -- permeability tensors and a request-serving loop.
 *
 * pm_tensor replaces the manual procedure of the reference README
 * ("Compute the permeability tensor", README.md:217-229): three solves of the
 * geometry rotated so that each axis in turn is the flow axis e3, their
 * (wk13, wk23, wk33) mapped back to physical axes.  Geometry arrays are
 * (nz, ny, nx) with x fastest (the .raw layout); the rotated arrays are
 * transpose(A, (2, 0, 1)) (flow along x) and transpose(A, (1, 2, 0)) (flow
 * along y).
 *
 * pm_serve has no reference counterpart (callers: permeability services that
 * would otherwise queue POREMAPS runs): a continuous-batching loop over the
 * batch engine.  Requests arrive over time, each expands into the three
 * rotated solves, is admitted when its device-memory footprint fits the
 * budget (first come, first served), shares the GPU with every other running
 * solve and is evicted when done.  Optionally the small requests run in a
 * second batch that gets about half the GPU time while both have work
 * (pm_serve_params.small_cells).
 */
#ifndef POREMAPS_AMD_SERVE_H
#define POREMAPS_AMD_SERVE_H

#include <stdint.h>

#include "poremaps_amd.h"
#include "poremaps_amd_batch.h"

#ifdef __cplusplus
extern "C" {
#endif

/* k[3*i + j]: permeability for velocity along axis i driven along axis j
 * (axes 0 = x, 1 = y, 2 = z), m^2, whole-domain (the wk columns). */
int pm_tensor(const pm_batch_params *p, const uint8_t *raw, int nx, int ny, int nz, int boundary_method,
              double voxel_size, double porosity, double k[9], pm_result res[3]);

typedef struct {
    const uint8_t *raw;      /* nx*ny*nz bytes, x fastest */
    int nx, ny, nz;
    int boundary_method;
    double voxel_size;
    double porosity;         /* -1: computed */
    double eps;              /* stopping tolerance; <= 0: the batch's eps */
    double arrival_s;        /* seconds after the start of pm_serve */
    int dom_interest[6];     /* sub-domain x0 x1 y0 y1 z0 z1 of the physical axes (interior
                                indices, inclusive; all 0: none), rotated with each solve */
} pm_request;

typedef struct {
    double k[9];
    double arrival_s, admit_s, done_s;
    int sweeps[3];
    int state[3];            /* PM_SAMPLE_* of the three solves */
    double k_dom[9];         /* with a dom_interest box: the sub-domain tensor, k[3*i + j]
                                from each solve's (k13, k23, k33) (src/evaluation.cc:54-211,
                                pressure gradient over the box along the flow axis) */
} pm_response;

/* Called when a request completes and at least once a second while solves
 * run: completed requests, solves on the GPU, seconds since the start.  A
 * nonzero return stops serving (PM_E_ABORTED). */
typedef int (*pm_serve_progress_fn)(int completed, int running, double t_s, void *user);

typedef struct {
    pm_batch_params batch;
    int realtime;            /* 1: requests become visible at arrival_s; 0: all at once */
    pm_serve_progress_fn progress;  /* may be NULL */
    void *user;
    int max_solves;          /* most solves on the GPU at once (0: limited by memory only);
                                a cap trades throughput under overload for shorter waits of
                                the requests already admitted */
    int64_t small_cells;     /* requests of at most this many voxels run in a second batch
                                (the memory budget holds for both together) that is stepped, while both
                                hold solves, for about as long as each step of the first: a
                                short request then advances at the pace of the short ones,
                                not of the largest solve.  0: one batch */
} pm_serve_params;

typedef struct {
    double wall_s;           /* first arrival to last completion */
    int completed;
    double latency_p50_s, latency_p99_s, latency_mean_s, latency_max_s;
    double solves_per_hour;  /* three solves per request */
    int max_running;         /* most solves sharing the GPU at once */
    int64_t peak_bytes;      /* most device memory held by admitted solves */
    int intervals;           /* batch steps */
    double admit_s;          /* host time spent admitting requests (rotation, upload, first sweep) */
    double step_s;           /* host time spent in batch steps (GPU intervals) */
    int rejected;            /* requests answered PM_SAMPLE_REJECTED (not in the latency figures) */
} pm_serve_stats;

int pm_serve(const pm_serve_params *p, const pm_request *req, int n, pm_response *resp, pm_serve_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
