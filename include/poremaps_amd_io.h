/*
 * poremaps_amd_io.h -- the reference's file formats, host side (no GPU).
 *
 * Replaces src/read_problem_params.cc:50-156 (ordered .inp), the host half of
 * src/output.cc:53-230 (interior raw dumps, log rows) and the byte read of
 * src/geometry.cc:58-126.
 */
#ifndef POREMAPS_AMD_IO_H
#define POREMAPS_AMD_IO_H

#include <stdint.h>

#include "poremaps_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Values of the .inp that are not solver parameters. */
typedef struct {
    int dom_decomposition[3];   /* the blocks of the reference mode's iterates and the
                                   domain_decomp_ file (pm_rank_map) */
    char geometry_file[400];    /* src/constants.h:46 file_name_char */
    char log_file[400];
    int write_output[4];        /* vels, press, nhood, domain decomposition */
} pm_inp_extra;

const char *pm_io_last_error(void);

/* Parse the 14 ordered lines (src/read_problem_params.cc:75-90); keys must
 * appear in the reference's order.  Sanity warnings go to stderr as in the
 * reference (:96-104, :144-146). */
int pm_read_inp(const char *path, pm_params *p, pm_inp_extra *x);

/* Read exactly n geometry bytes. */
int pm_read_raw_u8(const char *path, uint8_t *out, int64_t n);

/* Write n values at offset 0 of path (opened read-write, created if missing,
 * not truncated: src/output.cc:79).  rescale != 0 multiplies by scale. */
int pm_write_raw_f64(const char *path, const double *v, int64_t n, double scale, int rescale);
int pm_write_raw_i32(const char *path, const int32_t *v, int64_t n);

/* The rank of every voxel under dom_decomposition dims (x, y, z; 0 counts as
 * 1), nx*ny*nz int32 values, x fastest: the contents of the reference's
 * domain_decomp_ file (src/output.cc:156-190 with src/geometry.cc:129-171
 * and MPICH's Cartesian rank order, z fastest). */
int pm_rank_map(int nx, int ny, int nz, const int dims[3], int32_t *out);

/* Append one log row; firstline != 0 writes the header first. */
int pm_write_log_row(const char *path, const pm_log_row *row, int firstline);

#ifdef __cplusplus
}
#endif
#endif
