#ifndef PHASE2_H5_H
#define PHASE2_H5_H

/*
 * Read-back of ph2run's per-step output for callers without HDF5
 * bindings (libphase2_h5.so, beside the simul.h5 layer of
 * include/ph2run/data.h).  No counterpart in the reference.
 */

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The rows of /<grp>/values in the simul.h5 file at path: up to cap
 * (re, im) pairs into out and the row count into *nrows.  0, or -1 (no
 * such file, group or dataset, or more than cap rows). */
int ph2_h5_read_values(const char *path, const char *grp, double *out,
	size_t cap, size_t *nrows);

#ifdef __cplusplus
}
#endif

#endif /* PHASE2_H5_H */
