#ifndef PHASE2_STEP_WRITER_H
#define PHASE2_STEP_WRITER_H

/* Per-step result sink of the algorithms (replaces
 * include/phase2/step_writer.h:21-24 of the reference).  write() returns 0
 * or -1; a NULL writer keeps results in memory only. */

#include <stddef.h>

struct phase2_step_writer {
	void *ctx;
	int (*write)(void *ctx, size_t step_idx, _Complex double z);
};

#endif /* PHASE2_STEP_WRITER_H */
