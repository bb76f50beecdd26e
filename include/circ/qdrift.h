#ifndef QDRIFT_H
#define QDRIFT_H

/* qDRIFT driver (replaces include/circ/qdrift.h:52-61 of the reference).
 * Per sample: prepare, draw `depth` terms from the |c| CDF (input order) with
 * one xoshiro256** stream that continues across samples, apply each as
 * exp(i asin(step_size) sign(c) P), measure.  seed 0 means "the seed of the
 * previous qdrift instance in this process" (initially a fixed constant),
 * as in the reference. */

#include <stddef.h>

#include "phase2/circ.h"
#include "phase2/step_writer.h"
#include "prob.h"
#include "xoshiro256ss.h"

struct qdrift_data {
	size_t samples;
	size_t depth;
	double step_size;
	uint64_t seed;
};

struct qdrift_ranct {
	struct circ_hamil hm_ran;
	struct prob_cdf cdf;
};

struct qdrift {
	struct circ ct;
	struct qdrift_data dt;
	struct qdrift_ranct ranct;
	struct xoshiro256ss rng;
	struct phase2_step_writer *sw;
};

int qdrift_init(struct qdrift *qd, const struct qdrift_data *dt,
	struct circ_hamil hm, enum stprep_kind sp_kind, const void *sp_data,
	struct phase2_step_writer *sw);
void qdrift_free(struct qdrift *qd);
int qdrift_simul(struct qdrift *qd);

#endif /* QDRIFT_H */
