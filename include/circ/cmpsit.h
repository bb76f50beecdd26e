#ifndef CMPSIT_H
#define CMPSIT_H

/* Composite (partially randomised) symmetric Trotter driver (replaces
 * include/circ/cmpsit.h:61-71 of the reference).  The top `length` terms by
 * |c| (lex-sorted) are applied deterministically at angle_det; the rest
 * are sampled `depth` times per half-step at angle_rand.  Each step is a
 * forward half-sweep of one draw and a reverse half-sweep of a fresh draw,
 * both at omega = 0.5.  seed 0 behaves as in qdrift.h. */

#include <stddef.h>

#include "phase2.h"
#include "phase2/step_writer.h"
#include "xoshiro256ss.h"

struct cmpsit_data {
	size_t samples;
	size_t steps;
	size_t length;
	size_t depth;
	double angle_det;
	double angle_rand;
	uint64_t seed;
};

struct cmpsit_ranct {
	struct circ_hamil hm_det, hm_ran, hm_smpl;
	struct prob_cdf cdf;
	double lambda_r;
	size_t depth;
	double angle_rand;
};

struct cmpsit {
	struct circ ct;
	struct cmpsit_data dt;
	struct cmpsit_ranct ranct;
	struct xoshiro256ss rng;
	struct phase2_step_writer *sw;
};

int cmpsit_init(struct cmpsit *cp, const struct cmpsit_data *dt,
	struct circ_hamil hm, enum stprep_kind sp_kind, const void *sp_data,
	struct phase2_step_writer *sw);
void cmpsit_free(struct cmpsit *cp);
int cmpsit_simul(struct cmpsit *cp);

#endif /* CMPSIT_H */
