#ifndef TROTT2_H
#define TROTT2_H

/* Symmetric (Strang) 2nd-order Trotter driver (replaces
 * include/circ/trott2.h:33-41 of the reference): per step a forward sweep
 * at delta/2 then a reverse sweep at delta/2. */

#include "phase2/circ.h"
#include "phase2/step_writer.h"

struct trott2_data {
	size_t steps;
	double delta;
};

struct trott2 {
	struct circ ct;
	struct trott2_data dt;
	struct phase2_step_writer *sw;
};

int trott2_init(struct trott2 *t2, const struct trott2_data *dt,
	struct circ_hamil hm, enum stprep_kind sp_kind, const void *sp_data,
	struct phase2_step_writer *sw);
void trott2_free(struct trott2 *t2);
int trott2_simul(struct trott2 *t2);

#endif /* TROTT2_H */
