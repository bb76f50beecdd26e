#ifndef TROTT_H
#define TROTT_H

/* 1st-order Lie-Trotter driver (replaces include/circ/trott.h:32-40 of the
 * reference).  trott_init lex-sorts the adopted Hamiltonian; each step
 * sweeps it forward at omega = delta and records circ_measure. */

#include "phase2/circ.h"
#include "phase2/step_writer.h"

struct trott_data {
	size_t steps;
	double delta;
};

struct trott {
	struct circ ct;
	struct trott_data dt;
	struct phase2_step_writer *sw;
};

int trott_init(struct trott *tt, const struct trott_data *dt,
	struct circ_hamil hm, enum stprep_kind sp_kind, const void *sp_data,
	struct phase2_step_writer *sw);
void trott_free(struct trott *tt);
int trott_simul(struct trott *tt);

#endif /* TROTT_H */
