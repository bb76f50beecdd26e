/*
 * Product-formula drivers of include/circ/trott.h and include/circ/trott2.h
 * (behaviour of circ/trott.c:21-65 and circ/trott2.c:21-74 of the
 * reference).  Both orders share one core:
 *
 *   adopt   circ_init takes the Hamiltonian and the state payload, one
 *           result slot per step; the terms are put in lexicographic order
 *           (the order the sweeps apply them in);
 *   evolve  prepare the initial state once, then per step advance the
 *           register by one step of the formula, take the overlap with the
 *           initial state and hand it to the step writer, if any.
 *
 * A first-order step is one forward sweep at delta; a second-order
 * (Strang) step is a forward sweep at delta/2 followed by the reverse sweep
 * at delta/2, which makes the step symmetric.
 */
#define LOG_SUBSYS "trotter"

#include "log.h"
#include "circ/trott.h"
#include "circ/trott2.h"

static int adopt(struct circ *ct, struct circ_hamil hm, size_t steps,
	enum stprep_kind sp_kind, const void *sp_data, const char *who)
{
	if (circ_init(ct, hm, sp_kind, sp_data, steps) < 0) {
		log_error("%s: cannot set up the circuit (%zu steps)", who,
			steps);
		return -1;
	}
	circ_hamil_sort_lex(&ct->hm);
	return 0;
}

/* Orders of the product formula the core knows. */
enum formula { LIE_TROTTER = 1, STRANG = 2 };

static int advance(struct circ *ct, enum formula f, double delta)
{
	if (f == LIE_TROTTER)
		return circ_step(ct, &ct->hm, delta);
	const double half = delta / 2.0;
	if (circ_step(ct, &ct->hm, half) < 0)
		return -1;
	return circ_step_reverse(ct, &ct->hm, half);
}

static int evolve(struct circ *ct, enum formula f, double delta,
	struct phase2_step_writer *sw, const char *who)
{
	if (circ_prepst(ct) < 0) {
		log_error("%s: initial state not prepared", who);
		return -1;
	}
	_Complex double *ov = ct->vals.z;
	for (size_t n = 0; n < ct->vals.len; n++) {
		log_debug("%s: step %zu of %zu", who, n + 1, ct->vals.len);
		if (advance(ct, f, delta) < 0) {
			log_error("%s: sweep of step %zu did not complete", who,
				n);
			return -1;
		}
		ov[n] = circ_measure(ct);
		if (sw && sw->write(sw->ctx, n, ov[n]) < 0) {
			log_error("%s: step writer refused step %zu", who, n);
			return -1;
		}
	}
	/* a device call inside a void register call sets the sticky flag */
	return qreg_error(&ct->reg);
}

int trott_init(struct trott *tt, const struct trott_data *dt,
	struct circ_hamil hm, enum stprep_kind sp_kind, const void *sp_data,
	struct phase2_step_writer *sw)
{
	if (adopt(&tt->ct, hm, dt->steps, sp_kind, sp_data, "trott") < 0)
		return -1;
	tt->dt = *dt;
	tt->sw = sw;
	return 0;
}

void trott_free(struct trott *tt)
{
	circ_free(&tt->ct);
}

int trott_simul(struct trott *tt)
{
	return evolve(&tt->ct, LIE_TROTTER, tt->dt.delta, tt->sw, "trott");
}

int trott2_init(struct trott2 *t2, const struct trott2_data *dt,
	struct circ_hamil hm, enum stprep_kind sp_kind, const void *sp_data,
	struct phase2_step_writer *sw)
{
	if (adopt(&t2->ct, hm, dt->steps, sp_kind, sp_data, "trott2") < 0)
		return -1;
	t2->dt = *dt;
	t2->sw = sw;
	return 0;
}

void trott2_free(struct trott2 *t2)
{
	circ_free(&t2->ct);
}

int trott2_simul(struct trott2 *t2)
{
	return evolve(&t2->ct, STRANG, t2->dt.delta, t2->sw, "trott2");
}
