/*
 * Strang 2nd-order Trotter driver (behaviour of circ/trott2.c:21-74 of the
 * reference): per step a forward sweep at delta/2, a reverse sweep at
 * delta/2, then measure.
 */
#define LOG_SUBSYS "trott2"

#include "log.h"
#include "circ/trott2.h"

int trott2_init(struct trott2 *t2, const struct trott2_data *dt,
	struct circ_hamil hm, enum stprep_kind sp_kind, const void *sp_data,
	struct phase2_step_writer *sw)
{
	if (circ_init(&t2->ct, hm, sp_kind, sp_data, dt->steps) < 0) {
		log_error("trott2_init: circ_init failed");
		return -1;
	}
	t2->dt = *dt;
	t2->sw = sw;
	circ_hamil_sort_lex(&t2->ct.hm);
	return 0;
}

void trott2_free(struct trott2 *t2)
{
	circ_free(&t2->ct);
}

int trott2_simul(struct trott2 *t2)
{
	struct circ *ct = &t2->ct;
	const double half = t2->dt.delta / 2.0;
	if (circ_prepst(ct) < 0) {
		log_error("trott2_simul: state preparation failed");
		return -1;
	}
	for (size_t s = 0; s < ct->vals.len; s++) {
		if (circ_step(ct, &ct->hm, half) < 0 ||
		    circ_step_reverse(ct, &ct->hm, half) < 0) {
			log_error("trott2_simul: step %zu failed", s);
			return -1;
		}
		ct->vals.z[s] = circ_measure(ct);
		if (t2->sw && t2->sw->write(t2->sw->ctx, s, ct->vals.z[s]) < 0) {
			log_error("trott2_simul: write_step %zu failed", s);
			return -1;
		}
	}
	return qreg_error(&ct->reg);
}
