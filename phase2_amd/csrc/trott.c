/*
 * 1st-order Trotter driver (behaviour of circ/trott.c:21-65 of the
 * reference): lex-sort at init; per step one forward sweep at delta, then
 * measure and hand the overlap to the step writer.
 */
#define LOG_SUBSYS "trott"

#include "log.h"
#include "circ/trott.h"

int trott_init(struct trott *tt, const struct trott_data *dt,
	struct circ_hamil hm, enum stprep_kind sp_kind, const void *sp_data,
	struct phase2_step_writer *sw)
{
	if (circ_init(&tt->ct, hm, sp_kind, sp_data, dt->steps) < 0) {
		log_error("trott_init: circ_init failed");
		return -1;
	}
	tt->dt = *dt;
	tt->sw = sw;
	circ_hamil_sort_lex(&tt->ct.hm);
	return 0;
}

void trott_free(struct trott *tt)
{
	circ_free(&tt->ct);
}

int trott_simul(struct trott *tt)
{
	struct circ *ct = &tt->ct;
	if (circ_prepst(ct) < 0) {
		log_error("trott_simul: state preparation failed");
		return -1;
	}
	for (size_t s = 0; s < ct->vals.len; s++) {
		if (circ_step(ct, &ct->hm, tt->dt.delta) < 0) {
			log_error("trott_simul: step %zu failed", s);
			return -1;
		}
		ct->vals.z[s] = circ_measure(ct);
		if (tt->sw && tt->sw->write(tt->sw->ctx, s, ct->vals.z[s]) < 0) {
			log_error("trott_simul: write_step %zu failed", s);
			return -1;
		}
	}
	return qreg_error(&ct->reg);
}
