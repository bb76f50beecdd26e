/*
 * qDRIFT driver (behaviour of circ/qdrift.c:30-133 of the reference):
 * CDF over |c| in input order; one xoshiro256** stream continuing across
 * samples; per sample prepare, draw `depth` terms with cf = signof(c),
 * sweep at omega = asin(step_size), measure.  A zero seed takes the seed
 * of the previous instance (initially the reference's constant).
 */
#define LOG_SUBSYS "qdrift"

#include <math.h>
#include <stdint.h>

#include "log.h"
#include "circ/qdrift.h"
#include "internal.h"

static uint64_t SEED = UINT64_C(0xeccd9dcc749fcdca);

int qdrift_init(struct qdrift *qd, const struct qdrift_data *dt,
	struct circ_hamil hm, enum stprep_kind sp_kind, const void *sp_data,
	struct phase2_step_writer *sw)
{
	if (circ_init(&qd->ct, hm, sp_kind, sp_data, dt->samples) < 0)
		return -1;
	qd->dt = *dt;
	qd->sw = sw;
	qd->ranct.hm_ran.terms = NULL;
	qd->ranct.cdf.y = NULL;
	if (prob_cdf_init(&qd->ranct.cdf, qd->ct.hm.len) < 0 ||
	    circ_hamil_init(&qd->ranct.hm_ran, qd->ct.hm.qb, dt->depth) < 0) {
		log_error("qdrift_init: alloc failed");
		prob_cdf_free(&qd->ranct.cdf);
		circ_free(&qd->ct);
		return -1;
	}
	prob_cdf_from_array_strided(&qd->ranct.cdf, &qd->ct.hm.terms[0].cf,
		sizeof qd->ct.hm.terms[0], NULL);
	if (qd->dt.seed != 0)
		SEED = qd->dt.seed;
	else
		qd->dt.seed = SEED;
	xoshiro256ss_init(&qd->rng, qd->dt.seed);
	return 0;
}

void qdrift_free(struct qdrift *qd)
{
	circ_hamil_free(&qd->ranct.hm_ran);
	prob_cdf_free(&qd->ranct.cdf);
	circ_free(&qd->ct);
}

static void draw(struct qdrift *qd)
{
	struct circ_hamil *r = &qd->ranct.hm_ran;
	for (size_t k = 0; k < r->len; k++) {
		const double u = xoshiro256ss_dbl01(&qd->rng);
		const size_t i = prob_cdf_inverse(&qd->ranct.cdf, u);
		r->terms[k].cf = signof(qd->ct.hm.terms[i].cf);
		r->terms[k].op = qd->ct.hm.terms[i].op;
	}
}

int qdrift_simul(struct qdrift *qd)
{
	struct circ *ct = &qd->ct;
	const double omega = asin(qd->dt.step_size);
	for (size_t s = 0; s < ct->vals.len; s++) {
		if (circ_prepst(ct) < 0)
			return -1;
		draw(qd);
		if (circ_step(ct, &qd->ranct.hm_ran, omega) < 0) {
			log_error("qdrift_simul: sample %zu failed", s);
			return -1;
		}
		ct->vals.z[s] = circ_measure(ct);
		if (qd->sw && qd->sw->write(qd->sw->ctx, s, ct->vals.z[s]) < 0) {
			log_error("qdrift_simul: write_step %zu failed", s);
			return -1;
		}
	}
	return qreg_error(&ct->reg);
}
