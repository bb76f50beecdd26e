/*
 * Composite integrator driver (behaviour of circ/cmpsit.c:30-235 of the
 * reference).  Init orders a copy of H by |c| descending (stable), keeps
 * the first `length` terms as the deterministic part (lex-sorted) and
 * builds the CDF over the rest.  Per sample: prepare; per step a forward
 * half-sweep of one composite draw and a reverse half-sweep of a fresh
 * draw, both at omega = 0.5; then measure.  Deterministic terms carry
 * cf*angle_det, sampled ones signof(cf)*angle_rand.
 */
#define LOG_SUBSYS "cmpsit"

#include <math.h>
#include <stdbool.h>
#include <stdint.h>
#include <string.h>

#include "log.h"
#include "circ/cmpsit.h"
#include "internal.h"

static uint64_t SEED = UINT64_C(0xafb424901446f21f);

static int cmp_abs_desc(const struct circ_hamil_term *a,
	const struct circ_hamil_term *b)
{
	const double x = fabs(a->cf), y = fabs(b->cf);
	return x > y ? -1 : (x < y ? 1 : 0);
}

void cmpsit_ranct_free(struct cmpsit_ranct *r)
{
	circ_hamil_free(&r->hm_det);
	circ_hamil_free(&r->hm_ran);
	circ_hamil_free(&r->hm_smpl);
	prob_cdf_free(&r->cdf);
}

int cmpsit_ranct_init(struct cmpsit_ranct *r, const struct circ_hamil *hm,
	const struct cmpsit_data *dt)
{
	memset(r, 0, sizeof *r);
	if (dt->length > hm->len) {
		log_error("cmpsit: length %zu > %zu terms", dt->length, hm->len);
		return -1;
	}
	struct circ_hamil tmp;
	if (circ_hamil_init(&tmp, hm->qb, hm->len) < 0)
		return -1;
	memcpy(tmp.terms, hm->terms, sizeof *tmp.terms * hm->len);
	if (circ_terms_sort(tmp.terms, tmp.len, cmp_abs_desc) < 0)
		goto err;
	const size_t nran = hm->len - dt->length;
	if (circ_hamil_init(&r->hm_det, hm->qb, dt->length) < 0 ||
	    circ_hamil_init(&r->hm_ran, hm->qb, nran) < 0 ||
	    circ_hamil_init(&r->hm_smpl, hm->qb, dt->length + dt->depth) < 0)
		goto err;
	memcpy(r->hm_det.terms, tmp.terms, sizeof *tmp.terms * dt->length);
	circ_hamil_sort_lex(&r->hm_det);
	memcpy(r->hm_ran.terms, tmp.terms + dt->length,
		sizeof *tmp.terms * nran);
	if (prob_cdf_init(&r->cdf, nran) < 0)
		goto err;
	prob_cdf_from_array_strided(&r->cdf, &r->hm_ran.terms[0].cf,
		sizeof r->hm_ran.terms[0], &r->lambda_r);
	r->depth = dt->depth;
	r->angle_rand = dt->angle_rand;
	circ_hamil_free(&tmp);
	return 0;
err:
	circ_hamil_free(&tmp);
	cmpsit_ranct_free(r);
	return -1;
}

int cmpsit_init(struct cmpsit *cp, const struct cmpsit_data *dt,
	struct circ_hamil hm, enum stprep_kind sp_kind, const void *sp_data,
	struct phase2_step_writer *sw)
{
	if (circ_init(&cp->ct, hm, sp_kind, sp_data, dt->samples) < 0)
		return -1;
	cp->dt = *dt;
	cp->sw = sw;
	if (cmpsit_ranct_init(&cp->ranct, &cp->ct.hm, dt) < 0) {
		circ_free(&cp->ct);
		return -1;
	}
	if (cp->dt.seed != 0)
		SEED = cp->dt.seed;
	else
		cp->dt.seed = SEED;
	xoshiro256ss_init(&cp->rng, cp->dt.seed);
	return 0;
}

void cmpsit_free(struct cmpsit *cp)
{
	cmpsit_ranct_free(&cp->ranct);
	circ_free(&cp->ct);
}

/* One composite draw into hm_smpl: deterministic part, then depth samples
 * (circ/cmpsit.c:151-177). */
static void draw(struct cmpsit *cp)
{
	struct cmpsit_ranct *r = &cp->ranct;
	const size_t len = cp->dt.length;
	for (size_t k = 0; k < len; k++) {
		r->hm_smpl.terms[k].op = r->hm_det.terms[k].op;
		r->hm_smpl.terms[k].cf = r->hm_det.terms[k].cf * cp->dt.angle_det;
	}
	for (size_t k = len; k < len + r->depth; k++) {
		const double u = xoshiro256ss_dbl01(&cp->rng);
		const size_t i = prob_cdf_inverse(&r->cdf, u);
		r->hm_smpl.terms[k].op = r->hm_ran.terms[i].op;
		r->hm_smpl.terms[k].cf = signof(r->hm_ran.terms[i].cf) *
			r->angle_rand;
	}
}

int cmpsit_simul(struct cmpsit *cp)
{
	struct circ *ct = &cp->ct;
	for (size_t smp = 0; smp < ct->vals.len; smp++) {
		if (circ_prepst(ct) < 0)
			return -1;
		for (size_t s = 0; s < cp->dt.steps; s++) {
			for (int rev = 0; rev < 2; rev++) {
				draw(cp);
				const int rc = rev
					? circ_step_reverse(ct, &cp->ranct.hm_smpl,
						0.5)
					: circ_step(ct, &cp->ranct.hm_smpl, 0.5);
				if (rc < 0) {
					log_error("cmpsit_simul: half-step failed");
					return -1;
				}
			}
		}
		ct->vals.z[smp] = circ_measure(ct);
		if (cp->sw && cp->sw->write(cp->sw->ctx, smp, ct->vals.z[smp]) <
			0) {
			log_error("cmpsit_simul: write_step %zu failed", smp);
			return -1;
		}
	}
	return qreg_error(&ct->reg);
}
