/*
 * Circuit layer (behaviour of phase2/circ.c:42-361 of the reference; API of
 * include/phase2/circ.h).  Differences that the single-GPU register allows:
 *
 *  - circ_measure for a multidet state gathers all determinant amplitudes
 *    with one device gather and one copy, then sums a*conj(cf) on the host
 *    in determinant order (the reference's order, so the overlap is
 *    bit-identical); the reference reads one amplitude per collective
 *    call (phase2/circ.c:338-350).
 *  - the coefficient-matrix state is expanded and measured on the device
 *    (state_prep_coeff.c); no host shadow is read or written.
 */
#define LOG_SUBSYS "circ"

#include <complex.h>
#include <math.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

#include "log.h"
#include "phase2/circ.h"
#include "phase2/paulis.h"
#include "phase2/qreg_hip.h"
#include "phase2/state_prep_coeff.h"
#include "phase2_hip.h"

#include "circ_cache.h"
#include "internal.h"

struct circ_measure_plan {
	uint64_t *idx;	/* multidet indices, masked to the register */
	double *amp;	/* gathered amplitudes, 2 per determinant */
	size_t n;
};

int circ_hamil_init(struct circ_hamil *hm, uint32_t qb, size_t len)
{
	hm->terms = malloc(sizeof *hm->terms * (len ? len : 1));
	if (!hm->terms)
		return -1;
	hm->qb = qb;
	hm->len = len;
	return 0;
}

void circ_hamil_free(struct circ_hamil *hm)
{
	free(hm->terms);
	hm->terms = NULL;
	hm->len = 0;
}

/* Stable merge sort of terms.  The reference sorts with glibc qsort, which
 * is a stable merge sort in glibc 2.35 for arrays that fit in memory; a
 * stable sort gives the same order for equal keys. */
static void merge_sort(struct circ_hamil_term *t, struct circ_hamil_term *tmp,
	size_t n, term_cmp cmp)
{
	if (n < 2)
		return;
	const size_t h = n / 2;
	merge_sort(t, tmp, h, cmp);
	merge_sort(t + h, tmp, n - h, cmp);
	size_t i = 0, j = h, k = 0;
	while (i < h && j < n)
		tmp[k++] = cmp(&t[j], &t[i]) < 0 ? t[j++] : t[i++];
	while (i < h)
		tmp[k++] = t[i++];
	while (j < n)
		tmp[k++] = t[j++];
	memcpy(t, tmp, sizeof *t * n);
}

int circ_terms_sort(struct circ_hamil_term *t, size_t n, term_cmp cmp)
{
	struct circ_hamil_term *tmp = malloc(sizeof *tmp * (n ? n : 1));
	if (!tmp)
		return -1;
	merge_sort(t, tmp, n, cmp);
	free(tmp);
	return 0;
}

static int cmp_lex(const struct circ_hamil_term *a,
	const struct circ_hamil_term *b)
{
	return paulis_cmp(a->op, b->op);
}

void circ_hamil_sort_lex(struct circ_hamil *hm)
{
	if (circ_terms_sort(hm->terms, hm->len, cmp_lex) < 0)
		log_error("circ_hamil_sort_lex: out of memory");
}

int circ_muldet_init(struct circ_muldet *md, size_t len)
{
	md->dets = malloc(sizeof *md->dets * (len ? len : 1));
	if (!md->dets)
		return -1;
	md->len = len;
	return 0;
}

void circ_muldet_free(struct circ_muldet *md)
{
	free(md->dets);
	md->dets = NULL;
	md->len = 0;
}

void data_coeff_matrix_free(struct data_coeff_matrix *cm)
{
	if (!cm)
		return;
	if (cm->blocks) {
		for (size_t k = 0; k < cm->n_components; k++) {
			free((void *)cm->blocks[k].C_alpha);
			free((void *)cm->blocks[k].C_beta);
		}
		free(cm->blocks);
	}
	free((void *)cm->C_alpha);
	free((void *)cm->C_beta);
	memset(cm, 0, sizeof *cm);
}

int circ_values_init(struct circ_values *vals, size_t len)
{
	vals->z = malloc(sizeof *vals->z * (len ? len : 1));
	if (!vals->z)
		return -1;
	vals->len = len;
	return 0;
}

void circ_values_free(struct circ_values *vals)
{
	free(vals->z);
	vals->z = NULL;
	vals->len = 0;
}

static void free_payload(struct circ *ct)
{
	if (ct->stprep_kind == STPREP_MULTIDET) {
		circ_muldet_free(&ct->md);
	} else if (ct->stprep_kind == STPREP_COEFF_MATRIX) {
		data_coeff_matrix_free(&ct->cm);
		state_prep_coeff_scratch_free(&ct->sp_scratch);
	}
}

static void plan_free(void *v)
{
	struct circ_measure_plan *p = v;
	if (!p)
		return;
	free(p->idx);
	free(p->amp);
	free(p);
}

int circ_init(struct circ *ct, struct circ_hamil hm, enum stprep_kind sp_kind,
	const void *sp_data, size_t vals_len)
{
	memset(&ct->md, 0, sizeof ct->md);
	memset(&ct->cm, 0, sizeof ct->cm);
	memset(&ct->sp_scratch, 0, sizeof ct->sp_scratch);
	ct->cache = NULL;
	ct->hm = hm;
	ct->stprep_kind = sp_kind;

	switch (sp_kind) {
	case STPREP_MULTIDET:
		ct->md = *(const struct circ_muldet *)sp_data;
		break;
	case STPREP_COEFF_MATRIX:
		ct->cm = *(const struct data_coeff_matrix *)sp_data;
		if (state_prep_coeff_scratch_init(&ct->sp_scratch,
			    ct->cm.n_sites, ct->cm.n_alpha, ct->cm.n_beta) < 0) {
			log_error("circ_init: sp_scratch init failed");
			data_coeff_matrix_free(&ct->cm);
			circ_hamil_free(&ct->hm);
			return -1;
		}
		break;
	default:
		log_error("circ_init: unknown state-prep kind %d", sp_kind);
		circ_hamil_free(&ct->hm);
		return -1;
	}
	if (qreg_init(&ct->reg, ct->hm.qb) < 0) {
		log_error("circ_init: qreg_init failed");
		goto err;
	}
	ct->cache = circ_cache_init((int)ct->reg.qb_hi, (int)ct->reg.qb_lo);
	if (!ct->cache) {
		log_error("circ_init: cache_init failed");
		goto err_reg;
	}
	if (circ_values_init(&ct->vals, vals_len) < 0) {
		log_error("circ_init: values_init failed (len=%zu)", vals_len);
		goto err_cache;
	}
	return 0;

err_cache:
	circ_cache_free(ct->cache);
	ct->cache = NULL;
err_reg:
	qreg_free(&ct->reg);
err:
	free_payload(ct);
	circ_hamil_free(&ct->hm);
	return -1;
}

void circ_free(struct circ *ct)
{
	circ_values_free(&ct->vals);
	circ_hamil_free(&ct->hm);
	free_payload(ct);
	circ_cache_free(ct->cache);
	ct->cache = NULL;
	qreg_free(&ct->reg);	/* with it the overlap plan */
}

int circ_prepst(struct circ *ct)
{
	if (ct->stprep_kind == STPREP_MULTIDET) {
		/* an index beyond the register aliases silently in the
		 * reference (phase2/qreg.c:35-47); here it is an error */
		for (size_t d = 0; d < ct->md.len; d++)
			if (ct->md.dets[d].idx >= ct->reg.namp) {
				log_error("circ_prepst: determinant %zu index "
					  "%llu beyond %u qubits", d,
					(unsigned long long)ct->md.dets[d].idx,
					ct->reg.qb_lo);
				return -1;
			}
		qreg_zero(&ct->reg);
		for (size_t d = 0; d < ct->md.len; d++)
			qreg_setamp(&ct->reg, ct->md.dets[d].idx,
				ct->md.dets[d].cf);
		return qreg_error(&ct->reg);
	}
	if (ct->stprep_kind == STPREP_COEFF_MATRIX)
		return state_prep_coeff_expand_all(&ct->reg, &ct->sp_scratch,
			&ct->cm);
	return -1;
}

static void flush_rot(struct paulis hi, const struct paulis *lo,
	double *phis, size_t n, void *data)
{
	qreg_paulirot((struct qreg *)data, hi, lo, phis, n);
}

/* Forward or reverse sweep, phi_k = omega * cf_k, batched by the cache
 * (phase2/circ.c:256-292). */
static int step_generic(struct circ *ct, const struct circ_hamil *hm,
	double omega, bool reverse)
{
	for (size_t k = 0; k < hm->len; k++) {
		const struct circ_hamil_term *t =
			&hm->terms[reverse ? hm->len - 1 - k : k];
		const double phi = omega * t->cf;
		if (circ_cache_insert(ct->cache, t->op, phi) == 0)
			continue;
		circ_cache_flush(ct->cache, flush_rot, &ct->reg);
		if (circ_cache_insert(ct->cache, t->op, phi) < 0)
			return -1;
	}
	circ_cache_flush(ct->cache, flush_rot, &ct->reg);
	return qreg_error(&ct->reg);
}

int circ_step(struct circ *ct, const struct circ_hamil *hm, double omega)
{
	return step_generic(ct, hm, omega, false);
}

int circ_step_reverse(struct circ *ct, const struct circ_hamil *hm,
	double omega)
{
	return step_generic(ct, hm, omega, true);
}

/* The gather plan of the multidet overlap, built on the first measure and
 * kept with the register. */
static struct circ_measure_plan *plan_multidet(struct circ *ct)
{
	struct circ_measure_plan *p = qreg_hip_aux(&ct->reg);
	if (p)
		return p;
	const size_t n = ct->md.len;
	for (size_t d = 0; d < n; d++)
		if (ct->md.dets[d].idx >= ct->reg.namp) {
			log_error("circ_measure: determinant %zu beyond the "
				  "register", d);
			return NULL;
		}
	p = calloc(1, sizeof *p);
	if (!p)
		return NULL;
	p->idx = malloc(sizeof *p->idx * (n ? n : 1));
	p->amp = malloc(sizeof *p->amp * 2 * (n ? n : 1));
	if (!p->idx || !p->amp || qreg_hip_set_aux(&ct->reg, p,
		    plan_free) < 0) {
		plan_free(p);
		return NULL;
	}
	for (size_t d = 0; d < n; d++)
		p->idx[d] = ct->md.dets[d].idx;
	p->n = n;
	return p;
}

_Complex double circ_measure(struct circ *ct)
{
	if (ct->stprep_kind == STPREP_MULTIDET) {
		struct circ_measure_plan *p = plan_multidet(ct);
		if (!p) {
			log_error("circ_measure: no overlap plan");
			return 0.0;
		}
		if (ph2hip_gather(qreg_hip_device_ptr(&ct->reg),
			    ct->reg.qb_lo, p->idx, p->n, p->amp) < 0) {
			log_error("circ_measure: %s", ph2hip_strerror());
			return 0.0;
		}
		_Complex double pr = 0.0;
		for (size_t d = 0; d < p->n; d++) {
			const _Complex double a =
				CMPLX(p->amp[2 * d], p->amp[2 * d + 1]);
			pr += a * conj(ct->md.dets[d].cf);
		}
		return pr;
	}
	if (ct->stprep_kind == STPREP_COEFF_MATRIX) {
		const struct data_coeff_matrix *cm = &ct->cm;
		_Complex double pr = 0.0;
		if (cm->n_components == 0) {
			_Complex double z = 0.0;
			if (state_prep_coeff_inner(&ct->reg, &ct->sp_scratch,
				    cm->C_alpha,
				    cm->closed_shell ? NULL : cm->C_beta, 1.0,
				    cm->tapered, &z) < 0)
				return 0.0;
			return z;
		}
		for (size_t k = 0; k < cm->n_components; k++) {
			const struct data_coeff_block *b = &cm->blocks[k];
			_Complex double z = 0.0;
			if (state_prep_coeff_inner(&ct->reg, &ct->sp_scratch,
				    b->C_alpha,
				    cm->closed_shell ? NULL : b->C_beta, b->cf,
				    cm->tapered, &z) < 0)
				z = 0.0;
			pr += z;
		}
		return pr;
	}
	return 0.0;
}
