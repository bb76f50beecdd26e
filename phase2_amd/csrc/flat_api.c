/*
 * Flat-argument entry points (include/phase2_flat.h) over the phase2 API.
 */
#define _GNU_SOURCE /* sincos */
#define LOG_SUBSYS "flat"

#include <complex.h>
#include <stdlib.h>
#include <string.h>

#include "log.h"
#include "phase2.h"
#include "phase2/qreg_hip.h"
#include "phase2_flat.h"
#include "phase2_hip.h"
#include "circ/cmpsit.h"
#include "circ/qdrift.h"
#include "circ/trott.h"
#include "circ/trott2.h"
#include "internal.h"

#define FLAT_SEED UINT64_C(0x3c6ef372fe94f82b)

int ph2_init(void)
{
	struct world_info wd;
	if (world_info(&wd) == WORLD_READY)
		return 0;
	return world_init(NULL, NULL, FLAT_SEED) == WORLD_READY ? 0 : -1;
}

const char *ph2_strerror(void)
{
	return ph2hip_strerror();
}

void ph2_set_default(int mode, uint32_t tile_bits, uint32_t low_bits)
{
	qreg_hip_set_default(mode, tile_bits, low_bits);
}

static int make_hamil(struct circ_hamil *hm, uint32_t nqb, size_t n,
	const uint64_t *xm, const uint64_t *zm, const double *cf)
{
	if (circ_hamil_init(hm, nqb, n) < 0)
		return -1;
	for (size_t k = 0; k < n; k++) {
		hm->terms[k].cf = cf[k];
		hm->terms[k].op.pak[0] = xm[k];
		hm->terms[k].op.pak[1] = zm[k];
	}
	return 0;
}

static double *dup(const double *src, size_t n)
{
	double *d = malloc(sizeof *d * (n ? n : 1));
	if (d && n)
		memcpy(d, src, sizeof *d * n);
	return d;
}

/* Payload for circ_init (which adopts it): copies of the caller's arrays. */
static int make_prep(const struct ph2_prep *p, uint32_t nqb,
	struct circ_muldet *md, struct data_coeff_matrix *cm,
	enum stprep_kind *kind, const void **data)
{
	if (p->kind == 1) {
		if (circ_muldet_init(md, p->ndet) < 0)
			return -1;
		for (size_t d = 0; d < p->ndet; d++) {
			md->dets[d].idx = p->didx[d];
			md->dets[d].cf = CMPLX(p->dcf[2 * d], p->dcf[2 * d + 1]);
		}
		*kind = STPREP_MULTIDET;
		*data = md;
		return 0;
	}
	if (p->kind != 2)
		return -1;
	memset(cm, 0, sizeof *cm);
	cm->nqb = nqb;
	cm->n_sites = p->n_sites;
	cm->n_alpha = p->n_alpha;
	cm->n_beta = p->n_beta;
	cm->closed_shell = p->closed_shell;
	cm->tapered = p->tapered;
	cm->C_alpha = dup(p->Ca, (size_t)p->n_sites * p->n_alpha);
	if (!cm->C_alpha)
		return -1;
	if (!p->closed_shell) {
		cm->C_beta = dup(p->Cb, (size_t)p->n_sites * p->n_beta);
		if (!cm->C_beta)
			return -1;
	}
	*kind = STPREP_COEFF_MATRIX;
	*data = cm;
	return 0;
}

static void results(const struct circ *ct, double *out, double *state)
{
	for (size_t i = 0; i < ct->vals.len; i++) {
		out[2 * i] = creal(ct->vals.z[i]);
		out[2 * i + 1] = cimag(ct->vals.z[i]);
	}
	if (state)
		ph2hip_d2h(state, qreg_hip_device_ptr((struct qreg *)&ct->reg),
			ct->reg.namp * 16);
}

#define SETUP()                                                           \
	struct circ_hamil hm;                                             \
	struct circ_muldet md;                                            \
	struct data_coeff_matrix cm;                                      \
	enum stprep_kind kind;                                            \
	const void *data;                                                 \
	if (ph2_init() < 0 || make_hamil(&hm, nqb, nterms, xm, zm, cf) < 0) \
		return -1;                                                \
	if (make_prep(prep, nqb, &md, &cm, &kind, &data) < 0) {           \
		circ_hamil_free(&hm);                                     \
		return -1;                                                \
	}

int ph2_trott(int algo, uint32_t nqb, size_t nterms, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct ph2_prep *prep,
	double delta, size_t steps, double *out, double *state)
{
	SETUP();
	int rc;
	if (algo == 1) {
		struct trott tt;
		struct trott_data dt = { .steps = steps, .delta = delta };
		if (trott_init(&tt, &dt, hm, kind, data, NULL) < 0)
			return -1;
		rc = trott_simul(&tt);
		results(&tt.ct, out, state);
		trott_free(&tt);
	} else {
		struct trott2 t2;
		struct trott2_data dt = { .steps = steps, .delta = delta };
		if (trott2_init(&t2, &dt, hm, kind, data, NULL) < 0)
			return -1;
		rc = trott2_simul(&t2);
		results(&t2.ct, out, state);
		trott2_free(&t2);
	}
	return rc;
}

int ph2_qdrift(uint32_t nqb, size_t nterms, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct ph2_prep *prep,
	size_t samples, size_t depth, double step_size, uint64_t seed,
	double *out, double *state)
{
	SETUP();
	struct qdrift qd;
	struct qdrift_data dt = { .samples = samples, .depth = depth,
		.step_size = step_size, .seed = seed };
	if (qdrift_init(&qd, &dt, hm, kind, data, NULL) < 0)
		return -1;
	const int rc = qdrift_simul(&qd);
	results(&qd.ct, out, state);
	qdrift_free(&qd);
	return rc;
}

int ph2_cmpsit(uint32_t nqb, size_t nterms, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct ph2_prep *prep,
	size_t samples, size_t steps, size_t length, size_t depth,
	double angle_det, double angle_rand, uint64_t seed, double *out,
	double *state)
{
	SETUP();
	struct cmpsit cp;
	struct cmpsit_data dt = { .samples = samples, .steps = steps,
		.length = length, .depth = depth, .angle_det = angle_det,
		.angle_rand = angle_rand, .seed = seed };
	if (cmpsit_init(&cp, &dt, hm, kind, data, NULL) < 0)
		return -1;
	const int rc = cmpsit_simul(&cp);
	results(&cp.ct, out, state);
	cmpsit_free(&cp);
	return rc;
}

int ph2_rotate(double *state, uint32_t nqb, size_t n, const uint64_t *xm,
	const uint64_t *zm, const double *phi, int mode, uint32_t tile_bits,
	uint32_t low_bits)
{
	if (ph2_init() < 0)
		return -1;
	struct qreg reg;
	if (qreg_init(&reg, nqb) < 0)
		return -1;
	qreg_hip_configure(&reg, mode, tile_bits, low_bits);
	void *d = qreg_hip_device_ptr(&reg);
	int rc = ph2hip_h2d(d, state, reg.namp * 16);
	struct paulis codes[1024];
	const struct paulis hi = { .pak = { 0, 0 } };
	for (size_t k0 = 0; rc == 0 && k0 < n; k0 += 1024) {
		const size_t nb = n - k0 < 1024 ? n - k0 : 1024;
		for (size_t k = 0; k < nb; k++) {
			codes[k].pak[0] = xm[k0 + k];
			codes[k].pak[1] = zm[k0 + k];
		}
		qreg_paulirot(&reg, hi, codes, phi + k0, nb);
		rc = qreg_error(&reg);
	}
	if (rc == 0)
		rc = ph2hip_d2h(state, d, reg.namp * 16);
	qreg_free(&reg);
	return rc;
}

int ph2_expand(uint32_t nqb, uint32_t n_sites, uint32_t n_alpha,
	uint32_t n_beta, int closed_shell, int tapered, const double *Ca,
	const double *Cb, double *state)
{
	if (ph2_init() < 0)
		return -1;
	struct qreg reg;
	struct state_prep_coeff_scratch sc;
	if (qreg_init(&reg, nqb) < 0)
		return -1;
	if (state_prep_coeff_scratch_init(&sc, n_sites, n_alpha, n_beta) < 0) {
		qreg_free(&reg);
		return -1;
	}
	struct data_coeff_matrix cm;
	memset(&cm, 0, sizeof cm);
	cm.nqb = nqb;
	cm.n_sites = n_sites;
	cm.n_alpha = n_alpha;
	cm.n_beta = n_beta;
	cm.closed_shell = closed_shell;
	cm.tapered = tapered;
	cm.C_alpha = Ca;
	cm.C_beta = closed_shell ? NULL : Cb;
	int rc = state_prep_coeff_expand_all(&reg, &sc, &cm);
	if (rc == 0)
		rc = ph2hip_d2h(state, qreg_hip_device_ptr(&reg), reg.namp * 16);
	state_prep_coeff_scratch_free(&sc);
	qreg_free(&reg);
	return rc;
}

int ph2_inner(uint32_t nqb, uint32_t n_sites, uint32_t n_alpha,
	uint32_t n_beta, int closed_shell, int tapered, const double *Ca,
	const double *Cb, const double *state, double *out)
{
	if (ph2_init() < 0)
		return -1;
	struct qreg reg;
	struct state_prep_coeff_scratch sc;
	if (qreg_init(&reg, nqb) < 0)
		return -1;
	if (state_prep_coeff_scratch_init(&sc, n_sites, n_alpha, n_beta) < 0) {
		qreg_free(&reg);
		return -1;
	}
	int rc = ph2hip_h2d(qreg_hip_device_ptr(&reg), state, reg.namp * 16);
	_Complex double z = 0.0;
	if (rc == 0)
		rc = state_prep_coeff_inner(&reg, &sc, Ca,
			closed_shell ? NULL : Cb, 1.0, tapered, &z);
	out[0] = creal(z);
	out[1] = cimag(z);
	state_prep_coeff_scratch_free(&sc);
	qreg_free(&reg);
	return rc;
}

static int cmp_lex_terms(const struct circ_hamil_term *a,
	const struct circ_hamil_term *b)
{
	return paulis_cmp(a->op, b->op);
}

int ph2_sort_lex(size_t n, const uint64_t *xm, const uint64_t *zm,
	uint64_t *perm)
{
	struct circ_hamil_term *t = malloc(sizeof *t * (n ? n : 1));
	if (!t)
		return -1;
	for (size_t k = 0; k < n; k++) {
		t[k].cf = (double)k;
		t[k].op.pak[0] = xm[k];
		t[k].op.pak[1] = zm[k];
	}
	const int rc = circ_terms_sort(t, n, cmp_lex_terms);
	for (size_t k = 0; k < n; k++)
		perm[k] = (uint64_t)t[k].cf;
	free(t);
	return rc;
}

void *ph2_reg_create(uint32_t nqb, int mode, uint32_t tile_bits,
	uint32_t low_bits)
{
	if (ph2_init() < 0)
		return NULL;
	struct qreg *reg = malloc(sizeof *reg);
	if (!reg)
		return NULL;
	if (qreg_init(reg, nqb) < 0) {
		free(reg);
		return NULL;
	}
	qreg_hip_configure(reg, mode, tile_bits, low_bits);
	return reg;
}

void ph2_reg_free(void *r)
{
	if (!r)
		return;
	qreg_free(r);
	free(r);
}

int ph2_reg_configure(void *r, int mode, uint32_t tile_bits,
	uint32_t low_bits)
{
	qreg_hip_configure(r, mode, tile_bits, low_bits);
	return 0;
}

int ph2_reg_prep_multidet(void *r, size_t ndet, const uint64_t *didx,
	const double *dcf)
{
	struct qreg *reg = r;
	qreg_zero(reg);
	for (size_t d = 0; d < ndet; d++)
		qreg_setamp(reg, didx[d], CMPLX(dcf[2 * d], dcf[2 * d + 1]));
	return qreg_error(reg);
}

int ph2_reg_rotate(void *r, size_t n, const uint64_t *xm, const uint64_t *zm,
	const double *phi)
{
	struct qreg *reg = r;
	struct paulis codes[1024];
	const struct paulis hi = { .pak = { 0, 0 } };
	for (size_t k0 = 0; k0 < n; k0 += 1024) {
		const size_t nb = n - k0 < 1024 ? n - k0 : 1024;
		for (size_t k = 0; k < nb; k++) {
			codes[k].pak[0] = xm[k0 + k];
			codes[k].pak[1] = zm[k0 + k];
		}
		qreg_paulirot(reg, hi, codes, phi + k0, nb);
		if (qreg_error(reg))
			return -1;
	}
	return 0;
}

int ph2_reg_rotate_sched(void *r, size_t n, const uint64_t *xm,
	const uint64_t *zm, const double *phi, const uint32_t *pass_len,
	const uint32_t *tile_bits, const uint32_t *low_bits, size_t npass)
{
	struct qreg *reg = r;
	struct ph2hip_rot *rot = malloc(sizeof *rot * (n ? n : 1));
	if (!rot)
		return -1;
	for (size_t k = 0; k < n; k++) {
		rot[k].x = xm[k];
		rot[k].z = zm[k];
		sincos(phi[k], &rot[k].s, &rot[k].c);
	}
	const int rc = ph2hip_rotate_fused(qreg_hip_device_ptr(reg), reg->qb_lo,
		rot, n, pass_len, tile_bits, npass, low_bits);
	free(rot);
	return rc;
}

/* sum_d amp[idx_d] * conj(cf_d) in determinant order, as circ_measure. */
int ph2_reg_overlap_multidet(void *r, size_t ndet, const uint64_t *didx,
	const double *dcf, double *out)
{
	struct qreg *reg = r;
	uint64_t *idx = malloc(sizeof *idx * (ndet ? ndet : 1));
	double *a = malloc(sizeof *a * 2 * (ndet ? ndet : 1));
	int rc = -1;
	if (idx && a) {
		/* indices beyond the register are refused by the gather */
		for (size_t d = 0; d < ndet; d++)
			idx[d] = didx[d];
		rc = ph2hip_gather(qreg_hip_device_ptr(reg), reg->qb_lo, idx,
			ndet, a);
	}
	if (rc == 0) {
		_Complex double pr = 0.0;
		for (size_t d = 0; d < ndet; d++)
			pr += CMPLX(a[2 * d], a[2 * d + 1]) *
				conj(CMPLX(dcf[2 * d], dcf[2 * d + 1]));
		out[0] = creal(pr);
		out[1] = cimag(pr);
	}
	free(idx);
	free(a);
	return rc;
}

int ph2_reg_norm2(void *r, double *out)
{
	struct qreg *reg = r;
	return ph2hip_norm2(qreg_hip_device_ptr(reg), reg->qb_lo, out);
}

int ph2_reg_read(void *r, double *state)
{
	struct qreg *reg = r;
	return ph2hip_d2h(state, qreg_hip_device_ptr(reg), reg->namp * 16);
}

int ph2_reg_gather(void *r, size_t n, const uint64_t *idx, double *out)
{
	struct qreg *reg = r;
	return ph2hip_gather(qreg_hip_device_ptr(reg), reg->qb_lo, idx, n, out);
}

int ph2_reg_fill_hash(void *r, uint64_t seed)
{
	struct qreg *reg = r;
	return ph2hip_fill_hash(qreg_hip_device_ptr(reg), reg->namp, seed);
}

int ph2_reg_count_unequal(void *r1, void *r2, uint64_t *count)
{
	struct qreg *a = r1, *b = r2;
	if (a->namp != b->namp)
		return -1;
	return ph2hip_count_unequal(qreg_hip_device_ptr(a),
		qreg_hip_device_ptr(b), a->namp, count);
}
