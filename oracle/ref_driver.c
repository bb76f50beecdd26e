/*
 * ref_driver.c -- flat C entry points over the reference phase2 CPU backend
 * (BACKEND=qreg, one MPI rank), compiled together with the reference sources
 * by oracle/Makefile.ref into oracle/_ref/libphase2_ref.so.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is never part of the shipped library;
 * it exists so that tests/golden/make_golden.py can drive the reference
 * algorithms from Python (ctypes) on synthetic inputs and record their
 * outputs as fixtures, and so that bench.py can time the reference CPU path
 * as a baseline.  It includes the reference's public headers from
 * /root/reference/include at build time and links the reference objects.
 *
 * Reference call sites mirrored here:
 *   trott_init/trott_simul       circ/trott.c:21-65
 *   trott2_init/trott2_simul     circ/trott2.c:21-74
 *   qdrift_init/qdrift_simul     circ/qdrift.c:58-133
 *   cmpsit_init/cmpsit_simul     circ/cmpsit.c:108-235
 *   qreg_paulirot                phase2/qreg.c:135-142
 *   state_prep_coeff_expand_all  phase2/state_prep_coeff.c:276-325
 */
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "phase2.h"
#include "circ/trott.h"
#include "circ/trott2.h"
#include "circ/qdrift.h"
#include "circ/cmpsit.h"
#include "phase2/state_prep_coeff.h"
#include "det_small.h"

#define REFD_SEED UINT64_C(0x5eed0f0dd1ce5eed)

static int refd_world(void)
{
	struct world_info wd;
	if (world_info(&wd) == WORLD_READY)
		return 0;
	return world_init(NULL, NULL, REFD_SEED) == WORLD_READY ? 0 : -1;
}

static int build_hamil(struct circ_hamil *hm, uint32_t nqb, size_t nterms,
	const uint64_t *xm, const uint64_t *zm, const double *cf)
{
	if (circ_hamil_init(hm, nqb, nterms) < 0)
		return -1;
	for (size_t k = 0; k < nterms; k++) {
		hm->terms[k].cf = cf[k];
		hm->terms[k].op.pak[0] = xm[k];
		hm->terms[k].op.pak[1] = zm[k];
	}
	return 0;
}

/* Multidet payload: ndet (idx, re, im). */
static int build_muldet(struct circ_muldet *md, size_t ndet,
	const uint64_t *didx, const double *dcf)
{
	if (circ_muldet_init(md, ndet) < 0)
		return -1;
	for (size_t d = 0; d < ndet; d++) {
		md->dets[d].idx = didx[d];
		md->dets[d].cf = CMPLX(dcf[2 * d], dcf[2 * d + 1]);
	}
	return 0;
}

/* Coefficient-matrix payload (single block).  The circ adopts and frees
 * the arrays, so copy them. */
static int build_cm(struct data_coeff_matrix *cm, uint32_t nqb,
	uint32_t n_sites, uint32_t n_alpha, uint32_t n_beta, int closed_shell,
	int tapered, const double *Ca, const double *Cb)
{
	memset(cm, 0, sizeof *cm);
	cm->nqb = nqb;
	cm->n_sites = n_sites;
	cm->n_alpha = n_alpha;
	cm->n_beta = n_beta;
	cm->closed_shell = closed_shell;
	cm->tapered = tapered;
	size_t na = (size_t)n_sites * n_alpha, nb = (size_t)n_sites * n_beta;
	double *a = malloc(sizeof(double) * (na ? na : 1));
	if (!a)
		return -1;
	memcpy(a, Ca, sizeof(double) * na);
	cm->C_alpha = a;
	if (!closed_shell) {
		double *b = malloc(sizeof(double) * (nb ? nb : 1));
		if (!b)
			return -1;
		memcpy(b, Cb, sizeof(double) * nb);
		cm->C_beta = b;
	}
	return 0;
}

struct refd_prep {
	int kind; /* 1 multidet, 2 coeff matrix */
	size_t ndet;
	const uint64_t *didx;
	const double *dcf;
	uint32_t n_sites, n_alpha, n_beta;
	int closed_shell, tapered;
	const double *Ca, *Cb;
};

static int build_prep(const struct refd_prep *p, uint32_t nqb,
	struct circ_muldet *md, struct data_coeff_matrix *cm,
	enum stprep_kind *kind, const void **data)
{
	if (p->kind == 1) {
		if (build_muldet(md, p->ndet, p->didx, p->dcf) < 0)
			return -1;
		*kind = STPREP_MULTIDET;
		*data = md;
		return 0;
	}
	if (build_cm(cm, nqb, p->n_sites, p->n_alpha, p->n_beta,
		    p->closed_shell, p->tapered, p->Ca, p->Cb) < 0)
		return -1;
	*kind = STPREP_COEFF_MATRIX;
	*data = cm;
	return 0;
}

static void copy_out(const struct circ *ct, double *out, double *state)
{
	for (size_t i = 0; i < ct->vals.len; i++) {
		out[2 * i] = creal(ct->vals.z[i]);
		out[2 * i + 1] = cimag(ct->vals.z[i]);
	}
	if (state)
		memcpy(state, ct->reg.amp, sizeof(_Complex double) * ct->reg.namp);
}

/* algo: 1 trott, 2 trott2.  out: 2*steps doubles.  state: 2*2^nqb or NULL. */
int refd_trott(int algo, uint32_t nqb, size_t nterms, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct refd_prep *prep,
	double delta, size_t steps, double *out, double *state)
{
	if (refd_world() < 0)
		return -1;
	struct circ_hamil hm;
	struct circ_muldet md;
	struct data_coeff_matrix cm;
	enum stprep_kind kind;
	const void *data;
	if (build_hamil(&hm, nqb, nterms, xm, zm, cf) < 0)
		return -1;
	if (build_prep(prep, nqb, &md, &cm, &kind, &data) < 0)
		return -1;
	int rc = -1;
	if (algo == 1) {
		struct trott tt;
		struct trott_data dt = { .steps = steps, .delta = delta };
		if (trott_init(&tt, &dt, hm, kind, data, NULL) < 0)
			return -1;
		rc = trott_simul(&tt);
		copy_out(&tt.ct, out, state);
		trott_free(&tt);
	} else {
		struct trott2 t2;
		struct trott2_data dt = { .steps = steps, .delta = delta };
		if (trott2_init(&t2, &dt, hm, kind, data, NULL) < 0)
			return -1;
		rc = trott2_simul(&t2);
		copy_out(&t2.ct, out, state);
		trott2_free(&t2);
	}
	return rc;
}

int refd_qdrift(uint32_t nqb, size_t nterms, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct refd_prep *prep,
	size_t samples, size_t depth, double step_size, uint64_t seed,
	double *out, double *state)
{
	if (refd_world() < 0)
		return -1;
	struct circ_hamil hm;
	struct circ_muldet md;
	struct data_coeff_matrix cm;
	enum stprep_kind kind;
	const void *data;
	if (build_hamil(&hm, nqb, nterms, xm, zm, cf) < 0)
		return -1;
	if (build_prep(prep, nqb, &md, &cm, &kind, &data) < 0)
		return -1;
	struct qdrift qd;
	struct qdrift_data dt = { .samples = samples, .depth = depth,
		.step_size = step_size, .seed = seed };
	if (qdrift_init(&qd, &dt, hm, kind, data, NULL) < 0)
		return -1;
	int rc = qdrift_simul(&qd);
	copy_out(&qd.ct, out, state);
	qdrift_free(&qd);
	return rc;
}

int refd_cmpsit(uint32_t nqb, size_t nterms, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct refd_prep *prep,
	size_t samples, size_t steps, size_t length, size_t depth,
	double angle_det, double angle_rand, uint64_t seed, double *out,
	double *state)
{
	if (refd_world() < 0)
		return -1;
	struct circ_hamil hm;
	struct circ_muldet md;
	struct data_coeff_matrix cm;
	enum stprep_kind kind;
	const void *data;
	if (build_hamil(&hm, nqb, nterms, xm, zm, cf) < 0)
		return -1;
	if (build_prep(prep, nqb, &md, &cm, &kind, &data) < 0)
		return -1;
	struct cmpsit cp;
	struct cmpsit_data dt = { .samples = samples, .steps = steps,
		.length = length, .depth = depth, .angle_det = angle_det,
		.angle_rand = angle_rand, .seed = seed };
	if (cmpsit_init(&cp, &dt, hm, kind, data, NULL) < 0)
		return -1;
	int rc = cmpsit_simul(&cp);
	copy_out(&cp.ct, out, state);
	cmpsit_free(&cp);
	return rc;
}

/* Apply qreg_paulirot in blocks of at most 1024 codes (the circ_cache cap)
 * to a caller-supplied state (2*2^nqb doubles, updated in place). */
int refd_paulirot(uint32_t nqb, double *state, size_t ncodes,
	const uint64_t *xm, const uint64_t *zm, const double *phis)
{
	if (refd_world() < 0)
		return -1;
	struct qreg reg;
	if (qreg_init(&reg, nqb) < 0)
		return -1;
	memcpy(reg.amp, state, sizeof(_Complex double) * reg.namp);
	struct paulis codes[1024];
	const struct paulis hi = { .pak = { 0, 0 } };
	for (size_t k0 = 0; k0 < ncodes; k0 += 1024) {
		size_t nb = ncodes - k0 < 1024 ? ncodes - k0 : 1024;
		for (size_t k = 0; k < nb; k++) {
			codes[k].pak[0] = xm[k0 + k];
			codes[k].pak[1] = zm[k0 + k];
		}
		qreg_paulirot(&reg, hi, codes, phis + k0, nb);
	}
	memcpy(state, reg.amp, sizeof(_Complex double) * reg.namp);
	qreg_free(&reg);
	return 0;
}

/* Slater-Condon expansion into a fresh register; state gets 2*2^nqb. */
int refd_expand(uint32_t nqb, uint32_t n_sites, uint32_t n_alpha,
	uint32_t n_beta, int closed_shell, int tapered, const double *Ca,
	const double *Cb, double *state)
{
	if (refd_world() < 0)
		return -1;
	struct qreg reg;
	if (qreg_init(&reg, nqb) < 0)
		return -1;
	struct state_prep_coeff_scratch sc;
	if (state_prep_coeff_scratch_init(&sc, n_sites, n_alpha, n_beta) < 0)
		return -1;
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
	memcpy(state, reg.amp, sizeof(_Complex double) * reg.namp);
	state_prep_coeff_scratch_free(&sc);
	qreg_free(&reg);
	return rc;
}

/* Inner product <trial|state> with a coefficient-matrix trial. */
int refd_inner(uint32_t nqb, uint32_t n_sites, uint32_t n_alpha,
	uint32_t n_beta, int closed_shell, int tapered, const double *Ca,
	const double *Cb, const double *state, double *out)
{
	if (refd_world() < 0)
		return -1;
	struct qreg reg;
	if (qreg_init(&reg, nqb) < 0)
		return -1;
	memcpy(reg.amp, state, sizeof(_Complex double) * reg.namp);
	struct state_prep_coeff_scratch sc;
	if (state_prep_coeff_scratch_init(&sc, n_sites, n_alpha, n_beta) < 0)
		return -1;
	_Complex double z = 0;
	int rc = state_prep_coeff_inner(&reg, &sc, Ca, closed_shell ? NULL : Cb,
		1.0, tapered, &z);
	out[0] = creal(z);
	out[1] = cimag(z);
	state_prep_coeff_scratch_free(&sc);
	qreg_free(&reg);
	return rc;
}

/* xoshiro256** stream: n raw 64-bit outputs after init(seed). */
void refd_xoshiro(uint64_t seed, size_t n, uint64_t *out)
{
	struct xoshiro256ss rng;
	xoshiro256ss_init(&rng, seed);
	for (size_t i = 0; i < n; i++)
		out[i] = xoshiro256ss_next(&rng);
}

/* CDF inverse of ys over |w| (the qdrift sampler, lib/prob.c). */
int refd_cdf_inverse(size_t nw, const double *w, size_t ny, const double *ys,
	uint64_t *idx, double *lambda)
{
	struct prob_cdf cdf;
	if (prob_cdf_init(&cdf, nw) < 0)
		return -1;
	if (prob_cdf_from_array_strided(&cdf, w, sizeof(double), lambda) < 0)
		return -1;
	for (size_t i = 0; i < ny; i++)
		idx[i] = prob_cdf_inverse(&cdf, ys[i]);
	prob_cdf_free(&cdf);
	return 0;
}

/* Lex sort of Pauli codes (circ_hamil_sort_lex): returns the permutation
 * as the sorted position of each input (by reading back cf = input index). */
int refd_sort_lex(size_t n, const uint64_t *xm, const uint64_t *zm,
	uint64_t *perm)
{
	struct circ_hamil hm;
	if (circ_hamil_init(&hm, 64, n) < 0)
		return -1;
	for (size_t k = 0; k < n; k++) {
		hm.terms[k].cf = (double)k;
		hm.terms[k].op.pak[0] = xm[k];
		hm.terms[k].op.pak[1] = zm[k];
	}
	circ_hamil_sort_lex(&hm);
	for (size_t k = 0; k < n; k++)
		perm[k] = (uint64_t)hm.terms[k].cf;
	circ_hamil_free(&hm);
	return 0;
}

/* det_small over a row-major n x n matrix. */
double refd_det(const double *A, uint32_t n)
{
	return det_small(A, n);
}
