/*
 * Slater-Condon expansion and inner product on the device (API of
 * include/phase2/state_prep_coeff.h; behaviour of
 * phase2/state_prep_coeff.c:31-325 of the reference).
 *
 * The k-subset tables are built once per scratch (lexicographic order, as
 * combinations.c enumerates them) and mirrored on the device as occupation
 * masks.  Per block the host computes det(C[occ, :]) for every subset with
 * det_small -- the same LU and rounding as the reference -- uploads the
 * determinants, and one kernel forms weight*det_a*det_b (same product
 * order), prunes |c| < 1e-12 and writes or accumulates it at its index.
 * The inner product reduces on the device.  Determinants are cached while
 * the coefficient matrices are unchanged.
 */
#define LOG_SUBSYS "prep"

#include <complex.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "combinations.h"
#include "det_small.h"
#include "log.h"
#include "phase2/circ.h"
#include "phase2/qreg_hip.h"
#include "phase2/state_prep_coeff.h"
#include "phase2_hip.h"

struct sc_dev {
	struct ph2hip_sc *tab;
	double *ca, *cb;	/* copies of the matrices the dets came from */
	int valid;
	int cb_null;
};

/* The library's own tables of a scratch.  The public struct has the
 * reference's members only, so they live in one block in front of the
 * alpha tuples: sc->tup_a is the block's tail, and the head is found from
 * it. */
struct sc_priv {
	uint64_t *mask_a, *mask_b;	/* occupation bit masks, host */
	struct sc_dev *dev;		/* device mirror, built on first use */
	uint32_t tup[];			/* the alpha tuples (sc->tup_a) */
};

static struct sc_priv *priv_of(const struct state_prep_coeff_scratch *sc)
{
	if (!sc->tup_a)
		return NULL;
	return (struct sc_priv *)((char *)sc->tup_a -
		offsetof(struct sc_priv, tup));
}

static uint64_t n_choose_k(uint32_t n, uint32_t k)
{
	if (k > n)
		return 0;
	if (k > n - k)
		k = n - k;
	uint64_t r = 1;
	for (uint32_t i = 1; i <= k; i++)
		r = r * (n - k + i) / i;
	return r;
}

static int fill_subsets(uint32_t n, uint32_t k, uint32_t *tup,
	uint64_t *mask, size_t M)
{
	struct combo it;
	uint32_t cur[COMBINATIONS_MAX_K];
	const uint32_t kk = k ? k : 1;
	size_t m = 0;
	combinations_init(&it, n, k);
	while (combinations_next(&it, cur) == 0) {
		if (m >= M)
			return -1;
		uint64_t v = 0;
		for (uint32_t r = 0; r < k; r++) {
			tup[m * kk + r] = cur[r];
			v |= UINT64_C(1) << cur[r];
		}
		mask[m++] = v;
	}
	return m == M ? 0 : -1;
}

int state_prep_coeff_scratch_init(struct state_prep_coeff_scratch *sc,
	uint32_t n_sites, uint32_t n_alpha, uint32_t n_beta)
{
	memset(sc, 0, sizeof *sc);
	if (n_alpha > n_sites || n_beta > n_sites) {
		log_error("scratch_init: n_alpha=%u or n_beta=%u > n_sites=%u",
			n_alpha, n_beta, n_sites);
		return -1;
	}
	if (n_alpha > DET_SMALL_MAX_N || n_beta > DET_SMALL_MAX_N ||
	    n_sites > 32) {
		log_error("scratch_init: occupation above %u or n_sites=%u > 32",
			DET_SMALL_MAX_N, n_sites);
		return -1;
	}
	sc->n_sites = n_sites;
	sc->n_alpha = n_alpha;
	sc->n_beta = n_beta;
	sc->Ma = (size_t)n_choose_k(n_sites, n_alpha);
	sc->Mb = (size_t)n_choose_k(n_sites, n_beta);
	const uint32_t ka = n_alpha ? n_alpha : 1, kb = n_beta ? n_beta : 1;
	struct sc_priv *pv = calloc(1, sizeof *pv +
		sizeof *pv->tup * sc->Ma * ka);
	if (pv)
		sc->tup_a = pv->tup;
	sc->tup_b = malloc(sizeof *sc->tup_b * sc->Mb * kb);
	sc->det_a = malloc(sizeof *sc->det_a * sc->Ma);
	sc->det_b = malloc(sizeof *sc->det_b * sc->Mb);
	if (pv) {
		pv->mask_a = malloc(sizeof *pv->mask_a * sc->Ma);
		pv->mask_b = malloc(sizeof *pv->mask_b * sc->Mb);
	}
	if (!pv || !sc->tup_b || !sc->det_a || !sc->det_b ||
	    !pv->mask_a || !pv->mask_b) {
		log_error("scratch_init: alloc failed");
		goto err;
	}
	if (fill_subsets(n_sites, n_alpha, sc->tup_a, pv->mask_a, sc->Ma) < 0 ||
	    fill_subsets(n_sites, n_beta, sc->tup_b, pv->mask_b, sc->Mb) < 0)
		goto err;
	return 0;
err:
	state_prep_coeff_scratch_free(sc);
	return -1;
}

void state_prep_coeff_scratch_free(struct state_prep_coeff_scratch *sc)
{
	struct sc_priv *pv = priv_of(sc);
	if (pv) {
		struct sc_dev *d = pv->dev;
		if (d) {
			ph2hip_sc_destroy(d->tab);
			free(d->ca);
			free(d->cb);
			free(d);
		}
		free(pv->mask_a);
		free(pv->mask_b);
		free(pv);
	}
	free(sc->tup_b);
	free(sc->det_a);
	free(sc->det_b);
	memset(sc, 0, sizeof *sc);
}

/* det(C[occ, :]) for each subset; C row-major (n_sites, k). */
static void subset_dets(uint32_t k, const uint32_t *tup, const double *C,
	double *dets, size_t M)
{
	double sub[DET_SMALL_MAX_N * DET_SMALL_MAX_N];
	for (size_t m = 0; m < M; m++) {
		if (k == 0) {
			dets[m] = 1.0;
			continue;
		}
		for (uint32_t r = 0; r < k; r++)
			memcpy(&sub[r * k], &C[(size_t)tup[m * k + r] * k],
				sizeof(double) * k);
		dets[m] = det_small(sub, k);
	}
}

/* Device tables with the determinants of (C_alpha, C_beta) loaded. */
static struct ph2hip_sc *device_tables(struct state_prep_coeff_scratch *sc,
	const double *C_alpha, const double *C_beta)
{
	struct sc_priv *pv = priv_of(sc);
	if (!pv)
		return NULL;
	struct sc_dev *d = pv->dev;
	const size_t na = (size_t)sc->n_sites * sc->n_alpha;
	const size_t nb = (size_t)sc->n_sites * sc->n_beta;
	if (!d) {
		d = calloc(1, sizeof *d);
		if (!d)
			return NULL;
		d->ca = malloc(sizeof(double) * (na ? na : 1));
		d->cb = malloc(sizeof(double) * (nb ? nb : 1));
		if (!d->ca || !d->cb ||
		    ph2hip_sc_create(&d->tab, pv->mask_a, sc->Ma, pv->mask_b,
			    sc->Mb) < 0) {
			free(d->ca);
			free(d->cb);
			free(d);
			return NULL;
		}
		pv->dev = d;
	}
	const double *Cb = C_beta ? C_beta : C_alpha;
	const int same = d->valid && memcmp(d->ca, C_alpha, sizeof(double) * na)
		== 0 && (C_beta == NULL) == d->cb_null &&
		(C_beta == NULL || memcmp(d->cb, C_beta, sizeof(double) * nb)
			== 0);
	if (same)
		return d->tab;
	subset_dets(sc->n_alpha, sc->tup_a, C_alpha, sc->det_a, sc->Ma);
	subset_dets(sc->n_beta, sc->tup_b, Cb, sc->det_b, sc->Mb);
	if (ph2hip_sc_set_dets(d->tab, sc->det_a, sc->det_b) < 0)
		return NULL;
	memcpy(d->ca, C_alpha, sizeof(double) * na);
	if (C_beta)
		memcpy(d->cb, C_beta, sizeof(double) * nb);
	d->cb_null = C_beta == NULL;
	d->valid = 1;
	return d->tab;
}

int state_prep_coeff_expand(struct qreg *reg,
	struct state_prep_coeff_scratch *sc, const double *C_alpha,
	const double *C_beta, double weight, int tapered, int accumulate)
{
	struct ph2hip_sc *tab = device_tables(sc, C_alpha, C_beta);
	if (!tab) {
		log_error("expand: device tables: %s", ph2hip_strerror());
		return -1;
	}
	if (ph2hip_sc_scatter(qreg_hip_device_ptr(reg), reg->qb_lo, tab,
		    sc->n_sites, weight, tapered, accumulate) < 0) {
		log_error("expand: %s", ph2hip_strerror());
		return -1;
	}
	return 0;
}

int state_prep_coeff_expand_all(struct qreg *reg,
	struct state_prep_coeff_scratch *sc, const struct data_coeff_matrix *cm)
{
	qreg_zero(reg);
	if (qreg_error(reg))
		return -1;
	if (cm->n_components == 0)
		return state_prep_coeff_expand(reg, sc, cm->C_alpha,
			cm->closed_shell ? NULL : cm->C_beta, 1.0, cm->tapered, 0);
	for (size_t k = 0; k < cm->n_components; k++) {
		const struct data_coeff_block *b = &cm->blocks[k];
		if (state_prep_coeff_expand(reg, sc, b->C_alpha,
			    cm->closed_shell ? NULL : b->C_beta, b->cf,
			    cm->tapered, 1) < 0) {
			log_error("expand_all: block %zu failed", k);
			return -1;
		}
	}
	return 0;
}

int state_prep_coeff_inner(struct qreg *reg,
	struct state_prep_coeff_scratch *sc, const double *C_alpha,
	const double *C_beta, double weight, int tapered, _Complex double *out)
{
	struct ph2hip_sc *tab = device_tables(sc, C_alpha, C_beta);
	double r[2];
	if (!tab || ph2hip_sc_inner(qreg_hip_device_ptr(reg), reg->qb_lo, tab,
			    sc->n_sites, weight, tapered, r) < 0) {
		log_error("inner: %s", ph2hip_strerror());
		return -1;
	}
	*out = CMPLX(r[0], r[1]);
	return 0;
}
