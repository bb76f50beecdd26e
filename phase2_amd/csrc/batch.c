/*
 * Batched qDRIFT and composite samples over many independent registers
 * (include/phase2_batch.h).
 *
 * Every register follows its reference driver exactly: qdrift_simul
 * (circ/qdrift.c:105-133) or cmpsit_simul (circ/cmpsit.c:208-235).  It has
 * its own |c| CDF and its own xoshiro256** stream continuing across its
 * samples, applies the same rotations in the same order, and measures
 * sum a*conj(cf) in determinant order.
 *
 * The draws of every register are made on the device before the first
 * round (ph2hip_sample_cdf, one stream per register, bit-identical to the
 * host generator), and copied back for pass planning.  A register's work
 * is a sequence of rounds: a qDRIFT sample is one round of `depth` draws; a
 * composite sample is 2*steps rounds, one per half-step (the deterministic
 * terms, then `depth` draws; forward, then reversed).  Round r of every
 * register that has one runs together: registers starting a sample are
 * prepared, the rotations of all of them run as fused passes, registers
 * ending a sample are measured.  The registers share one device buffer.
 */
#define _GNU_SOURCE /* sincos */
#define LOG_SUBSYS "batch"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "log.h"
#include "circ/cmpsit.h"
#include "phase2/qreg_hip.h"
#include "phase2_batch.h"
#include "phase2_flat.h"
#include "phase2_hip.h"
#include "prob.h"
#include "xoshiro256ss.h"
#include "internal.h"

#define QDRIFT_DEFAULT_SEED UINT64_C(0xeccd9dcc749fcdca)
#define CMPSIT_DEFAULT_SEED UINT64_C(0xafb424901446f21f)
#define ALIGN_AMPS 256	/* register starts on 4 KiB boundaries */
#define TUNE_REGS 16

enum kind { KIND_QDRIFT, KIND_CMPSIT };

struct breg {
	enum kind kind;
	uint32_t nqb;
	uint64_t off;		/* first amplitude in the shared buffer */
	size_t out_off;		/* first overlap in out */
	size_t ndet;
	const uint64_t *didx;
	const double *dcf;	/* interleaved complex */
	size_t samples, rps;	/* samples; rounds per sample */
	size_t depth;		/* draws per round */
	/* qDRIFT: the terms in input order, phi = omega * signof(cf) */
	const uint64_t *xm, *zm;
	const double *cf;
	double omega;
	struct prob_cdf cdf;
	/* composite: the split of H (cmpsit_ranct_init) */
	struct circ_hamil hm;
	struct cmpsit_ranct ct;
	double angle_det;
	uint64_t seed;
	uint32_t *draws;	/* samples * rps * depth, drawn on the device */
	struct ph2hip_rot *rot;	/* this round's rotations */
	size_t nrot;
	struct ph2_sched sch;
};

struct batch {
	struct breg *b;
	size_t n;
	void *amp;
	uint64_t namp;
	uint64_t *idx;		/* prep / measure index scratch */
	double *val;
	size_t ndet_tot;
	struct ph2hip_batch_job *jobs;
	uint64_t launches, passes;
	double draw_seconds;
};

static double now_s(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void batch_free(struct batch *B)
{
	if (B->b) {
		for (size_t i = 0; i < B->n; i++) {
			struct breg *g = &B->b[i];
			prob_cdf_free(&g->cdf);
			if (g->kind == KIND_CMPSIT) {
				cmpsit_ranct_free(&g->ct);
				circ_hamil_free(&g->hm);
			}
			free(g->draws);
			free(g->rot);
			ph2_sched_free(&g->sch);
		}
	}
	free(B->b);
	free(B->idx);
	free(B->val);
	free(B->jobs);
	if (B->amp)
		ph2hip_free(B->amp);
	memset(B, 0, sizeof *B);
}

static size_t rounds_of(const struct breg *g)
{
	return g->samples * g->rps;
}

static int reg_qdrift(struct breg *g, const struct ph2_batch_reg *r)
{
	g->kind = KIND_QDRIFT;
	g->nqb = r->nqb;
	g->ndet = r->ndet;
	g->didx = r->didx;
	g->dcf = r->dcf;
	g->samples = r->samples;
	g->rps = 1;
	g->depth = r->depth;
	g->xm = r->xm;
	g->zm = r->zm;
	g->cf = r->cf;
	g->omega = asin(r->step_size);
	g->seed = r->seed ? r->seed : QDRIFT_DEFAULT_SEED;
	if (prob_cdf_init(&g->cdf, r->nterms) < 0)
		return -1;
	if (prob_cdf_from_array_strided(&g->cdf, r->cf, sizeof(double),
		    NULL) < 0) {
		log_error("batch: qdrift weights sum below DBL_EPSILON");
		return -1;
	}
	g->nrot = r->depth;
	return 0;
}

static int reg_cmpsit(struct breg *g, const struct ph2_batch_cmpsit_reg *r)
{
	g->kind = KIND_CMPSIT;
	g->nqb = r->nqb;
	g->ndet = r->ndet;
	g->didx = r->didx;
	g->dcf = r->dcf;
	g->samples = r->samples;
	/* a sample of no steps is prepared and measured in one round */
	g->rps = r->steps ? 2 * r->steps : 1;
	g->depth = r->steps ? r->depth : 0;
	g->angle_det = r->angle_det;
	g->seed = r->seed ? r->seed : CMPSIT_DEFAULT_SEED;
	if (r->length > r->nterms || (r->depth && r->length == r->nterms)) {
		log_error("batch: cmpsit length %zu of %zu terms (depth %zu)",
			r->length, r->nterms, r->depth);
		return -1;
	}
	if (circ_hamil_init(&g->hm, r->nqb, r->nterms) < 0)
		return -1;
	for (size_t k = 0; k < r->nterms; k++) {
		g->hm.terms[k].cf = r->cf[k];
		g->hm.terms[k].op.pak[0] = r->xm[k];
		g->hm.terms[k].op.pak[1] = r->zm[k];
	}
	const struct cmpsit_data dt = {
		.samples = r->samples, .steps = r->steps, .length = r->length,
		.depth = r->depth, .angle_det = r->angle_det,
		.angle_rand = r->angle_rand, .seed = g->seed };
	if (cmpsit_ranct_init(&g->ct, &g->hm, &dt) < 0)
		return -1;
	g->nrot = r->steps ? r->length + r->depth : 0;
	return 0;
}

/* Layout in the shared buffer, round buffers, and every draw of every
 * register made on the device in one launch. */
static int batch_layout(struct batch *B)
{
	uint64_t off = 0;
	size_t out_off = 0, nstream = 0;
	for (size_t i = 0; i < B->n; i++) {
		struct breg *g = &B->b[i];
		g->off = off;
		g->out_off = out_off;
		off += ((UINT64_C(1) << g->nqb) + ALIGN_AMPS - 1) / ALIGN_AMPS *
			ALIGN_AMPS;
		out_off += g->samples;
		B->ndet_tot += g->ndet;
		g->rot = malloc(sizeof *g->rot * (g->nrot ? g->nrot : 1));
		const size_t nd = rounds_of(g) * g->depth;
		g->draws = malloc(sizeof *g->draws * (nd ? nd : 1));
		if (!g->rot || !g->draws)
			return -1;
		nstream += nd > 0;
	}
	B->namp = off;
	B->idx = malloc(sizeof *B->idx * (B->ndet_tot ? B->ndet_tot : 1));
	B->val = malloc(sizeof *B->val * 2 * (B->ndet_tot ? B->ndet_tot : 1));
	if (!B->idx || !B->val)
		return -1;
	if (ph2hip_malloc(&B->amp, B->namp * 16) < 0) {
		log_error("batch: device alloc of %llu amplitudes: %s",
			(unsigned long long)B->namp, ph2hip_strerror());
		return -1;
	}
	if (nstream == 0)
		return 0;
	const double t0 = now_s();
	const double **tab = malloc(sizeof *tab * nstream);
	size_t *len = malloc(sizeof *len * nstream);
	size_t *nd = malloc(sizeof *nd * nstream);
	uint64_t *st = malloc(sizeof *st * 4 * nstream);
	size_t *first = malloc(sizeof *first * (nstream + 1));
	uint32_t *out = NULL;
	int rc = -1;
	if (!tab || !len || !nd || !st || !first)
		goto done;
	size_t m = 0, tot = 0;
	for (size_t i = 0; i < B->n; i++) {
		const struct breg *g = &B->b[i];
		const size_t k = rounds_of(g) * g->depth;
		if (k == 0)
			continue;
		const struct prob_cdf *c = g->kind == KIND_QDRIFT ? &g->cdf :
			&g->ct.cdf;
		struct xoshiro256ss rng;
		xoshiro256ss_init(&rng, g->seed);
		tab[m] = c->y;
		len[m] = c->len;
		nd[m] = k;
		memcpy(st + 4 * m, rng.s, sizeof rng.s);
		first[m] = tot;
		tot += k;
		m++;
	}
	out = malloc(sizeof *out * tot);
	if (!out)
		goto done;
	if (ph2hip_sample_cdf(nstream, tab, len, st, nd, out) < 0) {
		log_error("batch: device draws: %s", ph2hip_strerror());
		goto done;
	}
	m = 0;
	for (size_t i = 0; i < B->n; i++) {
		struct breg *g = &B->b[i];
		const size_t k = rounds_of(g) * g->depth;
		if (k == 0)
			continue;
		memcpy(g->draws, out + first[m], sizeof *out * k);
		m++;
	}
	rc = 0;
done:
	free(tab);
	free(len);
	free(nd);
	free(st);
	free(first);
	free(out);
	B->draw_seconds = now_s() - t0;
	return rc;
}

static int active(const struct breg *g, size_t r)
{
	return r < rounds_of(g);
}

static int starts_sample(const struct breg *g, size_t r)
{
	return active(g, r) && r % g->rps == 0;
}

static int ends_sample(const struct breg *g, size_t r)
{
	return active(g, r) && r % g->rps == g->rps - 1;
}

/* Zero the registers that start a sample in round r, then place their
 * determinants (circ_prepst, phase2/circ.c:215-224). */
static int prep_round(struct batch *B, size_t r, size_t lim)
{
	int all = lim == B->n;
	for (size_t i = 0; i < lim && all; i++)
		all = starts_sample(&B->b[i], r);
	if (all) {
		if (ph2hip_zero(B->amp, B->namp) < 0)
			return -1;
	} else {
		for (size_t i = 0; i < lim; i++) {
			const struct breg *g = &B->b[i];
			if (starts_sample(g, r) &&
			    ph2hip_zero((char *)B->amp + 16 * g->off,
				    UINT64_C(1) << g->nqb) < 0)
				return -1;
		}
	}
	size_t m = 0;
	for (size_t i = 0; i < lim; i++) {
		const struct breg *g = &B->b[i];
		if (!starts_sample(g, r))
			continue;
		for (size_t d = 0; d < g->ndet; d++) {
			B->idx[m] = g->off + g->didx[d];
			B->val[2 * m] = g->dcf[2 * d];
			B->val[2 * m + 1] = g->dcf[2 * d + 1];
			m++;
		}
	}
	return m ? ph2hip_scatter_set(B->amp, B->namp, B->idx, B->val, m) : 0;
}

static void set_rot(struct ph2hip_rot *o, const struct paulis op, double phi)
{
	o->x = op.pak[0];
	o->z = op.pak[1];
	sincos(phi, &o->s, &o->c);
}

/* The rotations of round r of one register, from its device draws.
 * qDRIFT (qdrift.c:95-103 and circ_step at omega = asin(step)):
 * phi = omega * signof(c_k).  Composite (cmpsit.c:151-177, 218-232): the
 * deterministic terms with cf * angle_det, then the draws with
 * signof(c_k) * angle_rand, phi = 0.5 * cf, reversed on the second
 * half-step. */
static void fill_round(struct breg *g, size_t r)
{
	const uint32_t *dw = g->draws + r * g->depth;
	if (g->kind == KIND_QDRIFT) {
		for (size_t d = 0; d < g->depth; d++) {
			const uint32_t k = dw[d];
			struct paulis op = { { g->xm[k], g->zm[k] } };
			set_rot(&g->rot[d], op, g->omega * signof(g->cf[k]));
		}
		return;
	}
	const struct cmpsit_ranct *ct = &g->ct;
	const size_t len = ct->hm_det.len, n = g->nrot;
	const int rev = (r % g->rps) & 1;
	for (size_t j = 0; j < n; j++) {
		struct paulis op;
		double cf;
		if (j < len) {
			op = ct->hm_det.terms[j].op;
			cf = ct->hm_det.terms[j].cf * g->angle_det;
		} else {
			const struct circ_hamil_term *t =
				&ct->hm_ran.terms[dw[j - len]];
			op = t->op;
			cf = signof(t->cf) * ct->angle_rand;
		}
		set_rot(&g->rot[rev ? n - 1 - j : j], op, 0.5 * cf);
	}
}

static int run_round(struct batch *B, size_t r, size_t lim, int mode,
	uint32_t tb, uint32_t low)
{
	size_t nj = 0;
	for (size_t i = 0; i < lim; i++) {
		struct breg *g = &B->b[i];
		if (!active(g, r) || g->nrot == 0)
			continue;
		const uint32_t nqb = g->nqb;
		void *a = (char *)B->amp + g->off * 16;
		const int small = nqb < PH2HIP_TILE_MIN;
		if (mode == PH2_BATCH_PLAIN_EACH || small) {
			if (ph2hip_rotate_plain(a, nqb, g->rot, g->nrot) < 0)
				return -1;
			B->launches += g->nrot;
			continue;
		}
		const uint32_t t = tb < nqb ? tb : nqb;
		if (ph2_sched_greedy(g->rot, g->nrot, nqb, t, low, &g->sch) < 0)
			return -1;
		B->passes += g->sch.npass;
		if (mode == PH2_BATCH_FUSED_EACH || nqb < tb) {
			if (ph2hip_rotate_fused(a, nqb, g->rot, g->nrot,
				    g->sch.len, g->sch.tile_bits, g->sch.npass,
				    g->sch.low_bits) < 0)
				return -1;
			B->launches += g->sch.npass;
			continue;
		}
		struct ph2hip_batch_job *J = &B->jobs[nj++];
		J->amp_off = g->off;
		J->nqb = nqb;
		J->rot = g->rot;
		J->n = g->nrot;
		J->pass_len = g->sch.len;
		J->npass = g->sch.npass;
	}
	if (nj) {
		uint64_t nl = 0;
		if (ph2hip_rotate_batched(B->amp, B->namp, B->jobs, nj, tb, low, 1,
			    &nl) < 0)
			return -1;
		B->launches += nl;
	}
	return 0;
}

/* Host half of a batched round: greedy schedules of every active register
 * and the pass descriptors of the whole round (no device work). */
static int plan_round(struct batch *B, size_t r, uint32_t tb, uint32_t low,
	struct ph2hip_batch_plan **plan)
{
	size_t nj = 0;
	for (size_t i = 0; i < B->n; i++) {
		struct breg *g = &B->b[i];
		if (!active(g, r) || g->nrot == 0)
			continue;
		const uint32_t nqb = g->nqb;
		if (ph2_sched_greedy(g->rot, g->nrot, nqb, tb, low, &g->sch) < 0)
			return -1;
		B->passes += g->sch.npass;
		struct ph2hip_batch_job *J = &B->jobs[nj++];
		J->amp_off = g->off;
		J->nqb = nqb;
		J->rot = g->rot;
		J->n = g->nrot;
		J->pass_len = g->sch.len;
		J->npass = g->sch.npass;
	}
	return ph2hip_batch_plan_build(B->jobs, nj, tb, low, B->namp, plan);
}

/* Registers that end a sample in round r: gather their determinants'
 * amplitudes, sum a*conj(cf) in determinant order (circ.c:338-350). */
static int measure_round(struct batch *B, size_t r, double *out)
{
	size_t m = 0;
	for (size_t i = 0; i < B->n; i++) {
		const struct breg *g = &B->b[i];
		if (!ends_sample(g, r))
			continue;
		for (size_t d = 0; d < g->ndet; d++)
			B->idx[m++] = g->off + g->didx[d];
	}
	if (m && ph2hip_gather_abs(B->amp, B->namp, B->idx, m, B->val) < 0)
		return -1;
	m = 0;
	for (size_t i = 0; i < B->n; i++) {
		const struct breg *g = &B->b[i];
		if (!ends_sample(g, r))
			continue;
		_Complex double pr = 0.0;
		for (size_t d = 0; d < g->ndet; d++, m++)
			pr += CMPLX(B->val[2 * m], B->val[2 * m + 1]) *
				conj(CMPLX(g->dcf[2 * d], g->dcf[2 * d + 1]));
		const size_t smp = r / g->rps;
		out[2 * (g->out_off + smp)] = creal(pr);
		out[2 * (g->out_off + smp) + 1] = cimag(pr);
	}
	return 0;
}

/* Autotuner: time the first round of the first TUNE_REGS registers under
 * each candidate tile/low setting (every setting gives the same state) and
 * keep the fastest. */
static int autotune(struct batch *B, int mode, uint32_t *tb, uint32_t *low)
{
	static const uint32_t TB[] = { 11, 12, 13 };
	static const uint32_t LOW[] = { 3, 4, 5 };
	const size_t lim = B->n < TUNE_REGS ? B->n : TUNE_REGS;
	uint32_t minq = 64;
	for (size_t i = 0; i < lim; i++)
		minq = B->b[i].nqb < minq ? B->b[i].nqb : minq;
	double best = 1e300;
	*tb = 12;
	*low = 3;
	for (size_t a = 0; a < sizeof TB / sizeof *TB; a++) {
		if (TB[a] > minq)
			continue;
		for (size_t c = 0; c < sizeof LOW / sizeof *LOW; c++) {
			if (prep_round(B, 0, lim) < 0 || ph2hip_synchronize() < 0)
				return -1;
			const double t0 = now_s();
			if (run_round(B, 0, lim, mode, TB[a], LOW[c]) < 0 ||
			    ph2hip_synchronize() < 0)
				return -1;
			const double dt = now_s() - t0;
			if (dt < best) {
				best = dt;
				*tb = TB[a];
				*low = LOW[c];
			}
		}
	}
	B->launches = B->passes = 0;
	return 0;
}

static void fill_all(struct batch *B, size_t r)
{
	for (size_t i = 0; i < B->n; i++)
		if (active(&B->b[i], r))
			fill_round(&B->b[i], r);
}

/* The rounds of every register, in order (shared by both drivers). */
static int batch_run(struct batch *B, int mode, uint32_t tile_bits,
	uint32_t low_bits, double *out, struct ph2_batch_stats *stats)
{
	size_t rounds = 0;
	for (size_t i = 0; i < B->n; i++)
		rounds = rounds_of(&B->b[i]) > rounds ? rounds_of(&B->b[i]) :
			rounds;
	int rc = 0;
	double tune = 0.0;
	/* round 0 is filled once; the tuner replays it */
	fill_all(B, 0);
	if (tile_bits == 0 && mode != PH2_BATCH_PLAIN_EACH && rounds) {
		const double t0 = now_s();
		rc = autotune(B, mode, &tile_bits, &low_bits);
		tune = now_s() - t0;
	} else if (tile_bits == 0) {
		tile_bits = 12;
		low_bits = 3;
	}
	if (ph2hip_synchronize() < 0)
		rc = -1;
	/* Batched rounds are pipelined: while the device runs round r, the
	 * host fills round r+1 from the draws and builds its plan (the
	 * rotations do not depend on the state). */
	int pipelined = mode == PH2_BATCH_FUSED_BATCHED;
	for (size_t i = 0; i < B->n && pipelined; i++)
		if (B->b[i].nqb < tile_bits)
			pipelined = 0;   /* narrow registers take other paths */
	double prep = 0.0;
	const double t0 = now_s();
	struct ph2hip_batch_plan *plan = NULL, *next = NULL;
	if (pipelined && rc == 0 && rounds) {
		const double tp = now_s();
		rc = plan_round(B, 0, tile_bits, low_bits, &next);
		prep += now_s() - tp;
	}
	for (size_t r = 0; r < rounds && rc == 0; r++) {
		if (!pipelined) {
			const double tp = now_s();
			if (r > 0)
				fill_all(B, r);
			prep += now_s() - tp;
			rc = prep_round(B, r, B->n);
			if (!rc)
				rc = run_round(B, r, B->n, mode, tile_bits,
					low_bits);
			if (!rc)
				rc = measure_round(B, r, out);
			continue;
		}
		plan = next;
		next = NULL;
		uint64_t nl = 0;
		rc = prep_round(B, r, B->n);
		if (!rc)
			rc = ph2hip_batch_plan_launch(B->amp, plan, 1, &nl);
		B->launches += nl;
		if (!rc && r + 1 < rounds) {
			const double tp = now_s();
			fill_all(B, r + 1);
			rc = plan_round(B, r + 1, tile_bits, low_bits, &next);
			prep += now_s() - tp;
		}
		if (!rc)
			rc = measure_round(B, r, out);   /* waits for round r */
		ph2hip_batch_plan_free(plan);
		plan = NULL;
	}
	if (ph2hip_synchronize() < 0)
		rc = -1;
	ph2hip_batch_plan_free(plan);
	ph2hip_batch_plan_free(next);
	if (rc < 0)
		log_error("batch: %s", ph2hip_strerror());
	if (stats) {
		memset(stats, 0, sizeof *stats);
		stats->seconds = now_s() - t0;
		stats->tune_seconds = tune;
		stats->prep_seconds = prep;
		stats->draw_seconds = B->draw_seconds;
		for (size_t i = 0; i < B->n; i++) {
			const struct breg *g = &B->b[i];
			stats->amp_updates += (double)rounds_of(g) *
				(double)g->nrot *
				(double)(UINT64_C(1) << g->nqb);
		}
		stats->launches = B->launches;
		stats->passes = B->passes;
		stats->tile_bits = tile_bits;
		stats->low_bits = low_bits;
	}
	return rc;
}

static int batch_alloc(struct batch *B, size_t n)
{
	memset(B, 0, sizeof *B);
	B->n = n;
	B->b = calloc(n ? n : 1, sizeof *B->b);
	B->jobs = calloc(n ? n : 1, sizeof *B->jobs);
	return B->b && B->jobs ? 0 : -1;
}

static int check_reg(size_t i, uint32_t nqb, size_t nterms, size_t ndet,
	const uint64_t *didx)
{
	if (nqb == 0 || nqb > 40 || nterms == 0 || nterms > UINT32_MAX) {
		log_error("batch: register %zu: nqb=%u nterms=%zu", i, nqb,
			nterms);
		return -1;
	}
	/* an index beyond the register would alias a lower one (or another
	 * register of the shared buffer): refused, as for one register */
	for (size_t d = 0; d < ndet; d++)
		if (didx[d] >> nqb) {
			log_error("batch: register %zu: determinant %zu index "
				"%llu beyond %u qubits", i, d,
				(unsigned long long)didx[d], nqb);
			return -1;
		}
	return 0;
}

int ph2_batch_qdrift(const struct ph2_batch_reg *regs, size_t nreg, int mode,
	uint32_t tile_bits, uint32_t low_bits, double *out,
	struct ph2_batch_stats *stats)
{
	if (ph2_init() < 0)
		return -1;
	struct batch B;
	int rc = batch_alloc(&B, nreg);
	for (size_t i = 0; i < nreg && rc == 0; i++) {
		rc = check_reg(i, regs[i].nqb, regs[i].nterms, regs[i].ndet,
			regs[i].didx);
		if (rc == 0)
			rc = reg_qdrift(&B.b[i], &regs[i]);
	}
	if (rc == 0)
		rc = batch_layout(&B);
	if (rc == 0)
		rc = batch_run(&B, mode, tile_bits, low_bits, out, stats);
	batch_free(&B);
	return rc;
}

int ph2_batch_cmpsit(const struct ph2_batch_cmpsit_reg *regs, size_t nreg,
	int mode, uint32_t tile_bits, uint32_t low_bits, double *out,
	struct ph2_batch_stats *stats)
{
	if (ph2_init() < 0)
		return -1;
	struct batch B;
	int rc = batch_alloc(&B, nreg);
	for (size_t i = 0; i < nreg && rc == 0; i++) {
		rc = check_reg(i, regs[i].nqb, regs[i].nterms, regs[i].ndet,
			regs[i].didx);
		if (rc == 0)
			rc = reg_cmpsit(&B.b[i], &regs[i]);
	}
	if (rc == 0)
		rc = batch_layout(&B);
	if (rc == 0)
		rc = batch_run(&B, mode, tile_bits, low_bits, out, stats);
	batch_free(&B);
	return rc;
}
