/*
 * Batched qDRIFT over many independent registers (include/phase2_batch.h).
 *
 * Every register follows qdrift_simul (circ/qdrift.c:105-133) exactly: its
 * own |c| CDF over the Hamiltonian in input order, its own xoshiro256**
 * stream continuing across its samples, per sample: prepare the multidet
 * state, draw `depth` terms (cf = signof(c), phi = asin(step)*cf, c/s by
 * sincos as the reference), rotate, measure sum a*conj(cf) in determinant
 * order.  The registers share one device buffer; a round (sample s of every
 * register that has one) is prepared, rotated and measured together.
 */
#define _GNU_SOURCE /* sincos */
#define LOG_SUBSYS "batch"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "log.h"
#include "phase2/qreg_hip.h"
#include "phase2_batch.h"
#include "phase2_flat.h"
#include "phase2_hip.h"
#include "prob.h"
#include "xoshiro256ss.h"
#include "internal.h"

#define QDRIFT_DEFAULT_SEED UINT64_C(0xeccd9dcc749fcdca)
#define ALIGN_AMPS 256	/* register starts on 4 KiB boundaries */
#define TUNE_REGS 16

struct breg {
	const struct ph2_batch_reg *r;
	uint64_t off;		/* first amplitude in the shared buffer */
	size_t out_off;		/* first overlap in out */
	struct prob_cdf cdf;
	struct xoshiro256ss rng;
	double omega;
	struct ph2hip_rot *rot;	/* this round's draws */
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
			prob_cdf_free(&B->b[i].cdf);
			free(B->b[i].rot);
			ph2_sched_free(&B->b[i].sch);
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

static int batch_init(struct batch *B, const struct ph2_batch_reg *regs,
	size_t n)
{
	memset(B, 0, sizeof *B);
	B->n = n;
	B->b = calloc(n ? n : 1, sizeof *B->b);
	B->jobs = calloc(n ? n : 1, sizeof *B->jobs);
	if (!B->b || !B->jobs)
		return -1;
	uint64_t off = 0;
	size_t out_off = 0;
	for (size_t i = 0; i < n; i++) {
		const struct ph2_batch_reg *r = &regs[i];
		struct breg *g = &B->b[i];
		if (r->nqb == 0 || r->nqb > 40 || r->nterms == 0) {
			log_error("batch: register %zu: nqb=%u nterms=%zu", i,
				r->nqb, r->nterms);
			return -1;
		}
		g->r = r;
		g->off = off;
		g->out_off = out_off;
		off += ((UINT64_C(1) << r->nqb) + ALIGN_AMPS - 1) /
			ALIGN_AMPS * ALIGN_AMPS;
		out_off += r->samples;
		B->ndet_tot += r->ndet;
		if (prob_cdf_init(&g->cdf, r->nterms) < 0)
			return -1;
		prob_cdf_from_array_strided(&g->cdf, r->cf, sizeof(double),
			NULL);
		xoshiro256ss_init(&g->rng, r->seed ? r->seed
						   : QDRIFT_DEFAULT_SEED);
		g->omega = asin(r->step_size);
		g->rot = malloc(sizeof *g->rot * (r->depth ? r->depth : 1));
		if (!g->rot)
			return -1;
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
	return 0;
}

static int active(const struct breg *g, size_t s)
{
	return g->r->samples > s;
}

/* Zero everything, then place the determinants of the active registers
 * (circ_prepst, phase2/circ.c:215-224). */
static int prep_round(struct batch *B, size_t s, size_t lim)
{
	if (ph2hip_zero(B->amp, B->namp) < 0)
		return -1;
	size_t m = 0;
	for (size_t i = 0; i < lim; i++) {
		const struct breg *g = &B->b[i];
		if (!active(g, s))
			continue;
		const uint64_t mask = (UINT64_C(1) << g->r->nqb) - 1;
		for (size_t d = 0; d < g->r->ndet; d++) {
			B->idx[m] = g->off + (g->r->didx[d] & mask);
			B->val[2 * m] = g->r->dcf[2 * d];
			B->val[2 * m + 1] = g->r->dcf[2 * d + 1];
			m++;
		}
	}
	return ph2hip_scatter_set(B->amp, B->namp, B->idx, B->val, m);
}

/* depth draws of one register (qdrift.c:95-103, then circ_step's
 * phi = omega * cf and the backend's sincos). */
static void draw(struct breg *g)
{
	const struct ph2_batch_reg *r = g->r;
	for (size_t d = 0; d < r->depth; d++) {
		const double u = xoshiro256ss_dbl01(&g->rng);
		const size_t k = prob_cdf_inverse(&g->cdf, u);
		const double phi = g->omega * signof(r->cf[k]);
		g->rot[d].x = r->xm[k];
		g->rot[d].z = r->zm[k];
		sincos(phi, &g->rot[d].s, &g->rot[d].c);
	}
}

static int run_round(struct batch *B, size_t s, size_t lim, int mode,
	uint32_t tb, uint32_t low)
{
	size_t nj = 0;
	for (size_t i = 0; i < lim; i++) {
		struct breg *g = &B->b[i];
		if (!active(g, s) || g->r->depth == 0)
			continue;
		const uint32_t nqb = g->r->nqb;
		void *a = (char *)B->amp + g->off * 16;
		const int small = nqb < PH2HIP_TILE_MIN;
		if (mode == PH2_BATCH_PLAIN_EACH || small) {
			if (ph2hip_rotate_plain(a, nqb, g->rot, g->r->depth) < 0)
				return -1;
			B->launches += g->r->depth;
			continue;
		}
		const uint32_t t = tb < nqb ? tb : nqb;
		if (ph2_sched_greedy(g->rot, g->r->depth, nqb, t, low, &g->sch) <
			0)
			return -1;
		B->passes += g->sch.npass;
		if (mode == PH2_BATCH_FUSED_EACH || nqb < tb) {
			if (ph2hip_rotate_fused(a, nqb, g->rot, g->r->depth,
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
		J->n = g->r->depth;
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
static int plan_round(struct batch *B, size_t s, uint32_t tb, uint32_t low,
	struct ph2hip_batch_plan **plan)
{
	size_t nj = 0;
	for (size_t i = 0; i < B->n; i++) {
		struct breg *g = &B->b[i];
		if (!active(g, s) || g->r->depth == 0)
			continue;
		const uint32_t nqb = g->r->nqb;
		if (ph2_sched_greedy(g->rot, g->r->depth, nqb, tb, low, &g->sch) <
			0)
			return -1;
		B->passes += g->sch.npass;
		struct ph2hip_batch_job *J = &B->jobs[nj++];
		J->amp_off = g->off;
		J->nqb = nqb;
		J->rot = g->rot;
		J->n = g->r->depth;
		J->pass_len = g->sch.len;
		J->npass = g->sch.npass;
	}
	return ph2hip_batch_plan_build(B->jobs, nj, tb, low, B->namp, plan);
}

static int measure_round(struct batch *B, size_t s, double *out)
{
	size_t m = 0;
	for (size_t i = 0; i < B->n; i++) {
		const struct breg *g = &B->b[i];
		if (!active(g, s))
			continue;
		const uint64_t mask = (UINT64_C(1) << g->r->nqb) - 1;
		for (size_t d = 0; d < g->r->ndet; d++)
			B->idx[m++] = g->off + (g->r->didx[d] & mask);
	}
	if (ph2hip_gather_abs(B->amp, B->namp, B->idx, m, B->val) < 0)
		return -1;
	m = 0;
	for (size_t i = 0; i < B->n; i++) {
		const struct breg *g = &B->b[i];
		if (!active(g, s))
			continue;
		_Complex double pr = 0.0;
		for (size_t d = 0; d < g->r->ndet; d++, m++)
			pr += CMPLX(B->val[2 * m], B->val[2 * m + 1]) *
				conj(CMPLX(g->r->dcf[2 * d], g->r->dcf[2 * d + 1]));
		out[2 * (g->out_off + s)] = creal(pr);
		out[2 * (g->out_off + s) + 1] = cimag(pr);
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
		minq = B->b[i].r->nqb < minq ? B->b[i].r->nqb : minq;
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

int ph2_batch_qdrift(const struct ph2_batch_reg *regs, size_t nreg, int mode,
	uint32_t tile_bits, uint32_t low_bits, double *out,
	struct ph2_batch_stats *stats)
{
	if (ph2_init() < 0)
		return -1;
	struct batch B;
	if (batch_init(&B, regs, nreg) < 0) {
		batch_free(&B);
		return -1;
	}
	size_t rounds = 0;
	for (size_t i = 0; i < nreg; i++)
		rounds = regs[i].samples > rounds ? regs[i].samples : rounds;
	int rc = 0;
	double tune = 0.0;
	/* round 0 is drawn once; the tuner replays those draws */
	for (size_t i = 0; i < nreg; i++)
		if (active(&B.b[i], 0))
			draw(&B.b[i]);
	if (tile_bits == 0 && mode != PH2_BATCH_PLAIN_EACH && rounds) {
		const double t0 = now_s();
		rc = autotune(&B, mode, &tile_bits, &low_bits);
		tune = now_s() - t0;
	} else if (tile_bits == 0) {
		tile_bits = 12;
		low_bits = 3;
	}
	if (ph2hip_synchronize() < 0)
		rc = -1;
	/* Batched rounds are pipelined: while the device runs round s, the
	 * host draws round s+1 and builds its plan (the draws do not depend on
	 * the state, and each register's stream stays in order). */
	int pipelined = mode == PH2_BATCH_FUSED_BATCHED;
	for (size_t i = 0; i < nreg && pipelined; i++)
		if (regs[i].nqb < tile_bits)
			pipelined = 0;   /* narrow registers take other paths */
	double prep = 0.0;
	const double t0 = now_s();
	struct ph2hip_batch_plan *plan = NULL, *next = NULL;
	if (pipelined && rc == 0 && rounds) {
		const double tp = now_s();
		rc = plan_round(&B, 0, tile_bits, low_bits, &next);
		prep += now_s() - tp;
	}
	for (size_t s = 0; s < rounds && rc == 0; s++) {
		if (!pipelined) {
			const double tp = now_s();
			if (s > 0)
				for (size_t i = 0; i < nreg; i++)
					if (active(&B.b[i], s))
						draw(&B.b[i]);
			prep += now_s() - tp;
			rc = prep_round(&B, s, nreg);
			if (!rc)
				rc = run_round(&B, s, nreg, mode, tile_bits,
					low_bits);
			if (!rc)
				rc = measure_round(&B, s, out);
			continue;
		}
		plan = next;
		next = NULL;
		uint64_t nl = 0;
		rc = prep_round(&B, s, nreg);
		if (!rc)
			rc = ph2hip_batch_plan_launch(B.amp, plan, 1, &nl);
		B.launches += nl;
		if (!rc && s + 1 < rounds) {
			const double tp = now_s();
			for (size_t i = 0; i < nreg; i++)
				if (active(&B.b[i], s + 1))
					draw(&B.b[i]);
			rc = plan_round(&B, s + 1, tile_bits, low_bits, &next);
			prep += now_s() - tp;
		}
		if (!rc)
			rc = measure_round(&B, s, out);   /* waits for round s */
		ph2hip_batch_plan_free(plan);
		plan = NULL;
	}
	if (ph2hip_synchronize() < 0)
		rc = -1;
	ph2hip_batch_plan_free(plan);
	ph2hip_batch_plan_free(next);
	if (rc == 0 && ph2hip_synchronize() < 0)
		rc = -1;
	if (rc < 0)
		log_error("ph2_batch_qdrift: %s", ph2hip_strerror());
	if (stats) {
		memset(stats, 0, sizeof *stats);
		stats->seconds = now_s() - t0;
		stats->tune_seconds = tune;
		stats->prep_seconds = prep;
		for (size_t i = 0; i < nreg; i++)
			stats->amp_updates += (double)regs[i].samples *
				(double)regs[i].depth *
				(double)(UINT64_C(1) << regs[i].nqb);
		stats->launches = B.launches;
		stats->passes = B.passes;
		stats->tile_bits = tile_bits;
		stats->low_bits = low_bits;
	}
	batch_free(&B);
	return rc;
}
