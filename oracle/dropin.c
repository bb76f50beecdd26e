/*
 * Drop-in check (test infrastructure, never shipped): flat entry points that
 * run the reference's own algorithm drivers -- circ/trott.c, circ/trott2.c,
 * circ/qdrift.c and circ/cmpsit.c, compiled unchanged against this
 * library's include/ by oracle/Makefile.ref -- on top of libphase2_hip.so.
 * The caller side is what ph2run does with them (ph2run/ph2run.c:327-360):
 * build a circ_hamil and a state-prep payload, *_init (which adopts both),
 * *_simul with a step writer, read ct.vals, *_free.
 *
 * Every dropin_* returns 0 or -1 and writes the per-step overlaps as
 * interleaved doubles to out, the step writer's copy of them to wrote and
 * the final state (2^nqb interleaved complex) to state if it is non-NULL.
 */
#include <complex.h>
#include <stdlib.h>
#include <string.h>

#include "phase2.h"
#include "phase2_flat.h"
#include "phase2_hip.h"
#include "phase2/qreg_hip.h"
#include "circ/cmpsit.h"
#include "circ/qdrift.h"
#include "circ/trott.h"
#include "circ/trott2.h"

struct recorder {
	double *z;
	size_t cap;
	size_t calls;
};

static int record(void *ctx, size_t step, _Complex double z)
{
	struct recorder *r = ctx;
	if (step >= r->cap)
		return -1;
	r->z[2 * step] = creal(z);
	r->z[2 * step + 1] = cimag(z);
	r->calls++;
	return 0;
}

static int hamil(struct circ_hamil *hm, uint32_t nqb, size_t n,
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

static double *copy_of(const double *src, size_t n)
{
	double *d = malloc(sizeof *d * (n ? n : 1));
	if (d && n)
		memcpy(d, src, sizeof *d * n);
	return d;
}

/* The payload circ_init adopts: a multidet list or one coefficient-matrix
 * block, copied from the caller's arrays. */
struct payload {
	struct circ_muldet md;
	struct data_coeff_matrix cm;
	enum stprep_kind kind;
	const void *data;
};

static int payload(struct payload *p, const struct ph2_prep *in,
	uint32_t nqb)
{
	memset(p, 0, sizeof *p);
	if (in->kind == 1) {
		if (circ_muldet_init(&p->md, in->ndet) < 0)
			return -1;
		for (size_t d = 0; d < in->ndet; d++) {
			p->md.dets[d].idx = in->didx[d];
			p->md.dets[d].cf = CMPLX(in->dcf[2 * d],
				in->dcf[2 * d + 1]);
		}
		p->kind = STPREP_MULTIDET;
		p->data = &p->md;
		return 0;
	}
	p->cm.nqb = nqb;
	p->cm.n_sites = in->n_sites;
	p->cm.n_alpha = in->n_alpha;
	p->cm.n_beta = in->n_beta;
	p->cm.closed_shell = in->closed_shell;
	p->cm.tapered = in->tapered;
	p->cm.C_alpha = copy_of(in->Ca, (size_t)in->n_sites * in->n_alpha);
	if (!in->closed_shell)
		p->cm.C_beta = copy_of(in->Cb, (size_t)in->n_sites * in->n_beta);
	p->kind = STPREP_COEFF_MATRIX;
	p->data = &p->cm;
	return p->cm.C_alpha ? 0 : -1;
}

static void results(struct circ *ct, double *out, double *state)
{
	for (size_t i = 0; i < ct->vals.len; i++) {
		out[2 * i] = creal(ct->vals.z[i]);
		out[2 * i + 1] = cimag(ct->vals.z[i]);
	}
	if (state)
		ph2hip_d2h(state, qreg_hip_device_ptr(&ct->reg),
			ct->reg.namp * 16);
}

int dropin_trott(int order, uint32_t nqb, size_t n, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct ph2_prep *prep,
	double delta, size_t steps, double *out, double *wrote, double *state)
{
	struct circ_hamil hm;
	struct payload p;
	struct recorder rec = { wrote, steps, 0 };
	struct phase2_step_writer sw = { &rec, record };
	if (ph2_init() < 0 || hamil(&hm, nqb, n, xm, zm, cf) < 0)
		return -1;
	if (payload(&p, prep, nqb) < 0)
		return -1;
	int rc = -1;
	if (order == 1) {
		struct trott tt;
		struct trott_data dt = { .steps = steps, .delta = delta };
		if (trott_init(&tt, &dt, hm, p.kind, p.data, &sw) < 0)
			return -1;
		rc = trott_simul(&tt);
		results(&tt.ct, out, state);
		trott_free(&tt);
	} else {
		struct trott2 t2;
		struct trott2_data dt = { .steps = steps, .delta = delta };
		if (trott2_init(&t2, &dt, hm, p.kind, p.data, &sw) < 0)
			return -1;
		rc = trott2_simul(&t2);
		results(&t2.ct, out, state);
		trott2_free(&t2);
	}
	return rc == 0 && rec.calls == steps ? 0 : -1;
}

int dropin_qdrift(uint32_t nqb, size_t n, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct ph2_prep *prep,
	size_t samples, size_t depth, double step_size, uint64_t seed,
	double *out, double *wrote, double *state)
{
	struct circ_hamil hm;
	struct payload p;
	struct recorder rec = { wrote, samples, 0 };
	struct phase2_step_writer sw = { &rec, record };
	if (ph2_init() < 0 || hamil(&hm, nqb, n, xm, zm, cf) < 0)
		return -1;
	if (payload(&p, prep, nqb) < 0)
		return -1;
	struct qdrift qd;
	struct qdrift_data dt = { .samples = samples, .depth = depth,
		.step_size = step_size, .seed = seed };
	if (qdrift_init(&qd, &dt, hm, p.kind, p.data, &sw) < 0)
		return -1;
	const int rc = qdrift_simul(&qd);
	results(&qd.ct, out, state);
	qdrift_free(&qd);
	return rc == 0 && rec.calls == samples ? 0 : -1;
}

int dropin_cmpsit(uint32_t nqb, size_t n, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct ph2_prep *prep,
	size_t samples, size_t steps, size_t length, size_t depth,
	double angle_det, double angle_rand, uint64_t seed, double *out,
	double *wrote, double *state)
{
	struct circ_hamil hm;
	struct payload p;
	struct recorder rec = { wrote, samples, 0 };
	struct phase2_step_writer sw = { &rec, record };
	if (ph2_init() < 0 || hamil(&hm, nqb, n, xm, zm, cf) < 0)
		return -1;
	if (payload(&p, prep, nqb) < 0)
		return -1;
	struct cmpsit cs;
	struct cmpsit_data dt = { .samples = samples, .steps = steps,
		.length = length, .depth = depth, .angle_det = angle_det,
		.angle_rand = angle_rand, .seed = seed };
	if (cmpsit_init(&cs, &dt, hm, p.kind, p.data, &sw) < 0)
		return -1;
	const int rc = cmpsit_simul(&cs);
	results(&cs.ct, out, state);
	cmpsit_free(&cs);
	return rc == 0 && rec.calls == samples ? 0 : -1;
}
