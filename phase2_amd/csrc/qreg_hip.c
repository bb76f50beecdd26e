/*
 * BACKEND=hip register (replaces phase2/qreg.c, phase2/qreg_qreg.c and
 * phase2/qreg_cuda.c of the reference on one GPU).
 *
 * One hipMalloc buffer holds the whole register; with one process there is
 * no hi/lo split (qb_hi = 0), no partner exchange and no mix/add passes:
 * the reference's single-rank flush (self-exchange, mix with bm = 1, rotate,
 * add 0) reduces exactly to in-order rotations, which is what
 * qreg_paulirot applies here.  cos/sin come from the host libm per term as
 * in phase2/qreg_qreg.c:189, through sincos(): gcc folds the reference's
 * adjacent cos() and sin() of one angle into a single glibc sincos() call,
 * whose results differ from separate cos()/sin() calls by one ulp for some
 * angles, so calling sincos() explicitly is what reproduces the reference.
 *
 * The void entry points cannot return errors; a failed device call logs
 * and sets a sticky flag read by qreg_error().
 */
#define _GNU_SOURCE /* sincos */
#define LOG_SUBSYS "qreg"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "log.h"
#include "phase2/qreg.h"
#include "phase2/qreg_hip.h"
#include "phase2/world.h"
#include "phase2_hip.h"

struct qreg_hip {
	void *damp;
	int err;
	int mode;
	uint32_t tile_bits, low_bits;
	struct ph2hip_rot *rot;
	size_t rot_cap;
	struct ph2_sched sched;
	/* state a caller above the register keeps with it (the multidet
	 * overlap plan of circ.c), released by qreg_free */
	void *aux;
	void (*aux_free)(void *);
	/* qreg_setamp writes not yet on the device, in call order */
	uint64_t *qidx;
	double *qval;
	size_t nq, qcap;
	/* host copy of amplitudes [win_lo, win_lo + win_n) as the device
	 * holds them with the queued writes applied; win_n = 0: none */
	double *win;
	uint64_t win_lo, win_n;
};

/* One amplitude moved by itself costs a synchronous device copy (about
 * 10 µs), so a loop over a register's amplitudes through qreg_setamp /
 * qreg_getamp, as the reference's callers and tests write, would run at
 * 10^5 amplitudes per second.  Writes are queued on the host and land in
 * one scatter (later writes to an index win) before anything else touches
 * the device buffer; reads are served from a host copy of an aligned
 * window of AMP_WIN amplitudes, taken after the queue has landed, kept
 * current by later qreg_setamp calls and dropped by every operation that
 * may change the buffer. */
#define AMP_WIN (UINT64_C(1) << 14)  /* 256 KiB */
#define AMP_QMAX ((size_t)1 << 20)   /* 24 MiB of queued writes */

static int g_mode = -1;
/* tile 0: auto_tile(); low 3 measured best of 3..5 for the greedy cut with
 * the half-tile kernel (33 qubits 0.7%, 26 qubits 3.5%, composite
 * integrator at 30 qubits 7% faster than 4) */
static uint32_t g_tile = 0, g_low = 3;
static ph2_sched_fn g_sched_fn;
static void *g_sched_ctx;

/* Tile width when none is set.
 *
 * Registers whose repeated sweeps replay from a captured graph (at most
 * ph2hip_replay_max_qb() qubits, 20 by default) are bound by each pass's
 * latency, not by launches.  Measured per step of a lex-ordered 470-611
 * term sweep replayed 64 times, tile 10 is best from 12 to 18 qubits and
 * tile 11 at 19 and 20.  Against the rule below, at 14, 16, 18 and 20
 * qubits: 0.287 vs 0.322, 0.318 vs 0.339, 0.390 vs 0.507 and 0.696 vs
 * 0.765 ms.  Never above nqb - 2.
 *
 * Otherwise: 13 from 26 qubits (the half-tile form at tile 13 sweeps 26
 * qubits 4.7% faster than at tile 12, and a pass still has 8192 tiles), 12
 * from 17, nqb - 5 below.  Small registers are latency-bound (a few dozen
 * workgroups per pass); without replay, at 16 qubits a whole Trotter run
 * takes 0.476 ms per step at tile 11, 0.503 at 10, 0.532 at 9 and 0.564 at
 * 12. */
static uint32_t auto_tile(uint32_t nqb)
{
	uint32_t t;
	if (nqb <= ph2hip_replay_max_qb()) {
		t = nqb >= 19 ? 11 : 10;
		if (t + 2 > nqb)
			t = nqb > 2 ? nqb - 2 : 0;
	} else {
		t = nqb >= 26 ? 13 : nqb > 16 ? 12 : (nqb > 5 ? nqb - 5 : 0);
	}
	return t < PH2HIP_TILE_MIN ? PH2HIP_TILE_MIN : t;
}

static void load_defaults(void)
{
	if (g_mode >= 0)
		return;
	g_mode = QREG_HIP_FUSED;
	const char *m = getenv("PHASE2_HIP_MODE");
	if (m && strcmp(m, "plain") == 0)
		g_mode = QREG_HIP_PLAIN;
	const char *t = getenv("PHASE2_HIP_TILE");
	if (t) {
		const long v = strtol(t, NULL, 10);
		if (v >= PH2HIP_TILE_MIN && v <= PH2HIP_TILE_MAX)
			g_tile = (uint32_t)v;
	}
	const char *l = getenv("PHASE2_HIP_LOW");
	if (l) {
		const long v = strtol(l, NULL, 10);
		if (v >= 0 && v <= PH2HIP_TILE_MAX)
			g_low = (uint32_t)v;
	}
}

void qreg_hip_set_default(int mode, uint32_t tile_bits, uint32_t low_bits)
{
	load_defaults();
	g_mode = mode;
	g_tile = tile_bits;
	g_low = low_bits;
}

void qreg_hip_configure(struct qreg *reg, int mode, uint32_t tile_bits,
	uint32_t low_bits)
{
	struct qreg_hip *h = reg->backend;
	if (!h)
		return;
	h->mode = mode;
	h->tile_bits = tile_bits;
	h->low_bits = low_bits;
}

void ph2_sched_install(ph2_sched_fn fn, void *ctx)
{
	g_sched_fn = fn;
	g_sched_ctx = ctx;
}

/* Put the queued writes on the device and drop the read window: called
 * before anything else reads or changes the device buffer. */
static int settle(struct qreg_hip *h, uint64_t namp)
{
	h->win_n = 0;
	if (h->nq == 0)
		return 0;
	const int rc = ph2hip_scatter_set(h->damp, namp, h->qidx, h->qval,
		h->nq);
	h->nq = 0;
	if (rc < 0) {
		h->err = -1;
		log_error("qreg_setamp: %s", ph2hip_strerror());
	}
	return rc;
}

/* The caller may read or change the buffer through the pointer, so the
 * queued writes land first and the read window goes. */
void *qreg_hip_device_ptr(struct qreg *reg)
{
	struct qreg_hip *h = reg->backend;
	if (!h)
		return NULL;
	settle(h, reg->namp);
	return h->damp;
}

static void set_err(struct qreg *reg, const char *what)
{
	struct qreg_hip *h = reg->backend;
	if (h)
		h->err = -1;
	log_error("%s: %s", what, ph2hip_strerror());
}

int qreg_error(const struct qreg *reg)
{
	struct qreg_hip *h = reg->backend;
	if (!h)
		return -1;
	/* a queued write fails, if at all, when it lands: land it now */
	if (h->nq) {
		const uint64_t lo = h->win_lo, n = h->win_n;
		settle(h, reg->namp);
		h->win_lo = lo;
		h->win_n = n;
	}
	return h->err;
}

void qreg_clear_error(struct qreg *reg)
{
	struct qreg_hip *h = reg->backend;
	if (h)
		h->err = 0;
}

int qreg_init(struct qreg *reg, uint32_t qb)
{
	memset(reg, 0, sizeof *reg);
	if (world_info(&reg->wd) != WORLD_READY) {
		log_error("qreg_init: world not ready");
		return -1;
	}
	if (qb == 0 || qb > 40) {
		log_error("qreg_init: nqb=%u outside 1..40", qb);
		return -1;
	}
	load_defaults();
	struct qreg_hip *h = calloc(1, sizeof *h);
	if (!h) {
		log_error("qreg_init: alloc backend");
		return -1;
	}
	const uint64_t namp = UINT64_C(1) << qb;
	if (ph2hip_malloc(&h->damp, namp * 16) < 0) {
		log_error("qreg_init: device alloc of %llu bytes: %s",
			(unsigned long long)(namp * 16), ph2hip_strerror());
		free(h);
		return -1;
	}
	if (qb <= QREG_SHADOW_MAX_QB) {
		reg->amp = malloc(sizeof(_Complex double) * namp);
		if (!reg->amp) {
			log_error("qreg_init: host shadow alloc");
			ph2hip_free(h->damp);
			free(h);
			return -1;
		}
	}
	h->mode = g_mode;
	h->tile_bits = g_tile;
	h->low_bits = g_low;
	reg->qb_lo = qb;
	reg->qb_hi = 0;
	reg->buf = NULL;
	reg->namp = namp;
	reg->backend = h;
	return 0;
}

void *qreg_hip_aux(struct qreg *reg)
{
	const struct qreg_hip *h = reg->backend;
	return h ? h->aux : NULL;
}

int qreg_hip_set_aux(struct qreg *reg, void *aux, void (*aux_free)(void *))
{
	struct qreg_hip *h = reg->backend;
	if (!h)
		return -1;
	if (h->aux && h->aux_free)
		h->aux_free(h->aux);
	h->aux = aux;
	h->aux_free = aux_free;
	return 0;
}

void qreg_free(struct qreg *reg)
{
	struct qreg_hip *h = reg->backend;
	if (h) {
		if (h->aux && h->aux_free)
			h->aux_free(h->aux);
		ph2hip_free(h->damp);
		free(h->rot);
		ph2_sched_free(&h->sched);
		free(h->qidx);
		free(h->qval);
		free(h->win);
		free(h);
	}
	free(reg->amp);
	reg->amp = NULL;
	reg->backend = NULL;
}

/* The reference masks an index to the register (phase2/qreg.c:35-47), so
 * on one rank an index beyond 2^qb silently aliases a lower one; here it is
 * an error (sticky flag, nothing read or written). */
static int in_range(struct qreg *reg, uint64_t i, const char *what)
{
	if (i < reg->namp)
		return 1;
	struct qreg_hip *h = reg->backend;
	h->err = -1;
	log_error("%s: index %llu beyond %u qubits", what,
		(unsigned long long)i, reg->qb_lo);
	return 0;
}

void qreg_getamp(struct qreg *reg, uint64_t i, _Complex double *z)
{
	struct qreg_hip *h = reg->backend;
	*z = 0.0;
	if (!in_range(reg, i, "qreg_getamp"))
		return;
	if (!h->win_n || i - h->win_lo >= h->win_n) {
		if (settle(h, reg->namp) < 0)
			return;
		if (!h->win && !(h->win = malloc(AMP_WIN * 16))) {
			log_error("qreg_getamp: alloc");
			h->err = -1;
			return;
		}
		const uint64_t lo = i & ~(AMP_WIN - 1);
		const uint64_t n = reg->namp - lo < AMP_WIN ? reg->namp - lo :
							       AMP_WIN;
		if (ph2hip_d2h(h->win, (const char *)h->damp + lo * 16,
			    n * 16) < 0) {
			set_err(reg, "qreg_getamp");
			return;
		}
		h->win_lo = lo;
		h->win_n = n;
	}
	const double *v = h->win + 2 * (i - h->win_lo);
	*z = CMPLX(v[0], v[1]);
}

void qreg_setamp(struct qreg *reg, uint64_t i, _Complex double z)
{
	struct qreg_hip *h = reg->backend;
	if (!in_range(reg, i, "qreg_setamp"))
		return;
	if (h->nq == h->qcap) {
		if (h->qcap >= AMP_QMAX) {
			/* the window already holds these writes: keep it */
			const uint64_t lo = h->win_lo, n = h->win_n;
			if (settle(h, reg->namp) < 0)
				return;
			h->win_lo = lo;
			h->win_n = n;
		} else {
			const size_t cap = h->qcap ? 2 * h->qcap : 1024;
			uint64_t *qi = realloc(h->qidx, sizeof *qi * cap);
			if (qi)
				h->qidx = qi;
			double *qv = realloc(h->qval, 2 * sizeof *qv * cap);
			if (qv)
				h->qval = qv;
			if (!qi || !qv) {
				log_error("qreg_setamp: alloc");
				h->err = -1;
				return;
			}
			h->qcap = cap;
		}
	}
	h->qidx[h->nq] = i;
	h->qval[2 * h->nq] = creal(z);
	h->qval[2 * h->nq + 1] = cimag(z);
	h->nq++;
	if (h->win_n && i - h->win_lo < h->win_n) {
		h->win[2 * (i - h->win_lo)] = creal(z);
		h->win[2 * (i - h->win_lo) + 1] = cimag(z);
	}
}

void qreg_zero(struct qreg *reg)
{
	struct qreg_hip *h = reg->backend;
	h->nq = 0; /* every queued write would be overwritten */
	h->win_n = 0;
	if (ph2hip_zero(h->damp, reg->namp) < 0)
		set_err(reg, "qreg_zero");
}

void qreg_sync_host_to_device(struct qreg *reg)
{
	struct qreg_hip *h = reg->backend;
	if (!reg->amp) {
		log_error("qreg_sync_host_to_device: no host shadow above %d "
			  "qubits", QREG_SHADOW_MAX_QB);
		h->err = -1;
		return;
	}
	h->nq = 0; /* every queued write would be overwritten */
	h->win_n = 0;
	if (ph2hip_h2d(h->damp, reg->amp, reg->namp * 16) < 0)
		set_err(reg, "qreg_sync_host_to_device");
}

void qreg_sync_device_to_host(struct qreg *reg)
{
	struct qreg_hip *h = reg->backend;
	if (!reg->amp) {
		log_error("qreg_sync_device_to_host: no host shadow above %d "
			  "qubits", QREG_SHADOW_MAX_QB);
		h->err = -1;
		return;
	}
	if (settle(h, reg->namp) < 0)
		return;
	if (ph2hip_d2h(reg->amp, h->damp, reg->namp * 16) < 0)
		set_err(reg, "qreg_sync_device_to_host");
}

void qreg_paulirot(struct qreg *reg, struct paulis code_hi,
	const struct paulis *codes_lo, const double *phis, size_t ncodes)
{
	struct qreg_hip *h = reg->backend;
	if (code_hi.pak[0] | code_hi.pak[1]) {
		log_error("qreg_paulirot: hi code on a single-GPU register");
		h->err = -1;
		return;
	}
	if (ncodes == 0)
		return;
	if (settle(h, reg->namp) < 0)
		return;
	if (ncodes > h->rot_cap) {
		struct ph2hip_rot *r = realloc(h->rot, sizeof *r * ncodes);
		if (!r) {
			log_error("qreg_paulirot: alloc");
			h->err = -1;
			return;
		}
		h->rot = r;
		h->rot_cap = ncodes;
	}
	for (size_t k = 0; k < ncodes; k++) {
		h->rot[k].x = codes_lo[k].pak[0];
		h->rot[k].z = codes_lo[k].pak[1];
		sincos(phis[k], &h->rot[k].s, &h->rot[k].c);
	}
	const uint32_t nqb = reg->qb_lo;
	int rc;
	if (h->mode == QREG_HIP_PLAIN || nqb < PH2HIP_TILE_MIN) {
		rc = ph2hip_rotate_plain(h->damp, nqb, h->rot, ncodes);
	} else {
		rc = -1;
		if (g_sched_fn)
			rc = g_sched_fn(g_sched_ctx, h->rot, ncodes, nqb,
				&h->sched);
		if (rc < 0)
			rc = ph2_sched_greedy(h->rot, ncodes, nqb,
				h->tile_bits ? h->tile_bits : auto_tile(nqb),
				h->low_bits, &h->sched);
		if (rc == 0)
			rc = ph2hip_rotate_fused(h->damp, nqb, h->rot, ncodes,
				h->sched.len, h->sched.tile_bits, h->sched.npass,
				h->sched.low_bits);
	}
	if (rc < 0)
		set_err(reg, "qreg_paulirot");
}
