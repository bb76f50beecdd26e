/*
 * Learned fusion schedule (include/phase2_learned.h).
 */
#define LOG_SUBSYS "sched"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "log.h"
#include "phase2_learned.h"

#define MAXLEN PH2HIP_PASS_MAX
#define CACHE_SLOTS 4096

struct cand {
	uint32_t i, j, tb, low;
};

struct entry {
	uint64_t key;
	uint32_t nqb, n;
	struct ph2_sched s;
};

static struct {
	struct ph2hip_mlp *mlp;
	uint32_t low, min_qb;
	struct entry *tab;
	uint64_t hits, misses;
	float *feat, *pred;
	struct cand *cand;
	size_t cap;
} LM;

/* candidate tile widths: both have the half-tile form (tile 11 was rarely
 * chosen once tile 12 had it) */
static const uint32_t TILES[] = { 12, 13 };
#define NLOW 3	/* contiguous low bits low, low+1, low+2 */

void ph2_learned_features(const struct ph2hip_rot *rot, size_t n,
	uint32_t tile_bits, const uint32_t *shape, float *f)
{
	uint64_t xu = 0, zu = 0;
	for (size_t k = 0; k < n; k++) {
		xu |= rot[k].x;
		zu |= rot[k].z;
	}
	for (uint32_t b = 0; b < PH2_LM_NQ; b++) {
		f[b] = (float)(xu >> b & 1);
		f[PH2_LM_NQ + b] = (float)(zu >> b & 1);
	}
	f[2 * PH2_LM_NQ + 0] = (float)n / 48.0f;
	f[2 * PH2_LM_NQ + 1] = (float)shape[2] / 48.0f;
	f[2 * PH2_LM_NQ + 2] = (float)shape[1] / 13.0f;
	f[2 * PH2_LM_NQ + 3] = (float)shape[0] / 13.0f;
	f[2 * PH2_LM_NQ + 4] = (float)tile_bits / 13.0f;
	f[2 * PH2_LM_NQ + 5] = (float)(shape[3] & 1u);
	f[2 * PH2_LM_NQ + 6] = (float)(shape[3] >> 1 & 1u);
	f[2 * PH2_LM_NQ + 7] = (float)(shape[3] >> 2 & 1u);
}

static uint64_t block_key(const struct ph2hip_rot *rot, size_t n,
	uint32_t nqb)
{
	uint64_t h = UINT64_C(0xcbf29ce484222325) ^ nqb ^ ((uint64_t)n << 8);
	for (size_t k = 0; k < n; k++) {
		h = (h ^ rot[k].x) * UINT64_C(0x100000001b3);
		h = (h ^ rot[k].z) * UINT64_C(0x100000001b3);
	}
	return h | 1; /* 0 marks an empty slot */
}

static void cache_clear(void)
{
	if (!LM.tab)
		return;
	for (size_t k = 0; k < CACHE_SLOTS; k++) {
		ph2_sched_free(&LM.tab[k].s);
		LM.tab[k].key = 0;
	}
}

static int sched_copy(struct ph2_sched *dst, const struct ph2_sched *src)
{
	if (ph2_sched_reserve(dst, src->npass) < 0)
		return -1;
	memcpy(dst->len, src->len, sizeof *dst->len * src->npass);
	memcpy(dst->tile_bits, src->tile_bits,
		sizeof *dst->tile_bits * src->npass);
	memcpy(dst->low_bits, src->low_bits,
		sizeof *dst->low_bits * src->npass);
	dst->npass = src->npass;
	return 0;
}

static int reserve(size_t m)
{
	if (m <= LM.cap)
		return 0;
	free(LM.feat);
	free(LM.pred);
	free(LM.cand);
	LM.feat = malloc(sizeof(float) * PH2_LM_NIN * m);
	LM.pred = malloc(sizeof(float) * m);
	LM.cand = malloc(sizeof(struct cand) * m);
	LM.cap = (LM.feat && LM.pred && LM.cand) ? m : 0;
	return LM.cap ? 0 : -1;
}

static inline int topbit(uint64_t v)
{
	return 63 - __builtin_clzll(v);
}

/* Candidate passes from every start, for every tile width and contiguous
 * low-bit count: the end just before each term that widens the span, the
 * longest feasible end, and the single term. */
static size_t candidates(const struct ph2hip_rot *rot, size_t n,
	uint32_t nqb, uint32_t low0, struct cand *out, size_t cap)
{
	size_t m = 0;
	for (size_t i = 0; i < n; i++) {
		for (size_t c = 0; c < NLOW * sizeof TILES / sizeof *TILES; c++) {
			const size_t t = c / NLOW;
			const uint32_t dl = (uint32_t)(c % NLOW);
			const uint32_t tb = TILES[t] < nqb ? TILES[t] : nqb;
			/* narrow registers clamp both widths to nqb: once */
			if (tb < PH2HIP_TILE_MIN || (t > 0 && tb ==
				    (TILES[t - 1] < nqb ? TILES[t - 1] : nqb)))
				continue;
			/* one tile bit at least is left for the flip masks */
			const uint32_t low = low0 + dl < tb ? low0 + dl : tb - 1;
			if (dl > 0 && low == (low0 + dl - 1 < tb ?
				    low0 + dl - 1 : tb - 1))
				continue;
			const uint64_t lowmask = (UINT64_C(1) << low) - 1;
			const uint32_t gmax = tb - low;
			uint64_t basis[64];
			uint32_t g = 0;
			size_t j = i;
			for (; j < n && j - i < MAXLEN; j++) {
				uint64_t v = rot[j].x & ~lowmask;
				for (uint32_t b = 0; b < g && v; b++)
					if (v >> topbit(basis[b]) & 1)
						v ^= basis[b];
				if (v) {
					if (g == gmax)
						break;
					if (j > i + 1 && m < cap)
						out[m++] = (struct cand){ (uint32_t)i,
							(uint32_t)j, tb, low };
					const int p = topbit(v);
					for (uint32_t b = 0; b < g; b++)
						if (basis[b] >> p & 1)
							basis[b] ^= v;
					basis[g++] = v;
				}
			}
			if (j > i + 1 && m < cap)
				out[m++] = (struct cand){ (uint32_t)i, (uint32_t)j,
					tb, low };
			if (m < cap)
				out[m++] = (struct cand){ (uint32_t)i,
					(uint32_t)(i + 1), tb, low };
		}
	}
	return m;
}

/* A candidate the scheduler considers: it fits its tile, and it is not a
 * multi-group tile-13 pass without the half-tile form (such a pass runs
 * one 128 KiB workgroup per CU; random cuts, the training data, never
 * produce one, so its prediction would be a guess; every start keeps its
 * tile-12 candidates).  shape gets ph2hip_pass_shape's four entries. */
static int keep_candidate(const struct ph2hip_rot *rot, struct cand cd,
	uint32_t *shape)
{
	if (ph2hip_pass_shape(rot + cd.i, cd.j - cd.i, cd.tb, cd.low,
		    shape) < 0)
		return 0;
	return !(cd.tb == 13 && shape[2] > 1 && !(shape[3] & 4u));
}

long ph2_learned_candidates(const struct ph2hip_rot *rot, size_t n,
	uint32_t nqb, uint32_t low_bits, uint32_t *out, size_t cap)
{
	const size_t m0 = n * 2 * NLOW * (MAXLEN / 4 + 3) + 16;
	struct cand *c = malloc(sizeof *c * m0);
	if (!c)
		return -1;
	const size_t m = candidates(rot, n, nqb, low_bits, c, m0);
	size_t k = 0;
	for (size_t q = 0; q < m; q++) {
		uint32_t shape[4];
		if (!keep_candidate(rot, c[q], shape))
			continue;
		if (k == cap) {
			free(c);
			return -1;
		}
		out[4 * k] = c[q].i;
		out[4 * k + 1] = c[q].j;
		out[4 * k + 2] = c[q].tb;
		out[4 * k + 3] = c[q].low;
		k++;
	}
	free(c);
	return (long)k;
}

static double to_ms(float y, uint32_t nqb)
{
	return exp((double)y + ((double)nqb - 30.0) * 0.69314718055994530942);
}

int ph2_learned_schedule(const struct ph2hip_rot *rot, size_t n, uint32_t nqb,
	struct ph2_sched *out)
{
	if (!LM.mlp)
		return -1;
	const uint64_t key = block_key(rot, n, nqb);
	struct entry *e = &LM.tab[key % CACHE_SLOTS];
	if (e->key == key && e->nqb == nqb && e->n == n) {
		LM.hits++;
		return sched_copy(out, &e->s);
	}
	LM.misses++;
	const size_t cap = n * 2 * NLOW * (MAXLEN / 4 + 3) + 16;
	if (reserve(cap) < 0)
		return -1;
	size_t m = candidates(rot, n, nqb, LM.low, LM.cand, cap);
	/* shapes and features; drop candidates that do not fit */
	size_t k = 0;
	for (size_t c = 0; c < m; c++) {
		uint32_t shape[4];
		const struct cand cd = LM.cand[c];
		if (!keep_candidate(rot, cd, shape))
			continue;
		LM.cand[k] = cd;
		ph2_learned_features(rot + cd.i, cd.j - cd.i, cd.tb, shape,
			LM.feat + k * PH2_LM_NIN);
		k++;
	}
	m = k;
	if (m == 0 || ph2hip_mlp_forward(LM.mlp, LM.feat, m, LM.pred) < 0)
		return -1;
	/* shortest path over cut points, in predicted milliseconds */
	double *best = malloc(sizeof(double) * (n + 1));
	size_t *from = malloc(sizeof(size_t) * (n + 1));
	if (!best || !from) {
		free(best);
		free(from);
		return -1;
	}
	for (size_t j = 0; j <= n; j++) {
		best[j] = 1e300;
		from[j] = SIZE_MAX;
	}
	best[0] = 0.0;
	/* candidates are grouped by ascending start */
	for (size_t c = 0; c < m; c++) {
		const struct cand cd = LM.cand[c];
		const double v = best[cd.i] + to_ms(LM.pred[c], nqb);
		if (v < best[cd.j]) {
			best[cd.j] = v;
			from[cd.j] = c;
		}
	}
	int rc = -1;
	if (from[n] != SIZE_MAX) {
		size_t np = 0;
		for (size_t j = n; j > 0; j = LM.cand[from[j]].i)
			np++;
		struct ph2_sched s = { 0 };
		if (ph2_sched_reserve(&s, np) == 0) {
			s.npass = np;
			size_t p = np;
			for (size_t j = n; j > 0; j = LM.cand[from[j]].i) {
				const struct cand cd = LM.cand[from[j]];
				p--;
				s.len[p] = cd.j - cd.i;
				s.tile_bits[p] = cd.tb;
				s.low_bits[p] = cd.low;
			}
			rc = sched_copy(out, &s);
			if (rc == 0) {
				ph2_sched_free(&e->s);
				e->s = s;
				e->key = key;
				e->nqb = nqb;
				e->n = (uint32_t)n;
			} else {
				ph2_sched_free(&s);
			}
		} else {
			ph2_sched_free(&s);
		}
	}
	free(best);
	free(from);
	return rc;
}

int ph2_learned_predict(const struct ph2hip_rot *rot, size_t n, uint32_t nqb,
	const struct ph2_sched *s, double *ms)
{
	if (!LM.mlp || reserve(s->npass + 1) < 0)
		return -1;
	size_t k0 = 0;
	for (size_t p = 0; p < s->npass; p++) {
		uint32_t shape[4];
		if (k0 + s->len[p] > n ||
		    ph2hip_pass_shape(rot + k0, s->len[p], s->tile_bits[p],
			    s->low_bits[p], shape) < 0)
			return -1;
		ph2_learned_features(rot + k0, s->len[p], s->tile_bits[p], shape,
			LM.feat + p * PH2_LM_NIN);
		k0 += s->len[p];
	}
	if (ph2hip_mlp_forward(LM.mlp, LM.feat, s->npass, LM.pred) < 0)
		return -1;
	for (size_t p = 0; p < s->npass; p++)
		ms[p] = to_ms(LM.pred[p], nqb);
	return 0;
}

static int hook(void *ctx, const struct ph2hip_rot *rot, size_t n,
	uint32_t nqb, struct ph2_sched *out)
{
	(void)ctx;
	if (nqb < LM.min_qb)
		return -1; /* greedy */
	return ph2_learned_schedule(rot, n, nqb, out);
}

void ph2_learned_uninstall(void)
{
	ph2_sched_install(NULL, NULL);
	cache_clear();
	free(LM.tab);
	free(LM.feat);
	free(LM.pred);
	free(LM.cand);
	ph2hip_mlp_destroy(LM.mlp);
	memset(&LM, 0, sizeof LM);
}

int ph2_learned_install(const float *params, uint32_t nhid, uint32_t low_bits,
	uint32_t min_qb)
{
	ph2_learned_uninstall();
	if (ph2hip_mlp_create(&LM.mlp, PH2_LM_NIN, nhid, 8192, 1) < 0 ||
	    ph2hip_mlp_set_params(LM.mlp, params) < 0) {
		log_error("learned schedule: %s", ph2hip_mlp_strerror());
		ph2_learned_uninstall();
		return -1;
	}
	LM.tab = calloc(CACHE_SLOTS, sizeof *LM.tab);
	if (!LM.tab) {
		ph2_learned_uninstall();
		return -1;
	}
	LM.low = low_bits;
	LM.min_qb = min_qb;
	ph2_sched_install(hook, NULL);
	return 0;
}

void ph2_learned_stats(uint64_t *hits, uint64_t *misses)
{
	*hits = LM.hits;
	*misses = LM.misses;
}
