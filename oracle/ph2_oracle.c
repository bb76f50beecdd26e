/*
 * ph2_oracle.c -- CPU restatement of the phase2 state-vector hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP backend is compared
 * with; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * load it (as oracle/liboracle.so).  The shipped library never links it.
 *
 * It restates the single-rank (one MPI rank, qb_hi = 0) semantics of the
 * reference CPU backend, where every circ_cache flush self-exchanges the
 * slab, mixes with bm = 1 (a = (x+x)/2 = x exactly, b = 0), rotates and adds
 * 0 (phase2/qreg.c:124-142, phase2/qreg_qreg.c:83-198).  The resulting state
 * is that of sequential, in-place, per-term rotations, which is what
 * orc_rotate computes.
 *
 * Pinning (see tests/test_oracle.py): every function here is checked against
 * golden vectors written by the reference itself (oracle/_ref, built from the
 * reference sources by oracle/Makefile.ref) and against closed forms.
 *
 * Arithmetic: built with -std=c11 -ffp-contract=off, so no product is fused
 * into an add; each component update is round(round(c*x) +- round(s*y)), the
 * same as the reference's _Complex double expression evaluated by gcc.
 */
#define _GNU_SOURCE /* sincos */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Pauli algebra (include/phase2/paulis.h:131-146)                    */
/* ------------------------------------------------------------------ */

/* Phase exponent r4 of P|i> = i^r4 |i ^ X>: Y count plus twice the number
 * of set bits of i under Z (include/phase2/paulis.h:137-144). */
static inline int orc_r4(uint64_t xm, uint64_t zm, uint64_t i)
{
	return (__builtin_popcountll(xm & zm) +
		2 * __builtin_popcountll(i & zm)) & 3;
}

/* One rotation exp(i*phi*P) applied in place to a (2*2^nqb doubles).
 * Pair (i, j = i^X) is updated once, from the smaller index i, with
 * z = i^r4(i) (phase2/qreg_qreg.c:149-160):
 *     a[i] <- c*a[i] + i*conj(s*z)*a[j]
 *     a[j] <- c*a[j] + i*(s*z)*a[i]
 * For X = 0 the second assignment overwrites the first, so the diagonal
 * update is a[i] <- c*a[i] + i*(s*z)*a[i]. */
static void orc_rotate_one(double *a, uint64_t namp, uint64_t xm,
	uint64_t zm, double c, double s)
{
	for (uint64_t i = 0; i < namp; i++) {
		const uint64_t j = i ^ xm;
		if (j < i)
			continue;
		const int r4 = orc_r4(xm, zm, i);
		const double xr = a[2 * i], xi = a[2 * i + 1];
		const double yr = a[2 * j], yi = a[2 * j + 1];
		/* u = i*conj(s z) multiplies the partner of i, v = i*(s z)
		 * multiplies the partner of j.
		 *   r4=0: u = +is, v = +is     r4=1: u = +s,  v = -s
		 *   r4=2: u = -is, v = -is     r4=3: u = -s,  v = +s   */
		double ur, ui, vr, vi; /* u*y and v*x, component-wise */
		switch (r4) {
		case 0:
			ur = -(s * yi); ui = s * yr;
			vr = -(s * xi); vi = s * xr;
			break;
		case 1:
			ur = s * yr; ui = s * yi;
			vr = -(s * xr); vi = -(s * xi);
			break;
		case 2:
			ur = s * yi; ui = -(s * yr);
			vr = s * xi; vi = -(s * xr);
			break;
		default:
			ur = -(s * yr); ui = -(s * yi);
			vr = s * xr; vi = s * xi;
			break;
		}
		if (j == i) {
			a[2 * i] = c * xr + vr;
			a[2 * i + 1] = c * xi + vi;
		} else {
			a[2 * i] = c * xr + ur;
			a[2 * i + 1] = c * xi + ui;
			a[2 * j] = c * yr + vr;
			a[2 * j + 1] = c * yi + vi;
		}
	}
}

/* cos and sin of one angle as the reference obtains them: gcc turns the
 * adjacent cos()/sin() of phase2/qreg_qreg.c:189 into one glibc sincos()
 * call (it differs from separate calls by one ulp for some angles). */
static inline void orc_cs(double phi, double *c, double *s)
{
	sincos(phi, s, c);
}

/* Sequential rotations (phase2/qreg_qreg.c:181-198 at one rank). */
void orc_rotate(double *a, uint32_t nqb, size_t n, const uint64_t *xm,
	const uint64_t *zm, const double *phi)
{
	const uint64_t namp = UINT64_C(1) << nqb;
	double c, s;
	for (size_t k = 0; k < n; k++) {
		orc_cs(phi[k], &c, &s);
		orc_rotate_one(a, namp, xm[k], zm[k], c, s);
	}
}

/* ------------------------------------------------------------------ */
/* Lexicographic order (phase2/paulis.c:190-205, phase2/circ.c:60-78) */
/* ------------------------------------------------------------------ */

/* Per-qubit rank I<X<Y<Z from the (flip, phase) bits. */
static inline int qrank(uint64_t x, uint64_t z, int q)
{
	const int xb = (int)(x >> q & 1), zb = (int)(z >> q & 1);
	return zb ? (xb ? 2 : 3) : xb;
}

static int lex_cmp(uint64_t xa, uint64_t za, uint64_t xb, uint64_t zb)
{
	if (xa == xb && za == zb)
		return 0;
	for (int q = 63; q >= 0; q--) {
		const int ra = qrank(xa, za, q), rb = qrank(xb, zb, q);
		if (ra != rb)
			return ra < rb ? -1 : 1;
	}
	return 0;
}

struct orc_term {
	double cf;
	uint64_t x, z;
};

/* Stable merge sort.  glibc 2.35 qsort (the reference's sort) is a merge
 * sort for arrays that fit in memory, so ties keep input order there too. */
static void msort_terms(struct orc_term *t, struct orc_term *tmp, size_t n,
	int (*cmp)(const struct orc_term *, const struct orc_term *))
{
	if (n < 2)
		return;
	const size_t h = n / 2;
	msort_terms(t, tmp, h, cmp);
	msort_terms(t + h, tmp, n - h, cmp);
	size_t i = 0, j = h, k = 0;
	while (i < h && j < n)
		tmp[k++] = cmp(&t[j], &t[i]) < 0 ? t[j++] : t[i++];
	while (i < h)
		tmp[k++] = t[i++];
	while (j < n)
		tmp[k++] = t[j++];
	memcpy(t, tmp, sizeof *t * n);
}

static int cmp_lex(const struct orc_term *a, const struct orc_term *b)
{
	return lex_cmp(a->x, a->z, b->x, b->z);
}

/* |cf| descending (circ/cmpsit.c:37-48). */
static int cmp_abs_desc(const struct orc_term *a, const struct orc_term *b)
{
	const double x = fabs(a->cf), y = fabs(b->cf);
	return x > y ? -1 : (x < y ? 1 : 0);
}

static int sort_terms(struct orc_term *t, size_t n,
	int (*cmp)(const struct orc_term *, const struct orc_term *))
{
	struct orc_term *tmp = malloc(sizeof *tmp * (n ? n : 1));
	if (!tmp)
		return -1;
	msort_terms(t, tmp, n, cmp);
	free(tmp);
	return 0;
}

/* perm[k] = input index of the k-th term in lex order. */
int orc_sort_lex(size_t n, const uint64_t *xm, const uint64_t *zm,
	uint64_t *perm)
{
	struct orc_term *t = malloc(sizeof *t * (n ? n : 1));
	if (!t)
		return -1;
	for (size_t k = 0; k < n; k++) {
		t[k].cf = (double)k;
		t[k].x = xm[k];
		t[k].z = zm[k];
	}
	int rc = sort_terms(t, n, cmp_lex);
	for (size_t k = 0; k < n; k++)
		perm[k] = (uint64_t)t[k].cf;
	free(t);
	return rc;
}

/* ------------------------------------------------------------------ */
/* xoshiro256** (lib/xoshiro256ss.c:34-85; Blackman & Vigna 2018)    */
/* ------------------------------------------------------------------ */

struct orc_rng {
	uint64_t s[4];
};

static inline uint64_t rotl64(uint64_t v, int k)
{
	return (v << k) | (v >> (64 - k));
}

static uint64_t rng_next(struct orc_rng *g)
{
	const uint64_t out = rotl64(g->s[1] * 5, 7) * 9;
	const uint64_t t = g->s[1] << 17;
	g->s[2] ^= g->s[0];
	g->s[3] ^= g->s[1];
	g->s[1] ^= g->s[2];
	g->s[0] ^= g->s[3];
	g->s[2] ^= t;
	g->s[3] = rotl64(g->s[3], 45);
	return out;
}

/* splitmix64 fills the four words, then 128 outputs are discarded. */
static void rng_seed(struct orc_rng *g, uint64_t seed)
{
	uint64_t sm = seed;
	for (int w = 0; w < 4; w++) {
		uint64_t v = (sm += UINT64_C(0x9e3779b97f4a7c15));
		v = (v ^ (v >> 30)) * UINT64_C(0xbf58476d1ce4e5b9);
		v = (v ^ (v >> 27)) * UINT64_C(0x94d049bb133111eb);
		g->s[w] = v ^ (v >> 31);
	}
	for (int k = 0; k < 128; k++)
		rng_next(g);
}

/* Uniform double in [0,1): top 53 bits times 2^-53
 * (include/xoshiro256ss.h:58-59). */
static inline double rng_dbl01(struct orc_rng *g)
{
	return (double)(rng_next(g) >> 11) * 0x1.0p-53;
}

void orc_xoshiro(uint64_t seed, size_t n, uint64_t *out)
{
	struct orc_rng g;
	rng_seed(&g, seed);
	for (size_t i = 0; i < n; i++)
		out[i] = rng_next(&g);
}

/* ------------------------------------------------------------------ */
/* Discrete CDF and its inverse (lib/prob.c:39-78)                     */
/* ------------------------------------------------------------------ */

/* y[i] = running sum of |w_k| / lambda, added term by term. */
static int cdf_build(double *y, size_t n, const double *w, size_t stride,
	double *lambda_out)
{
	double lambda = 0.0;
	for (size_t i = 0; i < n; i++) {
		y[i] = fabs(*(const double *)((const char *)w + i * stride));
		lambda += y[i];
	}
	if (lambda < DBL_EPSILON)
		return -1;
	double acc = 0.0;
	for (size_t i = 0; i < n; i++) {
		acc += y[i] / lambda;
		y[i] = acc;
	}
	if (lambda_out)
		*lambda_out = lambda;
	return 0;
}

/* Binary descent over halving strides, then a linear walk forward while
 * y[i] <= u, clamped at n-1. */
static size_t cdf_inverse(const double *y, size_t n, double u)
{
	size_t i = 0;
	for (size_t d = n / 2; d > 0; d /= 2)
		if (y[i + d] <= u)
			i += d;
	while (i + 1 < n && y[i] <= u)
		i++;
	return i;
}

int orc_cdf_inverse(size_t nw, const double *w, size_t ny, const double *ys,
	uint64_t *idx, double *lambda)
{
	double *y = malloc(sizeof *y * (nw ? nw : 1));
	if (!y || nw == 0)
		return free(y), -1;
	if (cdf_build(y, nw, w, sizeof(double), lambda) < 0)
		return free(y), -1;
	for (size_t k = 0; k < ny; k++)
		idx[k] = cdf_inverse(y, nw, ys[k]);
	free(y);
	return 0;
}

/* Sign with a DBL_EPSILON dead zone (circ/internal.h:26-32). */
static inline double orc_signof(double v)
{
	if (fabs(v) < DBL_EPSILON)
		return 0.0;
	return v < 0 ? -1.0 : 1.0;
}

/* ------------------------------------------------------------------ */
/* Slater-Condon (phase2/state_prep_coeff.c:31-274, lib/det_small.c)   */
/* ------------------------------------------------------------------ */

/* Determinant by Gaussian elimination with partial pivoting, product of
 * pivots with the swap sign; rows below the pivot are updated only when
 * their multiplier is non-zero (lib/det_small.c:26-74). */
double orc_det(const double *A, uint32_t n)
{
	if (n == 0)
		return 1.0;
	if (n > 32)
		return 0.0;
	double m[32 * 32];
	memcpy(m, A, sizeof(double) * n * n);
	double det = 1.0;
	for (uint32_t k = 0; k < n; k++) {
		uint32_t p = k;
		double best = fabs(m[k * n + k]);
		for (uint32_t r = k + 1; r < n; r++)
			if (fabs(m[r * n + k]) > best) {
				best = fabs(m[r * n + k]);
				p = r;
			}
		if (best == 0.0)
			return 0.0;
		if (p != k) {
			for (uint32_t c = k; c < n; c++) {
				double t = m[k * n + c];
				m[k * n + c] = m[p * n + c];
				m[p * n + c] = t;
			}
			det = -det;
		}
		const double piv = m[k * n + k];
		det *= piv;
		const double inv = 1.0 / piv;
		for (uint32_t r = k + 1; r < n; r++) {
			const double f = m[r * n + k] * inv;
			if (f == 0.0)
				continue;
			for (uint32_t c = k + 1; c < n; c++)
				m[r * n + c] -= f * m[k * n + c];
		}
	}
	return det;
}

/* All k-subsets of {0..n-1} in lexicographic order, as occupation masks
 * (itertools.combinations order). */
static size_t subsets(uint32_t n, uint32_t k, uint64_t *masks,
	uint32_t *tuples)
{
	uint32_t c[34];
	for (uint32_t i = 0; i < k; i++)
		c[i] = i;
	size_t m = 0;
	for (;;) {
		uint64_t v = 0;
		for (uint32_t i = 0; i < k; i++) {
			v |= UINT64_C(1) << c[i];
			if (tuples)
				tuples[m * k + i] = c[i];
		}
		masks[m++] = v;
		if (k == 0)
			break;
		int j = (int)k - 1;
		while (j >= 0 && c[j] == n - k + (uint32_t)j)
			j--;
		if (j < 0)
			break;
		c[j]++;
		for (uint32_t i = (uint32_t)j + 1; i < k; i++)
			c[i] = c[i - 1] + 1;
	}
	return m;
}

static uint64_t choose(uint32_t n, uint32_t k)
{
	if (k > n)
		return 0;
	uint64_t r = 1;
	for (uint32_t i = 1; i <= k; i++)
		r = r * (n - k + i) / i;
	return r;
}

/* det(C[occ, :]) for every k-subset; C row-major (n_sites, k). */
static void subset_dets(uint32_t k, const double *C,
	const uint32_t *tuples, size_t M, double *dets)
{
	double sub[32 * 32];
	for (size_t t = 0; t < M; t++) {
		if (k == 0) {
			dets[t] = 1.0;
			continue;
		}
		for (uint32_t r = 0; r < k; r++)
			for (uint32_t cc = 0; cc < k; cc++)
				sub[r * k + cc] = C[tuples[t * k + r] * k + cc];
		dets[t] = orc_det(sub, k);
	}
}

/* Index transform of a tapered state: drop bits 0 and n_sites. */
static inline uint64_t taper(uint64_t v, uint32_t n_sites)
{
	const uint64_t lo = (v >> 1) & ((UINT64_C(1) << (n_sites - 1)) - 1);
	return lo | ((v >> (n_sites + 1)) << (n_sites - 1));
}

struct sc_tables {
	size_t Ma, Mb;
	uint64_t *ma, *mb;
	double *da, *db;
};

static void sc_free(struct sc_tables *t)
{
	free(t->ma);
	free(t->mb);
	free(t->da);
	free(t->db);
}

static int sc_build(struct sc_tables *t, uint32_t n_sites, uint32_t na,
	uint32_t nb, const double *Ca, const double *Cb)
{
	memset(t, 0, sizeof *t);
	if (na > n_sites || nb > n_sites || na > 32 || nb > 32)
		return -1;
	t->Ma = choose(n_sites, na);
	t->Mb = choose(n_sites, nb);
	t->ma = malloc(sizeof(uint64_t) * t->Ma);
	t->mb = malloc(sizeof(uint64_t) * t->Mb);
	t->da = malloc(sizeof(double) * t->Ma);
	t->db = malloc(sizeof(double) * t->Mb);
	uint32_t *tup = malloc(sizeof(uint32_t) *
		(t->Ma * (na ? na : 1) + t->Mb * (nb ? nb : 1)));
	if (!t->ma || !t->mb || !t->da || !t->db || !tup) {
		free(tup);
		sc_free(t);
		return -1;
	}
	uint32_t *tb = tup + t->Ma * (na ? na : 1);
	subsets(n_sites, na, t->ma, tup);
	subsets(n_sites, nb, t->mb, tb);
	subset_dets(na, Ca, tup, t->Ma, t->da);
	subset_dets(nb, Cb ? Cb : Ca, tb, t->Mb, t->db);
	free(tup);
	return 0;
}

/* Scatter weight*det_a*det_b into a (set or accumulate), skipping
 * |c| < 1e-12 (phase2/state_prep_coeff.c:181-222). */
static void sc_scatter(const struct sc_tables *t, uint32_t n_sites,
	double weight, int tapered, int accumulate, double *a, uint64_t namp)
{
	for (size_t i = 0; i < t->Ma; i++)
		for (size_t j = 0; j < t->Mb; j++) {
			const double cf = weight * t->da[i] * t->db[j];
			if (fabs(cf) < 1.0e-12)
				continue;
			uint64_t idx = t->ma[i] | (t->mb[j] << n_sites);
			if (tapered)
				idx = taper(idx, n_sites);
			idx &= namp - 1;
			if (accumulate)
				a[2 * idx] += cf;
			else
				a[2 * idx] = cf;
			if (!accumulate)
				a[2 * idx + 1] = 0.0;
		}
}

/* sum over the support of cf * a[idx], two real accumulators
 * (phase2/state_prep_coeff.c:224-274). */
static void sc_inner(const struct sc_tables *t, uint32_t n_sites,
	double weight, int tapered, const double *a, uint64_t namp,
	double *out)
{
	double ar = 0.0, ai = 0.0;
	for (size_t i = 0; i < t->Ma; i++)
		for (size_t j = 0; j < t->Mb; j++) {
			const double cf = weight * t->da[i] * t->db[j];
			if (fabs(cf) < 1.0e-12)
				continue;
			uint64_t idx = t->ma[i] | (t->mb[j] << n_sites);
			if (tapered)
				idx = taper(idx, n_sites);
			idx &= namp - 1;
			ar += cf * a[2 * idx];
			ai += cf * a[2 * idx + 1];
		}
	out[0] = ar;
	out[1] = ai;
}

int orc_expand(uint32_t nqb, uint32_t n_sites, uint32_t na, uint32_t nb,
	int closed_shell, int tapered, const double *Ca, const double *Cb,
	double *state)
{
	struct sc_tables t;
	if (sc_build(&t, n_sites, na, nb, Ca, closed_shell ? NULL : Cb) < 0)
		return -1;
	const uint64_t namp = UINT64_C(1) << nqb;
	memset(state, 0, sizeof(double) * 2 * namp);
	sc_scatter(&t, n_sites, 1.0, tapered, 0, state, namp);
	sc_free(&t);
	return 0;
}

int orc_inner(uint32_t nqb, uint32_t n_sites, uint32_t na, uint32_t nb,
	int closed_shell, int tapered, const double *Ca, const double *Cb,
	const double *state, double *out)
{
	struct sc_tables t;
	if (sc_build(&t, n_sites, na, nb, Ca, closed_shell ? NULL : Cb) < 0)
		return -1;
	sc_inner(&t, n_sites, 1.0, tapered, state, UINT64_C(1) << nqb, out);
	sc_free(&t);
	return 0;
}

/* ------------------------------------------------------------------ */
/* Algorithm drivers (circ/trott.c, trott2.c, qdrift.c, cmpsit.c)     */
/* ------------------------------------------------------------------ */

/* Initial-state description shared with the GPU flat API. */
struct orc_prep {
	int kind; /* 1 multidet, 2 coefficient matrix (single block) */
	size_t ndet;
	const uint64_t *didx;
	const double *dcf; /* interleaved re, im */
	uint32_t n_sites, n_alpha, n_beta;
	int closed_shell, tapered;
	const double *Ca, *Cb;
};

struct orc_run {
	uint32_t nqb;
	uint64_t namp;
	double *a;
	const struct orc_prep *prep;
	struct sc_tables sc;
};

static int run_open(struct orc_run *r, uint32_t nqb, const struct orc_prep *p)
{
	r->nqb = nqb;
	r->namp = UINT64_C(1) << nqb;
	r->prep = p;
	r->a = malloc(sizeof(double) * 2 * r->namp);
	if (!r->a)
		return -1;
	memset(&r->sc, 0, sizeof r->sc);
	if (p->kind == 2 && sc_build(&r->sc, p->n_sites, p->n_alpha,
				p->n_beta, p->Ca,
				p->closed_shell ? NULL : p->Cb) < 0)
		return -1;
	return 0;
}

static void run_close(struct orc_run *r)
{
	free(r->a);
	sc_free(&r->sc);
}

/* circ_prepst (phase2/circ.c:215-233). */
static void run_prep(struct orc_run *r)
{
	const struct orc_prep *p = r->prep;
	memset(r->a, 0, sizeof(double) * 2 * r->namp);
	if (p->kind == 1) {
		for (size_t d = 0; d < p->ndet; d++) {
			const uint64_t i = p->didx[d] & (r->namp - 1);
			r->a[2 * i] = p->dcf[2 * d];
			r->a[2 * i + 1] = p->dcf[2 * d + 1];
		}
	} else {
		sc_scatter(&r->sc, p->n_sites, 1.0, p->tapered, 0, r->a,
			r->namp);
	}
}

/* circ_measure (phase2/circ.c:338-361): multidet sums a*conj(cf) in
 * determinant order. */
static void run_measure(struct orc_run *r, double *out)
{
	const struct orc_prep *p = r->prep;
	if (p->kind == 1) {
		double pr = 0.0, pi = 0.0;
		for (size_t d = 0; d < p->ndet; d++) {
			const uint64_t i = p->didx[d] & (r->namp - 1);
			const double ar = r->a[2 * i], ai = r->a[2 * i + 1];
			const double cr = p->dcf[2 * d], ci = p->dcf[2 * d + 1];
			pr += ar * cr + ai * ci;
			pi += ai * cr - ar * ci;
		}
		out[0] = pr;
		out[1] = pi;
	} else {
		sc_inner(&r->sc, p->n_sites, 1.0, p->tapered, r->a, r->namp,
			out);
	}
}

/* circ_step_generic (phase2/circ.c:256-292): phi_k = omega * cf_k. */
static void run_sweep(struct orc_run *r, const struct orc_term *t, size_t n,
	double omega, int reverse)
{
	for (size_t k = 0; k < n; k++) {
		const struct orc_term *q = &t[reverse ? n - 1 - k : k];
		const double phi = omega * q->cf;
		double c, s;
		orc_cs(phi, &c, &s);
		orc_rotate_one(r->a, r->namp, q->x, q->z, c, s);
	}
}

static struct orc_term *load_terms(size_t n, const uint64_t *xm,
	const uint64_t *zm, const double *cf)
{
	struct orc_term *t = malloc(sizeof *t * (n ? n : 1));
	if (!t)
		return NULL;
	for (size_t k = 0; k < n; k++) {
		t[k].cf = cf[k];
		t[k].x = xm[k];
		t[k].z = zm[k];
	}
	return t;
}

/* algo 1: trott (circ/trott.c:44-65); algo 2: trott2 (circ/trott2.c:44-74).
 * out: 2*steps doubles; state (optional): final 2*2^nqb doubles. */
int orc_trott(int algo, uint32_t nqb, size_t nterms, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct orc_prep *prep,
	double delta, size_t steps, double *out, double *state)
{
	struct orc_term *t = load_terms(nterms, xm, zm, cf);
	struct orc_run r;
	if (!t || run_open(&r, nqb, prep) < 0)
		return free(t), -1;
	sort_terms(t, nterms, cmp_lex);
	run_prep(&r);
	for (size_t s = 0; s < steps; s++) {
		if (algo == 1) {
			run_sweep(&r, t, nterms, delta, 0);
		} else {
			run_sweep(&r, t, nterms, delta / 2.0, 0);
			run_sweep(&r, t, nterms, delta / 2.0, 1);
		}
		run_measure(&r, out + 2 * s);
	}
	if (state)
		memcpy(state, r.a, sizeof(double) * 2 * r.namp);
	run_close(&r);
	free(t);
	return 0;
}

/* qDRIFT (circ/qdrift.c:58-133): CDF over |c| in input order; one RNG
 * stream across samples; per sample prep, depth draws with
 * cf = signof(c_k), sweep at omega = asin(step_size), measure. */
int orc_qdrift(uint32_t nqb, size_t nterms, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct orc_prep *prep,
	size_t samples, size_t depth, double step_size, uint64_t seed,
	double *out, double *state, uint64_t *drawn)
{
	struct orc_term *t = load_terms(nterms, xm, zm, cf);
	struct orc_term *smp = malloc(sizeof *smp * (depth ? depth : 1));
	double *y = malloc(sizeof *y * (nterms ? nterms : 1));
	struct orc_run r;
	if (!t || !smp || !y || nterms == 0 || run_open(&r, nqb, prep) < 0)
		return free(t), free(smp), free(y), -1;
	cdf_build(y, nterms, &t[0].cf, sizeof t[0], NULL);
	struct orc_rng g;
	rng_seed(&g, seed);
	const double omega = asin(step_size);
	for (size_t s = 0; s < samples; s++) {
		run_prep(&r);
		for (size_t d = 0; d < depth; d++) {
			const size_t k = cdf_inverse(y, nterms, rng_dbl01(&g));
			smp[d].cf = orc_signof(t[k].cf);
			smp[d].x = t[k].x;
			smp[d].z = t[k].z;
			if (drawn)
				drawn[s * depth + d] = k;
		}
		run_sweep(&r, smp, depth, omega, 0);
		run_measure(&r, out + 2 * s);
	}
	if (state)
		memcpy(state, r.a, sizeof(double) * 2 * r.namp);
	run_close(&r);
	free(t);
	free(smp);
	free(y);
	return 0;
}

/* Composite integrator (circ/cmpsit.c:37-235).  Terms ordered by |c|
 * descending; the first `length` form the deterministic part (then lex
 * sorted, angle cf*angle_det), the rest a CDF sampled `depth` times per
 * half-step (angle signof(c)*angle_rand).  Per sample: prep; per step a
 * forward half-sweep of one draw and a reverse half-sweep of a fresh draw,
 * both at omega = 0.5; then measure. */
int orc_cmpsit(uint32_t nqb, size_t nterms, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct orc_prep *prep,
	size_t samples, size_t steps, size_t length, size_t depth,
	double angle_det, double angle_rand, uint64_t seed, double *out,
	double *state)
{
	if (length > nterms || nterms - length == 0)
		return -1;
	struct orc_term *t = load_terms(nterms, xm, zm, cf);
	const size_t nr = nterms - length;
	struct orc_term *smp = malloc(sizeof *smp * (length + depth + 1));
	double *y = malloc(sizeof *y * nr);
	struct orc_run r;
	if (!t || !smp || !y || run_open(&r, nqb, prep) < 0)
		return free(t), free(smp), free(y), -1;
	sort_terms(t, nterms, cmp_abs_desc);
	struct orc_term *det = t, *ran = t + length;
	if (cdf_build(y, nr, &ran[0].cf, sizeof ran[0], NULL) < 0)
		return free(t), free(smp), free(y), run_close(&r), -1;
	sort_terms(det, length, cmp_lex);
	struct orc_rng g;
	rng_seed(&g, seed);
	for (size_t smp_i = 0; smp_i < samples; smp_i++) {
		run_prep(&r);
		for (size_t s = 0; s < steps; s++) {
			for (int rev = 0; rev < 2; rev++) {
				for (size_t k = 0; k < length; k++) {
					smp[k].x = det[k].x;
					smp[k].z = det[k].z;
					smp[k].cf = det[k].cf * angle_det;
				}
				for (size_t d = 0; d < depth; d++) {
					const size_t k = cdf_inverse(y, nr,
						rng_dbl01(&g));
					smp[length + d].x = ran[k].x;
					smp[length + d].z = ran[k].z;
					smp[length + d].cf =
						orc_signof(ran[k].cf) * angle_rand;
				}
				run_sweep(&r, smp, length + depth, 0.5, rev);
			}
		}
		run_measure(&r, out + 2 * smp_i);
	}
	if (state)
		memcpy(state, r.a, sizeof(double) * 2 * r.namp);
	run_close(&r);
	free(t);
	free(smp);
	free(y);
	return 0;
}

/* phase2_run (phase2/phase2_run.c:115-174): basis state |idx>, terms in
 * the given order, phi_k = delta*coeffs[k]; returns amp[idx]. */
int orc_phase2_run(uint32_t nqb, uint64_t idx, size_t n, const uint64_t *xm,
	const uint64_t *zm, const double *coeffs, double delta, double *out)
{
	if (nqb == 0 || nqb > 30)
		return -1;
	const uint64_t namp = UINT64_C(1) << nqb;
	double *a = calloc(2 * namp, sizeof(double));
	if (!a)
		return -1;
	a[2 * idx] = 1.0;
	for (size_t k = 0; k < n; k++) {
		const double phi = delta * coeffs[k];
		double c, s;
		orc_cs(phi, &c, &s);
		orc_rotate_one(a, namp, xm[k], zm[k], c, s);
	}
	out[0] = a[2 * idx];
	out[1] = a[2 * idx + 1];
	free(a);
	return 0;
}
