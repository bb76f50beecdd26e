/*
 * Fusion schedules: how a block of consecutive rotations is cut into HBM
 * passes (no counterpart in the reference; the block itself is the
 * circ_cache batch of phase2/circ_cache.c:55-76).
 */
#include <stdint.h>
#include <stdlib.h>

#include "phase2/qreg_hip.h"
#include "phase2_hip.h"

void ph2_sched_free(struct ph2_sched *s)
{
	free(s->len);
	free(s->tile_bits);
	free(s->low_bits);
	s->len = s->tile_bits = s->low_bits = NULL;
	s->npass = s->cap = 0;
}

int ph2_sched_reserve(struct ph2_sched *s, size_t cap)
{
	if (cap <= s->cap)
		return 0;
	uint32_t *a = realloc(s->len, sizeof *a * cap);
	if (!a)
		return -1;
	s->len = a;
	uint32_t *b = realloc(s->tile_bits, sizeof *b * cap);
	if (!b)
		return -1;
	s->tile_bits = b;
	uint32_t *c = realloc(s->low_bits, sizeof *c * cap);
	if (!c)
		return -1;
	s->low_bits = c;
	s->cap = cap;
	return 0;
}

int ph2_sched_push(struct ph2_sched *s, uint32_t len, uint32_t tb,
	uint32_t low)
{
	if (s->npass == s->cap &&
	    ph2_sched_reserve(s, s->cap ? 2 * s->cap : 64) < 0)
		return -1;
	s->len[s->npass] = len;
	s->tile_bits[s->npass] = tb;
	s->low_bits[s->npass] = low;
	s->npass++;
	return 0;
}

static inline int topbit(uint64_t v)
{
	return 63 - __builtin_clzll(v);
}

/* 1 if the pass rot[0..n) in tiles of tb runs in the half-tile form or
 * needs no exchange (one micro-group), 0 if not, -1 if it does not fit. */
static int half_or_single(const struct ph2hip_rot *rot, size_t n, uint32_t tb,
	uint32_t low)
{
	uint32_t shape[4];
	if (ph2hip_pass_shape(rot, n, tb, low, shape) < 0)
		return -1;
	return shape[2] <= 1 || (shape[3] & 4u) != 0;
}

/* Append the pass rot[0..len) at tile 13, cut into the longest prefixes
 * that keep the half-tile form: without it a multi-group tile-13 pass
 * runs one 128 KiB workgroup per CU, measured 2.8x the time of a
 * half-tile pass over the same register. */
static int push_split13(struct ph2_sched *out, const struct ph2hip_rot *rot,
	uint32_t len, uint32_t low)
{
	while (len > 0) {
		uint32_t lo = 1, hi = len;   /* longest good prefix in [lo, hi] */
		const int whole = half_or_single(rot, len, 13, low);
		if (whole < 0)
			return -1;
		if (whole) {
			lo = len;
		} else {
			while (lo < hi) {
				const uint32_t mid = (lo + hi + 1) / 2;
				const int ok = half_or_single(rot, mid, 13, low);
				if (ok < 0)
					return -1;
				if (ok)
					lo = mid;
				else
					hi = mid - 1;
			}
		}
		if (ph2_sched_push(out, lo, 13, low) < 0)
			return -1;
		rot += lo;
		len -= lo;
	}
	return 0;
}

static int push(struct ph2_sched *out, const struct ph2hip_rot *rot,
	uint32_t len, uint32_t tb, uint32_t low)
{
	if (tb == 13 && (ph2hip_fused_half_mask() >> 13 & 1))
		return push_split13(out, rot, len, low);
	return ph2_sched_push(out, len, tb, low);
}

int ph2_sched_greedy(const struct ph2hip_rot *rot, size_t n, uint32_t nqb,
	uint32_t tile_bits, uint32_t low_bits, struct ph2_sched *out)
{
	out->npass = 0;
	if (tile_bits > nqb)
		tile_bits = nqb;
	if (tile_bits == 0)
		return -1;
	/* at least one tile bit is left for the flip masks, so every pass
	 * holds at least one term (gmax >= 1) */
	if (low_bits >= tile_bits)
		low_bits = tile_bits - 1;
	const uint64_t lowmask = (UINT64_C(1) << low_bits) - 1;
	const uint32_t gmax = tile_bits - low_bits;
	uint64_t basis[64];
	int piv[64];
	uint32_t g = 0, len = 0;
	for (size_t k = 0; k < n; k++) {
		uint64_t v = rot[k].x & ~lowmask;
		for (uint32_t b = 0; b < g && v; b++)
			if (v >> piv[b] & 1)
				v ^= basis[b];
		const int grows = v != 0;
		if (len == PH2HIP_PASS_MAX || (grows && g == gmax)) {
			if (push(out, rot + k - len, len, tile_bits, low_bits) <
				0)
				return -1;
			g = 0;
			len = 0;
			v = rot[k].x & ~lowmask;
		}
		if (v) {
			const int p = topbit(v);
			for (uint32_t b = 0; b < g; b++)
				if (basis[b] >> p & 1)
					basis[b] ^= v;
			basis[g] = v;
			piv[g] = p;
			g++;
		}
		len++;
	}
	if (len && push(out, rot + n - len, len, tile_bits, low_bits) < 0)
		return -1;
	return 0;
}
