/*
 * Lexicographic k-subsets of {0..n-1} (behaviour of lib/combinations.c:
 * 23-89 of the reference: the itertools.combinations order; k == 0 yields
 * one empty subset; k > n or k > COMBINATIONS_MAX_K yields none).
 */
#include <stdint.h>
#include <string.h>

#include "combinations.h"

void combinations_init(struct combo *c, uint32_t n, uint32_t k)
{
	memset(c, 0, sizeof *c);
	c->n = n;
	c->k = k;
	if (k > n || k > COMBINATIONS_MAX_K) {
		c->done = 1;
		return;
	}
	for (uint32_t i = 0; i < k; i++)
		c->state[i] = i;
}

int combinations_next(struct combo *c, uint32_t *out)
{
	if (c->done)
		return 1;
	const uint32_t n = c->n, k = c->k;
	if (c->started) {
		/* rightmost position that can still move up */
		int32_t j = (int32_t)k - 1;
		while (j >= 0 && c->state[j] == n - k + (uint32_t)j)
			j--;
		if (j < 0) {
			c->done = 1;
			return 1;
		}
		c->state[j]++;
		for (uint32_t i = (uint32_t)j + 1; i < k; i++)
			c->state[i] = c->state[i - 1] + 1;
	}
	c->started = 1;
	memcpy(out, c->state, sizeof(uint32_t) * k);
	if (k == 0 || k == n)
		c->done = 1;
	return 0;
}
