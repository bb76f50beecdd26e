#ifndef PHASE2_COMBINATIONS_H
#define PHASE2_COMBINATIONS_H

/* k-subsets of {0..n-1} in lexicographic order, the order of
 * itertools.combinations (replaces include/combinations.h of the
 * reference; same names).  combinations_next writes the next subset to out
 * and returns 0, or returns 1 once exhausted. */

#include <stdint.h>

#define COMBINATIONS_MAX_K (32u)

struct combo {
	uint32_t n;
	uint32_t k;
	uint32_t state[COMBINATIONS_MAX_K + 2];
	int done;
	int started;
};

void combinations_init(struct combo *c, uint32_t n, uint32_t k);
int combinations_next(struct combo *c, uint32_t *out);

#endif /* PHASE2_COMBINATIONS_H */
