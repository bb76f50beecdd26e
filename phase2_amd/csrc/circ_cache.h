#ifndef PHASE2_CIRC_CACHE_H
#define PHASE2_CIRC_CACHE_H

/* Batch cache of consecutive rotations sharing a hi-qubit code, at most
 * CIRC_CACHE_MAX terms (interface of phase2/circ_cache.h of the reference).
 * On one GPU qb_hi = 0, so every block is simply the next <= 1024 terms;
 * the block is what the fused sweep cuts into passes. */

#include <stddef.h>

#include "phase2/paulis.h"

#define CIRC_CACHE_MAX 1024

struct circ_cache;

struct circ_cache *circ_cache_init(int hi, int lo);
void circ_cache_free(struct circ_cache *c);
/* 0 if accepted; -1 if full or the hi code differs (flush, then retry). */
int circ_cache_insert(struct circ_cache *c, struct paulis pa, double phi);
void circ_cache_flush(struct circ_cache *c,
	void (*fn)(struct paulis hi, const struct paulis *lo, double *phis,
		size_t n, void *data),
	void *data);
size_t circ_cache_len(const struct circ_cache *c);

#endif /* PHASE2_CIRC_CACHE_H */
