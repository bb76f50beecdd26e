#ifndef XOSHIRO256SS_H
#define XOSHIRO256SS_H

/*
 * xoshiro256** 1.0 (D. Blackman, S. Vigna, "Scrambled linear pseudorandom
 * number generators", ACM TOMS 47(4), 2021) with the seeding of the phase2
 * library: four splitmix64 outputs of the seed fill the state, then 128
 * outputs are discarded (replaces include/xoshiro256ss.h:41-59 and
 * lib/xoshiro256ss.c:34-85 of the reference).
 */

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct xoshiro256ss {
	uint64_t s[4];
};

void xoshiro256ss_init(struct xoshiro256ss *rng, uint64_t seed);
uint64_t xoshiro256ss_next(struct xoshiro256ss *rng);
/* Advance by 2^128 and 2^192 outputs (the published jump polynomials). */
void xoshiro256ss_jump(struct xoshiro256ss *rng);
void xoshiro256ss_longjump(struct xoshiro256ss *rng);

/* Uniform double in [0, 1): the top 53 bits scaled by 2^-53. */
#define xoshiro256ss_dbl01(rng) \
	((double)(xoshiro256ss_next(rng) >> 11) * 0x1.0p-53)

#ifdef __cplusplus
}
#endif

#endif /* XOSHIRO256SS_H */
