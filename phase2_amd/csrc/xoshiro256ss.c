/*
 * xoshiro256** 1.0 (Blackman & Vigna, ACM TOMS 47(4), 2021; public-domain
 * algorithm) with the phase2 seeding: splitmix64 of the seed fills the four
 * state words, then 128 outputs are discarded (behaviour of
 * lib/xoshiro256ss.c:34-85 of the reference).
 */
#include <stddef.h>
#include <stdint.h>

#include "xoshiro256ss.h"

static inline uint64_t rotl(uint64_t v, int k)
{
	return (v << k) | (v >> (64 - k));
}

static uint64_t splitmix64_next(uint64_t *x)
{
	uint64_t v = (*x += UINT64_C(0x9e3779b97f4a7c15));
	v = (v ^ (v >> 30)) * UINT64_C(0xbf58476d1ce4e5b9);
	v = (v ^ (v >> 27)) * UINT64_C(0x94d049bb133111eb);
	return v ^ (v >> 31);
}

void xoshiro256ss_init(struct xoshiro256ss *rng, uint64_t seed)
{
	uint64_t x = seed;
	for (int w = 0; w < 4; w++)
		rng->s[w] = splitmix64_next(&x);
	for (int k = 0; k < 128; k++)
		(void)xoshiro256ss_next(rng);
}

uint64_t xoshiro256ss_next(struct xoshiro256ss *rng)
{
	uint64_t *s = rng->s;
	const uint64_t out = rotl(s[1] * 5, 7) * 9;
	const uint64_t t = s[1] << 17;
	s[2] ^= s[0];
	s[3] ^= s[1];
	s[1] ^= s[2];
	s[0] ^= s[3];
	s[2] ^= t;
	s[3] = rotl(s[3], 45);
	return out;
}

/* Jump: XOR of the states at the set bits of the published polynomial. */
static void jump_with(struct xoshiro256ss *rng, const uint64_t poly[4])
{
	uint64_t acc[4] = { 0, 0, 0, 0 };
	for (int w = 0; w < 4; w++)
		for (int b = 0; b < 64; b++) {
			if (poly[w] >> b & 1)
				for (int i = 0; i < 4; i++)
					acc[i] ^= rng->s[i];
			(void)xoshiro256ss_next(rng);
		}
	for (int i = 0; i < 4; i++)
		rng->s[i] = acc[i];
}

void xoshiro256ss_jump(struct xoshiro256ss *rng)
{
	static const uint64_t P[4] = { 0x180ec6d33cfd0aba, 0xd5a61266f0c9392c,
		0xa9582618e03fc9aa, 0x39abdc4529b1661c };
	jump_with(rng, P);
}

void xoshiro256ss_longjump(struct xoshiro256ss *rng)
{
	static const uint64_t P[4] = { 0x76e15d3efefdcbbf, 0xc5004e441c522fb3,
		0x77710069854ee241, 0x39109bb02acbe635 };
	jump_with(rng, P);
}
