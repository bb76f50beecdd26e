/*
 * Pauli-string algebra (behaviour of phase2/paulis.c:9-205 of the
 * reference; encoding and names as in include/phase2/paulis.h).
 */
#include <complex.h>
#include <stdint.h>

#include "phase2/paulis.h"

struct paulis paulis_new(void)
{
	struct paulis p = { .pak = { 0, 0 } };
	return p;
}

enum pauli_op paulis_get(struct paulis code, uint32_t n)
{
	const unsigned f = (unsigned)(code.pak[0] >> n & 1);
	const unsigned z = (unsigned)(code.pak[1] >> n & 1);
	if (!z)
		return f ? PAULI_X : PAULI_I;
	return f ? PAULI_Y : PAULI_Z;
}

void paulis_set(struct paulis *code, enum pauli_op op, uint32_t n)
{
	const uint64_t bit = UINT64_C(1) << n;
	uint64_t f, z;
	switch (op) {
	case PAULI_I: f = 0; z = 0; break;
	case PAULI_X: f = bit; z = 0; break;
	case PAULI_Y: f = bit; z = bit; break;
	case PAULI_Z: f = 0; z = bit; break;
	default: return;
	}
	code->pak[0] = (code->pak[0] & ~bit) | f;
	code->pak[1] = (code->pak[1] & ~bit) | z;
}

int paulis_eq(struct paulis a, struct paulis b)
{
	return a.pak[0] == b.pak[0] && a.pak[1] == b.pak[1];
}

void paulis_shl(struct paulis *code, uint32_t n)
{
	code->pak[0] <<= n;
	code->pak[1] <<= n;
}

void paulis_shr(struct paulis *code, uint32_t n)
{
	code->pak[0] >>= n;
	code->pak[1] >>= n;
}

/* Multiplication by i^r4 is exact; it is written as the component
 * permutation it amounts to. */
uint64_t paulis_effect(struct paulis code, uint64_t i, _Complex double *z)
{
	int r4;
	const uint64_t j = paulis_effect_raw(code, i, &r4);
	if (z) {
		const double re = creal(*z), im = cimag(*z);
		switch (r4) {
		case 1: *z = CMPLX(-im, re); break;
		case 2: *z = CMPLX(-re, -im); break;
		case 3: *z = CMPLX(im, -re); break;
		default: break;
		}
	}
	return j;
}

void paulis_split(struct paulis code, uint32_t qb_lo, uint32_t qb_hi,
	struct paulis *lo, struct paulis *hi)
{
	const uint64_t mlo = qb_lo >= 64 ? ~UINT64_C(0)
		: (UINT64_C(1) << qb_lo) - 1;
	const uint32_t top = qb_lo + qb_hi;
	const uint64_t mall = top >= 64 ? ~UINT64_C(0)
		: (UINT64_C(1) << top) - 1;
	const uint64_t mhi = mall & ~mlo;
	lo->pak[0] = code.pak[0] & mlo;
	lo->pak[1] = code.pak[1] & mlo;
	hi->pak[0] = code.pak[0] & mhi;
	hi->pak[1] = code.pak[1] & mhi;
}

/* Rank of a qubit's operator in the order I < X < Y < Z. */
static int op_rank(struct paulis p, uint32_t q)
{
	return (int)paulis_get(p, q);
}

int paulis_cmp(struct paulis a, struct paulis b)
{
	if (paulis_eq(a, b))
		return 0;
	/* Only qubits where the strings differ can decide; scan from the top
	 * differing bit down. */
	uint64_t diff = (a.pak[0] ^ b.pak[0]) | (a.pak[1] ^ b.pak[1]);
	const uint32_t q = 63 - (uint32_t)__builtin_clzll(diff);
	const int ra = op_rank(a, q), rb = op_rank(b, q);
	return ra < rb ? -1 : 1;
}
