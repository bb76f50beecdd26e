#ifndef PHASE2_PAULIS_H
#define PHASE2_PAULIS_H

/*
 * Pauli strings on up to 64 qubits in symplectic form
 * (replaces include/phase2/paulis.h:37-171 of the reference; same names,
 * same encoding):
 *
 *   pak[0]  flip mask  -- qubit k carries X or Y when bit k is set
 *   pak[1]  phase mask -- qubit k carries Z or Y when bit k is set
 *
 * so (flip, phase) = (0,0) I, (1,0) X, (1,1) Y, (0,1) Z.
 */

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pauli_op {
	PAULI_I = 0,
	PAULI_X = 1,
	PAULI_Y = 2,
	PAULI_Z = 3,
};

struct paulis {
	uint64_t pak[2];
};

struct paulis paulis_new(void);
enum pauli_op paulis_get(struct paulis code, uint32_t n);
void paulis_set(struct paulis *code, enum pauli_op op, uint32_t n);
int paulis_eq(struct paulis code1, struct paulis code2);
void paulis_shl(struct paulis *code, uint32_t n);
void paulis_shr(struct paulis *code, uint32_t n);

/* P|i> = z|j>: returns j = i ^ flip and, if z is not NULL, multiplies *z by
 * i^r4 with r4 from paulis_effect_raw. */
uint64_t paulis_effect(struct paulis code, uint64_t i, _Complex double *z);

/* Integer core: j = i ^ flip, *r4_out = (popcount(flip & phase) +
 * 2*popcount(i & phase)) mod 4.  The GPU kernels use the same formula with
 * popcount(flip & phase) hoisted per term and only the parity of
 * popcount(i & phase) per amplitude. */
static inline uint64_t paulis_effect_raw(struct paulis code, uint64_t i,
	int *r4_out)
{
	const int y = __builtin_popcountll(code.pak[0] & code.pak[1]);
	const int m = __builtin_popcountll(i & code.pak[1]);
	*r4_out = (y + 2 * m) & 3;
	return i ^ code.pak[0];
}

/* lo keeps qubits [0, qb_lo), hi keeps [qb_lo, qb_lo + qb_hi). */
void paulis_split(struct paulis code, uint32_t qb_lo, uint32_t qb_hi,
	struct paulis *lo, struct paulis *hi);

/* Lexicographic order, qubit 63 first, I < X < Y < Z; -1, 0 or +1. */
int paulis_cmp(struct paulis a, struct paulis b);

#ifdef __cplusplus
}
#endif

#endif /* PHASE2_PAULIS_H */
