#ifndef PHASE2_STATE_PREP_COEFF_H
#define PHASE2_STATE_PREP_COEFF_H

/*
 * Slater-Condon expansion of a coefficient-matrix trial state into the
 * register, and the matching inner product (replaces
 * include/phase2/state_prep_coeff.h:41-72 of the reference; same names and
 * argument meaning).
 *
 * A block (C_alpha, C_beta, weight) contributes
 *     c(a, b) = weight * det(C_alpha[occ_a, :]) * det(C_beta[occ_b, :])
 * at index  sum_{i in occ_a} 2^i + sum_{i in occ_b} 2^(n_sites + i)
 * (bits 0 and n_sites dropped when tapered), for every pair of
 * lexicographic k-subsets; |c| < 1e-12 is skipped.  C is row-major
 * (n_sites, n_occ).  C_beta == NULL means C_beta = C_alpha.
 *
 * On the HIP backend the scatter and the inner product run on the device:
 * the scratch holds the subset tables (occupation masks) and determinants,
 * mirrored in device memory, and nothing reads the host shadow.  Indices
 * beyond the register are rejected (-1) instead of aliasing.
 */

#include <stddef.h>
#include <stdint.h>

#include "phase2/qreg.h"

#ifdef __cplusplus
extern "C" {
#endif

struct data_coeff_matrix;

/* The reference's layout.  tup_a points into a block that
 * state_prep_coeff_scratch_init allocates with the library's private
 * tables in front of it (occupation masks, device mirror); only
 * state_prep_coeff_scratch_free releases it. */
struct state_prep_coeff_scratch {
	uint32_t n_sites, n_alpha, n_beta;
	uint32_t *tup_a, *tup_b;	/* Ma x max(n_alpha,1), Mb x max(n_beta,1) */
	double *det_a, *det_b;		/* host determinants of the last block */
	size_t Ma, Mb;
};

int state_prep_coeff_scratch_init(struct state_prep_coeff_scratch *sc,
	uint32_t n_sites, uint32_t n_alpha, uint32_t n_beta);
void state_prep_coeff_scratch_free(struct state_prep_coeff_scratch *sc);

int state_prep_coeff_expand(struct qreg *reg,
	struct state_prep_coeff_scratch *sc, const double *C_alpha,
	const double *C_beta, double weight, int tapered, int accumulate);

int state_prep_coeff_expand_all(struct qreg *reg,
	struct state_prep_coeff_scratch *sc,
	const struct data_coeff_matrix *cm);

int state_prep_coeff_inner(struct qreg *reg,
	struct state_prep_coeff_scratch *sc, const double *C_alpha,
	const double *C_beta, double weight, int tapered, _Complex double *out);

#ifdef __cplusplus
}
#endif

#endif /* PHASE2_STATE_PREP_COEFF_H */
