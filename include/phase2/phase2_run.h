#ifndef PHASE2_PHASE2_RUN_H
#define PHASE2_PHASE2_RUN_H

/*
 * <psi| prod_k exp(i delta coeffs[k] P_k) |psi> for a basis state psi
 * (replaces include/phase2/phase2_run.h:38-40 of the reference, the symbol
 * the Python package binds).  pauli_strs are space-separated tokens such as
 * "X0 Z1 Y2"; psi_str is a '0'/'1' string whose length is the qubit count
 * (1..30), character k giving qubit k.  Terms apply in the given order.
 * Returns 0, or -1 on invalid input or allocation failure.
 */

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

int phase2_run(const char **pauli_strs, const double *coeffs, size_t nterms,
	double delta, const char *psi_str, double *out_re, double *out_im);

#ifdef __cplusplus
}
#endif

#endif /* PHASE2_PHASE2_RUN_H */
