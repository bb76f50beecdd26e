#ifndef PHASE2_AMD_INTERNAL_H
#define PHASE2_AMD_INTERNAL_H

/* Helpers shared by the library sources (not installed). */

#include <float.h>
#include <math.h>
#include <stddef.h>

#include "phase2/circ.h"

typedef int (*term_cmp)(const struct circ_hamil_term *,
	const struct circ_hamil_term *);

/* One private pointer a layer above the register keeps with it (qreg_hip.c):
 * set_aux replaces (and releases through its aux_free) the previous one;
 * qreg_free releases the last.  Returns -1 if the register has no backend. */
void *qreg_hip_aux(struct qreg *reg);
int qreg_hip_set_aux(struct qreg *reg, void *aux, void (*aux_free)(void *));

/* Stable sort of a term array (circ.c). */
int circ_terms_sort(struct circ_hamil_term *t, size_t n, term_cmp cmp);

/* The composite sampler's split of H (cmpsit.c): the `length` largest |c|
 * terms in lex order (hm_det), the rest by |c| descending (hm_ran) with its
 * CDF, and room for one draw (hm_smpl).  Shared with the batched driver. */
struct cmpsit_ranct;
struct cmpsit_data;
int cmpsit_ranct_init(struct cmpsit_ranct *r, const struct circ_hamil *hm,
	const struct cmpsit_data *dt);
void cmpsit_ranct_free(struct cmpsit_ranct *r);

/* Sign with a DBL_EPSILON dead zone: -1, +1, or 0 for |a| < DBL_EPSILON
 * (behaviour of circ/internal.h:26-32 of the reference). */
static inline double signof(double a)
{
	if (fabs(a) < DBL_EPSILON)
		return 0.0;
	return a < 0.0 ? -1.0 : 1.0;
}

#endif /* PHASE2_AMD_INTERNAL_H */
