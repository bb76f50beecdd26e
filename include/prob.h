#ifndef PROB_H
#define PROB_H

/*
 * Discrete CDF over |w| and inverse-CDF sampling (replaces include/prob.h
 * and lib/prob.c:17-78 of the reference; same names and conventions).
 */

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

struct prob_cdf {
	double *y;
	size_t len;
};

/* -1 for len == 0 or allocation failure. */
int prob_cdf_init(struct prob_cdf *cdf, size_t len);
void prob_cdf_free(struct prob_cdf *cdf);

/* y[i] = sum_{k<=i} |w_k| / lambda, lambda = sum |w_k|, reading cdf->len
 * weights at byte stride `stride` from base.  -1 when lambda < DBL_EPSILON
 * (y then holds the raw |w_k|).  *out_lambda gets lambda if non-NULL. */
int prob_cdf_from_array_strided(struct prob_cdf *cdf, const double *base,
	size_t stride, double *out_lambda);

/* Smallest i with y[i] > y0, clamped to len - 1. */
size_t prob_cdf_inverse(const struct prob_cdf *cdf, double y0);

#ifdef __cplusplus
}
#endif

#endif /* PROB_H */
