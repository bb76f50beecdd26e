/*
 * Discrete CDF over |w| and its inverse (behaviour of lib/prob.c:17-78 of
 * the reference: same normalisation order and the same binary-then-linear
 * search, so sampled indices agree bit for bit).
 */
#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>

#include "prob.h"

int prob_cdf_init(struct prob_cdf *cdf, size_t len)
{
	if (len == 0)
		return -1;
	cdf->y = malloc(sizeof *cdf->y * len);
	if (!cdf->y)
		return -1;
	cdf->len = len;
	return 0;
}

void prob_cdf_free(struct prob_cdf *cdf)
{
	free(cdf->y);
	cdf->y = NULL;
	cdf->len = 0;
}

int prob_cdf_from_array_strided(struct prob_cdf *cdf, const double *base,
	size_t stride, double *out_lambda)
{
	const char *p = (const char *)base;
	double lambda = 0.0;
	for (size_t i = 0; i < cdf->len; i++) {
		cdf->y[i] = fabs(*(const double *)(p + i * stride));
		lambda += cdf->y[i];
	}
	if (lambda < DBL_EPSILON)
		return -1;
	double run = 0.0;
	for (size_t i = 0; i < cdf->len; i++) {
		run += cdf->y[i] / lambda;
		cdf->y[i] = run;
	}
	if (out_lambda)
		*out_lambda = lambda;
	return 0;
}

size_t prob_cdf_inverse(const struct prob_cdf *cdf, double y0)
{
	const double *y = cdf->y;
	size_t i = 0;
	for (size_t d = cdf->len / 2; d > 0; d /= 2)
		if (y[i + d] <= y0)
			i += d;
	while (i + 1 < cdf->len && y[i] <= y0)
		i++;
	return i;
}
