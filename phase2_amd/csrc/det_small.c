/*
 * Small dense determinant: Gaussian elimination with partial pivoting
 * (behaviour of lib/det_small.c:26-74 of the reference: same pivot choice,
 * same update order, rows with a zero multiplier skipped), so values agree
 * bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "det_small.h"

double det_small(const double *A, uint32_t n)
{
	if (n > DET_SMALL_MAX_N)
		return 0.0;
	if (n == 0)
		return 1.0;
	double m[DET_SMALL_MAX_N * DET_SMALL_MAX_N];
	memcpy(m, A, sizeof(double) * n * n);
	double det = 1.0;
	for (uint32_t col = 0; col < n; col++) {
		uint32_t piv = col;
		double big = fabs(m[col * n + col]);
		for (uint32_t r = col + 1; r < n; r++) {
			const double v = fabs(m[r * n + col]);
			if (v > big) {
				big = v;
				piv = r;
			}
		}
		if (big == 0.0)
			return 0.0;
		if (piv != col) {
			for (uint32_t c = col; c < n; c++) {
				const double t = m[col * n + c];
				m[col * n + c] = m[piv * n + c];
				m[piv * n + c] = t;
			}
			det = -det;
		}
		const double p = m[col * n + col];
		det *= p;
		const double inv = 1.0 / p;
		for (uint32_t r = col + 1; r < n; r++) {
			const double f = m[r * n + col] * inv;
			if (f == 0.0)
				continue;
			for (uint32_t c = col + 1; c < n; c++)
				m[r * n + c] -= f * m[col * n + c];
		}
	}
	return det;
}
