#ifndef PHASE2_DET_SMALL_H
#define PHASE2_DET_SMALL_H

/* Determinant of a row-major n x n matrix by elimination with partial
 * pivoting; 1 for n = 0, 0 for a zero pivot column or n > DET_SMALL_MAX_N
 * (replaces include/det_small.h of the reference). */

#include <stdint.h>

#define DET_SMALL_MAX_N (32u)

double det_small(const double *A, uint32_t n);

#endif /* PHASE2_DET_SMALL_H */
