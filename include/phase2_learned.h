#ifndef PHASE2_LEARNED_H
#define PHASE2_LEARNED_H

/*
 * Learned fusion schedule (no counterpart in the reference; the greedy cut
 * of ph2_sched_greedy is the fallback).  An FP32 perceptron (ph2hip_mlp_*)
 * predicts the time of a candidate fused pass from
 *   f[0..40)   bit k of the OR of the pass's flip masks
 *   f[40..80)  bit k of the OR of its phase masks
 *   f[80]      terms / 48        f[81] micro-groups / 48
 *   f[82]      gathered dims / 13  f[83] contiguous low bits / 13
 *   f[84]      tile bits / 13
 *   f[85]      first micro-group reads HBM directly (0/1)
 *   f[86]      last micro-group writes HBM directly (0/1)
 *   f[87]      half-tile form (0/1)
 * as y = ln(ms) - (nqb - 30) ln 2 (time per 2^30 amplitudes).  For a block of
 * terms the scheduler enumerates passes [i, j) that end where the next term
 * would widen the tile (or at the longest feasible end), for tile widths 12
 * and 13 and for low_bits, low_bits + 1 and low_bits + 2 contiguous low
 * bits, predicts them all in one batch on the GPU and picks the cut, tile
 * and low bits of every pass with the smallest predicted total by dynamic
 * programming.  Schedules are cached
 * per block (flip and phase masks, width), so repeated Trotter sweeps pay
 * for scheduling once.  The order of terms never changes, so the state stays
 * bit-identical to the plain sweep.
 */

#include <stddef.h>
#include <stdint.h>

#include "phase2/qreg_hip.h"
#include "phase2_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PH2_LM_NQ 40
#define PH2_LM_NIN (2 * PH2_LM_NQ + 8)

/* Feature vector (PH2_LM_NIN floats) of the pass rot[0..n) in tiles of
 * tile_bits with the given shape (ph2hip_pass_shape, 4 entries). */
void ph2_learned_features(const struct ph2hip_rot *rot, size_t n,
	uint32_t tile_bits, const uint32_t *shape, float *f);

/* Load parameters (packed as ph2hip_mlp) of a PH2_LM_NIN -> nhid -> nhid
 * -> 1 network, and install the learned schedule for fused registers of at
 * least min_qb qubits (narrower ones keep the greedy cut).  low_bits is the
 * smallest contiguous low part a pass may request. */
int ph2_learned_install(const float *params, uint32_t nhid, uint32_t low_bits,
	uint32_t min_qb);
void ph2_learned_uninstall(void);

/* Learned schedule of one block (needs ph2_learned_install). */
int ph2_learned_schedule(const struct ph2hip_rot *rot, size_t n, uint32_t nqb,
	struct ph2_sched *out);

/* The candidate passes the scheduler chooses from for the block rot[0..n)
 * (low_bits as for ph2_learned_install; no network needed): 4 words each,
 * first term, end (exclusive), tile bits, contiguous low bits, grouped by
 * ascending first term.  Returns the count, or -1 if cap is too small.
 * For evaluations that time every candidate. */
long ph2_learned_candidates(const struct ph2hip_rot *rot, size_t n,
	uint32_t nqb, uint32_t low_bits, uint32_t *out, size_t cap);

/* Predicted milliseconds of each pass of a schedule (ms[npass]). */
int ph2_learned_predict(const struct ph2hip_rot *rot, size_t n, uint32_t nqb,
	const struct ph2_sched *s, double *ms);

/* Cache statistics: hits and misses since install. */
void ph2_learned_stats(uint64_t *hits, uint64_t *misses);

#ifdef __cplusplus
}
#endif

#endif /* PHASE2_LEARNED_H */
