#ifndef PHASE2_QREG_HIP_H
#define PHASE2_QREG_HIP_H

/*
 * HIP-backend controls beside the reference qreg API (no counterpart in the
 * reference).  They select how qreg_paulirot applies a block of rotations;
 * every mode leaves the state bit-identical to the plain sweep.
 *
 *   QREG_HIP_PLAIN  one HBM pass per term (ph2hip_rotate_plain)
 *   QREG_HIP_FUSED  passes of consecutive terms applied in LDS tiles
 *                   (ph2hip_rotate_fused), cut greedily or by the installed
 *                   schedule function
 *
 * Defaults come from PHASE2_HIP_MODE (plain|fused), PHASE2_HIP_TILE (tile
 * bits, 9..13) and PHASE2_HIP_LOW (contiguous low bits); otherwise fused,
 * automatic tile width and 3.  Tile width 0 (the automatic one) is 13 from
 * 26 qubits, 12 from 17, nqb - 5 (at least 9) below.  Registers narrower than 9 qubits always use the plain
 * sweep.
 */

#include <stddef.h>
#include <stdint.h>

#include "phase2/qreg.h"
#include "phase2_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum qreg_hip_mode {
	QREG_HIP_PLAIN = 0,
	QREG_HIP_FUSED = 1,
};

/* Process-wide defaults for registers created afterwards. */
void qreg_hip_set_default(int mode, uint32_t tile_bits, uint32_t low_bits);
/* Per-register setting. */
void qreg_hip_configure(struct qreg *reg, int mode, uint32_t tile_bits,
	uint32_t low_bits);

/* Fusion schedule of one block of terms: npass passes covering the terms
 * in order, pass p taking len[p] terms in tiles of tile_bits[p] bits with
 * at least low_bits[p] contiguous low bits. */
struct ph2_sched {
	size_t npass;
	uint32_t *len;        /* terms per pass */
	uint32_t *tile_bits;  /* tile width per pass */
	uint32_t *low_bits;   /* requested contiguous low bits per pass */
	size_t cap;
};

void ph2_sched_free(struct ph2_sched *s);
/* Room for cap passes; append one pass. 0 or -1 (out of memory). */
int ph2_sched_reserve(struct ph2_sched *s, size_t cap);
int ph2_sched_push(struct ph2_sched *s, uint32_t len, uint32_t tile_bits,
	uint32_t low_bits);

/* Greedy cut: extend each pass while its flip masks span at most
 * tile_bits - low_bits dimensions beyond the low bits and it holds at most
 * PH2HIP_PASS_MAX terms. */
int ph2_sched_greedy(const struct ph2hip_rot *rot, size_t n, uint32_t nqb,
	uint32_t tile_bits, uint32_t low_bits, struct ph2_sched *out);

/* A schedule function replaces the greedy cut for every fused register
 * (NULL restores greedy).  It must return 0 and a schedule covering all n
 * terms, or -1 to fall back to greedy for this block. */
typedef int (*ph2_sched_fn)(void *ctx, const struct ph2hip_rot *rot,
	size_t n, uint32_t nqb, struct ph2_sched *out);
void ph2_sched_install(ph2_sched_fn fn, void *ctx);

/* Device pointer of the register's amplitudes (for the batched and
 * benchmark entry points). */
void *qreg_hip_device_ptr(struct qreg *reg);

#ifdef __cplusplus
}
#endif

#endif /* PHASE2_QREG_HIP_H */
