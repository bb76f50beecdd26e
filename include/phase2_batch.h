#ifndef PHASE2_BATCH_H
#define PHASE2_BATCH_H

/*
 * Batched qDRIFT and composite samples: many independent registers (uneven
 * widths), each with its own Hamiltonian, initial multidet state, seed and
 * sampler parameters, run together on one GPU.  No counterpart in the
 * reference, which runs one qdrift or cmpsit instance at a time
 * (circ/qdrift.c:105-133, circ/cmpsit.c:208-235); each register's overlaps
 * equal what qdrift_simul or cmpsit_simul produce for it alone with the same
 * seed (bit for bit).
 *
 * Every draw of every register (its own xoshiro256** stream over its own
 * |c| CDF) is made on the device before the first round, by
 * ph2hip_sample_cdf.  A register's work is a sequence of rounds: a qDRIFT
 * sample is one round of `depth` draws; a composite sample is 2*steps
 * rounds, one per half-step (the `length` largest terms in lex order times
 * angle_det, then `depth` draws times angle_rand, at omega 0.5; forward,
 * then reversed).  In round r every register that has one is prepared if a
 * sample starts there and measured if one ends there; the rotations of all
 * registers run as
 *   PH2_BATCH_PLAIN_EACH    one plain launch per rotation per register,
 *   PH2_BATCH_FUSED_EACH    fused passes, launched register by register,
 *   PH2_BATCH_FUSED_BATCHED fused passes of all registers in one launch per
 *                           pass index, the round captured in a hipGraph;
 *                           the host plans round r+1 from the draws while
 *                           the device runs round r.
 * tile_bits == 0 asks the autotuner to pick tile and low bits by timing
 * candidates on the first round of a subset of the registers.
 */

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct ph2_batch_reg {
	uint32_t nqb;
	size_t nterms;
	const uint64_t *xm, *zm;
	const double *cf;
	size_t ndet;
	const uint64_t *didx;
	const double *dcf;	/* interleaved complex */
	size_t samples, depth;
	double step_size;
	uint64_t seed;		/* 0: the qdrift default seed */
};

enum ph2_batch_mode {
	PH2_BATCH_PLAIN_EACH = 0,
	PH2_BATCH_FUSED_EACH = 1,
	PH2_BATCH_FUSED_BATCHED = 2,
};

struct ph2_batch_stats {
	double seconds;		/* whole run, tuning excluded */
	double tune_seconds;
	double prep_seconds;	/* host rotation lists and pass building
				 * (overlaps the device in the batched mode) */
	double draw_seconds;	/* the device draws, before the first round */
	double amp_updates;	/* sum over registers of samples*depth*2^nqb */
	uint64_t launches;	/* rotation kernel launches */
	uint64_t passes;	/* fused passes (sum over registers) */
	uint32_t tile_bits, low_bits;
};

/* out: sum of samples complex overlaps (interleaved), register-major.
 * Returns 0 or -1. */
int ph2_batch_qdrift(const struct ph2_batch_reg *regs, size_t nreg, int mode,
	uint32_t tile_bits, uint32_t low_bits, double *out,
	struct ph2_batch_stats *stats);

/* A composite-sampler register (cmpsit_data of circ/cmpsit.h:35-43 plus
 * its Hamiltonian and multidet state). */
struct ph2_batch_cmpsit_reg {
	uint32_t nqb;
	size_t nterms;
	const uint64_t *xm, *zm;
	const double *cf;
	size_t ndet;
	const uint64_t *didx;
	const double *dcf;	/* interleaved complex */
	size_t samples, steps, length, depth;
	double angle_det, angle_rand;
	uint64_t seed;		/* 0: the cmpsit default seed */
};

/* Composite samples of every register (out as for ph2_batch_qdrift).
 * Returns 0 or -1 (also when length leaves no terms to sample from). */
int ph2_batch_cmpsit(const struct ph2_batch_cmpsit_reg *regs, size_t nreg,
	int mode, uint32_t tile_bits, uint32_t low_bits, double *out,
	struct ph2_batch_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* PHASE2_BATCH_H */
