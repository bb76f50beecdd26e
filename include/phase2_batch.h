#ifndef PHASE2_BATCH_H
#define PHASE2_BATCH_H

/*
 * Batched qDRIFT: many independent registers (uneven widths), each with its
 * own Hamiltonian, initial multidet state, seed and qdrift_data, run
 * together on one GPU.  No counterpart in the reference, which runs one
 * qdrift instance at a time (circ/qdrift.c:105-133); each register's
 * overlaps equal what qdrift_init/qdrift_simul produce for it alone with
 * the same seed (bit for bit).
 *
 * Per sample round every active register is prepared, draws `depth` terms
 * from its own xoshiro256** stream and |c| CDF (input order), and is
 * measured; the rotations of all registers run as
 *   PH2_BATCH_PLAIN_EACH    one plain launch per rotation per register,
 *   PH2_BATCH_FUSED_EACH    fused passes, launched register by register,
 *   PH2_BATCH_FUSED_BATCHED fused passes of all registers in one launch per
 *                           pass index, the round captured in a hipGraph;
 *                           the host draws and plans round s+1 while the
 *                           device runs round s.
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
	double prep_seconds;	/* host draws and pass building (overlaps the
				 * device in the batched mode) */
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

#ifdef __cplusplus
}
#endif

#endif /* PHASE2_BATCH_H */
