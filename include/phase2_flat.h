#ifndef PHASE2_FLAT_H
#define PHASE2_FLAT_H

/*
 * Flat-argument entry points over the phase2 C API for FFI callers
 * (ctypes): arrays and sizes in, overlaps and optionally the final state
 * out.  No counterpart in the reference beyond phase2_run
 * (include/phase2/phase2_run.h:38-40); each wraps the reference-shaped API
 * (circ_hamil + state-prep payload + trott/trott2/qdrift/cmpsit *_init,
 * *_simul, *_free) exactly as a C caller such as ph2run would use it
 * (ph2run/ph2run.c:327-360).
 *
 * Terms: flip masks xm, phase masks zm, coefficients cf (n each).
 * States: 2^nqb complex128 as interleaved doubles.  Every function returns 0
 * or -1; ph2_strerror() describes the last device error.
 */

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Initial state: kind 1 = multidet (ndet indices and interleaved complex
 * coefficients), kind 2 = single-block coefficient matrix. */
struct ph2_prep {
	int kind;
	size_t ndet;
	const uint64_t *didx;
	const double *dcf;
	uint32_t n_sites, n_alpha, n_beta;
	int closed_shell, tapered;
	const double *Ca, *Cb;
};

int ph2_init(void);
const char *ph2_strerror(void);

/* Fused-sweep defaults for registers created afterwards (qreg_hip.h). */
void ph2_set_default(int mode, uint32_t tile_bits, uint32_t low_bits);

/* algo 1 = trott, 2 = trott2; out: 2*steps doubles; state may be NULL. */
int ph2_trott(int algo, uint32_t nqb, size_t nterms, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct ph2_prep *prep,
	double delta, size_t steps, double *out, double *state);

int ph2_qdrift(uint32_t nqb, size_t nterms, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct ph2_prep *prep,
	size_t samples, size_t depth, double step_size, uint64_t seed,
	double *out, double *state);

int ph2_cmpsit(uint32_t nqb, size_t nterms, const uint64_t *xm,
	const uint64_t *zm, const double *cf, const struct ph2_prep *prep,
	size_t samples, size_t steps, size_t length, size_t depth,
	double angle_det, double angle_rand, uint64_t seed, double *out,
	double *state);

/* Rotations exp(i phi_k P_k) in order on a host state (copied to the
 * device and back), through qreg_paulirot in blocks of at most 1024 with
 * the given mode and tile settings. */
int ph2_rotate(double *state, uint32_t nqb, size_t n, const uint64_t *xm,
	const uint64_t *zm, const double *phi, int mode, uint32_t tile_bits,
	uint32_t low_bits);

/* Slater-Condon expansion into a zeroed register / inner product with a
 * host state (state_prep_coeff_expand_all / state_prep_coeff_inner). */
int ph2_expand(uint32_t nqb, uint32_t n_sites, uint32_t n_alpha,
	uint32_t n_beta, int closed_shell, int tapered, const double *Ca,
	const double *Cb, double *state);
int ph2_inner(uint32_t nqb, uint32_t n_sites, uint32_t n_alpha,
	uint32_t n_beta, int closed_shell, int tapered, const double *Ca,
	const double *Cb, const double *state, double *out);

/* Lexicographic order of terms (circ_hamil_sort_lex): perm[k] is the input
 * index of the k-th term. */
int ph2_sort_lex(size_t n, const uint64_t *xm, const uint64_t *zm,
	uint64_t *perm);

/* Device-resident register handle (a struct qreg): create, prepare a
 * multidet state, apply rotations in <=1024-term qreg_paulirot blocks
 * without host copies, measure <psi0|psi> against a multidet, squared norm,
 * copy out.  For benchmarks and batched callers. */
void *ph2_reg_create(uint32_t nqb, int mode, uint32_t tile_bits,
	uint32_t low_bits);
void ph2_reg_free(void *reg);
int ph2_reg_configure(void *reg, int mode, uint32_t tile_bits,
	uint32_t low_bits);
int ph2_reg_prep_multidet(void *reg, size_t ndet, const uint64_t *didx,
	const double *dcf);
int ph2_reg_rotate(void *reg, size_t n, const uint64_t *xm,
	const uint64_t *zm, const double *phi);
/* Rotations with an explicit fused schedule (pass lengths, per-pass tile
 * bits and contiguous low bits), bypassing the scheduler; for timing
 * records and tests. */
int ph2_reg_rotate_sched(void *reg, size_t n, const uint64_t *xm,
	const uint64_t *zm, const double *phi, const uint32_t *pass_len,
	const uint32_t *tile_bits, const uint32_t *low_bits, size_t npass);
int ph2_reg_overlap_multidet(void *reg, size_t ndet, const uint64_t *didx,
	const double *dcf, double *out);
int ph2_reg_norm2(void *reg, double *out);
int ph2_reg_read(void *reg, double *state);
/* out[2k], out[2k+1] = amp[idx[k]] (indices below 2^nqb). */
int ph2_reg_gather(void *reg, size_t n, const uint64_t *idx, double *out);
/* The reproducible dense state of ph2hip_fill_hash, and the count of
 * components in which two registers of one width differ. */
int ph2_reg_fill_hash(void *reg, uint64_t seed);
int ph2_reg_count_unequal(void *reg1, void *reg2, uint64_t *count);

#ifdef __cplusplus
}
#endif

#endif /* PHASE2_FLAT_H */
