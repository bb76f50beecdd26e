#ifndef PHASE2_HIP_H
#define PHASE2_HIP_H

/*
 * Thin C ABI between the C host library and the hand-written gfx950 HIP
 * kernels (phase2_amd/csrc/kernels.hip).  Plain pointers and sizes only;
 * device pointers are opaque `void *`.  Every function returns 0 on success
 * and -1 on failure, with the reason in ph2hip_strerror().  All work is
 * ordered on one HIP stream per process (default priority); functions that
 * return host data synchronise that stream.
 *
 * Amplitudes are complex128, interleaved (re, im), 16 B each; a register of
 * nqb qubits is 2^nqb amplitudes, indexed by the basis state with qubit k
 * at bit k.
 *
 * Rotation terms arrive with c = cos(phi) and s = sin(phi) computed by the
 * host libm, as the reference CPU backend does (phase2/qreg_qreg.c:189), so
 * every kernel performs the same IEEE operations as
 * phase2/qreg_qreg.c:149-160: per component round(round(c*x) +- round(s*y)),
 * no fused multiply-add.  All kernels are FP64.
 */

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* -- device and memory --------------------------------------------------- */

/* Select the device and create the stream (idempotent).  Replaces the
 * device selection of phase2/world_cuda.c:21-43. */
int ph2hip_init(int device);
void ph2hip_shutdown(void);
const char *ph2hip_strerror(void);
int ph2hip_synchronize(void);

/* Replace cudaMalloc/cudaFree of phase2/qreg_cuda.c:34-37 (one buffer, no
 * exchange buffer). */
int ph2hip_malloc(void **dptr, size_t bytes);
int ph2hip_free(void *dptr);
int ph2hip_h2d(void *dst, const void *src, size_t bytes);
int ph2hip_d2h(void *dst, const void *src, size_t bytes);
int ph2hip_d2d(void *dst, const void *src, size_t bytes);

/* Zero namp amplitudes with a fill kernel (phase2/qreg_cuda.c:96). */
int ph2hip_zero(void *amp, uint64_t namp);

/* -- rotations ------------------------------------------------------------ */

/* One Pauli rotation exp(i phi P): flip mask x, phase mask z, c = cos(phi),
 * s = sin(phi). */
struct ph2hip_rot {
	uint64_t x, z;
	double c, s;
};

/* Plain sweep: one HBM pass (one kernel launch) per term, in order.
 * Replaces kernelPauliRot/kernelMix/kernelAdd of
 * phase2/qreg_cuda_lo.cu:34-131 on one GPU. */
int ph2hip_rotate_plain(void *amp, uint32_t nqb, const struct ph2hip_rot *rot,
	size_t n);

/* Plain-kernel variant: 2, 4 or 8 pairs per thread, plus 16 for
 * non-temporal stores (default 2).  All variants give the same state. */
int ph2hip_set_plain_form(int form);

/* Fused sweep: the terms are applied in order, pass by pass; pass p covers
 * the next pass_len[p] terms and runs as one HBM pass over LDS tiles of
 * 2^tile_bits[p] amplitudes, each tile a coset of a subspace that contains
 * the low low_bits[p] index bits and every flip mask of the pass.  The state
 * is bit-identical to ph2hip_rotate_plain.  Fails (-1, nothing applied) if a
 * pass does not fit its tile (span too large), tile_bits is outside
 * [PH2HIP_TILE_MIN, PH2HIP_TILE_MAX] or above nqb, or a pass holds more than
 * PH2HIP_PASS_MAX terms.  No counterpart in the reference. */
#define PH2HIP_TILE_MIN 9
#define PH2HIP_TILE_MAX 13
#define PH2HIP_PASS_MAX 48
int ph2hip_rotate_fused(void *amp, uint32_t nqb, const struct ph2hip_rot *rot,
	size_t n, const uint32_t *pass_len, const uint32_t *tile_bits,
	size_t npass, const uint32_t *low_bits);

/* Kernel form of the fused pass per tile size: bit TB of pipe_mask selects
 * the persistent, software-pipelined form (persistent_wgs workgroups; 0
 * keeps the current count, initially two per compute unit) instead of one
 * workgroup per tile.  Initial mask from PHASE2_HIP_PIPE, else 0. */
int ph2hip_set_fused_form(uint32_t pipe_mask, uint64_t persistent_wgs);

/* Whether the first and last micro-group of a fused pass may move their
 * cosets straight between HBM and registers (when their lanes cover whole
 * 256-byte runs), skipping the LDS staging of the tile; on by default.
 * Every setting gives the same state. */
int ph2hip_set_fused_direct(int on);

/* Bit tb of tile_mask: fused passes of tile width tb take the half-tile
 * form where they have one (the exchanges between micro-groups run in two
 * rounds through half a tile of LDS, so twice the workgroups fit a CU;
 * first and last groups move their cosets between HBM and registers).
 * Tiles 12 and 13 have that form (PH2HIP_HALF_TILES; 256- and 512-thread
 * workgroups, 32 and 64 KiB of LDS); other bits are an error.  Batched
 * launches use it too.
 * Default: the PHASE2_HIP_HALF environment variable, else
 * PH2HIP_HALF_TILES.  Every
 * setting gives the same state.  No counterpart in the reference (its
 * CUDA backend has no fused pass). */
#define PH2HIP_HALF_TILES ((1u << 12) | (1u << 13))
int ph2hip_set_fused_half(uint32_t tile_mask);
/* The current mask. */
uint32_t ph2hip_fused_half_mask(void);

/* Persistent form of the half-tile pass at tile width 12 or 13: each
 * workgroup walks tiles and loads the next tile's first-group cosets (amps
 * = 8 or 16 of the 16 amplitudes of each coset) into registers while it
 * rotates the current one; wgs_per_cu workgroups per compute unit.  amps 0
 * returns to one workgroup per tile.  Default from PHASE2_HIP_PREFETCH12 /
 * PHASE2_HIP_PREFETCH13 ("amps:wgs_per_cu"), else off.  Every setting
 * gives the same state. */
int ph2hip_set_fused_prefetch(uint32_t tile_bits, uint32_t amps,
	uint32_t wgs_per_cu);

/* Replay of repeated fused sweeps on registers of at most max_qb qubits
 * (default 20, or PHASE2_HIP_REPLAY_QB; 0 turns it off): a
 * ph2hip_rotate_fused call whose arguments (register, terms with their
 * c and s, schedule) all equal an earlier call's -- a Trotter step's
 * circ_cache block from the second step on -- has its launches captured
 * into a hipGraph on its second occurrence and replayed from the third,
 * so the passes of a block go out as one submission.  The state is the
 * same whatever the setting.  Not while profiling.  Setting it drops the
 * cached graphs.  ph2hip_replay_hits counts replays.  No counterpart in the
 * reference. */
int ph2hip_set_replay(uint32_t max_qb);
uint64_t ph2hip_replay_hits(void);
/* The current limit (replay on for registers of at most this many qubits). */
uint32_t ph2hip_replay_max_qb(void);


/* Shape of the fused pass over rot[0..n) without launching it (shape has 4
 * entries): shape[0] = contiguous low bits L, shape[1] = gathered
 * dimensions g, shape[2] = micro-groups, shape[3] = which of the first
 * (bit 0) and last (bit 1) micro-groups move their cosets straight between
 * HBM and registers, and bit 2 for the half-tile form.  -1 if the terms do
 * not fit one tile. */
int ph2hip_pass_shape(const struct ph2hip_rot *rot, size_t n,
	uint32_t tile_bits, uint32_t low_bits, uint32_t *shape);

/* Micro-groups of that pass (for inspection and host-side tests): per group
 * PH2HIP_GROUP_WORDS words, the basis vectors u[0..3] (tile coordinates),
 * the tile bit of each coset-index bit jpos[0..11], the exchange split bit
 * xh and coset-index bit xk towards the next group (half-tile form), the
 * first term and the term count.  Returns the number of groups, -1 if the
 * terms do not fit one tile or cap (in groups) is too small. */
#define PH2HIP_GROUP_WORDS 20
long ph2hip_pass_groups(const struct ph2hip_rot *rot, size_t n,
	uint32_t tile_bits, uint32_t low_bits, uint32_t *out, size_t cap);

/* Dimension of span{x_k & ~(2^low_bits - 1)} over GF(2) for k in [0, n):
 * the number of tile bits beyond low_bits a pass over these terms needs. */
uint32_t ph2hip_span_rank(const struct ph2hip_rot *rot, size_t n,
	uint32_t low_bits);

/* -- batched registers ------------------------------------------------------ */

/* One register inside a shared device buffer: 2^nqb amplitudes starting at
 * amplitude amp_off, its rotations in order, and their fused schedule
 * (npass passes of pass_len[p] terms, all in tiles of the call's size). */
struct ph2hip_batch_job {
	uint64_t amp_off;
	uint32_t nqb;
	const struct ph2hip_rot *rot;
	size_t n;
	const uint32_t *pass_len;
	size_t npass;
};

/* Apply every job's rotations, bit-identical to ph2hip_rotate_plain on each
 * register alone.  Launch p applies pass p of every job that has one, so a
 * batch costs max(npass) launches instead of sum(n); with use_graph the
 * launches go out as one submission of the round graph below (not while
 * profiling).  *launches (if non-NULL) gets the launch count.
 * Returns when the work is complete.  No counterpart in the reference
 * (which runs samples one after another, circ/qdrift.c:115-131). */
int ph2hip_rotate_batched(void *amp, uint64_t namp_total,
	const struct ph2hip_batch_job *jobs, size_t njobs, uint32_t tile_bits,
	uint32_t low_bits, int use_graph, uint64_t *launches);

/* The same in two halves, so that a caller can prepare the next batch on
 * the host while the device runs this one.  plan_build validates the jobs
 * and builds every pass descriptor on the host (no device work; the jobs'
 * arrays may change once it returns); plan_launch uploads the plan and
 * queues its launches (the round graph with use_graph) behind the work already
 * on the library stream, and returns without waiting.  A plan may be freed
 * after ph2hip_synchronize (or any call that waits for the stream). */
struct ph2hip_batch_plan;
int ph2hip_batch_plan_build(const struct ph2hip_batch_job *jobs, size_t njobs,
	uint32_t tile_bits, uint32_t low_bits, uint64_t namp_total,
	struct ph2hip_batch_plan **plan);
int ph2hip_batch_plan_launch(void *amp, struct ph2hip_batch_plan *plan,
	int use_graph, uint64_t *launches);
void ph2hip_batch_plan_free(struct ph2hip_batch_plan *plan);

/* Launches of batched rounds (use_graph, not while profiling) go through
 * one graph of kernel nodes -- two per pass index, the half-tile form then
 * the whole-tile form -- built once and replayed every round with each
 * node's work items set anew (hipGraphExecKernelNodeSetParams; a node with
 * no work exits at once); it is rebuilt only when a round has more pass
 * indices than it holds or another tile width.  Counts of graph builds and
 * replays since the process started. */
void ph2hip_round_graph_stats(uint64_t *builds, uint64_t *replays);

/* amp[idx[k]] = (val[2k], val[2k+1]) with idx < namp_total (host arrays);
 * for repeated indices the last entry wins, as sequential qreg_setamp
 * calls would leave it. */
int ph2hip_scatter_set(void *amp, uint64_t namp_total, const uint64_t *idx,
	const double *val, size_t n);

/* out[2k], out[2k+1] = amp[idx[k]], idx < namp_total. */
int ph2hip_gather_abs(const void *amp, uint64_t namp_total,
	const uint64_t *idx, size_t n, double *out);

/* -- overlap and Slater-Condon -------------------------------------------- */

/* out[2k], out[2k+1] = amp[idx[k]] (host arrays; idx < 2^nqb checked).
 * Replaces the per-determinant cudaMemcpy of phase2/qreg_cuda.c:66-74. */
int ph2hip_gather(const void *amp, uint32_t nqb, const uint64_t *idx,
	size_t n, double *out);

/* ph2hip_gather without any device allocation: the indices and gathered
 * amplitudes pass through a fixed device array of the library (24 B per
 * entry for PH2HIP_GATHER_NOALLOC_MAX entries, 1.5 MiB on each device the
 * library runs on, held until the process ends) instead of buffers
 * allocated per call; no hipMalloc or hipFree.  Same results as
 * ph2hip_gather.  n above PH2HIP_GATHER_NOALLOC_MAX fails (-1, nothing
 * read).  Every call on one device shares that array: calls that overlap
 * on a device (several host threads) race.  No counterpart in the
 * reference; for callers that cannot allocate device memory (per-step
 * overlap loops and harnesses that forbid device allocation). */
#define PH2HIP_GATHER_NOALLOC_MAX 65536
int ph2hip_gather_noalloc(const void *amp, uint32_t nqb, const uint64_t *idx,
	size_t n, double *out);

/* Device-resident Slater-Condon tables: occupation masks and determinants
 * of the alpha and beta k-subsets (host arrays copied to the device). */
struct ph2hip_sc;
int ph2hip_sc_create(struct ph2hip_sc **sc, const uint64_t *mask_a,
	size_t Ma, const uint64_t *mask_b, size_t Mb);
int ph2hip_sc_set_dets(struct ph2hip_sc *sc, const double *det_a,
	const double *det_b);
void ph2hip_sc_destroy(struct ph2hip_sc *sc);

/* amp[idx(a,b)] = (c, 0) or amp[idx(a,b)].re += c for every pair with
 * c = weight*det_a[a]*det_b[b], |c| >= 1e-12 (phase2/state_prep_coeff.c:
 * 181-222).  Fails if any index would exceed 2^nqb. */
int ph2hip_sc_scatter(void *amp, uint32_t nqb, const struct ph2hip_sc *sc,
	uint32_t n_sites, double weight, int tapered, int accumulate);

/* out = sum_{a,b} c(a,b) * amp[idx(a,b)] as (re, im), FP64, reduced in a
 * fixed tree order (phase2/state_prep_coeff.c:224-274). */
int ph2hip_sc_inner(const void *amp, uint32_t nqb, const struct ph2hip_sc *sc,
	uint32_t n_sites, double weight, int tapered, double *out);

/* -- device-side sampler ---------------------------------------------------- */

/* The draws of the qDRIFT and composite samplers, made on the device:
 * stream k continues its xoshiro256** state state[4k .. 4k+3] (advanced in
 * place) for ndraw[k] draws u = (next >> 11) * 2^-53 and inverts its
 * cumulative table cdf[k][0 .. len[k]) (host arrays, the table
 * prob_cdf_from_array_strided builds) for each; out receives the term
 * indices, stream after stream.  The same indices and final states as
 * xoshiro256ss_dbl01 + prob_cdf_inverse on the host (lib/xoshiro256ss.c:
 * 72-85, include/xoshiro256ss.h:58-59, lib/prob.c:67-78), whose calls
 * circ/qdrift.c:95-103 and circ/cmpsit.c:165-172 make per draw.  One
 * 64-lane workgroup per stream.  No counterpart in the reference. */
int ph2hip_sample_cdf(size_t nstream, const double *const *cdf,
	const size_t *len, uint64_t *state, const size_t *ndraw, uint32_t *out);

/* -- measurement ----------------------------------------------------------- */

/* Per-launch timing of the rotation kernels: while enabled, every launch of
 * ph2hip_rotate_plain / ph2hip_rotate_fused is bracketed by HIP events on
 * the library stream; ph2hip_prof_read synchronises and returns launch
 * counts and summed kernel milliseconds per kind since the last reset.
 * Kinds: 0 plain pair pass, 1 plain diagonal pass, 2 fused pass, 3 batched
 * fused pass. */
#define PH2HIP_PROF_KINDS 4
struct ph2hip_prof {
	uint64_t launches[PH2HIP_PROF_KINDS];
	double ms[PH2HIP_PROF_KINDS];
	uint64_t terms[PH2HIP_PROF_KINDS];
	uint64_t amps[PH2HIP_PROF_KINDS];	/* each read and written once */
};
int ph2hip_prof_enable(int on);
int ph2hip_prof_read(struct ph2hip_prof *out, int reset);

/* Per-launch records kept while profiling is enabled (the newest
 * PH2HIP_REC_MAX).  For fused passes: tile bits, contiguous low bits L,
 * gathered dimensions g, terms, micro-groups, the form (direct: bit 0 and
 * bit 1 when the first and last micro-groups move their cosets straight
 * between HBM and registers, bit 2 for the half-tile form), the index bits
 * the tile gathers (pivot mask), the union of the pass's flip and phase
 * masks.  For every launch: the amplitudes it reads and writes once, and
 * the kernel time.  For batched launches ngroups holds the number of work
 * items and direct bit 2 the half-tile form.  ph2hip_prof_records copies up
 * to cap records (oldest first, after draining pending events) and returns
 * how many, or -1. */
#define PH2HIP_REC_MAX 65536
struct ph2hip_launch_rec {
	uint32_t kind, nqb, tile_bits, low_bits;
	uint32_t gdim, nterms, ngroups, direct;
	uint64_t pivots, xunion, zunion, amps;
	double ms;
};
long ph2hip_prof_records(struct ph2hip_launch_rec *out, size_t cap,
	int reset);

/* Events on the library stream for caller-side timing. */
int ph2hip_mark(int slot);
int ph2hip_elapsed_ms(int slot_start, int slot_stop, double *ms);

/* sum_i |amp[i]|^2 (norm-drift reporting), fixed reduction order. */
int ph2hip_norm2(const void *amp, uint32_t nqb, double *out);

/* Checks at full register width without host copies (no counterpart in the
 * reference).  fill_hash writes a reproducible dense state: with
 * g = 0x9e3779b97f4a7c15 and m the splitmix64 output mix,
 *   amp[i] = (u(m(seed + (2i+1) g)), u(m(seed + (2i+2) g))),
 *   u(h) = (h >> 11) * 2^-52 - 1  (exact, in [-1, 1)).
 * count_unequal counts the components (re and im separately) in which two
 * registers differ as IEEE doubles (-0 == +0); it waits for the result. */
int ph2hip_fill_hash(void *amp, uint64_t namp, uint64_t seed);
int ph2hip_count_unequal(const void *a, const void *b, uint64_t namp,
	uint64_t *count);

/* -- cost-model network ----------------------------------------------------- */

/* FP32 perceptron nin -> nhid -> nhid -> 1 with ReLU (mlp.hip): forward,
 * backward and Adam run as HIP kernels on their own stream.  Parameters are
 * packed W1[nhid][nin], b1, W2[nhid][nhid], b2, W3[nhid], b3; the loss is
 * 0.5 * mean (y - t)^2.  Batches are host arrays of at most max_batch rows
 * (forward splits larger inputs).  No counterpart in the reference. */
struct ph2hip_mlp;
int ph2hip_mlp_create(struct ph2hip_mlp **m, uint32_t nin, uint32_t nhid,
	uint32_t max_batch, uint64_t seed);
void ph2hip_mlp_destroy(struct ph2hip_mlp *m);
size_t ph2hip_mlp_nparams(const struct ph2hip_mlp *m);
int ph2hip_mlp_set_params(struct ph2hip_mlp *m, const float *params);
int ph2hip_mlp_get_params(struct ph2hip_mlp *m, float *params);
int ph2hip_mlp_forward(struct ph2hip_mlp *m, const float *x, size_t batch,
	float *y);
int ph2hip_mlp_grad(struct ph2hip_mlp *m, const float *x, const float *t,
	size_t batch, float *loss, float *grad);
int ph2hip_mlp_train_step(struct ph2hip_mlp *m, const float *x,
	const float *t, size_t batch, float lr, float *loss);
const char *ph2hip_mlp_strerror(void);

#ifdef __cplusplus
}
#endif

#endif /* PHASE2_HIP_H */
