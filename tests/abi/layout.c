/*
 * Prints the size of every public struct of the phase2 C API and the
 * offset and size of each of its members, one "name value" line each.
 * Compiled once against the reference's include/ and once against this
 * library's include/ (tests/test_abi_layout.py); identical output means a
 * caller compiled against either header lays the structs out the same.
 */
#include <stddef.h>
#include <stdio.h>

#include "phase2.h"
#include "circ/cmpsit.h"
#include "circ/qdrift.h"
#include "circ/trott.h"
#include "circ/trott2.h"
#include "phase2/state_prep_coeff.h"
#include "xoshiro256ss.h"

#define SZ(T) printf("sizeof(" #T ") %zu\n", sizeof(T))
#define M(T, m) printf(#T "." #m " %zu %zu\n", offsetof(T, m), \
	sizeof(((T *)0)->m))

int main(void)
{
	SZ(struct paulis);
	M(struct paulis, pak);
	SZ(struct world_info);
	M(struct world_info, stat);
	M(struct world_info, size);
	M(struct world_info, rank);
	M(struct world_info, seed);
	SZ(struct qreg);
	M(struct qreg, wd);
	M(struct qreg, qb_lo);
	M(struct qreg, qb_hi);
	M(struct qreg, amp);
	M(struct qreg, buf);
	M(struct qreg, namp);
	M(struct qreg, msg_count);
	M(struct qreg, reqs_snd);
	M(struct qreg, reqs_rcv);
	M(struct qreg, nreqs);
	M(struct qreg, backend);
	SZ(struct circ_hamil_term);
	M(struct circ_hamil_term, cf);
	M(struct circ_hamil_term, op);
	SZ(struct circ_hamil);
	M(struct circ_hamil, qb);
	M(struct circ_hamil, terms);
	M(struct circ_hamil, len);
	SZ(struct circ_muldet);
	M(struct circ_muldet, dets);
	M(struct circ_muldet, len);
	SZ(*((struct circ_muldet *)0)->dets);
	SZ(struct circ_values);
	M(struct circ_values, z);
	M(struct circ_values, len);
	SZ(struct data_coeff_block);
	M(struct data_coeff_block, cf);
	M(struct data_coeff_block, C_alpha);
	M(struct data_coeff_block, C_beta);
	SZ(struct data_coeff_matrix);
	M(struct data_coeff_matrix, nqb);
	M(struct data_coeff_matrix, n_sites);
	M(struct data_coeff_matrix, n_alpha);
	M(struct data_coeff_matrix, n_beta);
	M(struct data_coeff_matrix, closed_shell);
	M(struct data_coeff_matrix, tapered);
	M(struct data_coeff_matrix, n_components);
	M(struct data_coeff_matrix, C_alpha);
	M(struct data_coeff_matrix, C_beta);
	M(struct data_coeff_matrix, blocks);
	SZ(struct state_prep_coeff_scratch);
	M(struct state_prep_coeff_scratch, n_sites);
	M(struct state_prep_coeff_scratch, n_alpha);
	M(struct state_prep_coeff_scratch, n_beta);
	M(struct state_prep_coeff_scratch, tup_a);
	M(struct state_prep_coeff_scratch, tup_b);
	M(struct state_prep_coeff_scratch, det_a);
	M(struct state_prep_coeff_scratch, det_b);
	M(struct state_prep_coeff_scratch, Ma);
	M(struct state_prep_coeff_scratch, Mb);
	SZ(struct circ);
	M(struct circ, hm);
	M(struct circ, md);
	M(struct circ, cache);
	M(struct circ, vals);
	M(struct circ, reg);
	M(struct circ, stprep_kind);
	M(struct circ, cm);
	M(struct circ, sp_scratch);
	SZ(struct phase2_step_writer);
	M(struct phase2_step_writer, ctx);
	M(struct phase2_step_writer, write);
	SZ(struct xoshiro256ss);
	M(struct xoshiro256ss, s);
	SZ(struct prob_cdf);
	M(struct prob_cdf, y);
	M(struct prob_cdf, len);
	SZ(struct trott_data);
	M(struct trott_data, steps);
	M(struct trott_data, delta);
	SZ(struct trott);
	M(struct trott, ct);
	M(struct trott, dt);
	M(struct trott, sw);
	SZ(struct trott2_data);
	M(struct trott2_data, steps);
	M(struct trott2_data, delta);
	SZ(struct trott2);
	M(struct trott2, ct);
	M(struct trott2, dt);
	M(struct trott2, sw);
	SZ(struct qdrift_data);
	M(struct qdrift_data, samples);
	M(struct qdrift_data, depth);
	M(struct qdrift_data, step_size);
	M(struct qdrift_data, seed);
	SZ(struct qdrift);
	M(struct qdrift, ct);
	M(struct qdrift, dt);
	M(struct qdrift, ranct);
	M(struct qdrift, rng);
	M(struct qdrift, sw);
	M(struct qdrift_ranct, hm_ran);
	M(struct qdrift_ranct, cdf);
	SZ(struct cmpsit_data);
	M(struct cmpsit_data, samples);
	M(struct cmpsit_data, steps);
	M(struct cmpsit_data, length);
	M(struct cmpsit_data, depth);
	M(struct cmpsit_data, angle_det);
	M(struct cmpsit_data, angle_rand);
	M(struct cmpsit_data, seed);
	SZ(struct cmpsit);
	M(struct cmpsit, ct);
	M(struct cmpsit, dt);
	M(struct cmpsit, ranct);
	M(struct cmpsit, rng);
	M(struct cmpsit, sw);
	M(struct cmpsit_ranct, hm_det);
	M(struct cmpsit_ranct, hm_ran);
	M(struct cmpsit_ranct, hm_smpl);
	M(struct cmpsit_ranct, cdf);
	M(struct cmpsit_ranct, lambda_r);
	M(struct cmpsit_ranct, depth);
	M(struct cmpsit_ranct, angle_rand);
	return 0;
}
