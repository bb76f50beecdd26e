#ifndef PHASE2_CIRC_H
#define PHASE2_CIRC_H

/*
 * Circuit context: Hamiltonian, initial-state payload, register, batch
 * cache and per-step results (replaces include/phase2/circ.h:47-299 of the
 * reference; same types, names, ownership and return conventions).
 *
 * Every struct has the reference's members in the reference's order (the
 * same layout and size).  circ_init adopts hm and the state-prep payload
 * by value and frees them itself on error; circ_free releases everything;
 * the *_free helpers are idempotent.  Functions return 0 or -1 (with
 * log_error).
 */

#include <stddef.h>
#include <stdint.h>

#include "phase2/paulis.h"
#include "phase2/qreg.h"
#include "phase2/state_prep_coeff.h"

#ifdef __cplusplus
extern "C" {
#endif

struct circ_hamil {
	uint32_t qb;
	struct circ_hamil_term {
		double cf;
		struct paulis op;
	} *terms;
	size_t len;
};

struct circ_muldet {
	struct {
		uint64_t idx;
		_Complex double cf;
	} *dets;
	size_t len;
};

struct circ_values {
	_Complex double *z;
	size_t len;
};

enum stprep_kind {
	STPREP_MULTIDET = 1,
	STPREP_COEFF_MATRIX = 2,
};

struct data_coeff_block {
	double cf;
	const double *C_alpha;
	const double *C_beta;
};

struct data_coeff_matrix {
	uint32_t nqb;
	uint32_t n_sites;
	uint32_t n_alpha;
	uint32_t n_beta;
	int closed_shell;
	int tapered;
	size_t n_components;
	const double *C_alpha;
	const double *C_beta;
	struct data_coeff_block *blocks;
};

struct circ_cache;

struct circ {
	struct circ_hamil hm;
	struct circ_muldet md;
	struct circ_cache *cache;
	struct circ_values vals;
	struct qreg reg;
	enum stprep_kind stprep_kind;
	struct data_coeff_matrix cm;
	struct state_prep_coeff_scratch sp_scratch;
};

int circ_hamil_init(struct circ_hamil *hm, uint32_t qb, size_t len);
void circ_hamil_free(struct circ_hamil *hm);
void circ_hamil_sort_lex(struct circ_hamil *hm);

int circ_muldet_init(struct circ_muldet *md, size_t len);
void circ_muldet_free(struct circ_muldet *md);

void data_coeff_matrix_free(struct data_coeff_matrix *cm);

int circ_values_init(struct circ_values *vals, size_t len);
void circ_values_free(struct circ_values *vals);

int circ_init(struct circ *ct, struct circ_hamil hm, enum stprep_kind sp_kind,
	const void *sp_data, size_t vals_len);
void circ_free(struct circ *ct);

int circ_prepst(struct circ *ct);
int circ_step(struct circ *ct, const struct circ_hamil *hm, double omega);
int circ_step_reverse(struct circ *ct, const struct circ_hamil *hm,
	double omega);
_Complex double circ_measure(struct circ *ct);

#ifdef __cplusplus
}
#endif

#endif /* PHASE2_CIRC_H */
