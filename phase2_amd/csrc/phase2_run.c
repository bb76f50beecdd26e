/*
 * phase2_run: the entry the Python package binds (behaviour of
 * phase2/phase2_run.c:15-174 of the reference).  Parses "X0 Z1 ..." term
 * strings and a '0'/'1' basis string, applies the rotations in the given
 * order through the batch cache, and returns amp[idx] = <psi|U|psi>.
 */
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "phase2/paulis.h"
#include "phase2/phase2_run.h"
#include "phase2/qreg.h"
#include "phase2/world.h"

#include "circ_cache.h"

#define PH2RUN_SEED UINT64_C(0xa7c3e1b9d4f20586)
#define PH2RUN_MAX_QB 30

static int own_world;

static void release_world(void)
{
	if (own_world)
		world_free();
}

static int world_up(void)
{
	struct world_info wd;
	if (world_info(&wd) == WORLD_READY)
		return 0;
	if (world_init(NULL, NULL, PH2RUN_SEED) != WORLD_READY)
		return -1;
	own_world = 1;
	atexit(release_world);
	return 0;
}

static int parse_basis(const char *s, uint32_t nqb, uint64_t *out)
{
	uint64_t v = 0;
	for (uint32_t k = 0; k < nqb; k++) {
		if (s[k] == '1')
			v |= UINT64_C(1) << k;
		else if (s[k] != '0')
			return -1;
	}
	*out = v;
	return 0;
}

/* Blank-separated tokens "<op><qubit>", op in IXYZ.  Tokens are split on
 * blanks, tabs and newlines as strtok(" \t\n") splits them, and the
 * qubit field is read as strtoul(field, &end, 10) reads it in the
 * reference (phase2/phase2_run.c:60-103): leading white space (only \v,
 * \f and \r can be left inside a token) is skipped, then an optional
 * sign and at least one digit; the value is taken modulo 2^64 and negated
 * as an unsigned number after a '-', so "X-18446744073709551615" is qubit
 * 1; a magnitude of 2^64 or more saturates (ERANGE) and fails the qubit <
 * nqb check; nothing may follow the digits.  Strings of 4096 characters or
 * more are rejected, as there. */
static int parse_pauli(const char *s, uint32_t nqb, struct paulis *out)
{
	if (strlen(s) >= 4096)
		return -1;
	*out = paulis_new();
	const char *p = s;
	for (;;) {
		while (*p == ' ' || *p == '\t' || *p == '\n')
			p++;
		if (!*p)
			return 0;
		enum pauli_op op;
		switch (*p) {
		case 'I': op = PAULI_I; break;
		case 'X': op = PAULI_X; break;
		case 'Y': op = PAULI_Y; break;
		case 'Z': op = PAULI_Z; break;
		default: return -1;
		}
		p++;
		while (*p == '\v' || *p == '\f' || *p == '\r')
			p++;
		int neg = 0;
		if (*p == '+' || *p == '-')
			neg = *p++ == '-';
		if (*p < '0' || *p > '9')
			return -1;
		uint64_t q = 0;
		int range = 0;	/* magnitude >= 2^64 */
		while (*p >= '0' && *p <= '9') {
			const uint64_t d = (uint64_t)(*p - '0');
			if (q > UINT64_MAX / 10 ||
			    (q == UINT64_MAX / 10 && d > UINT64_MAX % 10))
				range = 1;
			q = q * 10 + d;
			p++;
		}
		if (*p && *p != ' ' && *p != '\t' && *p != '\n')
			return -1;
		if (range)
			q = UINT64_MAX;
		else if (neg)
			q = (uint64_t)0 - q;
		if (q >= nqb)
			return -1;
		paulis_set(out, op, (uint32_t)q);
	}
}

static void flush_cb(struct paulis hi, const struct paulis *lo, double *phis,
	size_t n, void *data)
{
	qreg_paulirot((struct qreg *)data, hi, lo, phis, n);
}

int phase2_run(const char **pauli_strs, const double *coeffs, size_t nterms,
	double delta, const char *psi_str, double *out_re, double *out_im)
{
	if (!psi_str)
		return -1;
	const size_t len = strlen(psi_str);
	if (len == 0 || len > PH2RUN_MAX_QB)
		return -1;
	const uint32_t nqb = (uint32_t)len;
	uint64_t idx;
	if (world_up() < 0 || parse_basis(psi_str, nqb, &idx) < 0)
		return -1;
	struct qreg reg;
	if (qreg_init(&reg, nqb) < 0)
		return -1;
	qreg_zero(&reg);
	qreg_setamp(&reg, idx, 1.0);
	struct circ_cache *cache = circ_cache_init((int)reg.qb_hi,
		(int)reg.qb_lo);
	int rc = -1;
	if (!cache)
		goto out;
	for (size_t k = 0; k < nterms; k++) {
		struct paulis code;
		if (!pauli_strs[k] || parse_pauli(pauli_strs[k], nqb, &code) < 0)
			goto out;
		const double phi = delta * coeffs[k];
		if (circ_cache_insert(cache, code, phi) != 0) {
			circ_cache_flush(cache, flush_cb, &reg);
			if (circ_cache_insert(cache, code, phi) < 0)
				goto out;
		}
	}
	circ_cache_flush(cache, flush_cb, &reg);
	_Complex double z;
	qreg_getamp(&reg, idx, &z);
	if (qreg_error(&reg) == 0) {
		*out_re = creal(z);
		*out_im = cimag(z);
		rc = 0;
	}
out:
	circ_cache_free(cache);
	qreg_free(&reg);
	return rc;
}
