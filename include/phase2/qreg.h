#ifndef PHASE2_QREG_H
#define PHASE2_QREG_H

/*
 * Quantum register on one GPU (replaces include/phase2/qreg.h:25-85 of the
 * reference, BACKEND=hip).  The canonical state is one device buffer of
 * 2^qb complex128 amplitudes (interleaved re, im; 16 B each), allocated
 * with hipMalloc; there is no partner-exchange buffer (qb_hi = 0, buf NULL).
 *
 * The struct has the reference's members in the reference's order, so code
 * compiled against either header sees the same layout.
 *
 * amp is a host shadow of the device state, used only by the bulk
 * qreg_sync_* transfers.  It is allocated at qreg_init for registers of at
 * most QREG_SHADOW_MAX_QB qubits and is NULL above that, so a 33-qubit
 * register needs no 128 GiB of host memory.
 */

#include <stddef.h>
#include <stdint.h>

#include "phase2/paulis.h"
#include "phase2/world.h"

#ifdef __cplusplus
extern "C" {
#endif

#define QREG_MAX_WIDTH (64)
#define QREG_SHADOW_MAX_QB (28)

struct qreg {
	struct world_info wd;

	uint32_t qb_lo, qb_hi;

	_Complex double *amp, *buf;
	uint64_t namp;

	/* The reference's MPI exchange bookkeeping (msg_count, the request
	 * arrays, nreqs).  One GPU exchanges nothing: always 0 / NULL.  Kept
	 * so that the struct -- and every struct that embeds it, struct circ
	 * and the algorithm drivers -- has the reference's layout and size
	 * (the request arrays are MPI_Request pointers there). */
	int msg_count;
	void *reqs_snd, *reqs_rcv;
	size_t nreqs;

	/* Backend handle (struct qreg_hip, private to qreg_hip.c). */
	void *backend;
};

/* 0 on success; -1 (with log_error) if the world is not ready, qb is 0 or
 * larger than QREG_MAX_WIDTH, or the device allocation fails. */
int qreg_init(struct qreg *reg, uint32_t qb);
void qreg_free(struct qreg *reg);

void qreg_getamp(struct qreg *reg, uint64_t i, _Complex double *z);
void qreg_setamp(struct qreg *reg, uint64_t i, _Complex double z);
void qreg_zero(struct qreg *reg);

/* Whole-register copies between the host shadow and the device state;
 * no-ops (with log_error) when the register has no shadow. */
void qreg_sync_host_to_device(struct qreg *reg);
void qreg_sync_device_to_host(struct qreg *reg);

/* Apply exp(i phis[k] code_hi(x)codes_lo[k]) for k = 0..ncodes-1 in order.
 * code_hi must be the identity on one GPU (qb_hi = 0). */
void qreg_paulirot(struct qreg *reg, struct paulis code_hi,
	const struct paulis *codes_lo, const double *phis, size_t ncodes);

/* Sticky error flag of the void entry points above: 0 while every device
 * call has succeeded, else -1.  qreg_clear_error resets it. */
int qreg_error(const struct qreg *reg);
void qreg_clear_error(struct qreg *reg);

#ifdef __cplusplus
}
#endif

#endif /* PHASE2_QREG_H */
