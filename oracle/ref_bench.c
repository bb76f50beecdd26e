/*
 * ref_bench.c -- times the reference CPU backend's rotation sweep
 * (qreg_paulirot, phase2/qreg.c:135-142 with phase2/qreg_qreg.c) under
 * mpirun, for the CPU baseline printed by bench.py.  TEST/BENCH
 * INFRASTRUCTURE ONLY: built by oracle/Makefile.ref from the reference
 * sources into oracle/_ref/ref_bench.
 *
 *   mpirun -n P ref_bench TERMS_FILE NQB
 *
 * TERMS_FILE holds K records {uint64 flip, uint64 phase, double phi}.  The
 * register starts as |0...0>; the K rotations are applied in blocks of at
 * most 1024 (the circ_cache cap) with the identity hi code split off as
 * circ_cache does.  Rank 0 prints one JSON line.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mpi.h"

#include "phase2.h"

struct rec {
	uint64_t x, z;
	double phi;
};

int main(int argc, char **argv)
{
	if (argc < 3) {
		fprintf(stderr, "usage: ref_bench TERMS_FILE NQB\n");
		return 2;
	}
	if (world_init(&argc, &argv, UINT64_C(0x1234567887654321)) !=
	    WORLD_READY)
		return 1;
	const uint32_t nqb = (uint32_t)atoi(argv[2]);
	FILE *f = fopen(argv[1], "rb");
	if (!f)
		return 1;
	fseek(f, 0, SEEK_END);
	const long bytes = ftell(f);
	fseek(f, 0, SEEK_SET);
	const size_t K = (size_t)bytes / sizeof(struct rec);
	struct rec *r = malloc(sizeof *r * (K ? K : 1));
	if (!r || fread(r, sizeof *r, K, f) != K)
		return 1;
	fclose(f);

	struct qreg reg;
	if (qreg_init(&reg, nqb) < 0)
		return 1;
	qreg_zero(&reg);
	qreg_setamp(&reg, 0, 1.0);

	struct paulis lo[1024], hi;
	double phis[1024];
	MPI_Barrier(MPI_COMM_WORLD);
	const double t0 = MPI_Wtime();
	size_t k = 0;
	while (k < K) {
		size_t nb = 0;
		struct paulis hi0 = paulis_new();
		while (k < K && nb < 1024) {
			struct paulis p = { .pak = { r[k].x, r[k].z } }, l, h;
			paulis_split(p, reg.qb_lo, reg.qb_hi, &l, &h);
			if (nb == 0)
				hi0 = h;
			else if (!paulis_eq(h, hi0))
				break;
			lo[nb] = l;
			phis[nb] = r[k].phi;
			nb++;
			k++;
		}
		hi = hi0;
		qreg_paulirot(&reg, hi, lo, phis, nb);
	}
	MPI_Barrier(MPI_COMM_WORLD);
	const double t1 = MPI_Wtime();
	_Complex double z0;
	qreg_getamp(&reg, 0, &z0);
	if (reg.wd.rank == 0) {
		const double secs = t1 - t0;
		const double upd = (double)K * (double)(UINT64_C(1) << nqb);
		printf("{\"ranks\": %d, \"nqb\": %u, \"terms\": %zu, "
		       "\"seconds\": %.6f, \"amp_updates_per_s\": %.6e, "
		       "\"amp0_re\": %.17g}\n",
			reg.wd.size, nqb, K, secs, upd / secs, __real__ z0);
		fflush(stdout);
	}
	qreg_free(&reg);
	free(r);
	world_free();
	return 0;
}
