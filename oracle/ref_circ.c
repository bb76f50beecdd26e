/*
 * ref_circ.c -- times a whole reference circuit run (trott_simul,
 * qdrift_simul; circ/trott.c:44-65, circ/qdrift.c:105-133) on the reference
 * CPU backend under mpirun, for the CPU baselines printed by bench.py.
 * TEST/BENCH INFRASTRUCTURE ONLY: built by oracle/Makefile.ref from the
 * reference sources into oracle/_ref/ref_circ.
 *
 *   mpirun -n P ref_circ JOB_FILE
 *
 * JOB_FILE (little endian):
 *   u32 algo (1 trott, 3 qdrift), u32 nqb,
 *   u64 nterms, u64 ndet, u64 steps, u64 samples, u64 depth, u64 seed,
 *   f64 delta (trott) or step_size (qdrift),
 *   nterms x {u64 flip, u64 phase, f64 cf},
 *   ndet x {u64 idx, f64 re, f64 im}.
 * Rank 0 prints one JSON line with the wall time between barriers around
 * init + simul and the first overlap.
 */
#include <complex.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mpi.h"

#include "phase2.h"
#include "circ/trott.h"
#include "circ/qdrift.h"

struct hdr {
	uint32_t algo, nqb;
	uint64_t nterms, ndet, steps, samples, depth, seed;
	double delta;
};

struct term {
	uint64_t x, z;
	double cf;
};

struct det {
	uint64_t idx;
	double re, im;
};

int main(int argc, char **argv)
{
	if (argc < 2) {
		fprintf(stderr, "usage: ref_circ JOB_FILE\n");
		return 2;
	}
	if (world_init(&argc, &argv, UINT64_C(0x5eed0f0dd1ce5eed)) !=
	    WORLD_READY)
		return 1;
	FILE *f = fopen(argv[1], "rb");
	if (!f)
		return 1;
	struct hdr h;
	if (fread(&h, sizeof h, 1, f) != 1)
		return 1;
	struct term *t = malloc(sizeof *t * (h.nterms ? h.nterms : 1));
	struct det *d = malloc(sizeof *d * (h.ndet ? h.ndet : 1));
	if (!t || !d || fread(t, sizeof *t, h.nterms, f) != h.nterms ||
	    fread(d, sizeof *d, h.ndet, f) != h.ndet)
		return 1;
	fclose(f);

	struct circ_hamil hm;
	struct circ_muldet md;
	if (circ_hamil_init(&hm, h.nqb, h.nterms) < 0 ||
	    circ_muldet_init(&md, h.ndet) < 0)
		return 1;
	for (uint64_t k = 0; k < h.nterms; k++) {
		hm.terms[k].cf = t[k].cf;
		hm.terms[k].op.pak[0] = t[k].x;
		hm.terms[k].op.pak[1] = t[k].z;
	}
	for (uint64_t k = 0; k < h.ndet; k++) {
		md.dets[k].idx = d[k].idx;
		md.dets[k].cf = CMPLX(d[k].re, d[k].im);
	}

	MPI_Barrier(MPI_COMM_WORLD);
	const double t0 = MPI_Wtime();
	_Complex double z0 = 0;
	int rc = -1;
	uint64_t rotations = 0;
	if (h.algo == 1) {
		struct trott tt;
		struct trott_data dt = { .steps = h.steps, .delta = h.delta };
		if (trott_init(&tt, &dt, hm, STPREP_MULTIDET, &md, NULL) == 0) {
			rc = trott_simul(&tt);
			z0 = tt.ct.vals.z[0];
			trott_free(&tt);
		}
		rotations = h.steps * h.nterms;
	} else if (h.algo == 3) {
		struct qdrift qd;
		struct qdrift_data dt = { .samples = h.samples,
			.depth = h.depth, .step_size = h.delta, .seed = h.seed };
		if (qdrift_init(&qd, &dt, hm, STPREP_MULTIDET, &md, NULL) == 0) {
			rc = qdrift_simul(&qd);
			z0 = qd.ct.vals.z[0];
			qdrift_free(&qd);
		}
		rotations = h.samples * h.depth;
	}
	MPI_Barrier(MPI_COMM_WORLD);
	const double t1 = MPI_Wtime();
	struct world_info wd;
	world_info(&wd);
	if (rc == 0 && wd.rank == 0) {
		const double secs = t1 - t0;
		const double upd = (double)rotations *
			(double)(UINT64_C(1) << h.nqb);
		printf("{\"ranks\": %d, \"nqb\": %u, \"rotations\": %llu, "
		       "\"seconds\": %.6f, \"amp_updates_per_s\": %.6e, "
		       "\"z0_re\": %.17g, \"z0_im\": %.17g}\n",
			wd.size, h.nqb, (unsigned long long)rotations, secs,
			upd / secs, creal(z0), cimag(z0));
		fflush(stdout);
	}
	/* hm and md were adopted by the circuit and freed with it. */
	free(t);
	free(d);
	world_free();
	return rc == 0 ? 0 : 1;
}
