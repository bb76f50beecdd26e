#ifndef PHASE2_WORLD_H
#define PHASE2_WORLD_H

/*
 * Process lifecycle (replaces include/phase2/world.h:16-77 of the
 * reference).  The HIP backend is single-process: one GPU holds the whole
 * register, so size is 1 and rank 0 and no MPI is involved.  world_init
 * brings up logging and the HIP device (PHASE2_HIP_DEVICE, default 0).
 */

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum world_stat {
	WORLD_UNDEF = -1,
	WORLD_READY = 0,
	WORLD_DONE = 1,
	WORLD_ERR = 2,
};

/* The backend number of a build (0 qreg, 2 CUDA in the reference; 3 HIP);
 * a consumer may set it on the command line, as the reference's headers
 * allow. */
#ifndef PHASE2_BACKEND
#define PHASE2_BACKEND 3
#endif
#define WORLD_BACKEND "HIP"

struct world_info {
	enum world_stat stat;
	int size;
	int rank;
	uint64_t seed;
};

/* argc/argv are accepted for signature compatibility and ignored; seed must
 * be non-zero.  Returns WORLD_READY or WORLD_ERR. */
int world_init(int *argc, char ***argv, uint64_t seed);
int world_free(void);
int world_info(struct world_info *wd);

#ifdef __cplusplus
}
#endif

#endif /* PHASE2_WORLD_H */
