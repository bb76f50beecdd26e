/*
 * Process lifecycle of the HIP backend (replaces phase2/world.c and
 * phase2/world_cuda.c of the reference): one process, one GPU.  world_init
 * starts logging and selects the device (PHASE2_HIP_DEVICE, default 0);
 * there is no MPI, so size is 1 and rank 0.
 */
#define LOG_SUBSYS "world"

#include <stdint.h>
#include <stdlib.h>

#include "log.h"
#include "phase2/world.h"
#include "phase2_hip.h"

static struct world_info WORLD = {
	.stat = WORLD_UNDEF,
	.size = 0,
	.rank = 0,
	.seed = UINT64_C(0x77dd8e60521fb661),
};

int world_init(int *argc, char ***argv, uint64_t seed)
{
	(void)argc;
	(void)argv;
	if (seed == 0 || log_init() < 0)
		return WORLD.stat = WORLD_ERR;
	if (ph2hip_init(-1) < 0) {
		log_error("world_init: HIP device: %s", ph2hip_strerror());
		return WORLD.stat = WORLD_ERR;
	}
	WORLD.size = 1;
	WORLD.rank = 0;
	WORLD.seed = seed;
	WORLD.stat = WORLD_READY;
	log_info("*** Init ***");
	log_info("Backend: %s", WORLD_BACKEND);
	return WORLD.stat;
}

int world_free(void)
{
	if (WORLD.stat == WORLD_READY) {
		ph2hip_shutdown();
		WORLD.stat = WORLD_DONE;
	}
	return WORLD.stat;
}

int world_info(struct world_info *wd)
{
	if (wd)
		*wd = WORLD;
	return WORLD.stat;
}
