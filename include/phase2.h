#ifndef PHASE2_H
#define PHASE2_H

/* Umbrella header of the HIP build of the phase2 library (replaces
 * include/phase2.h of the reference). */

#include <stddef.h>

#include "phase2/circ.h"
#include "phase2/paulis.h"
#include "phase2/qreg.h"
#include "phase2/step_writer.h"
#include "phase2/world.h"
#include "prob.h"

#endif /* PHASE2_H */
