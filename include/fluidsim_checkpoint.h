/*
 * fluidsim_checkpoint.h -- checkpoint and resume of an fs_* solver handle.
 *
 * No reference counterpart: the reference keeps its whole state in the
 * FluidSolver3D object and has no snapshot (src/FluidSim/FluidSolver3D.cuh:
 * 13-186).  A checkpoint is one binary file holding everything the next step
 * reads: the configuration, the grid size, the solid mask and labels, the
 * particles (positions, velocities, original ids, in their current order),
 * the rigid body (state and SDF), the clock (time, step count, current dt),
 * the last step's pressure (the next solve's initial guess, so that a resumed
 * run takes the same iterations) and, when the handle has one, the learned
 * preconditioner's parameters (its optimiser state is not stored: training is
 * not part of a step).  The CPU oracle reads and writes the same format.
 * Grid velocities and pressure are not stored: every step rebuilds them from
 * the particles.  The format is host-endian and versioned; loading checks the
 * header, and every size it reads against the bytes left in the file before
 * allocating, and fails with FS_ERR_IO on a mismatch (the handle is then
 * unchanged).
 */
#ifndef FLUIDSIM_CHECKPOINT_H
#define FLUIDSIM_CHECKPOINT_H

#include "fluidsim.h"

#ifdef __cplusplus
extern "C" {
#endif

int fs_save_checkpoint(void* s, const char* path);
/* Replaces the state of an existing handle (its configuration included). */
int fs_load_checkpoint(void* s, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* FLUIDSIM_CHECKPOINT_H */
