/*
 * fluidsim_noalloc.h -- an allocation-free variant of fs_solve_pressure_system.
 *
 * No counterpart in the reference (its pressure solve allocates device memory
 * for the matrix, right-hand side and cuDSS workspace on every step,
 * src/FluidSim/FluidSolver3D.cu:1815-2008).  For callers that cannot allocate
 * device memory: code that runs between HIP graph captures, tight serving
 * loops, harnesses that forbid hipMalloc.
 *
 * fs_solve_pressure_system_noalloc solves the same system with the same
 * kernels and gives the same bits as fs_solve_pressure_system on a handle with
 * default settings (multigrid-preconditioned CG, at most 500 iterations), but
 * needs no handle and calls no hipMalloc, hipFree or other allocating routine:
 * every device buffer (staging of labels, b and p, the multigrid hierarchy,
 * the CG vectors, partial sums and counters) is carved from one fixed
 * workspace, a file-scope device array of FS_NOALLOC_SCRATCH_BYTES that each
 * GPU the library runs on holds until the process ends.  It works on the
 * current device, on the default stream, and returns when the solution is in
 * p.
 *
 * Limits: W*H*D <= FS_NOALLOC_MAX_CELLS, and the workspace of the shape (about
 * 63 bytes per cell for a cube, more for a thin line of cells) must fit;
 * larger systems fail with FS_ERR_ARG and a message giving both sizes.
 * Calls that overlap on one GPU (from several threads) race on its
 * workspace: serialise them.  The workspace keeps the last call's contents;
 * every call clears what must start at zero.
 */
#ifndef FLUIDSIM_NOALLOC_H
#define FLUIDSIM_NOALLOC_H

#include <stdint.h>

#include "fluidsim.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FS_NOALLOC_MAX_CELLS (1 << 20)
#define FS_NOALLOC_SCRATCH_BYTES (96u << 20)

/* labels: uint8 W*H*D (FS_SOLID/FLUID/AIR/BODY), b: float W*H*D (0 off fluid),
 * p: double W*H*D out (0 off fluid); host memory, i-fastest. */
int fs_solve_pressure_system_noalloc(const uint8_t* labels, const float* b, double* p, int W, int H, int D,
                                     double tol, fs_solve_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* FLUIDSIM_NOALLOC_H */
