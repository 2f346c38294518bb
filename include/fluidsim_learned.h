/*
 * fluidsim_learned.h -- learned pressure preconditioner of the fs_* library.
 *
 * No reference counterpart: the reference solves the pressure system with the
 * cuDSS direct solver (src/FluidSim/FluidSolver3D.cu:1930-2008).  These entry
 * points train and run a small convolutional network z = M(r) that is used as
 * the preconditioner of flexible PCG when fs_config.precond ==
 * FS_PRECOND_LEARNED, with multigrid PCG as the fallback when the residual
 * stops decreasing (fs_config.stall_window iterations).
 *
 * Network (5 convolutions, 3 resolution levels, 16 channels, 3x3x3 kernels):
 *   A  full   [r/rms(r), fluid mask, air mask] -> 16 ReLU  (a0)
 *      pool 2x                                            (a1)
 *   B  half   16 -> 16                  ReLU          (b1)
 *      pool 2x                                            (b2)
 *   C  quarter 16 -> 16                 ReLU          (c2)
 *   D  half   up(c2) + b1 -> 16         ReLU          (d1)
 *   E  full   up(d1) + a0 -> 1                         z * mask * rms(r)
 * BF16 weights and activations, FP32 accumulation; B, C, D (and their input
 * gradients) run on MFMA.  Parameters are one FP32 vector in the order
 * A.w A.b B.w B.b C.w C.b D.w D.b E.w E.b, each conv weight laid out
 * [tap 27][cin][cout] with tap = (dx+1) + 3(dy+1) + 9(dz+1).
 */
#ifndef FLUIDSIM_LEARNED_H
#define FLUIDSIM_LEARNED_H

#include <stdint.h>

#include "fluidsim.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FSL_NUM_PARAMS (27 * 3 * 16 + 16 + 3 * (27 * 16 * 16 + 16) + 27 * 16 + 1)

/* Initialise the network parameters (He-style, counter-based RNG from seed). */
int fs_learned_init(void* s, uint64_t seed);
int fs_learned_get_params(void* s, float* params, int n);
int fs_learned_set_params(void* s, const float* params, int n);

/* z = M(r) on a W x H x D grid (labels select the fluid cells; z = 0 elsewhere). */
int fs_learned_apply(void* s, const uint8_t* labels, const float* r, float* z, int W, int H, int D);

/* Loss 0.5 d.(A d) / n_fluid with d = zn(b) - p / sqrt(p.Ap / n_fluid) on the
 * fluid cells (A the pressure matrix of the labels; zn = M(b)/rms(b) the
 * unscaled network output, rms over the fluid cells), and its gradient with
 * respect to the parameters (grad: FSL_NUM_PARAMS floats, may be null).  The
 * A-norm is the error norm that CG minimises; flexible PCG is invariant to the
 * scale of z, so the target is the shape of p.  Nothing is updated. */
int fs_learned_loss_grad(void* s, const uint8_t* labels, const float* b, const double* p, int W, int H, int D,
                         double* loss, float* grad);

/* One Adam step on the pair (b, p) (lr: learning rate).  loss: before the step. */
int fs_learned_train_step(void* s, const uint8_t* labels, const float* b, const double* p, int W, int H, int D,
                          float lr, double* loss);

/* One multigrid V-cycle z = V(r) from zero for the pressure system of the
 * labels (the reference's matrix; the plain solver's preconditioner), z on the
 * W x H x D grid (0 off the unknowns).  With r = b it gives the residual
 * b - A V(b) and the error A^-1 b - V(b) that FS_PRECOND_HYBRID's network is
 * trained on.  The handle's pressure grid is overwritten. */
int fs_mg_vcycle(void* s, const uint8_t* labels, const float* r, float* z, int W, int H, int D);

/* Select the pressure preconditioner of later solves (FS_PRECOND_*) and, when
 * stall_window > 0, the learned path's stall window.  FS_PRECOND_HYBRID:
 * z = V(r) + c N(r - A V(r)), a V-cycle corrected by the network applied to
 * the residual it leaves, c = (z2.r1) / (z2.A z2) the exact line search along
 * the network's output z2 (flexible PCG; multigrid fallback on a stall). */
int fs_set_preconditioner(void* s, int precond, int stall_window);

#ifdef __cplusplus
}
#endif
#endif /* FLUIDSIM_LEARNED_H */
