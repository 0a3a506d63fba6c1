/*
 * cube_batch.h -- batched ensemble runner: many independent boxes of the
 * differentiable PM (cube_dpm.h), of uneven sizes, advanced through forward,
 * loss gradient and adjoint together.  No counterpart in the reference
 * (one executable per universe there); for inference callers that need the
 * gradients of an ensemble at once.
 *
 * Each box runs   delta = forward(noise);  loss = mean((delta-target)^2);
 *                 delta_bar = 2 (delta-target)/nf^3;  noise_bar = adjoint.
 * mode 1 captures the whole ensemble once into one HIP graph (one branch per
 * box, each on its own stream, joined at the end) and replays it with a
 * single launch; mode 0 runs the same kernels box after box on one stream
 * (the per-box loop).  cube_batch_tune times candidate launch shapes of the
 * model's particle and mesh loops for every distinct box size and keeps the
 * fastest.  The I/O pointers are device pointers fixed by set_io; a captured
 * graph keeps using them (call set_io before the first mode-1 run, or again
 * to force a new capture).
 */
#ifndef CUBE_BATCH_H
#define CUBE_BATCH_H
#include <stdint.h>
#include "cube_dpm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cube_batch cube_batch;

int cube_batch_create(int32_t nbox, const cube_dpm_params *params, cube_batch **out);
void cube_batch_destroy(cube_batch *b);
/* the box handle, for cube_dpm_set_amplitude / cube_dpm_set_schedule */
int cube_batch_box(cube_batch *b, int32_t i, cube_dpm **box);
/* device pointers of box i: noise, noise_bar ng^3; target, delta nf^3;
   loss: one double */
int cube_batch_set_io(cube_batch *b, int32_t i, const float *noise, const float *target, float *delta,
                      float *noise_bar, double *loss);
/* shapes (optional, 2*nbox ints): the chosen grid and block of each box */
int cube_batch_tune(cube_batch *b, int32_t *shapes);
int cube_batch_run(cube_batch *b, int32_t mode);
int cube_batch_sync(cube_batch *b);

#ifdef __cplusplus
}
#endif
#endif
