/*
 * cube_dpm.h -- differentiable particle-mesh forward model and its adjoint
 * on one GPU.  No counterpart in the reference (which has no gradients):
 * this is the gradient form of the CUBE step, for inference and
 * optimisation callers that need d(loss)/d(initial white noise).
 *
 * Model (FP32 decompressed state, no fixed-point codecs, no PP):
 *   ICs      psi_d = (1/ng^3) IFFT[ i k_d/k^2 A(k) FFT(w) ]   on the ng^3
 *            particle lattice (w = white noise, A = amplitude per mode),
 *            x0 = q + cx * psi * (nf/ng),  p0 = cp * psi * (nf/ng)
 *            (q = lattice point, positions in mesh cells of the nf^3 mesh;
 *            the context holds each particle as its FP32 displacement x - q)
 *   force    F(x) = CIC-gather of (1/nf^3) IFFT[ i k/k^2 FFT(CIC-deposit(x)) ]
 *   step s   p += kA_s F(x);  x += d_s p;  p += kB_s F(x)      (KDK)
 *   output   delta = CIC-deposit(x_final) / (np/nf^3) - 1 on the nf^3 mesh
 * The schedule coefficients (cx, cp, kA_s, d_s, kB_s) are computed by the
 * caller (host FP64 cosmology), so any time stepping is expressible.
 *
 * The adjoint runs the transposed operators backwards in time: cotangent
 * scatter (the transpose of the gather), the Poisson operator with the
 * conjugate factor -i k/k^2, the CIC weight-gradient gather (the position
 * derivative of deposit and gather), the transposed drift and kicks, and
 * the transposed IC filter.  The forward either keeps the positions of
 * every step (checkpoint=1: 12 bytes per particle per step) or only the
 * ICs and recomputes (checkpoint=0: O(nsteps^2) force evaluations).
 *
 * Array arguments of forward/adjoint are DEVICE pointers on the context's
 * device (e.g. torch tensors' data_ptr()); the amplitude and schedule are
 * host arrays, copied.  Return codes are those of cube.h.
 */
#ifndef CUBE_DPM_H
#define CUBE_DPM_H
#include <stdint.h>
#include "cube.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t ng;         /* particles per dimension (lattice), power of two or 2^a 3^b */
  int32_t nf;         /* density / force mesh per dimension, a multiple of ng */
  int32_t nsteps;     /* KDK steps */
  int32_t checkpoint; /* 1: keep every step's positions; 0: recompute */
  int32_t device;
  int32_t reserved;
} cube_dpm_params;

typedef struct cube_dpm cube_dpm;

int cube_dpm_create(const cube_dpm_params *p, cube_dpm **out);
void cube_dpm_destroy(cube_dpm *d);

/* A(k) on the lattice half spectrum, ng*ng*(ng/2+1) floats, [kz][ky][kx] */
int cube_dpm_set_amplitude(cube_dpm *d, const float *amp);
/* ic[2] = {cx, cp}; steps[3*nsteps] = {kA_s, d_s, kB_s} */
int cube_dpm_set_schedule(cube_dpm *d, const double *ic, const double *steps);

/* noise: ng^3 floats (device); delta: nf^3 floats (device), written. */
int cube_dpm_forward(cube_dpm *d, const float *noise, float *delta);
/* delta_bar: nf^3 (device) = dL/d(delta); noise_bar: ng^3 (device), written.
   Needs the forward of the same noise just before. */
int cube_dpm_adjoint(cube_dpm *d, const float *delta_bar, float *noise_bar);
/* final positions (wrapped into [0, nf)) and momenta after the last
   forward, np*3 floats each (device; either may be NULL) */
int cube_dpm_state(cube_dpm *d, float *x, float *p);
/* wait for the context's stream */
int cube_dpm_sync(cube_dpm *d);

#ifdef __cplusplus
}
#endif
#endif
