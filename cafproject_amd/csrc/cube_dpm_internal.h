/*
 * cube_dpm_internal.h -- state of one differentiable-PM box, shared by
 * cube_dpm.hip (the model) and cube_batch.hip (the ensemble runner).
 */
#ifndef CUBE_DPM_INTERNAL_H
#define CUBE_DPM_INTERNAL_H
#include "cube_internal.h"
#include "../../include/cube_dpm.h"

struct cube_dpm {
  cube_dpm_params p;
  cube_ctx c; /* stream + reduction word for the FFT layer */
  int64_t np, ng3, nf3, sg, sf; /* sg/sf: spectrum sizes (pitched) */
  FFTPlan *plan_g, *plan_f;
  double ic[2];
  double *steps; /* host, 3*nsteps */
  float *amp;
  float *kern;          /* 3 * nf*nf*(nf/2+1), [d][kz][ky][kx] */
  float *xs;            /* checkpoints: (nsteps+1 or 1) * np*3 */
  float *x, *pm, *p0;   /* current positions, momenta; ICs' momenta */
  float *rho, *force, *S, *lam;
  unsigned long long *acc; /* 3*nf^3 fixed-point sums of the deposit / its transpose */
  unsigned *gmx;           /* max |g| of the transpose's cotangents (float bits) */
  float2 *spec_f, *work3_f;
  float2 *spec_g, *work3_g;
  float *psi;           /* 3*ng^3 */
  float *xbar, *pbar, *gvec;
  int has_forward;
  int grid, block;      /* launch shape of the particle and mesh loops (tunable) */
};

#endif
