/*
 * cube_batch.hip -- ensemble runner over differentiable-PM boxes
 * (include/cube_batch.h): per-box loop, one fork-join HIP graph for the
 * whole ensemble, and a launch-shape autotuner.
 *
 * The model's kernels are small for the small boxes (a 32^3-particle box
 * with a 64^3 mesh runs a few microseconds of work per launch over
 * hundreds of launches), so one box at a time leaves the GPU idle between
 * launches.  The graph removes the per-launch host cost and lets the
 * branches of all boxes run concurrently.  The graph holds kernel and
 * device-to-device copy nodes only (no memset nodes: buffers are cleared by
 * kernels).
 */
#include "cube_dpm_internal.h"
#include "../../include/cube_batch.h"
#include <stdlib.h>
#include <string.h>

#define MSE_B 256

struct BoxIO {
  const float *noise, *target;
  float *delta, *noise_bar;
  double *loss;
};

struct cube_batch {
  int nbox;
  int device;
  cube_dpm **box;
  BoxIO *io;
  float **dbar;         /* per box nf^3: the loss gradient */
  double **part;        /* per box MSE_B partial sums */
  hipStream_t s0;       /* the per-box loop stream and the capture origin */
  hipEvent_t fork, *join, t0, t1;
  hipGraph_t graph;
  hipGraphExec_t exec;
  int captured;
};

/* dbar = 2 (delta - target)/n; per-block partial sums of (delta-target)^2 */
__global__ __launch_bounds__(MSE_B) void k_mse(const float *delta, const float *target, int64_t n, float *dbar,
                                               double *part) {
  __shared__ double red[MSE_B];
  double s = 0;
  const float sc = 2.0f / (float)n;
  for (int64_t q = blockIdx.x * (int64_t)MSE_B + threadIdx.x; q < n; q += (int64_t)gridDim.x * MSE_B) {
    const float r = delta[q] - target[q];
    dbar[q] = sc * r;
    s += (double)r * (double)r;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = MSE_B / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
/* fixed-order total: loss = sum(part)/n */
__global__ void k_mse_total(const double *part, int nb, double inv_n, double *loss) {
  __shared__ double red[MSE_B];
  red[threadIdx.x] = threadIdx.x < nb ? part[threadIdx.x] : 0.0;
  __syncthreads();
  for (int o = MSE_B / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0] * inv_n;
}

/* forward, loss gradient, adjoint of box i on stream s */
static int box_run(cube_batch *b, int i, hipStream_t s) {
  cube_dpm *d = b->box[i];
  const BoxIO &io = b->io[i];
  if (!io.noise || !io.target || !io.delta || !io.noise_bar || !io.loss) return CUBE_ERR_STATE;
  const hipStream_t own = d->c.stream;
  d->c.stream = s;
  int st = cube_dpm_forward(d, io.noise, io.delta);
  if (!st) {
    hipLaunchKernelGGL(k_mse, dim3(MSE_B), dim3(MSE_B), 0, s, io.delta, io.target, d->nf3, b->dbar[i], b->part[i]);
    hipLaunchKernelGGL(k_mse_total, dim3(1), dim3(MSE_B), 0, s, b->part[i], MSE_B, 1.0 / (double)d->nf3, io.loss);
    if (hipGetLastError() != hipSuccess) st = CUBE_ERR_HIP;
  }
  if (!st) st = cube_dpm_adjoint(d, b->dbar[i], io.noise_bar);
  d->c.stream = own;
  return st;
}

extern "C" void cube_batch_destroy(cube_batch *b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (int i = 0; i < b->nbox; ++i) {
    if (b->box && b->box[i]) cube_dpm_destroy(b->box[i]);
    if (b->dbar && b->dbar[i]) (void)hipFree(b->dbar[i]);
    if (b->part && b->part[i]) (void)hipFree(b->part[i]);
    if (b->join && b->join[i]) (void)hipEventDestroy(b->join[i]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->t0) (void)hipEventDestroy(b->t0);
  if (b->t1) (void)hipEventDestroy(b->t1);
  if (b->s0) (void)hipStreamDestroy(b->s0);
  free(b->box);
  free(b->io);
  free(b->dbar);
  free(b->part);
  free(b->join);
  free(b);
}

extern "C" int cube_batch_create(int32_t nbox, const cube_dpm_params *params, cube_batch **out) {
  if (nbox < 1 || !params || !out) return CUBE_ERR_INVALID;
  cube_batch *b = (cube_batch *)calloc(1, sizeof(cube_batch));
  if (!b) return CUBE_ERR_ALLOC;
  b->nbox = nbox;
  b->device = params[0].device;
  b->box = (cube_dpm **)calloc(nbox, sizeof(cube_dpm *));
  b->io = (BoxIO *)calloc(nbox, sizeof(BoxIO));
  b->dbar = (float **)calloc(nbox, sizeof(float *));
  b->part = (double **)calloc(nbox, sizeof(double *));
  b->join = (hipEvent_t *)calloc(nbox, sizeof(hipEvent_t));
  int st = (b->box && b->io && b->dbar && b->part && b->join) ? CUBE_OK : CUBE_ERR_ALLOC;
  if (!st && hipSetDevice(b->device) != hipSuccess) st = CUBE_ERR_HIP;
  if (!st && hipStreamCreate(&b->s0) != hipSuccess) st = CUBE_ERR_HIP;
  if (!st && hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) st = CUBE_ERR_HIP;
  if (!st && hipEventCreate(&b->t0) != hipSuccess) st = CUBE_ERR_HIP;
  if (!st && hipEventCreate(&b->t1) != hipSuccess) st = CUBE_ERR_HIP;
  for (int i = 0; i < nbox && !st; ++i) {
    if (params[i].device != b->device) st = CUBE_ERR_INVALID;
    if (!st) st = cube_dpm_create(&params[i], &b->box[i]);
    const int64_t nf3 = (int64_t)params[i].nf * params[i].nf * params[i].nf;
    if (!st && hipMalloc((void **)&b->dbar[i], nf3 * sizeof(float)) != hipSuccess) st = CUBE_ERR_ALLOC;
    if (!st && hipMalloc((void **)&b->part[i], MSE_B * sizeof(double)) != hipSuccess) st = CUBE_ERR_ALLOC;
    if (!st && hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) != hipSuccess) st = CUBE_ERR_HIP;
  }
  if (st) {
    cube_batch_destroy(b);
    return st;
  }
  *out = b;
  return CUBE_OK;
}

extern "C" int cube_batch_box(cube_batch *b, int32_t i, cube_dpm **box) {
  if (!b || i < 0 || i >= b->nbox || !box) return CUBE_ERR_INVALID;
  *box = b->box[i];
  return CUBE_OK;
}

extern "C" int cube_batch_set_io(cube_batch *b, int32_t i, const float *noise, const float *target, float *delta,
                                 float *noise_bar, double *loss) {
  if (!b || i < 0 || i >= b->nbox) return CUBE_ERR_INVALID;
  b->io[i] = BoxIO{noise, target, delta, noise_bar, loss};
  b->captured = 0; /* the graph holds the old pointers */
  return CUBE_OK;
}

static int capture(cube_batch *b) {
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  CUBE_CHECK(hipStreamBeginCapture(b->s0, hipStreamCaptureModeRelaxed));
  int st = hipEventRecord(b->fork, b->s0) == hipSuccess ? CUBE_OK : CUBE_ERR_HIP;
  for (int i = 0; i < b->nbox && !st; ++i) {
    hipStream_t s = b->box[i]->c.stream;
    if (hipStreamWaitEvent(s, b->fork, 0) != hipSuccess) st = CUBE_ERR_HIP;
    if (!st) st = box_run(b, i, s);
    if (!st && hipEventRecord(b->join[i], s) != hipSuccess) st = CUBE_ERR_HIP;
    if (!st && hipStreamWaitEvent(b->s0, b->join[i], 0) != hipSuccess) st = CUBE_ERR_HIP;
  }
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->s0, &g);
  if (st) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess || !g) return CUBE_ERR_HIP;
  b->graph = g;
  CUBE_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->captured = 1;
  return CUBE_OK;
}

extern "C" int cube_batch_run(cube_batch *b, int32_t mode) {
  if (!b) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipSetDevice(b->device));
  if (mode == 0) {
    for (int i = 0; i < b->nbox; ++i) {
      int st = box_run(b, i, b->s0);
      if (st) return st;
    }
    return CUBE_OK;
  }
  if (mode != 1) return CUBE_ERR_INVALID;
  if (!b->captured) {
    /* a capture records launches without running them: make sure no
       earlier work of the boxes is still in flight on their own streams */
    CUBE_CHECK(hipDeviceSynchronize());
    int st = capture(b);
    if (st) return st;
  }
  CUBE_CHECK(hipGraphLaunch(b->exec, b->s0));
  return CUBE_OK;
}

extern "C" int cube_batch_sync(cube_batch *b) {
  if (!b) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipStreamSynchronize(b->s0));
  return CUBE_OK;
}

extern "C" int cube_batch_tune(cube_batch *b, int32_t *shapes) {
  if (!b) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipSetDevice(b->device));
  /* grid x block of the model's grid-stride loops: from one workgroup per
     CU to several per CU, at 2, 4, 8 and 16 waves per workgroup */
  static const int cand[][2] = {{256, 256},  {512, 256},  {1024, 256}, {2048, 256}, {4096, 256}, {8192, 256},
                                {1024, 128}, {2048, 128}, {4096, 128}, {512, 512},  {1024, 512}, {512, 1024}};
  const int nc = (int)(sizeof(cand) / sizeof(cand[0]));
  int *done = (int *)calloc(b->nbox, sizeof(int));
  if (!done) return CUBE_ERR_ALLOC;
  int st = CUBE_OK;
  for (int i = 0; i < b->nbox && !st; ++i) {
    if (done[i]) continue;
    cube_dpm *d = b->box[i];
    int best = 0;
    float best_ms = 1e30f;
    for (int q = 0; q < nc && !st; ++q) {
      d->grid = cand[q][0];
      d->block = cand[q][1];
      for (int rep = 0; rep < 2 && !st; ++rep) { /* the second run is timed */
        if (hipEventRecord(b->t0, b->s0) != hipSuccess) st = CUBE_ERR_HIP;
        if (!st) st = box_run(b, i, b->s0);
        if (!st && hipEventRecord(b->t1, b->s0) != hipSuccess) st = CUBE_ERR_HIP;
        if (!st && hipEventSynchronize(b->t1) != hipSuccess) st = CUBE_ERR_HIP;
      }
      float ms = 0.f;
      if (!st && hipEventElapsedTime(&ms, b->t0, b->t1) != hipSuccess) st = CUBE_ERR_HIP;
      if (!st && ms < best_ms) {
        best_ms = ms;
        best = q;
      }
    }
    /* every box of the same size takes the winner */
    for (int j = i; j < b->nbox; ++j)
      if (!done[j] && b->box[j]->p.ng == d->p.ng && b->box[j]->p.nf == d->p.nf &&
          b->box[j]->p.nsteps == d->p.nsteps && b->box[j]->p.checkpoint == d->p.checkpoint) {
        b->box[j]->grid = cand[best][0];
        b->box[j]->block = cand[best][1];
        done[j] = 1;
      }
  }
  free(done);
  if (st) return st;
  if (shapes)
    for (int i = 0; i < b->nbox; ++i) {
      shapes[2 * i] = b->box[i]->grid;
      shapes[2 * i + 1] = b->box[i]->block;
    }
  b->captured = 0; /* shapes are baked into a graph */
  return CUBE_OK;
}
