/*
 * cube_ctx.hip -- context lifetime, state transfer and the step entry points
 * of the C ABI (include/cube.h).
 */
#include "cube_internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <new>
#include <type_traits>

extern "C" const char *cube_strerror(int code) {
  switch (code) {
    case CUBE_OK: return "ok";
    case CUBE_ERR_TILE_OVERFLOW: return "too many particles in a tile+buffer: set tile_buffer larger";
    case CUBE_ERR_IMAGE_OVERFLOW: return "too many particles in the image+buffer: set image_buffer larger";
    case CUBE_ERR_PP_OVERFLOW: return "np_pp_max too small";
    case CUBE_ERR_CHECKSUM: return "particle checksum mismatch";
    case CUBE_ERR_COUNT: return "particle count mismatch";
    case CUBE_ERR_DISPLACEMENT: return "particle displacement larger than the buffer";
    case CUBE_ERR_HIP: return "HIP runtime error";
    case CUBE_ERR_ALLOC: return "device allocation failed";
    case CUBE_ERR_INVALID: return "invalid argument or unsupported size";
    case CUBE_ERR_STATE: return "call out of order";
  }
  return "unknown error";
}

__global__ void k_slots_zero(unsigned long long *s, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s[i] = 0ull;
}
__global__ __launch_bounds__(CUBE_SLOTS) void k_slots_fold(const unsigned long long *s, int add,
                                                          unsigned long long *d64, unsigned *d32) {
  __shared__ unsigned long long r[CUBE_SLOTS];
  const int t = threadIdx.x, reg = blockIdx.x;
  r[t] = s[reg * CUBE_SLOTS + t];
  __syncthreads();
  for (int o = CUBE_SLOTS / 2; o > 0; o >>= 1) {
    if (t < o) r[t] = add ? r[t] + r[t + o] : (r[t] > r[t + o] ? r[t] : r[t + o]);
    __syncthreads();
  }
  if (t == 0) {
    if (d64) d64[reg] = r[0];
    if (d32) d32[reg] = (unsigned)r[0];
  }
}
int cube_slots_zero(cube_ctx *c, unsigned long long *region, int nreg) {
  hipLaunchKernelGGL(k_slots_zero, dim3(1), dim3(256), 0, c->stream, region, nreg * CUBE_SLOTS);
  CUBE_LAUNCH_CHECK();
  return CUBE_OK;
}
int cube_slots_fold(cube_ctx *c, const unsigned long long *region, int nreg, int add, unsigned long long *d64,
                    unsigned *d32) {
  hipLaunchKernelGGL(k_slots_fold, dim3(nreg), dim3(CUBE_SLOTS), 0, c->stream, region, add, d64, d32);
  CUBE_LAUNCH_CHECK();
  return CUBE_OK;
}

static void free_all(cube_ctx *c) {
  if (c->scratch_in_force) /* carved out of force_f, not allocations of their own */
    c->dest = c->rank = c->xdest = c->xrank = nullptr, c->key = c->sortbuf = c->xkey = nullptr;
  void *ptrs[] = {c->xp, c->vp, c->xp2, c->vp2, c->pid, c->pid2, c->rhoc, c->rhoc2, c->vfield, c->vfield2,
                  c->cstart, c->cstart2, c->dest, c->key, c->sortbuf, c->fill, c->xdest, c->xkey, c->rank, c->xrank, c->fbflag,
                  c->tanf_tab, c->venc_tab, c->vdec0, c->vdec1, c->rho_f, c->spec_f, c->specd_f, c->force_f,
                  c->kern_f, c->rho_c, c->spec_c, c->specd_c, c->force_c, c->kern_c, c->dred,
                  c->ured, c->ired, c->scanbuf, c->vpart, c->slots, c->ppdefer, c->ppitems};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  cube_fft_destroy(c->plan_f);
  cube_fft_destroy(c->plan_c);
  for (int k = 0; k < 2 * CUBE_TEV; ++k)
    if (c->tev[k]) (void)hipEventDestroy(c->tev[k]);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_join2) (void)hipEventDestroy(c->ev_join2);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->side2) (void)hipStreamDestroy(c->side2);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

template <typename T> static int dalloc(T **p, size_t n) {
  if (n == 0) n = 1;
  if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return CUBE_ERR_ALLOC;
  }
  return CUBE_OK;
}
static int dalloc_bytes(void **p, size_t n) {
  if (hipMalloc(p, n ? n : 1) != hipSuccess) {
    *p = nullptr;
    return CUBE_ERR_ALLOC;
  }
  return CUBE_OK;
}

extern "C" int cube_ctx_create(const cube_params *p, cube_ctx **out) {
  if (!p || !out) return CUBE_ERR_INVALID;
  *out = nullptr;
  if (p->nc < 8 || p->nnt < 1 || p->nc % p->nnt || p->ncb < 1 || (p->izipx != 1 && p->izipx != 2) ||
      (p->izipv != 1 && p->izipv != 2) || p->np_nc < 1 || p->pp_range < 0 || p->pp_range > 4)
    return CUBE_ERR_INVALID;
  const int64_t nt = p->nc / p->nnt;
  if (nt < p->ncb) return CUBE_ERR_INVALID;
  /* the drift's within-cell order key keeps k_src-k_dst+2*ncb in 5 bits,
     and the coarse and fine meshes (nc, 4*nc) need built FFT sizes */
  if (p->ncb > 7 || !cube_fft_supported(p->nc) || !cube_fft_supported(4 * p->nc)) return CUBE_ERR_INVALID;
  cube_ctx *c = new (std::nothrow) cube_ctx;
  if (!c) return CUBE_ERR_ALLOC;
  memset(c, 0, sizeof(*c));
  c->p = *p;
  c->device = p->device;
  if (hipSetDevice(c->device) != hipSuccess) {
    delete c;
    return CUBE_ERR_HIP;
  }
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return CUBE_ERR_HIP;
  }
  if (hipStreamCreate(&c->side) != hipSuccess || hipStreamCreate(&c->side2) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join2, hipEventDisableTiming) != hipSuccess) {
    free_all(c);
    delete c;
    return CUBE_ERR_HIP;
  }
  c->nc = p->nc; c->nnt = p->nnt; c->nt = nt; c->ncb = p->ncb; c->ncell = 4;
  c->nf = c->nc * c->ncell; c->nh = c->nf / 2 + 1;
  c->ncells = c->nc * c->nc * c->nc;
  c->ntiles = c->nnt * c->nnt * c->nnt;
  c->nvbin = 1 << (8 * p->izipv);
  /* W/variables.f90:8-10 in real(4), as the reference evaluates them */
  c->np_image = (c->nc * p->np_nc) * (c->nc * p->np_nc) * (c->nc * p->np_nc);
  {
    const float r = (float)(nt + 2 * p->ncb) * 1.0f / (float)nt;
    const float r3 = r * (r * r);
    c->np_image_max = (int64_t)((float)c->np_image * r3 * p->image_buffer);
    c->np_tile_max = (int64_t)((float)(c->np_image / c->ntiles) * r3 * p->tile_buffer);
    c->np_pp_max = c->np_tile_max;
  }
  /* the disjoint state never holds buffer copies: the total particle count
     is conserved, so capacity only needs headroom for the rare duplicates */
  c->cap = c->np_image + c->np_image / 64 + 4096;
  c->xcap = c->np_image / 64 + 4096;
  int st = CUBE_OK;
  const size_t xb = (size_t)p->izipx * 3, vb = (size_t)p->izipv * 3;
  const int64_t nf3 = c->nf * c->nf * c->nf, nspec = (int64_t)cube_spec_pitch((int)c->nf) * c->nf * c->nf;
  const int64_t ho = c->nf / 2 + 1;
  /* Fine-mesh layout.  Full: three force spectra, the three components
     solved together.  Compact (for meshes whose full layout does not fit,
     e.g. 2048^3 in 288 GB): one force spectrum that the components pass
     through one after the other, and the drift's sort scratch (dead outside
     the drift and the initial count sort) inside the force mesh, behind
     the room the initial conditions stage their velocities in.  Both give
     the same bits. */
  const size_t scratch_b = (size_t)c->cap * (4 + 8 + 8 + 4) + (size_t)c->xcap * (8 + 4 + 8 + 4) + 4096;
  const size_t v_room = ((size_t)c->cap * 12 + 255) & ~(size_t)255;
  int compact = p->mem_layout == 2;
  if (p->mem_layout == 0) {
    size_t fr = 0, tot = 0;
    const size_t full = (size_t)c->cap * (2 * xb + 2 * vb + (p->has_pid ? 8 : 0)) + scratch_b +
                        (size_t)c->ncells * 56 + (size_t)nf3 * 16 + (size_t)nspec * 32 +
                        (size_t)cube_kernf_elems(ho) * 4 + (size_t)c->ncells * 64;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess && full > fr / 100 * 97) compact = 1;
  }
  c->fine_nspec = compact ? 1 : 3;
  st |= dalloc_bytes(&c->xp, c->cap * xb);
  st |= dalloc_bytes(&c->vp, c->cap * vb);
  st |= dalloc_bytes(&c->xp2, c->cap * xb);
  st |= dalloc_bytes(&c->vp2, c->cap * vb);
  if (p->has_pid) {
    st |= dalloc(&c->pid, c->cap);
    st |= dalloc(&c->pid2, c->cap);
  }
  st |= dalloc(&c->rhoc, c->ncells);
  st |= dalloc(&c->rhoc2, c->ncells);
  st |= dalloc(&c->vfield, 3 * c->ncells);
  st |= dalloc(&c->vfield2, 3 * c->ncells);
  st |= dalloc(&c->cstart, c->ncells + 1);
  st |= dalloc(&c->cstart2, c->ncells + 1);
  st |= dalloc(&c->fill, c->ncells);
  st |= dalloc(&c->ppdefer, c->ncells / 64 + 1);
  st |= dalloc(&c->ppitems, c->ncells / 64 + 2);
  st |= dalloc(&c->fbflag, c->ncells / 64 + 2); /* one per 64-cell finalize workgroup of the drift */
  st |= dalloc(&c->tanf_tab, c->nvbin);
  st |= dalloc(&c->venc_tab, c->nvbin / 2);
  st |= dalloc(&c->vdec0, c->nvbin);
  st |= dalloc(&c->vdec1, c->nvbin);
  st |= dalloc(&c->rho_f, nf3);
  st |= dalloc(&c->spec_f, nspec);
  st |= dalloc(&c->specd_f, c->fine_nspec * nspec);
  st |= dalloc(&c->force_f, 3 * nf3);
  c->scratch_in_force = compact && !st && v_room + scratch_b <= (size_t)nf3 * 12;
  if (c->scratch_in_force) {
    char *b = reinterpret_cast<char *>(c->force_f) + v_room;
    auto carve = [&b](auto **ptr, size_t n) {
      *ptr = reinterpret_cast<std::remove_reference_t<decltype(**ptr)> *>(b);
      b += (n * sizeof(**ptr) + 255) & ~(size_t)255;
    };
    carve(&c->key, c->cap);
    carve(&c->sortbuf, c->cap + c->xcap);
    carve(&c->xkey, c->xcap);
    carve(&c->dest, c->cap);
    carve(&c->rank, c->cap);
    carve(&c->xdest, c->xcap);
    carve(&c->xrank, c->xcap);
  } else {
    st |= dalloc(&c->dest, c->cap);
    st |= dalloc(&c->key, c->cap);
    st |= dalloc(&c->sortbuf, c->cap + c->xcap);
    st |= dalloc(&c->xdest, c->xcap);
    st |= dalloc(&c->xkey, c->xcap);
    st |= dalloc(&c->rank, c->cap);
    st |= dalloc(&c->xrank, c->xcap);
  }
  st |= dalloc(&c->kern_f, cube_kernf_elems(ho));
  const int64_t nc3 = c->ncells, nspec_c = (int64_t)cube_spec_pitch((int)c->nc) * c->nc * c->nc;
  st |= dalloc(&c->rho_c, nc3);
  st |= dalloc(&c->spec_c, nspec_c);
  st |= dalloc(&c->specd_c, 3 * nspec_c);
  st |= dalloc(&c->force_c, 3 * nc3);
  st |= dalloc(&c->kern_c, 3 * (c->nc / 2 + 1) * c->nc * c->nc); /* unpadded [d][kz][ky][kx] */
  st |= dalloc(&c->dred, 8192);
  st |= dalloc(&c->ured, 64);
  st |= dalloc(&c->slots, CUBE_NSLOTREG * CUBE_SLOTS);
  st |= dalloc(&c->ired, 4096);
  st |= dalloc(&c->scanbuf, c->ncells / 1024 + 4);
  st |= dalloc(&c->vpart, 3 * (c->ncells / 64 + 2));
  if (st) {
    free_all(c);
    delete c;
    return CUBE_ERR_ALLOC;
  }
  if ((st = cube_fft_create(c, (int)c->nf, &c->plan_f)) || (st = cube_fft_create(c, (int)c->nc, &c->plan_c)) ||
      (st = cube_power_scratch_setup(c->device))) {
    free_all(c);
    delete c;
    return st;
  }
  /* tan table: tan(pi*real(vp)/real(nvbin-1)) in real(4), computed with the
     host libm tanf exactly as the reference's compiled Fortran calls it
     (W/update_particle.f90:67) */
  {
    const int nv = c->nvbin;
    float *h = (float *)malloc(nv * sizeof(float));
    const float nvm1 = (float)(nv - 1);
    for (int i = 0; i < nv; ++i) {
      const int code = i - nv / 2;
      h[i] = tanf((CUBE_PI_F * (float)code) / nvm1);
    }
    hipError_t e = hipMemcpy(c->tanf_tab, h, nv * sizeof(float), hipMemcpyHostToDevice);
    free(h);
    /* code boundaries of the encoder: tan((k+0.5)*pi/(nvbin-1)), k < nv/2-1,
       in long double (the last one, beyond pi/2, is +inf) */
    const int kmax = nv / 2 - 1;
    double *tb = (double *)malloc((size_t)(kmax + 1) * sizeof(double));
    for (int k = 0; k < kmax; ++k)
      tb[k] = (double)tanl(((long double)k + 0.5L) * (long double)CUBE_PI_F / (long double)(nv - 1));
    tb[kmax] = HUGE_VAL;
    if (e == hipSuccess) e = hipMemcpy(c->venc_tab, tb, (size_t)(kmax + 1) * sizeof(double), hipMemcpyHostToDevice);
    free(tb);
    if (e != hipSuccess) {
      free_all(c);
      delete c;
      return CUBE_ERR_HIP;
    }
  }
  c->pp_ftemp = 0.0f;
  *out = c;
  return CUBE_OK;
}

extern "C" void cube_ctx_destroy(cube_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
}

extern "C" int cube_ctx_info(const cube_ctx *c, int64_t *o) {
  if (!c || !o) return CUBE_ERR_INVALID;
  o[0] = c->nt; o[1] = c->nf; o[2] = c->nh; o[3] = c->np_image; o[4] = c->np_image_max;
  o[5] = c->np_tile_max; o[6] = c->np_pp_max; o[7] = c->cap; o[8] = c->nplocal; o[9] = c->npglobal;
  o[10] = c->has_kernels; o[11] = c->has_state; o[12] = c->nc; o[13] = c->nnt; o[14] = c->ncb;
  o[15] = c->nvbin;
  return CUBE_OK;
}

extern "C" int cube_sync(cube_ctx *c) {
  if (!c) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipStreamSynchronize(c->stream));
  CUBE_CHECK(hipGetLastError());
  return CUBE_OK;
}

extern "C" double cube_host_atan(double x) { return cube_atan(x); }

extern "C" int cube_profile(cube_ctx *c, int on) {
  if (!c) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipSetDevice(c->device));
  if (on && !c->tev[0])
    for (int k = 0; k < 2 * CUBE_TEV; ++k) CUBE_CHECK(hipEventCreate(&c->tev[k]));
  if (!on) cube_timers_resolve(c);
  c->profiling = on ? 1 : 0;
  return CUBE_OK;
}

extern "C" int cube_timers(cube_ctx *c, double *ms, int64_t *calls, int reset) {
  if (!c) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipSetDevice(c->device));
  cube_timers_resolve(c);
  for (int i = 0; i < CUBE_NTIMERS; ++i) {
    if (ms) ms[i] = c->tms[i];
    if (calls) calls[i] = c->tcalls[i];
    if (reset) {
      c->tms[i] = 0;
      c->tcalls[i] = 0;
    }
  }
  return CUBE_OK;
}

/* ---- state transfer ---------------------------------------------------- */
extern "C" int cube_set_state(cube_ctx *c, int64_t nplocal, int64_t npglobal, const void *xp, const void *vp,
                              const int32_t *pid, const int32_t *rhoc, const float *vfield) {
  if (c) c->pm_ahead = 0; /* forces formed ahead by ext_pp_force no longer apply */
  if (!c || !xp || !vp || !rhoc || !vfield || nplocal < 0) return CUBE_ERR_INVALID;
  if (nplocal > c->cap) return CUBE_ERR_IMAGE_OVERFLOW;
  CUBE_CHECK(hipSetDevice(c->device));
  int64_t s = 0;
  for (int64_t q = 0; q < c->ncells; ++q) s += rhoc[q];
  if (s != nplocal) return CUBE_ERR_COUNT;
  CUBE_CHECK(hipMemcpy(c->xp, xp, nplocal * 3 * c->p.izipx, hipMemcpyHostToDevice));
  CUBE_CHECK(hipMemcpy(c->vp, vp, nplocal * 3 * c->p.izipv, hipMemcpyHostToDevice));
  if (c->p.has_pid) {
    if (pid) {
      CUBE_CHECK(hipMemcpy(c->pid, pid, nplocal * sizeof(int32_t), hipMemcpyHostToDevice));
    } else {
      CUBE_CHECK(hipMemset(c->pid, 0, nplocal * sizeof(int32_t)));
    }
  }
  CUBE_CHECK(hipMemcpy(c->rhoc, rhoc, c->ncells * sizeof(int32_t), hipMemcpyHostToDevice));
  CUBE_CHECK(hipMemcpy(c->vfield, vfield, 3 * c->ncells * sizeof(float), hipMemcpyHostToDevice));
  int64_t tot = 0;
  int st = cube_scan_counts(c, c->rhoc, c->cstart, c->ncells, &tot);
  if (st) return st;
  c->nplocal = nplocal;
  c->npglobal = npglobal;
  c->has_state = 1;
  return CUBE_OK;
}

extern "C" int cube_get_state(cube_ctx *c, int64_t *nplocal, void *xp, void *vp, int32_t *pid, int32_t *rhoc,
                              float *vfield) {
  if (!c) return CUBE_ERR_INVALID;
  if (!c->has_state) return CUBE_ERR_STATE;
  CUBE_CHECK(hipSetDevice(c->device));
  CUBE_CHECK(hipStreamSynchronize(c->stream));
  const int64_t n = c->nplocal;
  if (nplocal) *nplocal = n;
  if (xp) CUBE_CHECK(hipMemcpy(xp, c->xp, n * 3 * c->p.izipx, hipMemcpyDeviceToHost));
  if (vp) CUBE_CHECK(hipMemcpy(vp, c->vp, n * 3 * c->p.izipv, hipMemcpyDeviceToHost));
  if (pid && c->p.has_pid) CUBE_CHECK(hipMemcpy(pid, c->pid, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (rhoc) CUBE_CHECK(hipMemcpy(rhoc, c->rhoc, c->ncells * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (vfield) CUBE_CHECK(hipMemcpy(vfield, c->vfield, 3 * c->ncells * sizeof(float), hipMemcpyDeviceToHost));
  return CUBE_OK;
}

/* ---- kernels ----------------------------------------------------------- */
extern "C" int cube_init_kernels(cube_ctx *c, const float *wfxyzf, const float *wfxyzc) {
  if (c) c->pm_ahead = 0; /* forces formed ahead by ext_pp_force no longer apply */
  if (!c || !wfxyzf || !wfxyzc) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipSetDevice(c->device));
  int st = cube_green_init(c, wfxyzf, wfxyzc);
  if (st) return st;
  CUBE_CHECK(hipStreamSynchronize(c->stream));
  c->has_kernels = 1;
  return CUBE_OK;
}

/* ---- the step ------------------------------------------------------------ */
extern "C" int cube_update_x(cube_ctx *c, float dt_old, float dt, float vz_max, float sigma_vi,
                             cube_vstats *out) {
  if (c) c->pm_ahead = 0; /* forces formed ahead by ext_pp_force no longer apply */
  if (!c) return CUBE_ERR_INVALID;
  if (!c->has_state) return CUBE_ERR_STATE;
  CUBE_CHECK(hipSetDevice(c->device));
  return cube_drift(c, dt_old, dt, vz_max, sigma_vi, out);
}

/* buffer_grid: the buffered copies are implicit (a buffer cell is read from
   its periodic neighbour), so what remains of W/buffer_grid.f90 is the
   image-overflow check of redistribute_cdm (:137-154): the particles of all
   tiles' extended regions must fit np_image_max. */
extern "C" int cube_buffer_grid(cube_ctx *c) {
  if (!c) return CUBE_ERR_INVALID;
  if (!c->has_state) return CUBE_ERR_STATE;
  CUBE_CHECK(hipSetDevice(c->device));
  int64_t *sums = (int64_t *)malloc(c->ntiles * sizeof(int64_t));
  tbeg(c);
  int st = cube_tile_boxsum(c, c->rhoc, (int)c->ncb, sums);
  tend(c, CUBE_T_CHECKS);
  int64_t tot = 0;
  for (int64_t t = 0; t < c->ntiles; ++t) tot += sums[t];
  free(sums);
  if (st) return st;
  if ((double)tot / (double)c->np_image_max > 1.0) return CUBE_ERR_IMAGE_OVERFLOW;
  return CUBE_OK;
}
extern "C" int cube_buffer_x(cube_ctx *c) {
  if (!c) return CUBE_ERR_INVALID;
  return c->has_state ? CUBE_OK : CUBE_ERR_STATE;
}
extern "C" int cube_buffer_v(cube_ctx *c) {
  if (!c) return CUBE_ERR_INVALID;
  return c->has_state ? CUBE_OK : CUBE_ERR_STATE;
}

/* The fine deposit and force on stream fs and the coarse deposit and force
   on stream cs, both after the work already queued on the context stream;
   ev_join and ev_join2 mark their ends.  Nothing is read back: the mesh sums
   stay in dred and max |F|^2 in ured words 0 (fine) and 1 (coarse).  The
   two pipelines read only the positions and write disjoint buffers. */
static int pm_forces(cube_ctx *c, float mass_p, float a_mid, float dt, hipStream_t fs, hipStream_t cs) {
  const hipStream_t main = c->stream;
  CUBE_CHECK(hipEventRecord(c->ev_fork, main));
  int st = CUBE_OK;
  c->stream = fs;
  if (fs != main && hipStreamWaitEvent(fs, c->ev_fork, 0) != hipSuccess) st = CUBE_ERR_HIP;
  if (!st) {
    tbeg(c);
    st = cube_fine_deposit(c, mass_p, nullptr);
    tend(c, CUBE_T_FINE_DEPOSIT);
  }
  if (!st) {
    /* force_f gets the fine kick's per-term factors for this a_mid and dt */
    const float pre[2] = {a_mid, dt};
    tbeg(c);
    st = cube_fft_r2c_3d(c, c->plan_f, c->rho_f, c->spec_f);
    tend(c, CUBE_T_FINE_FFT);
    if (!st)
      st = cube_fft_green_c2r_3d(c, c->plan_f, c->spec_f, c->kern_f, 1, c->specd_f, c->force_f, (float)c->nf,
                                 nullptr, CUBE_T_FINE_GREEN, (unsigned *)c->ured,
                                 c->slots + CUBE_SLOT_F2_FINE * CUBE_SLOTS, pre, c->fine_nspec);
  }
  if (!st && hipEventRecord(c->ev_join, fs) != hipSuccess) st = CUBE_ERR_HIP;
  c->stream = cs;
  if (!st && hipStreamWaitEvent(cs, c->ev_fork, 0) != hipSuccess) st = CUBE_ERR_HIP;
  if (!st) {
    tbeg(c);
    st = cube_coarse_deposit(c, mass_p, nullptr);
    tend(c, CUBE_T_COARSE_DEPOSIT);
  }
  if (!st) {
    tbeg(c);
    st = cube_fft_r2c_3d(c, c->plan_c, c->rho_c, c->spec_c);
    if (!st)
      st = cube_fft_green_c2r_3d(c, c->plan_c, c->spec_c, c->kern_c, 0, c->specd_c, c->force_c, (float)c->nc,
                                 nullptr, -1, (unsigned *)c->ured + 1, c->slots + CUBE_SLOT_F2_COARSE * CUBE_SLOTS);
    tend(c, CUBE_T_COARSE_FORCE);
  }
  if (!st && hipEventRecord(c->ev_join2, cs) != hipSuccess) st = CUBE_ERR_HIP;
  c->stream = main;
  return st;
}

extern "C" int cube_pp_force(cube_ctx *c, float a_mid, float dt, float sigma_vi, float mass_p, float *f2_max_pp) {
  if (!c) return CUBE_ERR_INVALID;
  if (!c->has_state) return CUBE_ERR_STATE;
  CUBE_CHECK(hipSetDevice(c->device));
  /* W/ext_pp_force.f90:41-45: particles of coarse cells 0..nt+1 of a tile */
  int64_t *sums = (int64_t *)malloc(c->ntiles * sizeof(int64_t));
  tbeg(c);
  int st = cube_tile_boxsum(c, c->rhoc, 1, sums);
  tend(c, CUBE_T_CHECKS);
  int64_t mx = 0;
  for (int64_t t = 0; t < c->ntiles; ++t) mx = sums[t] > mx ? sums[t] : mx;
  free(sums);
  if (st) return st;
  if (mx > c->np_pp_max) return CUBE_ERR_PP_OVERFLOW;
  /* The PM forces of this step depend only on the positions (final after
     update_x) and on mass_p, not on the velocities PP updates: they are
     queued on the side streams first and run beside PP.  The call still
     returns with all its work done; particle_mesh uses them when nothing
     has changed the state in between and mass_p agrees. */
  c->pm_ahead = 0;
  if (c->has_kernels) {
    st = pm_forces(c, mass_p, a_mid, dt, c->side, c->side2);
    if (st) return st;
  }
  int64_t itest1 = 0;
  tbeg(c);
  st = cube_pp(c, a_mid, dt, sigma_vi, mass_p, f2_max_pp, &itest1);
  tend(c, CUBE_T_PP);
  if (c->has_kernels) {
    CUBE_CHECK(hipEventSynchronize(c->ev_join));
    CUBE_CHECK(hipEventSynchronize(c->ev_join2));
  }
  if (st) return st;
  if (itest1 != c->nplocal) return CUBE_ERR_COUNT;
  if (c->has_kernels) {
    c->pm_ahead = 1;
    c->pm_mass = mass_p;
    c->pm_amid = a_mid;
    c->pm_dt = dt;
  }
  return CUBE_OK;
}

extern "C" int cube_pm_fine_deposit(cube_ctx *c, float mass_p, double *testrhof) {
  if (c) c->pm_ahead = 0; /* forces formed ahead by ext_pp_force no longer apply */
  if (!c) return CUBE_ERR_INVALID;
  if (!c->has_state) return CUBE_ERR_STATE;
  CUBE_CHECK(hipSetDevice(c->device));
  tbeg(c);
  const int st = cube_fine_deposit(c, mass_p, testrhof);
  tend(c, CUBE_T_FINE_DEPOSIT);
  return st;
}
extern "C" int cube_pm_fine_force(cube_ctx *c, float *f2) {
  if (c) c->pm_ahead = 0; /* forces formed ahead by ext_pp_force no longer apply */
  if (!c) return CUBE_ERR_INVALID;
  if (!c->has_kernels) return CUBE_ERR_STATE;
  CUBE_CHECK(hipSetDevice(c->device));
  tbeg(c);
  int st = cube_fft_r2c_3d(c, c->plan_f, c->rho_f, c->spec_f);
  if (st) return st;
  tend(c, CUBE_T_FINE_FFT);
  return cube_fft_green_c2r_3d(c, c->plan_f, c->spec_f, c->kern_f, 1, c->specd_f, c->force_f, (float)c->nf, f2,
                               CUBE_T_FINE_GREEN, (unsigned *)c->ured,
                               c->slots + CUBE_SLOT_F2_FINE * CUBE_SLOTS, nullptr, c->fine_nspec);
}
extern "C" int cube_pm_fine_potential(cube_ctx *c, const float *wg1) {
  if (c) c->pm_ahead = 0; /* the force mesh is overwritten */
  if (!c || !wg1) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipSetDevice(c->device));
  return cube_fine_potential(c, wg1);
}
extern "C" int cube_pm_fine_kick(cube_ctx *c, float a_mid, float dt, float sigma_vi, float sigma_vi_new) {
  if (c) c->pm_ahead = 0; /* forces formed ahead by ext_pp_force no longer apply */
  if (!c) return CUBE_ERR_INVALID;
  if (!c->has_state) return CUBE_ERR_STATE;
  CUBE_CHECK(hipSetDevice(c->device));
  tbeg(c);
  const int st = cube_fine_kick(c, a_mid, dt, sigma_vi, sigma_vi_new, 1);
  tend(c, CUBE_T_FINE_KICK);
  return st;
}
extern "C" int cube_pm_coarse_deposit(cube_ctx *c, float mass_p, double *testrhoc) {
  if (c) c->pm_ahead = 0; /* forces formed ahead by ext_pp_force no longer apply */
  if (!c) return CUBE_ERR_INVALID;
  if (!c->has_state) return CUBE_ERR_STATE;
  CUBE_CHECK(hipSetDevice(c->device));
  tbeg(c);
  const int st = cube_coarse_deposit(c, mass_p, testrhoc);
  tend(c, CUBE_T_COARSE_DEPOSIT);
  return st;
}
extern "C" int cube_pm_coarse_force(cube_ctx *c, float *f2) {
  if (c) c->pm_ahead = 0; /* forces formed ahead by ext_pp_force no longer apply */
  if (!c) return CUBE_ERR_INVALID;
  if (!c->has_kernels) return CUBE_ERR_STATE;
  CUBE_CHECK(hipSetDevice(c->device));
  tbeg(c);
  int st = cube_fft_r2c_3d(c, c->plan_c, c->rho_c, c->spec_c);
  if (st) return st;
  st = cube_fft_green_c2r_3d(c, c->plan_c, c->spec_c, c->kern_c, 0, c->specd_c, c->force_c, (float)c->nc, f2, -1,
                             (unsigned *)c->ured + 1, c->slots + CUBE_SLOT_F2_COARSE * CUBE_SLOTS);
  tend(c, CUBE_T_COARSE_FORCE);
  return st;
}
extern "C" int cube_pm_coarse_kick(cube_ctx *c, float a_mid, float dt, float sigma_vi, float *vmax) {
  if (c) c->pm_ahead = 0; /* forces formed ahead by ext_pp_force no longer apply */
  if (!c) return CUBE_ERR_INVALID;
  if (!c->has_state) return CUBE_ERR_STATE;
  CUBE_CHECK(hipSetDevice(c->device));
  tbeg(c);
  const int st = cube_coarse_kick(c, a_mid, dt, sigma_vi, vmax);
  tend(c, CUBE_T_COARSE_KICK);
  return st;
}

/* particle_mesh, W/pm.f90:2-415 */
extern "C" int cube_particle_mesh(cube_ctx *c, float a_mid, float dt, float sigma_vi, float sigma_vi_new,
                                  float mass_p, cube_pm_out *out) {
  if (!c) return CUBE_ERR_INVALID;
  if (!c->has_state || !c->has_kernels) return CUBE_ERR_STATE;
  cube_pm_out o;
  memset(&o, 0, sizeof(o));
  int st;
  CUBE_CHECK(hipSetDevice(c->device));
  /* The coarse deposit and force read only the positions, so they run on a
     side stream beside the fine deposit and force (unless ext_pp_force
     already formed both beside PP); the two kicks, which both rewrite the
     velocities, stay in the reference's order on the main stream
     (W/pm.f90:2-415).  Scalars are read back at the end. */
  const int ahead = c->pm_ahead && c->pm_mass == mass_p && c->pm_amid == a_mid && c->pm_dt == dt;
  c->pm_ahead = 0;
  if (!ahead && (st = pm_forces(c, mass_p, a_mid, dt, c->stream, c->side))) return st;
  CUBE_CHECK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  CUBE_CHECK(hipStreamWaitEvent(c->stream, c->ev_join2, 0));
  /* the fine kick, sigma_vi = sigma_vi_new (W/pm.f90:193), the coarse kick:
     one pass over the particles, timed as the fine kick */
  tbeg(c);
  st = cube_pm_kicks(c, a_mid, dt, sigma_vi, sigma_vi_new, o.vmax); /* synchronises */
  tend(c, CUBE_T_FINE_KICK);
  if (st) return st;
  {
    double sums[2];
    unsigned f2[2];
    CUBE_CHECK(hipMemcpyAsync(&sums[0], c->dred + CUBE_DRED_FINE_SUM, sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    CUBE_CHECK(hipMemcpyAsync(&sums[1], c->dred + CUBE_DRED_COARSE_SUM, sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    CUBE_CHECK(hipMemcpyAsync(f2, c->ured, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    CUBE_CHECK(hipStreamSynchronize(c->stream));
    o.testrhof = sums[0];
    o.testrhoc = sums[1];
    memcpy(&o.f2_max_fine, &f2[0], sizeof(float));
    memcpy(&o.f2_max_coarse, &f2[1], sizeof(float));
  }
  /* W/pm.f90:395-398 in real(4) */
  const float GG = 1.0f / 6.0f / CUBE_PI_F;
  o.dt_fine = sqrtf(1.0f / (sqrtf(o.f2_max_fine) * a_mid * GG));
  o.dt_coarse = sqrtf((float)c->ncell / (sqrtf(o.f2_max_coarse) * a_mid * GG));
  float vm = o.vmax[0] > o.vmax[1] ? o.vmax[0] : o.vmax[1];
  vm = vm > o.vmax[2] ? vm : o.vmax[2];
  o.dt_vmax = 0.9f * 20.0f / vm;
  o.vz_max = o.vmax[2];
  if (out) *out = o;
  return CUBE_OK;
}

/* ---- fields ---------------------------------------------------------------- */
static int field_ptr(cube_ctx *c, int which, void **ptr, size_t *bytes) {
  const int64_t nf3 = c->nf * c->nf * c->nf, ho = c->nf / 2 + 1;
  switch (which) {
    case CUBE_FIELD_RHO_F: *ptr = c->rho_f; *bytes = nf3 * 4; return CUBE_OK;
    case CUBE_FIELD_FORCE_F: *ptr = c->force_f; *bytes = 3 * nf3 * 4; return CUBE_OK;
    case CUBE_FIELD_RHO_C: *ptr = c->rho_c; *bytes = c->ncells * 4; return CUBE_OK;
    case CUBE_FIELD_FORCE_C: *ptr = c->force_c; *bytes = 3 * c->ncells * 4; return CUBE_OK;
    case CUBE_FIELD_KERN_F: *ptr = c->kern_f; *bytes = 3 * ho * ho * ho * 4; return CUBE_OK;
    case CUBE_FIELD_KERN_C: *ptr = c->kern_c; *bytes = 3 * (c->nc / 2 + 1) * c->nc * c->nc * 4; return CUBE_OK;
    case CUBE_FIELD_PHI_F: *ptr = c->force_f; *bytes = nf3 * 4; return CUBE_OK;
  }
  return CUBE_ERR_INVALID;
}
/* kern_f crosses the boundary as the plain octant [d][kz][ky][kx] of
   (nf/2+1)^3 per component; on the device it is blocked (cube_kernf_index) */
static int kernf_transfer(cube_ctx *c, float *out, const float *in) {
  const int64_t ho = c->nf / 2 + 1, ne = cube_kernf_elems(ho);
  float *tmp = new (std::nothrow) float[ne];
  if (!tmp) return CUBE_ERR_ALLOC;
  int st = CUBE_OK;
  if (out) {
    if (hipMemcpy(tmp, c->kern_f, ne * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) st = CUBE_ERR_HIP;
  } else {
    memset(tmp, 0, ne * sizeof(float));
  }
  for (int d = 0; d < 3 && !st; ++d)
    for (int kz = 0; kz < ho; ++kz)
      for (int ky = 0; ky < ho; ++ky) {
        const int64_t r = ((d * ho + kz) * ho + ky) * ho;
        for (int kx = 0; kx < ho; ++kx) {
          const int64_t b = cube_kernf_index(ho, d, kz, ky, kx);
          if (out) out[r + kx] = tmp[b]; else tmp[b] = in[r + kx];
        }
      }
  if (!st && !out && hipMemcpy(c->kern_f, tmp, ne * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    st = CUBE_ERR_HIP;
  delete[] tmp;
  return st;
}
extern "C" int cube_get_field(cube_ctx *c, int which, void *host, size_t bytes) {
  if (!c || !host) return CUBE_ERR_INVALID;
  void *p;
  size_t n;
  int st = field_ptr(c, which, &p, &n);
  if (st) return st;
  if (bytes != n) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipSetDevice(c->device));
  CUBE_CHECK(hipStreamSynchronize(c->stream));
  if (which == CUBE_FIELD_KERN_F) return kernf_transfer(c, static_cast<float *>(host), nullptr);
  CUBE_CHECK(hipMemcpy(host, p, n, hipMemcpyDeviceToHost));
  return CUBE_OK;
}
extern "C" int cube_set_field(cube_ctx *c, int which, const void *host, size_t bytes) {
  if (c) c->pm_ahead = 0; /* forces formed ahead by ext_pp_force no longer apply */
  if (!c || !host) return CUBE_ERR_INVALID;
  void *p;
  size_t n;
  int st = field_ptr(c, which, &p, &n);
  if (st) return st;
  if (bytes != n || which == CUBE_FIELD_PHI_F) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipSetDevice(c->device));
  CUBE_CHECK(hipStreamSynchronize(c->stream));
  if (which == CUBE_FIELD_KERN_F) {
    if ((st = kernf_transfer(c, nullptr, static_cast<const float *>(host)))) return st;
  } else {
    CUBE_CHECK(hipMemcpy(p, host, n, hipMemcpyHostToDevice));
  }
  if (which == CUBE_FIELD_KERN_F || which == CUBE_FIELD_KERN_C) c->has_kernels = 1;
  return CUBE_OK;
}
