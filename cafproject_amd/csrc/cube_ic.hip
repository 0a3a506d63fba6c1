/*
 * cube_ic.hip -- Zel'dovich initial conditions and the count sort into the
 * disjoint checkpoint state, on the GPU: the counterpart of
 * CUBEnu/work/utilities/initial_conditions.f90 (WU/ below).
 *
 *  noise    : unit-variance Gaussian white noise on the particle lattice
 *             (ng = np_nc*nc per dim) from a counter-based generator
 *             (splitmix64 of seed and lattice index, Box-Muller) -- this
 *             project's own seeded stream; the reference's random_number
 *             stream is compiler-specific (WU/:282-313)
 *  filter   : delta_k = w_k sqrt(P(k)/cell^3), phi_k = -4 pi delta_k/k~^2
 *             with k~ the finite-difference wave number, DC = 0
 *             (WU/:329-349, :433-447)
 *  particles: x = q - grad(phi)/(8 pi ncell) scaled to the lattice spacing,
 *             v = -grad(phi)/(8 pi) vfactor(a), central differences
 *             (WU/:719-724, :754-757)
 *  sort     : cell = ceiling(x), cells in the disjoint order, particles in
 *             lattice order inside a cell; vfield = cell mean accumulated
 *             in real(4) (WU/:732-734); xp = floor(x/x_resolution) wrapped,
 *             vp = enc(v - vfield), pid from the fine-grid lattice index
 *             (WU/:752-762)
 */
#include "cube_internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

__device__ static inline uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
/* uniform in (0,1) with 53 bits */
__device__ static inline double u01(uint64_t h) { return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

__global__ void k_ic_noise(float *w, int64_t n, uint64_t seed) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t h1 = splitmix64(seed ^ splitmix64(2 * (uint64_t)i));
    const uint64_t h2 = splitmix64(seed ^ splitmix64(2 * (uint64_t)i + 1));
    const double r = sqrt(-2.0 * log(u01(h1)));
    w[i] = (float)(r * cos(6.283185307179586 * u01(h2)));
  }
}

/* spec: [kz][ky][kx] with kx in [0, ng/2]; lnk/lnp: P(k) table, k in h/Mpc */
__global__ void k_ic_filter(float2 *spec, int ng, double cell, const double *tab, int ntab, double lk0,
                            double dlk, double norm) {
  const int nh = ng / 2 + 1;
  const int64_t tot = (int64_t)nh * ng * ng;
  const double twopi = 6.283185307179586;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
    const int ix = (int)(q % nh), iy = (int)((q / nh) % ng), iz = (int)(q / ((int64_t)nh * ng));
    const int m[3] = {ix, iy <= ng / 2 ? iy : iy - ng, iz <= ng / 2 ? iz : iz - ng};
    double k2 = 0, kt2 = 0;
    for (int a = 0; a < 3; ++a) {
      const double kr = twopi * m[a] / ng; /* radians per lattice cell */
      k2 += (kr / cell) * (kr / cell);
      const double s = 2.0 * sin(0.5 * kr);
      kt2 += s * s;
    }
    const int64_t sq = ((int64_t)iz * ng + iy) * cube_spec_pitch(ng) + ix;
    float2 v = spec[sq];
    if (q == 0 || kt2 == 0.0) {
      v = make_float2(0.f, 0.f);
    } else {
      double lk = 0.5 * log(k2);
      double t = (lk - lk0) / dlk;
      if (t < 0) t = 0;
      if (t > ntab - 1.000001) t = ntab - 1.000001;
      const int i0 = (int)t;
      const double f = t - i0;
      const double lp = tab[i0] + (tab[i0 + 1] - tab[i0]) * f;
      const double amp = exp(0.5 * lp) / (cell * sqrt(cell)) * (-4.0 * 3.141592653589793 / kt2) * norm;
      v = make_float2((float)(v.x * amp), (float)(v.y * amp));
    }
    spec[sq] = v;
  }
}

/* lattice point -> position (global coarse units, double) and velocity */
__global__ void k_ic_particles(const float *phi, int ng, int np_nc, int nc, double psi_scale, double vf,
                               double *x, float *v) {
  const int64_t tot = (int64_t)ng * ng * ng;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < tot; p += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(p % ng), j = (int)((p / ng) % ng), k = (int)(p / ((int64_t)ng * ng));
    const int c[3] = {i, j, k};
    for (int a = 0; a < 3; ++a) {
      int cp[3] = {i, j, k}, cm[3] = {i, j, k};
      cp[a] = (c[a] + 1) % ng;
      cm[a] = (c[a] - 1 + ng) % ng;
      const float g = phi[((int64_t)cp[2] * ng + cp[1]) * ng + cp[0]] - phi[((int64_t)cm[2] * ng + cm[1]) * ng + cm[0]];
      const double psi = (double)g * psi_scale; /* fine cells */
      double xq = ((double)c[a] + 0.25) / np_nc + psi / 4.0;
      xq = fmod(xq, (double)nc);
      if (xq < 0) xq += nc;
      x[3 * p + a] = xq;
      v[3 * p + a] = (float)(psi * vf);
    }
  }
}

/* ---- count sort of continuous particles into the disjoint state ---- */
__global__ void k_xv_dest(const double *x, int64_t n, Geo g, int32_t *dest, uint64_t *key, int32_t *cnt) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    int gc[3];
    for (int a = 0; a < 3; ++a) {
      int c = (int)ceil(x[3 * p + a]) - 1; /* cell = ceiling(x), 1-based */
      gc[a] = wrapi(c, g.nc);
    }
    const int32_t d = (int32_t)cid_global(g, gc[0], gc[1], gc[2]);
    dest[p] = d;
    key[p] = (uint64_t)p;
    atomicAdd(cnt + d, 1);
  }
}

/* vfield = sum v / max(1,n), accumulated in real(4) in lattice order */
__global__ void k_xv_vfield(const float *v, const int32_t *cnt, const int64_t *start, const uint64_t *keys,
                            int64_t ncells, float *vfield) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int n = cnt[c];
  const int64_t s0 = start[c];
  float acc[3] = {0.f, 0.f, 0.f};
  for (int q = 0; q < n; ++q) {
    const int64_t p = (int64_t)(keys[s0 + q] & KEY_PMASK);
    for (int a = 0; a < 3; ++a) acc[a] = (float)((double)acc[a] + (double)v[3 * p + a]);
  }
  const float dn = (float)(n > 0 ? n : 1);
  for (int a = 0; a < 3; ++a) vfield[3 * c + a] = acc[a] / dn;
}

__global__ __launch_bounds__(256) void k_xv_resvar(const float *v, const int32_t *cnt, const int64_t *start,
                                                   const uint64_t *keys, const float *vfield, int64_t ncells,
                                                   double *part) {
  double s = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncells; c += (int64_t)gridDim.x * blockDim.x) {
    const int n = cnt[c];
    const int64_t s0 = start[c];
    for (int q = 0; q < n; ++q) {
      const int64_t p = (int64_t)(keys[s0 + q] & KEY_PMASK);
      for (int a = 0; a < 3; ++a) {
        const double r = (double)v[3 * p + a] - (double)vfield[3 * c + a];
        s += r * r;
      }
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

template <int IX, int IV>
__global__ void k_xv_encode(const double *x, const float *v, const int32_t *cnt, const int64_t *start,
                            const uint64_t *keys, const float *vfield, int64_t ncells, double venc, int ng,
                            int nf, int fine_per_lattice, typename IntOf<IX>::T *xp, typename IntOf<IV>::T *vp,
                            int32_t *pid) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int n = cnt[c];
  const int64_t s0 = start[c];
  const double res = IX == 1 ? 256.0 : 65536.0;
  for (int q = 0; q < n; ++q) {
    const int64_t p = (int64_t)(keys[s0 + q] & KEY_PMASK);
    for (int a = 0; a < 3; ++a) {
      xp[3 * (s0 + q) + a] = (typename IntOf<IX>::T)cube_wrap(IX, (int64_t)floor(x[3 * p + a] * res));
      vp[3 * (s0 + q) + a] =
          (typename IntOf<IV>::T)cube_encode_v(IV, venc, (double)v[3 * p + a] - (double)vfield[3 * c + a]);
    }
    if (pid && ng <= 0) {
      pid[s0 + q] = (int32_t)(p + 1); /* host-given particles: input index */
    } else if (pid) {
      /* pid = iq(1)+nf*iq(2)+nf^2*iq(3)+1 of the fine-grid lattice index (WU/:759-761) */
      const int64_t i = p % ng, j = (p / ng) % ng, k = p / ((int64_t)ng * ng);
      pid[s0 + q] = (int32_t)(fine_per_lattice * i + (int64_t)nf * (fine_per_lattice * j) +
                              (int64_t)nf * nf * (fine_per_lattice * k) + 1);
    }
  }
}

extern "C" int cube_sort_small_segments(cube_ctx *c, const int32_t *cnt, const int64_t *start, uint64_t *keys);

/* continuous particles already on the device -> current disjoint state */
static int state_from_device_xv(cube_ctx *c, const double *x, const float *v, int64_t n, int ng, float sigma_in,
                                float *sigma_out) {
  if (n > c->cap) return CUBE_ERR_IMAGE_OVERFLOW;
  const Geo g = geo_of(c);
  const unsigned nbc = (unsigned)((c->ncells + 255) / 256);
  hipLaunchKernelGGL(k_fill_zero_i32, dim3(1024), dim3(256), 0, c->stream, c->rhoc, c->ncells);
  hipLaunchKernelGGL(k_fill_zero_i32, dim3(1024), dim3(256), 0, c->stream, c->fill, c->ncells);
  CUBE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_xv_dest, dim3(4096), dim3(256), 0, c->stream, x, n, g, c->dest, c->key, c->rhoc);
  CUBE_LAUNCH_CHECK();
  int64_t total = 0;
  int st = cube_scan_counts(c, c->rhoc, c->cstart, c->ncells, &total);
  if (st) return st;
  if (total != n) return CUBE_ERR_COUNT;
  hipLaunchKernelGGL(k_place_keys, dim3(4096), dim3(256), 0, c->stream, c->dest, c->key, n, c->cstart, c->fill,
                     c->sortbuf);
  CUBE_LAUNCH_CHECK();
  if ((st = cube_sort_small_segments(c, c->rhoc, c->cstart, c->sortbuf))) return st;
  hipLaunchKernelGGL(k_xv_vfield, dim3(nbc), dim3(256), 0, c->stream, v, c->rhoc, c->cstart, c->sortbuf,
                     c->ncells, c->vfield);
  CUBE_LAUNCH_CHECK();
  float sigma = sigma_in;
  if (!(sigma > 0.f)) {
    hipLaunchKernelGGL(k_xv_resvar, dim3(1024), dim3(256), 0, c->stream, v, c->rhoc, c->cstart, c->sortbuf,
                       c->vfield, c->ncells, c->dred);
    CUBE_LAUNCH_CHECK();
    double part[1024];
    CUBE_CHECK(hipMemcpyAsync(part, c->dred, sizeof(part), hipMemcpyDeviceToHost, c->stream));
    CUBE_CHECK(hipStreamSynchronize(c->stream));
    double s = 0;
    for (int b = 0; b < 1024; ++b) s += part[b];
    sigma = (float)sqrt(s / (3.0 * (double)n));
  }
  const double venc = cube_vconst(sigma);
  const int nf = (int)c->nf, fpl = ng > 0 ? (int)(c->nf / ng) : 1;
  int32_t *pid = c->p.has_pid ? c->pid : nullptr;
  const int ix = c->p.izipx, iv = c->p.izipv;
#define ENC(IX, IV)                                                                                          \
  hipLaunchKernelGGL((k_xv_encode<IX, IV>), dim3(nbc), dim3(256), 0, c->stream, x, v, c->rhoc, c->cstart,       \
                     c->sortbuf, c->vfield, c->ncells, venc, ng, nf, fpl,                          \
                     (typename IntOf<IX>::T *)c->xp, (typename IntOf<IV>::T *)c->vp, pid)
  if (ix == 1 && iv == 1) ENC(1, 1);
  else if (ix == 1) ENC(1, 2);
  else if (iv == 1) ENC(2, 1);
  else ENC(2, 2);
#undef ENC
  CUBE_LAUNCH_CHECK();
  CUBE_CHECK(hipStreamSynchronize(c->stream));
  c->nplocal = n;
  c->npglobal = n;
  c->has_state = 1;
  if (sigma_out) *sigma_out = sigma;
  return CUBE_OK;
}

extern "C" int cube_set_state_xv(cube_ctx *c, int64_t n, const double *x, const float *v, float sigma_vi,
                                 float *sigma_out) {
  if (c) c->pm_ahead = 0; /* forces formed ahead by ext_pp_force no longer apply */
  if (!c || !x || !v || n < 0) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipSetDevice(c->device));
  /* stage through the (idle) fine-mesh buffers */
  const size_t xb = (size_t)n * 3 * sizeof(double), vb = (size_t)n * 3 * sizeof(float);
  const size_t nspec = (size_t)cube_spec_pitch((int)c->nf) * c->nf * c->nf;
  if (xb > c->fine_nspec * nspec * sizeof(float2) || vb > 3 * (size_t)c->nf * c->nf * c->nf * sizeof(float))
    return CUBE_ERR_INVALID;
  double *dx = (double *)c->specd_f;
  float *dv = c->force_f;
  CUBE_CHECK(hipMemcpy(dx, x, xb, hipMemcpyHostToDevice));
  CUBE_CHECK(hipMemcpy(dv, v, vb, hipMemcpyHostToDevice));
  return state_from_device_xv(c, dx, dv, n, 0, sigma_vi, sigma_out);
}

extern "C" int cube_ic_zeldovich(cube_ctx *c, const cube_ic_params *ip, float *sigma_out) {
  if (c) c->pm_ahead = 0; /* forces formed ahead by ext_pp_force no longer apply */
  if (!c || !ip || ip->ntab < 2 || !ip->lnp) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipSetDevice(c->device));
  const int ng = (int)(c->nc * c->p.np_nc);
  const int64_t n = (int64_t)ng * ng * ng;
  if (4 * c->nc % ng != 0 || n > c->cap) return CUBE_ERR_INVALID;
  /* x (3 doubles) staged in the force spectra, v (3 floats) in the force mesh */
  if ((size_t)n * 24 > (size_t)c->fine_nspec * cube_spec_pitch((int)c->nf) * c->nf * c->nf * sizeof(float2))
    return CUBE_ERR_INVALID;
  FFTPlan *plan = nullptr;
  int st = cube_fft_create(c, ng, &plan);
  if (st) return st;
  float *w = c->rho_f;           /* ng^3 <= nf^3 */
  float2 *spec = c->spec_f;
  double *tab = nullptr;
  CUBE_CHECK(hipMalloc(&tab, ip->ntab * sizeof(double)));
  CUBE_CHECK(hipMemcpy(tab, ip->lnp, ip->ntab * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_ic_noise, dim3(4096), dim3(256), 0, c->stream, w, n, ip->seed);
  if (hipGetLastError() != hipSuccess) st = CUBE_ERR_HIP;
  if (!st) st = cube_fft_forward_full(c, plan, w, spec);
  const double cell = ip->box / ng;
  if (!st) {
    hipLaunchKernelGGL(k_ic_filter, dim3(4096), dim3(256), 0, c->stream, spec, ng, cell, tab, ip->ntab, ip->lnk0,
                       ip->dlnk, 1.0 / (double)n);
    if (hipGetLastError() != hipSuccess) st = CUBE_ERR_HIP;
  }
  if (!st) st = cube_fft_inverse_full(c, plan, spec, w); /* phi on the lattice */
  double *dx = (double *)c->specd_f;
  float *dv = c->force_f;
  if (!st) {
    const double dx_fine = 4.0 / (double)c->p.np_nc;
    hipLaunchKernelGGL(k_ic_particles, dim3(4096), dim3(256), 0, c->stream, w, ng, (int)c->p.np_nc, (int)c->nc,
                       -dx_fine / (8.0 * 3.141592653589793), ip->vfactor, dx, dv);
    if (hipGetLastError() != hipSuccess) st = CUBE_ERR_HIP;
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(tab);
  cube_fft_destroy(plan);
  if (st) return st;
  return state_from_device_xv(c, dx, dv, n, ng, ip->sigma_vi, sigma_out);
}
