/*
 * cube_dpm.hip -- differentiable particle-mesh forward model and its
 * hand-written adjoint (include/cube_dpm.h).  No reference counterpart.
 *
 * A particle is held as its FP32 displacement d (in mesh cells) from its
 * lattice point q = (nf/ng) * (i mod ng, i/ng mod ng, i/ng^2); particles
 * keep their index for the whole run (no sorting), so the adjoint state is
 * per particle.  CIC convention: a particle at x = q + d touches mesh points
 * floor(x) and floor(x)+1 per axis (mod nf) with weights 1-f and f, formed
 * as q + floor(d) and f = d - floor(d): exact, and resolved to the
 * displacement's ulp rather than to the ulp of a position near nf, so far
 * fewer particles than with FP32 positions sit within rounding of a cell
 * edge, where the weights' slope (and so the gradient) jumps.
 *
 * Forward operators and their transposes:
 *   deposit  rho_c = sum_i W(x_i,c)          ^T: lam -> sum_c lam_c W(x_i,c)
 *   Poisson  F_d = (1/N^3) C2R[i K_d R2C(rho)]   ^T: K_d -> -K_d, summed over d
 *   gather   F_i = sum_c W(x_i,c) M_c         ^T (in M): S_c = sum_i W g_i
 *   and the position derivatives of deposit and gather use dW/dx.
 * K_d = k_d/k^2 (k = 2 pi m/N), zero at k=0 and on the Nyquist plane of
 * axis d, so every product spectrum stays Hermitian and the real
 * transforms are exact transposes of each other.
 */
#include "cube_dpm_internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PT 256

/* --------------------------- k-space factors --------------------------- */
__device__ __host__ static inline int fold(int m, int n) { return m <= n / 2 ? m : m - n; }
/* K_d = k_d/k^2 for the mode (mx, my, mz) of an n^3 mesh (mx in [0, n/2]) */
__device__ static inline float kfac(int d, int mx, int my, int mz, int n) {
  const int m[3] = {mx, fold(my, n), fold(mz, n)};
  if (2 * (m[d] < 0 ? -m[d] : m[d]) == n) return 0.f; /* Nyquist plane of axis d */
  const float s = 6.28318530717958648f / (float)n;
  const float k[3] = {s * m[0], s * m[1], s * m[2]};
  const float k2 = (k[0] * k[0] + k[1] * k[1]) + k[2] * k[2];
  return k2 > 0.f ? k[d] / k2 : 0.f;
}

__global__ void k_dpm_kern(float *kern, int n) {
  const int nh = n / 2 + 1;
  const int64_t tot = (int64_t)3 * n * n * nh;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
    const int kx = (int)(q % nh), ky = (int)((q / nh) % n), kz = (int)((q / ((int64_t)nh * n)) % n);
    const int d = (int)(q / ((int64_t)nh * n * n));
    kern[q] = kfac(d, kx, ky, kz, n);
  }
}

/* out = i K_d A in (sign=+1) on the lattice spectrum (pitched rows) */
__global__ void k_dpm_ic_mul(const float2 *in, float2 *out, const float *amp, int d, int n) {
  const int nh = n / 2 + 1, sp = cube_spec_pitch(n);
  const int64_t tot = (int64_t)n * n * nh;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
    const int kx = (int)(q % nh), ky = (int)((q / nh) % n), kz = (int)(q / ((int64_t)nh * n));
    const int64_t s = ((int64_t)kz * n + ky) * sp + kx;
    const float f = kfac(d, kx, ky, kz, n) * amp[q];
    const float2 v = in[s];
    out[s] = make_float2(-v.y * f, v.x * f);
  }
}
/* out = sum_d (-i) K_d A_d in_d: the transposed filter (amp = nullptr: A=1) */
__global__ void k_dpm_combine(const float2 *in3, int64_t plane, float2 *out, const float *amp, int n) {
  const int nh = n / 2 + 1, sp = cube_spec_pitch(n);
  const int64_t tot = (int64_t)n * n * nh;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
    const int kx = (int)(q % nh), ky = (int)((q / nh) % n), kz = (int)(q / ((int64_t)nh * n));
    const int64_t s = ((int64_t)kz * n + ky) * sp + kx;
    const float a = amp ? amp[q] : 1.f;
    float2 acc = make_float2(0.f, 0.f);
    for (int d = 0; d < 3; ++d) {
      const float f = kfac(d, kx, ky, kz, n) * a;
      const float2 v = in3[d * plane + s];
      /* (-i f) v = (f v.y, -f v.x) */
      acc.x += f * v.y;
      acc.y -= f * v.x;
    }
    out[s] = acc;
  }
}

/* --------------------------- particle kernels --------------------------- */
__device__ static inline float wrapf(float x, float n) {
  x = x - n * floorf(x / n);
  return x >= n ? x - n : x;
}
struct V3 { /* three interleaved components, read with one 12-byte load */
  float x, y, z;
};
struct Cic {
  int i[3][2];   /* mesh indices per axis (periodic) */
  float w[3][2]; /* weights */
};
/* the lattice point of particle q, in mesh cells: r*(q mod ng, ...) */
struct Lat {
  int ng, r;
};
__device__ static inline Cic cic_of(const float *dsp, int64_t q, Lat L, int n) {
  const int l[3] = {(int)(q % L.ng), (int)((q / L.ng) % L.ng), (int)(q / ((int64_t)L.ng * L.ng))};
  Cic c;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float fl = floorf(dsp[a]);
    const float f = dsp[a] - fl;
    int i0 = L.r * l[a] + (int)fl;
    i0 = ((i0 % n) + n) % n; /* always a valid mesh index, whatever d holds */
    c.i[a][0] = i0;
    c.i[a][1] = i0 + 1 == n ? 0 : i0 + 1;
    c.w[a][0] = 1.f - f;
    c.w[a][1] = f;
  }
  return c;
}

__global__ void k_dpm_ic_particles(const float *psi, int ng, int nf, float cx, float cp, float scale, float *x,
                                   float *p) {
  const int64_t n3 = (int64_t)ng * ng * ng;
  const float r = (float)nf / (float)ng;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n3; q += (int64_t)gridDim.x * blockDim.x) {
    const int l[3] = {(int)(q % ng), (int)((q / ng) % ng), (int)(q / ((int64_t)ng * ng))};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float s = psi[a * n3 + q] * scale * r;
      x[3 * q + a] = cx * s; /* displacement from the lattice point l*r */
      p[3 * q + a] = cp * s;
    }
  }
}

/*
 * The deposit and its transpose add particles' terms into mesh points from
 * many threads at once.  They add 64-bit fixed-point integers, whose sums do
 * not depend on the order of the atomics, and one pass converts to FP32
 * afterwards: the meshes, and so the whole model, are the same bits run to
 * run and in any launch shape or graph.  The deposit's terms are weights in
 * [0, 1], kept with 32 fraction bits; the transpose's terms w*g are scaled
 * by a power of two that puts max|g| in [2^39, 2^40) (found on the device
 * before the scatter), so every sum of up to 2^23 terms stays in range.
 * Each term is rounded once to the fixed-point grid, finer than FP32 at
 * the terms' scale; the conversion rounds the exact sum once.
 */
#define FIX_DEPOSIT 0x1p32
__global__ void k_dpm_deposit(const float *x, int64_t np, Lat L, int n, unsigned long long *acc) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
    const Cic c = cic_of(x + 3 * q, q, L, n);
#pragma unroll
    for (int z = 0; z < 2; ++z)
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int xx = 0; xx < 2; ++xx)
          atomicAdd(acc + ((int64_t)c.i[2][z] * n + c.i[1][y]) * n + c.i[0][xx],
                    (unsigned long long)llrint((double)(c.w[0][xx] * c.w[1][y] * c.w[2][z]) * FIX_DEPOSIT));
  }
}
/* the transpose's scale: 2^(40-e) for max|g| = m 2^e, m in [0.5, 1) */
__device__ static inline double fix_scale(const unsigned *gmx) {
  int e = 0;
  (void)frexp((double)__uint_as_float(*gmx), &e);
  return ldexp(1.0, 40 - e);
}
/* out = acc / scale (scale: FIX_DEPOSIT, or the transpose's from gmx) */
__global__ void k_dpm_from_fix(const unsigned long long *acc, int64_t n, const unsigned *gmx, float *out) {
  const double inv = gmx ? 1.0 / fix_scale(gmx) : 1.0 / FIX_DEPOSIT;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    out[q] = (float)((double)(long long)acc[q] * inv);
}
/* max |g| over 3*np values into *gmx (non-negative floats order as their bits) */
__global__ void k_dpm_absmax(const float *g, int64_t n, unsigned *gmx) {
  float m = 0.f;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(g[q]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(gmx, __float_as_uint(m));
}

/* p += coef * sum_c W F_c (F interleaved, 3 per mesh point) */
__global__ void k_dpm_kick(const float *x, float *p, int64_t np, const float *F, Lat L, int n, float coef) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
    const Cic c = cic_of(x + 3 * q, q, L, n);
    float f[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int z = 0; z < 2; ++z)
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int xx = 0; xx < 2; ++xx) {
          const float w = c.w[0][xx] * c.w[1][y] * c.w[2][z];
          const V3 m = *reinterpret_cast<const V3 *>(F + 3 * (((int64_t)c.i[2][z] * n + c.i[1][y]) * n + c.i[0][xx]));
          f[0] += w * m.x;
          f[1] += w * m.y;
          f[2] += w * m.z;
        }
#pragma unroll
    for (int a = 0; a < 3; ++a) p[3 * q + a] += coef * f[a];
  }
}

/* the displacement needs no wrap: the mesh index is taken mod nf */
__global__ void k_dpm_drift(float *x, const float *p, int64_t np, float coef) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < 3 * np; q += (int64_t)gridDim.x * blockDim.x)
    x[q] = x[q] + coef * p[q];
}
/* positions q + d wrapped into [0, nf), for cube_dpm_state */
__global__ void k_dpm_positions(const float *x, int64_t np, Lat L, int n, float *out) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
    const int l[3] = {(int)(q % L.ng), (int)((q / L.ng) % L.ng), (int)(q / ((int64_t)L.ng * L.ng))};
#pragma unroll
    for (int a = 0; a < 3; ++a) out[3 * q + a] = wrapf((float)(L.r * l[a]) + x[3 * q + a], (float)n);
  }
}

__global__ void k_dpm_delta(const float *rho, int64_t n3, float inv_nbar, float *delta) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n3; q += (int64_t)gridDim.x * blockDim.x)
    delta[q] = rho[q] * inv_nbar - 1.f;
}

/* S_d[c] += W(x_i,c) g_i,d  (SoA meshes) */
__global__ void k_dpm_scatter3(const float *x, const float *g, int64_t np, Lat L, int n, unsigned long long *S,
                               const unsigned *gmx) {
  const int64_t n3 = (int64_t)n * n * n;
  const double scale = fix_scale(gmx);
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
    const Cic c = cic_of(x + 3 * q, q, L, n);
    const float gv[3] = {g[3 * q], g[3 * q + 1], g[3 * q + 2]};
#pragma unroll
    for (int z = 0; z < 2; ++z)
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int xx = 0; xx < 2; ++xx) {
          const float w = c.w[0][xx] * c.w[1][y] * c.w[2][z];
          const int64_t m = ((int64_t)c.i[2][z] * n + c.i[1][y]) * n + c.i[0][xx];
#pragma unroll
          for (int a = 0; a < 3; ++a)
            atomicAdd(S + a * n3 + m, (unsigned long long)llrint((double)(w * gv[a]) * scale));
        }
  }
}

/* xbar_i += sum_c dW(x_i,c)/dx * (s_lam lam_c + g_i . M_c)   (g, M may be null) */
__global__ void k_dpm_grad(const float *x, int64_t np, Lat L, int n, const float *lam, float s_lam, const float *g,
                           const float *M, float *xbar) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
    const Cic c = cic_of(x + 3 * q, q, L, n);
    float gv[3] = {0.f, 0.f, 0.f};
    if (g) {
      gv[0] = g[3 * q];
      gv[1] = g[3 * q + 1];
      gv[2] = g[3 * q + 2];
    }
    float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int z = 0; z < 2; ++z)
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int xx = 0; xx < 2; ++xx) {
          const int64_t m = ((int64_t)c.i[2][z] * n + c.i[1][y]) * n + c.i[0][xx];
          float v = lam ? s_lam * lam[m] : 0.f;
          if (M) {
            const V3 mv = *reinterpret_cast<const V3 *>(M + 3 * m);
            v += (gv[0] * mv.x + gv[1] * mv.y) + gv[2] * mv.z;
          }
          /* dW/dx_a: the weight of axis a replaced by its slope -1 (lower) or +1 (upper) */
          const float sx = xx ? 1.f : -1.f, sy = y ? 1.f : -1.f, sz = z ? 1.f : -1.f;
          acc[0] += v * (sx * c.w[1][y] * c.w[2][z]);
          acc[1] += v * (c.w[0][xx] * sy * c.w[2][z]);
          acc[2] += v * (c.w[0][xx] * c.w[1][y] * sz);
        }
#pragma unroll
    for (int a = 0; a < 3; ++a) xbar[3 * q + a] += acc[a];
  }
}

__global__ void k_dpm_axpy3(float *dst, const float *src, int64_t n, float coef) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    dst[q] += coef * src[q];
}
__global__ void k_dpm_scale_to(float *dst, const float *src, int64_t n, float coef) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    dst[q] = coef * src[q];
}
__global__ void k_dpm_zero(float *p, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    p[q] = 0.f;
}
/* psibar_d (SoA lattice) = r (cx xbar_d + cp pbar_d) */
__global__ void k_dpm_psibar(const float *xbar, const float *pbar, int64_t np, float cx, float cp, float r,
                             float *psib) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x)
    for (int a = 0; a < 3; ++a) psib[a * np + q] = r * (cx * xbar[3 * q + a] + cp * pbar[3 * q + a]);
}

/* --------------------------------- host --------------------------------- */
#define GRID 2048
/* every grid-stride loop of the model runs with the box's tuned shape */
#define LAUNCH(k, ...)                                                                      \
  do {                                                                                      \
    hipLaunchKernelGGL(k, dim3(d->grid), dim3(d->block), 0, d->c.stream, __VA_ARGS__);      \
    CUBE_LAUNCH_CHECK();                                                                    \
  } while (0)

static Lat lat(const cube_dpm *d) { return Lat{d->p.ng, d->p.nf / d->p.ng}; }

static int zero(cube_dpm *d, float *p, int64_t n) {
  LAUNCH(k_dpm_zero, p, n);
  return CUBE_OK;
}

/* force mesh of positions x into d->force (interleaved 3 per point) */
static int force_of(cube_dpm *d, const float *x) {
  const int nf = d->p.nf;
  int st = zero(d, reinterpret_cast<float *>(d->acc), 2 * d->nf3);
  if (st) return st;
  LAUNCH(k_dpm_deposit, x, d->np, lat(d), nf, d->acc);
  LAUNCH(k_dpm_from_fix, d->acc, d->nf3, (const unsigned *)nullptr, d->rho);
  if ((st = cube_fft_r2c_3d(&d->c, d->plan_f, d->rho, d->spec_f))) return st;
  return cube_fft_green_c2r_3d(&d->c, d->plan_f, d->spec_f, d->kern, 0, d->work3_f, d->force, (float)nf, nullptr,
                               -1, (unsigned *)d->c.ured, d->c.slots);
}

/* one KDK step on (x, pm) */
static int step(cube_dpm *d, int s, int have_force) {
  int st;
  if (!have_force && (st = force_of(d, d->x))) return st;
  const double *c = d->steps + 3 * s;
  LAUNCH(k_dpm_kick, d->x, d->pm, d->np, d->force, lat(d), d->p.nf, (float)c[0]);
  LAUNCH(k_dpm_drift, d->x, d->pm, d->np, (float)c[1]);
  if ((st = force_of(d, d->x))) return st;
  LAUNCH(k_dpm_kick, d->x, d->pm, d->np, d->force, lat(d), d->p.nf, (float)c[2]);
  return CUBE_OK;
}

/* ICs from the noise into (x, pm) and the stored ICs */
static int ics(cube_dpm *d, const float *noise) {
  const int ng = d->p.ng;
  int st;
  if ((st = cube_fft_forward_full(&d->c, d->plan_g, noise, d->spec_g))) return st;
  for (int a = 0; a < 3; ++a) {
    LAUNCH(k_dpm_ic_mul, d->spec_g, d->work3_g, d->amp, a, ng);
    if ((st = cube_fft_inverse_full(&d->c, d->plan_g, d->work3_g, d->psi + a * d->ng3))) return st;
  }
  LAUNCH(k_dpm_ic_particles, d->psi, ng, d->p.nf, (float)d->ic[0], (float)d->ic[1], 1.f / (float)d->ng3, d->x,
         d->pm);
  CUBE_CHECK(hipMemcpyAsync(d->xs, d->x, d->np * 3 * sizeof(float), hipMemcpyDeviceToDevice, d->c.stream));
  CUBE_CHECK(hipMemcpyAsync(d->p0, d->pm, d->np * 3 * sizeof(float), hipMemcpyDeviceToDevice, d->c.stream));
  return CUBE_OK;
}

/* positions after s steps (from the checkpoints, or recomputed) */
static int positions_at(cube_dpm *d, int s, const float **out) {
  if (d->p.checkpoint) {
    *out = d->xs + (int64_t)s * d->np * 3;
    return CUBE_OK;
  }
  CUBE_CHECK(hipMemcpyAsync(d->x, d->xs, d->np * 3 * sizeof(float), hipMemcpyDeviceToDevice, d->c.stream));
  CUBE_CHECK(hipMemcpyAsync(d->pm, d->p0, d->np * 3 * sizeof(float), hipMemcpyDeviceToDevice, d->c.stream));
  for (int q = 0; q < s; ++q) {
    int st = step(d, q, 0);
    if (st) return st;
  }
  *out = d->x;
  return CUBE_OK;
}

/* xbar += VJP of F at x for the cotangent coef*pbar on F */
static int vjp_force(cube_dpm *d, const float *x, float coef) {
  const int nf = d->p.nf;
  int st;
  if ((st = force_of(d, x))) return st; /* M = F mesh at x */
  LAUNCH(k_dpm_scale_to, d->gvec, d->pbar, d->np * 3, coef);
  if ((st = zero(d, reinterpret_cast<float *>(d->acc), 6 * d->nf3))) return st;
  if ((st = zero(d, reinterpret_cast<float *>(d->gmx), 1))) return st;
  LAUNCH(k_dpm_absmax, d->gvec, 3 * d->np, d->gmx);
  LAUNCH(k_dpm_scatter3, x, d->gvec, d->np, lat(d), nf, d->acc, d->gmx);
  LAUNCH(k_dpm_from_fix, d->acc, 3 * d->nf3, d->gmx, d->S);
  for (int a = 0; a < 3; ++a)
    if ((st = cube_fft_forward_full(&d->c, d->plan_f, d->S + a * d->nf3, d->work3_f + a * d->sf))) return st;
  LAUNCH(k_dpm_combine, d->work3_f, d->sf, d->spec_f, (const float *)nullptr, nf);
  if ((st = cube_fft_inverse_full(&d->c, d->plan_f, d->spec_f, d->lam))) return st;
  const float s = 1.f / ((float)nf * (float)nf * (float)nf);
  LAUNCH(k_dpm_grad, x, d->np, lat(d), nf, d->lam, s, d->gvec, d->force, d->xbar);
  return CUBE_OK;
}

static int dalloc_f(float **p, int64_t n) {
  return hipMalloc((void **)p, (size_t)(n > 0 ? n : 1) * sizeof(float)) == hipSuccess ? CUBE_OK : CUBE_ERR_ALLOC;
}

extern "C" void cube_dpm_destroy(cube_dpm *d) {
  if (!d) return;
  void *ptrs[] = {d->amp, d->kern, d->xs, d->x, d->pm, d->p0, d->rho, d->force, d->S, d->lam, d->spec_f,
                  d->work3_f, d->spec_g, d->work3_g, d->psi, d->xbar, d->pbar, d->gvec, d->c.ured, d->c.slots,
                  d->acc, d->gmx};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  cube_fft_destroy(d->plan_g);
  cube_fft_destroy(d->plan_f);
  if (d->c.stream) (void)hipStreamDestroy(d->c.stream);
  free(d->steps);
  free(d);
}

extern "C" int cube_dpm_create(const cube_dpm_params *p, cube_dpm **out) {
  if (!p || !out || p->ng < 4 || p->nf < p->ng || p->nf % p->ng || p->nsteps < 1) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipSetDevice(p->device));
  cube_dpm *d = (cube_dpm *)calloc(1, sizeof(cube_dpm));
  if (!d) return CUBE_ERR_ALLOC;
  d->p = *p;
  d->c.device = p->device;
  d->grid = GRID;
  d->block = PT;
  d->np = d->ng3 = (int64_t)p->ng * p->ng * p->ng;
  d->nf3 = (int64_t)p->nf * p->nf * p->nf;
  d->sg = (int64_t)cube_spec_pitch(p->ng) * p->ng * p->ng;
  d->sf = (int64_t)cube_spec_pitch(p->nf) * p->nf * p->nf;
  d->steps = (double *)calloc(3 * p->nsteps, sizeof(double));
  int st = d->steps ? CUBE_OK : CUBE_ERR_ALLOC;
  if (!st && hipStreamCreate(&d->c.stream) != hipSuccess) st = CUBE_ERR_HIP;
  if (!st && hipMalloc((void **)&d->c.ured, 64 * sizeof(uint64_t)) != hipSuccess) st = CUBE_ERR_ALLOC;
  if (!st && hipMalloc((void **)&d->c.slots, CUBE_NSLOTREG * CUBE_SLOTS * sizeof(uint64_t)) != hipSuccess)
    st = CUBE_ERR_ALLOC;
  if (!st) st = cube_fft_create(&d->c, p->ng, &d->plan_g);
  if (!st) st = cube_fft_create(&d->c, p->nf, &d->plan_f);
  const int64_t np3 = d->np * 3;
  const int64_t ncp = p->checkpoint ? (int64_t)(p->nsteps + 1) : 1;
  if (!st) st |= dalloc_f(&d->amp, (int64_t)p->ng * p->ng * (p->ng / 2 + 1));
  if (!st) st |= dalloc_f(&d->kern, 3 * (int64_t)p->nf * p->nf * (p->nf / 2 + 1));
  if (!st) st |= dalloc_f(&d->xs, ncp * np3);
  if (!st) st |= dalloc_f(&d->x, np3);
  if (!st) st |= dalloc_f(&d->pm, np3);
  if (!st) st |= dalloc_f(&d->p0, np3);
  if (!st) st |= dalloc_f(&d->rho, d->nf3);
  if (!st) st |= dalloc_f(&d->force, 3 * d->nf3);
  if (!st) st |= dalloc_f(&d->S, 3 * d->nf3);
  if (!st) st |= dalloc_f(reinterpret_cast<float **>(&d->acc), 6 * d->nf3);
  if (!st) st |= dalloc_f(reinterpret_cast<float **>(&d->gmx), 1);
  if (!st) st |= dalloc_f(&d->lam, d->nf3);
  if (!st) st |= dalloc_f((float **)&d->spec_f, 2 * d->sf);
  if (!st) st |= dalloc_f((float **)&d->work3_f, 6 * d->sf);
  if (!st) st |= dalloc_f((float **)&d->spec_g, 2 * d->sg);
  if (!st) st |= dalloc_f((float **)&d->work3_g, 6 * d->sg);
  if (!st) st |= dalloc_f(&d->psi, 3 * d->ng3);
  if (!st) st |= dalloc_f(&d->xbar, np3);
  if (!st) st |= dalloc_f(&d->pbar, np3);
  if (!st) st |= dalloc_f(&d->gvec, np3);
  if (!st) {
    hipLaunchKernelGGL(k_dpm_kern, dim3(GRID), dim3(PT), 0, d->c.stream, d->kern, p->nf);
    if (hipGetLastError() != hipSuccess) st = CUBE_ERR_HIP;
  }
  if (!st && hipStreamSynchronize(d->c.stream) != hipSuccess) st = CUBE_ERR_HIP;
  if (st) {
    cube_dpm_destroy(d);
    return st;
  }
  *out = d;
  return CUBE_OK;
}

extern "C" int cube_dpm_set_amplitude(cube_dpm *d, const float *amp) {
  if (!d || !amp) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipSetDevice(d->p.device));
  const int ng = d->p.ng;
  CUBE_CHECK(hipMemcpy(d->amp, amp, (size_t)ng * ng * (ng / 2 + 1) * sizeof(float), hipMemcpyHostToDevice));
  return CUBE_OK;
}

extern "C" int cube_dpm_set_schedule(cube_dpm *d, const double *ic, const double *steps) {
  if (!d || !ic || !steps) return CUBE_ERR_INVALID;
  d->ic[0] = ic[0];
  d->ic[1] = ic[1];
  memcpy(d->steps, steps, 3 * d->p.nsteps * sizeof(double));
  return CUBE_OK;
}

extern "C" int cube_dpm_forward(cube_dpm *d, const float *noise, float *delta) {
  if (!d || !noise || !delta) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipSetDevice(d->p.device));
  int st = ics(d, noise);
  if (st) return st;
  if ((st = force_of(d, d->x))) return st;
  for (int s = 0; s < d->p.nsteps; ++s) {
    if ((st = step(d, s, 1))) return st;
    if (d->p.checkpoint)
      CUBE_CHECK(hipMemcpyAsync(d->xs + (int64_t)(s + 1) * d->np * 3, d->x, d->np * 3 * sizeof(float),
                                hipMemcpyDeviceToDevice, d->c.stream));
  }
  if ((st = zero(d, reinterpret_cast<float *>(d->acc), 2 * d->nf3))) return st;
  LAUNCH(k_dpm_deposit, d->x, d->np, lat(d), d->p.nf, d->acc);
  LAUNCH(k_dpm_from_fix, d->acc, d->nf3, (const unsigned *)nullptr, d->rho);
  LAUNCH(k_dpm_delta, d->rho, d->nf3, (float)((double)d->nf3 / (double)d->np), delta);
  d->has_forward = 1;
  return CUBE_OK;
}

extern "C" int cube_dpm_adjoint(cube_dpm *d, const float *delta_bar, float *noise_bar) {
  if (!d || !delta_bar || !noise_bar) return CUBE_ERR_INVALID;
  if (!d->has_forward) return CUBE_ERR_STATE;
  CUBE_CHECK(hipSetDevice(d->p.device));
  const int S = d->p.nsteps, nf = d->p.nf, ng = d->p.ng;
  const int64_t np3 = d->np * 3;
  int st;
  /* final density: delta = rho/nbar - 1 */
  const float *xS = nullptr;
  if ((st = positions_at(d, S, &xS))) return st;
  if ((st = zero(d, d->xbar, np3))) return st;
  if ((st = zero(d, d->pbar, np3))) return st;
  LAUNCH(k_dpm_grad, xS, d->np, lat(d), nf, delta_bar, (float)((double)d->nf3 / (double)d->np), (const float *)nullptr,
         (const float *)nullptr, d->xbar);
  for (int s = S - 1; s >= 0; --s) {
    const double *c = d->steps + 3 * s;
    LAUNCH(k_dpm_axpy3, d->pbar, d->xbar, np3, (float)c[1]); /* drift^T */
    const double coef = c[0] + (s > 0 ? d->steps[3 * (s - 1) + 2] : 0.0); /* kick A of s, kick B of s-1 */
    const float *xs = nullptr;
    if ((st = positions_at(d, s, &xs))) return st;
    if ((st = vjp_force(d, xs, (float)coef))) return st;
  }
  /* ICs^T: psibar = r (cx xbar + cp pbar); noise_bar = (1/ng^3) IFFT[sum_d -i K_d A FFT(psibar_d)] */
  LAUNCH(k_dpm_psibar, d->xbar, d->pbar, d->np, (float)d->ic[0], (float)d->ic[1], (float)nf / (float)ng, d->psi);
  for (int a = 0; a < 3; ++a)
    if ((st = cube_fft_forward_full(&d->c, d->plan_g, d->psi + a * d->ng3, d->work3_g + a * d->sg))) return st;
  LAUNCH(k_dpm_combine, d->work3_g, d->sg, d->spec_g, d->amp, ng);
  if ((st = cube_fft_inverse_full(&d->c, d->plan_g, d->spec_g, noise_bar))) return st;
  LAUNCH(k_dpm_scale_to, noise_bar, noise_bar, d->ng3, 1.f / (float)d->ng3);
  return CUBE_OK;
}

extern "C" int cube_dpm_state(cube_dpm *d, float *x, float *p) {
  if (!d) return CUBE_ERR_INVALID;
  if (!d->has_forward) return CUBE_ERR_STATE;
  CUBE_CHECK(hipSetDevice(d->p.device));
  const size_t b = d->np * 3 * sizeof(float);
  const float *xf = d->p.checkpoint ? d->xs + (int64_t)d->p.nsteps * d->np * 3 : d->x;
  if (x) LAUNCH(k_dpm_positions, xf, d->np, lat(d), d->p.nf, x); /* held as displacements */
  if (p) CUBE_CHECK(hipMemcpyAsync(p, d->pm, b, hipMemcpyDeviceToDevice, d->c.stream));
  return CUBE_OK;
}

extern "C" int cube_dpm_sync(cube_dpm *d) {
  if (!d) return CUBE_ERR_INVALID;
  CUBE_CHECK(hipStreamSynchronize(d->c.stream));
  return CUBE_OK;
}
