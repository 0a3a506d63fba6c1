/*
 * cube_power.hip -- matter power spectrum of the particle state
 * (cube_power in include/cube.h), the analysis of WU/cicpower.f90:96-187 and
 * WU/powerspectrum.f90:28-135 on the GPU: cell-centred CIC in FP64,
 * delta = rho/mean - 1, FP32 FFT, then per-mode |delta_k|^2 binned in
 * linear |k| bins with the sinc^4 CIC window removed.
 */
#include "cube_internal.h"
#include <math.h>
#include <string.h>
#include <mutex>

#define PW_T 256

/* cell-centred CIC of the disjoint state: x = global coarse position,
   mesh coordinate x*ng/nc - 0.5 (WU/cicpower.f90:113) */
template <int IX>
__global__ __launch_bounds__(PW_T) void k_pw_deposit(const typename IntOf<IX>::T *__restrict__ xp,
                                                     const int64_t *__restrict__ cstart, Geo g, int ng,
                                                     double *__restrict__ rho) {
  const int nt = g.nt;
  const int64_t row = blockIdx.x;
  const int j = (int)(row % nt), k = (int)((row / nt) % nt), t = (int)(row / ((int64_t)nt * nt));
  const int tt[3] = {t % g.nnt, (t / g.nnt) % g.nnt, t / (g.nnt * g.nnt)};
  __shared__ int64_t cs[520];
  const int64_t c0 = row * nt;
  for (int i = threadIdx.x; i <= nt; i += PW_T) cs[i] = cstart[c0 + i];
  __syncthreads();
  const double scale = (double)ng / (double)g.nc;
  for (int64_t p = cs[0] + threadIdx.x; p < cs[nt]; p += PW_T) {
    int lo = 0, hi = nt;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (cs[mid] <= p) lo = mid; else hi = mid;
    }
    const int cell[3] = {tt[0] * nt + lo, tt[1] * nt + j, tt[2] * nt + k};
    int i1[3];
    double w[3][2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double x = (double)cell[a] + cube_xoff(IX, (int32_t)xp[3 * p + a]);
      const double m = x * scale - 0.5;
      const double fl = floor(m);
      const double d2 = m - fl;
      i1[a] = (int)fl;
      w[a][0] = 1.0 - d2;
      w[a][1] = d2;
    }
#pragma unroll
    for (int cz = 0; cz < 2; ++cz)
#pragma unroll
      for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cx = 0; cx < 2; ++cx) {
          const int x = wrapi(i1[0] + cx, ng), y = wrapi(i1[1] + cy, ng), z = wrapi(i1[2] + cz, ng);
          atomicAdd(rho + ((int64_t)z * ng + y) * ng + x, w[0][cx] * w[1][cy] * w[2][cz]);
        }
  }
}

__global__ void k_pw_delta(const double *rho, int64_t n3, double inv_mean, float *delta) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n3; q += (int64_t)gridDim.x * blockDim.x)
    delta[q] = (float)(rho[q] * inv_mean - 1.0);
}

__device__ static inline double sincd(double k, int n) {
  if (k == 0.0) return 1.0;
  const double a = 3.14159265358979323846 * k / (double)n;
  return sin(a) / a;
}

/* bins 1..nbin of sum |k|, sum Delta^2, count (WU/powerspectrum.f90:60-133) */
__global__ __launch_bounds__(PW_T) void k_pw_bin(const float2 *__restrict__ spec, int ng, int nbin,
                                                 double *__restrict__ ksum, double *__restrict__ psum,
                                                 unsigned long long *__restrict__ cnt) {
  extern __shared__ double lsum[]; /* 2*(nbin+1) doubles, then nbin+1 counters */
  unsigned long long *lcnt = reinterpret_cast<unsigned long long *>(lsum + 2 * (nbin + 1));
  for (int i = threadIdx.x; i <= nbin; i += PW_T) {
    lsum[i] = 0.0;
    lsum[nbin + 1 + i] = 0.0;
    lcnt[i] = 0ull;
  }
  __syncthreads();
  const int ny = ng / 2, nh = ng / 2 + 1, sp = cube_spec_pitch(ng);
  const int64_t tot = (int64_t)ng * ng * nh;
  const double n3 = (double)ng * ng * ng;
  for (int64_t q = blockIdx.x * (int64_t)PW_T + threadIdx.x; q < tot; q += (int64_t)gridDim.x * PW_T) {
    const int ix = (int)(q % nh), iy = (int)((q / nh) % ng), iz = (int)(q / ((int64_t)nh * ng));
    bool skip = ix == 0 && iy == 0 && iz == 0;
    const bool edge = ix == 0 || ix == ny;
    skip = skip || (edge && iy > ny) || (edge && (iy == 0 || iy == ny) && iz > ny);
    if (skip) continue;
    const int kx = ix == ny ? -ny : ix, ky = iy < ny ? iy : iy - ng, kz = iz < ny ? iz : iz - ng;
    const double kr = sqrt((double)(kx * kx + ky * ky + kz * kz));
    const int ib = (int)rint(kr);
    if (ib < 1 || ib > nbin) continue;
    const float2 v = spec[((int64_t)iz * ng + iy) * sp + ix];
    const double re = (double)v.x / n3, im = (double)v.y / n3;
    const double s = sincd(kx, ng) * sincd(ky, ng) * sincd(kz, ng);
    const double amp = (re * re + im * im) / (s * s * s * s) * 4.0 * 3.14159265358979323846 * kr * kr * kr;
    atomicAdd(lsum + ib, kr);
    atomicAdd(lsum + nbin + 1 + ib, amp);
    atomicAdd(lcnt + ib, 1ull);
  }
  __syncthreads();
  for (int i = threadIdx.x; i <= nbin; i += PW_T) {
    if (lcnt[i]) {
      atomicAdd(ksum + i, lsum[i]);
      atomicAdd(psum + i, lsum[nbin + 1 + i]);
      atomicAdd(cnt + i, lcnt[i]);
    }
  }
}

__global__ void k_pw_zero(double *a, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    a[q] = 0.0;
}

/* Checks shared by both entry points. */
static int power_args(cube_ctx *c, int32_t ng, int32_t nbin, const double *ksum, const double *psum,
                      const int64_t *count) {
  if (!c || !ksum || !psum || !count || ng < 4 || ng > c->nf || c->nf % ng || nbin < 1) return CUBE_ERR_INVALID;
  if (!c->has_state) return CUBE_ERR_STATE;
  if (c->nt > 519) return CUBE_ERR_INVALID; /* k_pw_deposit stages nt+1 row offsets in LDS */
  /* the FP64 mesh lives in the force spectra */
  if ((size_t)ng * ng * ng * sizeof(double) >
      (size_t)c->fine_nspec * cube_spec_pitch((int)c->nf) * c->nf * c->nf * sizeof(float2))
    return CUBE_ERR_INVALID;
  return CUBE_OK;
}

/* Deposit, contrast, forward FFT and binning on the context's stream: the
   FP64 mesh in the force spectra, the FP32 delta in rho_f, its spectrum in
   spec_f; bins = 3*(nbin+1) doubles (sum |k|, sum Delta^2, count), cleared
   here by a fill kernel. */
static int power_device(cube_ctx *c, int32_t ng, int32_t nbin, float2 *tw, double *bins) {
  const int64_t n3 = (int64_t)ng * ng * ng;
  double *rho = reinterpret_cast<double *>(c->specd_f);
  hipLaunchKernelGGL(k_pw_zero, dim3(1024), dim3(256), 0, c->stream, rho, n3);
  hipLaunchKernelGGL(k_pw_zero, dim3(1), dim3(256), 0, c->stream, bins, (int64_t)3 * (nbin + 1));
  const int64_t nrows = c->ncells / c->nt;
  if (c->p.izipx == 1)
    hipLaunchKernelGGL(k_pw_deposit<1>, dim3((unsigned)nrows), dim3(PW_T), 0, c->stream, (const int8_t *)c->xp,
                       c->cstart, geo_of(c), ng, rho);
  else
    hipLaunchKernelGGL(k_pw_deposit<2>, dim3((unsigned)nrows), dim3(PW_T), 0, c->stream, (const int16_t *)c->xp,
                       c->cstart, geo_of(c), ng, rho);
  hipLaunchKernelGGL(k_pw_delta, dim3(1024), dim3(256), 0, c->stream, rho, n3, (double)n3 / (double)c->nplocal,
                     c->rho_f);
  if (hipGetLastError() != hipSuccess) return CUBE_ERR_HIP;
  int st = cube_fft_forward_full_tw(c, ng, tw, c->rho_f, c->spec_f);
  if (st) return st;
  const size_t lds = (size_t)(nbin + 1) * (2 * sizeof(double) + sizeof(unsigned long long));
  hipLaunchKernelGGL(k_pw_bin, dim3(512), dim3(PW_T), lds, c->stream, c->spec_f, ng, nbin, bins, bins + (nbin + 1),
                     (unsigned long long *)(bins + 2 * (nbin + 1)));
  return hipGetLastError() != hipSuccess ? CUBE_ERR_HIP : CUBE_OK;
}

extern "C" int cube_power(cube_ctx *c, int32_t ng, int32_t nbin, double *ksum, double *psum, int64_t *count) {
  if (c) c->pm_ahead = 0; /* forces formed ahead by ext_pp_force no longer apply */
  int st = power_args(c, ng, nbin, ksum, psum, count);
  if (st) return st;
  CUBE_CHECK(hipSetDevice(c->device));
  /* the bin sums in the force mesh, the twiddles in a plan of their own */
  double *bins = reinterpret_cast<double *>(c->force_f);
  if ((size_t)(3 * (nbin + 1)) * sizeof(double) > (size_t)3 * c->nf * c->nf * c->nf * sizeof(float))
    return CUBE_ERR_INVALID;
  FFTPlan *plan = nullptr;
  st = cube_fft_create(c, ng, &plan);
  if (st) return st;
  st = power_device(c, ng, nbin, cube_fft_plan_twiddles(plan), bins);
  if (!st) {
    double *h = new double[3 * (nbin + 1)];
    if (hipMemcpyAsync(h, bins, 3 * (nbin + 1) * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      st = CUBE_ERR_HIP;
    if (!st) {
      const unsigned long long *hc = reinterpret_cast<const unsigned long long *>(h + 2 * (nbin + 1));
      for (int b = 1; b <= nbin; ++b) {
        ksum[b - 1] = h[b];
        psum[b - 1] = h[nbin + 1 + b];
        count[b - 1] = (int64_t)hc[b];
      }
    }
    delete[] h;
  }
  (void)hipStreamSynchronize(c->stream);
  cube_fft_destroy(plan);
  return st;
}

/* ---------------- cube_power_noalloc ----------------
 * The same spectrum without any allocation in the call: the twiddle tables
 * of every built FFT size and the bin sums live in one file-scope device
 * array.  Host code never passes the array's name to a kernel or a copy
 * (on the host it is a stand-in, not a device address): the address of
 * each device's copy is taken once, at context creation on that device, and
 * kept per device.  Calls that overlap on one device share the array and
 * race; each call clears the bins it uses before the deposit. */
#define PW_TW_MAX 5632   /* float2 twiddles of all built sizes (5362) */
#define PW_NBIN_MAX 1774 /* nint(2048/2 * sqrt(3)): bins of the largest mesh */
#define PW_SCRATCH_DOUBLES (PW_TW_MAX + 3 * (PW_NBIN_MAX + 1))
#define PW_MAX_DEV 64

static __device__ __align__(256) double scratch[PW_SCRATCH_DOUBLES];

static std::mutex g_scratch_mu;
static double *g_scratch[PW_MAX_DEV];  /* device address of scratch, per device */
static int g_tw_off[2049];             /* float2 offset of size n's twiddles */

/* Once per device, with the device current: find scratch's address and
   write the twiddles of every built size into it. */
int cube_power_scratch_setup(int device) {
  if (device < 0 || device >= PW_MAX_DEV) return CUBE_ERR_INVALID;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return CUBE_OK;
  double *ptr = nullptr;
  if (hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return CUBE_ERR_HIP;
  float2 h[2048];
  int off = 0;
  for (int n = 1; n <= 2048; ++n) {
    if (!cube_fft_supported(n)) continue;
    if (off + n > PW_TW_MAX) return CUBE_ERR_INVALID;
    cube_fft_twiddles(n, h);
    if (hipMemcpy(reinterpret_cast<float2 *>(ptr) + off, h, n * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess)
      return CUBE_ERR_HIP;
    g_tw_off[n] = off;
    off += n;
  }
  g_scratch[device] = ptr;
  return CUBE_OK;
}

extern "C" int cube_power_noalloc(cube_ctx *c, int32_t ng, int32_t nbin, double *ksum, double *psum,
                                  int64_t *count) {
  if (c) c->pm_ahead = 0;
  int st = power_args(c, ng, nbin, ksum, psum, count);
  if (st) return st;
  if (nbin > PW_NBIN_MAX || ng > 2048) return CUBE_ERR_INVALID; /* beyond the scratch array */
  double *ptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ptr = (c->device >= 0 && c->device < PW_MAX_DEV) ? g_scratch[c->device] : nullptr;
  }
  if (!ptr) return CUBE_ERR_STATE; /* set up by cube_ctx_create on this device */
  CUBE_CHECK(hipSetDevice(c->device));
  float2 *tw = reinterpret_cast<float2 *>(ptr) + g_tw_off[ng];
  double *bins = ptr + PW_TW_MAX;
  st = power_device(c, ng, nbin, tw, bins);
  /* bins 1..nbin straight into the caller's arrays (the counts are
     unsigned 64-bit, the same bits as int64 below 2^63) */
  if (!st && (hipMemcpyAsync(ksum, bins + 1, nbin * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
              hipMemcpyAsync(psum, bins + (nbin + 1) + 1, nbin * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream) != hipSuccess ||
              hipMemcpyAsync(count, bins + 2 * (nbin + 1) + 1, nbin * sizeof(int64_t), hipMemcpyDeviceToHost,
                             c->stream) != hipSuccess))
    st = CUBE_ERR_HIP;
  if (hipStreamSynchronize(c->stream) != hipSuccess && !st) st = CUBE_ERR_HIP;
  return st;
}
