/*
 * cube_internal.h -- context layout and launch helpers of libcube_hip.so.
 *
 * Data layout in HBM (one image = the whole box on one GPU):
 *  - particles: the reference's *disjoint* state only (no buffer copies);
 *    cell-sorted in cell order cid = i + nt*(j + nt*(k + nt*tile)), tile =
 *    tx + nnt*(ty + nnt*tz), with xp/vp as 3 interleaved izipx/izipv-byte
 *    integers per particle and pid as int32;  cstart[cid] = first particle
 *    of the cell (int64, exclusive prefix sum of rhoc).  A buffer cell of a
 *    tile is read straight from the periodic neighbour's physical cell, so
 *    buffer_grid/buffer_x/buffer_v reduce to index arithmetic.
 *  - fine mesh: ONE periodic nf^3 mesh for the whole box (nf = 4 nc),
 *    x fastest; its half spectrum nf x nf rows of nf/2+1 complex, padded
 *    to cube_spec_pitch(nf).
 *  - coarse mesh nc^3 and its half spectrum.
 */
#ifndef CUBE_INTERNAL_H
#define CUBE_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include "../../include/cube.h"
#include "cube_math.h"

#define CUBE_CHECK(x)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) {                                                             \
      fprintf(stderr, "cube: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return CUBE_ERR_HIP;                                                              \
    }                                                                                   \
  } while (0)
#define CUBE_LAUNCH_CHECK() CUBE_CHECK(hipGetLastError())

/* key of a particle inside its destination cell, see cube_drift.hip */
#define KEY_PBITS 40
#define KEY_PMASK ((1ull << KEY_PBITS) - 1ull)

struct FFTPlan;  /* cube_fft.hip */

#define CUBE_TEV 64 /* stage timer event pairs held before they are resolved */

struct cube_ctx {
  cube_params p;
  int device;
  hipStream_t stream;
  /* geometry */
  int64_t nc, nnt, nt, ncb, ncell, nf, nh, ncells, ntiles;
  int64_t np_image, np_image_max, np_tile_max, np_pp_max, cap;
  int32_t nvbin;
  /* particle state (disjoint, cell-sorted) and its double buffer */
  void *xp, *vp, *xp2, *vp2;
  int32_t *pid, *pid2;
  int32_t *rhoc, *rhoc2;
  float *vfield, *vfield2;
  int64_t *cstart, *cstart2;  /* ncells+1 */
  int64_t nplocal, npglobal;
  int has_state;
  /* drift scratch */
  int32_t *dest;            /* cap: destination cell of each particle, -1 lost */
  uint64_t *key;            /* cap: order key inside the destination cell      */
  uint64_t *sortbuf;        /* cap + extra */
  int32_t *fill;            /* ncells */
  int32_t *xdest;           /* extra copies (duplicates in the reference)      */
  uint64_t *xkey;
  int32_t *rank, *xrank;    /* slot of each (extra) copy inside its new cell, arrival order */
  uint8_t *fbflag;          /* per finalize block: 1 = sorted and finished in LDS */
  int64_t xcap;
  /* codec tables */
  float *tanf_tab;          /* nvbin: tanf(pi*code/(nvbin-1)), index code+nvbin/2 */
  double *venc_tab;         /* nvbin/2: velocity code boundaries, cube_encode_v_tab */
  double *vdec0, *vdec1;    /* nvbin: decoded residual velocity for two sigmas  */
  /* fine mesh */
  float *rho_f;             /* nf^3 */
  float2 *spec_f;           /* pitch*nf*nf */
  float2 *specd_f;          /* fine_nspec * pitch*nf*nf */
  int fine_nspec;           /* force spectra held: 3, or 1 in the compact layout */
  int scratch_in_force;     /* the drift's sort scratch lives inside force_f */
  int32_t *ppdefer;         /* ncells/64: PP blocks deferred to the cell path */
  int32_t *ppitems;         /* ncells/64 + 1: their work items, then prefix sums */
  float *force_f;           /* 3*nf^3, (d fastest) */
  float *kern_f;            /* octant in 16-column blocks, cube_kernf_index */
  /* coarse mesh */
  float *rho_c;
  float2 *spec_c, *specd_c;
  float *force_c;           /* 3*nc^3 */
  float *kern_c;            /* 3*nc*nc*(nc/2+1) */
  int has_kernels;
  /* small device scratch for reductions */
  double *dred;             /* 8192 doubles: [0,4096) reductions of work on `stream`,
                               [4096,8192) of work on `side` */
  uint64_t *ured;           /* 64 words */
  int64_t *ired;            /* 4096 */
  int64_t *scanbuf;         /* ncells/1024 + 2 block sums of the scan */
  double *vpart;            /* 3 per finalize block of the drift (velocity sums) */
  /* host-side persistent state */
  float pp_ftemp;           /* running max of W/ext_pp_force.f90:123 */
  FFTPlan *plan_f, *plan_c;
  /* the PM force pipelines run on side streams beside other work: the
     coarse one beside the fine one in particle_mesh, and both beside PP
     when ext_pp_force computes them ahead (pm_ahead, for mass pm_mass) */
  hipStream_t side, side2;
  hipEvent_t ev_fork, ev_join, ev_join2;
  int pm_ahead;
  float pm_mass, pm_amid, pm_dt; /* force_f then holds the kick factors of this a_mid, dt */
  unsigned long long *slots; /* CUBE_NSLOTREG regions of CUBE_SLOTS words, see cube_slots_fold */
  /* stage timers: event pairs recorded without waiting, resolved when the
     ring is full or the timers are read */
  int profiling;
  hipEvent_t tev[2 * CUBE_TEV];
  int tid[CUBE_TEV];
  int tused, topen;
  double tms[CUBE_NTIMERS];
  int64_t tcalls[CUBE_NTIMERS];
};

/* add the finished event pairs to the stage totals (waits for them) */
static inline void cube_timers_resolve(cube_ctx *c) {
  for (int k = 0; k < c->tused; ++k) {
    (void)hipEventSynchronize(c->tev[2 * k + 1]);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->tev[2 * k], c->tev[2 * k + 1]);
    c->tms[c->tid[k]] += ms;
    c->tcalls[c->tid[k]] += 1;
  }
  c->tused = 0;
}
/* stage timing on the stream the stage runs on (no-ops unless profiling);
   host code opens at most one stage at a time */
static inline void tbeg(cube_ctx *c) {
  if (!c->profiling) return;
  if (c->tused == CUBE_TEV) cube_timers_resolve(c);
  c->topen = c->tused++;
  (void)hipEventRecord(c->tev[2 * c->topen], c->stream);
}
static inline void tend(cube_ctx *c, int id) {
  if (!c->profiling) return;
  c->tid[c->topen] = id;
  (void)hipEventRecord(c->tev[2 * c->topen + 1], c->stream);
}

/* --- common device helpers --- */
struct Geo {
  int nt, nnt, nc, ncb;
  __device__ __host__ int64_t cid(int i, int j, int k, int t) const {
    return (int64_t)i + (int64_t)nt * ((int64_t)j + (int64_t)nt * ((int64_t)k + (int64_t)nt * t));
  }
};
static inline Geo geo_of(const cube_ctx *c) {
  Geo g;
  g.nt = (int)c->nt; g.nnt = (int)c->nnt; g.nc = (int)c->nc; g.ncb = (int)c->ncb;
  return g;
}
__device__ static inline int wrapi(int a, int n) {
  a %= n;
  return a < 0 ? a + n : a;
}
/* cell id from global 0-based coarse coordinates (periodic) */
__device__ static inline int64_t cid_global(const Geo &g, int gx, int gy, int gz) {
  gx = wrapi(gx, g.nc); gy = wrapi(gy, g.nc); gz = wrapi(gz, g.nc);
  const int t = gx / g.nt + g.nnt * (gy / g.nt + g.nnt * (gz / g.nt));
  return g.cid(gx % g.nt, gy % g.nt, gz % g.nt, t);
}

/* loads/stores of 1- or 2-byte particle components */
template <int B> struct IntOf;
template <> struct IntOf<1> { typedef int8_t T; };
template <> struct IntOf<2> { typedef int16_t T; };

/* inclusive prefix sum over the 64 lanes of a wave (shuffles, no LDS) */
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    v += lane >= o ? u : 0;
  }
  return v;
}
/* inclusive prefix sum over a workgroup of NW waves (one barrier; `tot`
   gets the workgroup total; wsum is NW ints of LDS) */
template <int NW> __device__ __forceinline__ int block_incl_scan(int v, int *wsum, int &tot) {
  const int w = threadIdx.x >> 6;
  v = wave_incl_scan(v);
  if ((threadIdx.x & 63) == 63) wsum[w] = v;
  __syncthreads();
  int base = 0;
  tot = 0;
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    base += q < w ? wsum[q] : 0;
    tot += wsum[q];
  }
  return v + base;
}

/* cube_encode_v through the code boundaries T[k] = tan((k+0.5)*pi/(nvbin-1))
   for k < kmax and T[kmax] = +inf (venc_tab, built on the host in long
   double; pi is the real(4) pi of the reference): the code of |u| is the
   number of boundaries <= |u| (nint rounds half away from zero), and atan
   being odd makes the negative side its mirror.  An FP32 atan, within 0.01
   of the code, picks the candidate and two table reads settle it.  This is
   nint((nvbin-1)*atan(u)/pi) in exact arithmetic; the FP64 formula rounds
   to the same integer except within about 1e-11 of a half-integer. */
template <int IV>
__device__ __forceinline__ int32_t cube_encode_v_tab(double c, double v, const double *__restrict__ T) {
  constexpr int kmax = IV == 1 ? 127 : 32767;
  constexpr float scale = (float)((IV == 1 ? 255.0 : 65535.0) / (double)CUBE_PI_F);
  const double u = c * v;
  const double a = fabs(u);
  int k0 = (int)(scale * atanf((float)a) + 0.5f);
  k0 = k0 < kmax ? k0 : kmax;
  const double lo = k0 > 0 ? T[k0 - 1] : -1.0;
  const double hi = T[k0];
  int k = k0 - (a < lo ? 1 : 0) + (a >= hi ? 1 : 0);
  k = k < 0 ? 0 : (k > kmax ? kmax : k);
  return cube_wrap(IV, u < 0 ? -k : k);
}

/* a/b for a divisor fixed at compile time, y = RN(1/b): q0 = RN(a*y) and
   Markstein's correction RN(q0 + (a - b*q0)*y) with both steps fused.  For
   b = 6 and b = pi in real(4) this equals the IEEE quotient bit for bit for
   every 1e-30 <= |a| < 1e30 (checked on the CPU over every float of that
   range by tests/test_divide_const.py). */
template <int B> struct FixedDiv;
template <> struct FixedDiv<6> {
  static constexpr float b = 6.0f, y = 1.0f / 6.0f;
};
template <> struct FixedDiv<314> {
  static constexpr float b = CUBE_PI_F, y = 1.0f / CUBE_PI_F;
};
template <int B> __device__ __forceinline__ float div_fixed(float a) {
  constexpr float b = FixedDiv<B>::b, y = FixedDiv<B>::y;
  const float q0 = a * y;
  const float r = __builtin_fmaf(-q0, b, a);
  return __builtin_fmaf(r, y, q0);
}
/* ((a/6)/pi) for the kicks: when 6e-30 <= |a| < 1e30 both quotients stay in
   the checked range (a/6 lands in [1e-30, 1.7e29)); `ok` is cleared
   otherwise and the caller redoes the sum with divisions. */
__device__ __forceinline__ float div_6pi(float a, bool &ok) {
  const float m = fabsf(a);
  ok = ok && m >= 6e-30f && m < 1e30f;
  return div_fixed<314>(div_fixed<6>(a));
}

/* Half spectra of an n^3 real mesh are stored as rows of n/2+1 complex
   values padded to a multiple of 16 (128 bytes), so every row and every
   16-column batch of the strided FFT passes starts on a cache line. */
__host__ __device__ constexpr int cube_spec_pitch(int n) { return ((n / 2 + 1) + 15) / 16 * 16; }
/* kern_f storage.  The fine Green's function is kept for the octant
   0 <= kx, ky, kz <= nf/2 (odd in k_d, the other octants follow by sign) as,
   for each component d and ky, blocks of 16 kx columns that are each one
   contiguous [kz][16] run: element (d, kz, ky, kx) sits at
   (((d*ho + ky)*KB + kx/16)*ho + kz)*16 + kx%16, ho = nf/2+1, KB = ceil(ho/16).
   The z pass over 16 spectrum columns then streams its kernel values instead
   of reading 64-byte pieces of unaligned 513-float rows. */
__host__ __device__ constexpr int64_t cube_kernf_blocks(int64_t ho) { return (ho + 15) / 16; }
__host__ __device__ constexpr int64_t cube_kernf_elems(int64_t ho) { return 3 * ho * cube_kernf_blocks(ho) * ho * 16; }
__host__ __device__ constexpr int64_t cube_kernf_index(int64_t ho, int d, int kz, int ky, int kx) {
  return (((d * ho + ky) * cube_kernf_blocks(ho) + (kx >> 4)) * ho + kz) * 16 + (kx & 15);
}

/* Maxima and sums over a launch.  Atomics of many workgroups on one word
   serialise in the L2 (about 11 ns each: half a million workgroups wait
   about 6 ms), so a workgroup updates word blockIdx.x % CUBE_SLOTS of a
   region instead and cube_slots_fold folds the region afterwards (max and
   sum do not depend on the order). */
#define CUBE_SLOTS 256
enum { CUBE_SLOT_F2_FINE = 0, CUBE_SLOT_F2_COARSE = 1, CUBE_SLOT_PP = 2 /* f2, count */, CUBE_SLOT_VMAX = 4 /* x, y, z */,
       CUBE_NSLOTREG = 8 };
__device__ __forceinline__ unsigned long long *cube_slot(unsigned long long *region) {
  return region + blockIdx.x % CUBE_SLOTS;
}
int cube_slots_zero(cube_ctx *c, unsigned long long *region, int nreg);
/* region r folded into d64[r] (or the low 32 bits into d32[r]): max, or sum when add */
int cube_slots_fold(cube_ctx *c, const unsigned long long *region, int nreg, int add, unsigned long long *d64,
                    unsigned *d32);

/* launchers (one file each) */
int cube_drift(cube_ctx *c, float dt_old, float dt, float vz_max, float sigma_vi, cube_vstats *out);
int cube_vdec_table(cube_ctx *c, double *dst, float sigma);
int cube_scan_counts(cube_ctx *c, const int32_t *cnt, int64_t *start, int64_t n, int64_t *total);
int cube_tile_boxsum(cube_ctx *c, const int32_t *rhoc, int w, int64_t *host_out);

/* 1 when the line-FFT kernels are built for size n (12..2048, factors 2, 3, 5) */
int cube_fft_supported(int64_t n);
/* cube_power_noalloc's device scratch: address and twiddles, once per
   device, with that device current (cube_ctx_create calls it) */
int cube_power_scratch_setup(int device);
/* the n twiddles exp(-2 pi i m/n) every plan of size n holds (host) */
void cube_fft_twiddles(int n, float2 *h);
int cube_fft_create(cube_ctx *c, int n, FFTPlan **out);
void cube_fft_destroy(FFTPlan *p);
int cube_fft_r2c_3d(cube_ctx *c, FFTPlan *p, const float *in, float2 *out);
int cube_fft_green_c2r_3d(cube_ctx *c, FFTPlan *p, float2 *spec, const float *kern, int kern_octant,
                          float2 *work3, float *force3, float scale_div, float *f2max, int timer0,
                          unsigned *f2dev, unsigned long long *f2slots, const float *prescale = nullptr,
                          int ncomp_spec = 3);
int cube_fft_forward_full(cube_ctx *c, FFTPlan *p, const float *in, float2 *out);
float2 *cube_fft_plan_twiddles(FFTPlan *p);
/* the same with the size-n twiddles already on the device at tw (no plan) */
int cube_fft_forward_full_tw(cube_ctx *c, int n, float2 *tw, const float *in, float2 *out);
int cube_fft_inverse_full(cube_ctx *c, FFTPlan *p, float2 *spec, float *out);

int cube_fine_deposit(cube_ctx *c, float mass_p, double *testrhof);
/* device slots (in dred) of the deposits' mesh sums */
#define CUBE_DRED_FINE_SUM 4000
#define CUBE_DRED_COARSE_SUM (4096 + 4000)
int cube_fine_kick(cube_ctx *c, float a_mid, float dt, float sigma_vi, float sigma_vi_new, int sync);
int cube_coarse_deposit(cube_ctx *c, float mass_p, double *testrhoc);
int cube_coarse_kick(cube_ctx *c, float a_mid, float dt, float sigma_vi, float *vmax);
int cube_pm_kicks(cube_ctx *c, float a_mid, float dt, float sigma_vi, float sigma_vi_new, float *vmax);
int cube_green_init(cube_ctx *c, const float *wfxyzf, const float *wfxyzc);
/* phi of the current fine density into force_f[0, nf^3) (cube_mesh.hip) */
int cube_fine_potential(cube_ctx *c, const float *wg1);
int cube_pp(cube_ctx *c, float a_mid, float dt, float sigma_vi, float mass_p, float *f2_max_pp,
            int64_t *itest1);
int cube_vstats_run(cube_ctx *c, float sigma_vi, double *sums3);
__global__ void k_fill_zero_i32(int32_t *p, int64_t n);
__global__ void k_place_keys(const int32_t *dest, const uint64_t *key, int64_t n, const int64_t *start, int32_t *fill,
                             uint64_t *out);

#endif
