/*
 * cube_fft.hip -- 3D real FFTs of the periodic meshes with the Green's
 * function multiply fused in, replacing the reference's cube_fft
 * (W/cubefft.f90:15-25, FFTW r2c/c2r of the buffered fine tile) and
 * pencil_fft (W/pencil_fft.f90:31-60) for a single image.
 *
 * Conventions match FFTW's as the reference uses them: forward unnormalised
 * (e^{-i}), backward unnormalised (e^{+i}) followed by three divisions by N
 * in real(4) (W/pm.f90:123, W/pencil_fft.f90:58).
 *
 * Structure (mesh N^3, x fastest; half spectrum of NH=N/2+1 complex per
 * row, rows padded to cube_spec_pitch(N) so they start on cache lines):
 *   k_r2c_x   : rows along x, real N -> complex N/2 FFT + split -> NH
 *   k_c2c_r / k_c2c_s : lines along y or z, in place (three spectra in one
 *               launch for the inverse y pass)
 *   k_green_r / k_green_z : lines along z: forward FFT, then for d=0..2
 *               multiply by i*K_d(k) and inverse FFT, writing three spectra
 *   k_c2r_x3  : rows along x for the three spectra: merge -> complex N/2
 *               inverse FFT -> real, /N/N/N, write F as 3 floats per cell,
 *               and max |F|^2 (W/pm.f90:126,315)
 * Strided lines of power-of-two meshes 64..2048 (k_c2c_r, k_green_r) are
 * register-resident four-step FFTs over 16-column batches; other sizes and
 * the row passes use a Stockham autosort FFT in LDS (radix 8/4/2/3/5/7).
 */
#include "cube_internal.h"
#include <string.h>
#include <math.h>
#include <utility>

#define FFT_T 256

struct FFTPlan {
  int n;
  float2 *tw; /* n twiddles exp(-2 pi i m/n) */
};

/* ---------------- compile-time plans ---------------- */
__host__ __device__ constexpr int pick_radix(int rem) {
  return (rem % 8 == 0 && rem != 16) ? 8
         : (rem % 4 == 0)            ? 4
         : (rem % 2 == 0)            ? 2
         : (rem % 3 == 0)            ? 3
         : (rem % 5 == 0)            ? 5
         : (rem % 7 == 0)            ? 7
                                     : rem;
}
__host__ __device__ constexpr int radix_at(int L, int i) {
  int rem = L;
  for (int k = 0; k < i; ++k) rem /= pick_radix(rem);
  return rem > 1 ? pick_radix(rem) : 1;
}
__host__ __device__ constexpr int ns_at(int L, int i) { /* product of radices before pass i */
  int ns = 1, rem = L;
  for (int k = 0; k < i; ++k) {
    const int r = pick_radix(rem);
    ns *= r;
    rem /= r;
  }
  return ns;
}
__host__ __device__ constexpr int npass(int L) {
  int n = 0, rem = L;
  while (rem > 1) {
    rem /= pick_radix(rem);
    ++n;
  }
  return n;
}

/* ---------------- complex helpers ---------------- */
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
/* multiply by -i (forward, DIR=-1) or +i (inverse, DIR=+1) */
template <int DIR> __device__ __forceinline__ float2 rot(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int R, int DIR> struct DFT;
template <int DIR> struct DFT<2, DIR> {
  __device__ __forceinline__ static void run(float2 *v) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};
template <int DIR> struct DFT<4, DIR> {
  __device__ __forceinline__ static void run(float2 *v) {
    const float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const float2 t2 = cadd(v[1], v[3]), t3 = rot<DIR>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};
template <int DIR> struct DFT<8, DIR> {
  __device__ __forceinline__ static void run(float2 *v) {
    float2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    DFT<4, DIR>::run(e);
    DFT<4, DIR>::run(o);
    const float h = 0.70710678118654752f;
    /* W8^k = exp(DIR*2 pi i k/8) */
    const float2 w1 = make_float2(h, DIR * h), w3 = make_float2(-h, DIR * h);
    o[1] = cmul(o[1], w1);
    o[2] = rot<DIR>(o[2]);
    o[3] = cmul(o[3], w3);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = cadd(e[k], o[k]);
      v[k + 4] = csub(e[k], o[k]);
    }
  }
};
template <int DIR> struct DFT<3, DIR> {
  __device__ __forceinline__ static void run(float2 *v) {
    const float s = 0.86602540378443865f;
    const float2 a = v[0], b = v[1], c = v[2];
    const float2 sum = cadd(b, c);
    const float2 t = make_float2(a.x - 0.5f * sum.x, a.y - 0.5f * sum.y);
    const float2 d = csub(b, c);
    const float2 u = rot<DIR>(make_float2(s * d.x, s * d.y));
    v[0] = cadd(a, sum);
    v[1] = cadd(t, u);
    v[2] = csub(t, u);
  }
};
/* generic odd prime radix by direct summation */
template <int R, int DIR> struct DFTG {
  __device__ __forceinline__ static void run(float2 *v) {
    float2 out[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      float2 acc = v[0];
#pragma unroll
      for (int q = 1; q < R; ++q) {
        const int m = (k * q) % R;
        const float ang = DIR * 2.0f * 3.14159265358979f * (float)m / (float)R;
        acc = cadd(acc, cmul(v[q], make_float2(cosf(ang), sinf(ang))));
      }
      out[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = out[k];
  }
};
template <int DIR> struct DFT<5, DIR> : DFTG<5, DIR> {};
template <int DIR> struct DFT<7, DIR> : DFTG<7, DIR> {};

/*
 * One Stockham pass over a line of length L held in LDS at `buf` (unit
 * stride).  tl = thread index within the line, TL threads per line.
 * Pass with sub-transform size NS: butterfly j takes buf[j + q*L/R],
 * twiddles by exp(DIR 2 pi i k q/(NS R)), k=j%NS, and stores to
 * buf[(j/NS)*NS*R + k + q*NS].  Values are staged in registers so the pass
 * is in place; the caller's barrier pattern is inside.
 */
/* LDS index of element i of a line: one float2 of padding after every 32,
   so the Stockham stores of stride R and R*NS float2 spread over the 64
   banks of 4 bytes instead of piling on a few (the loads of consecutive
   elements stay conflict free); a line of L elements then spans
   L + L/32 + 1 float2 */
__device__ __forceinline__ int lpad(int i) { return i + (i >> 5); }
__host__ __device__ constexpr int lspan(int L) { return L + L / 32 + 1; }

template <int L, int R, int NS, int DIR, int TL>
__device__ __forceinline__ void stockham_pass(float2 *buf, int tl, const float2 *__restrict__ tw,
                                              int nt_tab) {
  constexpr int NB = L / R;
  constexpr int PER = (NB + TL - 1) / TL;
  float2 v[PER][R];
  /* the twiddles do not depend on the data: their (global, cached) loads
     are issued before the LDS reads and the barrier so their latency
     overlaps them */
  float2 wt[PER][R];
  const int twmul = nt_tab / (NS * R);
  if constexpr (NS > 1) {
#pragma unroll
    for (int b = 0; b < PER; ++b) {
      const int k = (tl + b * TL) % NS;
#pragma unroll
      for (int q = 1; q < R; ++q) wt[b][q] = tw[k * q * twmul];
    }
  }
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int j = tl + b * TL;
    if (PER * TL == NB || j < NB) {
#pragma unroll
      for (int q = 0; q < R; ++q) v[b][q] = buf[lpad(j + q * NB)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int j = tl + b * TL;
    if (PER * TL == NB || j < NB) {
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) {
          float2 w = wt[b][q];
          if (DIR > 0) w.y = -w.y;
          v[b][q] = cmul(v[b][q], w);
        }
      }
      DFT<R, DIR>::run(v[b]);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) buf[lpad(base + q * NS)] = v[b][q];
    }
  }
  __syncthreads();
}

template <int L, int DIR, int TL, int P>
__device__ __forceinline__ void fft_passes(float2 *buf, int tl, const float2 *tw, int nt_tab) {
  if constexpr (P < npass(L)) {
    stockham_pass<L, radix_at(L, P), ns_at(L, P), DIR, TL>(buf, tl, tw, nt_tab);
    fft_passes<L, DIR, TL, P + 1>(buf, tl, tw, nt_tab);
  }
}
/* full line FFT (natural order in and out) */
template <int L, int DIR, int TL>
__device__ __forceinline__ void line_fft(float2 *buf, int tl, const float2 *tw, int nt_tab) {
  fft_passes<L, DIR, TL, 0>(buf, tl, tw, nt_tab);
}

/* ---------------- batch sizes ---------------- */
template <int N> struct Cfg {
  static constexpr int L = N / 2;
  /* strided passes: B lines of length N */
  static constexpr int BS = N <= 512 ? 16 : (N <= 1024 ? 8 : 4);
  /* contiguous passes: B rows of complex length L */
  static constexpr int BC0 = 2048 / L;
  static constexpr int BC = BC0 < 1 ? 1 : (BC0 > 16 ? 16 : BC0);
};

/* ---------------- register-resident line FFTs ----------------
 * For power-of-two N = R1*R2 (64..2048) a strided line FFT keeps the line
 * in registers: a workgroup takes CB=16 adjacent columns (128-byte global
 * segments) and R1 threads per column, thread t holding x[t + R1*m],
 * m < R2.  Four-step: R2-point FFT in registers, twiddle by W_N^{t*q},
 * one transpose through LDS, R1-point FFTs in registers; the result comes
 * back in the same distribution (thread t holds X[t + R1*m]).  One LDS
 * round trip per line instead of one per radix pass, and 16-column
 * batches instead of 8 for the z lines.
 */
constexpr double ct_sin_q(double x) { /* Taylor series, x in [0, pi/2] */
  double term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double ct_cos_q(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
/* exp(DIR * 2 pi i E / M) rounded to float, exact at quarter turns */
template <int M, int E, int DIR> struct TwC {
  static constexpr int e = ((E % M) + M) % M;
  static constexpr int quad = (4 * e) / M;
  static constexpr double x = 1.57079632679489661923 * (double)(4 * e - quad * M) / (double)M;
  static constexpr double c = ct_cos_q(x), s = ct_sin_q(x);
  static constexpr double cr = quad == 0 ? c : quad == 1 ? -s : quad == 2 ? -c : s;
  static constexpr double ci = quad == 0 ? s : quad == 1 ? c : quad == 2 ? -s : -c;
  static constexpr float re = (float)cr;
  static constexpr float im = (float)(DIR * ci);
};

template <int M, int K, int DIR, int... Q>
__device__ __forceinline__ void tw_apply(float2 *a, std::integer_sequence<int, Q...>) {
  ((((K * Q) % M) != 0 ? (void)(a[Q] = cmul(a[Q], make_float2(TwC<M, K * Q, DIR>::re, TwC<M, K * Q, DIR>::im)))
                       : (void)0),
   ...);
}
/* butterfly J of one in-register Stockham pass (radix r, sub-size NS) */
template <int R, int DIR, int r, int NS, int J>
__device__ __forceinline__ void reg_bfly(const float2 *v, float2 *o) {
  constexpr int NB = R / r, k = J % NS;
  float2 a[r];
#pragma unroll
  for (int q = 0; q < r; ++q) a[q] = v[J + q * NB];
  if constexpr (NS > 1) tw_apply<NS * r, k, DIR>(a, std::make_integer_sequence<int, r>{});
  DFT<r, DIR>::run(a);
  constexpr int base = (J / NS) * NS * r + k;
#pragma unroll
  for (int q = 0; q < r; ++q) o[base + q * NS] = a[q];
}
template <int R, int DIR, int r, int NS, int... J>
__device__ __forceinline__ void reg_pass(const float2 *v, float2 *o, std::integer_sequence<int, J...>) {
  (reg_bfly<R, DIR, r, NS, J>(v, o), ...);
}
/* in-register FFT of R points, natural order in and out */
template <int R, int DIR, int P = 0>
__device__ __forceinline__ void reg_fft(float2 *v) {
  if constexpr (P < npass(R)) {
    float2 o[R];
    reg_pass<R, DIR, radix_at(R, P), ns_at(R, P)>(v, o, std::make_integer_sequence<int, R / radix_at(R, P)>{});
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = o[i];
    reg_fft<R, DIR, P + 1>(v);
  }
}

/* base + a byte offset known to fit 32 bits: lets the compiler keep the
   uniform base in scalar registers and the per-lane offset in one VGPR */
template <class T> __device__ __forceinline__ T *boff(T *base, uint32_t bytes) {
  return reinterpret_cast<T *>(reinterpret_cast<char *>(const_cast<void *>(static_cast<const void *>(base))) + bytes);
}
/* a wave-uniform pointer pinned to scalar registers; `salt` is an opaque
   zero that keeps separate uses (loads before an FFT, stores after it) from
   being merged into one long-lived vector address */
template <class T> __device__ __forceinline__ T *uni(T *p, int salt) {
  const uint64_t u = reinterpret_cast<uint64_t>(p) + (uint64_t)salt;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return reinterpret_cast<T *>(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ int opaque_zero() {
  int z = 0;
  asm volatile("s_mov_b32 %0, 0" : "=s"(z));
  return z;
}
/* the same zero, formed only once `after` has been computed: loads offset
   by it cannot be hoisted above the code that produces `after` */
__device__ __forceinline__ int opaque_zero_after(float after) {
  int z = 0;
  asm volatile("s_mov_b32 %0, 0 ; %1" : "=s"(z) : "v"(after));
  return z;
}

template <int N> struct RC {
  static constexpr bool ok = N == 64 || N == 128 || N == 256 || N == 512 || N == 1024 || N == 2048;
  static constexpr int R1 = N <= 128 ? 8 : (N <= 512 ? 16 : 32);
  static constexpr int R2 = N / R1;
  /* 16 columns (128-byte pieces) per workgroup; at N=2048 the transpose
     buffer is then 135 KB, one 512-thread workgroup per CU */
  static constexpr int CB = 16;
  static constexpr int NT = CB * R1;
  /* LDS transpose buffer S[c][q][t]: q stride R1+1, c stride = 64/CB
     (mod 64) so both the row writes and the column reads of a wave hit
     distinct banks */
  static constexpr int SQ = R1 + 1;
  static constexpr int SC0 = R2 * SQ;
  static constexpr int SC = SC0 + ((64 / CB - SC0) % 64 + 64) % 64;
  /* the twiddle table W_N^e, e < N, staged in LDS (up to 8 KB); at N=2048
     it is read from global memory (cached) to leave LDS to the transpose */
  static constexpr bool TWL = N <= 1024;
  /* the Green z-pass forms one force component per launch (the forward
     transform kept in the registers it is done in: 64 values per thread) */
  static constexpr bool ONE = N >= 2048;
};

/* line FFT of the CB columns of a workgroup; on entry v[m] = x[t + R1 m] of
   column c, on exit v[m] = X[t + R1 m].  twl: W_N^e, e < N, in LDS. */
template <int N, int DIR>
__device__ __forceinline__ void blk_fft(float2 *v, int c, int t, float *S, const float2 *twl) {
  using G = RC<N>;
  constexpr int R1 = G::R1, R2 = G::R2, J = R2 / R1, SQ = G::SQ;
  reg_fft<R2, DIR>(v);
  /* twiddles W_N^{t q} = W_N^{8 t qh} W_N^{t ql}, q = 8 qh + ql: 7 + R2/8-1
     table reads instead of R2-1 (fewer live registers, one extra rounding).
     The opaque zero offset stops the reads of successive calls in one
     kernel from being merged and held in registers between them. */
  const int t0 = opaque_zero();
  constexpr int QL = R2 < 8 ? R2 : 8, QH = R2 / QL;
  float2 lo[QL];
  lo[0] = make_float2(1.f, 0.f);
#pragma unroll
  for (int ql = 1; ql < QL; ++ql) {
    lo[ql] = twl[t0 + t * ql];
    if (DIR > 0) lo[ql].y = -lo[ql].y;
  }
#pragma unroll
  for (int qh = 0; qh < QH; ++qh) {
    float2 hi = make_float2(1.f, 0.f);
    if (qh > 0) {
      hi = twl[t0 + QL * t * qh];
      if (DIR > 0) hi.y = -hi.y;
    }
#pragma unroll
    for (int ql = 0; ql < QL; ++ql) {
      const int q = qh * QL + ql;
      if (q == 0) continue;
      const float2 w = qh == 0 ? lo[ql] : (ql == 0 ? hi : cmul(hi, lo[ql]));
      v[q] = cmul(v[q], w);
    }
  }
  float *Sc = S + c * G::SC;
  float2 a[R2];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < R2; ++q) Sc[q * SQ + t] = v[q].x;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < J; ++j)
#pragma unroll
    for (int p = 0; p < R1; ++p) a[j * R1 + p].x = Sc[(t + R1 * j) * SQ + p];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < R2; ++q) Sc[q * SQ + t] = v[q].y;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < J; ++j)
#pragma unroll
    for (int p = 0; p < R1; ++p) a[j * R1 + p].y = Sc[(t + R1 * j) * SQ + p];
#pragma unroll
  for (int j = 0; j < J; ++j) reg_fft<R1, DIR>(a + j * R1);
#pragma unroll
  for (int j = 0; j < J; ++j)
#pragma unroll
    for (int p = 0; p < R1; ++p) v[j + J * p] = a[j * R1 + p];
}

/* in-place strided line FFTs, register version: element (col, n) of line
   set (blockIdx.y, blockIdx.z) at a + z*cstride + y*ostride + n*stride + col.
   Requires R1*stride*8 < 2^32 (true for every mesh up to 2048^3). */
template <int N, int DIR>
__global__ __launch_bounds__(RC<N>::NT) void k_c2c_r(float2 *__restrict__ a, const float2 *__restrict__ tw,
                                                   int ncol, int64_t stride, int64_t ostride, int64_t cstride) {
  using G = RC<N>;
  constexpr int R1 = G::R1, R2 = G::R2;
  __shared__ float S[G::CB * G::SC];
  __shared__ float2 twls[G::TWL ? N : 1];
  const int tid = threadIdx.x, c = tid % G::CB, t = tid / G::CB;
  if constexpr (G::TWL)
    for (int e = tid; e < N; e += G::NT) twls[e] = tw[e];
  const float2 *twl = G::TWL ? twls : tw;
  const int col = blockIdx.x * G::CB + c;
  const bool ok = col < ncol;
  /* columns past the end compute on a copy of the last one and store nothing */
  const uint32_t lo = (uint32_t)(((int64_t)t * stride + (ok ? col : ncol - 1)) * 8);
  float2 *blk = a + (int64_t)blockIdx.z * cstride + (int64_t)blockIdx.y * ostride;
  float2 v[R2];
  const int s0 = opaque_zero();
#pragma unroll
  for (int m = 0; m < R2; ++m) v[m] = *boff(uni(blk + (int64_t)(R1 * m) * stride, s0), lo);
  __syncthreads();
  blk_fft<N, DIR>(v, c, t, S, twl);
  if (ok) {
    const int s1 = opaque_zero();
#pragma unroll
    for (int m = 0; m < R2; ++m) *boff(uni(blk + (int64_t)(R1 * m) * stride, s1), lo) = v[m];
  }
}

/* Green's function pass along z, register version of k_green_z; OCT selects
   the blocked octant kernel storage (fine mesh, cube_kernf_index) over the
   full one (coarse mesh).  A workgroup does the lines y and N-y, which read
   the same kernel block, for two adjacent column blocks. */
template <int N, bool OCT>
__global__ __launch_bounds__(RC<N>::NT) void k_green_r(const float2 *spec, float2 *outd,
                                                     const float *__restrict__ kern,
                                                     const float2 *__restrict__ tw, int d0, int d1) {
  using G = RC<N>;
  constexpr int NH = N / 2 + 1, H = N / 2, R1 = G::R1, R2 = G::R2, H1 = H + 1, SP = cube_spec_pitch(N);
  constexpr int64_t stride = (int64_t)N * SP, plane = (int64_t)SP * N * N;
  /* the forward z spectrum is needed by all three inverse transforms; at
     N=1024 (512 threads, 256 registers each) its upper half waits in LDS
     instead of registers, which would otherwise spill */
  static_assert(H % R1 == 0, "the octant offsets below take the side of H from m alone");
  constexpr int KL = N == 1024 ? R2 / 2 : 0;
  __shared__ float S[G::CB * G::SC];
  __shared__ float2 twls[G::TWL ? N : 1];
  __shared__ float2 keep[KL > 0 ? KL * G::NT : 1];
  const int tid = threadIdx.x, c = tid % G::CB, t = tid / G::CB;
  if constexpr (G::TWL)
    for (int e = tid; e < N; e += G::NT) twls[e] = tw[e];
  const float2 *twl = G::TWL ? twls : tw;
  /* one component per launch: the loop below runs once, so the forward
     spectrum v dies as w is formed from it and the two share registers */
  if constexpr (G::ONE) d1 = d0 + 1;
  const int yp = blockIdx.y;
  const int ny = (yp == 0 || yp == H) ? 1 : 2;
#pragma unroll 1
  for (int cb = 0; cb < 2; ++cb) {
#pragma unroll 1
    for (int iy = 0; iy < ny; ++iy) {
      const int y = iy == 0 ? yp : N - yp;
      const int kx0 = (2 * blockIdx.x + cb) * G::CB + c;
      const bool ok = kx0 < NH;
      const int kx = ok ? kx0 : NH - 1;
      const uint32_t lo = (uint32_t)(((int64_t)t * stride + kx) * 8);
      const float2 *src = spec + (int64_t)y * SP;
      float2 v[R2];
      const int s0 = opaque_zero();
      __syncthreads(); /* the previous line's keep/S reads are done */
#pragma unroll
      for (int m = 0; m < R2; ++m) v[m] = *boff(uni(src + (int64_t)(R1 * m) * stride, s0), lo);
      blk_fft<N, -1>(v, c, t, S, twl);
#pragma unroll
      for (int m = R2 - KL; m < R2; ++m) keep[(m - (R2 - KL)) * G::NT + tid] = v[m];
      const int ky = y <= H ? y : N - y;
      const float sy = y <= H ? 1.f : -1.f;
#pragma unroll 1
      for (int d = d0; d < d1; ++d) {
        /* at N=2048 the kernel loads wait for the forward transform, whose
           peak register use they would otherwise add to */
        const int z0 = G::ONE ? opaque_zero_after(v[R2 - 1].x) : opaque_zero();
        float2 w[R2];
#pragma unroll
        for (int m = 0; m < R2; ++m) {
          const int n = t + R1 * m;
          const float2 vm = m < R2 - KL ? v[m] : keep[(m - (R2 - KL)) * G::NT + tid];
          float K;
          if constexpr (OCT) {
            constexpr int KB = (int)cube_kernf_blocks(H1);
            /* kz = n for n <= H, N - n above; with H a multiple of R1 the
               side is fixed by m alone (n = H only at m = H/R1, t = 0, where
               both forms agree).  The part of the offset that depends on m
               goes into the uniform base, so the 64 loads share two
               per-lane offsets instead of holding 64 */
            const float *kb = kern + ((int64_t)d * H1 + ky) * KB * H1 * 16;
            const uint32_t kxo = (uint32_t)(((kx >> 4) * H1 * 16 + (kx & 15)) * 4);
            if (R1 * m + R1 - 1 <= H)
              K = *boff(uni(kb + 16 * R1 * m, z0), kxo + (uint32_t)(64 * t));
            else
              K = *boff(uni(kb + 16 * (N - R1 + 1 - R1 * m), z0), kxo + (uint32_t)(64 * (R1 - 1 - t)));
            if (d == 1) K *= sy;
            if (d == 2 && n > H) K = -K;
          } else { /* element ((d*N + n)*N + y)*NH + kx, n = t + R1*m: m in the base */
            const float *kb = uni(kern + (((int64_t)d * N + R1 * m) * N + y) * NH, z0);
            K = *boff(kb, (uint32_t)(((int64_t)t * N * NH + kx) * 4));
          }
          /* rho_f(::2)=-crho(2::2)*K, rho_f(2::2)=crho(::2)*K  (W/pm.f90:119-120) */
          w[m] = make_float2((-vm.y) * K, vm.x * K);
        }
        blk_fft<N, +1>(w, c, t, S, twl);
        if (ok) {
          float2 *dst = outd + (d - d0) * plane + (int64_t)y * SP;
          const int s1 = opaque_zero();
#pragma unroll
          for (int m = 0; m < R2; ++m) *boff(uni(dst + (int64_t)(R1 * m) * stride, s1), lo) = w[m];
        }
      }
    }
  }
}

/* ---------------- kernels ---------------- */
/* real rows (N floats) -> half spectra rows (NH complex) */
template <int N>
__global__ __launch_bounds__(FFT_T) void k_r2c_x(const float *__restrict__ in, float2 *__restrict__ out,
                                               const float2 *__restrict__ tw, int64_t nrows) {
  constexpr int L = N / 2, B = Cfg<N>::BC, TL = FFT_T / B, LP = lspan(L);
  __shared__ float2 sm[B * LP];
  const int tid = threadIdx.x, line = tid / TL, tl = tid % TL;
  const int64_t row0 = (int64_t)blockIdx.x * B;
  const float2 *in2 = reinterpret_cast<const float2 *>(in);
  /* every load of the rows in flight before the first LDS store */
  constexpr int IT = (B * L + FFT_T - 1) / FFT_T;
  float2 ld[IT];
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int idx = tid + i * FFT_T, l = idx / L, m = idx % L;
    ld[i] = ((B * L) % FFT_T == 0 || idx < B * L) && row0 + l < nrows ? in2[(row0 + l) * L + m]
                                                                     : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int idx = tid + i * FFT_T, l = idx / L, m = idx % L;
    if ((B * L) % FFT_T == 0 || idx < B * L) sm[l * LP + lpad(m)] = ld[i];
  }
  __syncthreads();
  line_fft<L, -1, TL>(sm + line * LP, tl, tw, N);
  /* split: X[k] = (Z[k]+conj(Z[L-k]))/2 + W^k (Z[k]-conj(Z[L-k]))/(2i) */
  for (int idx = tid; idx < B * (L + 1); idx += FFT_T) {
    const int l = idx / (L + 1), k = idx % (L + 1);
    const float2 zk = sm[l * LP + lpad((k == L ? 0 : k))];
    const float2 zc = sm[l * LP + lpad((k == 0 ? 0 : L - k))];
    const float2 e = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y - zc.y));
    const float2 o = make_float2(0.5f * (zk.y + zc.y), -0.5f * (zk.x - zc.x));
    const float2 w = tw[k];
    const float2 wo = cmul(w, o);
    if (row0 + l < nrows) out[(row0 + l) * cube_spec_pitch(N) + k] = cadd(e, wo);
  }
}

/* in-place strided line FFTs: element (col, n) at base + n*stride + col,
   col in [0, ncol); grid.x = column blocks, grid.y = outer index with
   outer stride `ostride` */
template <int N, int DIR>
__global__ __launch_bounds__(FFT_T) void k_c2c_s(float2 *__restrict__ a, const float2 *__restrict__ tw,
                                               int ncol, int64_t stride, int64_t ostride) {
  constexpr int B = Cfg<N>::BS, TL = FFT_T / B, NP = lspan(N);
  __shared__ float2 sm[B * NP];
  const int tid = threadIdx.x, line = tid / TL, tl = tid % TL;
  const int col0 = blockIdx.x * B;
  float2 *base = a + (int64_t)blockIdx.y * ostride + col0;
  /* every load in flight before the first LDS store */
  constexpr int IT = (B * N + FFT_T - 1) / FFT_T;
  float2 ld[IT];
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int idx = tid + i * FFT_T, l = idx % B, n = idx / B;
    ld[i] = ((B * N) % FFT_T == 0 || idx < B * N) && col0 + l < ncol ? base[n * stride + l] : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int idx = tid + i * FFT_T, l = idx % B, n = idx / B;
    if ((B * N) % FFT_T == 0 || idx < B * N) sm[l * NP + lpad(n)] = ld[i];
  }
  __syncthreads();
  line_fft<N, DIR, TL>(sm + line * NP, tl, tw, N);
  for (int idx = tid; idx < B * N; idx += FFT_T) {
    const int l = idx % B, n = idx / B;
    if (col0 + l < ncol) base[n * stride + l] = sm[l * NP + lpad(n)];
  }
}

/* Green's function pass along z.  spec: [z][y][x'] (x' < NH).  For each
   (x',y) line: forward FFT along z, then for d: out_d = IFFT_z(i K_d * .).
   kern: blocked octant storage of (N/2+1)^3 per component (octant=1, odd in
   the d-th wave number, cube_kernf_index) or full storage [d][kz][ky][kx] of
   N*N*NH (octant=0).  With one component, out may be spec itself: a
   workgroup reads its lines whole before it writes them (k_green_r
   likewise). */
template <int N>
__global__ __launch_bounds__(FFT_T) void k_green_z(const float2 *spec, float2 *outd,
                                                 const float *__restrict__ kern, int octant,
                                                 const float2 *__restrict__ tw, int d0, int d1) {
  constexpr int B = Cfg<N>::BS, TL = FFT_T / B, NP = lspan(N), NH = N / 2 + 1, H = N / 2, SP = cube_spec_pitch(N);
  constexpr int EPT = (B * N + FFT_T - 1) / FFT_T;
  __shared__ float2 sm[B * NP];
  const int tid = threadIdx.x, line = tid / TL, tl = tid % TL;
  const int col0 = blockIdx.x * B, y = blockIdx.y;
  const int64_t stride = (int64_t)N * SP;   /* between z planes */
  const int64_t plane = (int64_t)SP * N * N;
  const float2 *src = spec + (int64_t)y * SP + col0;
  for (int idx = tid; idx < B * N; idx += FFT_T) {
    const int l = idx % B, n = idx / B;
    sm[l * NP + lpad(n)] = (col0 + l < NH) ? src[n * stride + l] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  line_fft<N, -1, TL>(sm + line * NP, tl, tw, N);
  float2 keep[EPT];
#pragma unroll
  for (int s = 0; s < EPT; ++s) {
    const int idx = tid + s * FFT_T;
    if (idx < B * N) keep[s] = sm[(idx % B) * NP + lpad(idx / B)];
  }
  const int ky = y <= H ? y : N - y;
  const float sy = y <= H ? 1.f : -1.f;
  for (int d = d0; d < d1; ++d) {
    __syncthreads();
#pragma unroll
    for (int s = 0; s < EPT; ++s) {
      const int idx = tid + s * FFT_T;
      if (idx < B * N) {
        const int l = idx % B, n = idx / B, kx = col0 + l;
        float K = 0.f;
        if (kx < NH) {
          if (octant) {
            const int kz = n <= H ? n : N - n;
            const float sz = n <= H ? 1.f : -1.f;
            K = kern[cube_kernf_index(H + 1, d, kz, ky, kx)];
            if (d == 1) K *= sy;
            if (d == 2) K *= sz;
          } else {
            K = kern[(((int64_t)d * N + n) * N + y) * NH + kx];
          }
        }
        const float2 r = keep[s];
        /* rho_f(::2)=-crho(2::2)*K, rho_f(2::2)=crho(::2)*K  (W/pm.f90:119-120) */
        sm[l * NP + lpad(n)] = make_float2((-r.y) * K, r.x * K);
      }
    }
    __syncthreads();
    line_fft<N, +1, TL>(sm + line * NP, tl, tw, N);
    float2 *dst = outd + (d - d0) * plane + (int64_t)y * SP + col0;
    for (int idx = tid; idx < B * N; idx += FFT_T) {
      const int l = idx % B, n = idx / B;
      if (col0 + l < NH) dst[n * stride + l] = sm[l * NP + lpad(n)];
    }
  }
}

/* Z[k] = E + i O of rows row0..row0+B-1 of a half spectrum into the line
   buffers sm[l*LP + lpad(k)], k < N/2 (zero past nrows):
   E = X[k]+conj(X[L-k]), O = (X[k]-conj(X[L-k])) W^-k */
template <int N, int B>
__device__ __forceinline__ void c2r_merge_rows(const float2 *spec, float2 *sm, const float2 *__restrict__ tw,
                                               int64_t row0, int64_t nrows, int tid) {
  constexpr int L = N / 2, LP = lspan(L);
  for (int idx = tid; idx < B * L; idx += FFT_T) {
    const int l = idx / L, k = idx % L;
    float2 z = make_float2(0.f, 0.f);
    if (row0 + l < nrows) {
      const float2 a = spec[(row0 + l) * cube_spec_pitch(N) + k];
      const float2 cc = spec[(row0 + l) * cube_spec_pitch(N) + (L - k)];
      float2 w = tw[k];
      w.y = -w.y;
      const float2 e = make_float2(a.x + cc.x, a.y - cc.y);
      const float2 o = cmul(make_float2(a.x - cc.x, a.y + cc.y), w);
      z = make_float2(e.x - o.y, e.y + o.x);
    }
    sm[l * LP + lpad(k)] = z;
  }
}

/* half spectra rows -> real rows for the three components, scaled by
   three divisions by `div`; F written as 3 floats per cell; the slot region
   f2slots gets the max of (Fx^2+Fy^2)+Fz^2 as float bits. */
template <int N>
__global__ __launch_bounds__(FFT_T) void k_c2r_x3(const float2 *__restrict__ specd, float *__restrict__ force,
                                                const float2 *__restrict__ tw, int64_t nrows, float div,
                                                unsigned long long *__restrict__ f2slots, float pa, float pdt,
                                                int pre) {
  /* half the rows of the single-component kernels: three components in LDS
     at 2 rows per workgroup keep 6 workgroups per CU, which measures faster
     than 4 rows at 3 workgroups per CU despite idle threads in the passes,
     and faster than 4 rows one component at a time (every third float
     stored per pass: 11.4 against 8.7 ms at 1024^3) */
  constexpr int L = N / 2, BC = Cfg<N>::BC, B = BC >= 4 ? BC / 2 : BC, TL = FFT_T / B, LP = lspan(L);
  __shared__ float2 sm[3 * B * LP];
  __shared__ float red[FFT_T];
  /* a power-of-two divisor (every power-of-two mesh) has an exact inverse */
  const uint32_t db = __float_as_uint(div);
  const bool pow2 = (db & 0x807FFFFFu) == 0u && (db >> 23) > 0u && (db >> 23) < 200u;
  const float inv = 1.0f / div;
  const int tid = threadIdx.x, line = tid / TL, tl = tid % TL;
  const int64_t row0 = (int64_t)blockIdx.x * B;
  const int64_t plane = nrows * cube_spec_pitch(N);
  /* Z[k] = E + i O, E = X[k]+conj(X[L-k]), O = (X[k]-conj(X[L-k])) W^-k;
     every load of the three components in flight before the first store */
  constexpr int IT = (B * L + FFT_T - 1) / FFT_T;
  float2 xk[3][IT], xc[3][IT], wk[IT];
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int idx = tid + i * FFT_T, l = idx / L, k = idx % L;
    const bool in = ((B * L) % FFT_T == 0 || idx < B * L) && row0 + l < nrows;
    wk[i] = tw[in ? k : 0];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float2 *sp = specd + d * plane + (row0 + (in ? l : 0)) * cube_spec_pitch(N);
      xk[d][i] = sp[in ? k : 0];
      xc[d][i] = sp[in ? L - k : 0];
    }
  }
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int idx = tid + i * FFT_T, l = idx / L, k = idx % L;
    if ((B * L) % FFT_T != 0 && idx >= B * L) continue;
    const bool in = row0 + l < nrows;
    float2 w = wk[i];
    w.y = -w.y;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float2 z = make_float2(0.f, 0.f);
      if (in) {
        const float2 a = xk[d][i], c = xc[d][i];
        const float2 e = make_float2(a.x + c.x, a.y - c.y);
        const float2 o = cmul(make_float2(a.x - c.x, a.y + c.y), w);
        z = make_float2(e.x - o.y, e.y + o.x);
      }
      sm[(d * B + l) * LP + lpad(k)] = z;
    }
  }
  __syncthreads();
  for (int d = 0; d < 3; ++d) line_fft<L, +1, TL>(sm + (d * B + line) * LP, tl, tw, N);
  float m = 0.f;
  for (int idx = tid; idx < B * N; idx += FFT_T) {
    const int l = idx / N, x = idx % N;
    float f[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float2 z = sm[(d * B + l) * LP + lpad((x >> 1))];
      float v = (x & 1) ? z.y : z.x;
      if (pow2) { /* exact: v/2^k and v*2^-k round the same real number */
        v = v * inv;
        v = v * inv;
        v = v * inv;
      } else {
        v = v / div;
        v = v / div;
        v = v / div;
      }
      f[d] = v;
    }
    if (row0 + l < nrows) {
      struct F3 {
        float x, y, z;
      };
      float o[3] = {f[0], f[1], f[2]};
      if (pre) {
        /* the kick's factor of each term, (((F*a_mid)*dt)/6)/pi
           (W/pm.f90:146-153), formed once per mesh value instead of once
           per particle and corner; the IEEE quotient where div_6pi's
           range check fails */
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float a = (f[d] * pa) * pdt;
          bool ok = true;
          const float gq = div_6pi(a, ok);
          o[d] = ok ? gq : (a / 6.0f) / CUBE_PI_F;
        }
      }
      *reinterpret_cast<F3 *>(force + ((row0 + l) * N + x) * 3) = F3{o[0], o[1], o[2]}; /* one 12-byte store */
      const float f2 = (f[0] * f[0] + f[1] * f[1]) + f[2] * f[2];
      m = fmaxf(m, f2);
    }
  }
  red[tid] = m;
  __syncthreads();
  for (int s = FFT_T / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] = fmaxf(red[tid], red[tid + s]);
    __syncthreads();
  }
  if (tid == 0) atomicMax(cube_slot(f2slots), (unsigned long long)__float_as_uint(red[0]));
}

/* The compact fine mesh (one force spectrum at a time) forms the three
   components with the line FFTs and divisions of k_c2r_x3, so its cells end
   bit-identical to it, but stores components 0 and 1 as planes of their own
   (whole rows, no partial cache lines): k_c2r_xp writes one component's real
   rows, /div three times, at pitch opitch floats -- component 0 into the
   first third of the force mesh, component 1 in place over its own spectrum
   (each workgroup reads its rows whole before it writes them).  k_c2r_xf
   then transforms component 2 and writes the 3-float cells (with pre, the
   kick factors (((F*a_mid)*dt)/6)/pi) and max |F|^2, reading F0 and F1 back
   from those planes. */
template <int N>
__global__ __launch_bounds__(FFT_T) void k_c2r_xp(const float2 *spec, float *out, int64_t opitch,
                                                const float2 *__restrict__ tw, int64_t nrows, float div) {
  /* the rows per workgroup of k_c2r_x3, so the line FFTs are the same */
  constexpr int L = N / 2, BC = Cfg<N>::BC, B = BC >= 4 ? BC / 2 : BC, TL = FFT_T / B, LP = lspan(L);
  __shared__ float2 sm[B * LP];
  const uint32_t db = __float_as_uint(div);
  const bool pow2 = (db & 0x807FFFFFu) == 0u && (db >> 23) > 0u && (db >> 23) < 200u;
  const float inv = 1.0f / div;
  const int tid = threadIdx.x, line = tid / TL, tl = tid % TL;
  const int64_t row0 = (int64_t)blockIdx.x * B;
  c2r_merge_rows<N, B>(spec, sm, tw, row0, nrows, tid);
  __syncthreads();
  line_fft<L, +1, TL>(sm + line * LP, tl, tw, N);
  for (int idx = tid; idx < B * N; idx += FFT_T) {
    const int l = idx / N, x = idx % N;
    const float2 z = sm[l * LP + lpad((x >> 1))];
    float v = (x & 1) ? z.y : z.x;
    if (pow2) {
      v = v * inv;
      v = v * inv;
      v = v * inv;
    } else {
      v = v / div;
      v = v / div;
      v = v / div;
    }
    if (row0 + l < nrows) out[(row0 + l) * opitch + x] = v;
  }
}

/* Rows [rbeg, rend) of the last pass.  F0 lies in the first third of the
   force mesh, so a row's cells overwrite the F0 of rows 3r..3r+2: the host
   launches the rows in descending chunks [a, b) with 3a >= b, which only
   overwrite rows of earlier chunks, and every workgroup holds its own F0
   and F1 in registers before any of it stores (row 0's cells cover its own
   F0). */
template <int N>
__global__ __launch_bounds__(FFT_T) void k_c2r_xf(const float2 *__restrict__ spec, float *force, const float *F0,
                                                const float *__restrict__ F1, int64_t p1,
                                                const float2 *__restrict__ tw, int64_t rbeg, int64_t rend,
                                                float div, unsigned long long *__restrict__ f2slots, float pa,
                                                float pdt, int pre) {
  constexpr int L = N / 2, BC = Cfg<N>::BC, B = BC >= 4 ? BC / 2 : BC, TL = FFT_T / B, LP = lspan(L);
  constexpr int IT = (B * N + FFT_T - 1) / FFT_T;
  __shared__ float2 sm[B * LP];
  __shared__ float red[FFT_T];
  const uint32_t db = __float_as_uint(div);
  const bool pow2 = (db & 0x807FFFFFu) == 0u && (db >> 23) > 0u && (db >> 23) < 200u;
  const float inv = 1.0f / div;
  const int tid = threadIdx.x, line = tid / TL, tl = tid % TL;
  const int64_t row0 = rbeg + (int64_t)blockIdx.x * B;
  float f0[IT], f1[IT];
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int idx = tid + i * FFT_T, l = idx / N, x = idx % N;
    const bool in = ((B * N) % FFT_T == 0 || idx < B * N) && row0 + l < rend;
    f0[i] = in ? F0[(row0 + l) * N + x] : 0.f;
    f1[i] = in ? F1[(row0 + l) * p1 + x] : 0.f;
  }
  c2r_merge_rows<N, B>(spec, sm, tw, row0, rend, tid);
  __builtin_amdgcn_s_waitcnt(0); /* every F0 load of the workgroup has returned ... */
  __syncthreads();               /* ... before any of its stores */
  line_fft<L, +1, TL>(sm + line * LP, tl, tw, N);
  float m = 0.f;
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int idx = tid + i * FFT_T, l = idx / N, x = idx % N;
    if (((B * N) % FFT_T != 0 && idx >= B * N) || row0 + l >= rend) continue;
    const float2 z = sm[l * LP + lpad((x >> 1))];
    float v = (x & 1) ? z.y : z.x;
    if (pow2) {
      v = v * inv;
      v = v * inv;
      v = v * inv;
    } else {
      v = v / div;
      v = v / div;
      v = v / div;
    }
    const float f[3] = {f0[i], f1[i], v};
    float o[3] = {f[0], f[1], f[2]};
    if (pre) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const float a = (f[q] * pa) * pdt;
        bool ok = true;
        const float gq = div_6pi(a, ok);
        o[q] = ok ? gq : (a / 6.0f) / CUBE_PI_F;
      }
    }
    struct F3 {
      float x, y, z;
    };
    *reinterpret_cast<F3 *>(force + ((row0 + l) * N + x) * 3) = F3{o[0], o[1], o[2]};
    const float f2 = (f[0] * f[0] + f[1] * f[1]) + f[2] * f[2];
    m = fmaxf(m, f2);
  }
  red[tid] = m;
  __syncthreads();
  for (int q = FFT_T / 2; q > 0; q >>= 1) {
    if (tid < q) red[tid] = fmaxf(red[tid], red[tid + q]);
    __syncthreads();
  }
  if (tid == 0) atomicMax(cube_slot(f2slots), (unsigned long long)__float_as_uint(red[0]));
}

/* single-component c2r (no Green's function), for inverse_full */
template <int N>
__global__ __launch_bounds__(FFT_T) void k_c2r_x1(const float2 *__restrict__ spec, float *__restrict__ out,
                                                const float2 *__restrict__ tw, int64_t nrows) {
  constexpr int L = N / 2, B = Cfg<N>::BC, TL = FFT_T / B, LP = lspan(L);
  __shared__ float2 sm[B * LP];
  const int tid = threadIdx.x, line = tid / TL, tl = tid % TL;
  const int64_t row0 = (int64_t)blockIdx.x * B;
  for (int idx = tid; idx < B * L; idx += FFT_T) {
    const int l = idx / L, k = idx % L;
    float2 z = make_float2(0.f, 0.f);
    if (row0 + l < nrows) {
      const float2 xk = spec[(row0 + l) * cube_spec_pitch(N) + k];
      const float2 xc = spec[(row0 + l) * cube_spec_pitch(N) + (L - k)];
      const float2 e = make_float2(xk.x + xc.x, xk.y - xc.y);
      float2 w = tw[k];
      w.y = -w.y;
      const float2 o = cmul(make_float2(xk.x - xc.x, xk.y + xc.y), w);
      z = make_float2(e.x - o.y, e.y + o.x);
    }
    sm[l * LP + lpad(k)] = z;
  }
  __syncthreads();
  line_fft<L, +1, TL>(sm + line * LP, tl, tw, N);
  for (int idx = tid; idx < B * N; idx += FFT_T) {
    const int l = idx / N, x = idx % N;
    const float2 z = sm[l * LP + lpad((x >> 1))];
    if (row0 + l < nrows) out[(row0 + l) * N + x] = (x & 1) ? z.y : z.x;
  }
}

/* ---------------- host side ---------------- */
#define FFT_SIZES(X) \
  X(12) X(16) X(18) X(20) X(24) X(32) X(48) X(64) X(72) X(80) X(96) X(128) X(144) X(192) X(256) X(512) \
      X(576) X(1024) X(2048)

int cube_fft_supported(int64_t n) {
  switch (n) {
#define CASE(N) case N:
    FFT_SIZES(CASE)
#undef CASE
    return 1;
  }
  return 0;
}

void cube_fft_twiddles(int n, float2 *h) {
  for (int m = 0; m < n; ++m) {
    const double a = -2.0 * 3.14159265358979323846 * (double)m / (double)n;
    h[m] = make_float2((float)cos(a), (float)sin(a));
  }
}

int cube_fft_create(cube_ctx *c, int n, FFTPlan **out) {
  (void)c;
  if (!cube_fft_supported(n)) return CUBE_ERR_INVALID;
  FFTPlan *p = new FFTPlan;
  p->n = n;
  float2 *h = new float2[n];
  cube_fft_twiddles(n, h);
  if (hipMalloc(&p->tw, n * sizeof(float2)) != hipSuccess) {
    delete[] h;
    delete p;
    return CUBE_ERR_ALLOC;
  }
  CUBE_CHECK(hipMemcpy(p->tw, h, n * sizeof(float2), hipMemcpyHostToDevice));
  delete[] h;
  *out = p;
  return CUBE_OK;
}

void cube_fft_destroy(FFTPlan *p) {
  if (!p) return;
  (void)hipFree(p->tw);
  delete p;
}

template <int N> struct Launch {
  static int r2c(cube_ctx *c, FFTPlan *p, const float *in, float2 *out) {
    const int64_t rows = (int64_t)N * N;
    const int64_t nb = (rows + Cfg<N>::BC - 1) / Cfg<N>::BC;
    hipLaunchKernelGGL(k_r2c_x<N>, dim3((unsigned)nb), dim3(FFT_T), 0, c->stream, in, out, p->tw, rows);
    CUBE_LAUNCH_CHECK();
    return CUBE_OK;
  }
  /* ncomp line sets (component stride NH*N*N) of strided line FFTs */
  template <int DIR> static int strided(cube_ctx *c, FFTPlan *p, float2 *a, int64_t stride, int64_t ostride,
                                        int ncomp) {
    constexpr int NH = N / 2 + 1;
    const int64_t cstride = (int64_t)cube_spec_pitch(N) * N * N;
    if constexpr (RC<N>::ok) {
      dim3 g((NH + RC<N>::CB - 1) / RC<N>::CB, N, ncomp);
      hipLaunchKernelGGL((k_c2c_r<N, DIR>), g, dim3(RC<N>::NT), 0, c->stream, a, p->tw, NH, stride, ostride,
                         cstride);
      CUBE_LAUNCH_CHECK();
    } else {
      dim3 g((NH + Cfg<N>::BS - 1) / Cfg<N>::BS, N);
      for (int d = 0; d < ncomp; ++d) {
        hipLaunchKernelGGL((k_c2c_s<N, DIR>), g, dim3(FFT_T), 0, c->stream, a + d * cstride, p->tw, NH, stride,
                           ostride);
        CUBE_LAUNCH_CHECK();
      }
    }
    return CUBE_OK;
  }
  template <int DIR> static int c2c_y(cube_ctx *c, FFTPlan *p, float2 *a, int ncomp = 1) {
    constexpr int NH = N / 2 + 1;
    return strided<DIR>(c, p, a, (int64_t)cube_spec_pitch(N), (int64_t)cube_spec_pitch(N) * N, ncomp);
  }
  template <int DIR> static int c2c_z(cube_ctx *c, FFTPlan *p, float2 *a) {
    constexpr int NH = N / 2 + 1;
    return strided<DIR>(c, p, a, (int64_t)cube_spec_pitch(N) * N, (int64_t)cube_spec_pitch(N), 1);
  }
  /* components d0 <= d < d1, component d into work3 + (d-d0)*plane */
  static int green(cube_ctx *c, FFTPlan *p, const float2 *spec, const float *kern, int oct, float2 *work3,
                   int d0 = 0, int d1 = 3) {
    constexpr int NH = N / 2 + 1;
    if constexpr (RC<N>::ok) {
      /* pairs of column blocks x the line pairs (y, N-y), y = 0..N/2 */
      using G = RC<N>;
      dim3 g(((NH + G::CB - 1) / G::CB + 1) / 2, N / 2 + 1);
      const int64_t plane = (int64_t)cube_spec_pitch(N) * N * N;
      for (int e = d0; e < d1; e += G::ONE ? 1 : d1 - d0) {
        const int e1 = G::ONE ? e + 1 : d1;
        float2 *o = work3 + (e - d0) * plane;
        if (oct)
          hipLaunchKernelGGL((k_green_r<N, true>), g, dim3(G::NT), 0, c->stream, spec, o, kern, p->tw, e, e1);
        else
          hipLaunchKernelGGL((k_green_r<N, false>), g, dim3(G::NT), 0, c->stream, spec, o, kern, p->tw, e, e1);
        CUBE_LAUNCH_CHECK();
      }
    } else {
      dim3 g((NH + Cfg<N>::BS - 1) / Cfg<N>::BS, N);
      hipLaunchKernelGGL(k_green_z<N>, g, dim3(FFT_T), 0, c->stream, spec, work3, kern, oct, p->tw, d0, d1);
    }
    CUBE_LAUNCH_CHECK();
    return CUBE_OK;
  }
  static int c2r3(cube_ctx *c, FFTPlan *p, const float2 *work3, float *force, float div, unsigned long long *f2,
                  float pa, float pdt, int pre) {
    constexpr int BC = Cfg<N>::BC, B = BC >= 4 ? BC / 2 : BC;
    const int64_t rows = (int64_t)N * N;
    hipLaunchKernelGGL(k_c2r_x3<N>, dim3((unsigned)((rows + B - 1) / B)), dim3(FFT_T), 0, c->stream, work3,
                       force, p->tw, rows, div, f2, pa, pdt, pre);
    CUBE_LAUNCH_CHECK();
    return CUBE_OK;
  }
  /* one component's real rows at pitch opitch floats (out may be spec) */
  static int c2rp(cube_ctx *c, FFTPlan *p, const float2 *spec, float *out, int64_t opitch, float div) {
    constexpr int BC = Cfg<N>::BC, B = BC >= 4 ? BC / 2 : BC;
    const int64_t rows = (int64_t)N * N;
    hipLaunchKernelGGL(k_c2r_xp<N>, dim3((unsigned)((rows + B - 1) / B)), dim3(FFT_T), 0, c->stream, spec, out,
                       opitch, p->tw, rows, div);
    CUBE_LAUNCH_CHECK();
    return CUBE_OK;
  }
  /* the last pass, rows in descending chunks [a, b), a = ceil(b/3): a
     chunk's cells overwrite F0 rows >= 3a >= b only, whose cells are done */
  static int c2rf(cube_ctx *c, FFTPlan *p, const float2 *spec, float *force, const float *F1, int64_t p1,
                  float div, unsigned long long *f2, float pa, float pdt, int pre) {
    constexpr int BC = Cfg<N>::BC, B = BC >= 4 ? BC / 2 : BC;
    for (int64_t b = (int64_t)N * N; b > 0;) {
      int64_t a = (b + 2) / 3;
      if (a >= b) a = 0;
      hipLaunchKernelGGL(k_c2r_xf<N>, dim3((unsigned)((b - a + B - 1) / B)), dim3(FFT_T), 0, c->stream, spec,
                         force, force, F1, p1, p->tw, a, b, div, f2, pa, pdt, pre);
      CUBE_LAUNCH_CHECK();
      b = a;
    }
    return CUBE_OK;
  }
  static int c2r1(cube_ctx *c, FFTPlan *p, const float2 *spec, float *out) {
    const int64_t rows = (int64_t)N * N;
    hipLaunchKernelGGL(k_c2r_x1<N>, dim3((unsigned)((rows + Cfg<N>::BC - 1) / Cfg<N>::BC)), dim3(FFT_T), 0,
                       c->stream, spec, out, p->tw, rows);
    CUBE_LAUNCH_CHECK();
    return CUBE_OK;
  }
};

int cube_fft_r2c_3d(cube_ctx *c, FFTPlan *p, const float *in, float2 *out) {
  int st = CUBE_ERR_INVALID;
  switch (p->n) {
#define CASE(N)                                                           \
  case N:                                                                 \
    if ((st = Launch<N>::r2c(c, p, in, out))) return st;                  \
    return Launch<N>::c2c_y<-1>(c, p, out);
    FFT_SIZES(CASE)
#undef CASE
  }
  return st;
}

int cube_fft_forward_full(cube_ctx *c, FFTPlan *p, const float *in, float2 *out) {
  int st = cube_fft_r2c_3d(c, p, in, out);
  if (st) return st;
  switch (p->n) {
#define CASE(N) \
  case N:       \
    return Launch<N>::c2c_z<-1>(c, p, out);
    FFT_SIZES(CASE)
#undef CASE
  }
  return CUBE_ERR_INVALID;
}

float2 *cube_fft_plan_twiddles(FFTPlan *p) { return p->tw; }

int cube_fft_forward_full_tw(cube_ctx *c, int n, float2 *tw, const float *in, float2 *out) {
  if (!cube_fft_supported(n)) return CUBE_ERR_INVALID;
  FFTPlan p;
  p.n = n;
  p.tw = tw;
  return cube_fft_forward_full(c, &p, in, out);
}

int cube_fft_inverse_full(cube_ctx *c, FFTPlan *p, float2 *spec, float *out) {
  switch (p->n) {
#define CASE(N)                                                        \
  case N: {                                                            \
    int st;                                                            \
    if ((st = Launch<N>::c2c_z<+1>(c, p, spec))) return st;            \
    if ((st = Launch<N>::c2c_y<+1>(c, p, spec))) return st;            \
    return Launch<N>::c2r1(c, p, spec, out);                           \
  }
    FFT_SIZES(CASE)
#undef CASE
  }
  return CUBE_ERR_INVALID;
}

/* spec: output of cube_fft_r2c_3d (x and y transformed).  Produces the
   three force components (3 floats per cell) and max |F|^2; with prescale
   = {a_mid, dt} the cells hold the fine kick's (((F*a_mid)*dt)/6)/pi
   instead (max |F|^2 still of F).  work3 holds ncomp_spec spectra: 3 runs
   the three components together, 1 one after the other (the compact fine
   mesh), with bit-identical results: components 0 and 1 through work3 into
   planes (the first third of force3, then work3 itself), component 2
   through spec in place, so spec does not survive the call. */
int cube_fft_green_c2r_3d(cube_ctx *c, FFTPlan *p, float2 *spec, const float *kern, int kern_octant,
                          float2 *work3, float *force3, float div, float *f2max, int t0, unsigned *f2dev,
                          unsigned long long *f2slots, const float *prescale, int ncomp_spec) {
  const float pa = prescale ? prescale[0] : 0.f, pdt = prescale ? prescale[1] : 0.f;
  /* max |F|^2 as float bits: per workgroup into the slots, folded into the
     device word f2dev */
  int sst = cube_slots_zero(c, f2slots, 1);
  if (sst) return sst;
  const int64_t plane = (int64_t)cube_spec_pitch(p->n) * p->n * p->n;
  switch (p->n) {
#define CASE(N)                                                                   \
  case N: {                                                                       \
    int st;                                                                       \
    if (ncomp_spec >= 3) {                                                        \
      if (t0 >= 0) tbeg(c);                                                       \
      if ((st = Launch<N>::green(c, p, spec, kern, kern_octant, work3))) return st; \
      if (t0 >= 0) { tend(c, t0); tbeg(c); }                                      \
      if ((st = Launch<N>::c2c_y<+1>(c, p, work3, 3))) return st;                 \
      if (t0 >= 0) { tend(c, t0 + 1); tbeg(c); }                                  \
      if ((st = Launch<N>::c2r3(c, p, work3, force3, div, f2slots, pa, pdt, prescale != nullptr))) return st; \
      if (t0 >= 0) tend(c, t0 + 2);                                               \
    } else {                                                                      \
      /* component 2 goes through spec in place (its last reader) */              \
      for (int d = 0; d < 3; ++d) {                                               \
        float2 *sd = d < 2 ? work3 : spec;                                        \
        if (t0 >= 0) tbeg(c);                                                     \
        if ((st = Launch<N>::green(c, p, spec, kern, kern_octant, sd, d, d + 1))) return st; \
        if (t0 >= 0) { tend(c, t0); tbeg(c); }                                    \
        if ((st = Launch<N>::c2c_y<+1>(c, p, sd, 1))) return st;                  \
        if (t0 >= 0) { tend(c, t0 + 1); tbeg(c); }                                \
        if (d == 0) st = Launch<N>::c2rp(c, p, sd, force3, N, div);               \
        else if (d == 1)                                                          \
          st = Launch<N>::c2rp(c, p, sd, reinterpret_cast<float *>(work3), 2 * (int64_t)cube_spec_pitch(N), div); \
        else                                                                      \
          st = Launch<N>::c2rf(c, p, sd, force3, reinterpret_cast<const float *>(work3), 2 * (int64_t)cube_spec_pitch(N), \
                               div, f2slots, pa, pdt, prescale != nullptr);       \
        if (st) return st;                                                        \
        if (t0 >= 0) tend(c, t0 + 2);                                             \
      }                                                                           \
    }                                                                             \
    break;                                                                        \
  }
    FFT_SIZES(CASE)
#undef CASE
    default:
      return CUBE_ERR_INVALID;
  }
  if ((sst = cube_slots_fold(c, f2slots, 1, 0, nullptr, f2dev))) return sst;
  if (!f2max) return CUBE_OK; /* no readback: stays asynchronous (and capturable) */
  unsigned bits = 0;
  CUBE_CHECK(hipMemcpyAsync(&bits, f2dev, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  CUBE_CHECK(hipStreamSynchronize(c->stream));
  float f;
  memcpy(&f, &bits, sizeof(f));
  if (f2max) *f2max = f;
  return CUBE_OK;
}
