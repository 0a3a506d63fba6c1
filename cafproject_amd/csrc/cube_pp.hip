/*
 * cube_pp.hip -- ext_pp_force (W/ext_pp_force.f90:10-162) on the whole box.
 *
 * The reference builds, per tile, a head-of-chain linked list over fine
 * cells (hoc/ll, :53-75): particles of coarse cells 0..nt+1 are pushed in
 * (igz,igy,igx,lp) order, so a fine cell's chain lists its particles --
 * which all sit in one coarse cell -- in decreasing l.  Each physical
 * particle then sums the FP64 pair force over the (2*pp_range+1)^3
 * neighbouring fine cells in (kk,jj,ii) order, chain order inside a cell.
 *
 * Here the chain is implicit: the particles of a fine cell are those of its
 * coarse cell whose position code has the fine sub-cell index (top two bits
 * of the code for ncell=4), visited from the last particle of the coarse
 * cell backwards.  One thread per particle sums in exactly the reference
 * order, so the FP64 sums and the re-encoded velocities are bit-identical.
 */
#include "cube_internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>

#define PT 256

/* sqrt(x) for a normal x > 0 well inside the exponent range (here
   r^2 >= 2^-36): the reciprocal-square-root seed refined by Goldschmidt's
   coupled iteration (g -> sqrt(x), h -> 1/(2 sqrt(x))) and two Newton
   corrections of the residual x - g*g (Markstein), correctly rounded.  It
   is the sequence the compiler emits for sqrt() less the rescaling of
   extreme exponents and the special-case selects, which these operands
   never take, so the bits are sqrt()'s. */
__device__ __forceinline__ double pp_sqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = y * 0.5;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}
/* 1/d for a normal d well inside the exponent range (here 2^-45 < d <
   2^21): the reciprocal seed, two Newton steps e = 1 - d*y, y += y*e, and
   a final residual correction, correctly rounded; the compiler's division
   sequence for 1.0/d less its operand scaling and special-case fixup,
   which these operands never take, so the bits are 1.0/d's. */
__device__ __forceinline__ double pp_rcp(double d) {
  double y = __builtin_amdgcn_rcp(d);
  double e = fma(-d, y, 1.0);
  y = fma(y, e, y);
  e = fma(-d, y, 1.0);
  y = fma(y, e, y);
  const double r = fma(-d, y, 1.0);
  return fma(r, y, y);
}
/* x[a] / r3 for the three components, IEEE quotients either way.  With
   one_div a single division forms y = RN(1/r3) and each quotient is
   q0 = RN(x*y) corrected as RN(q0 + (x - r3*q0)*y) by two fused
   multiply-adds: for normal operands (here |x| < 64, r3 > 1e-3) that is
   the correctly rounded x/r3 (Markstein: y within half an ulp of 1/r3,
   q0 within one ulp of x/r3, the residual exact), bit for bit the
   division. */
__device__ static inline void quot3(const double x[3], double r3, double q[3], bool one_div) {
  if (one_div) {
    const double y = pp_rcp(r3);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double q0 = x[a] * y;
      q[a] = fma(fma(-r3, q0, x[a]), y, q0);
    }
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) q[a] = x[a] / r3;
  }
}

/* the pair force of a neighbour at x21 (W/ext_pp_force.f90:100-115), added
   to ft; every PP path forms it with this code */
__device__ __forceinline__ void pp_add(const double x21[3], double rsoft, double mass, int fdiv, double ft[3]) {
  const double r2 = (x21[0] * x21[0] + x21[1] * x21[1]) + x21[2] * x21[2];
  /* rmag = 0 (the particle itself) -> 1; pp_sqrt(0) is NaN and not taken */
  const double sq = pp_sqrt(r2);
  const double rmag = r2 == 0.0 ? 1.0 : sq;
  /* rcut = rmag/16 scales by a power of two, which commutes with rounding:
     rcut^3 = r3 * 2^-12 and rcut^5 = RN(rmag^2 * r3) * 2^-20 exactly (the
     products x*(x*x) and x^2*(x*x^2)), so 1.75*rcut^3 and 0.75*rcut^5 are single
     roundings of (1.75*2^-12)*r3 and (0.75*2^-20)*RN(rmag^2*r3) */
  const double rr = rmag * rmag, r3 = rmag * rr;
  const double pcut = (1.0 - (0x1.cp-12 * r3)) + (0x1.8p-21 * (rr * r3));
  double qt[3];
  quot3(x21, r3, qt, fdiv);
  /* at rmag <= rsoft the term is (...)*0.0 = +-0, which leaves the sum as
     it is (a sum that starts at +0 is never -0): a zero mass gives the same */
  const double m = rmag > rsoft ? mass : 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) ft[a] = ft[a] + (m * qt[a]) * pcut;
}

template <int IX, int IV>
__global__ __launch_bounds__(PT) void k_pp(const typename IntOf<IX>::T *__restrict__ xp,
                                           typename IntOf<IV>::T *__restrict__ vp,
                                           const int32_t *__restrict__ rhoc, const int64_t *__restrict__ cstart,
                                           const double *__restrict__ vdec, Geo g, int nvh, int pr, double rsoft,
                                           double mass, double a_mid, double dt, double venc,
                                           unsigned long long *__restrict__ out, int fdiv) {
  const int nt = g.nt;
  const int64_t row = blockIdx.x;
  const int j = (int)(row % nt), k = (int)((row / nt) % nt), t = (int)(row / ((int64_t)nt * nt));
  const int tz = t / (g.nnt * g.nnt), ty = (t / g.nnt) % g.nnt, tx = t % g.nnt;
  const int tt[3] = {tx, ty, tz};
  __shared__ int64_t cs[520];
  const int64_t c0 = row * nt;
  for (int i = threadIdx.x; i <= nt; i += PT) cs[i] = cstart[c0 + i];
  __syncthreads();
  const int sh = IX == 1 ? 6 : 14; /* fine sub-cell = code >> sh (ncell=4) */
  double f2m = 0;
  unsigned long long ntest = 0;
  for (int64_t p = cs[0] + threadIdx.x; p < cs[nt]; p += PT) {
    int lo = 0, hi = nt;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (cs[mid] <= p) lo = mid; else hi = mid;
    }
    const int loc[3] = {lo + 1, j + 1, k + 1};
    double x1[3], vreal[3], ft[3] = {0, 0, 0};
    int gf[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int32_t code = (int32_t)xp[3 * p + a];
      /* xvec1 = ncell*floor((ig-1)/4.) + ncell*(...)*x_resolution with ig the
         particle's fine cell; floor((ig-1)/4.) = coarse cell - 1 */
      x1[a] = (double)(4 * (loc[a] - 1)) + 4.0 * cube_xoff(IX, code);
      gf[a] = 4 * (loc[a] - 1) + (int)(cube_ucode(IX, code) >> sh) + 1; /* fine cell, 1-based, tile frame */
      vreal[a] = vdec[(int32_t)vp[3 * p + a] + nvh];
    }
    for (int kk = gf[2] - pr; kk <= gf[2] + pr; ++kk)
      for (int jj = gf[1] - pr; jj <= gf[1] + pr; ++jj)
        for (int ii = gf[0] - pr; ii <= gf[0] + pr; ++ii) {
          const int fc[3] = {ii, jj, kk};
          int cc[3], sub[3];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            cc[a] = (int)floorf((float)(fc[a] - 1) / 4.0f) + 1; /* coarse cell, tile frame */
            sub[a] = (fc[a] - 1) - 4 * (cc[a] - 1);
          }
          const int64_t cq = cid_global(g, tt[0] * nt + cc[0] - 1, tt[1] * nt + cc[1] - 1, tt[2] * nt + cc[2] - 1);
          const int np = rhoc[cq];
          const int64_t q0 = cstart[cq];
          for (int l = np - 1; l >= 0; --l) {
            const int64_t q = q0 + l;
            const int32_t c0x = (int32_t)xp[3 * q], c1x = (int32_t)xp[3 * q + 1], c2x = (int32_t)xp[3 * q + 2];
            if ((int)(cube_ucode(IX, c0x) >> sh) != sub[0] || (int)(cube_ucode(IX, c1x) >> sh) != sub[1] ||
                (int)(cube_ucode(IX, c2x) >> sh) != sub[2])
              continue;
            const int32_t cd[3] = {c0x, c1x, c2x};
            double x21[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) x21[a] = ((double)(4 * (cc[a] - 1)) + 4.0 * cube_xoff(IX, cd[a])) - x1[a];
            pp_add(x21, rsoft, mass, fdiv, ft);
          }
        }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      vreal[a] = vreal[a] + (((ft[a] * a_mid) * dt) / 6.0) / (double)CUBE_PI_F;
      vp[3 * p + a] = (typename IntOf<IV>::T)cube_encode_v(IV, venc, vreal[a]);
    }
    const double f2 = (ft[0] * ft[0] + ft[1] * ft[1]) + ft[2] * ft[2];
    f2m = f2 > f2m ? f2 : f2m;
    ++ntest;
  }
  __shared__ double red[PT];
  __shared__ unsigned long long cnt[PT];
  red[threadIdx.x] = f2m;
  cnt[threadIdx.x] = ntest;
  __syncthreads();
  for (int o = PT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
      cnt[threadIdx.x] += cnt[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicMax(cube_slot(out), (unsigned long long)__double_as_longlong(red[0]));
    atomicAdd(cube_slot(out + CUBE_SLOTS), cnt[0]);
  }
}

/*
 * Blocked PP: one workgroup per 4x4x4 coarse cells of a tile.  The
 * particles of the 6^3 coarse cells around it that fall in its fine-cell
 * neighbourhood (16+2*pr fine cells per axis) are staged in LDS, bucketed
 * per fine cell in chain order (decreasing l, the order of the reference's
 * hoc/ll lists), as exact fixed-point positions X = x * 2^15 in the tile
 * frame (x = ncell*(cell-1) + ncell*offset needs at most 10+15 bits), kept
 * relative to the region's corner (under 2^20 per axis) and packed 21 bits
 * per axis into one 64-bit word.  x2 - x1 = (X2 - X1) * 2^-15 is then the
 * same FP64 value the reference forms, and each particle sums its pairs in
 * the same order, so the result is bit-identical to the reference's walk.
 *
 * A fine cell lies in exactly one coarse cell, so one thread per staged
 * coarse cell counts its particles into the per-fine-cell table and, after
 * the scan, places them walking l downwards, with plain LDS updates.
 *
 * A block whose neighbourhood holds more than `maxs` particles -- a
 * clustered region -- is left to the cell path below: it goes to the list
 * `dlist` with its number of work items in `ditems`.
 */
#define PPB_T 256
#define PP_HEAVY 32 /* default heavy_at of the block kernel: staged cells above it are placed by their wave */
#define PPC_T 256
__device__ static inline int pp_fix(int IX, int cc, int32_t code) {
  return (cc - 1) * (4 * 32768) + (int)(2 * cube_ucode(IX, code) + 1) * (IX == 1 ? 256 : 1);
}
__device__ static inline uint64_t pp_pack(const int r[3]) {
  return (uint64_t)(uint32_t)r[0] | ((uint64_t)(uint32_t)r[1] << 21) | ((uint64_t)(uint32_t)r[2] << 42);
}
/* x21 = (X2 - X1) * 2^-15 from a staged word and the particle's own
   position x1 = X1 * 2^-15: X2 * 2^-15 - x1 is exact in FP64 (both are
   multiples of 2^-15 below 2^6), so one fused multiply-add per axis forms
   it */
__device__ __forceinline__ void pp_unpack_d(uint64_t w, const double x1[3], double x21[3]) {
  const int X2[3] = {(int)(w & 0x1FFFFF), (int)((w >> 21) & 0x1FFFFF), (int)(w >> 42)};
#pragma unroll
  for (int a = 0; a < 3; ++a) x21[a] = fma((double)X2[a], 0x1p-15, -x1[a]);
}
/* the kick by the summed force and the re-encoded velocity (:116-123) */
template <int IV>
__device__ __forceinline__ void pp_kick(typename IntOf<IV>::T *__restrict__ vp, int64_t p, double vreal[3],
                                        const double ft[3], double a_mid, double dt, double venc,
                                        const double *__restrict__ vtab, double &f2m) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    vreal[a] = vreal[a] + (((ft[a] * a_mid) * dt) / 6.0) / (double)CUBE_PI_F;
    vp[3 * p + a] = (typename IntOf<IV>::T)cube_encode_v_tab<IV>(venc, vreal[a], vtab);
  }
  const double f2 = (ft[0] * ft[0] + ft[1] * ft[1]) + ft[2] * ft[2];
  f2m = f2 > f2m ? f2 : f2m;
}
/* max |f|^2 and the particle count of a workgroup: wave shuffles, the
   waves through LDS, one global atomic pair (max and sum do not depend on
   the order) */
template <int NT>
__device__ __forceinline__ void pp_reduce(double f2m, unsigned long long ntest, unsigned long long *__restrict__ out) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    f2m = fmax(f2m, __shfl_xor(f2m, o, 64));
    ntest += __shfl_xor(ntest, o, 64);
  }
  __shared__ double wf2[NT / 64];
  __shared__ unsigned long long wn[NT / 64];
  if ((tid & 63) == 0) {
    wf2[tid >> 6] = f2m;
    wn[tid >> 6] = ntest;
  }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < NT / 64; ++w) {
      f2m = fmax(f2m, wf2[w]);
      ntest += wn[w];
    }
    atomicMax(cube_slot(out), (unsigned long long)__double_as_longlong(f2m));
    atomicAdd(cube_slot(out + CUBE_SLOTS), ntest);
  }
}

template <int IX, int IV>
__global__ __launch_bounds__(PPB_T) void k_pp_blk(const typename IntOf<IX>::T *__restrict__ xp,
                                                  typename IntOf<IV>::T *__restrict__ vp,
                                                  const int32_t *__restrict__ rhoc,
                                                  const int64_t *__restrict__ cstart,
                                                  const double *__restrict__ vdec, const double *__restrict__ vtab, Geo g, int nvh, int pr,
                                                  double rsoft, double mass, double a_mid, double dt, double venc,
                                                  int maxs, int fdiv, unsigned long long *__restrict__ out,
                                                  int32_t *__restrict__ dlist, int32_t *__restrict__ ditems,
                                                  unsigned long long *__restrict__ dcount, int heavy_at) {
  extern __shared__ uint64_t smem64[];
  const int RF = 16 + 2 * pr, nreg = RF * RF * RF;
  uint64_t *pos = smem64;                                            /* packed positions (maxs) */
  uint16_t *cur = reinterpret_cast<uint16_t *>(smem64 + maxs) + 1; /* per fine cell, cur[-1] = 0 */
  __shared__ int part[PPB_T / 64];
  __shared__ int64_t cst[16][5];
  __shared__ int rowoff[17];
  __shared__ int wide; /* a staged coarse cell too full for the 16-bit counts */
  const int tid = threadIdx.x, nt = g.nt, nb = nt / 4, nnt = g.nnt;
  const int sh = IX == 1 ? 6 : 14;
  double f2m = 0;
  unsigned long long ntest = 0;
  const int64_t b = blockIdx.x;
  const int bx = (int)(b % nb), by = (int)((b / nb) % nb), bz = (int)((b / ((int64_t)nb * nb)) % nb);
  const int t = (int)(b / ((int64_t)nb * nb * nb));
  const int tt[3] = {t % nnt, (t / nnt) % nnt, t / (nnt * nnt)};
  const int B0[3] = {4 * bx, 4 * by, 4 * bz};             /* cells B0+1..B0+4 (1-based, tile frame) */
  const int flo[3] = {4 * B0[0] + 1 - pr, 4 * B0[1] + 1 - pr, 4 * B0[2] + 1 - pr};
  /* tile-frame fixed-point coordinate of the region's low corner */
  const int X0[3] = {(flo[0] - 1) * 32768, (flo[1] - 1) * 32768, (flo[2] - 1) * 32768};

  for (int i = tid - 1; i < nreg; i += PPB_T) cur[i] = 0;
  if (tid == 0) wide = 0;
  if (tid < 16) {
    const int ry = tid % 4, rz = tid / 4;
    const int64_t c0 = g.cid(B0[0], B0[1] + ry, B0[2] + rz, t);
#pragma unroll
    for (int q = 0; q < 5; ++q) cst[tid][q] = cstart[c0 + q];
  }
  /* the staged coarse cells B0..B0+5 per axis (tile frame, 1-based) */
  const int q = tid;
  int sc[3] = {0, 0, 0}, np = 0;
  int64_t p0 = 0;
  if (q < 216) {
    sc[0] = B0[0] + q % 6;
    sc[1] = B0[1] + (q / 6) % 6;
    sc[2] = B0[2] + q / 36;
    const int64_t cq = cid_global(g, tt[0] * nt + sc[0] - 1, tt[1] * nt + sc[1] - 1, tt[2] * nt + sc[2] - 1);
    np = rhoc[cq];
    p0 = cstart[cq];
  }
  __syncthreads();
  if (np > 65535) wide = 1;
  /* the region cell of a particle of staged cell c, or -1 */
  auto region_in = [&](const int c[3], const int32_t code[3]) -> int {
    bool in = true;
    int r[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      r[a] = 4 * (c[a] - 1) + (int)(cube_ucode(IX, code[a]) >> sh) + 1 - flo[a];
      in = in && r[a] >= 0 && r[a] < RF;
    }
    return in ? (r[2] * RF + r[1]) * RF + r[0] : -1;
  };
  auto region = [&](const int32_t code[3]) -> int { return region_in(sc, code); };
  /* A staged cell of more than heavy_at particles would hold its wave for
     its whole length: its thread hands it to the wave, which counts and
     places 64 of its particles at a time after the light cells.  Lanes of
     one fine cell are matched by six ballots on the fine sub-cell index; a
     lane's rank among the lower lanes of its fine cell orders it as the
     serial walk does (the walk that places runs down l, lane 0 first). */
  const bool heavy = np > heavy_at;
  const int npl = heavy ? 0 : np;
  auto heavy_cells = [&](bool place) {
    const int lane = tid & 63;
    uint64_t hm = __ballot(heavy);
    while (hm) {
      const int src = __ffsll((unsigned long long)hm) - 1;
      hm &= hm - 1;
      const int hnp = __shfl(np, src, 64);
      const int64_t hp0 = (int64_t)(((uint64_t)(uint32_t)__shfl((int)(p0 >> 32), src, 64) << 32) |
                                    (uint64_t)(uint32_t)__shfl((int)p0, src, 64));
      const int hsc[3] = {__shfl(sc[0], src, 64), __shfl(sc[1], src, 64), __shfl(sc[2], src, 64)};
      for (int k = 0; 64 * k < hnp; ++k) {
        const int l = place ? hnp - 1 - 64 * k - lane : 64 * k + lane;
        int32_t code[3] = {0, 0, 0};
        int ri = -1;
        if (l >= 0 && l < hnp) {
#pragma unroll
          for (int a = 0; a < 3; ++a) code[a] = (int32_t)xp[3 * (hp0 + l) + a];
          ri = region_in(hsc, code);
        }
        const int sub = (int)((cube_ucode(IX, code[2]) >> sh) << 4 | (cube_ucode(IX, code[1]) >> sh) << 2 |
                              (cube_ucode(IX, code[0]) >> sh));
        uint64_t m = __ballot(ri >= 0);
#pragma unroll
        for (int b = 0; b < 6; ++b) {
          const uint64_t bb = __ballot((sub >> b) & 1);
          m &= ((sub >> b) & 1) ? bb : ~bb;
        }
        if (ri >= 0) {
          const int rank = __popcll(m & ((1ull << lane) - 1ull));
          if (!place) {
            if (rank == 0) cur[ri] = (uint16_t)(cur[ri] + __popcll(m));
          } else {
            const int slot = cur[ri] + rank;
            int r[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) r[a] = pp_fix(IX, hsc[a], code[a]) - X0[a];
            pos[slot] = pp_pack(r);
            if ((m >> lane) == 1ull) cur[ri] = (uint16_t)(slot + 1); /* the last lane of its fine cell */
          }
        }
      }
    }
  };
  /* the codes of particles l0, l0+d, l0+2d, l0+3d (clamped to [0, np)) in
     flight together */
  auto codes4 = [&](int l0, int d, int32_t c4[4][3]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int l = min(max(l0 + d * u, 0), np - 1);
#pragma unroll
      for (int a = 0; a < 3; ++a) c4[u][a] = (int32_t)xp[3 * (p0 + l) + a];
    }
  };
  /* count per fine cell: the cell's owner thread (or wave) only, no atomics */
  for (int l0 = 0; l0 < npl; l0 += 4) {
    int32_t c4[4][3];
    codes4(l0, 1, c4);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (l0 + u >= npl) break;
      const int ri = region(c4[u]);
      if (ri >= 0) cur[ri] = (uint16_t)(cur[ri] + 1);
    }
  }
  heavy_cells(false);
  __syncthreads();
  /* exclusive scan of the counts */
  const int chunk = (nreg + PPB_T - 1) / PPB_T;
  const int i0 = tid * chunk, i1 = min(i0 + chunk, nreg);
  int s = 0;
  for (int i = i0; i < i1; ++i) s += cur[i];
  int nstage;
  const int incl = block_incl_scan<PPB_T / 64>(s, part, nstage);
  s = incl - s;
  if (nstage > maxs || wide) { /* uniform: the block goes to the cell path */
    if (tid < 64) {
      /* work items: one per PPC_T particles of each of its 64 coarse cells */
      int items = (int)((cst[tid >> 2][(tid & 3) + 1] - cst[tid >> 2][tid & 3] + PPC_T - 1) / PPC_T);
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) items += __shfl_xor(items, o, 64);
      if (tid == 0) {
        const unsigned long long d = atomicAdd(dcount, 1ull);
        dlist[d] = (int32_t)b;
        ditems[d] = items;
      }
    }
    return;
  }
  for (int i = i0; i < i1; ++i) {
    const int cnt = cur[i];
    cur[i] = (uint16_t)s;
    s += cnt;
  }
  __syncthreads();
  /* chain order inside a fine cell is decreasing l: walk l downwards and
     append; afterwards cur[f] is the end of fine cell f's list */
  for (int l0 = npl - 1; l0 >= 0; l0 -= 4) {
    int32_t c4[4][3];
    codes4(l0, -1, c4);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (l0 - u < 0) break;
      const int ri = region(c4[u]);
      if (ri < 0) continue;
      const int slot = cur[ri];
      cur[ri] = (uint16_t)(slot + 1);
      int r[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) r[a] = pp_fix(IX, sc[a], c4[u][a]) - X0[a];
      pos[slot] = pp_pack(r);
    }
  }
  heavy_cells(true);
  if (tid == 0) {
    int acc = 0;
    for (int q2 = 0; q2 < 16; ++q2) {
      rowoff[q2] = acc;
      acc += (int)(cst[q2][4] - cst[q2][0]);
    }
    rowoff[16] = acc;
  }
  __syncthreads();
  /* fine cell f's list is pos[cur[f-1] .. cur[f]) */

  const int ntot = rowoff[16];
  for (int i = tid; i < ntot; i += PPB_T) {
    int row = 0; /* the last of the 16 rows with rowoff[row] <= i */
#pragma unroll
    for (int st = 8; st > 0; st >>= 1) row += rowoff[row + st] <= i ? st : 0;
    const int64_t p = cst[row][0] + (i - rowoff[row]);
    int cx = 0;
    while (cx < 3 && cst[row][cx + 1] <= p) ++cx;
    const int loc[3] = {B0[0] + 1 + cx, B0[1] + 1 + row % 4, B0[2] + 1 + row / 4};
    double vreal[3], ft[3] = {0, 0, 0};
    int gf[3];
    double x1[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int32_t code = (int32_t)xp[3 * p + a];
      x1[a] = (double)(pp_fix(IX, loc[a], code) - X0[a]) * 0x1p-15;
      gf[a] = 4 * (loc[a] - 1) + (int)(cube_ucode(IX, code) >> sh) + 1;
      vreal[a] = vdec[(int32_t)vp[3 * p + a] + nvh];
    }
    /* the 2pr+1 fine cells ii-pr..ii+pr of one (kk,jj) are adjacent in the
       region order, so their lists form one contiguous range of pos:
       (2pr+1)^2 ranges, walked kk-major then jj, hold every neighbour in
       the reference's (kk,jj,ii, chain) order.  One flat loop over the
       particle's pairs keeps a wave's lanes together: it runs the largest
       pair count of the wave, not 27 times the largest per-cell count. */
    const int RW = 2 * pr + 1;
    const int f00 = ((gf[2] - pr - flo[2]) * RF + (gf[1] - pr - flo[1])) * RF + (gf[0] - pr - flo[0]);
    /* row (kk, jj) starts at f00 + (kk*RF + jj)*RF: stepped by counters,
       not by dividing by the run-time RW */
    int n = 0;
    for (int kk = 0; kk < RW; ++kk)
      for (int jj = 0; jj < RW; ++jj) {
        const int f0 = f00 + (kk * RF + jj) * RF;
        n += (int)cur[f0 + RW - 1] - (int)cur[f0 - 1];
      }
    int f0 = f00 - RF, jr = -1, base = 0, cend = 0; /* pair t < cend lives at pos slot base + t */
    for (int t2 = 0; t2 < n; ++t2) {
      while (t2 >= cend) {
        if (++jr == RW) { /* next kk */
          jr = 0;
          f0 += (RF - RW + 1) * RF;
        } else {
          f0 += RF;
        }
        const int s0 = cur[f0 - 1], e = cur[f0 + RW - 1];
        base = s0 - cend;
        cend += e - s0;
      }
      double x21[3];
      pp_unpack_d(pos[base + t2], x1, x21);
      pp_add(x21, rsoft, mass, fdiv, ft);
    }
    pp_kick<IV>(vp, p, vreal, ft, a_mid, dt, venc, vtab, f2m);
    ++ntest;
  }
  pp_reduce<PPB_T>(f2m, ntest, out);
}

/* dlist's item counts to exclusive prefix sums in place, the total after
   the last (one workgroup) */
__global__ __launch_bounds__(1024) void k_pp_item_scan(int32_t *__restrict__ d,
                                                       const unsigned long long *__restrict__ dcount) {
  __shared__ int part[16];
  const int n = (int)*dcount, tid = threadIdx.x;
  const int chunk = (n + 1023) / 1024;
  const int i0 = min(tid * chunk, n), i1 = min(i0 + chunk, n);
  int s = 0;
  for (int i = i0; i < i1; ++i) s += d[i];
  int tot;
  const int incl = block_incl_scan<16>(s, part, tot);
  s = incl - s;
  for (int i = i0; i < i1; ++i) {
    const int v = d[i];
    d[i] = s;
    s += v;
  }
  if (tid == 0) d[n] = tot;
}

/*
 * The cell path, for the blocks of clustered regions: a work item is one
 * coarse cell C of a deferred block and up to PPC_T of its particles
 * (chunk c: particles c*PPC_T ...), so a dense block spreads over as many
 * workgroups as it has particles / PPC_T, and workgroups claim items from
 * a counter as they finish.  An item stages the (4+2*pr)^3 fine cells
 * around C, which lie in the 3^3 coarse cells C-1..C+1, in the staged
 * order above (region order of fine cells, chain order inside each).  A
 * particle's neighbours are a subsequence of that order in the
 * reference's (kk,jj,ii, chain) order, so when the region holds more than
 * `cap` particles it is staged `cap` slots at a time and each particle
 * carries its sums from one window to the next: the same pairs in the
 * same order, the same bits, in bounded LDS.
 *
 * Placement keeps chain order with wave-wide ranks: a wave takes 64
 * particles of a staged coarse cell in decreasing l, lanes with the same
 * fine cell find each other by ballots over the bits of its index, and a
 * lane's slot is the cell's running count plus the number of such lanes
 * before it.  Each staged coarse cell belongs to one wave, and each fine
 * cell to one coarse cell, so no two waves touch one count.
 */
template <int IX, int IV>
__global__ __launch_bounds__(PPC_T) void k_pp_cell(const typename IntOf<IX>::T *__restrict__ xp,
                                                   typename IntOf<IV>::T *__restrict__ vp,
                                                   const int32_t *__restrict__ rhoc,
                                                   const int64_t *__restrict__ cstart,
                                                   const double *__restrict__ vdec, const double *__restrict__ vtab, Geo g,
                                                   int nvh, int pr, double rsoft, double mass, double a_mid, double dt,
                                                   double venc, int cap, int fdiv, unsigned long long *__restrict__ out,
                                                   const int32_t *__restrict__ dlist, const int32_t *__restrict__ dpre,
                                                   const unsigned long long *__restrict__ dcount,
                                                   unsigned long long *__restrict__ claim) {
  extern __shared__ uint64_t smem64[];
  const int RF = 4 + 2 * pr, nreg = RF * RF * RF; /* nreg < 2048 for pr <= 4 */
  uint64_t *pos = smem64;                           /* one window of staged positions (cap) */
  int *off = reinterpret_cast<int *>(smem64 + cap); /* nreg + 1: first slot of each fine cell */
  int *cnt = off + nreg + 1;                        /* nreg: per fine cell, counted or placed so far */
  __shared__ int part[PPC_T / 64];
  __shared__ int s_item, s_blk, s_cell, s_chunk;
  __shared__ int own0[64];     /* first rank of each of C's 64 fine cells in the staged order */
  __shared__ int tgt[PPC_T];   /* this item's particles (l in C), in the staged order */
  __shared__ int snp[27];
  __shared__ int64_t sp0[27];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nt = g.nt, nb = nt / 4, nnt = g.nnt;
  const int sh = IX == 1 ? 6 : 14;
  const int ndef = (int)*dcount;
  const int total = dpre[ndef];
  double f2m = 0;
  unsigned long long ntest = 0;
  for (;;) {
    if (tid == 0) {
      s_item = (int)atomicAdd(claim, 1ull);
      s_cell = -1;
    }
    __syncthreads();
    const int item = s_item;
    if (item >= total) break; /* uniform */
    if (wv == 0) {
      /* the deferred block: dpre[w] <= item < dpre[w+1], 64 probes a round */
      int lo = 0, hi = ndef;
      while (hi - lo > 1) {
        const int qp = lo + 1 + (int)(((int64_t)(hi - lo - 1) * lane) / 64);
        const uint64_t m = __ballot(dpre[qp] <= item);
        const int k = __popcll(m);
        const int qlo = __shfl(qp, max(k - 1, 0), 64), qhi = __shfl(qp, min(k, 63), 64);
        if (k > 0) lo = qlo;
        hi = k < 64 ? qhi : hi;
      }
      /* the cell and chunk: scan of the 64 cells' chunk counts */
      const int b = dlist[lo];
      const int bx = b % nb, by = (b / nb) % nb, bz = (b / (nb * nb)) % nb, tb = b / (nb * nb * nb);
      const int64_t cl = g.cid(4 * bx + (lane & 3), 4 * by + ((lane >> 2) & 3), 4 * bz + (lane >> 4), tb);
      const int nch = (rhoc[cl] + PPC_T - 1) / PPC_T;
      const int incl = wave_incl_scan(nch), k = item - dpre[lo];
      if (k >= incl - nch && k < incl) {
        s_blk = b;
        s_cell = lane;
        s_chunk = k - (incl - nch);
      }
    }
    __syncthreads();
    const int b = s_blk, j = s_cell, c = s_chunk;
    if (j < 0) continue; /* uniform; cannot happen while cstart is the prefix sum of rhoc */
    const int bx = b % nb, by = (b / nb) % nb, bz = (b / (nb * nb)) % nb, tb = b / (nb * nb * nb);
    const int tt[3] = {tb % nnt, (tb / nnt) % nnt, tb / (nnt * nnt)};
    const int C[3] = {4 * bx + 1 + (j & 3), 4 * by + 1 + ((j >> 2) & 3), 4 * bz + 1 + (j >> 4)}; /* tile frame */
    const int flo[3] = {4 * (C[0] - 1) + 1 - pr, 4 * (C[1] - 1) + 1 - pr, 4 * (C[2] - 1) + 1 - pr};
    const int X0[3] = {(flo[0] - 1) * 32768, (flo[1] - 1) * 32768, (flo[2] - 1) * 32768};
    if (tid < 27) {
      const int64_t cq = cid_global(g, tt[0] * nt + C[0] - 2 + tid % 3, tt[1] * nt + C[1] - 2 + (tid / 3) % 3,
                                    tt[2] * nt + C[2] - 2 + tid / 9);
      snp[tid] = rhoc[cq];
      sp0[tid] = cstart[cq];
    }
    for (int i = tid; i < nreg; i += PPC_T) cnt[i] = 0;
    __syncthreads();
    /* staged coarse cell q = C - 1 + (q%3, q/3%3, q/9); its particle's
       region cell, or -1 */
    auto region = [&](int q, const int32_t code[3]) -> int {
      const int sq[3] = {C[0] - 1 + q % 3, C[1] - 1 + (q / 3) % 3, C[2] - 1 + q / 9};
      bool in = true;
      int r[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        r[a] = 4 * (sq[a] - 1) + (int)(cube_ucode(IX, code[a]) >> sh) + 1 - flo[a];
        in = in && r[a] >= 0 && r[a] < RF;
      }
      return in ? (r[2] * RF + r[1]) * RF + r[0] : -1;
    };
    /* count per fine cell (order-free: LDS atomics) */
    for (int q = 0; q < 27; ++q) {
      const int np = snp[q];
      const int64_t p0 = sp0[q];
      for (int l = tid; l < np; l += PPC_T) {
        const int32_t cd[3] = {(int32_t)xp[3 * (p0 + l)], (int32_t)xp[3 * (p0 + l) + 1], (int32_t)xp[3 * (p0 + l) + 2]};
        const int ri = region(q, cd);
        if (ri >= 0) atomicAdd(&cnt[ri], 1);
      }
    }
    __syncthreads();
    /* exclusive scan into off; off[nreg] = particles staged */
    const int chunk = (nreg + PPC_T - 1) / PPC_T;
    const int i0 = min(tid * chunk, nreg), i1 = min(i0 + chunk, nreg);
    int s = 0;
    for (int i = i0; i < i1; ++i) s += cnt[i];
    int nstage;
    const int incl = block_incl_scan<PPC_T / 64>(s, part, nstage);
    s = incl - s;
    for (int i = i0; i < i1; ++i) {
      off[i] = s;
      s += cnt[i];
    }
    if (tid == 0) off[nreg] = nstage;
    /* this thread's particle: number c*PPC_T + tid of C's particles taken
       in the staged order (fine cell by fine cell), so that a wave's lanes
       share fine cells and so their neighbour lists and pair counts */
    const bool act = c * PPC_T + tid < snp[13];
    int64_t p = 0;
    int f00 = 0;
    double x1[3] = {0, 0, 0};
    double vreal[3] = {0, 0, 0}, ft[3] = {0, 0, 0};
    const int RW = 2 * pr + 1;
    for (int w0 = 0; w0 < nstage; w0 += cap) {
      const int w1 = w0 + cap;
      __syncthreads(); /* off is complete; the last window's pairs are done */
      for (int i = tid; i < nreg; i += PPC_T) cnt[i] = 0;
      if (w0 == 0 && wv == 0) {
        /* C's fine cells are the region cells pr..pr+3 on each axis */
        const int ro = (((lane >> 4) + pr) * RF + ((lane >> 2) & 3) + pr) * RF + (lane & 3) + pr;
        const int no = off[ro + 1] - off[ro];
        own0[lane] = wave_incl_scan(no) - no;
      }
      __syncthreads();
      for (int q = wv; q < 27; q += PPC_T / 64) {
        const int np = snp[q];
        const int64_t p0 = sp0[q];
        const int sq[3] = {C[0] - 1 + q % 3, C[1] - 1 + (q / 3) % 3, C[2] - 1 + q / 9};
        for (int base = 0; base < np; base += 64) { /* uniform over the wave */
          const int l = np - 1 - (base + lane);
          int ri = -1;
          int32_t cd[3] = {0, 0, 0};
          if (l >= 0) {
#pragma unroll
            for (int a = 0; a < 3; ++a) cd[a] = (int32_t)xp[3 * (p0 + l) + a];
            ri = region(q, cd);
          }
          const unsigned key = (unsigned)(ri + 1);
          uint64_t m = ~0ull;
#pragma unroll
          for (int bit = 0; bit < 11; ++bit) {
            const uint64_t bb = __ballot((key >> bit) & 1u);
            m &= ((key >> bit) & 1u) ? bb : ~bb;
          }
          int done = 0;
          if (ri >= 0) {
            done = cnt[ri];
            const int slot = off[ri] + done + __popcll(m & ((1ull << lane) - 1ull));
            if (slot >= w0 && slot < w1) {
              int r[3];
#pragma unroll
              for (int a = 0; a < 3; ++a) r[a] = pp_fix(IX, sq[a], cd[a]) - X0[a];
              pos[slot - w0] = pp_pack(r);
            }
          }
          /* every lane has read its count: the group's last lane advances it */
          if (ri >= 0 && ((m >> lane) >> 1) == 0ull) cnt[ri] = done + __popcll(m);
          if (q == 13 && w0 == 0 && ri >= 0) {
            /* C's own particle: its rank among C's particles in staged order */
            const int rx = ri % RF - pr, ry = (ri / RF) % RF - pr, rz = ri / (RF * RF) - pr;
            const int k = own0[(rz * 4 + ry) * 4 + rx] + done + __popcll(m & ((1ull << lane) - 1ull)) - c * PPC_T;
            if (k >= 0 && k < PPC_T) tgt[k] = l;
          }
        }
      }
      __syncthreads();
      if (w0 == 0 && act) {
        p = sp0[13] + tgt[tid];
        int gf[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const int32_t code = (int32_t)xp[3 * p + a];
          x1[a] = (double)(pp_fix(IX, C[a], code) - X0[a]) * 0x1p-15;
          gf[a] = 4 * (C[a] - 1) + (int)(cube_ucode(IX, code) >> sh) + 1;
          vreal[a] = vdec[(int32_t)vp[3 * p + a] + nvh];
        }
        f00 = ((gf[2] - pr - flo[2]) * RF + (gf[1] - pr - flo[1])) * RF + (gf[0] - pr - flo[0]);
      }
      if (act) {
        /* the particle's (2pr+1)^2 row ranges clipped to [w0, w1), one flat
           loop as in the block kernel */
        int n = 0;
        for (int kk = 0; kk < RW; ++kk)
          for (int jj = 0; jj < RW; ++jj) {
            const int f0 = f00 + (kk * RF + jj) * RF;
            n += max(0, min(off[f0 + RW], w1) - max(off[f0], w0));
          }
        int f0 = f00 - RF, jr = -1, base = 0, cend = 0;
        for (int t2 = 0; t2 < n; ++t2) {
          while (t2 >= cend) {
            if (++jr == RW) {
              jr = 0;
              f0 += (RF - RW + 1) * RF;
            } else {
              f0 += RF;
            }
            const int s0 = max(off[f0], w0), e = min(off[f0 + RW], w1);
            base = s0 - w0 - cend;
            cend += max(e - s0, 0);
          }
          double x21[3];
          pp_unpack_d(pos[base + t2], x1, x21);
          pp_add(x21, rsoft, mass, fdiv, ft);
        }
      }
    }
    if (act) {
      pp_kick<IV>(vp, p, vreal, ft, a_mid, dt, venc, vtab, f2m);
      ++ntest;
    }
    __syncthreads(); /* the item's tables are read until here */
  }
  pp_reduce<PPC_T>(f2m, ntest, out);
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

template <int IX, int IV>
static int pp_impl(cube_ctx *c, float a_mid, float dt, float sigma_vi, float mass_p, unsigned long long *out) {
  const int pr = (int)c->p.pp_range;
  /* one division per pair (quot3); CUBE_PP_FASTDIV=0 takes the three
     divisions instead (the same bits) */
  const char *fd = getenv("CUBE_PP_FASTDIV");
  const int fdiv = !(fd && *fd && atoi(fd) == 0);
  if (c->nt % 4 == 0 && pr >= 0 && pr <= 4) {
    const int RF = 16 + 2 * pr, nreg = RF * RF * RF;
    const size_t table = (((size_t)nreg + 1) * 2 + 7) / 8 * 8; /* uint16 per fine cell, plus cur[-1] */
    /* block kernel: staged positions (8 bytes each) and the fine-cell table
       in about 28 KB, so five workgroups share a CU's LDS */
    int maxs = pr <= 1 ? 2048 : (pr == 2 ? 1024 : 512);
    /* cell path: a window of 4096 staged positions and two int tables
       (under 47 KB at pr = 4) */
    const int RC = 4 + 2 * pr, nregc = RC * RC * RC;
    int cap = 4096;
    /* other values move the split between the two paths and the cell
       path's window (small ones are used by the tests to cover every path);
       a window is bounded by the LDS one workgroup may take */
    int ldsmax = 65536;
    if (hipDeviceGetAttribute(&ldsmax, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess)
      ldsmax = 65536;
    if (const char *e = getenv("CUBE_PP_STAGE_MAX"); e && *e) maxs = std::max(0, std::min(maxs, atoi(e)));
    int heavy_at = PP_HEAVY;
    if (const char *e = getenv("CUBE_PP_HEAVY"); e && *e) heavy_at = std::max(0, atoi(e));
    if (const char *e = getenv("CUBE_PP_CELL_CAP"); e && *e)
      cap = std::max(1, std::min((ldsmax - 1024 - (2 * nregc + 1) * 4) / 8, atoi(e)));
    const int64_t nb = c->nt / 4, nblk = nb * nb * nb * c->ntiles;
    unsigned long long *dcount = (unsigned long long *)c->ured + 24, *claim = (unsigned long long *)c->ured + 25;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, dcount, 2);
    CUBE_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_pp_blk<IX, IV>), dim3((unsigned)nblk), dim3(PPB_T), (size_t)maxs * 8 + table, c->stream,
                       (const typename IntOf<IX>::T *)c->xp, (typename IntOf<IV>::T *)c->vp, c->rhoc, c->cstart,
                       c->vdec0, c->venc_tab, geo_of(c), c->nvbin / 2, pr, (double)c->p.rsoft, (double)mass_p, (double)a_mid,
                       (double)dt, cube_vconst(sigma_vi), maxs, fdiv, out, c->ppdefer, c->ppitems, dcount, heavy_at);
    CUBE_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pp_item_scan, dim3(1), dim3(1024), 0, c->stream, c->ppitems, dcount);
    CUBE_LAUNCH_CHECK();
    /* a fixed grid claims the items (the count stays on the device: no
       host round trip) */
    int ncu = 256;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess) ncu = 256;
    const size_t lds = (size_t)cap * 8 + (size_t)(2 * nregc + 1) * 4;
    hipLaunchKernelGGL((k_pp_cell<IX, IV>), dim3((unsigned)(4 * ncu)), dim3(PPC_T), lds, c->stream,
                       (const typename IntOf<IX>::T *)c->xp, (typename IntOf<IV>::T *)c->vp, c->rhoc, c->cstart,
                       c->vdec0, c->venc_tab, geo_of(c), c->nvbin / 2, pr, (double)c->p.rsoft, (double)mass_p,
                       (double)a_mid, (double)dt, cube_vconst(sigma_vi), cap, fdiv, out, c->ppdefer, c->ppitems, dcount,
                       claim);
    CUBE_LAUNCH_CHECK();
    return CUBE_OK;
  }
  const int64_t nrows = c->ncells / c->nt;
  hipLaunchKernelGGL((k_pp<IX, IV>), dim3((unsigned)nrows), dim3(PT), 0, c->stream,
                     (const typename IntOf<IX>::T *)c->xp, (typename IntOf<IV>::T *)c->vp, c->rhoc, c->cstart,
                     c->vdec0, geo_of(c), c->nvbin / 2, (int)c->p.pp_range, (double)c->p.rsoft, (double)mass_p,
                     (double)a_mid, (double)dt, cube_vconst(sigma_vi), out, fdiv);
  CUBE_LAUNCH_CHECK();
  return CUBE_OK;
}

int cube_pp(cube_ctx *c, float a_mid, float dt, float sigma_vi, float mass_p, float *f2_max_pp, int64_t *itest1) {
  int st = cube_vdec_table(c, c->vdec0, sigma_vi);
  if (st) return st;
  unsigned long long *out = (unsigned long long *)c->ured + 16, *osl = c->slots + CUBE_SLOT_PP * CUBE_SLOTS;
  if ((st = cube_slots_zero(c, osl, 2))) return st;
  const int x = c->p.izipx, v = c->p.izipv;
  if (x == 1 && v == 1) st = pp_impl<1, 1>(c, a_mid, dt, sigma_vi, mass_p, osl);
  else if (x == 1) st = pp_impl<1, 2>(c, a_mid, dt, sigma_vi, mass_p, osl);
  else if (v == 1) st = pp_impl<2, 1>(c, a_mid, dt, sigma_vi, mass_p, osl);
  else st = pp_impl<2, 2>(c, a_mid, dt, sigma_vi, mass_p, osl);
  if (st) return st;
  if ((st = cube_slots_fold(c, osl, 1, 0, out, nullptr)) || (st = cube_slots_fold(c, osl + CUBE_SLOTS, 1, 1, out + 1, nullptr)))
    return st;
  unsigned long long h[2];
  CUBE_CHECK(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  CUBE_CHECK(hipStreamSynchronize(c->stream));
  double f2;
  memcpy(&f2, &h[0], sizeof(f2));
  /* ftemp=max(ftemp,sum(f_tot**2)) with a `save`d real(4) ftemp that is
     never reset (W/ext_pp_force.f90:20,123): a running max over calls */
  const double prev = (double)c->pp_ftemp;
  c->pp_ftemp = (float)(f2 > prev ? f2 : prev);
  if (f2_max_pp) *f2_max_pp = c->pp_ftemp;
  if (itest1) *itest1 = (int64_t)h[1];
  return CUBE_OK;
}
