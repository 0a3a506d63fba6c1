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

__device__ static inline double powi3(double x) { return x * (x * x); }
__device__ static inline double powi5(double x) {
  const double x2 = x * x;
  return x2 * (x * x2);
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
    const double y = 1.0 / r3;
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
            double rmag = sqrt((x21[0] * x21[0] + x21[1] * x21[1]) + x21[2] * x21[2]);
            if (rmag == 0.0) rmag = 1.0;
            const double rcut = rmag / 16.0;
            const double pcut = (1.0 - (1.75 * powi3(rcut))) + (0.75 * powi5(rcut));
            const double r3 = powi3(rmag);
            double qt[3];
            quot3(x21, r3, qt, fdiv);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              double fp = (mass * qt[a]) * pcut;
              if (!(rmag > rsoft)) fp = fp * 0.0;
              ft[a] = ft[a] + fp;
            }
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
 * the same order, so the result is bit-identical to the unstaged walk.
 *
 * A fine cell lies in exactly one coarse cell, so one thread per staged
 * coarse cell counts its particles into the per-fine-cell table and, after
 * the scan, places them walking l downwards, with plain LDS updates.
 *
 * Two launches share the kernel.  The first, one workgroup per block with
 * a small staging area (several workgroups per CU), defers a block whose
 * neighbourhood holds more than `maxs` particles -- a clustered region --
 * to the list `dlist`.  The second walks that list with a staging area of
 * most of a CU's LDS (one workgroup per CU, `list`/`nlist`); a block that
 * overflows even that walks the global lists instead.  Every block's
 * particles are kicked exactly once, by the same arithmetic either way.
 */
#define PPB_T 256
__device__ static inline int pp_fix(int IX, int cc, int32_t code) {
  return (cc - 1) * (4 * 32768) + (int)(2 * cube_ucode(IX, code) + 1) * (IX == 1 ? 256 : 1);
}
__device__ static inline uint64_t pp_pack(const int r[3]) {
  return (uint64_t)(uint32_t)r[0] | ((uint64_t)(uint32_t)r[1] << 21) | ((uint64_t)(uint32_t)r[2] << 42);
}

template <int IX, int IV>
__global__ __launch_bounds__(PPB_T) void k_pp_blk(const typename IntOf<IX>::T *__restrict__ xp,
                                                  typename IntOf<IV>::T *__restrict__ vp,
                                                  const int32_t *__restrict__ rhoc,
                                                  const int64_t *__restrict__ cstart,
                                                  const double *__restrict__ vdec, const double *__restrict__ vtab, Geo g, int nvh, int pr,
                                                  double rsoft, double mass, double a_mid, double dt, double venc,
                                                  int maxs, int fdiv, unsigned long long *__restrict__ out,
                                                  int32_t *__restrict__ dlist, unsigned long long *__restrict__ dcount,
                                                  const int32_t *__restrict__ list,
                                                  const unsigned long long *__restrict__ nlist) {
  extern __shared__ uint64_t smem64[];
  const int RF = 16 + 2 * pr, nreg = RF * RF * RF;
  uint64_t *pos = smem64;                                            /* packed positions (maxs) */
  uint16_t *cur = reinterpret_cast<uint16_t *>(smem64 + maxs) + 1; /* per fine cell, cur[-1] = 0 */
  __shared__ int part[PPB_T / 64];
  __shared__ int64_t cst[16][5];
  __shared__ int rowoff[17];
  const int tid = threadIdx.x, nt = g.nt, nb = nt / 4, nnt = g.nnt;
  const int sh = IX == 1 ? 6 : 14;
  double f2m = 0;
  unsigned long long ntest = 0;
  /* first launch: block blockIdx.x; second launch: the deferred blocks */
  const int64_t nw = list ? (int64_t)*nlist : (int64_t)gridDim.x;
  for (int64_t w = blockIdx.x; w < nw; w += gridDim.x) {
  const int64_t b = list ? (int64_t)list[w] : w;
  const int bx = (int)(b % nb), by = (int)((b / nb) % nb), bz = (int)((b / ((int64_t)nb * nb)) % nb);
  const int t = (int)(b / ((int64_t)nb * nb * nb));
  const int tt[3] = {t % nnt, (t / nnt) % nnt, t / (nnt * nnt)};
  const int B0[3] = {4 * bx, 4 * by, 4 * bz};             /* cells B0+1..B0+4 (1-based, tile frame) */
  const int flo[3] = {4 * B0[0] + 1 - pr, 4 * B0[1] + 1 - pr, 4 * B0[2] + 1 - pr};
  /* tile-frame fixed-point coordinate of the region's low corner */
  const int X0[3] = {(flo[0] - 1) * 32768, (flo[1] - 1) * 32768, (flo[2] - 1) * 32768};

  for (int i = tid - 1; i < nreg; i += PPB_T) cur[i] = 0;
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
  /* the region cell of a particle of this thread's staged cell, or -1 */
  auto region = [&](const int32_t code[3]) -> int {
    bool in = true;
    int r[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      r[a] = 4 * (sc[a] - 1) + (int)(cube_ucode(IX, code[a]) >> sh) + 1 - flo[a];
      in = in && r[a] >= 0 && r[a] < RF;
    }
    return in ? (r[2] * RF + r[1]) * RF + r[0] : -1;
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
  /* count per fine cell: the cell's owner thread only, no atomics */
  for (int l0 = 0; l0 < np; l0 += 4) {
    int32_t c4[4][3];
    codes4(l0, 1, c4);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (l0 + u >= np) break;
      const int ri = region(c4[u]);
      if (ri >= 0) cur[ri] = (uint16_t)(cur[ri] + 1);
    }
  }
  __syncthreads();
  /* exclusive scan of the counts */
  const int chunk = (nreg + PPB_T - 1) / PPB_T;
  const int i0 = tid * chunk, i1 = min(i0 + chunk, nreg);
  int s = 0;
  for (int i = i0; i < i1; ++i) s += cur[i];
  int nstage;
  const int incl = block_incl_scan<PPB_T / 64>(s, part, nstage);
  const bool staged = nstage <= maxs; /* uniform over the block */
  s = incl - s;
  if (!staged && dlist) { /* uniform: the block goes to the second launch */
    if (tid == 0) dlist[atomicAdd(dcount, 1ull)] = (int32_t)b;
    __syncthreads();
    continue;
  }
  for (int i = i0; i < i1; ++i) {
    const int cnt = cur[i];
    cur[i] = (uint16_t)(staged ? s : 0);
    s += cnt;
  }
  __syncthreads();
  if (staged) {
    /* chain order inside a fine cell is decreasing l: walk l downwards and
       append; afterwards cur[f] is the end of fine cell f's list */
    for (int l0 = np - 1; l0 >= 0; l0 -= 4) {
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
  }
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
    int X1[3], gf[3];
    int32_t pcode[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int32_t code = (int32_t)xp[3 * p + a];
      pcode[a] = code;
      X1[a] = pp_fix(IX, loc[a], code) - X0[a];
      gf[a] = 4 * (loc[a] - 1) + (int)(cube_ucode(IX, code) >> sh) + 1;
      vreal[a] = vdec[(int32_t)vp[3 * p + a] + nvh];
    }
    if (staged) {
      /* the 2pr+1 fine cells ii-pr..ii+pr of one (kk,jj) are adjacent in
         the region order, so their lists form one contiguous range of pos:
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
      for (int t = 0; t < n; ++t) {
        while (t >= cend) {
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
        const uint64_t w = pos[base + t];
        const int X2[3] = {(int)(w & 0x1FFFFF), (int)((w >> 21) & 0x1FFFFF), (int)(w >> 42)};
        double x21[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) x21[a] = (double)(X2[a] - X1[a]) * (1.0 / 32768.0);
        double rmag = sqrt((x21[0] * x21[0] + x21[1] * x21[1]) + x21[2] * x21[2]);
        if (rmag == 0.0) rmag = 1.0;
        const double rcut = rmag / 16.0;
        const double pcut = (1.0 - (1.75 * powi3(rcut))) + (0.75 * powi5(rcut));
        const double r3 = powi3(rmag);
        double qt[3];
        quot3(x21, r3, qt, fdiv);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          double fp = (mass * qt[a]) * pcut;
          if (!(rmag > rsoft)) fp = fp * 0.0;
          ft[a] = ft[a] + fp;
        }
      }
    } else {
    double x1[3]; /* FP64 tile-frame position for the walk over the global lists */
#pragma unroll
    for (int a = 0; a < 3; ++a) x1[a] = (double)(4 * (loc[a] - 1)) + 4.0 * cube_xoff(IX, pcode[a]);
    for (int kk = gf[2] - pr; kk <= gf[2] + pr; ++kk)
      for (int jj = gf[1] - pr; jj <= gf[1] + pr; ++jj)
        for (int ii = gf[0] - pr; ii <= gf[0] + pr; ++ii) {
          {
            const int fc[3] = {ii, jj, kk};
            int cc[3], sub[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              cc[a] = (int)floorf((float)(fc[a] - 1) / 4.0f) + 1;
              sub[a] = (fc[a] - 1) - 4 * (cc[a] - 1);
            }
            const int64_t cq =
                cid_global(g, tt[0] * nt + cc[0] - 1, tt[1] * nt + cc[1] - 1, tt[2] * nt + cc[2] - 1);
            const int np = rhoc[cq];
            const int64_t q0 = cstart[cq];
            for (int l = np - 1; l >= 0; --l) {
              const int64_t q = q0 + l;
              const int32_t cd[3] = {(int32_t)xp[3 * q], (int32_t)xp[3 * q + 1], (int32_t)xp[3 * q + 2]};
              if ((int)(cube_ucode(IX, cd[0]) >> sh) != sub[0] || (int)(cube_ucode(IX, cd[1]) >> sh) != sub[1] ||
                  (int)(cube_ucode(IX, cd[2]) >> sh) != sub[2])
                continue;
              double x21[3];
#pragma unroll
              for (int a = 0; a < 3; ++a)
                x21[a] = ((double)(4 * (cc[a] - 1)) + 4.0 * cube_xoff(IX, cd[a])) - x1[a];
              double rmag = sqrt((x21[0] * x21[0] + x21[1] * x21[1]) + x21[2] * x21[2]);
              if (rmag == 0.0) rmag = 1.0;
              const double rcut = rmag / 16.0;
              const double pcut = (1.0 - (1.75 * powi3(rcut))) + (0.75 * powi5(rcut));
              const double r3 = powi3(rmag);
              double qt[3];
              quot3(x21, r3, qt, fdiv);
#pragma unroll
              for (int a = 0; a < 3; ++a) {
                double fp = (mass * qt[a]) * pcut;
                if (!(rmag > rsoft)) fp = fp * 0.0;
                ft[a] = ft[a] + fp;
              }
            }
          }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      vreal[a] = vreal[a] + (((ft[a] * a_mid) * dt) / 6.0) / (double)CUBE_PI_F;
      vp[3 * p + a] = (typename IntOf<IV>::T)cube_encode_v_tab<IV>(venc, vreal[a], vtab);
    }
    const double f2 = (ft[0] * ft[0] + ft[1] * ft[1]) + ft[2] * ft[2];
    f2m = f2 > f2m ? f2 : f2m;
    ++ntest;
  }
  __syncthreads(); /* the block's tables are read until here */
  }
  /* max |f|^2 and the particle count: wave shuffles, the waves through
     LDS, one global atomic pair per workgroup (max and sum do not depend
     on the order) */
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    f2m = fmax(f2m, __shfl_xor(f2m, o, 64));
    ntest += __shfl_xor(ntest, o, 64);
  }
  __shared__ double wf2[PPB_T / 64];
  __shared__ unsigned long long wn[PPB_T / 64];
  if ((tid & 63) == 0) {
    wf2[tid >> 6] = f2m;
    wn[tid >> 6] = ntest;
  }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < PPB_T / 64; ++w) {
      f2m = fmax(f2m, wf2[w]);
      ntest += wn[w];
    }
    atomicMax(cube_slot(out), (unsigned long long)__double_as_longlong(f2m));
    atomicAdd(cube_slot(out + CUBE_SLOTS), ntest);
  }
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
  const int fdiv = !(fd && atoi(fd) == 0);
  if (c->nt % 4 == 0 && pr >= 0 && pr <= 4) {
    const int RF = 16 + 2 * pr, nreg = RF * RF * RF;
    const size_t table = (((size_t)nreg + 1) * 2 + 7) / 8 * 8; /* uint16 per fine cell, plus cur[-1] */
    /* first launch: staged positions (8 bytes each) and the fine-cell
       table in about 28 KB, so five workgroups share a CU's LDS */
    int maxs = pr <= 1 ? 2048 : (pr == 2 ? 1024 : 512);
    /* second launch: most of the LDS a workgroup may take, for the blocks
       the first one deferred (uint16 offsets cap it at 65535) */
    int lds = 65536;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess) lds = 65536;
    int maxb = (int)std::min<int64_t>(16384, ((int64_t)lds - 4096 - (int64_t)table) / 8);
    /* lower caps send more workgroups down the second launch or the
       unstaged walk (used by the tests to cover every path in one run) */
    if (const char *e = getenv("CUBE_PP_STAGE_MAX")) maxs = std::min(maxs, atoi(e));
    if (const char *e = getenv("CUBE_PP_BIG_MAX")) maxb = std::min(maxb, atoi(e));
    const bool two = maxb > maxs;
    const int64_t nb = c->nt / 4, nblk = nb * nb * nb * c->ntiles;
    unsigned long long *dcount = (unsigned long long *)c->ured + 24;
    if (two) {
      hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, dcount, 1);
      CUBE_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL((k_pp_blk<IX, IV>), dim3((unsigned)nblk), dim3(PPB_T), (size_t)maxs * 8 + table, c->stream,
                       (const typename IntOf<IX>::T *)c->xp, (typename IntOf<IV>::T *)c->vp, c->rhoc, c->cstart,
                       c->vdec0, c->venc_tab, geo_of(c), c->nvbin / 2, pr, (double)c->p.rsoft, (double)mass_p, (double)a_mid,
                       (double)dt, cube_vconst(sigma_vi), maxs, fdiv, out, two ? c->ppdefer : nullptr, dcount, nullptr,
                       nullptr);
    CUBE_LAUNCH_CHECK();
    if (two) {
      /* a fixed grid of one workgroup per CU-slot walks however many blocks
         were deferred (read on the device: no host round trip) */
      int ncu = 256;
      if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess) ncu = 256;
      const unsigned grid = (unsigned)std::min<int64_t>(nblk, 2 * (int64_t)ncu);
      hipLaunchKernelGGL((k_pp_blk<IX, IV>), dim3(grid), dim3(PPB_T), (size_t)maxb * 8 + table, c->stream,
                         (const typename IntOf<IX>::T *)c->xp, (typename IntOf<IV>::T *)c->vp, c->rhoc, c->cstart,
                         c->vdec0, c->venc_tab, geo_of(c), c->nvbin / 2, pr, (double)c->p.rsoft, (double)mass_p,
                         (double)a_mid, (double)dt, cube_vconst(sigma_vi), maxb, fdiv, out, nullptr, nullptr,
                         c->ppdefer, dcount);
      CUBE_LAUNCH_CHECK();
    }
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
