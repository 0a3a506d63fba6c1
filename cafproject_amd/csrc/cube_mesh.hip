/*
 * cube_mesh.hip -- CIC deposits, force gathers (kicks) and the Green's
 * functions of particle_mesh (W/pm.f90) and kernel_f/kernel_c
 * (W/kernel_f.f90, W/kernel_c.f90) for a single image.
 *
 * Deposits are deterministic gathers: one thread per mesh cell walks the
 * few source coarse cells that can reach it in the reference's own visit
 * order (ilayer, k, j, i, l) with nlayer=3, re-evaluates each particle's
 * real(4) CIC weights with the reference's expression in the owning tile's
 * frame, and adds them in that order -- so every physical cell holds the
 * bit pattern the reference's layered OpenMP loop produces for it.
 */
#include "cube_internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MT 256

/* per-axis candidate source coarse cells of a mesh cell, in visit order
   restricted to one axis: cells with different k are ordered by
   (ilayer(k), k) with ilayer = (k - k0) mod 3 */
struct Src {
  int n;
  int c[3];
};
__device__ static inline void order_z(Src &s, int k0) {
  /* sort s.c[0..n) by ((c-k0) mod 3, c), n <= 3 */
  for (int a = 1; a < s.n; ++a) {
    const int x = s.c[a];
    const int kx = ((x - k0) % 3 + 3) % 3;
    int b = a - 1;
    while (b >= 0) {
      const int y = s.c[b];
      const int ky = ((y - k0) % 3 + 3) % 3;
      if (ky > kx || (ky == kx && y > x)) {
        s.c[b + 1] = y;
        --b;
      } else {
        break;
      }
    }
    s.c[b + 1] = x;
  }
}

/* the coarse CIC coordinate ((cell-1)+(x offset))-0.5 of W/pm.f90:213,335
   rounded to real(4); c = cell-1 (tile frame or global).  With 1-byte codes
   c + (u+0.5)/256 has at most 24 significant bits for c < 2^15, so the
   real(4) sums are exact and round as the real(8) form does; 2-byte codes
   keep the real(8) form. */
template <int IX> __device__ __forceinline__ float coarse_tempx(int c, int32_t code) {
  if constexpr (IX == 1)
    return ((float)c + ((float)cube_ucode(IX, code) + 0.5f) * 0x1p-8f) - 0.5f;
  else
    return (float)(((double)c + cube_xoff(IX, code)) - 0.5);
}
/* ---------------- coarse CIC deposit, W/pm.f90:205-267 ---------------- */
template <int IX>
__global__ __launch_bounds__(MT) void k_coarse_deposit(const typename IntOf<IX>::T *__restrict__ xp,
                                                       const int32_t *__restrict__ rhoc,
                                                       const int64_t *__restrict__ cstart, Geo g, float mass,
                                                       float *__restrict__ r3, double *__restrict__ part) {
  const int nc = g.nc;
  const int64_t nc3 = (int64_t)nc * nc * nc;
  double tsum = 0;
  for (int64_t m = blockIdx.x * (int64_t)MT + threadIdx.x; m < nc3; m += (int64_t)gridDim.x * MT) {
    const int gg[3] = {(int)(m % nc), (int)((m / nc) % nc), (int)(m / ((int64_t)nc * nc))};
    int T[3], lm[3];
    Src s[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      T[a] = gg[a] / g.nt;
      lm[a] = gg[a] - T[a] * g.nt + 1;
      s[a].n = 3; s[a].c[0] = lm[a] - 1; s[a].c[1] = lm[a]; s[a].c[2] = lm[a] + 1;
    }
    order_z(s[2], 0);
    float acc = 0.f;
    for (int qz = 0; qz < 3; ++qz)
      for (int qy = 0; qy < 3; ++qy)
        for (int qx = 0; qx < 3; ++qx) {
          const int ci[3] = {s[0].c[qx], s[1].c[qy], s[2].c[qz]};
          const int64_t cc = cid_global(g, T[0] * g.nt + ci[0] - 1, T[1] * g.nt + ci[1] - 1, T[2] * g.nt + ci[2] - 1);
          const int np = rhoc[cc];
          const int64_t p0 = cstart[cc];
          for (int l = 0; l < np; ++l) {
            float w[3];
            bool hit = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              const int32_t code = (int32_t)xp[3 * (p0 + l) + a];
              const float tempx = coarse_tempx<IX>(ci[a] - 1, code);
              const int i1 = (int)floorf(tempx) + 1;
              const float d1 = (float)i1 - tempx;
              if (i1 == lm[a]) w[a] = d1;
              else if (i1 + 1 == lm[a]) w[a] = 1.0f - d1;
              else hit = false;
            }
            if (hit) acc = acc + w[0] * w[1] * w[2] * mass;
          }
        }
    r3[m] = acc;
    tsum += (double)acc;
  }
  __shared__ double red[MT];
  red[threadIdx.x] = tsum;
  __syncthreads();
  for (int o = MT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

/* sum of nb partials in a fixed order (thread t takes b = t mod 64, then a
   tree over the threads), so the sum repeats bit for bit */
__global__ __launch_bounds__(64) void k_sum_parts1(const double *part, int nb, double *out) {
  __shared__ double r[64];
  const int t = threadIdx.x;
  double s = 0;
  for (int b = t; b < nb; b += 64) s += part[b];
  r[t] = s;
  __syncthreads();
  for (int o = 32; o > 0; o >>= 1) {
    if (t < o) r[t] += r[t + o];
    __syncthreads();
  }
  if (t == 0) *out = r[0];
}

/*
 * Fine deposit, one thread per coarse cell C owning its 4x4x4 fine cells.
 * Every particle that can reach them lives in the 2x2x2 source cells
 * {C-1,C} per axis.  The thread walks those cells in the reference's visit
 * order -- z by (ilayer,k), then y, then x, then l -- and adds each hit into
 * a per-thread LDS accumulator of its 64 fine cells.  Restricted to one fine
 * cell this is exactly that cell's reference order, so the sums are
 * bit-identical; each particle is read by the 8 threads whose blocks its
 * cloud can touch instead of by all 64+ fine cells around it.
 */
/* global coordinates of cell cg of the nc^3 x-fastest mesh (32-bit
   divisions while the index fits, which is every mesh up to 1290^3) */
__device__ __forceinline__ void gcoords(int64_t cg, int nc, int gg[3]) {
  if (cg <= 0x7fffffff) {
    const uint32_t c = (uint32_t)cg, n = (uint32_t)nc, q = c / n, z = q / n;
    gg[0] = (int)(c - q * n);
    gg[1] = (int)(q - z * n);
    gg[2] = (int)z;
  } else {
    gg[0] = (int)(cg % nc);
    gg[1] = (int)((cg / nc) % nc);
    gg[2] = (int)(cg / ((int64_t)nc * nc));
  }
}
#define FD_T 128
__device__ __forceinline__ int rl_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float rl_f(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
template <int IX>
__global__ __launch_bounds__(FD_T) void k_fine_deposit_cc(const typename IntOf<IX>::T *__restrict__ xp,
                                                          const int32_t *__restrict__ rhoc,
                                                          const int64_t *__restrict__ cstart, Geo g, float mass,
                                                          float *__restrict__ rho, double *__restrict__ part,
                                                          int heavy_at) {
  __shared__ float acc[64 * FD_T]; /* element q of thread t at q*FD_T + t: conflict free */
  const int nc = g.nc, nf = 4 * nc;
  const int64_t nc3 = (int64_t)nc * nc * nc;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  double tsum = 0;
  for (int64_t cg = blockIdx.x * (int64_t)FD_T + t; cg - t < nc3; cg += (int64_t)gridDim.x * FD_T) {
    const bool live = cg < nc3;
    for (int q = 0; q < 64; ++q) acc[q * FD_T + t] = 0.f;
    int gg[3] = {0, 0, 0}, C[3] = {1, 1, 1}, za = 0, zb = 0;
    int npc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t p0c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool heavy = false;
    if (live) {
      gcoords(cg, nc, gg);
      int T[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        T[a] = gg[a] / g.nt;
        C[a] = gg[a] - T[a] * g.nt + 1; /* 1-based local coarse cell */
      }
      /* cell ids are linear in the local coordinates and the tile index, so
         a source below C on an axis is C's id plus that axis's step down:
         -stride inside the tile, the wrapped neighbour tile's last cell at
         the tile's low edge */
      const int64_t nt = g.nt, nt3 = nt * nt * nt;
      const int64_t cidC = g.cid(C[0] - 1, C[1] - 1, C[2] - 1, T[0] + g.nnt * (T[1] + g.nnt * T[2]));
      int64_t dm[3];
      {
        const int64_t stride[3] = {1, nt, nt * nt}, tstride[3] = {nt3, nt3 * g.nnt, nt3 * g.nnt * g.nnt};
#pragma unroll
        for (int a = 0; a < 3; ++a)
          dm[a] = C[a] > 1 ? -stride[a] : stride[a] * (nt - 1) + tstride[a] * (T[a] > 0 ? -1 : g.nnt - 1);
      }
      /* the two source z cells in (ilayer, k) order, ilayer = (k-k0) mod 3 */
      za = C[2] - 1;
      zb = C[2];
      {
        const int k0 = 2 - g.ncb;
        const int ka = ((za - k0) % 3 + 3) % 3, kb = ((zb - k0) % 3 + 3) % 3;
        if (ka > kb) {
          za = C[2];
          zb = C[2] - 1;
        }
      }
      /* counts and starts of the 8 source cells, all loads issued up front */
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int64_t cc = cidC + ((q & 1) ? 0 : dm[0]) + (((q >> 1) & 1) ? 0 : dm[1]) +
                           ((((q >> 2) ? zb : za) == C[2]) ? 0 : dm[2]);
        npc[q] = rhoc[cc];
        p0c[q] = cstart[cc];
      }
      heavy = ((npc[0] + npc[1]) + (npc[2] + npc[3])) + ((npc[4] + npc[5]) + (npc[6] + npc[7])) > heavy_at;
      if (!heavy)
#pragma unroll 1
      for (int qz = 0; qz < 2; ++qz)
#pragma unroll 1
        for (int qy = 0; qy < 2; ++qy)
#pragma unroll 1
          for (int qx = 0; qx < 2; ++qx) {
            const int ci[3] = {C[0] - 1 + qx, C[1] - 1 + qy, qz ? zb : za};
            const int qi = (qz * 2 + qy) * 2 + qx;
            int np = npc[0];
            int64_t p0 = p0c[0];
#pragma unroll
            for (int q = 1; q < 8; ++q) { /* register select, no dynamic indexing */
              np = qi == q ? npc[q] : np;
              p0 = qi == q ? p0c[q] : p0;
            }
            /* a source cell below C on an axis reaches C's 4^3 fine cells
               only through particles in its last fine sub-cell on that axis
               (code >> sh == 3), so those visits first mark the candidates
               of each 64-particle chunk in a mask and then visit only them,
               in increasing l.  Where the real(4) tempx is inexact (2-byte
               codes, nt > 128) the top code of a lower sub-cell can round up
               onto the boundary itself, and its weight past the boundary is
               then exactly 0: skipping it drops only +0 terms. */
            const int sh = IX == 1 ? 6 : 14;
            const bool ex = qx == 0, ey = qy == 0, ez = ci[2] == C[2] - 1;
            for (int lb = 0; lb < np; lb += 64) {
              const int nb = min(64, np - lb);
              uint64_t mask = nb == 64 ? ~0ull : ((1ull << nb) - 1ull);
              if (ex || ey || ez) {
                uint64_t m = 0;
                for (int l0 = 0; l0 < nb; l0 += 4) {
                  int32_t c4[4][3];
#pragma unroll
                  for (int u = 0; u < 4; ++u) {
                    const int64_t q = p0 + lb + min(l0 + u, nb - 1);
#pragma unroll
                    for (int a = 0; a < 3; ++a) c4[u][a] = (int32_t)xp[3 * q + a];
                  }
#pragma unroll
                  for (int u = 0; u < 4; ++u) {
                    const bool pass = (!ex || (cube_ucode(IX, c4[u][0]) >> sh) == 3) &&
                                      (!ey || (cube_ucode(IX, c4[u][1]) >> sh) == 3) &&
                                      (!ez || (cube_ucode(IX, c4[u][2]) >> sh) == 3);
                    m |= (uint64_t)(pass ? 1 : 0) << (l0 + u);
                  }
                }
                mask &= m;
              }
              /* two candidates per round so that both code loads are in flight */
              while (mask) {
                int lq[2];
                lq[0] = lb + __builtin_ctzll(mask);
                mask &= mask - 1ull;
                const bool two = mask != 0ull;
                lq[1] = two ? lb + __builtin_ctzll(mask) : lq[0];
                if (two) mask &= mask - 1ull;
                int32_t cq[2][3];
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                  for (int a = 0; a < 3; ++a) cq[u][a] = (int32_t)xp[3 * (p0 + lq[u]) + a];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                  if (u == 1 && !two) break;
                  int r[3];
                  float d1[3];
#pragma unroll
                  for (int a = 0; a < 3; ++a) {
                    /* ncell*(i-1)+ncell*(x offset) rounded to real(4) (W/pm.f90:63):
                       both terms are exact floats, so one float add rounds it */
                    /* (u+0.5)*4*2^-8izipx: exact in real(4) as in real(8) */
                    const float off4 = ((float)cube_ucode(IX, cq[u][a]) + 0.5f) * (IX == 1 ? 0x1p-6f : 0x1p-14f);
                    const float tempx = (float)(4 * (ci[a] - 1)) + off4;
                    const int i1 = (int)floorf(tempx) + 1;
                    d1[a] = (float)i1 - tempx;
                    r[a] = i1 - 4 * (C[a] - 1) - 1; /* 0-based in C's block */
                  }
#pragma unroll
                  for (int cz = 0; cz < 2; ++cz) {
                    const int z = r[2] + cz;
                    if (z < 0 || z > 3) continue;
                    const float wz = cz ? 1.0f - d1[2] : d1[2];
#pragma unroll
                    for (int cy = 0; cy < 2; ++cy) {
                      const int y = r[1] + cy;
                      if (y < 0 || y > 3) continue;
                      const float wy = cy ? 1.0f - d1[1] : d1[1];
#pragma unroll
                      for (int cx = 0; cx < 2; ++cx) {
                        const int x = r[0] + cx;
                        if (x < 0 || x > 3) continue;
                        const float wx = cx ? 1.0f - d1[0] : d1[0];
                        float &sacc = acc[((z * 4 + y) * 4 + x) * FD_T + t];
                        sacc = sacc + wx * wy * wz * mass;
                      }
                    }
                  }
                }
              }
            }
          }
    }
    /* a heavy lane's cell is formed by its whole wave: lane f owns fine
       cell f = (z*4+y)*4+x of it and adds, in the same order and with the
       same expression as the lane would, the terms of the candidates that
       64 lanes decode at a time */
    uint64_t hm = __ballot(live && heavy);
    while (hm) {
      const int L = __builtin_ctzll(hm);
      hm &= hm - 1;
      const int LC[3] = {rl_i(C[0], L), rl_i(C[1], L), rl_i(C[2], L)};
      const int Lza = rl_i(za, L), Lzb = rl_i(zb, L);
      const int fx = lane & 3, fy = (lane >> 2) & 3, fz = lane >> 4;
      float sacc = 0.f;
      for (int qi = 0; qi < 8; ++qi) {
        const int qx = qi & 1, qy = (qi >> 1) & 1, qz = qi >> 2;
        const int ci[3] = {LC[0] - 1 + qx, LC[1] - 1 + qy, qz ? Lzb : Lza};
        int np = 0;
        int64_t p0 = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) { /* the owner's counts and starts */
          const int nq = rl_i(npc[q], L);
          const int64_t pq = (int64_t)(((uint64_t)(uint32_t)rl_i((int)(p0c[q] >> 32), L) << 32) |
                                       (uint64_t)(uint32_t)rl_i((int)(p0c[q] & 0xffffffff), L));
          np = qi == q ? nq : np;
          p0 = qi == q ? pq : p0;
        }
        const int sh = IX == 1 ? 6 : 14;
        const bool ex = qx == 0, ey = qy == 0, ez = ci[2] == LC[2] - 1;
        for (int lb = 0; lb < np; lb += 64) {
          const int l = lb + lane;
          bool pass = false;
          float d1v[3] = {0.f, 0.f, 0.f}, d2v[3] = {0.f, 0.f, 0.f};
          int rp = 0;
          if (l < np) {
            int32_t cq[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) cq[a] = (int32_t)xp[3 * (p0 + l) + a];
            pass = (!ex || (cube_ucode(IX, cq[0]) >> sh) == 3) && (!ey || (cube_ucode(IX, cq[1]) >> sh) == 3) &&
                   (!ez || (cube_ucode(IX, cq[2]) >> sh) == 3);
            if (pass) {
              int r[3];
#pragma unroll
              for (int a = 0; a < 3; ++a) {
                const float off4 = ((float)cube_ucode(IX, cq[a]) + 0.5f) * (IX == 1 ? 0x1p-6f : 0x1p-14f);
                const float tempx = (float)(4 * (ci[a] - 1)) + off4;
                const int i1 = (int)floorf(tempx) + 1;
                d1v[a] = (float)i1 - tempx;
                d2v[a] = 1.0f - d1v[a];
                r[a] = i1 - 4 * (LC[a] - 1) - 1;
              }
              rp = ((r[0] + 1) & 7) | (((r[1] + 1) & 7) << 3) | (((r[2] + 1) & 7) << 6);
            }
          }
          uint64_t m = __ballot(pass);
          while (m) { /* candidates in increasing l, read from their lanes */
            const int i = __builtin_ctzll(m);
            m &= m - 1;
            const int ri = rl_i(rp, i);
            const int cx = fx - ((ri & 7) - 1), cy = fy - (((ri >> 3) & 7) - 1), cz = fz - (((ri >> 6) & 7) - 1);
            const float ax0 = rl_f(d1v[0], i), ax1 = rl_f(d2v[0], i), ay0 = rl_f(d1v[1], i), ay1 = rl_f(d2v[1], i);
            const float az0 = rl_f(d1v[2], i), az1 = rl_f(d2v[2], i);
            if ((unsigned)cx <= 1u && (unsigned)cy <= 1u && (unsigned)cz <= 1u) {
              const float wx = cx ? ax1 : ax0, wy = cy ? ay1 : ay0, wz = cz ? az1 : az0;
              sacc = sacc + wx * wy * wz * mass;
            }
          }
        }
      }
      acc[lane * FD_T + wv * 64 + L] = sacc;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (live) {
      const int64_t f0x = 4 * (int64_t)gg[0], f0y = 4 * (int64_t)gg[1], f0z = 4 * (int64_t)gg[2];
      for (int z = 0; z < 4; ++z)
        for (int y = 0; y < 4; ++y) {
          float4 v;
          v.x = acc[((z * 4 + y) * 4 + 0) * FD_T + t];
          v.y = acc[((z * 4 + y) * 4 + 1) * FD_T + t];
          v.z = acc[((z * 4 + y) * 4 + 2) * FD_T + t];
          v.w = acc[((z * 4 + y) * 4 + 3) * FD_T + t];
          *reinterpret_cast<float4 *>(rho + ((f0z + z) * nf + (f0y + y)) * nf + f0x) = v;
          tsum += ((double)v.x + (double)v.y) + ((double)v.z + (double)v.w);
        }
    }
  }
  __syncthreads();
  double *red = reinterpret_cast<double *>(acc);
  red[t] = tsum;
  __syncthreads();
  for (int o = FD_T / 2; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = red[0];
}

/*
 * Coarse deposit, one thread per 2x2x2 block of mesh cells (m0..m0+1 per
 * axis, within one tile since nt is even): the sources are the 4 cells
 * m0-1..m0+2 per axis, walked in the reference order (ilayer=k mod 3, k,
 * j, i, l), with the 8 sums in registers.  A particle that misses a cell
 * adds +0.0, which leaves a sum of non-negative terms unchanged.
 *
 * The 16 source rows (k, j) are gathered in that order.  A lane whose
 * block's sources pass `heavy_at` particles (a clustered region) stops at
 * a row boundary, and its whole wave finishes the block: 64 lanes form 64
 * particles' 8 terms at once (the same expression), and 8 lanes add them
 * to the partial sums in particle order -- the same additions in the same
 * order, so the same bits, without one lane walking thousands of particles
 * while the other 63 wait.  The owner adds the 8 sums to its total in the
 * same place of its sequence as a light block's.
 */
template <int IX>
__device__ __forceinline__ void cd_weights(const int ci[3], const int m0[3], const int32_t code[3], float w0[3],
                                           float w1[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float tempx = coarse_tempx<IX>(ci[a] - 1, code[a]);
    const int i1 = (int)floorf(tempx) + 1;
    const float d1 = (float)i1 - tempx, d2 = 1.0f - d1;
    const int r = i1 - m0[a]; /* corner 1 relative to the block */
    w0[a] = r == 0 ? d1 : (r == -1 ? d2 : 0.f);
    w1[a] = r == 1 ? d1 : (r == 0 ? d2 : 0.f);
  }
}

template <int IX>
__global__ __launch_bounds__(MT) void k_coarse_deposit_b2(const typename IntOf<IX>::T *__restrict__ xp,
                                                          const int32_t *__restrict__ rhoc,
                                                          const int64_t *__restrict__ cstart, Geo g, float mass,
                                                          float *__restrict__ r3, double *__restrict__ part,
                                                          int heavy_at) {
  const int nc = g.nc, nb = nc / 2;
  const int64_t nb3 = (int64_t)nb * nb * nb;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __shared__ float cw[MT / 64][64 * 9]; /* a wave's 64 particles x 8 terms, rows of 9: no bank conflicts */
  __shared__ float sfin[MT / 64][8];
  double tsum = 0;
  /* the loop is uniform over a wave: lane l takes block b0 + l */
  for (int64_t b0 = blockIdx.x * (int64_t)MT + wv * 64; b0 < nb3; b0 += (int64_t)gridDim.x * MT) {
    const int64_t b = b0 + lane;
    const bool live = b < nb3;
    int gg[3] = {0, 0, 0}, T[3] = {0, 0, 0}, m0[3] = {1, 1, 1}, zs[4] = {0, 0, 0, 0};
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int rrow = 16; /* first source row the lane has not gathered */
    if (live) {
      gg[0] = 2 * (int)(b % nb);
      gg[1] = 2 * (int)((b / nb) % nb);
      gg[2] = 2 * (int)(b / ((int64_t)nb * nb));
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        T[a] = gg[a] / g.nt;
        m0[a] = gg[a] - T[a] * g.nt + 1;
      }
      zs[0] = m0[2] - 1;
      zs[1] = m0[2];
      zs[2] = m0[2] + 1;
      zs[3] = m0[2] + 2;
      /* order by (k mod 3, k): insertion sort of 4 */
      for (int a = 1; a < 4; ++a) {
        const int x = zs[a];
        int bb = a - 1;
        while (bb >= 0 && ((zs[bb] % 3) > (x % 3) || ((zs[bb] % 3) == (x % 3) && zs[bb] > x))) {
          zs[bb + 1] = zs[bb];
          --bb;
        }
        zs[bb + 1] = x;
      }
      int cum = 0;
      for (int row = 0; row < 16; ++row) {
        const int qz = row >> 2, sy = m0[1] - 1 + (row & 3);
        /* the row's 4 source cells: counts and starts loaded together */
        int npr[4];
        int64_t p0r[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int64_t cc = cid_global(g, T[0] * g.nt + m0[0] - 2 + u, T[1] * g.nt + sy - 1, T[2] * g.nt + zs[qz] - 1);
          npr[u] = rhoc[cc];
          p0r[u] = cstart[cc];
        }
        cum += (npr[0] + npr[1]) + (npr[2] + npr[3]);
        if (cum > heavy_at) {
          rrow = row;
          break;
        }
#pragma unroll 1
        for (int sx = m0[0] - 1; sx <= m0[0] + 2; ++sx) {
          const int ci[3] = {sx, sy, zs[qz]};
          const int ui = sx - (m0[0] - 1);
          int np = npr[0];
          int64_t p0 = p0r[0];
#pragma unroll
          for (int u = 1; u < 4; ++u) { /* register select, no dynamic indexing */
            np = ui == u ? npr[u] : np;
            p0 = ui == u ? p0r[u] : p0;
          }
          /* four particles' codes in flight at once (the last one repeated) */
          for (int l0 = 0; l0 < np; l0 += 4) {
            int32_t code4[4][3];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const int64_t q = p0 + min(l0 + u, np - 1);
#pragma unroll
              for (int a = 0; a < 3; ++a) code4[u][a] = (int32_t)xp[3 * q + a];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              if (l0 + u >= np) break;
              float w0[3], w1[3];
              cd_weights<IX>(ci, m0, code4[u], w0, w1);
#pragma unroll
              for (int q = 0; q < 8; ++q) {
                const float wx = (q & 1) ? w1[0] : w0[0], wy = (q & 2) ? w1[1] : w0[1], wz = (q & 4) ? w1[2] : w0[2];
                s[q] = s[q] + wx * wy * wz * mass;
              }
            }
          }
        }
      }
    }
    /* the wave finishes each heavy lane's block from its row rrow on */
    uint64_t hm = __ballot(live && rrow < 16);
    while (hm) {
      const int L = __builtin_ctzll(hm);
      hm &= hm - 1;
      const int Lgg[3] = {rl_i(gg[0], L), rl_i(gg[1], L), rl_i(gg[2], L)};
      const int LT[3] = {rl_i(T[0], L), rl_i(T[1], L), rl_i(T[2], L)};
      const int Lm0[3] = {rl_i(m0[0], L), rl_i(m0[1], L), rl_i(m0[2], L)};
      const int Lzs[4] = {rl_i(zs[0], L), rl_i(zs[1], L), rl_i(zs[2], L), rl_i(zs[3], L)};
      const int Lrow = rl_i(rrow, L);
      float acc = 0.f; /* lane q < 8: sum q, from the owner's partial sum */
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float v = rl_f(s[q], L);
        acc = lane == q ? v : acc;
      }
      for (int row = Lrow; row < 16; ++row) {
        const int qz = row >> 2, sy = Lm0[1] - 1 + (row & 3);
        for (int u = 0; u < 4; ++u) {
          const int sx = Lm0[0] - 1 + u;
          const int64_t cc = cid_global(g, LT[0] * g.nt + sx - 1, LT[1] * g.nt + sy - 1, LT[2] * g.nt + Lzs[qz] - 1);
          const int np = rhoc[cc];
          const int64_t p0 = cstart[cc];
          const int ci[3] = {sx, sy, Lzs[qz]};
          for (int base = 0; base < np; base += 64) {
            const int l = base + lane;
            if (l < np) {
              int32_t code[3];
#pragma unroll
              for (int a = 0; a < 3; ++a) code[a] = (int32_t)xp[3 * (p0 + l) + a];
              float w0[3], w1[3];
              cd_weights<IX>(ci, Lm0, code, w0, w1);
#pragma unroll
              for (int q = 0; q < 8; ++q) {
                const float wx = (q & 1) ? w1[0] : w0[0], wy = (q & 2) ? w1[1] : w0[1], wz = (q & 4) ? w1[2] : w0[2];
                cw[wv][lane * 9 + q] = wx * wy * wz * mass;
              }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane < 8) {
              const int m = min(64, np - base);
              for (int i = 0; i < m; ++i) acc = acc + cw[wv][i * 9 + lane];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          }
        }
      }
      if (lane < 8) {
        const int x = Lgg[0] + (lane & 1), y = Lgg[1] + ((lane >> 1) & 1), z = Lgg[2] + ((lane >> 2) & 1);
        r3[((int64_t)z * nc + y) * nc + x] = acc;
        sfin[wv][lane] = acc;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (lane == L)
        for (int q = 0; q < 8; ++q) tsum += (double)sfin[wv][q];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (live && rrow == 16) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int x = gg[0] + (q & 1), y = gg[1] + ((q >> 1) & 1), z = gg[2] + ((q >> 2) & 1);
        r3[((int64_t)z * nc + y) * nc + x] = s[q];
        tsum += (double)s[q];
      }
    }
  }
  __shared__ double red[MT];
  red[threadIdx.x] = tsum;
  __syncthreads();
  for (int o = MT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

#define DEP_BLOCKS 2048
template <int IX> static int fine_dep(cube_ctx *c, float mass, double *t) {
  /* a coarse cell whose 8 source cells hold more particles than this is
     formed by its lane's whole wave (CUBE_FD_HEAVY) */
  int heavy_at = 512;
  if (const char *e = getenv("CUBE_FD_HEAVY"); e && *e) heavy_at = std::max(0, atoi(e));
  hipLaunchKernelGGL(k_fine_deposit_cc<IX>, dim3(DEP_BLOCKS), dim3(FD_T), 0, c->stream,
                     (const typename IntOf<IX>::T *)c->xp, c->rhoc, c->cstart, geo_of(c), mass, c->rho_f, c->dred,
                     heavy_at);
  CUBE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_parts1, dim3(1), dim3(64), 0, c->stream, c->dred, DEP_BLOCKS,
                     c->dred + CUBE_DRED_FINE_SUM);
  CUBE_LAUNCH_CHECK();
  if (t) {
    CUBE_CHECK(hipMemcpyAsync(t, c->dred + CUBE_DRED_FINE_SUM, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CUBE_CHECK(hipStreamSynchronize(c->stream));
  }
  return CUBE_OK;
}
int cube_fine_deposit(cube_ctx *c, float mass_p, double *testrhof) {
  return c->p.izipx == 1 ? fine_dep<1>(c, mass_p, testrhof) : fine_dep<2>(c, mass_p, testrhof);
}
template <int IX> static int coarse_dep(cube_ctx *c, float mass, double *t) {
  /* partial sums in the upper half of dred, apart from the fine deposit's,
     so that the two deposits may run at the same time */
  double *part = c->dred + 4096;
  /* source particles a lane gathers alone before its wave takes the rest
     of its block (CUBE_CD_HEAVY) */
  int heavy_at = 1024;
  if (const char *e = getenv("CUBE_CD_HEAVY"); e && *e) heavy_at = std::max(0, atoi(e));
  if (c->nt % 2 == 0)
    hipLaunchKernelGGL(k_coarse_deposit_b2<IX>, dim3(DEP_BLOCKS), dim3(MT), 0, c->stream,
                       (const typename IntOf<IX>::T *)c->xp, c->rhoc, c->cstart, geo_of(c), mass, c->rho_c, part,
                       heavy_at);
  else
    hipLaunchKernelGGL(k_coarse_deposit<IX>, dim3(DEP_BLOCKS), dim3(MT), 0, c->stream,
                       (const typename IntOf<IX>::T *)c->xp, c->rhoc, c->cstart, geo_of(c), mass, c->rho_c, part);
  CUBE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_parts1, dim3(1), dim3(64), 0, c->stream, part, DEP_BLOCKS,
                     c->dred + CUBE_DRED_COARSE_SUM);
  CUBE_LAUNCH_CHECK();
  if (t) {
    CUBE_CHECK(hipMemcpyAsync(t, c->dred + CUBE_DRED_COARSE_SUM, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CUBE_CHECK(hipStreamSynchronize(c->stream));
  }
  return CUBE_OK;
}
int cube_coarse_deposit(cube_ctx *c, float mass_p, double *testrhoc) {
  return c->p.izipx == 1 ? coarse_dep<1>(c, mass_p, testrhoc) : coarse_dep<2>(c, mass_p, testrhoc);
}

/* ---------------- kicks ---------------- */
#define KT 256
struct F3 {
  float x, y, z;
};

/* G = (((F*a_mid)*dt)/6)/pi per mesh value, the factor of every kick term
   (W/pm.f90:341-348); the coarse kick reads G so each coarse cell's value is
   formed once instead of by every particle whose cloud touches it */
__global__ void k_kick_prescale(const float *__restrict__ F, int64_t n, float a_mid, float dt, float *__restrict__ G) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const float a = (F[q] * a_mid) * dt;
    bool ok = true;
    const float v = div_6pi(a, ok);
    G[q] = ok ? v : (a / 6.0f) / CUBE_PI_F;
  }
}

/* One kick of one particle: vreal[a] += the CIC gather of the mesh over
   the 8 corners in the order of W/pm.f90:146-153 (fine: force_f with
   ((((F*a_mid)*dt)/6)/pi)*wx*wy*wz per term; coarse: the prescaled G with
   G*wx*wy*wz, W/pm.f90:341-348).  loc: 1-based cell in the tile, tt: tile. */
template <int IX, bool FINE, bool PRE = false>
__device__ __forceinline__ void kick_particle(const int32_t code[3], const int loc[3], const int tt[3],
                                              const Geo &g, const float *__restrict__ force, float a_mid,
                                              float dt, double vreal[3]) {
  /* PRE: the fine mesh already holds (((F*a_mid)*dt)/6)/pi per value */
  constexpr bool DIV = FINE && !PRE;
  const int nt = g.nt, nfine = 4 * g.nc, nft = 4 * nt;
  int i1[3];
  float d1[3], d2[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float tempx;
    if (FINE) /* ncell*((/i,j,k/)-1)+ncell*(...)*x_resolution, tile frame (W/pm.f90:138):
                 both terms are exact floats, so one float add rounds it */
      tempx = (float)(4 * (loc[a] - 1)) + ((float)cube_ucode(IX, code[a]) + 0.5f) * (IX == 1 ? 0x1p-6f : 0x1p-14f);
    else /* ((/itx,../)-1)*nt+((/i,../)-1)+(...)*x_resolution-0.5 (W/pm.f90:335) */
      tempx = coarse_tempx<IX>(tt[a] * nt + (loc[a] - 1), code[a]);
    i1[a] = (int)floorf(tempx) + 1;
    d1[a] = (float)i1[a] - tempx;
    d2[a] = 1.0f - d1[a];
  }
  /* 0-based periodic mesh coordinates of the two CIC planes per axis.
     tempx lies in (4(loc-1), 4 loc] (fine) or (g-1/2, g+1/2] around the
     global cell g (coarse), so the first plane is in [0, nfine] or
     [-1, nc-1] and one compare wraps each plane (no integer modulo) */
  const int n = FINE ? nfine : g.nc;
  int m1[3], m2[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int v = FINE ? tt[a] * nft + i1[a] - 1 : i1[a] - 1;
    m1[a] = v < 0 ? v + n : (v >= n ? v - n : v);
    m2[a] = v + 1 >= n ? v + 1 - n : v + 1;
  }
  /* corner order of W/pm.f90:146-153 */
  const int cx[8] = {0, 1, 0, 0, 0, 1, 1, 1}, cy[8] = {0, 0, 1, 0, 1, 0, 1, 1}, cz[8] = {0, 0, 0, 1, 1, 1, 0, 1};
  const double v0[3] = {vreal[0], vreal[1], vreal[2]};
  bool ok = true;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int X = cx[q] ? m2[0] : m1[0], Y = cy[q] ? m2[1] : m1[1], Z = cz[q] ? m2[2] : m1[2];
    const float wx = cx[q] ? d2[0] : d1[0], wy = cy[q] ? d2[1] : d1[1], wz = cz[q] ? d2[2] : d1[2];
    /* the corner's three components in one 12-byte load */
    const F3 Fv = *reinterpret_cast<const F3 *>(force + 3 * (((int64_t)Z * n + Y) * n + X));
    const float F[3] = {Fv.x, Fv.y, Fv.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float gq = DIV ? div_6pi((F[a] * a_mid) * dt, ok) : F[a];
      const float inc = ((gq * wx) * wy) * wz;
      vreal[a] = vreal[a] + (double)inc;
    }
  }
  if (DIV && !ok) { /* a quotient outside the checked range: redo with divisions */
#pragma unroll
    for (int a = 0; a < 3; ++a) vreal[a] = v0[a];
    for (int q = 0; q < 8; ++q) {
      const int X = cx[q] ? m2[0] : m1[0], Y = cy[q] ? m2[1] : m1[1], Z = cz[q] ? m2[2] : m1[2];
      const float wx = cx[q] ? d2[0] : d1[0], wy = cy[q] ? d2[1] : d1[1], wz = cz[q] ? d2[2] : d1[2];
      const float *F = force + 3 * (((int64_t)Z * n + Y) * n + X);
      for (int a = 0; a < 3; ++a) {
        const float inc = ((((((F[a] * a_mid) * dt) / 6.0f) / CUBE_PI_F) * wx) * wy) * wz;
        vreal[a] = vreal[a] + (double)inc;
      }
    }
  }
}

/* max of |v + v_c| over the workgroup: wave shuffles, then the waves
   through LDS, one atomic per component into its slot region of
   vmax_bits (W/pm.f90:349) */
__device__ __forceinline__ void vmax_reduce(double vm[3], unsigned long long *__restrict__ vmax_bits) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int a = 0; a < 3; ++a) vm[a] = fmax(vm[a], __shfl_xor(vm[a], o, 64));
  __shared__ double red[3][KT / 64];
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; ++a) red[a][threadIdx.x >> 6] = vm[a];
  __syncthreads();
  if (threadIdx.x < 3) {
    double m = red[threadIdx.x][0];
    for (int w = 1; w < KT / 64; ++w) m = fmax(m, red[threadIdx.x][w]);
    atomicMax(cube_slot(vmax_bits + threadIdx.x * CUBE_SLOTS), (unsigned long long)__double_as_longlong(m));
  }
}

/* kicks, one workgroup per cell row (tile, k, j): particles of the row are
   contiguous; each thread finds its cell by binary search in LDS.
   NKICK = 1: the one kick of FINE (decode with vdec, encode with venc);
   NKICK = 2: the fine kick (vdec, venc) from the prescaled fine mesh that
   the step's force pipeline leaves in force_f, then the coarse one on the
   re-decoded code (vdec2, venc2) in the same pass, which is the step's
   sequence W/pm.f90:129-159 then 322-354 with the intermediate codes kept
   in registers instead of going through vp. */
#define KICK_ROW 2048 /* particles of a row whose cells the kick tables in LDS (else a binary search) */
template <int IX, int IV, bool FINE, int NKICK>
__global__ __launch_bounds__(KT) void k_kick(const typename IntOf<IX>::T *__restrict__ xp,
                                             typename IntOf<IV>::T *__restrict__ vp,
                                             const int64_t *__restrict__ cstart, const float *__restrict__ vfield,
                                             const float *__restrict__ force, const double *__restrict__ vdec,
                                             Geo g, int nvh, float a_mid, float dt, double venc,
                                             unsigned long long *__restrict__ vmax_bits,
                                             const float *__restrict__ force2, const double *__restrict__ vdec2,
                                             double venc2) {
  const int nt = g.nt;
  const int64_t row = blockIdx.x;
  const int j = (int)(row % nt), k = (int)((row / nt) % nt), t = (int)(row / ((int64_t)nt * nt));
  const int tz = t / (g.nnt * g.nnt), ty = (t / g.nnt) % g.nnt, tx = t % g.nnt;
  __shared__ int64_t cs[520];
  __shared__ uint16_t lcell[KICK_ROW]; /* slot -> cell of the row, when it fits */
  const int64_t c0 = row * nt;
  for (int i = threadIdx.x; i <= nt; i += KT) cs[i] = cstart[c0 + i];
  __syncthreads();
  const bool tab = cs[nt] - cs[0] <= KICK_ROW; /* uniform */
  if (tab) {
    for (int i = threadIdx.x; i < nt; i += KT)
      for (int64_t q = cs[i]; q < cs[i + 1]; ++q) lcell[q - cs[0]] = (uint16_t)i;
    __syncthreads();
  }
  constexpr bool VMAX = !FINE || NKICK == 2; /* the coarse kick reports vmax */
  double vm[3] = {0, 0, 0};
  for (int64_t p = cs[0] + threadIdx.x; p < cs[nt]; p += KT) {
    int lo = 0, hi = nt;
    if (tab) {
      lo = lcell[p - cs[0]];
    } else {
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cs[mid] <= p) lo = mid; else hi = mid;
      }
    }
    const int loc[3] = {lo + 1, j + 1, k + 1};
    const int tt[3] = {tx, ty, tz};
    int32_t code[3];
    double vreal[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      code[a] = (int32_t)xp[3 * p + a];
      vreal[a] = vdec[(int32_t)vp[3 * p + a] + nvh];
    }
    kick_particle<IX, FINE, FINE && NKICK == 2>(code, loc, tt, g, force, a_mid, dt, vreal);
    if (NKICK == 2) {
      /* the code the fine kick would store, decoded again for the coarse kick */
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int32_t v1 = (int32_t)(typename IntOf<IV>::T)cube_encode_v(IV, venc, vreal[a]);
        vreal[a] = vdec2[v1 + nvh];
      }
      kick_particle<IX, false>(code, loc, tt, g, force2, a_mid, dt, vreal);
    }
    const double ve = NKICK == 2 ? venc2 : venc;
    F3 vc = {0.f, 0.f, 0.f};
    if (VMAX) vc = *reinterpret_cast<const F3 *>(vfield + 3 * (c0 + lo)); /* one 12-byte load */
    const float vcf[3] = {vc.x, vc.y, vc.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (VMAX) {
        const double av = fabs(vreal[a] + (double)vcf[a]);
        vm[a] = av > vm[a] ? av : vm[a];
      }
      vp[3 * p + a] = (typename IntOf<IV>::T)cube_encode_v(IV, ve, vreal[a]);
    }
  }
  if (VMAX) vmax_reduce(vm, vmax_bits);
}

template <int IX, int IV, bool FINE, int NKICK>
static int kick_impl(cube_ctx *c, const float *force, float a_mid, float dt, double venc,
                     unsigned long long *vmb, const float *force2, double venc2) {
  const int64_t nrows = c->ncells / c->nt;
  hipLaunchKernelGGL((k_kick<IX, IV, FINE, NKICK>), dim3((unsigned)nrows), dim3(KT), 0, c->stream,
                     (const typename IntOf<IX>::T *)c->xp, (typename IntOf<IV>::T *)c->vp, c->cstart, c->vfield, force,
                     c->vdec0, geo_of(c), c->nvbin / 2, a_mid, dt, venc, vmb, force2, c->vdec1, venc2);
  CUBE_LAUNCH_CHECK();
  return CUBE_OK;
}
template <bool FINE, int NKICK = 1>
static int kick(cube_ctx *c, const float *force, float a_mid, float dt, double venc, unsigned long long *vmb,
                const float *force2 = nullptr, double venc2 = 0.0) {
  const int x = c->p.izipx, v = c->p.izipv;
  if (x == 1 && v == 1) return kick_impl<1, 1, FINE, NKICK>(c, force, a_mid, dt, venc, vmb, force2, venc2);
  if (x == 1) return kick_impl<1, 2, FINE, NKICK>(c, force, a_mid, dt, venc, vmb, force2, venc2);
  if (v == 1) return kick_impl<2, 1, FINE, NKICK>(c, force, a_mid, dt, venc, vmb, force2, venc2);
  return kick_impl<2, 2, FINE, NKICK>(c, force, a_mid, dt, venc, vmb, force2, venc2);
}

/* fine kick, W/pm.f90:129-159: decode with sigma_vi, encode with sigma_vi_new */
int cube_fine_kick(cube_ctx *c, float a_mid, float dt, float sigma_vi, float sigma_vi_new, int sync) {
  int st = cube_vdec_table(c, c->vdec0, sigma_vi);
  if (st) return st;
  st = kick<true>(c, c->force_f, a_mid, dt, cube_vconst(sigma_vi_new), nullptr);
  if (st) return st;
  if (sync) CUBE_CHECK(hipStreamSynchronize(c->stream));
  return CUBE_OK;
}

/* The step's two kicks in one pass over the particles (W/pm.f90:129-159,
   then :193 and :322-387): the fine kick decodes with sigma_vi and encodes
   with sigma_vi_new, the coarse kick decodes and encodes with sigma_vi_new;
   vmax (W/pm.f90:349) as cube_coarse_kick.  force_f holds the fine kick's
   factors (((F*a_mid)*dt)/6)/pi, formed by the step's force pipeline for
   this a_mid and dt.  Equal to cube_fine_kick then
   cube_coarse_kick(sigma_vi_new) on the plain force bit for bit. */
int cube_pm_kicks(cube_ctx *c, float a_mid, float dt, float sigma_vi, float sigma_vi_new, float *vmax) {
  int st = cube_vdec_table(c, c->vdec0, sigma_vi);
  if (!st) st = cube_vdec_table(c, c->vdec1, sigma_vi_new);
  if (st) return st;
  unsigned long long *vmb = (unsigned long long *)c->ured + 8, *vsl = c->slots + CUBE_SLOT_VMAX * CUBE_SLOTS;
  if ((st = cube_slots_zero(c, vsl, 3))) return st;
  float *G = reinterpret_cast<float *>(c->specd_c);
  hipLaunchKernelGGL(k_kick_prescale, dim3(2048), dim3(256), 0, c->stream, c->force_c, 3 * c->ncells, a_mid, dt, G);
  CUBE_LAUNCH_CHECK();
  const double venc = cube_vconst(sigma_vi_new);
  st = kick<true, 2>(c, c->force_f, a_mid, dt, venc, vsl, G, venc);
  if (st) return st;
  if ((st = cube_slots_fold(c, vsl, 3, 0, vmb, nullptr))) return st;
  unsigned long long h[3];
  CUBE_CHECK(hipMemcpyAsync(h, vmb, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  CUBE_CHECK(hipStreamSynchronize(c->stream));
  for (int a = 0; a < 3; ++a) {
    double d;
    memcpy(&d, &h[a], sizeof(d));
    if (vmax) vmax[a] = (float)d; /* vmax is real(4) (W/variables.f90:35) */
  }
  return CUBE_OK;
}

/* coarse kick, W/pm.f90:322-387: decode and encode with sigma_vi */
int cube_coarse_kick(cube_ctx *c, float a_mid, float dt, float sigma_vi, float *vmax) {
  int st = cube_vdec_table(c, c->vdec0, sigma_vi);
  if (st) return st;
  unsigned long long *vmb = (unsigned long long *)c->ured + 8, *vsl = c->slots + CUBE_SLOT_VMAX * CUBE_SLOTS;
  if ((st = cube_slots_zero(c, vsl, 3))) return st;
  /* G in the coarse spectra scratch, free once force_c is formed */
  float *G = reinterpret_cast<float *>(c->specd_c);
  hipLaunchKernelGGL(k_kick_prescale, dim3(2048), dim3(256), 0, c->stream, c->force_c, 3 * c->ncells, a_mid, dt, G);
  CUBE_LAUNCH_CHECK();
  st = kick<false>(c, G, a_mid, dt, cube_vconst(sigma_vi), vsl);
  if (st) return st;
  if ((st = cube_slots_fold(c, vsl, 3, 0, vmb, nullptr))) return st;
  unsigned long long h[3];
  CUBE_CHECK(hipMemcpyAsync(h, vmb, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  CUBE_CHECK(hipStreamSynchronize(c->stream));
  for (int a = 0; a < 3; ++a) {
    double d;
    memcpy(&d, &h[a], sizeof(d));
    if (vmax) vmax[a] = (float)d; /* vmax is real(4) (W/variables.f90:35) */
  }
  return CUBE_OK;
}

/* ---------------- Green's functions ---------------- */
/* real-space fine kernel of component d on the nf^3 mesh: the 16^3 table at
   offsets 0..15, odd-mirrored in axis d and even in the others
   (W/kernel_f.f90:25-31), corner |o|<=pp_range zeroed when PP is on */
__global__ void k_fine_kernel_rs(float *rs, int n, const float *tab, int d, int pp_on, int pp_range) {
  const int64_t n3 = (int64_t)n * n * n;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n3; q += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(q % n), y = (int)((q / n) % n), z = (int)(q / ((int64_t)n * n));
    const int o[3] = {x <= n / 2 ? x : x - n, y <= n / 2 ? y : y - n, z <= n / 2 ? z : z - n};
    float v = 0.f;
    if (o[0] >= -15 && o[0] <= 15 && o[1] >= -15 && o[1] <= 15 && o[2] >= -15 && o[2] <= 15) {
      const int a0 = o[0] < 0 ? -o[0] : o[0], a1 = o[1] < 0 ? -o[1] : o[1], a2 = o[2] < 0 ? -o[2] : o[2];
      const bool zero = pp_on && a0 <= pp_range && a1 <= pp_range && a2 <= pp_range;
      if (!zero) {
        v = tab[((a2 * 16 + a1) * 16 + a0) * 3 + d];
        if (o[d] < 0) v = -v;
      }
    }
    rs[q] = v;
  }
}
/* keep Im of the positive-frequency octant 0..n/2 of component d in the
   blocked storage of cube_kernf_index (columns past n/2 of the last block
   hold zero) */
__global__ void k_extract_octant(const float2 *spec, int n, float *kern, int d) {
  const int h = n / 2 + 1;
  const int64_t kb = cube_kernf_blocks(h), tot = (int64_t)h * kb * h * 16;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
    const int kz = (int)((q / 16) % h), b = (int)((q / (16 * h)) % kb), ky = (int)(q / (16 * h * kb));
    const int kx = b * 16 + (int)(q % 16);
    kern[(int64_t)d * tot + q] = kx < h ? spec[((int64_t)kz * n + ky) * cube_spec_pitch(n) + kx].y : 0.f;
  }
}
/* uncorrected coarse kernel -r/r^3 in fine units (W/kernel_c.f90:16-28) and
   the table-corrected one (:31-72), component d */
__global__ void k_coarse_kernel_rs(float *rs_plain, float *rs_corr, int nc, const float *ctab, int d) {
  const int64_t n3 = (int64_t)nc * nc * nc;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n3; q += (int64_t)gridDim.x * blockDim.x) {
    const int ii[3] = {(int)(q % nc), (int)((q / nc) % nc), (int)(q / ((int64_t)nc * nc))};
    float rx[3];
    for (int a = 0; a < 3; ++a) rx[a] = (float)(4 * ((ii[a] + nc / 2) % nc - nc / 2));
    const float r = sqrtf((rx[0] * rx[0] + rx[1] * rx[1]) + rx[2] * rx[2]);
    const float v = r == 0.f ? 0.f : -rx[d] / (r * (r * r));
    rs_plain[q] = v;
    /* octant copies of the 4^3 table: low block 0..3 or high block nc-3..nc-1
       (mirrored index nc-i), sign flipped for component d in a high block */
    int ti[3];
    bool in = true;
    float sgn = 1.f;
    for (int a = 0; a < 3; ++a) {
      if (ii[a] < 4) ti[a] = ii[a];
      else if (ii[a] >= nc - 3) {
        ti[a] = nc - ii[a];
        if (a == d) sgn = -sgn;
      } else in = false;
    }
    rs_corr[q] = in ? sgn * ctab[((ti[2] * 4 + ti[1]) * 4 + ti[0]) * 3 + d] : v;
  }
}
/* LRCKCORR (W/kernel_c.f90:301-325), all real(4) */
__global__ void k_lrckcorr(const float2 *spec_corr, const float2 *spec_plain, int nc, float *kern, int d) {
  const int nh = nc / 2 + 1;
  const int64_t tot = (int64_t)nh * nc * nc;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
    const int ix = (int)(q % nh), iy = (int)((q / nh) % nc), iz = (int)(q / ((int64_t)nh * nc));
    const int ki[3] = {(ix + nc / 2) % nc - nc / 2, (iy + nc / 2) % nc - nc / 2, (iz + nc / 2) % nc - nc / 2};
    float kx[3], ks[3];
    for (int a = 0; a < 3; ++a) {
      kx[a] = (float)ki[a];
      ks[a] = 2.0f * sinf(CUBE_PI_F * kx[a] / (float)nc);
    }
    const float kr = sqrtf((kx[0] * kx[0] + kx[1] * kx[1]) + kx[2] * kx[2]);
    const int64_t sq = ((int64_t)iz * nc + iy) * cube_spec_pitch(nc) + ix;
    const float kc = spec_corr[sq].y;
    const float cu = spec_plain[sq].y;
    float out = kc;
    /* for nc < 24 the |k|<=8 sphere reaches the Nyquist plane of axis d,
       where Im FFT of the odd kernel is exactly zero and the reference
       divides by it (NaN); those modes keep the uncorrected value */
    if (!(kr > 8.0f || ki[d] == 0 || 2 * ki[d] == -nc)) {
      const float ssum = (ks[0] * ks[0] + ks[1] * ks[1]) + ks[2] * ks[2];
      out = ((((kc * 0.25f) * CUBE_PI_F) * ks[d]) / ssum) / cu;
    }
    kern[(int64_t)d * tot + q] = out;
  }
}

int cube_green_init(cube_ctx *c, const float *wfxyzf, const float *wfxyzc) {
  float *dtab = nullptr, *dctab = nullptr;
  CUBE_CHECK(hipMalloc(&dtab, 16 * 16 * 16 * 3 * sizeof(float)));
  CUBE_CHECK(hipMalloc(&dctab, 4 * 4 * 4 * 3 * sizeof(float)));
  CUBE_CHECK(hipMemcpy(dtab, wfxyzf, 16 * 16 * 16 * 3 * sizeof(float), hipMemcpyHostToDevice));
  CUBE_CHECK(hipMemcpy(dctab, wfxyzc, 4 * 4 * 4 * 3 * sizeof(float), hipMemcpyHostToDevice));
  int st = CUBE_OK;
  const int nf = (int)c->nf, nc = (int)c->nc;
  for (int d = 0; d < 3 && !st; ++d) {
    hipLaunchKernelGGL(k_fine_kernel_rs, dim3(2048), dim3(256), 0, c->stream, c->rho_f, nf, dtab, d, c->p.pp_on,
                       c->p.pp_range);
    if (hipGetLastError() != hipSuccess) st = CUBE_ERR_HIP;
    if (!st) st = cube_fft_forward_full(c, c->plan_f, c->rho_f, c->spec_f);
    if (!st) {
      hipLaunchKernelGGL(k_extract_octant, dim3(2048), dim3(256), 0, c->stream, c->spec_f, nf, c->kern_f, d);
      if (hipGetLastError() != hipSuccess) st = CUBE_ERR_HIP;
    }
  }
  for (int d = 0; d < 3 && !st; ++d) {
    float *plain = c->rho_c, *corr = (float *)c->specd_c; /* specd_c is large enough for nc^3 floats */
    float2 *sp_plain = c->spec_c, *sp_corr = c->specd_c + (int64_t)cube_spec_pitch(nc) * nc * nc;
    hipLaunchKernelGGL(k_coarse_kernel_rs, dim3(1024), dim3(256), 0, c->stream, plain, corr, nc, dctab, d);
    if (hipGetLastError() != hipSuccess) st = CUBE_ERR_HIP;
    if (!st) st = cube_fft_forward_full(c, c->plan_c, plain, sp_plain);
    /* corr lives in specd_c plane 0; transform it into plane 1 */
    if (!st) st = cube_fft_forward_full(c, c->plan_c, corr, sp_corr);
    if (!st) {
      hipLaunchKernelGGL(k_lrckcorr, dim3(1024), dim3(256), 0, c->stream, sp_corr, sp_plain, nc, c->kern_c, d);
      if (hipGetLastError() != hipSuccess) st = CUBE_ERR_HIP;
    }
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(dtab);
  (void)hipFree(dctab);
  return st;
}

/* ---------------- fine potential ----------------
 * The reference's main path never forms a potential (its fine force kernel
 * goes from i*K*rho_k straight to forces); it ships the fine potential
 * kernel table CUBEnu/kernels/wg1.ascii (CUBEnu/kernels/README:8-9), 16^3
 * values at offsets 0..15.  Mirrored evenly into the periodic nf^3 mesh
 * (zero beyond offset 15, the near corner zeroed with PP on as in the force
 * kernel) its spectrum is real, and phi = C2R(Re K * R2C(rho_f)) / nf^3,
 * divided by nf three times in real(4) like the force (W/pm.f90:122-123). */
__global__ void k_fine_pot_rs(float *rs, int n, const float *tab, int pp_on, int pp_range) {
  const int64_t n3 = (int64_t)n * n * n;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n3; q += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(q % n), y = (int)((q / n) % n), z = (int)(q / ((int64_t)n * n));
    const int a0 = min(x, n - x), a1 = min(y, n - y), a2 = min(z, n - z);
    float v = 0.f;
    if (a0 <= 15 && a1 <= 15 && a2 <= 15 && !(pp_on && a0 <= pp_range && a1 <= pp_range && a2 <= pp_range))
      v = tab[(a2 * 16 + a1) * 16 + a0];
    rs[q] = v;
  }
}
__global__ void k_mul_re(const float2 *spec, const float2 *kspec, float2 *out, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const float2 s = spec[q];
    const float k = kspec[q].x;
    out[q] = make_float2(s.x * k, s.y * k);
  }
}
__global__ void k_div3(float *a, int64_t n3, float n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n3; q += (int64_t)gridDim.x * blockDim.x) {
    float v = a[q];
    v = v / n;
    v = v / n;
    v = v / n;
    a[q] = v;
  }
}

int cube_fine_potential(cube_ctx *c, const float *wg1) {
  const int nf = (int)c->nf;
  const int64_t nf3 = (int64_t)nf * nf * nf, ns = (int64_t)cube_spec_pitch(nf) * nf * nf;
  float *dtab = nullptr;
  CUBE_CHECK(hipMalloc(&dtab, 16 * 16 * 16 * sizeof(float)));
  int st = CUBE_OK;
  if (hipMemcpy(dtab, wg1, 16 * 16 * 16 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) st = CUBE_ERR_HIP;
  /* rho_k in spec_f; the kernel in force_f's first nf^3 floats, its
     spectrum in force spectrum 0, the product back in spec_f; phi into
     force_f's first nf^3 floats */
  if (!st) st = cube_fft_forward_full(c, c->plan_f, c->rho_f, c->spec_f);
  if (!st) {
    hipLaunchKernelGGL(k_fine_pot_rs, dim3(2048), dim3(256), 0, c->stream, c->force_f, nf, dtab, c->p.pp_on,
                       c->p.pp_range);
    if (hipGetLastError() != hipSuccess) st = CUBE_ERR_HIP;
  }
  if (!st) st = cube_fft_forward_full(c, c->plan_f, c->force_f, c->specd_f);
  if (!st) {
    hipLaunchKernelGGL(k_mul_re, dim3(2048), dim3(256), 0, c->stream, c->spec_f, c->specd_f, c->spec_f, ns);
    if (hipGetLastError() != hipSuccess) st = CUBE_ERR_HIP;
  }
  if (!st) st = cube_fft_inverse_full(c, c->plan_f, c->spec_f, c->force_f);
  if (!st) {
    hipLaunchKernelGGL(k_div3, dim3(2048), dim3(256), 0, c->stream, c->force_f, nf3, (float)nf);
    if (hipGetLastError() != hipSuccess) st = CUBE_ERR_HIP;
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(dtab);
  return st;
}
