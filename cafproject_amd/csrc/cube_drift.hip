/*
 * cube_drift.hip -- update_x (W/update_particle.f90:16-256) on the whole box.
 *
 * The reference drifts tile by tile over the tile's extended region
 * (physical + ncb buffer cells, whose particles are copies of neighbouring
 * tiles), counts particles into new cells (pass 1), builds prefix sums
 * (spine_tile), and scatters them (pass 2) in the visit order
 * (ilayer, k, j, i, l) with ilayer = (k-(1-ncb)) mod nlayer; only particles
 * landing in physical cells survive.  Both passes decode with the same
 * arithmetic, so a particle's new cell is fixed by the coordinates of the
 * copy that lands.
 *
 * Here there are no buffer copies.  Each particle evaluates, per axis, the
 * copies the reference would hold of it (home coordinate c, and c-nt in the
 * next tile if c>nt-ncb, and c+nt in the previous tile if c<=ncb), with the
 * reference's FP64 expression ceiling(xq+dt_mid*v/ncell) in each copy's own
 * tile frame, and keeps the copies that land in [1,nt] (normally exactly
 * one).  Its order inside the destination cell is the 64-bit key
 *   ilayer(9) | k_src-k_dst+2ncb (5) | j.. (5) | i.. (5) | particle index (40)
 * which, for a fixed destination cell, sorts exactly like (ilayer,k,j,i,l).
 * Counting (atomics on integers), an exclusive scan, atomic slot placement
 * and a per-cell sort of the keys give the reference order independently of
 * the atomics' timing; a per-cell pass then accumulates vfield_new in that
 * order (FP64 add, FP32 store, as the reference's real(4) vfield_new) and
 * re-quantises x and v.
 */
#include "cube_internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>

#define DT 256
static const double WEIGHT_V = (double)0.1f; /* real(8),parameter :: weight_v=0.1 */

enum { DERR_RANGE = 1, DERR_KEY = 2, DERR_XOVER = 8 };

/* ---------------- small utilities ---------------- */
__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_i64(int64_t *p, int64_t n, int64_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
static int fill_i32(cube_ctx *c, int32_t *p, int64_t n, int32_t v) {
  hipLaunchKernelGGL(k_fill_i32, dim3(1024), dim3(256), 0, c->stream, p, n, v);
  CUBE_LAUNCH_CHECK();
  return CUBE_OK;
}
static int fill_i64(cube_ctx *c, int64_t *p, int64_t n, int64_t v) {
  hipLaunchKernelGGL(k_fill_i64, dim3(64), dim3(256), 0, c->stream, p, n, v);
  CUBE_LAUNCH_CHECK();
  return CUBE_OK;
}

/* decoded residual velocity table: vdec[code+nvbin/2] =
   tan(pi*real(code)/real(nvbin-1)) / (sqrt(pi/2)/(sigma*vrel_boost)) */
__global__ void k_vdec(const float *tanf_tab, double *vdec, int nv, double vc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) vdec[i] = (double)tanf_tab[i] / vc;
}
int cube_vdec_table(cube_ctx *c, double *dst, float sigma) {
  const int nv = c->nvbin;
  hipLaunchKernelGGL(k_vdec, dim3((nv + 255) / 256), dim3(256), 0, c->stream, c->tanf_tab, dst, nv,
                     cube_vconst(sigma));
  CUBE_LAUNCH_CHECK();
  return CUBE_OK;
}

/* ---------------- exclusive scan of int32 counts -> int64 ---------------- */
#define SCAN_B 1024
__global__ __launch_bounds__(SCAN_B) void k_scan_blocks(const int32_t *cnt, int64_t n, int64_t *bsum) {
  __shared__ int64_t s[SCAN_B];
  const int64_t i = blockIdx.x * (int64_t)SCAN_B + threadIdx.x;
  s[threadIdx.x] = i < n ? cnt[i] : 0;
  __syncthreads();
  for (int o = SCAN_B / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = s[0];
}
__global__ __launch_bounds__(SCAN_B) void k_scan_partials(int64_t *bsum, int64_t nb, int64_t *total) {
  /* exclusive scan of nb block sums, one block, chunked */
  __shared__ int64_t s[SCAN_B];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < nb; base += SCAN_B) {
    const int64_t i = base + threadIdx.x;
    const int64_t v = i < nb ? bsum[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < SCAN_B; o <<= 1) {
      const int64_t t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == SCAN_B - 1) carry += s[SCAN_B - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}
__global__ __launch_bounds__(SCAN_B) void k_scan_final(const int32_t *cnt, int64_t n, const int64_t *boff,
                                                     int64_t *start, const int64_t *total) {
  __shared__ int64_t s[SCAN_B];
  const int64_t i = blockIdx.x * (int64_t)SCAN_B + threadIdx.x;
  const int64_t v = i < n ? cnt[i] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < SCAN_B; o <<= 1) {
    const int64_t t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) start[i] = boff[blockIdx.x] + s[threadIdx.x] - v;
  if (i == n - 1) start[n] = *total;
}
int cube_scan_counts(cube_ctx *c, const int32_t *cnt, int64_t *start, int64_t n, int64_t *total) {
  const int64_t nb = (n + SCAN_B - 1) / SCAN_B;
  int64_t *bsum = c->scanbuf;   /* nb + 1 entries */
  int64_t *tot = c->scanbuf + nb;
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(SCAN_B), 0, c->stream, cnt, n, bsum);
  CUBE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(SCAN_B), 0, c->stream, bsum, nb, tot);
  CUBE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_final, dim3((unsigned)nb), dim3(SCAN_B), 0, c->stream, cnt, n, bsum, start, tot);
  CUBE_LAUNCH_CHECK();
  if (total) {
    CUBE_CHECK(hipMemcpyAsync(total, tot, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  }
  CUBE_CHECK(hipStreamSynchronize(c->stream));
  return CUBE_OK;
}

/* ---------------- per-tile sums over the extended box ---------------- */
/* out[t] = sum of rhoc over local cells [1-w, nt+w]^3 of tile t (periodic) */
__global__ void k_tile_boxsum(const int32_t *rhoc, Geo g, int w, unsigned long long *out) {
  const int t = blockIdx.y;
  const int tz = t / (g.nnt * g.nnt), ty = (t / g.nnt) % g.nnt, tx = t % g.nnt;
  /* (nt+2w)^3 < 2^31 cells per box: 32-bit index arithmetic; the tile's
     own cells by their id, the buffer shell through cid_global */
  const int e = g.nt + 2 * w, nt = g.nt;
  const uint32_t vol = (uint32_t)e * e * e, ue = (uint32_t)e;
  int64_t s = 0;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < vol; q += gridDim.x * blockDim.x) {
    const uint32_t qe = q / ue;
    const int i = (int)(q - qe * ue) - w, j = (int)(qe % ue) - w, k = (int)(qe / ue) - w;
    const bool in = i >= 0 && i < nt && j >= 0 && j < nt && k >= 0 && k < nt;
    s += rhoc[in ? g.cid(i, j, k, t) : cid_global(g, tx * nt + i, ty * nt + j, tz * nt + k)];
  }
  __shared__ int64_t r[256];
  r[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) r[threadIdx.x] += r[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out + t, (unsigned long long)r[0]);
}
int cube_tile_boxsum(cube_ctx *c, const int32_t *rhoc, int w, int64_t *host_out) {
  int st = fill_i64(c, c->ired, c->ntiles, 0);
  if (st) return st;
  hipLaunchKernelGGL(k_tile_boxsum, dim3(256, (unsigned)c->ntiles), dim3(256), 0, c->stream, rhoc, geo_of(c), w,
                     (unsigned long long *)c->ired);
  CUBE_LAUNCH_CHECK();
  CUBE_CHECK(hipMemcpyAsync(host_out, c->ired, c->ntiles * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  CUBE_CHECK(hipStreamSynchronize(c->stream));
  return CUBE_OK;
}

/* ---------------- pass 1: destination cell and order key ---------------- */
/* three interleaved components, read with one 12-byte load */
struct V3f {
  float x, y, z;
};

struct DriftP {
  Geo g;
  int nlayer;
  int nvh;        /* nvbin/2 */
  float dt_mid;
  double xres;
  double inv_xres4; /* 1/(xres*ncell), a power of two: multiplying by it is the division, exactly */
  int agg_cap;    /* largest row (particles) whose slots are aggregated in LDS */
};
#define AGG_CAP 4096

template <int IX, int IV>
__global__ __launch_bounds__(DT) void k_drift_dest(const typename IntOf<IX>::T *__restrict__ xp,
                                                   const typename IntOf<IV>::T *__restrict__ vp,
                                                   const float *__restrict__ vfield,
                                                   const int64_t *__restrict__ cstart,
                                                   const double *__restrict__ vdec, DriftP P,
                                                   int32_t *__restrict__ dest, uint64_t *__restrict__ key,
                                                   int32_t *__restrict__ cnt, int32_t *__restrict__ rank,
                                                   int32_t *__restrict__ xdest, uint64_t *__restrict__ xkey,
                                                   int32_t *__restrict__ xrank, int64_t xcap,
                                                   unsigned long long *__restrict__ counters) {
  /* counters: [0] extra copies, [1] lost, [2] error flags */
  const Geo g = P.g;
  const int nt = g.nt, ncb = g.ncb;
  const int64_t row = blockIdx.x;               /* j + nt*(k + nt*t) */
  const int j = (int)(row % nt), k = (int)((row / nt) % nt), t = (int)(row / ((int64_t)nt * nt));
  const int tz = t / (g.nnt * g.nnt), ty = (t / g.nnt) % g.nnt, tx = t % g.nnt;
  __shared__ int64_t cs[520];
  /* Slots inside destination cells: a main copy landing in the 3x3 rows of
     cells around this row (same tile) takes a local rank from an LDS
     counter, and each such cell then reserves its range with one global
     atomic; other copies take their slot with a global atomic directly. */
  extern __shared__ int dyn[];
  const int nwin = 9 * nt;
  int *lcnt = dyn;               /* nwin counters, then the reserved bases */
  /* per particle of the row: window cell << 19 | local rank, or ~0u */
  uint32_t *pslot = reinterpret_cast<uint32_t *>(dyn + nwin);
  const int64_t c0 = row * nt;
  for (int i = threadIdx.x; i <= nt; i += DT) cs[i] = cstart[c0 + i];
  for (int i = threadIdx.x; i < nwin; i += DT) lcnt[i] = 0;
  __syncthreads();
  const int64_t pbeg = cs[0], pend = cs[nt];
  const bool agg = pend - pbeg <= P.agg_cap;   /* uniform over the block */
  if (agg) { /* pslot first holds each particle's cell of the row */
    for (int c = threadIdx.x; c < nt; c += DT)
      for (int64_t q = cs[c]; q < cs[c + 1]; ++q) pslot[q - pbeg] = (uint32_t)c;
    __syncthreads();
  }
  for (int64_t p = pbeg + threadIdx.x; p < pend; p += DT) {
    /* cell of p: largest i with cs[i] <= p */
    int lo = 0, hi = nt;
    if (agg) {
      lo = (int)pslot[p - pbeg];
      pslot[p - pbeg] = ~0u;
    } else {
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cs[mid] <= p) lo = mid; else hi = mid;
      }
    }
    const int i = lo;
    const int home[3] = {i + 1, j + 1, k + 1};   /* 1-based tile-local */
    const int tile[3] = {tx, ty, tz};
    int lcoord[3][3], lig[3][3], loff[3][3], nl[3];
    int err = 0;
    const V3f vcell = *reinterpret_cast<const V3f *>(vfield + 3 * (c0 + i)); /* one 12-byte load */
    const float vcf[3] = {vcell.x, vcell.y, vcell.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int32_t xc = (int32_t)xp[3 * p + a];
      const int32_t vc = (int32_t)vp[3 * p + a];
      const double off = ((double)cube_ucode(IX, xc) + 0.5) * P.xres;
      const double vreal = vdec[vc + P.nvh] + (double)vcf[a];
      const double dx = ((double)P.dt_mid * vreal) / 4.0;
      nl[a] = 0;
      int cand[3] = {home[a], 0, 0}, coff[3] = {0, 0, 0}, nc = 1;
      if (home[a] >= nt - ncb + 1) { cand[nc] = home[a] - nt; coff[nc] = 1; ++nc; }
      if (home[a] <= ncb) { cand[nc] = home[a] + nt; coff[nc] = -1; ++nc; }
      for (int q = 0; q < nc; ++q) {
        const double xq = ((double)cand[q] - 1.0) + off;
        const double ig = ceil(xq + dx);
        if (ig < 1 - 2 * ncb || ig > nt + 2 * ncb) { err |= DERR_RANGE; continue; }
        const int igi = (int)ig;
        if (igi >= 1 && igi <= nt) {
          lcoord[a][nl[a]] = cand[q];
          lig[a][nl[a]] = igi;
          loff[a][nl[a]] = coff[q];
          ++nl[a];
        }
      }
    }
    const int ncopy = nl[0] * nl[1] * nl[2];
    if (ncopy == 0) {
      dest[p] = -1;
      atomicAdd(counters + 1, 1ull);
    }
    int w = 0;
    for (int qz = 0; qz < nl[2]; ++qz)
      for (int qy = 0; qy < nl[1]; ++qy)
        for (int qx = 0; qx < nl[0]; ++qx, ++w) {
          const int q[3] = {qx, qy, qz};
          int dt_[3], r[3];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            dt_[a] = wrapi(tile[a] + loff[a][q[a]], g.nnt);
            r[a] = lcoord[a][q[a]] - lig[a][q[a]] + 2 * ncb;
            if (r[a] < 0 || r[a] > 31) err |= DERR_KEY;
          }
          const int il = (lcoord[2][qz] - (1 - ncb)) % P.nlayer;
          const int32_t dc = (int32_t)g.cid(lig[0][qx] - 1, lig[1][qy] - 1, lig[2][qz] - 1,
                                            dt_[0] + g.nnt * (dt_[1] + g.nnt * dt_[2]));
          const uint64_t kk = ((uint64_t)il << 55) | ((uint64_t)(r[2] & 31) << 50) |
                              ((uint64_t)(r[1] & 31) << 45) | ((uint64_t)(r[0] & 31) << 40) | (uint64_t)p;
          /* slots are taken in arrival order; the per-cell sort restores
             the reference order */
          if (w == 0) {
            dest[p] = dc;
            key[p] = kk;
            const int ddy = lig[1][qy] - 1 - j, ddz = lig[2][qz] - 1 - k;
            if (agg && dt_[0] == tx && dt_[1] == ty && dt_[2] == tz && ddy >= -1 && ddy <= 1 && ddz >= -1 &&
                ddz <= 1) {
              const int wi = ((ddz + 1) * 3 + (ddy + 1)) * nt + (lig[0][qx] - 1);
              const int lr = atomicAdd(lcnt + wi, 1);
              pslot[p - pbeg] = ((uint32_t)wi << 19) | (uint32_t)lr;
            } else {
              rank[p] = atomicAdd(cnt + dc, 1);
            }
          } else {
            const unsigned long long x = atomicAdd(counters + 0, 1ull);
            if ((int64_t)x < xcap) {
              xdest[x] = dc;
              xkey[x] = kk;
              xrank[x] = atomicAdd(cnt + dc, 1);
            } else {
              err |= DERR_XOVER;
            }
          }
        }
    if (err) atomicOr(counters + 2, (unsigned long long)err);
  }
  if (!agg) return; /* uniform */
  __syncthreads();
  /* reserve each window cell's range: one global atomic per cell */
  for (int wi = threadIdx.x; wi < nwin; wi += DT) {
    const int n = lcnt[wi];
    if (n > 0) {
      const int y = j + (wi / nt) % 3 - 1, z = k + (wi / nt) / 3 - 1;
      lcnt[wi] = atomicAdd(cnt + g.cid(wi % nt, y, z, t), n);
    }
  }
  __syncthreads();
  for (int64_t p = pbeg + threadIdx.x; p < pend; p += DT) {
    const uint32_t s = pslot[p - pbeg];
    if (s != ~0u) rank[p] = lcnt[s >> 19] + (int)(s & 0x7ffffu);
  }
}

/* ---------------- placement ---------------- */
__global__ void k_place(const int32_t *__restrict__ dest, const uint64_t *__restrict__ key, int64_t n,
                        const int64_t *__restrict__ start, int32_t *__restrict__ fill, uint64_t *__restrict__ out) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = dest[p];
    if (d < 0) continue;
    const int32_t s = atomicAdd(fill + d, 1);
    out[start[d] + s] = key[p];
  }
}

/* scatter to the slot taken at counting time: no second round of atomics */
__global__ void k_place_ranked(const int32_t *__restrict__ dest, const int32_t *__restrict__ rank,
                               const uint64_t *__restrict__ key, int64_t n, const int64_t *__restrict__ start,
                               uint64_t *__restrict__ out) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = dest[p];
    if (d < 0) continue;
    out[start[d] + rank[p]] = key[p];
  }
}

/* ---------------- per-cell sort ---------------- */
#define SMALL_SORT 48
#define MED_SORT 1024
#define BIG_SORT 16384
#define FB_CELLS 64      /* cells per finalize workgroup (the context sizes fbflag/vpart for it) */
/* slots staged in LDS per finalize pass: 640 (10 per cell against the 8 of
   an average cell) keeps the workgroup at ~22 KB of LDS, 7 per CU; a block
   with more slots takes the chunked path with the global key sort */
#define FB_CH 640
#define FB_LDS_SORT 64   /* largest cell sorted inside the finalize workgroup */
/* flag[b] = 1 when finalize block b (cells b*FB_CELLS...) holds at most FB_CH
   particles and no cell above FB_LDS_SORT: it sorts its cells in LDS */
__global__ void k_fb_flags(const int32_t *__restrict__ cnt, const int64_t *__restrict__ start, int64_t ncells,
                           uint8_t *__restrict__ flag) {
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t c0 = b * FB_CELLS;
  if (c0 >= ncells) return;
  const int64_t c1 = min(c0 + FB_CELLS, ncells);
  int mx = 0;
  for (int64_t c = c0; c < c1; ++c) mx = max(mx, (int)cnt[c]);
  flag[b] = (start[c1] - start[c0] <= FB_CH && mx <= FB_LDS_SORT) ? 1 : 0;
}
__global__ void k_sort_small(const int32_t *__restrict__ cnt, const int64_t *__restrict__ start, int64_t ncells,
                             uint64_t *__restrict__ a, int32_t *__restrict__ biglist,
                             unsigned long long *__restrict__ counters, const uint8_t *__restrict__ skip) {
  const int64_t cidx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (cidx >= ncells) return;
  if (skip && skip[cidx / FB_CELLS]) return; /* sorted by its finalize workgroup */
  const int n = cnt[cidx];
  if (n <= 1) return;
  if (n > MED_SORT) { /* big cells from the front of the list */
    const unsigned long long q = atomicAdd(counters + 3, 1ull);
    biglist[q] = (int32_t)cidx;
    return;
  }
  if (n > SMALL_SORT) { /* medium cells from the back */
    const unsigned long long q = atomicAdd(counters + 4, 1ull);
    biglist[ncells - 1 - (int64_t)q] = (int32_t)cidx;
    return;
  }
  uint64_t v[SMALL_SORT];
  uint64_t *s = a + start[cidx];
  for (int q = 0; q < SMALL_SORT; ++q)
    if (q < n) v[q] = s[q];
  /* insertion sort in registers (n <= 48) */
  for (int q = 1; q < SMALL_SORT; ++q) {
    if (q >= n) break;
    const uint64_t x = v[q];
    int r = q - 1;
    while (r >= 0 && v[r] > x) {
      v[r + 1] = v[r];
      --r;
    }
    v[r + 1] = x;
  }
  for (int q = 0; q < SMALL_SORT; ++q)
    if (q < n) s[q] = v[q];
}

/* bitonic sort of sk[0..m) in LDS, m a power of two */
__device__ static void lds_bitonic(uint64_t *sk, int m) {
  for (int size = 2; size <= m; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int q = threadIdx.x; q < m; q += blockDim.x) {
        const int o = q ^ stride;
        if (o > q) {
          const bool up = (q & size) == 0;
          const uint64_t x = sk[q], y = sk[o];
          if ((x > y) == up) {
            sk[q] = y;
            sk[o] = x;
          }
        }
      }
      __syncthreads();
    }
}
/* number of elements of the ascending run r[0..n) below x */
__device__ static int lower_bound_u64(const uint64_t *r, int n, uint64_t x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (r[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/* One wave per medium cell (SMALL_SORT < n <= MED_SORT), four per
   workgroup: a bitonic sort of the cell's keys (padded to a power of two,
   at least 64) in the wave's 8 KB of LDS, the stages ordered by the wave
   itself -- no workgroup barriers, and 32 KB per workgroup so that five
   share a CU where a big cell's workgroup takes 128 KB. */
__global__ __launch_bounds__(256) void k_sort_med(const int32_t *__restrict__ cnt, const int64_t *__restrict__ start,
                                                  const int32_t *__restrict__ list, int64_t nlist,
                                                  uint64_t *__restrict__ a) {
  __shared__ uint64_t sm[4 * MED_SORT];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t i = (int64_t)blockIdx.x * 4 + wv;
  if (i >= nlist) return; /* the whole wave */
  uint64_t *sk = sm + wv * MED_SORT;
  const int32_t cidx = list[i];
  const int n = cnt[cidx];
  uint64_t *s = a + start[cidx];
  int m = 64;
  while (m < n) m <<= 1;
  for (int q = lane; q < m; q += 64) sk[q] = q < n ? s[q] : ~0ull;
  for (int size = 2; size <= m; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int q = lane; q < m; q += 64) { /* each pair (q, q^stride) has one owner */
        const int o = q ^ stride;
        if (o > q) {
          const bool up = (q & size) == 0;
          const uint64_t x = sk[q], y = sk[o];
          if ((x > y) == up) {
            sk[q] = y;
            sk[o] = x;
          }
        }
      }
    }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int q = lane; q < n; q += 64) s[q] = sk[q];
}

/* One workgroup per big cell.  Up to BIG_SORT keys: a bitonic sort in LDS.
   Larger cells (a collapsed halo core can hold more particles than LDS
   takes): BIG_SORT-key runs sorted in LDS, then pairwise merges of runs in
   global memory between the cell's segment and the same range of tmp (the
   placement keys, dead by now).  Keys within a cell are distinct (the low
   bits are the particle index), so an element's merged position is its
   index in its run plus its rank in the other run. */
__global__ __launch_bounds__(1024) void k_sort_big(const int32_t *__restrict__ cnt,
                                                   const int64_t *__restrict__ start,
                                                   const int32_t *__restrict__ biglist, uint64_t *__restrict__ a,
                                                   uint64_t *__restrict__ tmp) {
  extern __shared__ uint64_t sk[];
  const int32_t cidx = biglist[blockIdx.x];
  const int n = cnt[cidx];
  uint64_t *s = a + start[cidx];
  if (n <= BIG_SORT) {
    int m = 1;
    while (m < n) m <<= 1;
    for (int q = threadIdx.x; q < m; q += blockDim.x) sk[q] = q < n ? s[q] : ~0ull;
    __syncthreads();
    lds_bitonic(sk, m);
    for (int q = threadIdx.x; q < n; q += blockDim.x) s[q] = sk[q];
    return;
  }
  for (int c0 = 0; c0 < n; c0 += BIG_SORT) {
    const int len = min(BIG_SORT, n - c0);
    for (int q = threadIdx.x; q < BIG_SORT; q += blockDim.x) sk[q] = q < len ? s[c0 + q] : ~0ull;
    __syncthreads();
    lds_bitonic(sk, BIG_SORT);
    for (int q = threadIdx.x; q < len; q += blockDim.x) s[c0 + q] = sk[q];
    __syncthreads();
  }
  uint64_t *src = s, *dst = tmp + start[cidx];
  for (int w = BIG_SORT; w < n; w <<= 1) {
    for (int lo = 0; lo < n; lo += 2 * w) {
      const int mid = min(lo + w, n), hi = min(lo + 2 * w, n);
      for (int q = lo + (int)threadIdx.x; q < hi; q += blockDim.x) {
        const uint64_t x = src[q];
        const int pos = q < mid ? (q - lo) + lower_bound_u64(src + mid, hi - mid, x)
                                : (q - mid) + lower_bound_u64(src + lo, mid - lo, x);
        dst[lo + pos] = x;
      }
    }
    __syncthreads();
    uint64_t *t = src;
    src = dst;
    dst = t;
  }
  if (src != s) {
    for (int q = threadIdx.x; q < n; q += blockDim.x) s[q] = src[q];
  }
}

/* ---------------- per-cell finalize ---------------- */
/*
 * Finalize by slot ranges: a workgroup owns FB_CELLS consecutive new cells
 * and the contiguous slot range of their particles.  Lanes gather the
 * decoded velocities of consecutive slots into LDS (one lane per slot), one
 * thread per cell then accumulates vfield_new serially in slot order (the
 * reference's order), and lanes write the re-quantised x, v and pid of
 * consecutive slots, so the stores are coalesced.  The velocity statistics
 * of the new state (W/update_particle.f90:172-224) are reduced on the way:
 * per workgroup in a fixed tree, then over workgroups in order.
 */
template <int IX, int IV>
__global__ __launch_bounds__(256) void k_drift_final2(
    const typename IntOf<IX>::T *__restrict__ xp, const typename IntOf<IV>::T *__restrict__ vp,
    const int32_t *__restrict__ pid, const float *__restrict__ vfield_old, const int64_t *__restrict__ start,
    const uint64_t *__restrict__ keys, const uint8_t *__restrict__ flag, const double *__restrict__ vdec, DriftP P,
    double venc, typename IntOf<IX>::T *__restrict__ xp2, typename IntOf<IV>::T *__restrict__ vp2,
    int32_t *__restrict__ pid2, float *__restrict__ vfield_new, int64_t ncells, double *__restrict__ part) {
  __shared__ int64_t cst[FB_CELLS + 1];
  __shared__ double vr[FB_CH][3]; /* decoded velocities; the final reduction reuses it */
  __shared__ uint64_t pix[FB_CH]; /* sorted keys, then source particle indices */
  __shared__ uint8_t lcell[FB_CH]; /* slot -> local cell */
  __shared__ int cgeo[FB_CELLS][4]; /* local i, j, k and tile of each cell */
  __shared__ float vn[FB_CELLS][3];
  const Geo g = P.g;
  const int nt = g.nt, ncb = g.ncb, tid = threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.x * FB_CELLS;
  const int ncl = (int)min((int64_t)FB_CELLS, ncells - c0);
  for (int i = tid; i <= ncl; i += 256) cst[i] = start[c0 + i];
  if (tid < ncl) {
    const int64_t cidx = c0 + tid;
    cgeo[tid][0] = (int)(cidx % nt);
    cgeo[tid][1] = (int)((cidx / nt) % nt);
    cgeo[tid][2] = (int)((cidx / ((int64_t)nt * nt)) % nt);
    cgeo[tid][3] = (int)(cidx / ((int64_t)nt * nt * nt));
  }
  __syncthreads();
  const int64_t S0 = cst[0], S1 = cst[ncl];
  /* every slot fits (and every cell is small): sort the keys of each cell
     here and keep the gathers in LDS for the write pass */
  const bool one = flag[blockIdx.x] != 0;
  if (one) {
    for (int64_t s = S0 + tid; s < S1; s += 256) pix[s - S0] = keys[s];
    if (tid < ncl) {
      const int lo = (int)(cst[tid] - S0), hi = (int)(cst[tid + 1] - S0);
      for (int q = lo; q < hi; ++q) lcell[q] = (uint8_t)tid;
    }
    __syncthreads();
    /* each key's place in its cell is the number of smaller keys there
       (unique: they carry the particle index); every lane ranks its own
       slots at once instead of one lane per cell sorting serially */
    constexpr int PER = (FB_CH + 255) / 256;
    uint64_t kx[PER];
    int dst[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int q = tid + 256 * u;
      dst[u] = -1;
      if (q < (int)(S1 - S0)) {
        const int c = lcell[q], lo = (int)(cst[c] - S0), hi = (int)(cst[c + 1] - S0);
        const uint64_t x = pix[q];
        int r = 0;
        for (int w = lo; w < hi; ++w) r += pix[w] < x ? 1 : 0;
        kx[u] = x;
        dst[u] = lo + r;
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u)
      if (dst[u] >= 0) pix[dst[u]] = kx[u];
    __syncthreads();
  }
  auto cell_of = [&](int64_t s) -> int {
    if (one) return lcell[s - S0];
    int lo = 0, hi = ncl;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (cst[mid] <= s) lo = mid; else hi = mid;
    }
    return lo;
  };

  /* slot -> (local cell, source particle, decoded velocity) */
  auto gather = [&](int64_t s, int &lc, int64_t &p, double v[3]) {
    lc = cell_of(s);
    const int i = cgeo[lc][0], j = cgeo[lc][1], k = cgeo[lc][2], t = cgeo[lc][3];
    const uint64_t kk = one ? pix[s - S0] : keys[s];
    p = (int64_t)(kk & KEY_PMASK);
    const int rx = (int)((kk >> 40) & 31), ry = (int)((kk >> 45) & 31), rz = (int)((kk >> 50) & 31);
    /* source cell = destination + (r - 2ncb), in the destination tile frame */
    const int sx = i + rx - 2 * ncb, sy = j + ry - 2 * ncb, sz = k + rz - 2 * ncb;
    int64_t sc;
    if (sx >= 0 && sx < nt && sy >= 0 && sy < nt && sz >= 0 && sz < nt) {
      sc = g.cid(sx, sy, sz, t);
    } else {
      const int tz = t / (g.nnt * g.nnt), ty = (t / g.nnt) % g.nnt, tx = t % g.nnt;
      sc = cid_global(g, tx * nt + sx, ty * nt + sy, tz * nt + sz);
    }
    const V3f vo = *reinterpret_cast<const V3f *>(vfield_old + 3 * sc); /* one 12-byte load */
    const float vof[3] = {vo.x, vo.y, vo.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = vdec[(int32_t)vp[3 * p + a] + P.nvh] + (double)vof[a];
  };

  /* vfield_new: thread (component aa, cell ca) sums its cell's slots in
     slot order -- three independent serial chains per cell, so a crowded
     cell holds its workgroup a third as long */
  const int ca = tid & 63, aa = tid >> 6;
  const bool accum = aa < 3 && ca < ncl;
  float acc = 0.f;
  if (accum) acc = (float)((double)vfield_old[3 * (c0 + ca) + aa] * WEIGHT_V);
  for (int64_t ch = S0; ch < S1; ch += FB_CH) {
    const int64_t ce = min(ch + FB_CH, S1);
    for (int64_t s = ch + tid; s < ce; s += 256) {
      int lc;
      int64_t p;
      double v[3];
      gather(s, lc, p, v);
      vr[s - ch][0] = v[0];
      vr[s - ch][1] = v[1];
      vr[s - ch][2] = v[2];
      pix[s - ch] = (uint64_t)p;
    }
    __syncthreads();
    if (accum) {
      const int64_t lo = max(cst[ca], ch), hi = min(cst[ca + 1], ce);
      for (int64_t q = lo; q < hi; ++q) acc = (float)((double)acc + vr[q - ch][aa]);
    }
    __syncthreads();
  }
  if (accum) {
    const double n = (double)(cst[ca + 1] - cst[ca]);
    vn[ca][aa] = (float)((double)acc / (n + WEIGHT_V));
    vfield_new[3 * (c0 + ca) + aa] = vn[ca][aa];
  }
  __syncthreads();
  double sc = 0, sr = 0, ss = 0;
  if (tid < ncl) {
    const float v0 = vn[tid][0], v1 = vn[tid][1], v2 = vn[tid][2];
    sc = (double)((v0 * v0 + v1 * v1) + v2 * v2);
  }
  for (int64_t s = S0 + tid; s < S1; s += 256) {
    int lc;
    int64_t p;
    double v[3];
    if (one) {
      lc = lcell[s - S0];
      p = (int64_t)pix[s - S0];
      v[0] = vr[s - S0][0];
      v[1] = vr[s - S0][1];
      v[2] = vr[s - S0][2];
    } else {
      gather(s, lc, p, v);
    }
    double r[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      /* xp_new=xp+nint(dt_mid*vreal/(x_resolution*ncell)) (:115) */
      const double qx = ((double)P.dt_mid * v[a]) * P.inv_xres4;
      xp2[3 * s + a] = (typename IntOf<IX>::T)cube_wrap(IX, (int64_t)xp[3 * p + a] + (int64_t)cube_nint(qx));
      /* vp_new=nint(...atan(...*(vreal-vfield_new))...) (:116-117) */
      const int32_t code = cube_encode_v(IV, venc, v[a] - (double)vn[lc][a]);
      vp2[3 * s + a] = (typename IntOf<IV>::T)code;
      r[a] = vdec[(int32_t)(typename IntOf<IV>::T)code + P.nvh];
    }
    if (pid) pid2[s] = pid[p];
    sr += (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
    r[0] += (double)vn[lc][0];
    r[1] += (double)vn[lc][1];
    r[2] += (double)vn[lc][2];
    ss += (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
  }
  __syncthreads(); /* vr is free: reduce through it */
  double(*red)[256] = reinterpret_cast<double(*)[256]>(&vr[0][0]);
  red[0][tid] = sc;
  red[1][tid] = sr;
  red[2][tid] = ss;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o)
      for (int a = 0; a < 3; ++a) red[a][tid] += red[a][tid + o];
    __syncthreads();
  }
  if (tid < 3) part[3 * blockIdx.x + tid] = red[tid][0];
}

/* fixed-order sum of nb triples: thread t sums a contiguous run of blocks,
   then a tree over the threads */
__global__ __launch_bounds__(1024) void k_sum_parts_tree(const double *__restrict__ part, int64_t nb,
                                                         double *__restrict__ out) {
  __shared__ double red[3][1024];
  const int t = threadIdx.x;
  /* thread t sums the partials b = t mod 1024 (a fixed order, so the sums
     repeat bit for bit; neighbouring lanes read neighbouring partials) */
  double s[3] = {0, 0, 0};
  for (int64_t b = t; b < nb; b += 1024)
#pragma unroll
    for (int a = 0; a < 3; ++a) s[a] += part[3 * b + a];
  for (int a = 0; a < 3; ++a) red[a][t] = s[a];
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (t < o)
      for (int a = 0; a < 3; ++a) red[a][t] += red[a][t + o];
    __syncthreads();
  }
  if (t < 3) out[t] = red[t][0];
}

/* ---------------- velocity statistics (:172-224) ---------------- */
#define VS_BLOCKS 1024
template <int IV>
__global__ __launch_bounds__(256) void k_vstats(const typename IntOf<IV>::T *__restrict__ vp,
                                                const float *__restrict__ vfield, const int32_t *__restrict__ cnt,
                                                const int64_t *__restrict__ start, const double *__restrict__ vdec,
                                                int nvh, int64_t ncells, double *__restrict__ part) {
  double sc = 0, sr = 0, ss = 0;
  for (int64_t cidx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; cidx < ncells;
       cidx += (int64_t)gridDim.x * blockDim.x) {
    const float v0 = vfield[3 * cidx], v1 = vfield[3 * cidx + 1], v2 = vfield[3 * cidx + 2];
    sc += (double)((v0 * v0 + v1 * v1) + v2 * v2);
    const int n = cnt[cidx];
    const int64_t s0 = start[cidx];
    for (int q = 0; q < n; ++q) {
      double r[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) r[a] = vdec[(int32_t)vp[3 * (s0 + q) + a] + nvh];
      sr += (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
      r[0] += (double)v0;
      r[1] += (double)v1;
      r[2] += (double)v2;
      ss += (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
    }
  }
  __shared__ double red[3][256];
  red[0][threadIdx.x] = sc;
  red[1][threadIdx.x] = sr;
  red[2][threadIdx.x] = ss;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int a = 0; a < 3; ++a) red[a][threadIdx.x] += red[a][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}
__global__ void k_sum_parts(const double *part, int nb, double *out) {
  /* fixed-order final sum of the block partials */
  if (threadIdx.x < 3) {
    double s = 0;
    for (int b = 0; b < nb; ++b) s += part[3 * b + threadIdx.x];
    out[threadIdx.x] = s;
  }
}
int cube_vstats_run(cube_ctx *c, float sigma_vi, double *sums3) {
  int st = cube_vdec_table(c, c->vdec0, sigma_vi);
  if (st) return st;
  const int nvh = c->nvbin / 2;
  if (c->p.izipv == 1)
    hipLaunchKernelGGL(k_vstats<1>, dim3(VS_BLOCKS), dim3(256), 0, c->stream, (const int8_t *)c->vp, c->vfield,
                       c->rhoc, c->cstart, c->vdec0, nvh, c->ncells, c->dred);
  else
    hipLaunchKernelGGL(k_vstats<2>, dim3(VS_BLOCKS), dim3(256), 0, c->stream, (const int16_t *)c->vp, c->vfield,
                       c->rhoc, c->cstart, c->vdec0, nvh, c->ncells, c->dred);
  CUBE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, c->stream, c->dred, VS_BLOCKS, c->dred + 3 * VS_BLOCKS);
  CUBE_LAUNCH_CHECK();
  CUBE_CHECK(hipMemcpyAsync(sums3, c->dred + 3 * VS_BLOCKS, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CUBE_CHECK(hipStreamSynchronize(c->stream));
  return CUBE_OK;
}

/* ---------------- helpers shared with the IC sort ---------------- */
__global__ void k_fill_zero_i32(int32_t *p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}
__global__ void k_place_keys(const int32_t *dest, const uint64_t *key, int64_t n, const int64_t *start, int32_t *fill,
                             uint64_t *out) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = dest[p];
    if (d < 0) continue;
    const int32_t s = atomicAdd(fill + d, 1);
    out[start[d] + s] = key[p];
  }
}
/* sort every cell's key segment ascending (small ones in registers, medium
   ones with a wave's bitonic sort, big ones with a workgroup's, huge ones by
   merging sorted runs);
   c->key, dead once the keys are placed, is the merge buffer */
extern "C" int cube_sort_small_segments(cube_ctx *c, const int32_t *cnt, const int64_t *start, uint64_t *keys) {
  unsigned long long *ctr = (unsigned long long *)c->ured;
  int st = fill_i64(c, (int64_t *)ctr + 3, 2, 0);
  if (st) return st;
  const unsigned nbc = (unsigned)((c->ncells + 255) / 256);
  int32_t *biglist = (int32_t *)c->dest;
  hipLaunchKernelGGL(k_sort_small, dim3(nbc), dim3(256), 0, c->stream, cnt, start, c->ncells, keys, biglist, ctr,
                     (const uint8_t *)nullptr);
  CUBE_LAUNCH_CHECK();
  unsigned long long hc[8];
  CUBE_CHECK(hipMemcpy(hc, ctr, sizeof(hc), hipMemcpyDeviceToHost));
  if (hc[4]) {
    hipLaunchKernelGGL(k_sort_med, dim3((unsigned)((hc[4] + 3) / 4)), dim3(256), 0, c->stream, cnt, start,
                       biglist + c->ncells - (int64_t)hc[4], (int64_t)hc[4], keys);
    CUBE_LAUNCH_CHECK();
  }
  if (hc[3]) {
    hipLaunchKernelGGL(k_sort_big, dim3((unsigned)hc[3]), dim3(1024), BIG_SORT * sizeof(uint64_t), c->stream, cnt,
                       start, biglist, keys, c->key);
    CUBE_LAUNCH_CHECK();
  }
  return CUBE_OK;
}

/* ---------------- driver ---------------- */
template <int IX, int IV> static int drift_impl(cube_ctx *c, DriftP P, double venc, int64_t *nxtra,
                                                unsigned long long *hc) {
  typedef typename IntOf<IX>::T TX;
  typedef typename IntOf<IV>::T TV;
  unsigned long long *ctr = (unsigned long long *)c->ured;
  int st;
  if ((st = fill_i32(c, c->rhoc2, c->ncells, 0))) return st;
  if ((st = fill_i64(c, (int64_t *)ctr, 8, 0))) return st;
  const int64_t nrows = c->ncells / c->nt;
  const size_t dyn = (size_t)(9 * c->nt + P.agg_cap) * sizeof(int);
  hipLaunchKernelGGL((k_drift_dest<IX, IV>), dim3((unsigned)nrows), dim3(DT), dyn, c->stream, (const TX *)c->xp,
                     (const TV *)c->vp, c->vfield, c->cstart, c->vdec0, P, c->dest, c->key, c->rhoc2, c->rank,
                     c->xdest, c->xkey, c->xrank, c->xcap, ctr);
  CUBE_LAUNCH_CHECK();
  int64_t total = 0;
  if ((st = cube_scan_counts(c, c->rhoc2, c->cstart2, c->ncells, &total))) return st;
  CUBE_CHECK(hipMemcpy(hc, ctr, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  if (hc[2] & (DERR_RANGE | DERR_KEY)) return CUBE_ERR_DISPLACEMENT;
  if (hc[2] & DERR_XOVER) return CUBE_ERR_IMAGE_OVERFLOW;
  if (total > c->cap) return CUBE_ERR_IMAGE_OVERFLOW;
  *nxtra = (int64_t)hc[0];
  hipLaunchKernelGGL(k_place_ranked, dim3(4096), dim3(256), 0, c->stream, c->dest, c->rank, c->key, c->nplocal,
                     c->cstart2, c->sortbuf);
  CUBE_LAUNCH_CHECK();
  if (hc[0]) {
    hipLaunchKernelGGL(k_place_ranked, dim3(64), dim3(256), 0, c->stream, c->xdest, c->xrank, c->xkey,
                       (int64_t)hc[0], c->cstart2, c->sortbuf);
    CUBE_LAUNCH_CHECK();
  }
  const unsigned nbc = (unsigned)((c->ncells + 255) / 256);
  const int64_t nfb = (c->ncells + FB_CELLS - 1) / FB_CELLS;
  hipLaunchKernelGGL(k_fb_flags, dim3((unsigned)((nfb + 255) / 256)), dim3(256), 0, c->stream, c->rhoc2, c->cstart2,
                     c->ncells, c->fbflag);
  CUBE_LAUNCH_CHECK();
  int32_t *biglist = (int32_t *)c->dest; /* dest is dead after placement */
  hipLaunchKernelGGL(k_sort_small, dim3(nbc), dim3(256), 0, c->stream, c->rhoc2, c->cstart2, c->ncells, c->sortbuf,
                     biglist, ctr, c->fbflag);
  CUBE_LAUNCH_CHECK();
  CUBE_CHECK(hipMemcpy(hc, ctr, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  if (hc[4]) {
    hipLaunchKernelGGL(k_sort_med, dim3((unsigned)((hc[4] + 3) / 4)), dim3(256), 0, c->stream, c->rhoc2,
                       c->cstart2, biglist + c->ncells - (int64_t)hc[4], (int64_t)hc[4], c->sortbuf);
    CUBE_LAUNCH_CHECK();
  }
  if (hc[3]) {
    hipLaunchKernelGGL(k_sort_big, dim3((unsigned)hc[3]), dim3(1024), BIG_SORT * sizeof(uint64_t), c->stream,
                       c->rhoc2, c->cstart2, biglist, c->sortbuf, c->key); /* key is dead after placement */
    CUBE_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL((k_drift_final2<IX, IV>), dim3((unsigned)nfb), dim3(256), 0, c->stream, (const TX *)c->xp,
                     (const TV *)c->vp, c->p.has_pid ? c->pid : nullptr, c->vfield, c->cstart2, c->sortbuf,
                     c->fbflag, c->vdec0, P, venc, (TX *)c->xp2, (TV *)c->vp2, c->pid2, c->vfield2, c->ncells,
                     c->vpart);
  CUBE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_parts_tree, dim3(1), dim3(1024), 0, c->stream, c->vpart, nfb, c->dred);
  CUBE_LAUNCH_CHECK();
  CUBE_CHECK(hipStreamSynchronize(c->stream));
  c->nplocal = total;
  return CUBE_OK;
}

int cube_drift(cube_ctx *c, float dt_old, float dt, float vz_max, float sigma_vi, cube_vstats *out) {
  /* dt_mid=(dt_old+dt)/2; nlayer=2*ceiling(dt_mid*vz_max/ncell)+1 (real(4)) */
  const float dt_mid = (dt_old + dt) / 2.0f;
  const int nlayer = 2 * (int)ceilf(dt_mid * vz_max / 4.0f) + 1;
  if (nlayer < 1 || nlayer > 511) return CUBE_ERR_DISPLACEMENT;
  /* tile overflow check (W/update_particle.f90:86-94): the copies in each
     tile's extended region must fit np_tile_max */
  int64_t *sums = (int64_t *)malloc(c->ntiles * sizeof(int64_t));
  tbeg(c);
  int st = cube_tile_boxsum(c, c->rhoc, (int)c->ncb, sums);
  tend(c, CUBE_T_CHECKS);
  double ov = 0;
  int64_t mx = 0;
  for (int64_t t = 0; t < c->ntiles; ++t) {
    const double o = (double)((float)sums[t] / (float)c->np_tile_max);
    ov = o > ov ? o : ov;
    mx = sums[t] > mx ? sums[t] : mx;
  }
  free(sums);
  if (st) return st;
  if (mx > c->np_tile_max) return CUBE_ERR_TILE_OVERFLOW;
  tbeg(c);
  if ((st = cube_vdec_table(c, c->vdec0, sigma_vi))) return st;
  DriftP P;
  P.g = geo_of(c);
  P.nlayer = nlayer;
  P.nvh = c->nvbin / 2;
  P.dt_mid = dt_mid;
  P.xres = c->p.izipx == 1 ? 1.0 / 256.0 : 1.0 / 65536.0;
  P.inv_xres4 = c->p.izipx == 1 ? 64.0 : 16384.0;
  P.agg_cap = AGG_CAP;
  /* a lower cap sends more rows down the all-global slot path (used by the
     tests to cover both paths in one run) */
  if (const char *e = getenv("CUBE_DRIFT_AGG_CAP"); e && *e) P.agg_cap = std::min(AGG_CAP, atoi(e));
  const double venc = cube_vconst(sigma_vi);
  int64_t nxtra = 0;
  unsigned long long hc[8];
  const int64_t np_old = c->nplocal;
  if (c->p.izipx == 1 && c->p.izipv == 1) st = drift_impl<1, 1>(c, P, venc, &nxtra, hc);
  else if (c->p.izipx == 1) st = drift_impl<1, 2>(c, P, venc, &nxtra, hc);
  else if (c->p.izipv == 1) st = drift_impl<2, 1>(c, P, venc, &nxtra, hc);
  else st = drift_impl<2, 2>(c, P, venc, &nxtra, hc);
  if (st) return st;
  tend(c, CUBE_T_DRIFT);
  /* the new state becomes current */
  void *tx = c->xp; c->xp = c->xp2; c->xp2 = tx;
  void *tv = c->vp; c->vp = c->vp2; c->vp2 = tv;
  int32_t *tp = c->pid; c->pid = c->pid2; c->pid2 = tp;
  int32_t *tr = c->rhoc; c->rhoc = c->rhoc2; c->rhoc2 = tr;
  float *tf = c->vfield; c->vfield = c->vfield2; c->vfield2 = tf;
  int64_t *ts = c->cstart; c->cstart = c->cstart2; c->cstart2 = ts;
  /* std of the velocity field, FP64 sums (:171-224), reduced by the
     finalize kernel of the drift */
  double s3[3];
  CUBE_CHECK(hipMemcpy(s3, c->dred, 3 * sizeof(double), hipMemcpyDeviceToHost));
  const double npg = (double)c->npglobal;
  cube_vstats v;
  memset(&v, 0, sizeof(v));
  v.sum_c = s3[0];
  v.sum_r = s3[1];
  v.sum_s = s3[2];
  v.std_vsim = sqrt(s3[2] / npg);
  v.std_vsim_c = sqrt(s3[0] / (double)c->nc / (double)c->nc / (double)c->nc);
  v.std_vsim_res = sqrt(s3[1] / npg);
  v.sigma_vi_new = (float)(v.std_vsim_res / (double)sqrtf(3.0f));
  v.nlayer = nlayer;
  v.nplocal = c->nplocal;
  v.lost = (int64_t)hc[1];
  v.overhead_tile = ov;
  (void)np_old;
  if (out) *out = v;
  if (c->nplocal != c->npglobal) return CUBE_ERR_COUNT;
  return CUBE_OK;
}
