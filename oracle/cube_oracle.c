/*
 * cube_oracle.c -- CPU restatement of the CUBE particle-mesh step.
 * TEST INFRASTRUCTURE ONLY (parity oracle and CPU baseline); see cube_oracle.h.
 *
 * Every routine cites the reference lines it restates (paths are relative to
 * the reference repository, W/ = CUBEnu/work/main/).  Floating-point
 * expressions keep the reference's evaluation order and kinds: real(4) work is
 * done in float, real(8) in double, mixed expressions promote exactly where
 * Fortran does.  Build with -ffp-contract=off (no FMA contraction), as the
 * reference's gfortran -O3 build on x86-64 has none.
 */
#include "cube_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <omp.h>

/* ---- constants (W/parameters.f90) ------------------------------------- */
#define PI_F 3.14159274101257324f            /* pi=4*atan(1.), real(4) */
static const double VREL_BOOST = 2.5;         /* W/parameters.f90:123 */
/* real(8),parameter :: weight_v=0.1 -- a default-real literal, so it is
   (double)0.1f (W/update_particle.f90:25) */
static const double WEIGHT_V = (double)0.1f;

/* ---- index helpers ----------------------------------------------------- */
#define LB(c) (1 - (c)->ncb)
static inline int64_t tile_id(const ocube *c, int64_t tx, int64_t ty, int64_t tz) {
  return (tx - 1) + c->nnt * ((ty - 1) + c->nnt * (tz - 1));
}
static inline int64_t cell(const ocube *c, int64_t i, int64_t j, int64_t k, int64_t t) {
  const int64_t b = LB(c), n = c->nte;
  return (i - b) + n * ((j - b) + n * ((k - b) + n * t));
}
#define RHOC(i, j, k, t) c->rhoc[cell(c, i, j, k, t)]
#define VF(d, i, j, k, t) c->vfield[3 * cell(c, i, j, k, t) + (d)]
static inline int64_t brow(const ocube *c, int64_t j, int64_t k, int64_t t) {
  const int64_t b = LB(c), n = c->nte;
  return (j - b) + n * ((k - b) + n * t);
}
static inline int64_t prow(const ocube *c, int64_t j, int64_t k, int64_t t) {
  return (j - 1) + c->nt * ((k - 1) + c->nt * t);
}
/* sum(rhoc(i:,j,k,t)) */
static inline int64_t row_suffix(const ocube *c, int64_t i, int64_t j, int64_t k, int64_t t) {
  int64_t s = 0;
  for (int64_t ii = i; ii <= c->nt + c->ncb; ii++) s += RHOC(ii, j, k, t);
  return s;
}
#define XP(d, ip) c->xp[3 * ((ip) - 1) + (d)]
#define VP(d, ip) c->vp[3 * ((ip) - 1) + (d)]

/* ---- integer kinds ------------------------------------------------------ */
static inline int32_t wrap_kind(int64_t bytes, int64_t v) {
  if (bytes == 1) return (int32_t)(int8_t)(uint8_t)(v & 0xff);
  return (int32_t)(int16_t)(uint16_t)(v & 0xffff);
}
/* Fortran nint: round half away from zero */
static inline int64_t f_nint(double x) { return (int64_t)llround(x); }

/* ---- codecs (spec ms_caf.tex:69-74, 114-119) ---------------------------- */
/* (int(xp+ishift,izipx)+rshift)*x_resolution = (unsigned(xp)+0.5)*2^-8izipx,
   W/update_particle.f90:66, W/parameters.f90:16-19,119 */
double oc_decode_x(int64_t izipx, int32_t code) {
  const int64_t ishift = -((int64_t)1 << (izipx * 8 - 1));
  const double rshift = 0.5 - (double)ishift;
  const double xres = 1.0 / (double)((int64_t)1 << (izipx * 8));
  return ((double)wrap_kind(izipx, (int64_t)code + ishift) + rshift) * xres;
}
static inline double x_res(int64_t izipx) { return 1.0 / (double)((int64_t)1 << (izipx * 8)); }

float oc_tanf_code(int64_t izipv, int32_t code) {
  const float nvm1 = (float)(((int64_t)1 << (8 * izipv)) - 1);
  return tanf((PI_F * (float)code) / nvm1);
}
/* sqrt(pi/2)/(sigma_vi*vrel_boost): real(4)/real(8) -> real(8) */
static inline double vconst(float sigma) {
  return (double)sqrtf(PI_F / 2.0f) / ((double)sigma * VREL_BOOST);
}
/* vreal=tan(pi*real(vp)/real(nvbin-1))/(sqrt(pi/2)/(sigma_vi*vrel_boost)),
   W/update_particle.f90:67 */
double oc_decode_v(int64_t izipv, int32_t code, float sigma) {
  return (double)oc_tanf_code(izipv, code) / vconst(sigma);
}
/* nint(real(nvbin-1)*atan(sqrt(pi/2)/(sigma*vrel_boost)*vreal)/pi,kind=izipv),
   W/update_particle.f90:117 */
int32_t oc_encode_v(int64_t izipv, double v, float sigma) {
  const double nvm1 = (double)(float)(((int64_t)1 << (8 * izipv)) - 1);
  const double y = (nvm1 * atan(vconst(sigma) * v)) / (double)PI_F;
  return wrap_kind(izipv, f_nint(y));
}
/* xp_new=xp+nint(dt_mid*vreal/(x_resolution*ncell)) with integer wrap,
   W/update_particle.f90:115; the caller passes dtv = dt_mid*vreal (ncell=4) */
int32_t oc_encode_x_drift(int64_t izipx, int32_t code, double dtv) {
  const double q = dtv / (x_res(izipx) * 4.0);
  return wrap_kind(izipx, (int64_t)code + f_nint(q));
}

/* ---- create / destroy ---------------------------------------------------- */
int oc_create(ocube *c, int64_t nc, int64_t nnt, int64_t ncb, int64_t izipx,
              int64_t izipv, int64_t np_nc, int64_t pp_on, int64_t pp_range,
              float rsoft, float image_buffer, float tile_buffer, int64_t has_pid) {
  memset(c, 0, sizeof(*c));
  if (nc <= 0 || nnt <= 0 || nc % nnt || ncb < 1 || (izipx != 1 && izipx != 2) ||
      (izipv != 1 && izipv != 2))
    return OC_ERR_INVALID;
  c->nc = nc; c->nnt = nnt; c->nt = nc / nnt; c->ncb = ncb; c->ncell = 4;
  c->nte = c->nt + 2 * ncb; c->nft = c->nt * c->ncell; c->nfb = ncb * c->ncell;
  c->nfe = c->nft + 2 * c->nfb; c->nf_cutoff = 16;
  c->izipx = izipx; c->izipv = izipv; c->np_nc = np_nc;
  c->pp_on = pp_on; c->pp_range = pp_range; c->rsoft = rsoft;
  c->image_buffer = image_buffer; c->tile_buffer = tile_buffer; c->has_pid = has_pid;
  c->nvbin = (int64_t)1 << (8 * izipv);
  /* W/variables.f90:8-10, evaluated in real(4) as the reference does */
  c->np_image = (nc * np_nc) * (nc * np_nc) * (nc * np_nc);
  {
    float r = (float)c->nte * 1.0f / (float)c->nt;
    float r3 = r * (r * r);
    c->np_image_max = (int64_t)((float)c->np_image * r3 * image_buffer);
    c->np_tile_max = (int64_t)((float)(c->np_image / (nnt * nnt * nnt)) * r3 * tile_buffer);
    c->np_pp_max = c->np_tile_max;
  }
  const int64_t ncells = c->nte * c->nte * c->nte * nnt * nnt * nnt;
  const int64_t nrows_b = c->nte * c->nte * nnt * nnt * nnt;
  const int64_t nrows_p = c->nt * c->nt * nnt * nnt * nnt;
  c->xp = calloc(3 * c->np_image_max, sizeof(int16_t));
  c->vp = calloc(3 * c->np_image_max, sizeof(int16_t));
  c->pid = has_pid ? calloc(c->np_image_max, sizeof(int32_t)) : NULL;
  c->rhoc = calloc(ncells, sizeof(int32_t));
  c->vfield = calloc(3 * ncells, sizeof(float));
  c->idx_b_l = calloc(nrows_b, sizeof(int64_t));
  c->idx_b_r = calloc(nrows_b, sizeof(int64_t));
  c->ppl0 = calloc(nrows_p, sizeof(int64_t)); c->pplr = calloc(nrows_p, sizeof(int64_t));
  c->pprl = calloc(nrows_p, sizeof(int64_t)); c->ppr0 = calloc(nrows_p, sizeof(int64_t));
  c->ppl = calloc(nrows_p, sizeof(int64_t)); c->ppr = calloc(nrows_p, sizeof(int64_t));
  if (!c->xp || !c->vp || (has_pid && !c->pid) || !c->rhoc || !c->vfield || !c->idx_b_l ||
      !c->idx_b_r || !c->ppl0 || !c->pplr || !c->pprl || !c->ppr0 || !c->ppl || !c->ppr) {
    oc_destroy(c);
    return OC_ERR_ALLOC;
  }
  return OC_OK;
}

void oc_destroy(ocube *c) {
  free(c->xp); free(c->vp); free(c->pid); free(c->rhoc); free(c->vfield);
  free(c->idx_b_l); free(c->idx_b_r); free(c->ppl0); free(c->pplr); free(c->pprl);
  free(c->ppr0); free(c->ppl); free(c->ppr);
  memset(c, 0, sizeof(*c));
}

/* ---- spine_image, W/variables.f90:141-211 ------------------------------- */
void oc_spine_image(ocube *c) {
  const int64_t nt = c->nt, ncb = c->ncb, nnt = c->nnt, lb = LB(c);
  const int64_t ntiles = nnt * nnt * nnt;
  int64_t *ctile = malloc(ntiles * sizeof(int64_t)), *ctile_p = malloc(ntiles * sizeof(int64_t));
  int64_t nsum = 0, nsum_p = 0;
  for (int64_t t = 0; t < ntiles; t++) {   /* ihz, ihy, ihx order == t order */
    ctile[t] = nsum; ctile_p[t] = nsum_p;
    for (int64_t k = lb; k <= nt + ncb; k++)
      for (int64_t j = lb; j <= nt + ncb; j++)
        for (int64_t i = lb; i <= nt + ncb; i++) {
          nsum += RHOC(i, j, k, t);
          if (i >= 1 && i <= nt && j >= 1 && j <= nt && k >= 1 && k <= nt) nsum_p += RHOC(i, j, k, t);
        }
  }
  for (int64_t t = 0; t < ntiles; t++) {
    nsum = ctile[t];
    for (int64_t k = lb; k <= nt + ncb; k++)
      for (int64_t j = lb; j <= nt + ncb; j++) {
        c->idx_b_l[brow(c, j, k, t)] = nsum;
        for (int64_t i = lb; i <= nt + ncb; i++) nsum += RHOC(i, j, k, t);
        c->idx_b_r[brow(c, j, k, t)] = nsum;
      }
  }
  for (int64_t t = 0; t < ntiles; t++) {
    nsum_p = ctile_p[t];
    for (int64_t k = 1; k <= nt; k++)
      for (int64_t j = 1; j <= nt; j++) {
        const int64_t r = prow(c, j, k, t);
        int64_t np_phy = 0, sr = 0, srl = 0, sl = 0, slr = 0;
        for (int64_t i = 1; i <= nt; i++) np_phy += RHOC(i, j, k, t);
        for (int64_t i = nt + 1; i <= nt + ncb; i++) sr += RHOC(i, j, k, t);
        for (int64_t i = nt - ncb + 1; i <= nt; i++) srl += RHOC(i, j, k, t);
        for (int64_t i = lb; i <= 0; i++) sl += RHOC(i, j, k, t);
        for (int64_t i = 1; i <= ncb; i++) slr += RHOC(i, j, k, t);
        c->ppl[r] = nsum_p;
        nsum_p += np_phy;
        c->ppr[r] = nsum_p;
        c->ppr0[r] = c->idx_b_r[brow(c, j, k, t)] - sr;
        c->pprl[r] = c->ppr0[r] - srl;
        c->ppl0[r] = c->idx_b_l[brow(c, j, k, t)] + sl;
        c->pplr[r] = c->ppl0[r] + slr;
      }
  }
  free(ctile); free(ctile_p);
}

/* ---- disjoint state transfer (W/particle_initialization.f90:34-82) ------- */
int oc_set_disjoint(ocube *c, int64_t nplocal, const int16_t *xp, const int16_t *vp,
                    const int32_t *pid, const int32_t *rhoc_phys, const float *vfield_phys) {
  if (nplocal > c->np_image_max) return OC_ERR_IMAGE_OVERFLOW;
  const int64_t nt = c->nt, ntiles = c->nnt * c->nnt * c->nnt;
  memset(c->xp, 0, 3 * c->np_image_max * sizeof(int16_t));
  memset(c->vp, 0, 3 * c->np_image_max * sizeof(int16_t));
  memcpy(c->xp, xp, 3 * nplocal * sizeof(int16_t));
  memcpy(c->vp, vp, 3 * nplocal * sizeof(int16_t));
  if (c->has_pid) {
    memset(c->pid, 0, c->np_image_max * sizeof(int32_t));
    if (pid) memcpy(c->pid, pid, nplocal * sizeof(int32_t));
  }
  memset(c->rhoc, 0, c->nte * c->nte * c->nte * ntiles * sizeof(int32_t));
  memset(c->vfield, 0, 3 * c->nte * c->nte * c->nte * ntiles * sizeof(float));
  int64_t n = 0;
  for (int64_t t = 0; t < ntiles; t++)
    for (int64_t k = 1; k <= nt; k++)
      for (int64_t j = 1; j <= nt; j++)
        for (int64_t i = 1; i <= nt; i++, n++) {
          RHOC(i, j, k, t) = rhoc_phys[n];
          for (int d = 0; d < 3; d++) VF(d, i, j, k, t) = vfield_phys[3 * n + d];
        }
  c->nplocal = nplocal;
  oc_spine_image(c);
  return OC_OK;
}

/* Gather the physical particles in disjoint order.  Works on both the
   disjoint layout (after update_xp) and the buffered one (after buffer_grid):
   it walks physical rows with a freshly computed spine. */
int oc_get_disjoint(const ocube *c_in, int16_t *xp, int16_t *vp, int32_t *pid,
                    int32_t *rhoc_phys, float *vfield_phys) {
  ocube *c = (ocube *)c_in;
  const int64_t nt = c->nt, ntiles = c->nnt * c->nnt * c->nnt;
  int64_t n = 0, m = 0;
  for (int64_t t = 0; t < ntiles; t++)
    for (int64_t k = 1; k <= nt; k++)
      for (int64_t j = 1; j <= nt; j++) {
        const int64_t r = brow(c, j, k, t);
        for (int64_t i = 1; i <= nt; i++, n++) {
          const int64_t np = RHOC(i, j, k, t);
          const int64_t nzero = c->idx_b_r[r] - row_suffix(c, i, j, k, t);
          if (rhoc_phys) rhoc_phys[n] = (int32_t)np;
          if (vfield_phys) for (int d = 0; d < 3; d++) vfield_phys[3 * n + d] = VF(d, i, j, k, t);
          for (int64_t l = 1; l <= np; l++, m++) {
            const int64_t ip = nzero + l;
            if (xp) for (int d = 0; d < 3; d++) xp[3 * m + d] = XP(d, ip);
            if (vp) for (int d = 0; d < 3; d++) vp[3 * m + d] = VP(d, ip);
            if (pid && c->has_pid) pid[m] = c->pid[ip - 1];
          }
        }
      }
  return OC_OK;
}

/* ---- buffer_np / buffer_vc, W/buffer_grid.f90:28-122 --------------------- */
/* For nn=1 every coindexed reference is this image, so each face copy takes
   the periodic neighbour tile (wrapping over nnt).  Passes run x, y, z, each
   over the full extent of the other two axes, so corners are transitive. */
static void buffer_faces(ocube *c, int is_v) {
  const int64_t nt = c->nt, ncb = c->ncb, nnt = c->nnt, lb = LB(c);
  for (int axis = 0; axis < 3; axis++) {
    for (int64_t tz = 1; tz <= nnt; tz++)
      for (int64_t ty = 1; ty <= nnt; ty++)
        for (int64_t tx = 1; tx <= nnt; tx++) {
          const int64_t t = tile_id(c, tx, ty, tz);
          int64_t tm[3] = {tx, ty, tz}, tp[3] = {tx, ty, tz};
          tm[axis] = (tm[axis] - 2 + nnt) % nnt + 1;   /* left neighbour */
          tp[axis] = tp[axis] % nnt + 1;               /* right neighbour */
          const int64_t t_left = tile_id(c, tm[0], tm[1], tm[2]);
          const int64_t t_right = tile_id(c, tp[0], tp[1], tp[2]);
          for (int64_t a = lb; a <= nt + ncb; a++)
            for (int64_t b = lb; b <= nt + ncb; b++)
              for (int64_t s = 0; s < ncb; s++) {
                /* left buffer cell index lb+s <- neighbour nt-ncb+1+s;
                   right buffer nt+1+s <- neighbour 1+s */
                int64_t dl[3], sl[3], dr[3], sr[3];
                int64_t o0 = a, o1 = b;
                int64_t *v[4] = {dl, sl, dr, sr};
                for (int q = 0; q < 4; q++) {
                  int64_t oth = 0;
                  for (int ax = 0; ax < 3; ax++) {
                    if (ax == axis) continue;
                    v[q][ax] = (oth++ == 0) ? o0 : o1;
                  }
                }
                dl[axis] = lb + s; sl[axis] = nt - ncb + 1 + s;
                dr[axis] = nt + 1 + s; sr[axis] = 1 + s;
                if (!is_v) {
                  RHOC(dl[0], dl[1], dl[2], t) = RHOC(sl[0], sl[1], sl[2], t_left);
                  RHOC(dr[0], dr[1], dr[2], t) = RHOC(sr[0], sr[1], sr[2], t_right);
                } else {
                  for (int d = 0; d < 3; d++) {
                    VF(d, dl[0], dl[1], dl[2], t) = VF(d, sl[0], sl[1], sl[2], t_left);
                    VF(d, dr[0], dr[1], dr[2], t) = VF(d, sr[0], sr[1], sr[2], t_right);
                  }
                }
              }
        }
  }
}

/* ---- redistribute_cdm, W/buffer_grid.f90:124-209 ------------------------- */
static int redistribute_cdm(ocube *c) {
  const int64_t nt = c->nt, ntiles = c->nnt * c->nnt * c->nnt;
  const int64_t ncells = c->nte * c->nte * c->nte * ntiles;
  int64_t tot = 0;
  for (int64_t q = 0; q < ncells; q++) tot += c->rhoc[q];
  c->overhead_image = (double)tot / (double)c->np_image_max;
  if (c->overhead_image > 1.0) return OC_ERR_IMAGE_OVERFLOW;
  int64_t check0 = 0, check1 = 0;
  for (int64_t q = 0; q < 3 * c->np_image_max; q++) check0 += c->xp[q];
  const int64_t nshift = c->np_image_max - c->nplocal;
  memmove(c->xp + 3 * nshift, c->xp, 3 * c->nplocal * sizeof(int16_t));
  memset(c->xp, 0, 3 * nshift * sizeof(int16_t));
  memmove(c->vp + 3 * nshift, c->vp, 3 * c->nplocal * sizeof(int16_t));
  memset(c->vp, 0, 3 * nshift * sizeof(int16_t));
  if (c->has_pid) {
    memmove(c->pid + nshift, c->pid, c->nplocal * sizeof(int32_t));
    memset(c->pid, 0, nshift * sizeof(int32_t));
  }
  for (int64_t q = 0; q < 3 * c->np_image_max; q++) check1 += c->xp[q];
  if (check0 != check1) return OC_ERR_CHECKSUM;
  oc_spine_image(c);
  for (int64_t t = 0; t < ntiles; t++)
    for (int64_t iz = 1; iz <= nt; iz++)
      for (int64_t iy = 1; iy <= nt; iy++) {
        const int64_t r = prow(c, iy, iz, t);
        const int64_t dst = c->ppl0[r], src = nshift + c->ppl[r], len = c->ppr0[r] - c->ppl0[r];
        memmove(c->xp + 3 * dst, c->xp + 3 * src, 3 * len * sizeof(int16_t));
        memmove(c->vp + 3 * dst, c->vp + 3 * src, 3 * len * sizeof(int16_t));
        if (c->has_pid) memmove(c->pid + dst, c->pid + src, len * sizeof(int32_t));
      }
  return OC_OK;
}

int oc_buffer_grid(ocube *c) {
  buffer_faces(c, 0);
  buffer_faces(c, 1);
  return redistribute_cdm(c);
}

/* ---- buffer_xvp, W/buffer_particle.f90:27-357 ---------------------------- */
static void copy_range(ocube *c, int dox, int dov, int64_t dst0, int64_t dst1, int64_t src0) {
  /* x(:, dst0+1:dst1) = x(:, src0+1:src0+dst1-dst0) */
  const int64_t len = dst1 - dst0;
  if (len <= 0) return;
  if (dox) memmove(c->xp + 3 * dst0, c->xp + 3 * src0, 3 * len * sizeof(int16_t));
  if (dov) {
    memmove(c->vp + 3 * dst0, c->vp + 3 * src0, 3 * len * sizeof(int16_t));
    if (c->has_pid) memmove(c->pid + dst0, c->pid + src0, len * sizeof(int32_t));
  }
}

void oc_buffer_xvp(ocube *c, int dox, int dov) {
  const int64_t nt = c->nt, ncb = c->ncb, nnt = c->nnt, lb = LB(c);
  /* x- : every tile's left buffer from its left neighbour's last ncb cells */
  for (int64_t tz = 1; tz <= nnt; tz++)
    for (int64_t ty = 1; ty <= nnt; ty++)
      for (int64_t tx = 1; tx <= nnt; tx++) {
        const int64_t t = tile_id(c, tx, ty, tz);
        const int64_t tl = tile_id(c, (tx - 2 + nnt) % nnt + 1, ty, tz);
        for (int64_t iz = 1; iz <= nt; iz++)
          for (int64_t iy = 1; iy <= nt; iy++)
            copy_range(c, dox, dov, c->idx_b_l[brow(c, iy, iz, t)], c->ppl0[prow(c, iy, iz, t)],
                       c->pprl[prow(c, iy, iz, tl)]);
      }
  /* x+ */
  for (int64_t tz = 1; tz <= nnt; tz++)
    for (int64_t ty = 1; ty <= nnt; ty++)
      for (int64_t tx = 1; tx <= nnt; tx++) {
        const int64_t t = tile_id(c, tx, ty, tz);
        const int64_t tr = tile_id(c, tx % nnt + 1, ty, tz);
        for (int64_t iz = 1; iz <= nt; iz++)
          for (int64_t iy = 1; iy <= nt; iy++)
            copy_range(c, dox, dov, c->ppr0[prow(c, iy, iz, t)], c->idx_b_r[brow(c, iy, iz, t)],
                       c->ppl0[prow(c, iy, iz, tr)]);
      }
  /* y- : rows 1-ncb..0 of plane iz from rows nt-ncb+1..nt of the y-neighbour */
  for (int64_t tz = 1; tz <= nnt; tz++)
    for (int64_t ty = 1; ty <= nnt; ty++)
      for (int64_t tx = 1; tx <= nnt; tx++) {
        const int64_t t = tile_id(c, tx, ty, tz);
        const int64_t tl = tile_id(c, tx, (ty - 2 + nnt) % nnt + 1, tz);
        for (int64_t iz = 1; iz <= nt; iz++)
          copy_range(c, dox, dov, c->idx_b_l[brow(c, lb, iz, t)], c->idx_b_r[brow(c, 0, iz, t)],
                     c->idx_b_l[brow(c, nt - ncb + 1, iz, tl)]);
      }
  /* y+ */
  for (int64_t tz = 1; tz <= nnt; tz++)
    for (int64_t ty = 1; ty <= nnt; ty++)
      for (int64_t tx = 1; tx <= nnt; tx++) {
        const int64_t t = tile_id(c, tx, ty, tz);
        const int64_t tr = tile_id(c, tx, ty % nnt + 1, tz);
        for (int64_t iz = 1; iz <= nt; iz++)
          copy_range(c, dox, dov, c->idx_b_l[brow(c, nt + 1, iz, t)],
                     c->idx_b_r[brow(c, nt + ncb, iz, t)], c->idx_b_l[brow(c, 1, iz, tr)]);
      }
  /* z- : planes 1-ncb..0 from planes nt-ncb+1..nt of the z-neighbour */
  for (int64_t tz = 1; tz <= nnt; tz++)
    for (int64_t ty = 1; ty <= nnt; ty++)
      for (int64_t tx = 1; tx <= nnt; tx++) {
        const int64_t t = tile_id(c, tx, ty, tz);
        const int64_t tl = tile_id(c, tx, ty, (tz - 2 + nnt) % nnt + 1);
        copy_range(c, dox, dov, c->idx_b_l[brow(c, lb, lb, t)], c->idx_b_r[brow(c, nt + ncb, 0, t)],
                   c->idx_b_l[brow(c, lb, nt - ncb + 1, tl)]);
      }
  /* z+ */
  for (int64_t tz = 1; tz <= nnt; tz++)
    for (int64_t ty = 1; ty <= nnt; ty++)
      for (int64_t tx = 1; tx <= nnt; tx++) {
        const int64_t t = tile_id(c, tx, ty, tz);
        const int64_t tr = tile_id(c, tx, ty, tz % nnt + 1);
        copy_range(c, dox, dov, c->idx_b_l[brow(c, lb, nt + 1, t)],
                   c->idx_b_r[brow(c, nt + ncb, nt + ncb, t)], c->idx_b_l[brow(c, lb, 1, tr)]);
      }
}

/* ---- update_xp, W/update_particle.f90:16-256 ----------------------------- */
/* 1 when no particle of tile t moves more than (nlayer-1)/2 cells in z
   during the drift, the premise of the reference's layered parallel loops */
static int drift_layers_disjoint(ocube *c, int64_t t, float dt_mid, double vc, int64_t nlayer) {
  const int64_t nt = c->nt, ncb = c->ncb, lb = LB(c), ncell = c->ncell, h = (nlayer - 1) / 2;
  int ok = 1;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : ok)
  for (int64_t k = lb; k <= nt + ncb; k++)
    for (int64_t j = lb; j <= nt + ncb; j++)
      for (int64_t i = lb; i <= nt + ncb; i++) {
        const int64_t np = RHOC(i, j, k, t);
        const int64_t nzero = c->idx_b_r[brow(c, j, k, t)] - row_suffix(c, i, j, k, t);
        for (int64_t l = 1; l <= np; l++) {
          const int64_t ip = nzero + l;
          const double xq = ((double)k - 1.0) + oc_decode_x(c->izipx, XP(2, ip));
          double vr = (double)oc_tanf_code(c->izipv, VP(2, ip)) / vc;
          vr = vr + (double)VF(2, i, j, k, t);
          const int64_t ig = (int64_t)ceil(xq + ((double)dt_mid * vr) / (double)ncell);
          if (ig - k > h || k - ig > h) ok = 0;
        }
      }
  return ok;
}

int oc_update_xp(ocube *c, float dt_old, float dt, float vz_max, float sigma_vi,
                 double *stats, float *sigma_vi_new) {
  const int64_t nt = c->nt, ncb = c->ncb, nnt = c->nnt, lb = LB(c), ncell = c->ncell;
  const int64_t ne = nt + 4 * ncb, lb2 = 1 - 2 * ncb;   /* rhoce(1-2ncb:nt+2ncb)^3 */
  const float dt_mid = (dt_old + dt) / 2.0f;
  const int64_t nlayer = 2 * (int64_t)ceilf(dt_mid * vz_max / (float)ncell) + 1;
  const double xres = x_res(c->izipx);
  const double vc = vconst(sigma_vi);
  int status = OC_OK;
  int32_t *rhoce = calloc(ne * ne * ne, sizeof(int32_t));
  int32_t *rholocal = calloc(ne * ne * ne, sizeof(int32_t));
  float *vfn = calloc(3 * ne * ne * ne, sizeof(float));
  int64_t *idx_ex_r = calloc(ne * ne, sizeof(int64_t));
  int64_t *pp_l = calloc(nt * nt, sizeof(int64_t)), *pp_r = calloc(nt * nt, sizeof(int64_t));
  int64_t *ppe_l = calloc(nt * nt, sizeof(int64_t)), *ppe_r = calloc(nt * nt, sizeof(int64_t));
  int16_t *xp_new = calloc(3 * c->np_tile_max, sizeof(int16_t));
  int16_t *vp_new = calloc(3 * c->np_tile_max, sizeof(int16_t));
  int32_t *pid_new = c->has_pid ? calloc(c->np_tile_max, sizeof(int32_t)) : NULL;
#define E(i, j, k) (((i) - lb2) + ne * (((j) - lb2) + ne * ((k) - lb2)))
  int64_t np_prev = 0;
  c->overhead_tile = 0;
  c->lost = 0;
  for (int64_t tz = 1; tz <= nnt && status == OC_OK; tz++)
    for (int64_t ty = 1; ty <= nnt && status == OC_OK; ty++)
      for (int64_t tx = 1; tx <= nnt && status == OC_OK; tx++) {
        const int64_t t = tile_id(c, tx, ty, tz);
        memset(rhoce, 0, ne * ne * ne * sizeof(int32_t));
        memset(rholocal, 0, ne * ne * ne * sizeof(int32_t));
        memset(vfn, 0, 3 * ne * ne * ne * sizeof(float));
        for (int64_t k = lb; k <= nt + ncb; k++)
          for (int64_t j = lb; j <= nt + ncb; j++)
            for (int64_t i = lb; i <= nt + ncb; i++)
              for (int d = 0; d < 3; d++)
                vfn[3 * E(i, j, k) + d] = (float)((double)VF(d, i, j, k, t) * WEIGHT_V);
        memset(xp_new, 0, 3 * c->np_tile_max * sizeof(int16_t));
        memset(vp_new, 0, 3 * c->np_tile_max * sizeof(int16_t));
        if (pid_new) memset(pid_new, 0, c->np_tile_max * sizeof(int32_t));
        /* The reference runs each layer's k slabs in parallel (OpenMP,
           W/update_particle.f90:55-58): slabs nlayer apart reach disjoint
           destination planes when no particle moves more than
           (nlayer-1)/2 cells in z, so every destination cell still sums in
           (ilayer,k,j,i,l) order.  The oracle checks that premise per tile
           and runs the slabs serially where it fails. */
        const int par = drift_layers_disjoint(c, t, dt_mid, vc, nlayer);
        /* pass 1: count and vfield, layered (ilayer,k,j,i,l) order (:55-78) */
        for (int64_t ilayer = 0; ilayer < nlayer && status == OC_OK; ilayer++) {
          int st1 = OC_OK;
#pragma omp parallel for schedule(dynamic, 1) reduction(max : st1) if (par)
          for (int64_t k = lb + ilayer; k <= nt + ncb; k += nlayer)
            for (int64_t j = lb; j <= nt + ncb; j++)
              for (int64_t i = lb; i <= nt + ncb; i++) {
                const int64_t np = RHOC(i, j, k, t);
                const int64_t nzero = c->idx_b_r[brow(c, j, k, t)] - row_suffix(c, i, j, k, t);
                const int64_t ijk[3] = {i, j, k};
                for (int64_t l = 1; l <= np; l++) {
                  const int64_t ip = nzero + l;
                  int64_t ig[3];
                  double vreal[3];
                  int bad = 0;
                  for (int d = 0; d < 3; d++) {
                    const double xq = ((double)ijk[d] - 1.0) + oc_decode_x(c->izipx, XP(d, ip));
                    vreal[d] = (double)oc_tanf_code(c->izipv, VP(d, ip)) / vc;
                    vreal[d] = vreal[d] + (double)VF(d, i, j, k, t);
                    const double deltax = ((double)dt_mid * vreal[d]) / (double)ncell;
                    ig[d] = (int64_t)ceil(xq + deltax);
                    if (ig[d] < lb2 || ig[d] > nt + 2 * ncb) bad = 1;
                  }
                  if (bad) {
                    st1 = OC_ERR_DISPLACEMENT;
                    continue;
                  }
                  const int64_t e = E(ig[0], ig[1], ig[2]);
                  rhoce[e] += 1;
                  for (int d = 0; d < 3; d++)
                    vfn[3 * e + d] = (float)((double)vfn[3 * e + d] + vreal[d]);
                }
              }
          if (st1 != OC_OK) status = st1;
        }
        if (status != OC_OK) break;
        /* vfield_new/(rhoce+weight_v) (:81-83) */
        for (int64_t q = 0; q < ne * ne * ne; q++)
          for (int d = 0; d < 3; d++)
            vfn[3 * q + d] = (float)((double)vfn[3 * q + d] / ((double)rhoce[q] + WEIGHT_V));
        /* spine_tile (W/variables.f90:96-139) */
        {
          int64_t nsum = 0;
          for (int64_t igz = lb2; igz <= nt + 2 * ncb; igz++)
            for (int64_t igy = lb2; igy <= nt + 2 * ncb; igy++) {
              for (int64_t igx = lb2; igx <= nt + 2 * ncb; igx++) nsum += rhoce[E(igx, igy, igz)];
              idx_ex_r[(igy - lb2) + ne * (igz - lb2)] = nsum;
            }
          for (int64_t igz = 1; igz <= nt; igz++)
            for (int64_t igy = 1; igy <= nt; igy++) {
              int64_t s = 0;
              for (int64_t igx = nt + 1; igx <= nt + 2 * ncb; igx++) s += rhoce[E(igx, igy, igz)];
              ppe_r[(igy - 1) + nt * (igz - 1)] = idx_ex_r[(igy - lb2) + ne * (igz - lb2)] - s;
            }
          nsum = 0;
          for (int64_t igz = 1; igz <= nt; igz++)
            for (int64_t igy = 1; igy <= nt; igy++) {
              const int64_t r = (igy - 1) + nt * (igz - 1);
              int64_t np_phy = 0;
              for (int64_t igx = 1; igx <= nt; igx++) np_phy += rhoce[E(igx, igy, igz)];
              pp_l[r] = nsum + 1;
              nsum += np_phy;
              pp_r[r] = nsum;
              ppe_l[r] = ppe_r[r] - np_phy + 1;
            }
        }
        const int64_t ntile_tot = idx_ex_r[(nt + 2 * ncb - lb2) + ne * (nt + 2 * ncb - lb2)];
        {
          const double ov = (double)((float)ntile_tot / (float)c->np_tile_max);
          if (ov > c->overhead_tile) c->overhead_tile = ov;
        }
        if (ntile_tot > c->np_tile_max) { status = OC_ERR_TILE_OVERFLOW; break; }
        /* pass 2: scatter with re-quantisation (:97-126), layered as pass 1 */
        for (int64_t ilayer = 0; ilayer < nlayer; ilayer++)
#pragma omp parallel for schedule(dynamic, 1) if (par)
          for (int64_t k = lb + ilayer; k <= nt + ncb; k += nlayer)
            for (int64_t j = lb; j <= nt + ncb; j++)
              for (int64_t i = lb; i <= nt + ncb; i++) {
                const int64_t np = RHOC(i, j, k, t);
                const int64_t nzero = c->idx_b_r[brow(c, j, k, t)] - row_suffix(c, i, j, k, t);
                const int64_t ijk[3] = {i, j, k};
                for (int64_t l = 1; l <= np; l++) {
                  const int64_t ip = nzero + l;
                  int64_t ig[3];
                  double vreal[3];
                  for (int d = 0; d < 3; d++) {
                    const double xq = ((double)ijk[d] - 1.0) + oc_decode_x(c->izipx, XP(d, ip));
                    vreal[d] = (double)oc_tanf_code(c->izipv, VP(d, ip)) / vc;
                    vreal[d] = vreal[d] + (double)VF(d, i, j, k, t);
                    const double deltax = ((double)dt_mid * vreal[d]) / (double)ncell;
                    ig[d] = (int64_t)ceil(xq + deltax);
                  }
                  const int64_t e = E(ig[0], ig[1], ig[2]);
                  rholocal[e] += 1;
                  int64_t suf = 0;
                  for (int64_t ii = ig[0]; ii <= nt + 2 * ncb; ii++) suf += rhoce[E(ii, ig[1], ig[2])];
                  const int64_t idx = idx_ex_r[(ig[1] - lb2) + ne * (ig[2] - lb2)] - suf + rholocal[e];
                  for (int d = 0; d < 3; d++) {
                    const double q = ((double)dt_mid * vreal[d]) / (xres * (double)ncell);
                    xp_new[3 * (idx - 1) + d] = (int16_t)wrap_kind(c->izipx, (int64_t)XP(d, ip) + f_nint(q));
                    const double vr = vreal[d] - (double)vfn[3 * e + d];
                    vp_new[3 * (idx - 1) + d] = (int16_t)oc_encode_v(c->izipv, vr, sigma_vi);
                  }
                  if (pid_new) pid_new[idx - 1] = c->pid[ip - 1];
                }
              }
        /* delete buffer particles: compact the physical rows (:130-142) */
        for (int64_t k = 1; k <= nt; k++)
          for (int64_t j = 1; j <= nt; j++) {
            const int64_t r = (j - 1) + nt * (k - 1);
            const int64_t len = pp_r[r] - pp_l[r] + 1;
            if (len <= 0) continue;
            memmove(c->xp + 3 * (np_prev + pp_l[r] - 1), xp_new + 3 * (ppe_l[r] - 1), 3 * len * sizeof(int16_t));
            memmove(c->vp + 3 * (np_prev + pp_l[r] - 1), vp_new + 3 * (ppe_l[r] - 1), 3 * len * sizeof(int16_t));
            if (pid_new) memmove(c->pid + (np_prev + pp_l[r] - 1), pid_new + (ppe_l[r] - 1), len * sizeof(int32_t));
          }
        np_prev += pp_r[(nt - 1) + nt * (nt - 1)];
        /* update rhoc and vfield of this tile (:144-147) */
        for (int64_t k = lb; k <= nt + ncb; k++)
          for (int64_t j = lb; j <= nt + ncb; j++)
            for (int64_t i = lb; i <= nt + ncb; i++) {
              const int phys = i >= 1 && i <= nt && j >= 1 && j <= nt && k >= 1 && k <= nt;
              RHOC(i, j, k, t) = phys ? rhoce[E(i, j, k)] : 0;
              for (int d = 0; d < 3; d++) VF(d, i, j, k, t) = phys ? vfn[3 * E(i, j, k) + d] : 0.0f;
            }
      }
#undef E
  free(rhoce); free(rholocal); free(vfn); free(idx_ex_r); free(pp_l); free(pp_r);
  free(ppe_l); free(ppe_r); free(xp_new); free(vp_new); free(pid_new);
  if (status != OC_OK) return status;
  c->nplocal = np_prev;
  memset(c->xp + 3 * c->nplocal, 0, 3 * (c->np_image_max - c->nplocal) * sizeof(int16_t));
  memset(c->vp + 3 * c->nplocal, 0, 3 * (c->np_image_max - c->nplocal) * sizeof(int16_t));
  if (c->has_pid) memset(c->pid + c->nplocal, 0, (c->np_image_max - c->nplocal) * sizeof(int32_t));
  /* velocity statistics (:171-224), summed serially in loop order */
  oc_spine_image(c);
  double sum_c = 0, sum_r = 0, sum_s = 0;
  const int64_t ntiles = nnt * nnt * nnt;
  for (int64_t t = 0; t < ntiles; t++)
    for (int64_t k = 1; k <= nt; k++)
      for (int64_t j = 1; j <= nt; j++)
        for (int64_t i = 1; i <= nt; i++) {
          const float v0 = VF(0, i, j, k, t), v1 = VF(1, i, j, k, t), v2 = VF(2, i, j, k, t);
          sum_c = sum_c + (double)((v0 * v0 + v1 * v1) + v2 * v2);
          const int64_t np = RHOC(i, j, k, t);
          const int64_t nzero = c->idx_b_r[brow(c, j, k, t)] - row_suffix(c, i, j, k, t);
          for (int64_t l = 1; l <= np; l++) {
            const int64_t ip = nzero + l;
            double vr[3];
            for (int d = 0; d < 3; d++) vr[d] = (double)oc_tanf_code(c->izipv, VP(d, ip)) / vc;
            sum_r = sum_r + ((vr[0] * vr[0] + vr[1] * vr[1]) + vr[2] * vr[2]);
            for (int d = 0; d < 3; d++) vr[d] = vr[d] + (double)VF(d, i, j, k, t);
            sum_s = sum_s + ((vr[0] * vr[0] + vr[1] * vr[1]) + vr[2] * vr[2]);
          }
        }
  const double npg = (double)c->npglobal;
  const double std_vsim = sqrt(sum_s / npg);
  const double std_vsim_c = sqrt(sum_c / (double)c->nc / (double)c->nc / (double)c->nc);
  const double std_vsim_res = sqrt(sum_r / npg);
  if (stats) { stats[0] = sum_c; stats[1] = sum_r; stats[2] = sum_s;
               stats[3] = std_vsim; stats[4] = std_vsim_c; stats[5] = std_vsim_res; }
  if (sigma_vi_new) *sigma_vi_new = (float)(std_vsim_res / (double)sqrtf(3.0f));
  if (c->nplocal != c->npglobal) return OC_ERR_COUNT;
  return OC_OK;
}

/* ---- fine CIC deposit of one tile, W/pm.f90:49-108 ----------------------- */
double oc_fine_deposit(const ocube *c_in, int64_t itx, int64_t ity, int64_t itz,
                       float mass_p, float *rho_f) {
  ocube *c = (ocube *)c_in;
  const int64_t nt = c->nt, ncb = c->ncb, ncell = c->ncell, nfb = c->nfb, nfe = c->nfe;
  const int64_t t = tile_id(c, itx, ity, itz);
  const int64_t sx = nfe + 2;
#define RF(x, y, z) rho_f[((x) - 1) + sx * (((y) - 1) + nfe * ((z) - 1))]
  memset(rho_f, 0, sx * nfe * nfe * sizeof(float));
  /* layers of k slabs 3 apart touch disjoint planes of rho_f, so the slabs
     of a layer run in parallel as in the reference (W/pm.f90:52-55) and
     every cell sums in (ilayer,k,j,i,l) order */
  for (int64_t ilayer = 0; ilayer < 3; ilayer++)
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t k = 2 - ncb + ilayer; k <= nt + ncb - 1; k += 3)
      for (int64_t j = 2 - ncb; j <= nt + ncb - 1; j++)
        for (int64_t i = 2 - ncb; i <= nt + ncb - 1; i++) {
          const int64_t np = RHOC(i, j, k, t);
          const int64_t nzero = c->idx_b_r[brow(c, j, k, t)] - row_suffix(c, i, j, k, t);
          const int64_t ijk[3] = {i, j, k};
          for (int64_t l = 1; l <= np; l++) {
            const int64_t ip = nzero + l;
            int64_t i1[3], i2[3];
            float d1[3], d2[3];
            for (int d = 0; d < 3; d++) {
              const float tempx = (float)((double)(ncell * (ijk[d] - 1)) +
                                          (double)ncell * oc_decode_x(c->izipx, XP(d, ip)));
              i1[d] = (int64_t)floorf(tempx) + 1;
              i2[d] = i1[d] + 1;
              d1[d] = (float)i1[d] - tempx;
              d2[d] = 1.0f - d1[d];
              i1[d] += nfb; i2[d] += nfb;
            }
            RF(i1[0], i1[1], i1[2]) += d1[0] * d1[1] * d1[2] * mass_p;
            RF(i2[0], i1[1], i1[2]) += d2[0] * d1[1] * d1[2] * mass_p;
            RF(i1[0], i2[1], i1[2]) += d1[0] * d2[1] * d1[2] * mass_p;
            RF(i1[0], i1[1], i2[2]) += d1[0] * d1[1] * d2[2] * mass_p;
            RF(i1[0], i2[1], i2[2]) += d1[0] * d2[1] * d2[2] * mass_p;
            RF(i2[0], i1[1], i2[2]) += d2[0] * d1[1] * d2[2] * mass_p;
            RF(i2[0], i2[1], i1[2]) += d2[0] * d2[1] * d1[2] * mass_p;
            RF(i2[0], i2[1], i2[2]) += d2[0] * d2[1] * d2[2] * mass_p;
          }
        }
  double testrhof = 0;
  for (int64_t z = nfb + 1; z <= c->nft + nfb; z++)
    for (int64_t y = nfb + 1; y <= c->nft + nfb; y++)
      for (int64_t x = nfb + 1; x <= c->nft + nfb; x++) testrhof += (double)RF(x, y, z);
#undef RF
  return testrhof;
}

/* ---- fine kick of one tile, W/pm.f90:127-159 ----------------------------- */
void oc_fine_kick(ocube *c, int64_t itx, int64_t ity, int64_t itz, const float *force_f,
                  float a_mid, float dt, float sigma_vi, float sigma_vi_new) {
  const int64_t nt = c->nt, ncell = c->ncell, nfb = c->nfb, nf2 = c->nft + 2;
  const int64_t t = tile_id(c, itx, ity, itz);
  const double vc = vconst(sigma_vi);
#define FF(d, x, y, z) force_f[(d) + 3 * (((x) - nfb) + nf2 * (((y) - nfb) + nf2 * ((z) - nfb)))]
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t k = 1; k <= nt; k++)
    for (int64_t j = 1; j <= nt; j++)
      for (int64_t i = 1; i <= nt; i++) {
        const int64_t np = RHOC(i, j, k, t);
        const int64_t nzero = c->idx_b_r[brow(c, j, k, t)] - row_suffix(c, i, j, k, t);
        const int64_t ijk[3] = {i, j, k};
        for (int64_t l = 1; l <= np; l++) {
          const int64_t ip = nzero + l;
          int64_t i1[3], i2[3];
          float d1[3], d2[3];
          for (int d = 0; d < 3; d++) {
            const float tempx = (float)((double)(ncell * (ijk[d] - 1)) +
                                        (double)ncell * oc_decode_x(c->izipx, XP(d, ip)));
            i1[d] = (int64_t)floorf(tempx) + 1;
            i2[d] = i1[d] + 1;
            d1[d] = (float)i1[d] - tempx;
            d2[d] = 1.0f - d1[d];
            i1[d] += nfb; i2[d] += nfb;
          }
          double vreal[3];
          for (int d = 0; d < 3; d++) vreal[d] = (double)oc_tanf_code(c->izipv, VP(d, ip)) / vc;
          const int64_t *cx[8] = {i1, i2, i1, i1, i1, i2, i2, i2};
          const int64_t *cy[8] = {i1, i1, i2, i1, i2, i1, i2, i2};
          const int64_t *cz[8] = {i1, i1, i1, i2, i2, i2, i1, i2};
          const float *wx[8] = {d1, d2, d1, d1, d1, d2, d2, d2};
          const float *wy[8] = {d1, d1, d2, d1, d2, d1, d2, d2};
          const float *wz[8] = {d1, d1, d1, d2, d2, d2, d1, d2};
          for (int q = 0; q < 8; q++)
            for (int d = 0; d < 3; d++) {
              const float f = FF(d, cx[q][0], cy[q][1], cz[q][2]);
              const float inc = ((((((f * a_mid) * dt) / 6.0f) / PI_F) * wx[q][0]) * wy[q][1]) * wz[q][2];
              vreal[d] = vreal[d] + (double)inc;
            }
          for (int d = 0; d < 3; d++) VP(d, ip) = (int16_t)oc_encode_v(c->izipv, vreal[d], sigma_vi_new);
        }
      }
#undef FF
}

/* ---- coarse CIC deposit, W/pm.f90:202-267 -------------------------------- */
double oc_coarse_deposit(const ocube *c_in, float mass_p, float *r3) {
  ocube *c = (ocube *)c_in;
  const int64_t nt = c->nt, nnt = c->nnt, nc = c->nc, m = nt + 4;  /* r3t(-1:nt+2) */
  float *r3t = malloc(m * m * m * sizeof(float));
#define RT(x, y, z) r3t[((x) + 1) + m * (((y) + 1) + m * ((z) + 1))]
  memset(r3, 0, nc * nc * nc * sizeof(float));
  for (int64_t tz = 1; tz <= nnt; tz++)
    for (int64_t ty = 1; ty <= nnt; ty++)
      for (int64_t tx = 1; tx <= nnt; tx++) {
        const int64_t t = tile_id(c, tx, ty, tz);
        memset(r3t, 0, m * m * m * sizeof(float));
        for (int64_t ilayer = 0; ilayer < 3; ilayer++)
#pragma omp parallel for schedule(dynamic, 1)
          for (int64_t k = ilayer; k <= nt + 1; k += 3)
            for (int64_t j = 0; j <= nt + 1; j++)
              for (int64_t i = 0; i <= nt + 1; i++) {
                const int64_t np = RHOC(i, j, k, t);
                const int64_t nzero = c->idx_b_r[brow(c, j, k, t)] - row_suffix(c, i, j, k, t);
                const int64_t ijk[3] = {i, j, k};
                for (int64_t l = 1; l <= np; l++) {
                  const int64_t ip = nzero + l;
                  int64_t i1[3], i2[3];
                  float d1[3], d2[3];
                  for (int d = 0; d < 3; d++) {
                    const float tempx = (float)(((double)(ijk[d] - 1) + oc_decode_x(c->izipx, XP(d, ip))) - 0.5);
                    i1[d] = (int64_t)floorf(tempx) + 1;
                    i2[d] = i1[d] + 1;
                    d1[d] = (float)i1[d] - tempx;
                    d2[d] = 1.0f - d1[d];
                  }
                  RT(i1[0], i1[1], i1[2]) += d1[0] * d1[1] * d1[2] * mass_p;
                  RT(i2[0], i1[1], i1[2]) += d2[0] * d1[1] * d1[2] * mass_p;
                  RT(i1[0], i2[1], i1[2]) += d1[0] * d2[1] * d1[2] * mass_p;
                  RT(i1[0], i1[1], i2[2]) += d1[0] * d1[1] * d2[2] * mass_p;
                  RT(i1[0], i2[1], i2[2]) += d1[0] * d2[1] * d2[2] * mass_p;
                  RT(i2[0], i1[1], i2[2]) += d2[0] * d1[1] * d2[2] * mass_p;
                  RT(i2[0], i2[1], i1[2]) += d2[0] * d2[1] * d1[2] * mass_p;
                  RT(i2[0], i2[1], i2[2]) += d2[0] * d2[1] * d2[2] * mass_p;
                }
              }
        for (int64_t z = 1; z <= nt; z++)
          for (int64_t y = 1; y <= nt; y++)
            for (int64_t x = 1; x <= nt; x++)
              r3[((tx - 1) * nt + x - 1) + nc * (((ty - 1) * nt + y - 1) + nc * ((tz - 1) * nt + z - 1))] = RT(x, y, z);
      }
#undef RT
  free(r3t);
  double s = 0;
  for (int64_t q = 0; q < nc * nc * nc; q++) s += (double)r3[q];
  return s;
}

/* ---- coarse kick, W/pm.f90:319-387 --------------------------------------- */
void oc_coarse_kick(ocube *c, const float *force_c, float a_mid, float dt,
                    float sigma_vi, float *vmax) {
  const int64_t nt = c->nt, nnt = c->nnt, nc = c->nc, m = nc + 2;
  const double vc = vconst(sigma_vi);
  double vm[3] = {0, 0, 0};
#define FC(d, x, y, z) force_c[(d) + 3 * ((x) + m * ((y) + m * (z)))]
  for (int64_t tz = 1; tz <= nnt; tz++)
    for (int64_t ty = 1; ty <= nnt; ty++)
      for (int64_t tx = 1; tx <= nnt; tx++) {
        const int64_t t = tile_id(c, tx, ty, tz);
        const int64_t txyz[3] = {tx, ty, tz};
#pragma omp parallel for schedule(dynamic, 1) reduction(max : vm[:3])
        for (int64_t k = 1; k <= nt; k++)
          for (int64_t j = 1; j <= nt; j++)
            for (int64_t i = 1; i <= nt; i++) {
              const int64_t np = RHOC(i, j, k, t);
              const int64_t nzero = c->idx_b_r[brow(c, j, k, t)] - row_suffix(c, i, j, k, t);
              const int64_t ijk[3] = {i, j, k};
              for (int64_t l = 1; l <= np; l++) {
                const int64_t ip = nzero + l;
                int64_t i1[3], i2[3];
                float d1[3], d2[3];
                for (int d = 0; d < 3; d++) {
                  const float tempx = (float)(((double)((txyz[d] - 1) * nt + (ijk[d] - 1)) +
                                               oc_decode_x(c->izipx, XP(d, ip))) - 0.5);
                  i1[d] = (int64_t)floorf(tempx) + 1;
                  i2[d] = i1[d] + 1;
                  d1[d] = (float)i1[d] - tempx;
                  d2[d] = 1.0f - d1[d];
                }
                double vreal[3];
                for (int d = 0; d < 3; d++) vreal[d] = (double)oc_tanf_code(c->izipv, VP(d, ip)) / vc;
                const int64_t *cx[8] = {i1, i2, i1, i1, i1, i2, i2, i2};
                const int64_t *cy[8] = {i1, i1, i2, i1, i2, i1, i2, i2};
                const int64_t *cz[8] = {i1, i1, i1, i2, i2, i2, i1, i2};
                const float *wx[8] = {d1, d2, d1, d1, d1, d2, d2, d2};
                const float *wy[8] = {d1, d1, d2, d1, d2, d1, d2, d2};
                const float *wz[8] = {d1, d1, d1, d2, d2, d2, d1, d2};
                for (int q = 0; q < 8; q++)
                  for (int d = 0; d < 3; d++) {
                    const float f = FC(d, cx[q][0], cy[q][1], cz[q][2]);
                    const float inc = ((((((f * a_mid) * dt) / 6.0f) / PI_F) * wx[q][0]) * wy[q][1]) * wz[q][2];
                    vreal[d] = vreal[d] + (double)inc;
                  }
                for (int d = 0; d < 3; d++) {
                  const double a = fabs(vreal[d] + (double)VF(d, i, j, k, t));
                  if (a > vm[d]) vm[d] = a;
                  VP(d, ip) = (int16_t)oc_encode_v(c->izipv, vreal[d], sigma_vi);
                }
              }
            }
      }
#undef FC
  for (int d = 0; d < 3; d++) vmax[d] = (float)vm[d];
}

/* ---- ext_pp_force, W/ext_pp_force.f90:10-162 ----------------------------- */
static inline double powi3(double x) { return x * (x * x); }
static inline double powi5(double x) { const double x2 = x * x; return x2 * (x * x2); }

int oc_pp_force(ocube *c, float a_mid, float dt, float sigma_vi, float mass_p,
                float *f2_max_pp) {
  const int64_t nt = c->nt, nnt = c->nnt, ncell = c->ncell, nft = c->nft, pr = c->pp_range;
  const int64_t lo = 1 - ncell, hn = nft + 2 * ncell;   /* hoc(1-ncell:nft+ncell)^3 */
  const double vc = vconst(sigma_vi);
  int32_t *hoc = malloc(hn * hn * hn * sizeof(int32_t));
  int32_t *ll = malloc(c->np_pp_max * sizeof(int32_t));
#define HOC(x, y, z) hoc[((x) - lo) + hn * (((y) - lo) + hn * ((z) - lo))]
  c->itest1 = 0; c->npairs = 0;
  int status = OC_OK;
  for (int64_t tz = 1; tz <= nnt && status == OC_OK; tz++)
    for (int64_t ty = 1; ty <= nnt && status == OC_OK; ty++)
      for (int64_t tx = 1; tx <= nnt && status == OC_OK; tx++) {
        const int64_t t = tile_id(c, tx, ty, tz);
        int64_t s = 0;
        for (int64_t k = 0; k <= nt + 1; k++)
          for (int64_t j = 0; j <= nt + 1; j++)
            for (int64_t i = 0; i <= nt + 1; i++) s += RHOC(i, j, k, t);
        if (c->np_pp_max < s) { status = OC_ERR_PP_OVERFLOW; break; }
        int64_t sl = 0;
        for (int64_t i = LB(c); i <= -1; i++) sl += RHOC(i, 0, 0, t);
        const int64_t ip_offset = c->idx_b_l[brow(c, 0, 0, t)] + sl;
        memset(hoc, 0, hn * hn * hn * sizeof(int32_t));
        memset(ll, 0, c->np_pp_max * sizeof(int32_t));
        for (int64_t igz = 0; igz <= nt + 1; igz++)
          for (int64_t igy = 0; igy <= nt + 1; igy++)
            for (int64_t igx = 0; igx <= nt + 1; igx++) {
              const int64_t np = RHOC(igx, igy, igz, t);
              const int64_t nzero = c->idx_b_r[brow(c, igy, igz, t)] - row_suffix(c, igx, igy, igz, t);
              const int64_t ig[3] = {igx, igy, igz};
              for (int64_t lp = 1; lp <= np; lp++) {
                const int64_t ip1 = nzero + lp;
                int64_t iv[3];
                for (int d = 0; d < 3; d++) {
                  const double xv = (double)(ncell * (ig[d] - 1)) +
                                    (double)ncell * oc_decode_x(c->izipx, XP(d, ip1));
                  iv[d] = (int64_t)floor(xv) + 1;
                }
                const int32_t ipll1 = (int32_t)(ip1 - ip_offset);
                ll[ipll1 - 1] = HOC(iv[0], iv[1], iv[2]);
                HOC(iv[0], iv[1], iv[2]) = ipll1;
              }
            }
        /* particles only read positions of others and write their own
           velocity, so the fine cells run in parallel; the running max of
           f2 in real(4) (:123) is order-free (rounding is monotone) */
        int64_t n1 = 0, npr = 0;
        double fmx = (double)(float)c->pp_ftemp;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : n1, npr) reduction(max : fmx)
        for (int64_t igz = 1; igz <= nft; igz++)
          for (int64_t igy = 1; igy <= nft; igy++)
            for (int64_t igx = 1; igx <= nft; igx++) {
              int32_t ipll1 = HOC(igx, igy, igz);
              const int64_t g1[3] = {igx, igy, igz};
              while (ipll1 != 0) {
                n1++;
                const int64_t ip1 = ipll1 + ip_offset;
                double x1[3], vreal[3], f_tot[3] = {0, 0, 0};
                for (int d = 0; d < 3; d++) {
                  x1[d] = (double)(ncell * (int64_t)floorf((float)(g1[d] - 1) / 4.0f)) +
                          (double)ncell * oc_decode_x(c->izipx, XP(d, ip1));
                  vreal[d] = (double)oc_tanf_code(c->izipv, VP(d, ip1)) / vc;
                }
                for (int64_t kk = igz - pr; kk <= igz + pr; kk++)
                  for (int64_t jj = igy - pr; jj <= igy + pr; jj++)
                    for (int64_t ii = igx - pr; ii <= igx + pr; ii++) {
                      int32_t ipll2 = HOC(ii, jj, kk);
                      const int64_t g2[3] = {ii, jj, kk};
                      while (ipll2 != 0) {
                        npr++;
                        const int64_t ip2 = ipll2 + ip_offset;
                        double x21[3];
                        for (int d = 0; d < 3; d++) {
                          const double x2 = (double)(ncell * (int64_t)floorf((float)(g2[d] - 1) / 4.0f)) +
                                            (double)ncell * oc_decode_x(c->izipx, XP(d, ip2));
                          x21[d] = x2 - x1[d];
                        }
                        double rmag = sqrt((x21[0] * x21[0] + x21[1] * x21[1]) + x21[2] * x21[2]);
                        if (rmag == 0) rmag = 1.0;
                        const double rcut = rmag / (double)c->nf_cutoff;
                        const double pcut = (1.0 - (1.75 * powi3(rcut))) + (0.75 * powi5(rcut));
                        const double r3 = powi3(rmag);
                        for (int d = 0; d < 3; d++) {
                          double fp = ((double)mass_p * (x21[d] / r3)) * pcut;
                          if (!(rmag > (double)c->rsoft)) fp = fp * 0.0;
                          f_tot[d] = f_tot[d] + fp;
                        }
                        ipll2 = ll[ipll2 - 1];
                      }
                    }
                for (int d = 0; d < 3; d++) {
                  vreal[d] = vreal[d] + (((f_tot[d] * (double)a_mid) * (double)dt) / 6.0) / (double)PI_F;
                  VP(d, ip1) = (int16_t)oc_encode_v(c->izipv, vreal[d], sigma_vi);
                }
                const double f2 = (f_tot[0] * f_tot[0] + f_tot[1] * f_tot[1]) + f_tot[2] * f_tot[2];
                const double f2r = (double)(float)f2;
                if (f2r > fmx) fmx = f2r;
                ipll1 = ll[ipll1 - 1];
              }
            }
        c->itest1 += n1;
        c->npairs += npr;
        c->pp_ftemp = fmx;
      }
#undef HOC
  free(hoc); free(ll);
  if (f2_max_pp) *f2_max_pp = (float)c->pp_ftemp;
  if (status != OC_OK) return status;
  if (c->itest1 != c->nplocal) return OC_ERR_COUNT;
  return OC_OK;
}

/* ---- IC count sort, WU/initial_conditions.f90:684-786 -------------------- */
/* x, v: n x 3 (global coarse-cell units, velocities in simulation units).
   Particles are binned into cells ceiling(x) (1-based) with periodic wrap,
   kept in input order within a cell; vfield is the cell mean (empty: 0);
   xp=floor(x_local/x_resolution) wrapped, vp=enc(v-vfield). */
int oc_state_from_xv(ocube *c, int64_t n, const double *x, const double *v,
                     const int32_t *pid, float sigma_vi) {
  const int64_t nt = c->nt, nnt = c->nnt, nc = c->nc;
  const int64_t nphys = nc * nc * nc;
  int32_t *cnt = calloc(nphys, sizeof(int32_t));
  double *vsum = calloc(3 * nphys, sizeof(double));
  int64_t *cid = malloc(n * sizeof(int64_t));
  int64_t *start = malloc((nphys + 1) * sizeof(int64_t));
  int16_t *xp = malloc(3 * n * sizeof(int16_t)), *vp = malloc(3 * n * sizeof(int16_t));
  int32_t *pd = malloc(n * sizeof(int32_t));
  float *vf = malloc(3 * nphys * sizeof(float));
  const double xres = x_res(c->izipx);
  for (int64_t p = 0; p < n; p++) {
    int64_t g[3];
    for (int d = 0; d < 3; d++) {
      double xx = fmod(x[3 * p + d], (double)nc);
      if (xx < 0) xx += (double)nc;
      g[d] = (int64_t)ceil(xx);
      if (g[d] < 1) g[d] += nc;          /* x == 0 belongs to the last cell */
      if (g[d] > nc) g[d] -= nc;
    }
    const int64_t t = ((g[0] - 1) / nt) + nnt * (((g[1] - 1) / nt) + nnt * ((g[2] - 1) / nt));
    const int64_t li = (g[0] - 1) % nt, lj = (g[1] - 1) % nt, lk = (g[2] - 1) % nt;
    cid[p] = li + nt * (lj + nt * (lk + nt * t));
    cnt[cid[p]]++;
    for (int d = 0; d < 3; d++) vsum[3 * cid[p] + d] += v[3 * p + d];
  }
  start[0] = 0;
  for (int64_t q = 0; q < nphys; q++) start[q + 1] = start[q] + cnt[q];
  /* vfield: real(4) running sum in input order, then /merge(1,n,n==0)
     (WU/initial_conditions.f90:723,732-734) */
  {
    float *acc = calloc(3 * nphys, sizeof(float));
    for (int64_t p = 0; p < n; p++)
      for (int d = 0; d < 3; d++)
        acc[3 * cid[p] + d] = (float)((double)acc[3 * cid[p] + d] + v[3 * p + d]);
    for (int64_t q = 0; q < nphys; q++)
      for (int d = 0; d < 3; d++) vf[3 * q + d] = acc[3 * q + d] / (float)(cnt[q] ? cnt[q] : 1);
    free(acc);
  }
  (void)vsum;
  int64_t *fill = calloc(nphys, sizeof(int64_t));
  for (int64_t p = 0; p < n; p++) {
    const int64_t q = cid[p], slot = start[q] + fill[q]++;
    for (int d = 0; d < 3; d++) {
      double xx = fmod(x[3 * p + d], (double)nc);
      if (xx < 0) xx += (double)nc;
      xp[3 * slot + d] = (int16_t)wrap_kind(c->izipx, (int64_t)floor(xx / xres));
      vp[3 * slot + d] = (int16_t)oc_encode_v(c->izipv, v[3 * p + d] - (double)vf[3 * q + d], sigma_vi);
    }
    pd[slot] = pid ? pid[p] : (int32_t)(p + 1);
  }
  int st = oc_set_disjoint(c, n, xp, vp, pd, cnt, vf);
  c->npglobal = n;
  free(cnt); free(vsum); free(cid); free(start); free(xp); free(vp); free(pd); free(vf); free(fill);
  return st;
}

/* ---- handle helpers for the ctypes wrapper ------------------------------- */
ocube *oc_alloc(void) { return calloc(1, sizeof(ocube)); }
void oc_free(ocube *c) { if (c) { oc_destroy(c); free(c); } }
void oc_info(const ocube *c, int64_t *o) {
  o[0] = c->nt; o[1] = c->nte; o[2] = c->nft; o[3] = c->nfb; o[4] = c->nfe;
  o[5] = c->np_image_max; o[6] = c->np_tile_max; o[7] = c->np_pp_max;
  o[8] = c->nplocal; o[9] = c->npglobal; o[10] = c->itest1; o[11] = c->npairs;
  o[12] = c->nc; o[13] = c->nnt; o[14] = c->ncb; o[15] = c->np_image;
}
void oc_set_npglobal(ocube *c, int64_t n) { c->npglobal = n; }
double oc_get_overhead(const ocube *c, int which) { return which ? c->overhead_image : c->overhead_tile; }
float oc_get_pp_ftemp(const ocube *c) { return (float)c->pp_ftemp; }
void oc_set_pp_ftemp(ocube *c, float f) { c->pp_ftemp = (double)f; }
int16_t *oc_xp_ptr(ocube *c) { return c->xp; }
int16_t *oc_vp_ptr(ocube *c) { return c->vp; }
int32_t *oc_pid_ptr(ocube *c) { return c->pid; }
int32_t *oc_rhoc_ptr(ocube *c) { return c->rhoc; }
float *oc_vfield_ptr(ocube *c) { return c->vfield; }
int64_t *oc_spine_ptr(ocube *c, int which) {
  int64_t *a[8] = {c->idx_b_l, c->idx_b_r, c->ppl0, c->pplr, c->pprl, c->ppr0, c->ppl, c->ppr};
  return a[which];
}

/* ---- host time integration ---------------------------------------------- */
/* expansion of W/timestep.f90:137-177 (the branch built without
   -Dnew_expansion), kind for kind as gfortran -O3 evaluates it on x86-64:
     dt_x=dt0/2                      real(4)
     omHsq=4.0/9.0                   a real(4) quotient widened to real(8)
     a_x**(-3*wde)                   real exponent: a libm pow call (gcc
                                     expands pow(x,c) inline only for
                                     c in {-1,0,1,2} without -ffast-math)
     a_x**3, a_x**2                  integer powers: (x*x)*x and x*x
     dt_x**2, dt_x**3                real(4) products
     omega_m, omega_l, wde           real(4) parameters (W/parameters.f90:101-107)
   Mixed real(8)*real(4) operands promote where Fortran promotes them. */
static void oc_derivs(double a_x, float om, float ol, float wde, double omhsq, double *adot,
                      double *addot, double *atdot) {
  const float e3 = -3.0f * wde;                       /* -3*wde */
  const float c1 = (1.0f - om) - ol;                  /* 1.0-omega_m-omega_l */
  const float c2 = 1.5f * (1.0f - wde);               /* 1.5*(1.0-wde) */
  const float c3 = (1.5f * (2.0f - 3.0f * wde)) * (1.0f - wde);
  const double a3rlm = (pow(a_x, (double)e3) * (double)ol) / (double)om;
  const double arkm = (a_x * (double)c1) / (double)om;
  const double ax3 = (a_x * a_x) * a_x;
  *adot = sqrt((omhsq * ax3) * ((1.0 + arkm) + a3rlm));
  *addot = ((a_x * a_x) * omhsq) * ((1.5 + 2.0 * arkm) + (double)c2 * a3rlm);
  *atdot = ((a_x * *adot) * omhsq) * ((3.0 + 6.0 * arkm) + (double)c3 * a3rlm);
}

void oc_expansion(float a0, float dt0, float om, float ol, float wde, float *da1, float *da2) {
  const float dt_x = dt0 / 2.0f;
  const float dt2 = dt_x * dt_x;
  const float dt3 = dt2 * dt_x;
  const double omhsq = (double)(4.0f / 9.0f);
  double adot, addot, atdot;
  oc_derivs((double)a0, om, ol, wde, omhsq, &adot, &addot, &atdot);
  *da1 = (float)((adot * (double)dt_x + (addot * (double)dt2) / 2.0) + (atdot * (double)dt3) / 6.0);
  const float a1 = a0 + *da1;                         /* a_x=a0+da1, real(4) sum */
  oc_derivs((double)a1, om, ol, wde, omhsq, &adot, &addot, &atdot);
  *da2 = (float)((adot * (double)dt_x + (addot * (double)dt2) / 2.0) + (atdot * (double)dt3) / 6.0);
}

/* Vectorised form for the tests: n (a0, dt0) pairs. */
void oc_expansion_n(int64_t n, const float *a0, const float *dt0, float om, float ol, float wde,
                    float *da1, float *da2) {
  for (int64_t i = 0; i < n; i++) oc_expansion(a0[i], dt0[i], om, ol, wde, &da1[i], &da2[i]);
}

/* host threads of the oracle's parallel loops (OpenMP) */
void oc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oc_get_threads(void) { return omp_get_max_threads(); }

/* ---------------- FP32 mesh glue of the force solves ----------------
 * The element-wise steps around the FFTs of W/pm.f90:119-126 and :285-315,
 * threaded like the reference's OpenMP build: each is one float operation
 * per element, so the result does not depend on the thread count. */

/* rho_f(::2) = -crho(2::2)*kern, rho_f(2::2) = crho(::2)*kern
   (W/pm.f90:119-120, :281-282): complex64 interleaved in and out */
void oc_green_mul(const float *crho, const float *kern, float *out, int64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < n; ++q) {
    const float re = crho[2 * q], im = crho[2 * q + 1], k = kern[q];
    out[2 * q] = -im * k;
    out[2 * q + 1] = re * k;
  }
}

/* force_f(d,:,:,:) = r(nfb:nfe-nfb+1,...)/nfe/nfe/nfe (W/pm.f90:122-124):
   r is [z][y][x] n^3 (x fastest), out is [z][y][x][3] of m^3 with
   m = n - 2*lo, taking r from index lo on each axis */
void oc_force_extract(const float *r, int64_t n, int64_t lo, int64_t m, int64_t d, float *out) {
  const float fn = (float)n;
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < m; ++z)
    for (int64_t y = 0; y < m; ++y) {
      const float *src = r + ((z + lo) * n + (y + lo)) * n + lo;
      float *dst = out + ((z * m + y) * m) * 3 + d;
      for (int64_t x = 0; x < m; ++x) {
        float v = src[x];
        v = v / fn;
        v = v / fn;
        v = v / fn;
        dst[3 * x] = v;
      }
    }
}

/* max over cells of (F1*F1 + F2*F2) + F3*F3 in real(4) (W/pm.f90:126, :315) */
float oc_f2_max(const float *f, int64_t ncell) {
  float mx = 0.0f;
#pragma omp parallel for schedule(static) reduction(max : mx)
  for (int64_t q = 0; q < ncell; ++q) {
    const float a = f[3 * q], b = f[3 * q + 1], c = f[3 * q + 2];
    const float s = (a * a + b * b) + c * c;
    if (s > mx) mx = s;
  }
  return mx;
}
