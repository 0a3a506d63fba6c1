/*
 * cube.h -- C ABI of the MI355X CUBE particle-mesh step (libcube_hip.so).
 *
 * The reference has no FFI: its hot path is a set of parameterless Fortran
 * subroutines that mutate module state, called from the step loop of
 * CUBEnu/work/main/main.f90:102-118.  Each entry point below replaces one of
 * them for a single image (nn=1) covering the whole periodic box on one GPU,
 * keeping the call order and argument meaning:
 *
 *   call timestep       -> host (FP64 a(t)), W/timestep.f90:1-92
 *   call update_x       -> cube_update_x        (W/update_particle.f90:5-256)
 *   call buffer_grid    -> cube_buffer_grid     (W/buffer_grid.f90:5-209)
 *   call buffer_x       -> cube_buffer_x        (W/buffer_particle.f90:5-14)
 *   call ext_pp_force   -> cube_pp_force        (W/ext_pp_force.f90:10-162)
 *   call particle_mesh  -> cube_particle_mesh   (W/pm.f90:2-415)
 *   call buffer_v       -> cube_buffer_v        (W/buffer_particle.f90:16-25)
 *   call kernel_f/_c    -> cube_init_kernels    (W/kernel_f.f90, W/kernel_c.f90)
 *   particle_initialization / checkpoint
 *                       -> cube_set_state / cube_get_state
 *                          (W/particle_initialization.f90:34-82,
 *                           W/checkpoint.f90:38-44)
 * (W/ = CUBEnu/work/main/.)
 *
 * Conventions
 *  - Every entry point returns int: 0 = success, otherwise a CUBE_ERR_* code;
 *    each reference `stop` condition has its own code.  No exception or
 *    longjmp crosses the ABI.
 *  - The context owns all device memory; host buffers are copied, never kept.
 *  - Arrays use the reference's checkpoint layout (Fortran column-major):
 *      xp  : izipx-byte signed integers, 3 per particle (x fastest)
 *      vp  : izipv-byte signed integers, 3 per particle
 *      pid : int32 per particle (optional)
 *      rhoc: int32 per coarse cell, order (i,j,k,tile_x,tile_y,tile_z)
 *      vfield: 3 float per coarse cell, same cell order
 *    Particles are cell-sorted in that cell order.
 *  - Calls are synchronous for their scalar outputs; one host thread per
 *    context; not re-entrant.  Particle counts and indices are int64.
 */
#ifndef CUBE_H
#define CUBE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CUBE_OK = 0,
  CUBE_ERR_TILE_OVERFLOW = 1,  /* W/update_particle.f90:88-94   tile_buffer too small  */
  CUBE_ERR_IMAGE_OVERFLOW = 2, /* W/buffer_grid.f90:148-154     image_buffer too small */
  CUBE_ERR_PP_OVERFLOW = 3,    /* W/ext_pp_force.f90:41-45      np_pp_max too small    */
  CUBE_ERR_CHECKSUM = 4,       /* W/buffer_grid.f90:174-177     particle checksum      */
  CUBE_ERR_COUNT = 5,          /* W/update_particle.f90:245-250, W/ext_pp_force.f90:156-160 */
  CUBE_ERR_DISPLACEMENT = 6,   /* a particle moved past the buffer (rhoce bounds)     */
  CUBE_ERR_HIP = 10,           /* a HIP runtime call failed                           */
  CUBE_ERR_ALLOC = 11,         /* device allocation failed                            */
  CUBE_ERR_INVALID = 12,       /* bad argument or unsupported size                    */
  CUBE_ERR_STATE = 13          /* call out of order (no state / no kernels)           */
};

/* Compile-time constants of W/parameters.f90:8-123 that the path reads. */
typedef struct {
  int64_t nc;        /* coarse cells per dim of the (single) image       */
  int64_t nnt;       /* tiles per dim (fixes the reference visit order)   */
  int64_t ncb;       /* buffer depth in coarse cells (6)                  */
  int64_t np_nc;     /* particles per coarse cell per dim (np_image)      */
  int32_t izipx;     /* bytes per position component: 1 or 2             */
  int32_t izipv;     /* bytes per velocity component: 1 or 2             */
  int32_t has_pid;   /* carry particle ids                               */
  int32_t pp_on;     /* Extended_pp_force                                 */
  int32_t pp_range;  /* PP neighbour range in fine cells (<=4)            */
  float rsoft;       /* PP softening length (fine cells)                  */
  float image_buffer;/* reference image_buffer (1.5)                      */
  float tile_buffer; /* reference tile_buffer (2.5)                       */
  int32_t device;    /* HIP device ordinal                                */
  int32_t mem_layout;/* fine mesh layout: 0 = compact when the full one
                        does not fit the device's free memory, 1 = full,
                        2 = compact (one force spectrum at a time, drift
                        scratch inside the force mesh; same results)     */
} cube_params;

/* Outputs of update_x (W/update_particle.f90:172-250). */
typedef struct {
  double sum_c, sum_r, sum_s;              /* FP64 sums of v_c^2, v_res^2, v^2   */
  double std_vsim, std_vsim_c, std_vsim_res;
  float sigma_vi_new;                      /* std_vsim_res/sqrt(3.), real(4)     */
  int32_t nlayer;                          /* 2*ceiling(dt_mid*vz_max/ncell)+1   */
  int64_t nplocal;                         /* particles after the drift          */
  int64_t lost;                            /* particles that left every tile     */
  double overhead_tile;                    /* max tile fill / np_tile_max        */
} cube_vstats;

/* Outputs of particle_mesh (W/pm.f90:391-410). */
typedef struct {
  float f2_max_fine, f2_max_coarse;
  float dt_fine, dt_coarse, dt_vmax, vz_max;
  float vmax[3];
  float reserved;
  double testrhof, testrhoc;               /* total fine / coarse mass          */
} cube_pm_out;

typedef struct cube_ctx cube_ctx;

/* Context owning device memory, sized from the parameters. */
int cube_ctx_create(const cube_params *p, cube_ctx **out);
void cube_ctx_destroy(cube_ctx *ctx);
const char *cube_strerror(int code);
/* ints[0..15]: nt, nf, nh, np_image, np_image_max, np_tile_max, np_pp_max,
   capacity, nplocal, npglobal, ... ; used by the bindings */
int cube_ctx_info(const cube_ctx *ctx, int64_t *ints16);

/* kernel_f + kernel_c (W/initialize.f90:37-38).  wfxyzf: 16*16*16*3 floats
   of CUBEnu/kernels/wfxyzf.3.ascii, order (i,j,k) with i fastest, 3
   components per entry; wfxyzc: 4*4*4*3 floats of wfxyzc.2.ascii. */
int cube_init_kernels(cube_ctx *ctx, const float *wfxyzf, const float *wfxyzc);

/* particle_initialization: disjoint state in checkpoint layout. */
int cube_set_state(cube_ctx *ctx, int64_t nplocal, int64_t npglobal, const void *xp,
                   const void *vp, const int32_t *pid, const int32_t *rhoc,
                   const float *vfield);
/* checkpoint: copy the disjoint state back (any pointer may be NULL). */
int cube_get_state(cube_ctx *ctx, int64_t *nplocal, void *xp, void *vp, int32_t *pid,
                   int32_t *rhoc, float *vfield);

/* The step, in W/main.f90:102-118 order. */
int cube_update_x(cube_ctx *ctx, float dt_old, float dt, float vz_max, float sigma_vi,
                  cube_vstats *out);
int cube_buffer_grid(cube_ctx *ctx);
int cube_buffer_x(cube_ctx *ctx);
int cube_pp_force(cube_ctx *ctx, float a_mid, float dt, float sigma_vi, float mass_p,
                  float *f2_max_pp);
int cube_particle_mesh(cube_ctx *ctx, float a_mid, float dt, float sigma_vi,
                       float sigma_vi_new, float mass_p, cube_pm_out *out);
int cube_buffer_v(cube_ctx *ctx);

/* particle_mesh split into its stages (W/pm.f90:49-126, 129-159, 202-267,
   271-315, 319-387), for callers that interleave their own work and for
   per-stage parity checks. */
int cube_pm_fine_deposit(cube_ctx *ctx, float mass_p, double *testrhof);
int cube_pm_fine_force(cube_ctx *ctx, float *f2_max_fine);
int cube_pm_fine_kick(cube_ctx *ctx, float a_mid, float dt, float sigma_vi,
                      float sigma_vi_new);
/* Fine potential of the current fine density (after cube_pm_fine_deposit),
   read with CUBE_FIELD_PHI_F.  The reference's step never forms one; it
   ships the fine potential kernel CUBEnu/kernels/wg1.ascii (README:8-9):
   wg1 = its 16^3 values, (i,j,k) order with i fastest, mirrored evenly
   into the periodic nf^3 mesh; phi = C2R(K * R2C(rho_f)) / nf^3. */
int cube_pm_fine_potential(cube_ctx *ctx, const float *wg1);
int cube_pm_coarse_deposit(cube_ctx *ctx, float mass_p, double *testrhoc);
int cube_pm_coarse_force(cube_ctx *ctx, float *f2_max_coarse);
int cube_pm_coarse_kick(cube_ctx *ctx, float a_mid, float dt, float sigma_vi, float *vmax);

/* Initial conditions (the IC generator CUBEnu/work/utilities/
   initial_conditions.f90, on the GPU): seeded white noise on the particle
   lattice (np_nc*nc per dim), delta_k = w_k sqrt(P(k)/cell^3), phi_k =
   -4 pi delta_k/k~^2, Zel'dovich x and v from central differences of phi,
   then the count sort into the disjoint state (xp, vp, pid, rhoc, vfield).
   P(k) at the initial scale factor comes as a table of ln P on a uniform
   ln k grid (k in h/Mpc).  sigma_vi <= 0 measures it from the residual
   velocities; the value used is returned. */
typedef struct {
  uint64_t seed;
  double box;        /* Mpc/h */
  double vfactor;    /* velocity factor of the reference IC, vfactor(a) */
  float sigma_vi;    /* <= 0: measure */
  int32_t ntab;
  double lnk0, dlnk; /* ln k of the first entry, spacing */
  const double *lnp; /* ntab values of ln P(k) [(Mpc/h)^3] */
} cube_ic_params;
int cube_ic_zeldovich(cube_ctx *ctx, const cube_ic_params *ip, float *sigma_vi_out);
/* Count sort of continuous particles into the disjoint state: x in global
   coarse-cell units (n x 3 doubles), v in simulation units (n x 3 floats),
   cell = ceiling(x), input order kept inside a cell (WU/:684-786). */
int cube_set_state_xv(cube_ctx *ctx, int64_t n, const double *x, const float *v, float sigma_vi,
                      float *sigma_vi_out);

/* Mesh fields, for diagnostics and checks.  which: */
enum {
  CUBE_FIELD_RHO_F = 0,   /* nf^3 float, fine density, (x,y,z) x fastest           */
  CUBE_FIELD_FORCE_F = 1, /* nf^3 x 3 float, fine force, 3 components per cell;
                              after ext_pp_force or particle_mesh it holds the
                              fine kick's factors (((F*a_mid)*dt)/6)/pi
                              (cube_pm_fine_force forms F itself)             */
  CUBE_FIELD_RHO_C = 2,   /* nc^3 float, coarse density                             */
  CUBE_FIELD_FORCE_C = 3, /* nc^3 x 3 float, coarse force                           */
  CUBE_FIELD_KERN_F = 4,  /* 3 x (nf/2+1)^3 float, fine Green's function octant     */
  CUBE_FIELD_KERN_C = 5,  /* 3 x nc x nc x (nc/2+1) float, coarse Green's function  */
  CUBE_FIELD_PHI_F = 6    /* nf^3 float, fine potential of cube_pm_fine_potential
                              (read only; it shares the force mesh's memory, so
                              any later call that forms forces replaces it)   */
};
int cube_get_field(cube_ctx *ctx, int which, void *host, size_t bytes);
int cube_set_field(cube_ctx *ctx, int which, const void *host, size_t bytes);

/* Per-stage device time, measured with HIP events on the context's stream
   (the counterpart of the reference's per-stage system_clock prints, e.g.
   W/pm.f90:195-196, W/update_particle.f90:251-252).  Stage ids: */
enum {
  CUBE_T_DRIFT = 0,        /* update_x: destination, sort, re-quantise     */
  CUBE_T_VSTATS = 1,       /* update_x: velocity statistics                */
  CUBE_T_PP = 2,           /* ext_pp_force                                 */
  CUBE_T_FINE_DEPOSIT = 3, /* fine CIC deposit                             */
  CUBE_T_FINE_FFT = 4,     /* fine r2c x + c2c y                           */
  CUBE_T_FINE_GREEN = 5,   /* fine z pass: fft, i*K_d, inverse fft (x3)   */
  CUBE_T_FINE_INVY = 6,    /* fine inverse y passes (x3)                   */
  CUBE_T_FINE_C2R = 7,     /* fine c2r x of the three components + max F^2 */
  CUBE_T_FINE_KICK = 8,    /* fine force gather + kick; in particle_mesh
                              both kicks, which run as one pass            */
  CUBE_T_COARSE_DEPOSIT = 9,
  CUBE_T_COARSE_FORCE = 10,
  CUBE_T_COARSE_KICK = 11, /* the coarse kick called on its own            */
  CUBE_T_CHECKS = 12,      /* overflow checks (tile/image/pp box sums)     */
  CUBE_NTIMERS = 13
};
/* on=1 starts recording (adds a synchronisation per stage), on=0 stops. */
int cube_profile(cube_ctx *ctx, int on);
/* ms[CUBE_NTIMERS] accumulated milliseconds, calls[CUBE_NTIMERS] launches
   counted; reset=1 clears the accumulators after reading. */
int cube_timers(cube_ctx *ctx, double *ms, int64_t *calls, int reset);

/* Wait for all work of the context; returns CUBE_ERR_HIP on a device error. */
int cube_sync(cube_ctx *ctx);

/* Matter power spectrum of the current particle state, as the reference's
   analysis computes it (WU/cicpower.f90:96-187, WU/powerspectrum.f90:28-135):
   cell-centred CIC on an ng^3 mesh (ng divides 4*nc, ng <= 4*nc), delta =
   rho/mean - 1, delta_k = FFT/ng^3, the kx=0 and kx=ng/2 planes' Hermitian
   duplicates skipped, linear bins ibin = nint(|k|) (k in fundamental
   units), Delta^2 = |delta_k|^2 / sinc^4 * 4 pi |k|^3.  Writes, for bins
   1..nbin (nbin = nint(ng/2 * sqrt(3))), the per-bin sums of |k|, Delta^2
   and the mode counts; the caller divides.  Uses the fine-mesh buffers. */
int cube_power(cube_ctx *ctx, int32_t ng, int32_t nbin, double *ksum, double *psum, int64_t *count);

/* cube_power without any allocation in the call.  No counterpart in the
   reference: for callers that cannot allocate device memory (code under
   HIP graph capture, in-loop validation of a serving loop).  The twiddle
   tables and the bin sums live in a fixed device array of the library
   (about 88 KB per device, held until the process ends); meshes up to
   ng = 2048 and nbin <= 1774, else CUBE_ERR_INVALID.  Calls that overlap on
   one device share that array and race.  Same arguments and results as
   cube_power (its FP64 sums are atomic, so equal to rounding). */
int cube_power_noalloc(cube_ctx *ctx, int32_t ng, int32_t nbin, double *ksum, double *psum, int64_t *count);

/* Host-side twins of device math, exported for the bindings' checks. */
double cube_host_atan(double x);

#ifdef __cplusplus
}
#endif
#endif
