/*
 * cube_math.h -- fixed-point codecs of the CUBE particle state, shared by
 * host and device code (compiled with -ffp-contract=off everywhere, so each
 * expression rounds exactly as written).
 *
 * Position: x = (cell-1) + (unsigned(xp)+0.5)*2^-8izipx   (ms_caf.tex:69-74;
 *   W/update_particle.f90:66, W/parameters.f90:16-19,119).
 * Velocity: v = v_c + tan(pi*vp/(nvbin-1)) / (sqrt(pi/2)/(sigma*2.5))
 *           vp = nint((nvbin-1)*atan(sqrt(pi/2)/(sigma*2.5)*(v-v_c))/pi)
 *   (ms_caf.tex:114-119; W/update_particle.f90:67,117; W/pm.f90:145,158).
 * Kinds follow the reference: pi, the tan argument and the tan are real(4),
 * the divisor and everything after it real(8).
 */
#ifndef CUBE_MATH_H
#define CUBE_MATH_H
#include <stdint.h>

#if defined(__HIPCC__)
#define CUBE_HD __host__ __device__
#else
#define CUBE_HD
#endif

#define CUBE_PI_F 3.14159274101257324f /* pi=4*atan(1.) in real(4) */

/* round half away from zero (Fortran nint), exact in FP64 */
CUBE_HD static inline double cube_nint(double x) {
  const double t = __builtin_trunc(x);
  const double f = x - t; /* exact */
  if (f >= 0.5) return t + 1.0;
  if (f <= -0.5) return t - 1.0;
  return t;
}

/* two's-complement wrap to a 1- or 2-byte integer */
CUBE_HD static inline int32_t cube_wrap(int bytes, int64_t v) {
  return bytes == 1 ? (int32_t)(int8_t)(uint8_t)(v & 0xff) : (int32_t)(int16_t)(uint16_t)(v & 0xffff);
}

/* unsigned code of a position component */
CUBE_HD static inline uint32_t cube_ucode(int bytes, int32_t xp) {
  return bytes == 1 ? (uint32_t)(xp & 0xff) : (uint32_t)(xp & 0xffff);
}

/* (unsigned(xp)+0.5)*2^-8izipx, exact */
CUBE_HD static inline double cube_xoff(int bytes, int32_t xp) {
  const double res = bytes == 1 ? 1.0 / 256.0 : 1.0 / 65536.0;
  return ((double)cube_ucode(bytes, xp) + 0.5) * res;
}

/* sqrt(pi/2)/(sigma*vrel_boost): real(4)/real(8) with vrel_boost=2.5.
   sqrt(pi/2) in real(4) is the correctly rounded 1.2533141f. */
CUBE_HD static inline double cube_vconst(float sigma) {
  return (double)1.25331413745880127f / ((double)sigma * 2.5);
}

/*
 * atan in FP64 from +,-,*,/ only, so host and device agree bit for bit.
 * Method: the argument reduction of Cody & Waite as used by fdlibm's s_atan.c
 * (intervals split at 7/16, 11/16, 19/16, 39/16 with atan(1/2), atan(1),
 * atan(3/2), pi/2 stored as hi+lo pairs), and on |t|<7/16 the form
 * atan(t) = t - t*z*P(z), z=t^2, with P the degree-10 Chebyshev interpolant
 * of (t-atan t)/t^3 on [0,(7/16)^2] computed for this file in 60-digit
 * arithmetic (relative error of the interpolant 1.7e-17).
 */
CUBE_HD static inline double cube_atan(double x) {
  const double a0 = 0.33333333333333326, a1 = -0.1999999999998777, a2 = 0.1428571428314864,
               a3 = -0.11111110900560475, a4 = 0.09090900195560148, a5 = -0.07692087348767168,
               a6 = 0.06663239299107136, a7 = -0.05847724720645106, a8 = 0.05033942776485848,
               a9 = -0.03782521866106712, a10 = 0.017547235157290005;
  const double hi[4] = {0.4636476090008061, 0.7853981633974483, 0.982793723247329,
                        1.5707963267948966};
  const double lo[4] = {2.2698777452961687e-17, 3.061616997868383e-17, 1.3903311031230998e-17,
                        6.123233995736766e-17};
  const int neg = x < 0;
  double ax = neg ? -x : x;
  if (!(ax == ax)) return x;              /* NaN */
  if (ax >= 7.378697629483821e19) {       /* 2^66: atan = pi/2 */
    const double r = hi[3] + lo[3];
    return neg ? -r : r;
  }
  if (ax < 3.725290298461914e-09) return x; /* 2^-28 */
  /* the reduced argument t = num/den of the interval, chosen by selects so
     that one division serves every lane */
  double num = ax, den = 1.0, h = 0.0, l = 0.0;
  int id = -1;
  if (ax >= 2.4375) {
    id = 3;
    num = -1.0;
    den = ax;
  } else if (ax >= 1.1875) {
    id = 2;
    num = ax - 1.5;
    den = 1.0 + 1.5 * ax;
  } else if (ax >= 0.6875) {
    id = 1;
    num = ax - 1.0;
    den = ax + 1.0;
  } else if (ax >= 0.4375) {
    id = 0;
    num = 2.0 * ax - 1.0;
    den = 2.0 + ax;
  }
  if (id >= 0) {
    h = hi[id];
    l = lo[id];
  }
  const double t = num / den; /* ax/1 = ax exactly on the first interval */
  const double z = t * t;
  const double p = a0 + z * (a1 + z * (a2 + z * (a3 + z * (a4 + z * (a5 + z * (a6 + z * (a7 + z * (a8 + z * (a9 + z * a10)))))))));
  const double r = id < 0 ? t - t * (z * p) : h - ((t * (z * p) - l) - t);
  return neg ? -r : r;
}

/* vp code of a residual velocity: nint((nvbin-1)*atan(c*v)/pi) with
   c = cube_vconst(sigma); W/update_particle.f90:117 */
CUBE_HD static inline int32_t cube_encode_v(int bytes, double c, double v) {
  const double nvm1 = bytes == 1 ? 255.0 : 65535.0;
  const double y = (nvm1 * cube_atan(c * v)) / (double)CUBE_PI_F;
  return cube_wrap(bytes, (int64_t)cube_nint(y));
}

#endif
