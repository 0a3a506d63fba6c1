"""Host-side time integration: `timestep` and `expansion` of
CUBEnu/work/main/timestep.f90:1-92,133-177.

The reference keeps a, da, dt, a_mid in real(4) and evaluates the
third-order Taylor step of the Friedmann equation with real(8) internals.
Each sub-expression here has the kind gfortran gives it there (numpy
float32 for real(4), Python float for real(8)), so a run driven by this
module takes the same (a, dt) sequence as the reference given the same dt
constraints:

  dt_x = dt0/2, dt_x**2, dt_x**3      real(4)  (W/timestep.f90:143,158,175)
  omHsq = 4.0/9.0                     the real(4) quotient widened (:145,161)
  a_x**(-3*wde)                       a libm pow call (real exponent, :146,163)
  a_x**3, a_x**2                      (x*x)*x and x*x (integer powers, :150-156)
  a_x = a0 + da1                      a real(4) sum widened (:160)
  abs((a+da)/a_next-1) >= 1e-6        real(4) (:55)
"""
import math

import numpy as np

f32 = np.float32


def _derivs(a_x, om, ol, w, omhsq):
    """adot, addot, atdot at a_x (real(8)); om, ol, w real(4) parameters."""
    e3 = float(f32(-3) * w)                                   # -3*wde, real(4)
    c1 = float(f32(f32(f32(1) - om) - ol))                    # 1.0-omega_m-omega_l
    c2 = float(f32(1.5) * f32(f32(1) - w))                    # 1.5*(1.0-wde)
    c3 = float(f32(f32(1.5) * f32(f32(2) - f32(3) * w)) * f32(f32(1) - w))
    a3rlm = (math.pow(a_x, e3) * float(ol)) / float(om)
    arkm = (a_x * c1) / float(om)
    adot = math.sqrt((omhsq * ((a_x * a_x) * a_x)) * ((1.0 + arkm) + a3rlm))
    addot = ((a_x * a_x) * omhsq) * ((1.5 + 2.0 * arkm) + c2 * a3rlm)
    atdot = ((a_x * adot) * omhsq) * ((3.0 + 6.0 * arkm) + c3 * a3rlm)
    return adot, addot, atdot


def expansion(a0, dt0, omega_m, omega_l, wde=-1.0):
    """da1, da2 of one step: W/timestep.f90:137-177 (the default branch,
    without the radiation term).  a0, dt0 real(4); returns a real(4) pair."""
    om, ol, w = f32(omega_m), f32(omega_l), f32(wde)
    a0 = f32(a0)
    dt_x = f32(f32(dt0) / f32(2))
    dt2 = f32(dt_x * dt_x)
    dt3 = f32(dt2 * dt_x)
    omhsq = float(f32(4.0) / f32(9.0))
    d, d2, d3 = float(dt_x), float(dt2), float(dt3)
    adot, addot, atdot = _derivs(float(a0), om, ol, w, omhsq)
    da1 = f32((adot * d + (addot * d2) / 2.0) + (atdot * d3) / 6.0)
    a_x = float(f32(a0 + da1))
    adot, addot, atdot = _derivs(a_x, om, ol, w, omhsq)
    da2 = f32((adot * d + (addot * d2) / 2.0) + (atdot * d3) / 6.0)
    return da1, da2


class Clock:
    """a, t, tau, dt, dt_old, a_mid of W/variables.f90:22-23 and the
    checkpoint bookkeeping of W/timestep.f90."""

    def __init__(self, a, z_checkpoints, omega_m, omega_l, dt_max=1.0, ra_max=0.2):
        self.a = f32(a)
        self.t = f32(0)
        self.tau = f32(0)
        self.dt = f32(0)
        self.dt_old = f32(0)
        self.da = f32(0)
        self.a_mid = f32(a)
        self.om, self.ol = omega_m, omega_l
        self.dt_max, self.ra_max = f32(dt_max), f32(ra_max)
        self.z_checkpoint = [f32(z) for z in z_checkpoints]
        self.cur_checkpoint = 1  # index of the next checkpoint (W/main.f90:79)
        self.checkpoint_step = False
        self.final_step = False
        self.istep = 0

    def timestep(self, dt_fine=1000.0, dt_coarse=1000.0, dt_pp=1000.0, dt_vmax=1000.0):
        """W/timestep.f90:1-92 for one image."""
        self.dt_old = self.dt
        self.istep += 1
        a = self.a
        dt_e = self.dt_max
        ntry = 0
        while True:
            ntry += 1
            da1, da2 = expansion(a, dt_e, self.om, self.ol)
            da = f32(da1 + da2)
            ra = f32(da / f32(a + da))
            if ra > self.ra_max:
                dt_e = f32(dt_e * f32(self.ra_max / ra))
            else:
                break
            if ntry > 10:
                break
        dt = f32(min(dt_e, f32(dt_fine), f32(dt_coarse), f32(dt_pp), f32(dt_vmax)))
        da1, da2 = expansion(a, dt, self.om, self.ol)
        da = f32(da1 + da2)
        self.checkpoint_step = False
        z_next = self.z_checkpoint[self.cur_checkpoint - 1]
        a_next = f32(f32(1.0) / f32(1 + z_next))
        if da >= f32(a_next - a):
            self.checkpoint_step = True
            if self.cur_checkpoint == len(self.z_checkpoint):
                self.final_step = True
            n = 0
            while abs(f32(f32(f32(a + da) / a_next) - f32(1))) >= f32(1e-6) and n < 100:
                dt = f32(dt * f32(a_next - a) / da)
                da1, da2 = expansion(a, dt, self.om, self.ol)
                da = f32(da1 + da2)
                n += 1
        self.a_mid = f32(a + f32(da / f32(2)))
        self.tau = f32(self.tau + dt)
        self.t = f32(self.t + dt)
        self.a = f32(a + da)
        self.da = da
        self.dt = dt
        return self
