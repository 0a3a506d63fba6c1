! cube_clock.f90 -- host time integration of the Fortran driver.
!
! expansion of CUBEnu/work/main/timestep.f90:133-177 (the branch built
! without -Dnew_expansion), each sub-expression in the kind the reference
! gives it:
!   dt_x=dt0/2, dt_x**2, dt_x**3    real(4)                      (:143,158,175)
!   omHsq=4.0/9.0                   a real(4) quotient in real(8) (:145,161)
!   a_x**(-3*wde)                   real exponent: a pow call     (:146,163)
!   a_x**3, a_x**2                  integer powers                (:150-156)
!   a_x=a0+da1                      a real(4) sum                 (:160)
! omega_m, omega_l and wde are real(4), as the reference's parameters
! (W/parameters.f90:101-107).  bind(C) so that the CPU tests can call it.
module cube_clock
  use iso_c_binding
  implicit none
  private
  public :: expansion

contains

  subroutine expansion(a0, dt0, om, ol, wde, da1, da2) bind(C, name='cube_clock_expansion')
    real(c_float), value :: a0, dt0, om, ol, wde
    real(c_float), intent(out) :: da1, da2
    real(4) :: dt_x, dt2, dt3
    real(8) :: omhsq, adot, addot, atdot
    dt_x = dt0 / 2
    dt2 = dt_x**2
    dt3 = dt_x**3
    omhsq = 4.0 / 9.0
    call derivs(real(a0, 8), om, ol, wde, omhsq, adot, addot, atdot)
    da1 = real(adot * dt_x + (addot * dt2) / 2.0 + (atdot * dt3) / 6.0, 4)
    call derivs(real(a0 + da1, 8), om, ol, wde, omhsq, adot, addot, atdot)
    da2 = real(adot * dt_x + (addot * dt2) / 2.0 + (atdot * dt3) / 6.0, 4)
  end subroutine

  subroutine derivs(a_x, om, ol, wde, omhsq, adot, addot, atdot)
    real(8), intent(in) :: a_x, omhsq
    real(4), intent(in) :: om, ol, wde
    real(8), intent(out) :: adot, addot, atdot
    real(8) :: a3rlm, arkm, e3
    real(4) :: c1, c2, c3
    e3 = -3 * wde
    c1 = 1.0 - om - ol
    c2 = 1.5 * (1.0 - wde)
    c3 = 1.5 * (2.0 - 3.0 * wde) * (1.0 - wde)
    a3rlm = a_x**e3 * ol / om
    arkm = a_x * c1 / om
    adot = sqrt(omhsq * a_x**3 * (1.0 + arkm + a3rlm))
    addot = a_x**2 * omhsq * (1.5 + 2.0 * arkm + c2 * a3rlm)
    atdot = a_x * adot * omhsq * (3.0 + 6.0 * arkm + c3 * a3rlm)
  end subroutine

end module cube_clock
