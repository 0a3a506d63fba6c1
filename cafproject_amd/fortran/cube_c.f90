! cube_c.f90 -- ISO_C_BINDING interface to the C ABI of include/cube.h, for
! a Fortran driver that keeps the reference's call shape
! (CUBEnu/work/main/main.f90:102-118): each reference subroutine call
! becomes one function call on an opaque context, returning a status code.
module cube_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: CUBE_OK = 0

  ! cube_params: compile-time constants of W/parameters.f90:8-123
  type, bind(C) :: cube_params
    integer(c_int64_t) :: nc, nnt, ncb, np_nc
    integer(c_int32_t) :: izipx, izipv, has_pid, pp_on, pp_range
    real(c_float) :: rsoft, image_buffer, tile_buffer
    integer(c_int32_t) :: device, mem_layout
  end type

  ! outputs of update_x (W/update_particle.f90:172-250)
  type, bind(C) :: cube_vstats
    real(c_double) :: sum_c, sum_r, sum_s, std_vsim, std_vsim_c, std_vsim_res
    real(c_float) :: sigma_vi_new
    integer(c_int32_t) :: nlayer
    integer(c_int64_t) :: nplocal, lost
    real(c_double) :: overhead_tile
  end type

  ! outputs of particle_mesh (W/pm.f90:391-410)
  type, bind(C) :: cube_pm_out
    real(c_float) :: f2_max_fine, f2_max_coarse, dt_fine, dt_coarse, dt_vmax, vz_max
    real(c_float) :: vmax(3), reserved
    real(c_double) :: testrhof, testrhoc
  end type

  interface
    integer(c_int) function cube_ctx_create(p, ctx) bind(C, name="cube_ctx_create")
      import :: c_int, c_ptr, cube_params
      type(cube_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function
    subroutine cube_ctx_destroy(ctx) bind(C, name="cube_ctx_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function cube_strerror(code) bind(C, name="cube_strerror")
      import :: c_ptr, c_int
      integer(c_int), value :: code
    end function
    integer(c_int) function cube_init_kernels(ctx, wfxyzf, wfxyzc) bind(C, name="cube_init_kernels")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(in) :: wfxyzf(*), wfxyzc(*)
    end function
    integer(c_int) function cube_set_state(ctx, nplocal, npglobal, xp, vp, pid, rhoc, vfield) &
        bind(C, name="cube_set_state")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, xp, vp, pid, rhoc, vfield
      integer(c_int64_t), value :: nplocal, npglobal
    end function
    integer(c_int) function cube_get_state(ctx, nplocal, xp, vp, pid, rhoc, vfield) &
        bind(C, name="cube_get_state")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, xp, vp, pid, rhoc, vfield
      integer(c_int64_t), intent(out) :: nplocal
    end function
    integer(c_int) function cube_update_x(ctx, dt_old, dt, vz_max, sigma_vi, out) bind(C, name="cube_update_x")
      import :: c_int, c_ptr, c_float, cube_vstats
      type(c_ptr), value :: ctx
      real(c_float), value :: dt_old, dt, vz_max, sigma_vi
      type(cube_vstats), intent(out) :: out
    end function
    integer(c_int) function cube_buffer_grid(ctx) bind(C, name="cube_buffer_grid")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function cube_buffer_x(ctx) bind(C, name="cube_buffer_x")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function cube_buffer_v(ctx) bind(C, name="cube_buffer_v")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function cube_pp_force(ctx, a_mid, dt, sigma_vi, mass_p, f2_max_pp) bind(C, name="cube_pp_force")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      real(c_float), value :: a_mid, dt, sigma_vi, mass_p
      real(c_float), intent(out) :: f2_max_pp
    end function
    integer(c_int) function cube_particle_mesh(ctx, a_mid, dt, sigma_vi, sigma_vi_new, mass_p, out) &
        bind(C, name="cube_particle_mesh")
      import :: c_int, c_ptr, c_float, cube_pm_out
      type(c_ptr), value :: ctx
      real(c_float), value :: a_mid, dt, sigma_vi, sigma_vi_new, mass_p
      type(cube_pm_out), intent(out) :: out
    end function
  end interface

contains

  ! stop with the library's message, as the reference stops on its errors
  subroutine cube_check(st, what)
    integer(c_int), intent(in) :: st
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (st == CUBE_OK) return
    call c_f_pointer(cube_strerror(st), msg, [256])
    n = 0
    do while (n < 256)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    print *, 'cube: ', what, ' failed: ', msg(1:n)
    error stop 1
  end subroutine

end module cube_c
