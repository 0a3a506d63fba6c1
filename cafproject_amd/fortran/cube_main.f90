! cube_main.f90 -- Fortran driver of the GPU CUBE step.
!
! The step loop of CUBEnu/work/main/main.f90:102-118 with every reference
! subroutine replaced by its C-ABI counterpart (module cube_c):
!   timestep -> update_x -> buffer_grid -> buffer_x -> [ext_pp_force]
!   -> particle_mesh -> buffer_v
! State comes from a checkpoint in the reference's file format
! (W/particle_initialization.f90:22-82) and goes back to one
! (W/checkpoint.f90:12-55).  The host time integration (a(t) in FP64
! internals, a and dt in real(4), W/timestep.f90) stays here on the host.
!
! usage: cube_main <in_dir> <z_in> <out_dir> <z_out> <nsteps> <np_nc> <pp_on> <tables_dir>
! Prints one line per step with the bit patterns of a, dt, dt_old, a_mid,
! sigma_vi (so another driver can replay the same scalars).
program cube_main
  use iso_c_binding
  use cube_c
  use cube_clock
  implicit none

  ! sim_header (W/parameters.f90:136-160): 17 integer(8) + 22 real(4)
  integer(8) :: hi(17)
  real(4) :: hr(22)
  character(512) :: in_dir, out_dir, tables
  character(32) :: arg
  real(4) :: z_in, z_out
  integer :: nsteps, np_nc_arg, pp_arg, istep, u, ios
  integer(8) :: nplocal, npglobal, nc, nnt, ncb, n3
  integer(c_int) :: st
  type(cube_params), target :: p
  type(c_ptr) :: ctx
  type(cube_vstats) :: vs
  type(cube_pm_out) :: pm
  integer(1), allocatable, target :: xp1(:), vp1(:)
  integer(2), allocatable, target :: xp2(:), vp2(:)
  integer(4), allocatable, target :: pid(:), rhoc(:)
  real(4), allocatable, target :: vfield(:)
  real(c_float), target :: wf(3 * 16 * 16 * 16), wc(3 * 4 * 4 * 4)
  type(c_ptr) :: pxp, pvp, ppid
  logical :: has_pid
  ! clock, W/variables.f90:22-23
  real(4) :: a, t, tau, dt, dt_old, da, a_mid, om, ol, sigma_vi, vz_max, mass_p
  real(4) :: dt_fine, dt_coarse, dt_pp, dt_vmax, f2pp, gg

  call get_command_argument(1, in_dir)
  call get_command_argument(2, arg); read (arg, *) z_in
  call get_command_argument(3, out_dir)
  call get_command_argument(4, arg); read (arg, *) z_out
  call get_command_argument(5, arg); read (arg, *) nsteps
  call get_command_argument(6, arg); read (arg, *) np_nc_arg
  call get_command_argument(7, arg); read (arg, *) pp_arg
  call get_command_argument(8, tables)

  call read_tables(trim(tables), wf, wc)

  ! -- particle_initialization ------------------------------------------
  open (newunit=u, file=fname(in_dir, z_in, 'info'), access='stream', status='old', form='unformatted')
  read (u) hi, hr
  close (u)
  nplocal = hi(1); npglobal = hi(2); nnt = hi(11); ncb = hi(14)
  nc = hi(12) * nnt
  n3 = nc ** 3
  a = hr(1); t = hr(2); tau = hr(3)
  dt_pp = hr(4); dt_fine = hr(5); dt_coarse = hr(6); dt_vmax = hr(7)
  mass_p = hr(9); om = hr(13); ol = hr(14); sigma_vi = hr(18); vz_max = hr(22)
  allocate (rhoc(n3), vfield(3 * n3))
  if (hi(5) == 1) then
    allocate (xp1(3 * nplocal)); call read_i1(fname(in_dir, z_in, 'xp'), xp1); pxp = c_loc(xp1)
  else
    allocate (xp2(3 * nplocal)); call read_i2(fname(in_dir, z_in, 'xp'), xp2); pxp = c_loc(xp2)
  end if
  if (hi(6) == 1) then
    allocate (vp1(3 * nplocal)); call read_i1(fname(in_dir, z_in, 'vp'), vp1); pvp = c_loc(vp1)
  else
    allocate (vp2(3 * nplocal)); call read_i2(fname(in_dir, z_in, 'vp'), vp2); pvp = c_loc(vp2)
  end if
  open (newunit=u, file=fname(in_dir, z_in, 'np'), access='stream', status='old', form='unformatted')
  read (u) rhoc
  close (u)
  open (newunit=u, file=fname(in_dir, z_in, 'vc'), access='stream', status='old', form='unformatted')
  read (u) vfield
  close (u)
  open (newunit=u, file=fname(in_dir, z_in, 'id'), access='stream', status='old', form='unformatted', iostat=ios)
  has_pid = ios == 0
  ppid = c_null_ptr
  if (has_pid) then
    allocate (pid(nplocal))
    read (u) pid
    close (u)
    ppid = c_loc(pid)
  end if

  p%nc = nc; p%nnt = nnt; p%ncb = ncb; p%np_nc = np_nc_arg
  p%izipx = int(hi(5), 4); p%izipv = int(hi(6), 4)
  p%has_pid = merge(1, 0, has_pid); p%pp_on = pp_arg; p%pp_range = 1
  p%rsoft = 0.3; p%image_buffer = 1.5; p%tile_buffer = 2.5
  p%device = 0; p%mem_layout = 0
  call cube_check(cube_ctx_create(p, ctx), 'cube_ctx_create')
  call cube_check(cube_init_kernels(ctx, wf, wc), 'cube_init_kernels')
  call cube_check(cube_set_state(ctx, nplocal, npglobal, pxp, pvp, ppid, c_loc(rhoc), c_loc(vfield)), &
                  'cube_set_state')
  call cube_check(cube_buffer_grid(ctx), 'buffer_grid')
  call cube_check(cube_buffer_x(ctx), 'buffer_x')
  call cube_check(cube_buffer_v(ctx), 'buffer_v')

  ! -- the step loop (W/main.f90:102-118) --------------------------------
  gg = 1.0 / 6.0 / 3.14159265358979
  dt = 0.0
  do istep = 1, nsteps
    call timestep()
    st = cube_update_x(ctx, dt_old, dt, vz_max, sigma_vi, vs)
    call cube_check(st, 'update_x')
    call cube_check(cube_buffer_grid(ctx), 'buffer_grid')
    call cube_check(cube_buffer_x(ctx), 'buffer_x')
    if (pp_arg /= 0) then
      call cube_check(cube_pp_force(ctx, a_mid, dt, sigma_vi, mass_p, f2pp), 'ext_pp_force')
      dt_pp = 1.0 / (sqrt(f2pp) * a_mid * gg)
    end if
    call cube_check(cube_particle_mesh(ctx, a_mid, dt, sigma_vi, vs%sigma_vi_new, mass_p, pm), 'particle_mesh')
    sigma_vi = vs%sigma_vi_new
    call cube_check(cube_buffer_v(ctx), 'buffer_v')
    dt_fine = pm%dt_fine; dt_coarse = pm%dt_coarse; dt_vmax = pm%dt_vmax; vz_max = pm%vz_max
    write (*, '(a,i6,5(1x,z8.8),1x,i12)') 'step', istep, transfer(a, 0), transfer(dt, 0), &
      transfer(dt_old, 0), transfer(a_mid, 0), transfer(sigma_vi, 0), vs%nplocal
  end do

  ! -- checkpoint ---------------------------------------------------------
  call cube_check(cube_get_state(ctx, nplocal, pxp, pvp, ppid, c_loc(rhoc), c_loc(vfield)), 'cube_get_state')
  hi(1) = nplocal; hi(15) = hi(15) + nsteps
  hr(1) = a; hr(2) = t; hr(3) = tau; hr(4) = dt_pp; hr(5) = dt_fine; hr(6) = dt_coarse; hr(7) = dt_vmax
  hr(18) = sigma_vi; hr(22) = vz_max
  call execute_command_line('mkdir -p '//trim(out_dir)//'/image1')
  open (newunit=u, file=fname(out_dir, z_out, 'info'), access='stream', status='replace', form='unformatted')
  write (u) hi, hr
  close (u)
  open (newunit=u, file=fname(out_dir, z_out, 'xp'), access='stream', status='replace', form='unformatted')
  if (hi(5) == 1) then
    write (u) xp1(1:3 * nplocal)
  else
    write (u) xp2(1:3 * nplocal)
  end if
  close (u)
  open (newunit=u, file=fname(out_dir, z_out, 'vp'), access='stream', status='replace', form='unformatted')
  if (hi(6) == 1) then
    write (u) vp1(1:3 * nplocal)
  else
    write (u) vp2(1:3 * nplocal)
  end if
  close (u)
  open (newunit=u, file=fname(out_dir, z_out, 'np'), access='stream', status='replace', form='unformatted')
  write (u) rhoc
  close (u)
  open (newunit=u, file=fname(out_dir, z_out, 'vc'), access='stream', status='replace', form='unformatted')
  write (u) vfield
  close (u)
  if (has_pid) then
    open (newunit=u, file=fname(out_dir, z_out, 'id'), access='stream', status='replace', form='unformatted')
    write (u) pid(1:nplocal)
    close (u)
  end if
  call cube_ctx_destroy(ctx)

contains

  ! <dir>/image1/<z>_<kind>_1.bin, z written as (f7.3) and left-adjusted
  function fname(dir, z, kind) result(s)
    character(*), intent(in) :: dir, kind
    real(4), intent(in) :: z
    character(:), allocatable :: s
    character(7) :: zs
    write (zs, '(f7.3)') z
    s = trim(dir)//'/image1/'//trim(adjustl(zs))//'_'//kind//'_1.bin'
  end function

  subroutine read_i1(f, x)
    character(*), intent(in) :: f
    integer(1), intent(out) :: x(:)
    integer :: u
    open (newunit=u, file=f, access='stream', status='old', form='unformatted')
    read (u) x
    close (u)
  end subroutine

  subroutine read_i2(f, x)
    character(*), intent(in) :: f
    integer(2), intent(out) :: x(:)
    integer :: u
    open (newunit=u, file=f, access='stream', status='old', form='unformatted')
    read (u) x
    close (u)
  end subroutine

  ! the 16^3 fine and 4^3 coarse kernel tables, lines "i j k fx fy fz"
  ! checked for (i,j,k) order, as flat arrays with 3 components per entry
  subroutine read_tables(dir, wf, wc)
    character(*), intent(in) :: dir
    real(c_float), intent(out) :: wf(:), wc(:)
    integer :: u, i, j, k, ii, jj, kk, n
    real(4) :: f(3)
    open (newunit=u, file=dir//'/wfxyzf.3.ascii', status='old', action='read')
    n = 0
    do k = 1, 16
      do j = 1, 16
        do i = 1, 16
          read (u, *) ii, jj, kk, f
          if (ii /= i .or. jj /= j .or. kk /= k) then
            print *, 'wfxyzf.3.ascii: index mismatch at', i, j, k
            error stop 1
          end if
          wf(3 * n + 1:3 * n + 3) = f
          n = n + 1
        end do
      end do
    end do
    close (u)
    open (newunit=u, file=dir//'/wfxyzc.2.ascii', status='old', action='read')
    do n = 0, 63
      read (u, '(3i4,3e16.8)') ii, jj, kk, f
      wc(3 * n + 1:3 * n + 3) = f
    end do
    close (u)
  end subroutine

  ! W/timestep.f90:1-92 for one image, checkpoint landing at z_out
  subroutine timestep()
    real(4) :: dt_e, da1, da2, ra, a_next
    integer :: ntry, n
    dt_old = dt
    dt_e = 1.0
    ntry = 0
    do
      ntry = ntry + 1
      call expansion(a, dt_e, om, ol, -1.0, da1, da2)
      da = da1 + da2
      ra = da / (a + da)
      if (ra > 0.2) then
        dt_e = dt_e * (0.2 / ra)
      else
        exit
      end if
      if (ntry > 10) exit
    end do
    dt = min(dt_e, dt_fine, dt_coarse, dt_pp, dt_vmax)
    call expansion(a, dt, om, ol, -1.0, da1, da2)
    da = da1 + da2
    a_next = 1.0 / (1.0 + z_out)
    if (da >= a_next - a) then
      n = 0
      do while (abs((a + da) / a_next - 1) >= 1e-6 .and. n < 100)
        dt = dt * (a_next - a) / da
        call expansion(a, dt, om, ol, -1.0, da1, da2)
        da = da1 + da2
        n = n + 1
      end do
    end if
    a_mid = a + da / 2.0
    tau = tau + dt
    t = t + dt
    a = a + da
  end subroutine

end program cube_main
