! gturb_driver.f90 -- Fortran driver of the MI355X step library.
!
! It keeps the reference's run-directory conventions and control flow
! (src/main.f90): params.dat (src/inieuler.f90:14-20), curframe.dat
! (src/main.f90:37-39, 213-215), files/seed.NNN (src/main.f90:42-45,
! 207-210), restart velocity fields/u{x,y,z}.NNN as float32 unformatted
! records (src/read.f90:1-66, src/write.f90:1-54), and the time loop of
! src/main.f90:121-185 with its outputs:
!   every imix steps, of the state at time t before the step (:133-138):
!     files/fluxes.NNN  (src/diagnostics.f90:62-103)
!     files/spectra.NNN (src/diagnostics.f90:5-56)
!   every imix steps, after the step at t+dt (:152-161):
!     files/globalE.NNN, files/globalZ.NNN, 11 columns each
!     (src/diagnostics.f90:114-199), and the 't = , Eb = , E =' line
!   every iout steps except the last (:163-184): fields/u?.NNN.JJJ (writef2,
!     src/write.f90:56-109)
! The time step itself is the library's: gt_step runs the loop body of
! src/main.f90:124-150 on the GPU; the diagnostics are the library's device
! reductions (gt_fluxes, gt_spectra).  Numbers are written as es24.16 doubles
! (the reference's 'g' edit descriptor without a width is compiler-specific).
!
! A diagnostics step whose energy or dissipation sum is not finite ends the
! run: the library returns GT_ENONFINITE, the driver reports the step and
! time on stderr and stops with exit status 3, before writing anything of
! that step.
!
! Records of more than MAXSUB = 2147483639 bytes (a float32 field of 1024^3
! and above is 4 GiB, whose 4-byte record marker would wrap) are written and
! read as subrecords in gfortran's documented layout: every subrecord has a
! leading and a trailing 4-byte length, the leading one negative when
! another subrecord follows and the trailing one negative when one precedes.
! Smaller records are one plain record, byte for byte what the reference
! writes.  The environment variable GT_MAX_SUBRECORD lowers the limit (tests).
!
! The grid is a run-time argument (the reference fixes it at compile time in
! src/modules.f90:4):
!     gturb_mi355x NX NY NZ [variant=3|2] [dt2_mode=0|1] [precision=0|1]
! A fresh start (ifr = 0 in curframe.dat) needs fields/u*.000 written by the
! caller (the reference's init_noise uses an undefined seed).
module gturbo3d_api
  use iso_c_binding
  implicit none

  type, bind(C) :: gt_params
    integer(c_int) :: nx, ny, nz
    real(c_double) :: xlx, xly, xlz
    real(c_double) :: nu, alpnu, mu, alpmu
    real(c_double) :: dt, omega
    real(c_double) :: rtrunc
    integer(c_int) :: truncate
    real(c_double) :: famp, k1f, k2f
    integer(c_int) :: variant, dt2_mode, precision
  end type gt_params

  interface
    integer(c_int) function gt_create(p, ctx) bind(C, name="gt_create")
      import :: c_int, c_ptr, gt_params
      type(gt_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function gt_destroy(ctx) bind(C, name="gt_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function gt_init_velocity(ctx, ux, uy, uz) bind(C, name="gt_init_velocity")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ux(*), uy(*), uz(*)
    end function
    integer(c_int) function gt_step(ctx, n, seed) bind(C, name="gt_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      integer(c_int32_t), intent(inout) :: seed
    end function
    integer(c_int) function gt_get_state(ctx, bx, by, bz) bind(C, name="gt_get_state")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: bx(*), by(*), bz(*)
    end function
    integer(c_int) function gt_nonlinear(ctx, nx, ny, nz) bind(C, name="gt_nonlinear")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: nx(*), ny(*), nz(*)
    end function
    integer(c_int) function gt_velocity(ctx, ux, uy, uz) bind(C, name="gt_velocity")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ux(*), uy(*), uz(*)
    end function
    integer(c_int) function gt_forcing_table(ctx, nf, ikxf, ikyf, ikzf, ff, inpe, inpz, cap) &
        bind(C, name="gt_forcing_table")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: nf
      type(c_ptr), value :: ikxf, ikyf, ikzf, ff
      real(c_double), intent(out) :: inpe, inpz
      integer(c_int), value :: cap
    end function
    integer(c_int) function gt_spectra(ctx, sp, sp2d, hel, glob) bind(C, name="gt_spectra")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, sp, sp2d, hel, glob
    end function
    integer(c_int) function gt_fluxes(ctx, spf) bind(C, name="gt_fluxes")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: spf(*)
    end function
    integer(c_int) function gt_sync(ctx) bind(C, name="gt_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function gt_last_error() bind(C, name="gt_last_error")
      import :: c_ptr
    end function
  end interface

  integer(c_int), parameter :: GT_ENONFINITE = 6

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(gt_last_error(), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(0, '(a,a,i0,a,512a)') what, ' failed, code ', rc, ': ', msg(1:n)
    stop 1
  end subroutine check

end module gturbo3d_api

module driver_io
  use iso_c_binding
  implicit none
  integer(8), parameter :: MAXSUB = 2147483639_8  ! gfortran's largest subrecord (4-byte markers)
contains

  ! the subrecord limit in bytes: MAXSUB, or GT_MAX_SUBRECORD when it is smaller
  integer(8) function max_subrecord()
    character(len=32) :: v
    integer :: st
    integer(8) :: lim
    max_subrecord = MAXSUB
    call get_environment_variable('GT_MAX_SUBRECORD', v, status=st)
    if (st == 0) then
      read(v, *) lim
      if (lim >= 1 .and. lim < MAXSUB) max_subrecord = lim
    end if
  end function max_subrecord

  ! float32 physical field, one logical unformatted record (src/read.f90:10-22)
  subroutine read_field(path, nx, ny, nz, u)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nx, ny, nz
    real(c_double), intent(out) :: u(nx, ny, nz)
    real(4), allocatable, target :: w(:)
    integer(1), pointer :: wb(:)
    integer(8) :: nbytes, pos, slen
    integer(4) :: lead, trail
    integer :: lu, ios
    nbytes = 4_8 * int(nx, 8) * int(ny, 8) * int(nz, 8)
    allocate(w(nbytes / 4))
    call c_f_pointer(c_loc(w), wb, [nbytes])
    open(newunit=lu, file=path, access='stream', form='unformatted', action='read', status='old', iostat=ios)
    if (ios /= 0) then
      write(0, *) 'cannot open ', trim(path)
      stop 2
    end if
    pos = 0
    do
      read(lu, iostat=ios) lead
      if (ios /= 0) exit
      slen = abs(int(lead, 8))
      if (pos + slen > nbytes) then
        ios = 1
        exit
      end if
      if (slen > 0) read(lu, iostat=ios) wb(pos + 1:pos + slen)
      if (ios /= 0) exit
      read(lu, iostat=ios) trail
      if (ios /= 0 .or. abs(int(trail, 8)) /= slen) then
        ios = 1
        exit
      end if
      pos = pos + slen
      if (lead >= 0) exit
    end do
    close(lu)
    if (ios /= 0 .or. pos /= nbytes) then
      write(0, *) 'short or corrupt field file ', trim(path)
      stop 2
    end if
    u = reshape(dble(w), [nx, ny, nz])
    deallocate(w)
  end subroutine read_field

  ! float32 physical field as one logical record (src/write.f90:12-22), split
  ! into subrecords above max_subrecord() bytes
  subroutine write_field(path, nx, ny, nz, u)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nx, ny, nz
    real(c_double), intent(in) :: u(nx, ny, nz)
    real(4), allocatable, target :: w(:)
    integer(1), pointer :: wb(:)
    integer(8) :: nbytes, pos, slen, lim
    integer(4) :: lead, trail
    integer :: lu
    nbytes = 4_8 * int(nx, 8) * int(ny, 8) * int(nz, 8)
    allocate(w(nbytes / 4))
    w = reshape(real(u, 4), [nbytes / 4])
    call c_f_pointer(c_loc(w), wb, [nbytes])
    lim = max_subrecord()
    open(newunit=lu, file=path, access='stream', form='unformatted', status='replace')
    pos = 0
    do
      slen = min(lim, nbytes - pos)
      lead = int(slen, 4)
      trail = int(slen, 4)
      if (pos + slen < nbytes) lead = -lead  ! another subrecord follows
      if (pos > 0) trail = -trail            ! a subrecord precedes
      write(lu) lead
      if (slen > 0) write(lu) wb(pos + 1:pos + slen)
      write(lu) trail
      pos = pos + slen
      if (pos >= nbytes) exit
    end do
    close(lu)
    deallocate(w)
  end subroutine write_field

end module driver_io

program gturb_mi355x
  use iso_c_binding
  use gturbo3d_api
  use driver_io
  implicit none

  type(gt_params) :: p
  type(c_ptr) :: ctx
  integer :: nx, ny, nz, nxh, npas, iout, imix, ifr, jfr, ipas, nxt, nrun, nbin, lu, ib, rc
  integer(c_int32_t) :: seed
  integer(c_int) :: nf
  real(c_double) :: t, inpe, inpz
  real(c_double), allocatable :: ux(:, :, :), uy(:, :, :), uz(:, :, :)
  real(c_double), allocatable, target :: sp(:), sp2d(:), spf(:), glob(:)
  character(len=64) :: arg, nome
  integer(8) :: c0, c1, crate

  call system_clock(c0, crate)
  if (command_argument_count() < 3) then
    write(0, *) 'usage: gturb_mi355x NX NY NZ [variant=3|2] [dt2_mode=0|1] [precision=0|1]'
    stop 2
  end if
  call get_command_argument(1, arg); read(arg, *) nx
  call get_command_argument(2, arg); read(arg, *) ny
  call get_command_argument(3, arg); read(arg, *) nz
  p%variant = 3; p%dt2_mode = 0; p%precision = 0
  if (command_argument_count() >= 4) then
    call get_command_argument(4, arg); read(arg, *) p%variant
  end if
  if (command_argument_count() >= 5) then
    call get_command_argument(5, arg); read(arg, *) p%dt2_mode
  end if
  if (command_argument_count() >= 6) then
    call get_command_argument(6, arg); read(arg, *) p%precision
  end if
  p%nx = nx; p%ny = ny; p%nz = nz
  nxh = nx / 2 + 1
  nbin = nx / 2

  ! params.dat (src/inieuler.f90:14-20)
  open(newunit=lu, file='./params.dat', status='old', action='read')
  read(lu, *) p%xlx, p%xly, p%xlz
  read(lu, *) p%nu, p%alpnu, p%mu, p%alpmu
  read(lu, *) p%dt, p%omega
  read(lu, *) npas, iout, imix
  read(lu, *) p%rtrunc, p%truncate
  read(lu, *) p%famp, p%k1f, p%k2f
  close(lu)

  open(newunit=lu, file='./curframe.dat', status='old', action='read')
  read(lu, *) ifr, t
  close(lu)
  write(nome, "('./files/seed.',i3.3)") ifr
  open(newunit=lu, file=nome, status='old', action='read')
  read(lu, *) seed
  close(lu)

  call check(gt_create(p, ctx), 'gt_create')
  call check(gt_forcing_table(ctx, nf, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, inpe, inpz, 0), &
             'gt_forcing_table')
  write(6, '(a,3i6)') ' 3D-rotating Navier-Stokes, MI355X step, grid ', nx, ny, nz
  write(6, '(a,es12.4,a,es12.4,a,es12.4)') ' omega ', p%omega, ' nu ', p%nu, ' dt ', p%dt
  write(6, '(a,i0,a,es14.6)') ' forced modes ', nf, ', mean energy input ', inpe

  allocate(ux(nx, ny, nz), uy(nx, ny, nz), uz(nx, ny, nz))
  allocate(sp(nbin), sp2d(nbin), spf(nbin), glob(18))

  ! restart velocity -> b (src/main.f90:91-106)
  write(nome, "('./fields/ux.',i3.3)") ifr; call read_field(nome, nx, ny, nz, ux)
  write(nome, "('./fields/uy.',i3.3)") ifr; call read_field(nome, nx, ny, nz, uy)
  write(nome, "('./fields/uz.',i3.3)") ifr; call read_field(nome, nx, ny, nz, uz)
  call check(gt_init_velocity(ctx, ux, uy, uz), 'gt_init_velocity')

  ! output files (src/main.f90:50-61)
  write(nome, "('./files/globalE.',i3.3)") ifr
  open(unit=18, file=nome, form='formatted', status='unknown', position='append')
  write(nome, "('./files/globalZ.',i3.3)") ifr
  open(unit=19, file=nome, form='formatted', status='unknown', position='append')
  write(nome, "('./files/spectra.',i3.3)") ifr
  open(unit=20, file=nome, form='formatted', status='unknown', position='append')
  write(nome, "('./files/fluxes.',i3.3)") ifr
  open(unit=30, file=nome, form='formatted', status='unknown', position='append')

  ! time loop (src/main.f90:121-185): the steps between two output steps run
  ! as one gt_step call; an output step runs alone, with its diagnostics of
  ! time t before it and of t+dt after it
  jfr = 1
  ipas = 0
  do while (ipas < npas)
    nxt = npas
    if (imix > 0) nxt = min(nxt, (ipas / imix + 1) * imix)
    if (iout > 0) nxt = min(nxt, (ipas / iout + 1) * iout)
    nrun = nxt - 1 - ipas
    if (nrun > 0) then
      call check(gt_step(ctx, nrun, seed), 'gt_step')
      ipas = ipas + nrun
      t = t + dble(nrun) * p%dt
    end if
    ipas = ipas + 1
    if (is_every(ipas, imix)) then
      ! spectrum and flux at time t, before the step (src/main.f90:133-138)
      rc = gt_spectra(ctx, c_loc(sp), c_loc(sp2d), c_null_ptr, c_null_ptr)
      if (rc == GT_ENONFINITE) call nonfinite(ipas, t)
      call check(rc, 'gt_spectra')
      call check(gt_fluxes(ctx, spf), 'gt_fluxes')
      do ib = 1, nbin
        write(30, '(i8,es24.16)') ib, spf(ib)
      end do
      call flush(30)
      do ib = 1, nbin
        write(20, '(i8,2es24.16)') ib - 1, sp(ib), sp2d(ib)
      end do
      call flush(20)
    end if
    call check(gt_step(ctx, 1, seed), 'gt_step')
    t = t + p%dt
    if (is_every(ipas, imix)) then
      ! global quantities at t+dt (src/main.f90:152-161, src/diagnostics.f90:126-196)
      rc = gt_spectra(ctx, c_null_ptr, c_null_ptr, c_null_ptr, c_loc(glob))
      if (rc == GT_ENONFINITE) call nonfinite(ipas, t)
      call check(rc, 'gt_spectra')
      write(18, '(11es24.16)') t, glob(1:9), inpe
      call flush(18)
      write(6, '(a,es14.6,a,es14.6,a,es14.6)') ' t = ', t, ', Eb = ', sum(glob(4:9)) / inpe, &
        ', E = ', sum(glob(1:3))
      call flush(6)
      write(19, '(11es24.16)') t, glob(10:18), inpz
      call flush(19)
    end if
    if (is_every(ipas, iout) .and. ipas /= npas) then
      ! velocity snapshot (src/main.f90:163-184, writef2)
      call check(gt_velocity(ctx, ux, uy, uz), 'gt_velocity')
      write(nome, "('./fields/ux.',i3.3,'.',i3.3)") ifr, jfr; call write_field(nome, nx, ny, nz, ux)
      write(nome, "('./fields/uy.',i3.3,'.',i3.3)") ifr, jfr; call write_field(nome, nx, ny, nz, uy)
      write(nome, "('./fields/uz.',i3.3,'.',i3.3)") ifr, jfr; call write_field(nome, nx, ny, nz, uz)
      jfr = jfr + 1
      write(6, '(a,i0)') ' Saved v field at timestep = ', ipas
      call flush(6)
    end if
  end do

  ! final velocity, seed and frame (src/main.f90:188-215)
  ifr = ifr + 1
  call check(gt_velocity(ctx, ux, uy, uz), 'gt_velocity')
  write(nome, "('./fields/ux.',i3.3)") ifr; call write_field(nome, nx, ny, nz, ux)
  write(nome, "('./fields/uy.',i3.3)") ifr; call write_field(nome, nx, ny, nz, uy)
  write(nome, "('./fields/uz.',i3.3)") ifr; call write_field(nome, nx, ny, nz, uz)
  write(nome, "('./files/seed.',i3.3)") ifr
  open(newunit=lu, file=nome, status='replace')
  write(lu, *) seed
  close(lu)
  open(newunit=lu, file='./curframe.dat', status='replace')
  write(lu, *) ifr, t
  close(lu)
  close(18)
  close(19)
  close(20)
  close(30)
  call check(gt_destroy(ctx), 'gt_destroy')

  call system_clock(c1)
  write(6, '(a,es14.6)') ' Mean time per iteration (s) = ', dble(c1 - c0) / dble(crate) / dble(max(npas, 1))

contains

  logical function is_every(n, every)
    integer, intent(in) :: n, every
    is_every = every > 0
    if (is_every) is_every = mod(n, every) == 0
  end function is_every

  subroutine nonfinite(n, tt)
    integer, intent(in) :: n
    real(c_double), intent(in) :: tt
    write(0, '(a,i0,a,es14.6,a)') ' non-finite state at step ', n, ' (t = ', tt, '): run stopped'
    call flush(0)
    stop 3
  end subroutine nonfinite
end program gturb_mi355x
