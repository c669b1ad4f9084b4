! gturb_driver.f90 -- Fortran driver of the MI355X step library.
!
! It keeps the reference's run-directory conventions and control flow
! (src/main.f90): params.dat (src/inieuler.f90:14-20), curframe.dat
! (src/main.f90:37-39, 213-215), files/seed.NNN (src/main.f90:42-45,
! 207-210), restart velocity fields/u{x,y,z}.NNN as float32 unformatted
! records (src/read.f90:1-66, src/write.f90:1-54) and the diagnostics
! cadence imix (src/main.f90:134-138, 153-161).  The time step itself is the
! library's: gt_step runs the loop body of src/main.f90:124-150 on the GPU.
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
    integer(c_int) function gt_sync(ctx) bind(C, name="gt_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function gt_last_error() bind(C, name="gt_last_error")
      import :: c_ptr
    end function
  end interface

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
contains

  ! float32 physical field, one unformatted record (src/read.f90:10-22)
  subroutine read_field(path, nx, ny, nz, u)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nx, ny, nz
    real(c_double), intent(out) :: u(nx, ny, nz)
    real(4), allocatable :: w(:, :, :)
    integer :: lu, ios
    allocate(w(nx, ny, nz))
    open(newunit=lu, file=path, form='unformatted', action='read', status='old', iostat=ios)
    if (ios /= 0) then
      write(0, *) 'cannot open ', trim(path)
      stop 2
    end if
    read(lu, iostat=ios) w
    if (ios /= 0) then
      write(0, *) 'short or corrupt field file ', trim(path)
      stop 2
    end if
    close(lu)
    u = dble(w)
    deallocate(w)
  end subroutine read_field

  subroutine write_field(path, nx, ny, nz, u)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nx, ny, nz
    real(c_double), intent(in) :: u(nx, ny, nz)
    integer :: lu
    open(newunit=lu, file=path, form='unformatted', status='replace')
    write(lu) real(u, 4)
    close(lu)
  end subroutine write_field

end module driver_io

program gturb_mi355x
  use iso_c_binding
  use gturbo3d_api
  use driver_io
  implicit none

  type(gt_params) :: p
  type(c_ptr) :: ctx
  integer :: nx, ny, nz, nxh, npas, iout, imix, ifr, ipas, nchunk, nbin, lu, i, j, k, ib
  integer(c_int32_t) :: seed
  integer(c_int) :: nf
  real(c_double) :: t, inpe, inpz, tmp(9), fac, kk, k2, dkx, dky, dkz, scra, e1, e2
  real(c_double), allocatable :: ux(:, :, :), uy(:, :, :), uz(:, :, :), sp(:), sp2d(:)
  real(c_double), allocatable :: fkx(:), fky(:), fkz(:)
  complex(c_double_complex), allocatable :: bx(:, :, :), by(:, :, :), bz(:, :, :)
  character(len=64) :: arg, nome
  integer(8) :: c0, c1, crate
  real(c_double), parameter :: pi = 3.14159265358979323846d0

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
  allocate(bx(nxh, ny, nz), by(nxh, ny, nz), bz(nxh, ny, nz))
  allocate(sp(nbin), sp2d(nbin), fkx(nxh), fky(ny), fkz(nz))

  ! restart velocity -> b (src/main.f90:91-106)
  write(nome, "('./fields/ux.',i3.3)") ifr; call read_field(nome, nx, ny, nz, ux)
  write(nome, "('./fields/uy.',i3.3)") ifr; call read_field(nome, nx, ny, nz, uy)
  write(nome, "('./fields/uz.',i3.3)") ifr; call read_field(nome, nx, ny, nz, uz)
  call check(gt_init_velocity(ctx, ux, uy, uz), 'gt_init_velocity')

  ! wavenumbers for the host diagnostics (src/defk.f90:24-43)
  dkx = 2.0d0 * pi / p%xlx; dky = 2.0d0 * pi / p%xly; dkz = 2.0d0 * pi / p%xlz
  do i = 1, nxh
    fkx(i) = dble(i - 1) * dkx
  end do
  do j = 1, ny
    fky(j) = dble(merge(j - 1, j - ny - 1, j <= ny / 2 + 1)) * dky
  end do
  do k = 1, nz
    fkz(k) = dble(merge(k - 1, k - nz - 1, k <= nz / 2 + 1)) * dkz
  end do

  write(nome, "('./files/globalE.',i3.3)") ifr
  open(unit=18, file=nome, form='formatted', status='unknown', position='append')
  write(nome, "('./files/spectra.',i3.3)") ifr
  open(unit=20, file=nome, form='formatted', status='unknown', position='append')

  ipas = 0
  do while (ipas < npas)
    nchunk = min(imix - mod(ipas, imix), npas - ipas)
    call check(gt_step(ctx, nchunk, seed), 'gt_step')
    ipas = ipas + nchunk
    t = t + dble(nchunk) * p%dt
    if (mod(ipas, imix) == 0) then
      call check(gt_get_state(ctx, bx, by, bz), 'gt_get_state')
      ! energy spectrum sum fac k^2 |b|^2 (src/diagnostics.f90:22-35) and glob energy (:131-153)
      sp = 0.d0; sp2d = 0.d0; tmp = 0.d0
      do k = 1, nz
        do j = 1, ny
          do i = 1, nxh
            fac = 1.d0
            if (i == 1 .or. i == nxh) fac = 0.5d0
            k2 = fkx(i)**2 + fky(j)**2 + fkz(k)**2
            kk = sqrt(k2)
            scra = abs(bx(i, j, k))**2 + abs(by(i, j, k))**2 + abs(bz(i, j, k))**2
            ib = int(kk) + 1
            if (ib > 0 .and. ib < nbin) sp(ib) = sp(ib) + kk * kk * fac * scra
            if (k == 1 .and. ib > 0 .and. ib < nbin) sp2d(ib) = sp2d(ib) + kk * kk * fac * abs(bz(i, j, k))**2
            if (k2 > 0.d0) then
              tmp(1) = tmp(1) + fac * k2 * scra
              tmp(4) = tmp(4) + 2.d0 * p%nu * fac * k2 * scra * k2**p%alpnu
              tmp(7) = tmp(7) + 2.d0 * p%mu * fac * k2 * scra * k2**p%alpmu
            end if
          end do
        end do
      end do
      e1 = tmp(1)
      e2 = tmp(4) + tmp(7)
      write(18, '(4es24.16)') t, e1, e2, inpe
      call flush(18)
      do ib = 1, nbin
        write(20, '(3es24.16)') dble(ib - 1), sp(ib), sp2d(ib)
      end do
      call flush(20)
      write(6, '(a,es14.6,a,es12.4,a,es14.6)') ' t = ', t, ', Eb = ', e2 / inpe, ', E = ', e1
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
  close(20)
  call check(gt_destroy(ctx), 'gt_destroy')

  call system_clock(c1)
  write(6, '(a,es14.6)') ' Mean time per iteration (s) = ', dble(c1 - c0) / dble(crate) / dble(max(npas, 1))
end program gturb_mi355x
