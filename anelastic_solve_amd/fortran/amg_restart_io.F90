! Restart dumps from Fortran, for the round-trip checks of the host reader and
! writer (anelastic_solve_amd/restart.py): the same unformatted sequential
! record sequence as the reference's serial restart writer and reader
! (dns/Restart.F90:738-755 write_flow_small, :124-140 Greadfile).
!
!   amg_restart_io write <file> nx ny nz
!       writes a dump of ghosted extents (nx+2, ny+2, nz+2) filled with
!       exactly representable values: u(i,j,k) = i + 1000 j + 1000000 k + 0.25,
!       v = -u, w = u/2, p = u + 0.125, rho = 2u, scal1 = u - 7.5, header
!       (123, 4.5, 0.015625, 9.75, 1.25, -0.5, 1000, 0.71, 1.5)
!   amg_restart_io copy <in> <out> nx ny nz
!       reads <in> with the reader's size check (exit status 3 on a size
!       mismatch) and writes everything it read to <out>
program amg_restart_io
  implicit none
  integer, parameter :: r8 = 8, i4 = 4
  character(len=512) :: mode, f1, f2, arg
  integer(i4) :: nx, ny, nz, nxp2, nyp2, nzp2, n_r(3), nstep, i, j, k, s1, na
  real(r8) :: time, delt, g, rho_0, scal1_0, Re, Pr, Sc
  real(r8), allocatable :: u(:,:,:), v(:,:,:), w(:,:,:), p(:,:,:), rho(:,:,:), scal1(:,:,:)

  call get_command_argument(1, mode)
  na = 0
  if (trim(mode) == 'write') then
    call get_command_argument(2, f1)
    na = 3
  else if (trim(mode) == 'copy') then
    call get_command_argument(2, f1)
    call get_command_argument(3, f2)
    na = 4
  else
    write(0, '(a)') 'usage: amg_restart_io write <file> nx ny nz | copy <in> <out> nx ny nz'
    stop 2
  end if
  call get_command_argument(na, arg); read(arg, *) nx
  call get_command_argument(na + 1, arg); read(arg, *) ny
  call get_command_argument(na + 2, arg); read(arg, *) nz
  nxp2 = nx + 2; nyp2 = ny + 2; nzp2 = nz + 2
  allocate(u(nxp2, nyp2, nzp2), v(nxp2, nyp2, nzp2), w(nxp2, nyp2, nzp2), &
           p(nxp2, nyp2, nzp2), rho(nxp2, nyp2, nzp2), scal1(nxp2, nyp2, nzp2))

  if (trim(mode) == 'write') then
    do k = 1, nzp2
      do j = 1, nyp2
        do i = 1, nxp2
          u(i, j, k) = real(i, r8) + 1000.0_r8 * j + 1000000.0_r8 * k + 0.25_r8
        end do
      end do
    end do
    v = -u
    w = 0.5_r8 * u
    p = u + 0.125_r8
    rho = 2.0_r8 * u
    scal1 = u - 7.5_r8
    nstep = 123
    time = 4.5_r8; delt = 0.015625_r8; g = 9.75_r8; rho_0 = 1.25_r8; scal1_0 = -0.5_r8
    Re = 1000.0_r8; Pr = 0.71_r8; Sc = 1.5_r8
    call put(f1)
  else
    open(310, file=f1, form='unformatted', status='old', iostat=s1)
    if (s1 /= 0) then
      write(0, '(a)') 'cannot open '//trim(f1)
      stop 2
    end if
    read(310) nstep, time, delt, g, rho_0, scal1_0, Re, Pr, Sc
    read(310) n_r
    if (n_r(1) /= nxp2 .or. n_r(2) /= nyp2 .or. n_r(3) /= nzp2) then
      close(310)
      write(0, '(a,3i6)') 'data file size is', n_r
      stop 3
    end if
    read(310) u
    read(310) v
    read(310) w
    read(310) p
    read(310) rho
    read(310) scal1
    close(310)
    call put(f2)
  end if

contains

  subroutine put(fname)
    character(len=*), intent(in) :: fname
    integer :: s
    open(310, file=fname, form='unformatted', status='unknown', iostat=s)
    if (s /= 0) then
      write(0, '(a)') 'cannot create '//trim(fname)
      stop 2
    end if
    write(310) nstep, time, delt, g, rho_0, scal1_0, Re, Pr, Sc
    write(310) nxp2, nyp2, nzp2
    write(310) u
    write(310) v
    write(310) w
    write(310) p
    write(310) rho
    write(310) scal1
    close(310)
  end subroutine put

end program amg_restart_io
