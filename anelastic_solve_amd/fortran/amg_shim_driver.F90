! Test driver of the Fortran drop-in: calls mg_read_params / MG_init /
! mg_solver / mg_end through implicit interfaces, as the reference's test
! program does (mgMK/test/test.F90:67,431,451), on a host-owned grid.
!
!   amg_shim_driver layout
!       prints the byte sizes of the bind(C) types (checked against C).
!   amg_shim_driver params <file.ini>
!       reads an INI file through mg_read_params and prints the controls.
!   amg_shim_driver dirichlet <x1|x3> <fp64|mixed> <work.ini>
!       the closed-form case 'dirichletBCs' of mgMK/test/test.F90:150-189:
!       del^2 p = 2 on a uniform 32^3 unit box, periodic in the other two
!       directions, p = 1 and p = 2 at the ghost-cell centres of the two
!       Dirichlet faces, so p = z^2 + b1 z + c1 exactly for the 3-point
!       stencil.  Prints PASS/FAIL with the max error; exit status 0 on PASS.
!   amg_shim_driver relmax <fp64|mixed> <work.ini>
!       the same case (x3) with the reference's own stop rule only: the INI
!       has tolmax and no tolres, so mg_solver stops on the relative
!       solution change (mgMK/mg_solver.F90:127).  PASS when it converges
!       before maxcy with 0 < relmax <= tolmax and the closed form holds.
!   amg_shim_driver rho3d <work.ini>
!       MG_init with a background density r that varies horizontally:
!       prints the status (AMG_EUNSUPPORTED); needs no GPU.
program amg_shim_driver
  use iso_c_binding
  use amg_cabi
  use amg_mg_state
  use ntypes
  use Domain
  use Grid
  use Flow
  implicit none
  character(len=256) :: mode, a2, a3, a4
  integer :: stat

  call get_command_argument(1, mode)
  call get_command_argument(2, a2)
  call get_command_argument(3, a3)
  call get_command_argument(4, a4)
  select case (trim(mode))
  case ('layout')
     call layout()
  case ('params')
     call mg_read_params(trim(a2), stat)
     call print_params(stat)
  case ('dirichlet')
     call dirichlet(trim(a2), trim(a3), trim(a4), .false.)
  case ('relmax')
     call dirichlet('x3', trim(a2), trim(a3), .true.)
  case ('rho3d')
     call rho3d(trim(a2))
  case default
     write(0, '(a)') 'usage: amg_shim_driver layout | params <ini> | dirichlet <x1|x3> <fp64|mixed> <ini>'
     write(0, '(a)') '     | relmax <fp64|mixed> <ini> | rho3d <ini>'
     error stop 2
  end select

contains

  subroutine layout()
    type(amg_config) :: c
    type(amg_grid) :: g
    type(amg_stats) :: s
    type(amg_flowbc) :: b
    integer :: TB(3, 2, 12)
    real(c_double) :: VB(3, 2, 12)
    integer :: d, f, v
    print '(a,4(1x,i0))', 'LAYOUT', c_sizeof(c), c_sizeof(g), c_sizeof(s), c_sizeof(b)
    ! TB(d,f,v) = 100*d + 10*f + v lands at flat index (d-1) + 3*(f-1) + 6*(v-1)
    do v = 1, 12
       do f = 1, 2
          do d = 1, 3
             TB(d, f, v) = 100 * d + 10 * f + v
             VB(d, f, v) = -TB(d, f, v)
          end do
       end do
    end do
    b = amg_flowbc_from(TB, VB)
    print '(a,3(1x,i0),1x,f8.1)', 'FLOWBC', b%tb(1), b%tb(2 + 3 * 1 + 6 * 4), b%tb(72), b%vb(72)
  end subroutine layout

  subroutine print_params(st)
    integer, intent(in) :: st
    print '(a,1x,i0)', 'STAT', st
    if (st /= 0) return
    print '(a,5(1x,i0))', 'CTRL', mg_cfg%smoother, mg_cfg%maxcy, mg_cfg%stop_rule, &
         mg_cfg%precision, mg_cfg%strict
    print '(a,3(1x,i0))', 'LEVELS', mg_cfg%nlevels
    print '(a,es24.16)', 'TOL', mg_cfg%tol
    print '(a,6(1x,i0))', 'TBC', mg_cfg%tbc
    print '(a,6es24.16)', 'VBC', mg_cfg%vbc
    print '(a,16(1x,i0))', 'PRE', mg_cfg%pre_relax
    print '(a,16(1x,i0))', 'POST', mg_cfg%post_relax
  end subroutine print_params

  ! Uniform ghost-inclusive 1-D metrics on [0, 1] with the loadGrid
  ! conventions (utils/loadGrid.F90:79-90): index 1 and n+2 are ghosts,
  ! xc(i) = (xe(i)+xe(i-1))/2, dxc(i) = xe(i)-xe(i-1), dxe(i) = xc(i+1)-xc(i).
  subroutine uniform(n, e, c, de, dc, rde, rdc)
    integer, intent(in) :: n
    real(r8), allocatable, intent(out) :: e(:), c(:), de(:), dc(:), rde(:), rdc(:)
    integer :: i
    real(r8) :: h
    h = 1.0_r8 / n
    allocate(e(n + 2), c(n + 2), de(n + 2), dc(n + 2), rde(n + 2), rdc(n + 2))
    do i = 1, n + 2
       e(i) = (i - 1) * h
       c(i) = (i - 1.5_r8) * h
    end do
    de = h; dc = h; rde = 1.0_r8 / h; rdc = 1.0_r8 / h
  end subroutine uniform

  subroutine rho3d(ini)
    character(len=*), intent(in) :: ini
    integer, parameter :: n = 16
    integer :: i, st, u
    nx = n; ny = n; nz = n
    nxp2 = n + 2; nyp2 = n + 2; nzp2 = n + 2
    sx = 2; ex = n + 1; sy = 2; ey = n + 1; sz = 2; ez = n + 1
    xL = 1.0_r8; yL = 1.0_r8; zL = 1.0_r8
    call uniform(n, xe, xc, dxe, dxc, rdxe, rdxc)
    call uniform(n, ye, yc, dye, dyc, rdye, rdyc)
    call uniform(n, ze, zc, dze, dzc, rdze, rdzc)
    allocate(r(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1))
    do i = sx - 1, ex + 1
       r(i, :, :) = 1.0_r8 + 0.01_r8 * xc(i)
    end do
    open(newunit=u, file=ini, status='replace', action='write')
    write(u, '(a)') '[MGsolver]'
    write(u, '(a)') 'Smoother=LJ'
    write(u, '(a)') 'nxlevels=3'
    write(u, '(a)') 'nylevels=3'
    write(u, '(a)') 'nzlevels=3'
    write(u, '(a)') 'maxcy=20'
    write(u, '(a)') 'tolmax=1.0d-6'
    write(u, '(a)') 'PreRelax=2,2,2'
    write(u, '(a)') 'PstRelax=2,2,2'
    write(u, '(a)') '[BCs]'
    write(u, '(a)') 'X1=periodic'
    write(u, '(a)') 'X2=periodic'
    write(u, '(a)') 'X3=read'
    write(u, '(a)') '[X3min]'
    write(u, '(a)') 'PTYPE=Neumann'
    write(u, '(a)') 'PVALUE=0.d0'
    write(u, '(a)') '[X3max]'
    write(u, '(a)') 'PTYPE=Neumann'
    write(u, '(a)') 'PVALUE=0.d0'
    close(u)
    call mg_read_params(ini, st)
    if (st /= 0) error stop 3
    call MG_init(st)
    print '(a,1x,i0)', 'STAT', st
  end subroutine rho3d

  subroutine dirichlet(dir, prec, ini, relmax_only)
    character(len=*), intent(in) :: dir, prec, ini
    logical, intent(in) :: relmax_only
    integer, parameter :: n = 32
    real(r8), parameter :: a1 = 2.0_r8, va = 1.0_r8, vb = 2.0_r8
    real(r8), allocatable, target :: phi(:,:,:), src(:,:,:)
    real(r8) :: b1, c1, zlo, zhi, x, ex_p, err, pmax, relmax
    integer :: i, j, k, iter, st, u

    nx = n; ny = n; nz = n
    nxp2 = n + 2; nyp2 = n + 2; nzp2 = n + 2
    sx = 2; ex = n + 1; sy = 2; ey = n + 1; sz = 2; ez = n + 1
    xL = 1.0_r8; yL = 1.0_r8; zL = 1.0_r8
    call uniform(n, xe, xc, dxe, dxc, rdxe, rdxc)
    call uniform(n, ye, yc, dye, dyc, rdye, rdyc)
    call uniform(n, ze, zc, dze, dzc, rdze, rdzc)
    if (dir == 'x3') then
       ! a background density that is identically 1 exercises the column copy
       allocate(r(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1))
       r = 1.0_r8
    end if

    open(newunit=u, file=ini, status='replace', action='write')
    write(u, '(a)') '[MGsolver]'
    write(u, '(a)') 'Smoother=LJ'
    write(u, '(a)') 'Strict=yes'
    write(u, '(a)') 'Verbose=no'
    write(u, '(a)') 'nxlevels=4'
    write(u, '(a)') 'nylevels=4'
    write(u, '(a)') 'nzlevels=4'
    write(u, '(a)') 'maxcy=60'
    if (relmax_only) then
       write(u, '(a)') 'tolmax=1.0d-12'
    else
       write(u, '(a)') 'tolmax=1.0d-13'
       write(u, '(a)') 'tolres=1.0d-12'
    end if
    write(u, '(a)') 'Precision='//prec
    write(u, '(a)') 'PreRelax=2,2,2,2'
    write(u, '(a)') 'PstRelax=2,2,2,2'
    write(u, '(a)') '[BCs]'
    if (dir == 'x1') then
       write(u, '(a)') 'X1=read'
       write(u, '(a)') 'X2=periodic'
       write(u, '(a)') 'X3=periodic'
       write(u, '(a)') '[X1min]'
    else
       write(u, '(a)') 'X1=periodic'
       write(u, '(a)') 'X2=periodic'
       write(u, '(a)') 'X3=read'
       write(u, '(a)') '[X3min]'
    end if
    write(u, '(a)') 'PTYPE=Dirichlet'
    write(u, '(a)') 'PVALUE=1.d0'
    if (dir == 'x1') then
       write(u, '(a)') '[X1max]'
    else
       write(u, '(a)') '[X3max]'
    end if
    write(u, '(a)') 'PTYPE=Dirichlet'
    write(u, '(a)') 'PVALUE=2.d0'
    close(u)

    call mg_read_params(ini, st)
    if (st /= 0) error stop 3
    call MG_init(st)
    if (st /= 0) then
       print '(a,1x,i0)', 'MG_init failed', st
       error stop 4
    end if

    allocate(phi(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1), src(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1))
    src = a1
    phi = 0.5_r8 * (va + vb)
    call mg_solver(phi, src, iter, relmax)

    ! p(s) = a1/2 s^2 + b1 s + c1 with p = va, vb at the two ghost centres
    if (dir == 'x1') then
       zlo = xc(1); zhi = xc(nxp2)
    else
       zlo = zc(1); zhi = zc(nzp2)
    end if
    b1 = (va - vb) / (zlo - zhi) - 0.5_r8 * a1 * (zlo + zhi)
    c1 = 0.5_r8 * (va + vb) - 0.25_r8 * a1 * (zlo**2 + zhi**2) - 0.5_r8 * b1 * (zlo + zhi)
    err = 0.0_r8; pmax = 0.0_r8
    do k = sz, ez
       do j = sy, ey
          do i = sx, ex
             if (dir == 'x1') then
                x = xc(i)
             else
                x = zc(k)
             end if
             ex_p = 0.5_r8 * a1 * x * x + b1 * x + c1
             err = max(err, abs(phi(i, j, k) - ex_p))
             pmax = max(pmax, abs(ex_p))
          end do
       end do
    end do
    ! ghosts: periodic wrap in the other directions, the wall value at the
    ! Dirichlet ghost centres (mgMK/mg_bdry.F90:109-114,124-129)
    if (dir == 'x1') then
       err = max(err, maxval(abs(phi(sx-1, sy:ey, sz:ez) - va)), maxval(abs(phi(ex+1, sy:ey, sz:ez) - vb)))
       err = max(err, maxval(abs(phi(sx:ex, sy-1, sz:ez) - phi(sx:ex, ey, sz:ez))))
    else
       err = max(err, maxval(abs(phi(sx:ex, sy:ey, sz-1) - va)), maxval(abs(phi(sx:ex, sy:ey, ez+1) - vb)))
       err = max(err, maxval(abs(phi(sx-1, sy:ey, sz:ez) - phi(ex, sy:ey, sz:ez))))
    end if
    call mg_end(st)
    if (relmax_only .and. .not. (relmax > 0.0_r8 .and. relmax <= 1.0e-12_r8 .and. iter < 60)) then
       print '(a,1x,a,1x,a,1x,i0,2es12.4)', 'FAIL', dir, prec, iter, err / pmax, relmax
       error stop 1
    end if
    if (err / pmax <= 1.0e-9_r8 .and. iter >= 1 .and. iter <= 60 .and. st == 0) then
       print '(a,1x,a,1x,a,1x,i0,2es12.4)', 'PASS', dir, prec, iter, err / pmax, relmax
    else
       print '(a,1x,a,1x,a,1x,i0,2es12.4)', 'FAIL', dir, prec, iter, err / pmax, relmax
       error stop 1
    end if
  end subroutine dirichlet

end program amg_shim_driver
