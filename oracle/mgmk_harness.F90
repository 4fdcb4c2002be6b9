! mgmk_harness.F90 -- test infrastructure: a driver of the reference's own
! multigrid solver (mgMK, compiled from its Fortran sources by
! oracle/Makefile.ref into oracle/_ref/) on a uniform periodic-box grid.
! It is never part of the product: only tests/ (through fixtures written
! by tests/golden/make_mgmk_golden.py) and bench.py's CPU baseline run it.
!
! What it sets up, as the reference's test program does
! (mgMK/test/test.F90:47-142,431-457), with its own grid instead of grid
! files:
!   * MPI and the reference's 3-D decomposition (decomp/mpi_setup.f90) over
!     px*py*pz ranks;
!   * the module Grid of include/modules.F90:17-39 for uniform spacing
!     dx = L/n on [0, L] with the loadGrid conventions (index 1 and n+2 are
!     ghosts); xL = L, so the uniform spacing L/(n_pts-1) that mgdpfpde
!     takes (mgMK/mg_pfpde.F90:35-40, n_pts = n+1 from mgMK/mg_init.F90:
!     125-129) equals dx;
!   * Flow::r = 1 (the Boussinesq case; the reference's ptest never
!     allocates r, mgMK/mg_solver.F90:81-96 reads it);
!   * mg_read_params / MG_init / mg_solver exactly as the DNS calls them.
!
! Usage (every rank takes the same arguments):
!   mgmk_harness <ini> <nx> <ny> <nz> <px> <py> <pz> <Lx> <Ly> <Lz>
!                <rhs> <out> <mode> <count> <workdir>
!   rhs   : file of (nx+2)*(ny+2)*(nz+2) doubles, x fastest, ghosts included
!           (the Fortran array rhsf(sx-1:ex+1, ...) of the whole domain)
!   out   : file for the solution in the same layout, or '-'
!   mode  : 'solve' -- one mg_solver call with the INI's maxcy/tolmax;
!           'trace' -- <count> calls of one V-cycle each (maxcy = 1),
!                      warm-started, printing the relative residual after
!                      each (the residual rule of the GPU solver);
!           'time'  -- <count> timed calls of mg_solver from a zero guess,
!                      each a fixed number of V-cycles (maxcy from the INI,
!                      tolmax = 0 so it never stops early), timed with
!                      MPI_WTIME around the call as mgMK/test/test.F90:
!                      443-457 does.
!   workdir: where mgMK writes its per-level grid files (mg_write_grid).
! Rank 0 prints 'KEY value ...' lines on stdout.
program mgmk_harness
  use ntypes, only: i4, r8
  use mgVARS, only: maxcy, tolmax
  use Domain
  use Grid
  use Flow, only: r
  use IO, only: IOUT, mgDIR
  use dd
  implicit none
  character(len=512) :: ini, rhsfile, outfile, mode, workdir, arg
  integer(i4) :: np(3), n3(3), count, ok, iter, c, i, j, k, ierr
  real(r8) :: Ls(3), relmax, relres, t0, t, tmin, tsum, bmax
  real(r8), allocatable :: phi(:,:,:), src(:,:,:), work(:,:,:)

  call MPI_INIT(ierr)
  call MPI_COMM_RANK(commx1x2x3, myid, ierr)
  call MPI_COMM_SIZE(commx1x2x3, sizex1x2x3, ierr)
  numprocs = sizex1x2x3

  call get_command_argument(1, ini)
  do i = 1, 3
     call get_command_argument(1 + i, arg); read(arg, *) n3(i)
     call get_command_argument(4 + i, arg); read(arg, *) np(i)
     call get_command_argument(7 + i, arg); read(arg, *) Ls(i)
  end do
  call get_command_argument(11, rhsfile)
  call get_command_argument(12, outfile)
  call get_command_argument(13, mode)
  call get_command_argument(14, arg); read(arg, *) count
  call get_command_argument(15, workdir)

  IOUT = 0
  mgDIR = workdir
  nx = n3(1); ny = n3(2); nz = n3(3)
  nxp2 = nx + 2; nyp2 = ny + 2; nzp2 = nz + 2
  nxprocs = np(1); nyprocs = np(2); nzprocs = np(3)
  call uniform(nx, Ls(1), xe, xc, dxe, dxc, rdxe, rdxc); xL = Ls(1)
  call uniform(ny, Ls(2), ye, yc, dye, dyc, rdye, rdyc); yL = Ls(2)
  call uniform(nz, Ls(3), ze, zc, dze, dzc, rdze, rdzc); zL = Ls(3)

  call mg_read_params(trim(ini), ok)
  if (ok /= 0) call die('mg_read_params failed')
  call mpi_setup(ok)
  if (ok /= 0) call die('mpi_setup failed')
  allocate(r(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1))
  r = 1.0_r8
  call MG_init(ok)
  if (ok /= 0) call die('MG_init failed')

  allocate(phi(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1), src(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1))
  allocate(work(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1))
  call read_local(trim(rhsfile), src)
  bmax = gmax(maxval(abs(src(sx:ex, sy:ey, sz:ez))))

  select case (trim(mode))
  case ('solve')
     phi = 0.0_r8
     work = src
     call mg_solver(phi, work, iter, relmax)
     relres = residual(phi) / bmax
     if (myid == 0) print '(a,1x,i0,2es24.16)', 'SOLVE', iter, relmax, relres
  case ('trace')
     maxcy = 1
     tolmax = 0.0_r8
     phi = 0.0_r8
     do c = 1, count
        work = src
        call mg_solver(phi, work, iter, relmax)
        relres = residual(phi) / bmax
        if (myid == 0) print '(a,1x,i0,2es24.16)', 'CYCLE', c, relmax, relres
     end do
  case ('time')
     tolmax = 0.0_r8
     tmin = huge(1.0_r8)
     tsum = 0.0_r8
     do c = 1, count
        phi = 0.0_r8
        work = src
        call MPI_BARRIER(commx1x2x3, ierr)
        t0 = MPI_WTIME()
        call mg_solver(phi, work, iter, relmax)
        call MPI_BARRIER(commx1x2x3, ierr)
        t = MPI_WTIME() - t0
        tmin = min(tmin, t)
        tsum = tsum + t
        if (myid == 0) print '(a,1x,i0,es16.8)', 'TIME', c, t
     end do
     relres = residual(phi) / bmax
     if (myid == 0) print '(a,1x,i0,1x,i0,3es16.8,1x,i0)', 'TIMED', count, maxcy, tmin, &
          tsum / count, relres, sizex1x2x3
  case default
     call die('unknown mode')
  end select
  if (trim(outfile) /= '-') call write_global(trim(outfile), phi)
  call MPI_FINALIZE(ierr)

contains

  subroutine die(msg)
    character(len=*), intent(in) :: msg
    integer(i4) :: e
    write(0, '(a)') 'mgmk_harness: '//msg
    call MPI_ABORT(commx1x2x3, 1, e)
    stop 1
  end subroutine die

  ! Uniform ghost-inclusive metrics on [0, L] with the loadGrid conventions
  ! (utils/loadGrid.F90:79-90): xe = upper face, xc = centre, dxc = width,
  ! dxe = centre distance.
  subroutine uniform(n, L, e, cc, de, dc, rde, rdc)
    integer(i4), intent(in) :: n
    real(r8), intent(in) :: L
    real(r8), allocatable, intent(out) :: e(:), cc(:), de(:), dc(:), rde(:), rdc(:)
    integer(i4) :: q
    real(r8) :: h
    h = L / n
    allocate(e(n + 2), cc(n + 2), de(n + 2), dc(n + 2), rde(n + 2), rdc(n + 2))
    do q = 1, n + 2
       e(q) = (q - 1) * h
       cc(q) = (q - 1.5_r8) * h
    end do
    de = h; dc = h; rde = 1.0_r8 / h; rdc = 1.0_r8 / h
  end subroutine uniform

  real(r8) function gmax(v)
    real(r8), intent(in) :: v
    integer(i4) :: e
    call MPI_ALLREDUCE(v, gmax, 1, MPI_DOUBLE_PRECISION, MPI_MAX, commx1x2x3, e)
  end function gmax

  ! max |rhs - A p| over the interior, A the fine operator of mgdpfpde
  ! (mgMK/mg_pfpde.F90:35-57) with r = 1, on p with the ghosts mg_solver
  ! left in it (mgMK/mg_solver.F90:139).
  real(r8) function residual(p)
    real(r8), intent(in) :: p(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1)
    real(r8) :: cx, cy, cz, s, m
    integer(i4) :: ii, jj, kk
    cx = 1.0_r8 / (xL / nx)**2
    cy = 1.0_r8 / (yL / ny)**2
    cz = 1.0_r8 / (zL / nz)**2
    m = 0.0_r8
    do kk = sz, ez
       do jj = sy, ey
          do ii = sx, ex
             s = cx * (p(ii - 1, jj, kk) + p(ii + 1, jj, kk)) + cy * (p(ii, jj - 1, kk) + &
                 p(ii, jj + 1, kk)) + cz * (p(ii, jj, kk - 1) + p(ii, jj, kk + 1)) - &
                 2.0_r8 * (cx + cy + cz) * p(ii, jj, kk)
             m = max(m, abs(src(ii, jj, kk) - s))
          end do
       end do
    end do
    residual = gmax(m)
  end function residual

  ! this rank's block (with ghosts) of a whole-domain stream file
  subroutine read_local(fn, a)
    character(len=*), intent(in) :: fn
    real(r8), intent(out) :: a(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1)
    real(r8), allocatable :: row(:)
    integer(i4) :: u, kk, jj
    integer(8) :: pos
    allocate(row(nxp2))
    open(newunit=u, file=fn, access='stream', form='unformatted', status='old', action='read')
    do kk = sz - 1, ez + 1
       do jj = sy - 1, ey + 1
          pos = ((int(kk - 1, 8) * nyp2 + (jj - 1)) * nxp2) * 8 + 1
          read(u, pos=pos) row
          a(:, jj, kk) = row(sx - 1:ex + 1)
       end do
    end do
    close(u)
  end subroutine read_local

  ! the whole-domain solution (interior from each rank, ghosts from the
  ! ranks that own them) written by rank 0
  subroutine write_global(fn, a)
    character(len=*), intent(in) :: fn
    real(r8), intent(in) :: a(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1)
    real(r8), allocatable :: g(:,:,:), gs(:,:,:)
    integer(i4) :: u, e, i0, i1, j0, j1, k0, k1
    allocate(g(nxp2, nyp2, nzp2), gs(nxp2, nyp2, nzp2))
    g = 0.0_r8
    i0 = sx; i1 = ex; j0 = sy; j1 = ey; k0 = sz; k1 = ez
    if (sx == 2) i0 = 1
    if (ex == nx + 1) i1 = nx + 2
    if (sy == 2) j0 = 1
    if (ey == ny + 1) j1 = ny + 2
    if (sz == 2) k0 = 1
    if (ez == nz + 1) k1 = nz + 2
    g(i0:i1, j0:j1, k0:k1) = a(i0:i1, j0:j1, k0:k1)
    call MPI_REDUCE(g, gs, nxp2 * nyp2 * nzp2, MPI_DOUBLE_PRECISION, MPI_SUM, 0, commx1x2x3, e)
    if (myid == 0) then
       open(newunit=u, file=fn, access='stream', form='unformatted', status='replace', action='write')
       write(u) gs
       close(u)
    end if
  end subroutine write_global

end program mgmk_harness
