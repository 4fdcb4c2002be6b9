! Drop-in replacements for the reference's multigrid entry points.
!
! The host program calls these four external subroutines with implicit
! interfaces, exactly as it calls the reference's libmg.a
! (Makefile:56,124-145,205-206):
!
!   mg_read_params(infile, stat)          mgMK/mg_read_params.F90:1-218
!   MG_init(stat)                         mgMK/mg_init.F90:1-351
!   mg_solver(phif, rhsf, iter, relmax)   mgMK/mg_solver.F90:1-141
!   mg_end(ierr)                          mgMK/mg_end.F90:1-41
!
! Like the reference they read the host's module state: Domain (sizes and
! sweep ranges), Grid (1-D metrics) and Flow (background density r), all
! as declared in include/modules.F90:9-55.  In an integrated build they
! are compiled against the host's own modules; the test build uses the
! stand-alone replica in amg_host_modules.F90.
!
! One process owns the whole domain (sx=2, ex=nx+1, ... as the serial
! reference sets them, dns/Sturb.F90:99-104).  r must be horizontally
! uniform (the reference's set_densityBG makes it a function of z,
! dns/IC.F90:1085-1092); MG_init returns AMG_EUNSUPPORTED otherwise.
module amg_mg_state
  use iso_c_binding
  use amg_cabi
  implicit none
  type(amg_config), save :: mg_cfg
  type(c_ptr), save :: mg_handle = c_null_ptr
  logical, save :: mg_have_params = .false.
  real(c_double), allocatable, target, save :: mg_rho(:)
end module amg_mg_state


subroutine mg_read_params(infile, stat)
  use iso_c_binding
  use amg_cabi
  use amg_mg_state
  implicit none
  character(len=*), intent(in) :: infile
  integer, intent(out) :: stat
  character(kind=c_char, len=len_trim(infile) + 1) :: cname

  cname = trim(infile)//c_null_char
  call amg_config_default(mg_cfg)
  stat = int(amg_read_params(cname, mg_cfg))
  mg_have_params = (stat == AMG_OK)
  if (stat /= AMG_OK) write(0, '(a)') 'mg_read_params: '//trim(amg_message(int(stat, c_int)))
end subroutine mg_read_params


subroutine MG_init(stat)
  use iso_c_binding
  use amg_cabi
  use amg_mg_state
  use Domain, only: nx, ny, nz, sx, ex, sy, ey, sz, ez
  use Grid, only: xe, xc, dxe, dxc, ye, yc, dye, dyc, ze, zc, dze, dzc
  use Flow, only: r
  implicit none
  integer, intent(out) :: stat
  type(amg_grid) :: g
  integer :: k
  real(c_double) :: dev

  if (.not. mg_have_params) then
     write(0, '(a)') 'MG_init: mg_read_params has not succeeded'
     stat = AMG_EINVAL
     return
  end if
  if (ex - sx + 1 /= nx .or. ey - sy + 1 /= ny .or. ez - sz + 1 /= nz) then
     write(0, '(a)') 'MG_init: the solver owns the whole domain (one process per GPU)'
     stat = AMG_EUNSUPPORTED
     return
  end if
  if (c_associated(mg_handle)) stat = int(amg_destroy(mg_handle))
  mg_handle = c_null_ptr

  g%nx = nx; g%ny = ny; g%nz = nz
  g%xe = c_loc(xe(1)); g%xc = c_loc(xc(1)); g%dxe = c_loc(dxe(1)); g%dxc = c_loc(dxc(1))
  g%ye = c_loc(ye(1)); g%yc = c_loc(yc(1)); g%dye = c_loc(dye(1)); g%dyc = c_loc(dyc(1))
  g%ze = c_loc(ze(1)); g%zc = c_loc(zc(1)); g%dze = c_loc(dze(1)); g%dzc = c_loc(dzc(1))
  g%rho = c_null_ptr
  if (allocated(r)) then
     if (allocated(mg_rho)) deallocate(mg_rho)
     allocate(mg_rho(nz + 2))
     dev = 0.0_c_double
     do k = sz - 1, ez + 1
        mg_rho(k - sz + 2) = r(sx, sy, k)
        dev = max(dev, maxval(abs(r(sx:ex, sy:ey, k) - r(sx, sy, k))))
     end do
     if (dev > 1.0e-12_c_double * maxval(abs(mg_rho))) then
        write(0, '(a)') 'MG_init: background density r varies horizontally'
        stat = AMG_EUNSUPPORTED
        return
     end if
     g%rho = c_loc(mg_rho(1))
  end if
  stat = int(amg_create(mg_cfg, g, mg_handle))
  if (stat /= AMG_OK) write(0, '(a)') 'MG_init: '//trim(amg_message(int(stat, c_int)))
end subroutine MG_init


! phif: initial guess in, solution out with filled ghost layers.  rhsf is
! left intact (the reference overwrites it with the fine residual; its
! callers never read it afterwards).  iter = V-cycles run (maxcy+1 when the
! tolerance is not reached, after a WARNING line as in the reference);
! relmax = relative solution change of the last cycle.
subroutine mg_solver(phif, rhsf, iter, relmax)
  use iso_c_binding
  use amg_cabi
  use amg_mg_state
  use Domain, only: sx, ex, sy, ey, sz, ez
  implicit none
  real(c_double), intent(inout), target :: phif(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1)
  real(c_double), intent(inout), target :: rhsf(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1)
  integer, intent(out) :: iter
  real(c_double), intent(out) :: relmax
  integer(c_int) :: st, it

  if (.not. c_associated(mg_handle)) then
     write(0, '(a)') 'mg_solver: MG_init has not succeeded'
     error stop 1
  end if
  st = amg_solve_host(mg_handle, c_loc(phif), c_loc(rhsf), it, relmax, c_null_ptr)
  iter = int(it)
  if (st /= AMG_OK .and. st /= AMG_ENOTCONV) then
     write(0, '(a)') 'mg_solver: '//trim(amg_message(st))
     error stop 1
  end if
end subroutine mg_solver


subroutine mg_end(ierr)
  use iso_c_binding
  use amg_cabi
  use amg_mg_state
  implicit none
  integer, intent(out) :: ierr
  ierr = 0
  if (c_associated(mg_handle)) ierr = int(amg_destroy(mg_handle))
  mg_handle = c_null_ptr
  if (allocated(mg_rho)) deallocate(mg_rho)
end subroutine mg_end
