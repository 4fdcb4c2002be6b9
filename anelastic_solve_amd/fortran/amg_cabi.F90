! Fortran 2003 interoperable view of the solver's C ABI (include/amg.h).
!
! The derived types below have the memory layout of the C structs of the
! same names (bind(C) follows the companion C compiler's layout rules);
! the interfaces bind the extern "C" entry points by name.  Status codes
! are the AMG_* values of amg.h.
module amg_cabi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: AMG_OK = 0, AMG_EINVAL = -1, AMG_EBADSIZE = -2, &
       AMG_EBADBC = -3, AMG_ENOMEM = -4, AMG_ENOTCONV = -5, AMG_ENAN = -6, AMG_EHIP = -7, &
       AMG_EUNSUPPORTED = -8, AMG_EPARSE = -9, AMG_ENOGPU = -10, AMG_EBUSY = -11
  integer(c_int), parameter :: AMG_BC_PERIODIC = 0, AMG_BC_NEUMANN = 1, AMG_BC_DIRICHLET = 2
  integer(c_int), parameter :: AMG_SMOOTHER_RBPGS = 0, AMG_SMOOTHER_LJ = 2
  integer(c_int), parameter :: AMG_FP64 = 0, AMG_MIXED = 1
  integer(c_int), parameter :: AMG_STOP_RESIDUAL = 0, AMG_STOP_RELMAX = 1
  integer, parameter :: AMG_MAX_LEVELS = 16

  type, bind(C) :: amg_config
     integer(c_int) :: nlevels(3)
     integer(c_int) :: smoother
     integer(c_int) :: maxcy
     real(c_double) :: tol
     integer(c_int) :: stop_rule
     integer(c_int) :: pre_relax(AMG_MAX_LEVELS)
     integer(c_int) :: post_relax(AMG_MAX_LEVELS)
     integer(c_int) :: tbc(6)
     real(c_double) :: vbc(6)
     integer(c_int) :: precision
     integer(c_int) :: report_relmax
     integer(c_int) :: verbose
     integer(c_int) :: strict
     integer(c_int) :: use_graph
     integer(c_int) :: coarse_direct
     integer(c_int) :: pair_cycles
  end type amg_config

  type, bind(C) :: amg_grid
     integer(c_int) :: nx, ny, nz
     type(c_ptr) :: xe = c_null_ptr, xc = c_null_ptr, dxe = c_null_ptr, dxc = c_null_ptr
     type(c_ptr) :: ye = c_null_ptr, yc = c_null_ptr, dye = c_null_ptr, dyc = c_null_ptr
     type(c_ptr) :: ze = c_null_ptr, zc = c_null_ptr, dze = c_null_ptr, dzc = c_null_ptr
     type(c_ptr) :: rho = c_null_ptr
  end type amg_grid

  type, bind(C) :: amg_stats
     integer(c_int) :: iters
     integer(c_int) :: converged
     real(c_double) :: relres, relmax, res0, bnorm, ms_total, ms_cycles, bytes_cycles
     integer(c_int) :: nlevels
     integer(c_int) :: level_n(3, AMG_MAX_LEVELS)
  end type amg_stats

  ! TB(1:3,1:2,1:12) and VB(1:3,1:2,1:12) of module boundC, flattened in
  ! Fortran order (amg_flowbc_from copies them in).
  type, bind(C) :: amg_flowbc
     integer(c_int) :: tb(72)
     real(c_double) :: vb(72)
  end type amg_flowbc

  interface
     subroutine amg_config_default(cfg) bind(C, name="amg_config_default")
       import :: amg_config
       type(amg_config), intent(out) :: cfg
     end subroutine amg_config_default

     integer(c_int) function amg_read_params(infile, cfg) bind(C, name="amg_read_params")
       import :: c_int, c_char, amg_config
       character(kind=c_char), dimension(*), intent(in) :: infile
       type(amg_config), intent(inout) :: cfg
     end function amg_read_params

     integer(c_int) function amg_create(cfg, grid, handle) bind(C, name="amg_create")
       import :: c_int, c_ptr, amg_config, amg_grid
       type(amg_config), intent(in) :: cfg
       type(amg_grid), intent(in) :: grid
       type(c_ptr), intent(out) :: handle
     end function amg_create

     integer(c_int) function amg_destroy(handle) bind(C, name="amg_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: handle
     end function amg_destroy

     integer(c_int) function amg_solve_host(handle, phif, rhsf, iter, relmax, stats) &
          bind(C, name="amg_solve_host")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: handle, phif, rhsf, stats
       integer(c_int), intent(out) :: iter
       real(c_double), intent(out) :: relmax
     end function amg_solve_host

     integer(c_int) function amg_solve_device(handle, d_phif, d_rhsf, iter, relmax, stats) &
          bind(C, name="amg_solve_device")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: handle, d_phif, d_rhsf, stats
       integer(c_int), intent(out) :: iter
       real(c_double), intent(out) :: relmax
     end function amg_solve_device

     integer(c_int) function amg_psource_device(handle, u, v, w, src, dt) &
          bind(C, name="amg_psource_device")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: handle, u, v, w, src
       real(c_double), value :: dt
     end function amg_psource_device

     integer(c_int) function amg_pgrad_device(handle, p, u, v, w, dt) &
          bind(C, name="amg_pgrad_device")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: handle, p, u, v, w
       real(c_double), value :: dt
     end function amg_pgrad_device

     integer(c_int) function amg_divergence_device(handle, u, v, w, div, avg, maxabs, loc) &
          bind(C, name="amg_divergence_device")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: handle, u, v, w, div
       real(c_double), intent(out) :: avg, maxabs
       integer(c_int), intent(out) :: loc(3)
     end function amg_divergence_device

     integer(c_int) function amg_ghost_device(handle, var, vtype, bc) &
          bind(C, name="amg_ghost_device")
       import :: c_int, c_ptr, c_char, amg_flowbc
       type(c_ptr), value :: handle, var
       character(kind=c_char), intent(in) :: vtype(*)
       type(amg_flowbc), intent(in) :: bc
     end function amg_ghost_device

     integer(c_int) function amg_pgrad_ghost_device(handle, p, u, v, w, dt, bc) &
          bind(C, name="amg_pgrad_ghost_device")
       import :: c_int, c_ptr, c_double, amg_flowbc
       type(c_ptr), value :: handle, p, u, v, w
       real(c_double), value :: dt
       type(amg_flowbc), intent(in) :: bc
     end function amg_pgrad_ghost_device

     integer(c_int) function amg_cn_advance_device(handle, veltemp1, veltemp2, f1, f2, f3, rkdt, &
          var, p, bc, rRe, rPr, rSc) bind(C, name="amg_cn_advance_device")
       import :: c_int, c_ptr, c_double, amg_flowbc
       type(c_ptr), value :: handle, veltemp1, veltemp2, p
       real(c_double), value :: f1, f2, f3, rkdt
       integer(c_int), value :: var
       type(amg_flowbc), intent(in) :: bc
       real(c_double), value :: rRe, rPr, rSc
     end function amg_cn_advance_device

     type(c_ptr) function amg_strerror(code) bind(C, name="amg_strerror")
       import :: c_ptr, c_int
       integer(c_int), value :: code
     end function amg_strerror

     type(c_ptr) function amg_last_error() bind(C, name="amg_last_error")
       import :: c_ptr
     end function amg_last_error

     integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  ! Copies a NUL-terminated C string into a Fortran string (truncated to len(out)).
  subroutine amg_c_string(p, out)
    type(c_ptr), intent(in) :: p
    character(len=*), intent(out) :: out
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    out = ' '
    if (.not. c_associated(p)) return
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    do i = 1, min(n, len(out))
       out(i:i) = chars(i)
    end do
  end subroutine amg_c_string

  ! The flow BC tables of module boundC as the C struct amg_flowbc.
  function amg_flowbc_from(TB, VB) result(bc)
    integer, intent(in) :: TB(3, 2, 12)
    real(c_double), intent(in) :: VB(3, 2, 12)
    type(amg_flowbc) :: bc
    bc%tb = int(reshape(TB, [72]), c_int)
    bc%vb = reshape(VB, [72])
  end function amg_flowbc_from

  ! "<strerror(code)>: <last error message>" for diagnostics.
  function amg_message(code) result(msg)
    integer(c_int), intent(in) :: code
    character(len=400) :: msg
    character(len=80) :: name
    character(len=300) :: detail
    call amg_c_string(amg_strerror(code), name)
    call amg_c_string(amg_last_error(), detail)
    msg = trim(name)//': '//trim(detail)
  end function amg_message

end module amg_cabi
