! Stand-alone stand-in for the host program's modules, for building and
! testing the drop-in shim without the host.  Only the names the shim reads
! are declared, with the kinds and array shapes of include/modules.F90:2-55
! (ntypes, Domain, Grid, Flow).  An integrated build compiles the shim
! against the host's own modules instead and does not use this file.
module ntypes
  integer, parameter :: r8 = 8
  integer, parameter :: i4 = 4
end module ntypes

module Domain
  use ntypes
  integer(i4) :: nx, ny, nz, nxp2, nyp2, nzp2
  integer(i4) :: sx, ex, sy, ey, sz, ez
end module Domain

module Grid
  use ntypes
  real(r8) :: xL, yL, zL
  real(r8), allocatable, dimension(:), target :: xe, xc, dxe, dxc, rdxe, rdxc
  real(r8), allocatable, dimension(:), target :: ye, yc, dye, dyc, rdye, rdyc
  real(r8), allocatable, dimension(:), target :: ze, zc, dze, dzc, rdze, rdzc
end module Grid

module Flow
  use ntypes
  real(r8), allocatable, dimension(:,:,:) :: r
end module Flow
