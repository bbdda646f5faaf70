!-----------------------------------------------------------------------
! host_modules.f90 -- the module state a host program of the solvers and
! p' stencils shares with them, restated for the drop-in test host.
!
! The drop-ins (cappuccino-ghostcells-mpi_amd/fortran/cap_dropin.f90) read
! and write the host program's modules by name, so a host program must
! provide modules with these names, variables and kinds.  The reference
! provides them in modules_allocatable.f90 (types:1-13, parameters:15-109,
! indexes:111-132, geometry:147-159, coef:161-170, coefb:172-180,
! variables:274-298, title_mod:312-322, gradients:389-398, bc:208-244,
! inlet:380-387) and setind.f90:3-44; this file restates
! only the entities the drop-ins and the test driver use, so that the test
! host is built from this repository's sources alone.
!-----------------------------------------------------------------------
module types
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  integer, parameter :: prec = dp
end module types

module parameters
  use types
  implicit none
  integer, parameter :: nphi = 10
  integer, parameter :: iu = 1, iv = 2, iw = 3, ip = 4, ite = 5, ied = 6, ien = 7, ivis = 8, &
                        ivart = 9, icon = 10
  ! single-precision literals widened to prec, as the reference declares them
  real(prec), parameter :: small = 1e-30
  real(prec), parameter :: great = 1e+30
  integer :: nxyz, nxyza
  real(prec) :: alfa
  logical :: ltest
  integer :: icorr, ipcorr
  integer, dimension(nphi) :: nsw
  real(prec), dimension(nphi) :: sor, resor
  ! convection schemes and deferred-correction blending of fluxuvw
  logical :: lcds, lluds, lquds, lsmart, lavl, lmuscl, lumist, lgamma, lcds4
  logical :: flux_limiter
  real(dp), dimension(nphi) :: gds
end module parameters

module indexes
  use types
  use parameters
  implicit none
  integer :: ni, nj, nk
  integer :: nim, njm, nkm
  integer :: nimm, njmm, nkmm
  integer :: nij, nik, njk, nijk
  integer, dimension(:), allocatable :: li, lk
end module indexes

module geometry
  use types
  implicit none
  real(prec), dimension(:), allocatable :: x, y, z
  real(prec), dimension(:), allocatable :: fx, fy, fz
  real(prec), dimension(:), allocatable :: xc, yc, zc, vol
  real(prec), dimension(:), allocatable :: ar1x, ar1y, ar1z
  real(prec), dimension(:), allocatable :: ar2x, ar2y, ar2z
  real(prec), dimension(:), allocatable :: ar3x, ar3y, ar3z
end module geometry

module coef
  use types
  implicit none
  real(prec), dimension(:), allocatable :: ae, aw, an, as, at, ab, ap
  real(prec), dimension(:), allocatable :: sp, res, spv
  real(prec), dimension(:), allocatable :: su, sv, sw
  real(prec), dimension(:), allocatable :: apu, apv, apw
end module coef

module coefb
  use types
  implicit none
  real(prec), dimension(:), allocatable :: bb, bs, bw, bp, bn, be, bt
end module coefb

module variables
  use types
  implicit none
  real(prec), dimension(:), allocatable :: u, v, w
  real(prec), dimension(:), allocatable :: f1, f2, f3
  real(prec), dimension(:), allocatable :: p, pp, te, ed, t
  real(prec), dimension(:), allocatable :: vis, den
end module variables

module gradients
  use types
  implicit none
  ! gradu(3, nxyza): the component is the fastest index
  real(prec), dimension(:, :), allocatable :: gradu, gradv, gradw, gradp
end module gradients

module title_mod
  use types
  use parameters
  implicit none
  character(len=7), dimension(10) :: chvarSolver = (/'U      ', 'V      ', 'W      ', &
      'p      ', 'k      ', 'epsilon', 'Energy ', 'Visc   ', 'VarTemp', 'Conc   '/)
end module title_mod

module bc
  implicit none
  ! boundary type of each boundary face cell: 1 inlet, 2 outlet, 3 symmetry,
  ! 4 wall; lbw / lbe of nj*nk, lbs / lbn of ni*nk, lbb / lbt of ni*nj cells
  integer, dimension(:), allocatable :: lbw, lbe
  integer, dimension(:), allocatable :: lbs, lbn
  integer, dimension(:), allocatable :: lbb, lbt
end module bc

module inlet
  use types
  implicit none
  real(prec) :: uin, vin, win, tein, edin, tin
end module inlet

! Index helpers of the structured grid: the derived extents and the plane
! and row offsets li(i) = (i-1) nj, lk(k) = (k-1) ni nj of the padded
! j-fastest layout (the reference's setind.f90:3-44).
subroutine setind
  use types
  use parameters
  use indexes
  implicit none
  integer :: i, k
  nim = ni - 1; njm = nj - 1; nkm = nk - 1
  nimm = nim - 1; njmm = njm - 1; nkmm = nkm - 1
  nij = ni*nj; nik = ni*nk; njk = nj*nk; nijk = nij*nk
  do i = 1, ni
    li(i) = (i - 1)*nj
  end do
  do k = 1, nk
    lk(k) = (k - 1)*nij
  end do
end subroutine setind
