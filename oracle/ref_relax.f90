! Project-written harness for the reference's lattice relaxation (test
! infrastructure only).  Compiled by oracle/ref_build.sh against the
! reference's own Setup and Geometry modules, it builds the cell as
! Main.f90:119-154 does, applies LatticeRelaxation (Geometry.f90:279-1135)
! where Main.f90:156-160 does, and writes the coordinates before and after
! (raw little-endian doubles, ndim x 3, column-major) into the current
! directory: Coords.bin, CoordsRelaxed.bin.
program ref_relax
  use Setup
  use Geometry
  implicit none
  real(dp) :: aMoire, cs, sn, t1(2), t2(2), g1(2), g12(2), Rot(2,2)
  real(dp), allocatable :: Coords(:,:)
  integer(dp) :: n
  integer :: fu

  aMoire = 3.0_dp*ntheta**2 + 3.0_dp*ntheta + 1.0_dp
  g1  = (4.0_dp*pi/3.0_dp)/aMoire*(real(3*ntheta+1,dp)*a1+a2)
  g12 = (4.0_dp*pi/3.0_dp)/aMoire*(real(3*ntheta+2,dp)*a2-a1)
  cs = 1.0_dp-1.0_dp/(2.0_dp*aMoire)
  sn = sqrt(1.0_dp-cs**2)
  Rot = reshape([cos(0.5_dp*acos(cs)),-sin(0.5_dp*acos(cs)),sin(0.5_dp*acos(cs)),cos(0.5_dp*acos(cs))],[2,2])
  g1  = matmul(Rot,g1)
  g12 = matmul(Rot,g12)
  t1 = real( ntheta  ,dp)*a1 + real(  ntheta+1,dp)*a2
  t2 = real(-ntheta-1,dp)*a1 + real(2*ntheta+1,dp)*a2
  allocate(Coords(ndim,3))
  call WignerSeitzCell(Coords, t1, t2, cs, sn)
  do n=1,ndim
    Coords(n,1:2) = matmul(Rot, Coords(n,1:2))
  end do
  open(newunit=fu, file='Coords.bin', access='stream', form='unformatted', status='replace')
  write(fu) Coords
  close(fu)
  call LatticeRelaxation(Coords, g1, g12)
  open(newunit=fu, file='CoordsRelaxed.bin', access='stream', form='unformatted', status='replace')
  write(fu) Coords
  close(fu)
end program ref_relax
