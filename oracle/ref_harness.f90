! Project-written harness around the reference FORGE modules (test
! infrastructure only).  It is compiled by oracle/ref_build.sh against the
! reference's own Setup/Geometry/TightBinding/HartreeFock/lapack_routines
! modules and writes raw little-endian arrays (stream access) plus a text
! manifest into the current directory.  tests/golden/make_fixtures.py turns
! them into the committed fixtures.
!
! What each stage calls in the reference:
!   geometry     NumberNeighborCells/OrderNeighborCells (Geometry.f90:14-91),
!                SampleBZ (Geometry.f90:96-127), WignerSeitzCell
!                (Geometry.f90:132-273); g1/g12/rotation follow Main.f90:126-154
!   LongRange    LongRangeInteraction (HartreeFock.f90:14-53)
!   H_TB(k)      TightBindingHamiltonian (TightBinding.f90:13-55)
!   first diag   Solve (HartreeFock.f90:402-483), SortEnergies_1Spin (:885-925),
!                Fermi block as Main.f90:255-267, GetFock (:487-560)
!   energies     GetKineticEnergy/GetFockEnergy/GetHartreeHubbardEnergy
!   potential    as Main.f90:319-327
!   H_HF(k)      HamiltonianHartreeFock (HartreeFock.f90:57-127)
!   eigen        diagonalize -> zdiagRed (LapackRoutines.f90:381-433)
!   iteration 1  Solve, SortEnergies_1Spin, Fermi (Main.f90:400-412), GetFock,
!                OptimalStep (HartreeFock.f90:1148-1227), mixing (Main.f90:558)
program ref_harness
  use omp_lib
  use Setup
  use Geometry
  use TightBinding
  use HartreeFock
  use lapack_routines
  implicit none

  integer(dp) :: Nk, nnc, i, n, m, icount, nF, nOcc, nPart, kk, ivk1, ivk2, nb
  integer(dp), allocatable :: nUC1(:), nUC2(:), nMomComp(:,:), nMomFlat(:,:), nSortedMom(:,:)
  real(dp) :: aMoire, cs, sn, t1(2), t2(2), t3(2), g1(2), g12(2), Rot(2,2), tn(2,3), vk(2)
  real(dp) :: FermiE, DegFactor, Step, Ekin(1), EF(1), EH, EU, Etot, scal(8)
  real(dp), allocatable :: Coords(:,:), MomVal(:,:,:), LongRange(:,:), Potential(:), Bands(:,:)
  real(dp), allocatable :: Sorted(:), Density(:,:), DensityIn(:,:), DensitySub(:), Energies(:)
  complex(dp), allocatable :: zH(:,:), zFockBulk(:,:,:), zFock(:,:,:), zFockIn(:,:,:), zEig(:,:,:), zSorted(:,:)
  complex(dp), allocatable :: zRhoNew(:,:,:,:), zRhoIn(:,:,:,:)
  integer :: mu
  integer(dp) :: ksel(2)

  Nk = numk*numk
  nb = numb
  open(newunit=mu, file='manifest.txt', status='replace')

  ! ---------------- geometry
  call NumberNeighborCells(numI, nnc)
  allocate(nUC1(nnc), nUC2(nnc))
  call OrderNeighborCells(numI, nnc, nUC1, nUC2)
  aMoire = 3.0_dp*ntheta**2 + 3.0_dp*ntheta + 1.0_dp
  g1  = (4.0_dp*pi/3.0_dp)/aMoire*(real(3*ntheta+1,dp)*a1+a2)
  g12 = (4.0_dp*pi/3.0_dp)/aMoire*(real(3*ntheta+2,dp)*a2-a1)
  cs = 1.0_dp-1.0_dp/(2.0_dp*aMoire)
  sn = sqrt(1.0_dp-cs**2)
  Rot = reshape([cos(0.5_dp*acos(cs)),-sin(0.5_dp*acos(cs)),sin(0.5_dp*acos(cs)),cos(0.5_dp*acos(cs))],[2,2])
  g1  = matmul(Rot,g1)
  g12 = matmul(Rot,g12)
  allocate(nMomComp(Nk,2), nMomFlat(numk,numk), MomVal(numk,numk,2))
  call SampleBZ(nMomComp, nMomFlat, MomVal, numk, g1, g12)
  t1 = real( ntheta  ,dp)*a1 + real(  ntheta+1,dp)*a2
  t2 = real(-ntheta-1,dp)*a1 + real(2*ntheta+1,dp)*a2
  allocate(Coords(ndim,3))
  call WignerSeitzCell(Coords, t1, t2, cs, sn)
  do n=1,ndim
    Coords(n,1:2) = matmul(Rot, Coords(n,1:2))
  end do
  t1 = matmul(Rot,t1)
  t2 = matmul(Rot,t2)
  t3 = t2-t1
  tn = reshape([t1,t2,t3],[2,3])

  call dump_r('Coords', Coords, 3*ndim, mu, [ndim,3_dp])
  call dump_r('tn', tn, 6_dp, mu, [2_dp,3_dp])
  call dump_r('g1', g1, 2_dp, mu, [2_dp])
  call dump_r('g12', g12, 2_dp, mu, [2_dp])
  call dump_r('MomentaValues', MomVal, 2*Nk, mu, [numk,numk,2_dp])
  call dump_i('nMomentaComponents', nMomComp, 2*Nk, mu, [Nk,2_dp])
  call dump_i('nUnitCell_1', nUC1, nnc, mu, [nnc])
  call dump_i('nUnitCell_2', nUC2, nnc, mu, [nnc])
  scal = 0
  scal(1) = ndim; scal(2) = numk; scal(3) = nfilling; scal(4) = xi; scal(5) = nLower; scal(6) = nUpper
  scal(7) = nnc; scal(8) = nlayers
  call dump_r('params', scal, 8_dp, mu, [8_dp])

  ! ---------------- long-range (Hartree) kernel
  allocate(LongRange(ndim,ndim))
  call LongRangeInteraction(LongRange, Coords, ndim, t1, t2)
  call dump_r('LongRange', LongRange, ndim*ndim, mu, [ndim,ndim])

  ! ---------------- tight-binding H(k) at two k-points
  ksel = [2_dp, Nk]
  allocate(zH(ndim,ndim))
  do kk=1,2
    ivk1 = nMomComp(ksel(kk),1); ivk2 = nMomComp(ksel(kk),2)
    vk = MomVal(ivk1+1,ivk2+1,:)
    call TightBindingHamiltonian(zH, Delta, Coords, nUC1, nUC2, ndim, nnc, vk, tn)
    if (kk.eq.1) call dump_z('zH_TB_k2', zH, ndim*ndim, mu, [ndim,ndim])
    if (kk.eq.2) call dump_z('zH_TB_kN', zH, ndim*ndim, mu, [ndim,ndim])
  end do

  ! ---------------- first diagonalisation (rho = 0, potential = 0)
  allocate(zFockBulk(ndim,ndim,nnc), zFock(ndim,ndim,nnc), zFockIn(ndim,ndim,nnc))
  allocate(Potential(ndim), Bands(nb,Nk), zEig(ndim,nb,Nk), zSorted(ndim,nb*Nk), Sorted(nb*Nk))
  allocate(nSortedMom(nb*Nk,2), Density(ndim,1), DensityIn(ndim,1), DensitySub(ndim))
  zFockBulk = 0; zFock = 0; Potential = 0
  DensitySub = 0.5_dp
  call Solve(zFockBulk, zFock, Potential, alpha, Bands, zEig, Coords, MomVal, nMomComp, nLower, nUpper, &
             ndim, numk, Nk, nnc, nUC1, nUC2, tn)
  call dump_r('Bands0', Bands, nb*Nk, mu, [nb,Nk])
  call dump_z('zFockBulk0', zFockBulk, ndim*ndim*nnc, mu, [ndim,ndim,nnc])
  call SortEnergies_1Spin(ndim, numk, nb, Nk, nMomComp, zSorted, Sorted, nSortedMom, zEig, Bands)
  call dump_r('Sorted0', Sorted, nb*Nk, mu, [nb*Nk])
  call dump_i('nSortedMomenta0', nSortedMom, 2*nb*Nk, mu, [nb*Nk,2_dp])
  call fermi(Sorted, nF, FermiE, nOcc, nPart, DegFactor)
  scal = 0
  scal(1) = nF; scal(2) = FermiE; scal(3) = nOcc; scal(4) = nPart; scal(5) = DegFactor
  call dump_r('fermi0', scal, 5_dp, mu, [5_dp])
  zFock = 0
  call GetFock(zFock, zSorted, ndim, nb, numk, DegFactor, nOcc, nPart, Coords, MomVal, nnc, nUC1, nUC2, nSortedMom)
  call dump_z('zFockWin0', zFock, ndim*ndim*nnc, mu, [ndim,ndim,nnc])
  zFock = (zFock + zFockBulk)/cmplx(real(Nk,dp),0.0_dp,dp)
  call dump_z('rho0', zFock, ndim*ndim*nnc, mu, [ndim,ndim,nnc])

  do n=1,ndim
    Density(n,1) = real(zFock(n,n,1),dp)
  end do
  call GetKineticEnergy(Ekin(1), Coords, zFock, Delta, nUC1, nUC2, ndim, nnc, tn)
  call GetFockEnergy(EF(1), Coords, zFock, nUC1, nUC2, ndim, nnc, tn)
  call GetHartreeHubbardEnergy(EH, EU, LongRange, Density(:,1)-DensitySub, Density(:,1)-DensitySub, ndim)
  Etot = alpha*EF(1)*2.0_dp + Ekin(1)*2.0_dp + U*EU + alphaH*EH
  scal = 0
  scal(1) = Ekin(1); scal(2) = EF(1); scal(3) = EH; scal(4) = EU; scal(5) = Etot
  call dump_r('energies0', scal, 5_dp, mu, [5_dp])

  Potential = 0
  do n=1,ndim
    do m=1,ndim
      Potential(n) = Potential(n) + 2.0_dp*(Density(m,1)-DensitySub(m))*alphaH*LongRange(n,m)
    end do
  end do
  Potential = Potential + U*(Density(:,1)-DensitySub)
  call dump_r('Potential0', Potential, ndim, mu, [ndim])

  ! ---------------- Hartree-Fock H(k) and its eigenvalues
  allocate(Energies(ndim))
  do kk=1,2
    ivk1 = nMomComp(ksel(kk),1); ivk2 = nMomComp(ksel(kk),2)
    vk = MomVal(ivk1+1,ivk2+1,:)
    call HamiltonianHartreeFock(zH, Coords, Potential, alpha, Delta, zFock, nUC1, nUC2, ndim, nnc, vk, tn)
    if (kk.eq.1) call dump_z('zH_HF_k2', zH, ndim*ndim, mu, [ndim,ndim])
    if (kk.eq.2) call dump_z('zH_HF_kN', zH, ndim*ndim, mu, [ndim,ndim])
    Energies = 0
    call diagonalize(zH, Energies, 'V', 1_dp, nUpper)
    if (kk.eq.1) call dump_r('eig_HF_k2', Energies, nUpper, mu, [nUpper])
    if (kk.eq.2) call dump_r('eig_HF_kN', Energies, nUpper, mu, [nUpper])
  end do

  ! ---------------- one SCF iteration from rho0 (Main.f90:373-574, ODA)
  zFockIn = zFock
  DensityIn = Density
  call Solve(zFockBulk, zFock, Potential, alpha, Bands, zEig, Coords, MomVal, nMomComp, nLower, nUpper, &
             ndim, numk, Nk, nnc, nUC1, nUC2, tn)
  call dump_r('Bands1', Bands, nb*Nk, mu, [nb,Nk])
  call dump_z('zFockBulk1', zFockBulk, ndim*ndim*nnc, mu, [ndim,ndim,nnc])
  call SortEnergies_1Spin(ndim, numk, nb, Nk, nMomComp, zSorted, Sorted, nSortedMom, zEig, Bands)
  call dump_r('Sorted1', Sorted, nb*Nk, mu, [nb*Nk])
  call fermi(Sorted, nF, FermiE, nOcc, nPart, DegFactor)
  scal = 0
  scal(1) = nF; scal(2) = FermiE; scal(3) = nOcc; scal(4) = nPart; scal(5) = DegFactor
  call dump_r('fermi1', scal, 5_dp, mu, [5_dp])
  zFock = 0
  call GetFock(zFock, zSorted, ndim, nb, numk, DegFactor, nOcc, nPart, Coords, MomVal, nnc, nUC1, nUC2, nSortedMom)
  call dump_z('zFockWin1', zFock, ndim*ndim*nnc, mu, [ndim,ndim,nnc])
  zFock = (zFock + zFockBulk)/cmplx(real(Nk,dp),0.0_dp,dp)
  call dump_z('rhoNew1', zFock, ndim*ndim*nnc, mu, [ndim,ndim,nnc])
  do n=1,ndim
    Density(n,1) = real(zFock(n,n,1),dp)
  end do
  allocate(zRhoNew(ndim,ndim,nnc,1), zRhoIn(ndim,ndim,nnc,1))
  zRhoNew(:,:,:,1) = zFock
  zRhoIn(:,:,:,1) = zFockIn
  call OptimalStep(Step, EH, EF, EU, Ekin, zRhoNew, zRhoIn, Density, DensityIn, DensitySub, ndim, nnc, &
                   Coords, LongRange, Delta, nUC1, nUC2, tn)
  Etot = alpha*EF(1)*2.0_dp + Ekin(1)*2.0_dp + U*EU + alphaH*EH
  scal = 0
  scal(1) = Ekin(1); scal(2) = EF(1); scal(3) = EH; scal(4) = EU; scal(5) = Etot; scal(6) = Step
  call dump_r('oda1', scal, 6_dp, mu, [6_dp])
  zFock = Step*zFock + (1.0_dp-Step)*zFockIn
  call dump_z('rho1', zFock, ndim*ndim*nnc, mu, [ndim,ndim,nnc])
  close(mu)

contains

  ! Fermi level / occupation counts exactly as Main.f90:400-412 (numS=1).
  subroutine fermi(S, nF, EFermi, nOcc, nPart, Deg)
    real(dp), intent(in) :: S(:)
    integer(dp), intent(out) :: nF, nOcc, nPart
    real(dp), intent(out) :: EFermi, Deg
    integer(dp) :: ic
    nF = nint(real(numk*numk*NeutralityPoint,dp)+real(numk*numk,dp)/2.0_dp*nfilling)
    EFermi = (S(nF)+S(nF+1))/2.
    nOcc = 0; nPart = 0
    do ic=1,size(S)
      if ((S(ic) - EFermi).lt.-EnergyTolerance) then
        nOcc = nOcc + 1
      else if (abs(S(ic) - EFermi).lt.EnergyTolerance) then
        nOcc = nOcc + 1
        nPart = nPart + 1
      end if
    end do
    Deg = 1.0_dp/real(nPart,dp)*real(nF+nPart-nOcc,dp)
  end subroutine fermi

  subroutine dump_r(name, a, cnt, mu, dims)
    character(*), intent(in) :: name
    integer(dp), intent(in) :: cnt, dims(:)
    real(dp), intent(in) :: a(*)
    integer, intent(in) :: mu
    integer :: u
    open(newunit=u, file=name//'.bin', access='stream', form='unformatted', status='replace')
    write(u) a(1:cnt)
    close(u)
    write(mu,'(A,1X,A,*(1X,I0))') name, 'f8', dims
  end subroutine dump_r

  subroutine dump_z(name, a, cnt, mu, dims)
    character(*), intent(in) :: name
    integer(dp), intent(in) :: cnt, dims(:)
    complex(dp), intent(in) :: a(*)
    integer, intent(in) :: mu
    integer :: u
    open(newunit=u, file=name//'.bin', access='stream', form='unformatted', status='replace')
    write(u) a(1:cnt)
    close(u)
    write(mu,'(A,1X,A,*(1X,I0))') name, 'c16', dims
  end subroutine dump_z

  subroutine dump_i(name, a, cnt, mu, dims)
    character(*), intent(in) :: name
    integer(dp), intent(in) :: cnt, dims(:)
    integer(dp), intent(in) :: a(*)
    integer, intent(in) :: mu
    integer :: u
    open(newunit=u, file=name//'.bin', access='stream', form='unformatted', status='replace')
    write(u) a(1:cnt)
    close(u)
    write(mu,'(A,1X,A,*(1X,I0))') name, 'i8', dims
  end subroutine dump_i

end program ref_harness
