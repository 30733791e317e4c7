! forge_hartreefock.f90 -- drop-in shims for the reference's hot-path routines.
!
! Each subroutine keeps the name and the dummy-argument list of the reference
! routine it replaces, so a host written against the reference calls it
! unchanged; the body calls the MI355X engine through forge_c.  Arguments
! the engine already holds (the geometry and the model constants, uploaded
! once by forge_ctx_create) are compared with the context's copies: sizes
! always, alpha_, Delta_, Coords and tn when forge_attach was given them.  numS = 1 or 2 (OptimalStep).
!
!   LongRangeInteraction     HartreeFock.f90:14-53
!   HamiltonianHartreeFock   HartreeFock.f90:57-127
!   Solve                    HartreeFock.f90:402-483
!   GetFock                  HartreeFock.f90:487-560
!   SortEnergies_1Spin       HartreeFock.f90:885-925   (host: 20 Nk keys)
!   GetKineticEnergy         HartreeFock.f90:1012-1059
!   GetHartreeHubbardEnergy  HartreeFock.f90:1066-1092
!   GetFockEnergy            HartreeFock.f90:1098-1145
!   OptimalStep              HartreeFock.f90:1148-1227
!   TightBindingHamiltonian  TightBinding.f90:13-55
!   zdiagRed                 LapackRoutines.f90:381-433
!   z_vector_mul             LapackRoutines.f90:168-183   (A may hold k > 1 columns)
!   z_matrix_mul, c_matrix_mul  LapackRoutines.f90:208-236
!   z_matrix_overlap         LapackRoutines.f90:119-139
!   C3_RelatedPoints         Geometry.f90:1172-1223    (host)
!   C2_RelatedPoints         Geometry.f90:1140-1167    (host)
!   Solve_C3                 HartreeFock.f90:131-398
!   GetFock_C3               HartreeFock.f90:564-881
!   SortEnergies_2Spins      HartreeFock.f90:929-1004  (host)
!   ValleyTransform          TightBinding.f90:170-211
module forge_hartreefock
  use iso_c_binding
  use forge_c
  use forge_setup
  implicit none
  private
  public :: forge_attach, forge_check, LongRangeInteraction, HamiltonianHartreeFock, Solve, GetFock
  public :: C3_RelatedPoints, Solve_C3, GetFock_C3, SortEnergies_2Spins, C2_RelatedPoints
  public :: SortEnergies_1Spin, GetKineticEnergy, GetHartreeHubbardEnergy, GetFockEnergy
  public :: OptimalStep, TightBindingHamiltonian, zdiagRed
  public :: vector_mul, z_vector_mul, matrix_mul, z_matrix_mul, c_matrix_mul
  public :: matrix_overlap, z_matrix_overlap, ValleyTransform

  ! generics of the reference's LAPACK shim (LapackRoutines.f90:21-37)
  interface vector_mul
    module procedure z_vector_mul
  end interface
  interface matrix_mul
    module procedure z_matrix_mul, c_matrix_mul
  end interface
  interface matrix_overlap
    module procedure z_matrix_overlap
  end interface

  type(c_ptr), save :: ctx = c_null_ptr
  integer(dp), save :: ctx_ndim = -1, ctx_nnc = -1
  ! the model the context was built with (forge_params and forge_ctx_create's
  ! arrays): the shims compare the caller's alpha_, Delta_, Coords and tn
  ! with these and stop on a difference, since the engine uses its own copies
  logical, save :: ctx_model = .false.
  real(dp), save :: ctx_alpha = 0.0_dp, ctx_delta = 0.0_dp, ctx_tn(2, 3) = 0.0_dp
  real(dp), allocatable, save :: ctx_coords(:, :)


contains

  ! The engine context the shims call (created by the host with forge_ctx_create).
  ! alpha_, Delta_, Coords and tn: the values the context was created with.
  subroutine forge_attach(c, ndim_, nnc_, alpha_, Delta_, Coords, tn)
    type(c_ptr), intent(in) :: c
    integer(dp), intent(in) :: ndim_, nnc_
    real(dp), intent(in), optional :: alpha_, Delta_, Coords(:, :), tn(2, 3)
    ctx = c
    ctx_ndim = ndim_
    ctx_nnc = nnc_
    ctx_model = present(alpha_) .and. present(Delta_) .and. present(Coords) .and. present(tn)
    if (ctx_model) then
      ctx_alpha = alpha_
      ctx_delta = Delta_
      ctx_tn = tn
      if (allocated(ctx_coords)) deallocate(ctx_coords)
      allocate(ctx_coords(size(Coords, 1), size(Coords, 2)))
      ctx_coords = Coords
    end if
  end subroutine forge_attach

  ! Stops when an argument differs from the value the context holds.
  subroutine same_model(where, alpha_, Delta_, Coords, tn)
    character(*), intent(in) :: where
    real(dp), intent(in), optional :: alpha_, Delta_, Coords(:, :), tn(2, 3)
    character(16) :: what
    if (.not. ctx_model) return
    what = ''
    if (present(alpha_)) then
      if (alpha_ /= ctx_alpha) what = 'alpha'
    end if
    if (present(Delta_)) then
      if (Delta_ /= ctx_delta) what = 'Delta'
    end if
    if (present(tn)) then
      if (any(tn /= ctx_tn)) what = 'tn'
    end if
    if (present(Coords)) then
      if (any(shape(Coords) /= shape(ctx_coords))) then
        what = 'Coords'
      else if (any(Coords /= ctx_coords)) then
        what = 'Coords'
      end if
    end if
    if (what /= '') then
      write(*, *) 'forge: ', where, ': ', trim(what), ' differs from the value the engine context holds'
      error stop 2
    end if
  end subroutine same_model

  ! Stops with the engine's message on a negative return code.
  subroutine forge_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc < 0) then
      write(*, '(A,A,A,I0,A,A)') 'forge: ', where, ' failed (', rc, '): ', forge_error_message()
      error stop 2
    end if
  end subroutine forge_check

  subroutine need(ndim_, nnc_, where)
    integer(dp), intent(in) :: ndim_, nnc_
    character(*), intent(in) :: where
    if (.not. c_associated(ctx)) then
      write(*, *) 'forge: ', where, ' called before forge_attach'
      error stop 2
    end if
    if (ndim_ /= ctx_ndim .or. (nnc_ >= 0 .and. nnc_ /= ctx_nnc)) then
      write(*, *) 'forge: ', where, ': sizes', ndim_, nnc_, 'differ from the context', ctx_ndim, ctx_nnc
      error stop 2
    end if
  end subroutine need

  subroutine LongRangeInteraction(LongRange, Coords, ndim_, t1, t2)
    integer(dp), intent(in) :: ndim_
    real(dp), intent(in) :: t1(2), t2(2), Coords(ndim_, 3)
    real(dp), intent(inout) :: LongRange(ndim_, ndim_)
    call need(ndim_, -1_dp, 'LongRangeInteraction')
    call same_model('LongRangeInteraction', Coords=Coords)
    if (ctx_model) then
      if (any(t1 /= ctx_tn(:, 1)) .or. any(t2 /= ctx_tn(:, 2))) then
        write(*, *) 'forge: LongRangeInteraction: t1, t2 differ from the engine context'
        error stop 2
      end if
    end if
    call forge_check(forge_longrange(ctx, LongRange), 'LongRangeInteraction')
  end subroutine LongRangeInteraction

  subroutine HamiltonianHartreeFock(zH, Coords, Potential, alpha_, Delta_, zFock, nUnitCell_1, &
                                    nUnitCell_2, ndim_, numNeighborCells, vk, tn)
    integer(dp), intent(in) :: ndim_, numNeighborCells
    integer(dp), intent(in) :: nUnitCell_1(numNeighborCells), nUnitCell_2(numNeighborCells)
    real(dp), intent(in) :: tn(2, 3), vk(2), Coords(ndim_, 3)
    real(dp), intent(in) :: alpha_, Delta_, Potential(ndim_)
    complex(dp), intent(in) :: zFock(ndim_, ndim_, numNeighborCells)
    complex(dp), intent(out) :: zH(ndim_, ndim_)
    call need(ndim_, numNeighborCells, 'HamiltonianHartreeFock')
    call same_model('HamiltonianHartreeFock', alpha_=alpha_, Delta_=Delta_, Coords=Coords, tn=tn)
    call forge_check(forge_hamiltonian_hf(ctx, zH, Potential, zFock, vk), 'HamiltonianHartreeFock')
  end subroutine HamiltonianHartreeFock

  subroutine TightBindingHamiltonian(zH, Delta_, Coords, nUnitCell_1, nUnitCell_2, ndim_, &
                                     numNeighborCells, vk, tn)
    integer(dp), intent(in) :: ndim_, numNeighborCells
    integer(dp), intent(in) :: nUnitCell_1(numNeighborCells), nUnitCell_2(numNeighborCells)
    real(dp), intent(in) :: tn(2, 3), vk(2), Coords(ndim_, 3), Delta_
    complex(dp), intent(out) :: zH(ndim_, ndim_)
    call need(ndim_, numNeighborCells, 'TightBindingHamiltonian')
    call same_model('TightBindingHamiltonian', Delta_=Delta_, Coords=Coords, tn=tn)
    call forge_check(forge_tb_hamiltonian(ctx, zH, vk), 'TightBindingHamiltonian')
  end subroutine TightBindingHamiltonian

  ! Eigenpairs nLower..nUpper of the Hermitian matrix in A's lower triangle;
  ! A(:, 1:nUpper-nLower+1) receives the eigenvectors.  A failure count is
  ! reported as the reference reports zheevr's info, and the run continues.
  subroutine zdiagRed(A, eigenvalues, ev, nLower_, nUpper_)
    complex(8), intent(inout) :: A(:, :)
    real(8), intent(inout) :: eigenvalues(:)
    character(1), intent(in) :: ev
    integer(8), intent(in) :: nLower_, nUpper_
    integer(c_int32_t) :: info
    integer(c_int) :: rc
    rc = forge_zdiagred(A, int(size(A, 1), c_int64_t), eigenvalues, nLower_, nUpper_, info)
    call forge_check(rc, 'zdiagRed')
    if (info /= 0) write(*, *) 'zheev failed', info
  end subroutine zdiagRed

  ! Every k's window eigenvectors <- C sign(C^H V C) with V the valley
  ! operator at k.  The engine builds V from its own geometry, the stored
  ! cells and RotateLayers (forge_setup); RotMatrix is checked against the
  ! rotation the engine uses (by half the twist angle of the cell,
  ! Main.f90:133).
  subroutine ValleyTransform(zEigenvectors, Coords, MomentaValues, nMomentaComponents, numb_, ndim_, &
                             numk_, Nk_, numNeighborCells, nUnitCell_1, nUnitCell_2, tn, RotMatrix)
    integer(dp), intent(in) :: ndim_, numk_, Nk_, numb_, numNeighborCells
    integer(dp), intent(in) :: nMomentaComponents(Nk_, 2), nUnitCell_1(numNeighborCells)
    integer(dp), intent(in) :: nUnitCell_2(numNeighborCells)
    real(dp), intent(in) :: MomentaValues(numk_, numk_, 2), Coords(ndim_, 3), tn(2, 3), RotMatrix(2, 2)
    complex(dp), intent(inout) :: zEigenvectors(ndim_, numb_, Nk_)
    integer(c_int64_t), allocatable, target :: rot(:)
    real(dp) :: aM, h
    call need(ndim_, numNeighborCells, 'ValleyTransform')
    call same_model('ValleyTransform', Coords=Coords, tn=tn)
    aM = real(ndim_ / (2 * nlayers), dp)
    h = 0.5_dp * acos(1.0_dp - 1.0_dp / (2.0_dp * aM))
    if (maxval(abs(RotMatrix - reshape([cos(h), -sin(h), sin(h), cos(h)], [2, 2]))) > 1e-12_dp) then
      write(*, *) 'forge: ValleyTransform: RotMatrix differs from the rotation of the engine''s cell'
      error stop 2
    end if
    allocate(rot(nlayers))
    rot = RotateLayers
    call forge_check(forge_valley_transform(ctx, zEigenvectors, rot), 'ValleyTransform')
  end subroutine ValleyTransform

  ! C = C + z_factor A A^H (the reference's rank-1 update with one column of
  ! A; every column of A here)
  subroutine z_vector_mul(A, C, z_factor)
    complex(8), intent(in) :: A(:, :)
    complex(8), intent(inout), target :: C(:, :)
    complex(8), intent(in), optional, target :: z_factor
    type(c_ptr) :: zp
    complex(8), target :: zloc
    zp = c_null_ptr
    if (present(z_factor)) then
      zloc = z_factor
      zp = c_loc(zloc)
    end if
    call forge_check(forge_z_vector_mul(C, A, int(size(A, 1), c_int64_t), int(size(A, 2), c_int64_t), &
                                        zp), 'z_vector_mul')
  end subroutine z_vector_mul

  ! C = A B
  subroutine z_matrix_mul(A, B, C)
    complex(8), intent(in) :: A(:, :), B(:, :)
    complex(8), intent(inout) :: C(:, :)
    call forge_check(forge_z_matrix_mul(C, A, B, int(size(A, 1), c_int64_t), &
                                        int(size(A, 2), c_int64_t), int(size(B, 2), c_int64_t)), &
                     'z_matrix_mul')
  end subroutine z_matrix_mul

  subroutine c_matrix_mul(A, B, C)
    complex(4), intent(in) :: A(:, :), B(:, :)
    complex(4), intent(inout) :: C(:, :)
    call forge_check(forge_c_matrix_mul(C, A, B, int(size(A, 1), c_int64_t), &
                                        int(size(A, 2), c_int64_t), int(size(B, 2), c_int64_t)), &
                     'c_matrix_mul')
  end subroutine c_matrix_mul

  ! C = z_factor A^H B
  subroutine z_matrix_overlap(A, B, C, z_factor)
    complex(8), intent(in) :: A(:, :), B(:, :)
    complex(8), intent(out) :: C(:, :)
    complex(8), intent(in), optional :: z_factor
    type(c_ptr) :: zp
    complex(8), target :: zloc
    zp = c_null_ptr
    if (present(z_factor)) then
      zloc = z_factor
      zp = c_loc(zloc)
    end if
    call forge_check(forge_z_matrix_overlap(C, A, B, int(size(A, 1), c_int64_t), &
                                            int(size(A, 2), c_int64_t), zp), 'z_matrix_overlap')
  end subroutine z_matrix_overlap

  subroutine Solve(zFockBulk, zFock, Potential, alpha_, Bands, zEigenvectors, Coords, MomentaValues, &
                   nMomentaComponents, nLower_, nUpper_, ndim_, numk_, Nk_, numNeighborCells, &
                   nUnitCell_1, nUnitCell_2, tn)
    integer(dp), intent(in) :: ndim_, numk_, Nk_, nLower_, nUpper_, numNeighborCells
    integer(dp), intent(in) :: nMomentaComponents(Nk_, 2)
    integer(dp), intent(in) :: nUnitCell_1(numNeighborCells), nUnitCell_2(numNeighborCells)
    real(dp), intent(in) :: tn(2, 3), MomentaValues(numk_, numk_, 2), Coords(ndim_, 3)
    real(dp), intent(in) :: Potential(ndim_), alpha_
    complex(dp), intent(in) :: zFock(ndim_, ndim_, numNeighborCells)
    real(dp), intent(out) :: Bands(nUpper_ - nLower_ + 1, Nk_)
    complex(dp), intent(out) :: zEigenvectors(ndim_, nUpper_ - nLower_ + 1, Nk_)
    complex(dp), intent(out) :: zFockBulk(ndim_, ndim_, numNeighborCells)
    integer(c_int) :: rc
    call need(ndim_, numNeighborCells, 'Solve')
    call same_model('Solve', alpha_=alpha_, Coords=Coords, tn=tn)
    rc = forge_solve(ctx, zFockBulk, zFock, Potential, Bands, zEigenvectors)
    call forge_check(rc, 'Solve')
    if (rc > 0) write(*, *) 'zheev failed', rc
  end subroutine Solve

  subroutine GetFock(zFock, zSortedEigenvectors, ndim_, numb_, numk_, DegFactor, OccStates, &
                     PartOccStates, Coords, MomentaValues, numNeighborCells, nUnitCell_1, &
                     nUnitCell_2, nSortedMomenta)
    integer(dp), intent(in) :: ndim_, numb_, numk_, OccStates, PartOccStates, numNeighborCells
    integer(dp), intent(in) :: nUnitCell_1(numNeighborCells), nUnitCell_2(numNeighborCells)
    integer(dp), intent(in) :: nSortedMomenta(numb_ * numk_ * numk_, 2)
    real(dp), intent(in) :: Coords(ndim_, 3), MomentaValues(numk_, numk_, 2), DegFactor
    complex(dp), intent(in) :: zSortedEigenvectors(ndim_, numb_ * numk_ * numk_)
    complex(dp), intent(inout) :: zFock(ndim_, ndim_, numNeighborCells)
    call need(ndim_, numNeighborCells, 'GetFock')
    call forge_check(forge_getfock(ctx, zFock, zSortedEigenvectors, nSortedMomenta, DegFactor, &
                                   OccStates, PartOccStates), 'GetFock')
  end subroutine GetFock

  ! C3_RelatedPoints(nC3pairs, Coords, ndim), Geometry.f90:1172-1223 (host code
  ! of the library; prints the reference's message when the cell is not
  ! C3-symmetric).
  subroutine C3_RelatedPoints(nC3pairs, Coords, ndim_)
    integer(dp), intent(in) :: ndim_
    real(dp), intent(in) :: Coords(ndim_, 3)
    integer(dp), intent(inout) :: nC3pairs(ndim_)
    integer(dp) :: cnt
    cnt = forge_c3_pairs(Coords, ndim_, int(nlayers, c_int64_t), nC3pairs)
    if (cnt < ndim_) write(*, *) 'ERROR C3. cnt=', cnt
  end subroutine C3_RelatedPoints

  ! C2_RelatedPoints(nC2pairs, Coords, ndim), Geometry.f90:1140-1167 (host
  ! code of the library; prints the reference's message when the cell is not
  ! C2-symmetric).
  subroutine C2_RelatedPoints(nC2pairs, Coords, ndim_)
    integer(dp), intent(in) :: ndim_
    real(dp), intent(in) :: Coords(ndim_, 3)
    integer(dp), intent(inout) :: nC2pairs(ndim_)
    integer(dp) :: cnt
    cnt = forge_c2_pairs(Coords, ndim_, int(nlayers, c_int64_t), nC2pairs)
    if (cnt < ndim_ / 2) write(*, *) 'ERROR C2. cnt=', cnt
  end subroutine C2_RelatedPoints

  subroutine Solve_C3(zFockBulk, zFock, Potential, alpha_, Bands, zEigenvectors, Coords, &
                      MomentaValues, nMomentaComponents, nLower_, nUpper_, ndim_, RotateLayers_, &
                      numk_, Nk_, numNeighborCells, nUnitCell_1, nUnitCell_2, tn, g1, g12, nC3pairs)
    integer(dp), intent(in) :: ndim_, numk_, Nk_, nLower_, nUpper_, numNeighborCells
    integer(dp), intent(in) :: nC3pairs(ndim_), RotateLayers_(*)
    integer(dp), intent(in) :: nMomentaComponents(Nk_, 2)
    integer(dp), intent(in) :: nUnitCell_1(numNeighborCells), nUnitCell_2(numNeighborCells)
    real(dp), intent(in) :: tn(2, 3), MomentaValues(numk_, numk_, 2), Coords(ndim_, 3)
    real(dp), intent(in) :: Potential(ndim_), g1(2), g12(2), alpha_
    complex(dp), intent(in) :: zFock(ndim_, ndim_, numNeighborCells)
    real(dp), intent(out) :: Bands(nUpper_ - nLower_ + 1, Nk_)
    complex(dp), intent(out) :: zEigenvectors(ndim_, nUpper_ - nLower_ + 1, Nk_)
    complex(dp), intent(out) :: zFockBulk(ndim_, ndim_, numNeighborCells)
    integer(c_int) :: rc
    call need(ndim_, numNeighborCells, 'Solve_C3')
    call same_model('Solve_C3', alpha_=alpha_, Coords=Coords, tn=tn)
    rc = forge_solve_c3(ctx, zFockBulk, zFock, Potential, Bands, zEigenvectors, RotateLayers_, &
                        nC3pairs)
    call forge_check(rc, 'Solve_C3')
    if (rc > 0) write(*, *) 'zheev failed', rc
  end subroutine Solve_C3

  subroutine GetFock_C3(zFock, zSortedEigenvectors, ndim_, RotateLayers_, numb_, numk_, DegFactor, &
                        OccStates, PartOccStates, Coords, MomentaValues, numNeighborCells, &
                        nUnitCell_1, nUnitCell_2, nSortedMomenta, nC3pairs)
    integer(dp), intent(in) :: ndim_, numb_, numk_, OccStates, PartOccStates, numNeighborCells
    integer(dp), intent(in) :: RotateLayers_(*), nC3pairs(ndim_)
    integer(dp), intent(in) :: nUnitCell_1(numNeighborCells), nUnitCell_2(numNeighborCells)
    integer(dp), intent(in) :: nSortedMomenta(numb_ * numk_ * numk_, 2)
    real(dp), intent(in) :: Coords(ndim_, 3), MomentaValues(numk_, numk_, 2), DegFactor
    complex(dp), intent(in) :: zSortedEigenvectors(ndim_, numb_ * numk_ * numk_)
    complex(dp), intent(inout) :: zFock(ndim_, ndim_, numNeighborCells)
    call need(ndim_, numNeighborCells, 'GetFock_C3')
    call forge_check(forge_getfock_c3(ctx, zFock, zSortedEigenvectors, nSortedMomenta, DegFactor, &
                                      OccStates, PartOccStates, RotateLayers_, nC3pairs), &
                     'GetFock_C3')
  end subroutine GetFock_C3

  ! Window states of all k flattened band-major (index (b-1) Nk + k) and
  ! sorted by energy (heap sort); vectors and momenta follow their energies.
  subroutine SortEnergies_1Spin(ndim_, numk_, numb_, Nk_, nMomentaComponents, zSortedEigenvectors, &
                                SortedEnergies_1Spin, nSortedMomenta, zEigenvectors, Bands)
    integer(dp), intent(in) :: numb_, numk_, Nk_, ndim_
    integer(dp), intent(in) :: nMomentaComponents(Nk_, 2)
    real(dp), intent(in) :: Bands(numb_, Nk_)
    complex(dp), intent(in) :: zEigenvectors(ndim_, numb_, Nk_)
    integer(dp), intent(inout) :: nSortedMomenta(numb_ * numk_ * numk_, 2)
    real(dp), intent(inout) :: SortedEnergies_1Spin(numb_ * numk_ * numk_)
    complex(dp), intent(inout) :: zSortedEigenvectors(ndim_, numb_ * numk_ * numk_)
    integer(dp) :: N, i, b, k, ic
    integer(dp), allocatable :: idx(:)
    real(dp), allocatable :: key(:)
    N = numb_ * Nk_
    allocate(idx(N), key(N))
    do b = 1, numb_
      do k = 1, Nk_
        ic = (b - 1) * Nk_ + k
        key(ic) = Bands(b, k)
        idx(ic) = ic
      end do
    end do
    call heap_sort(key, idx, N)
    do i = 1, N
      b = (idx(i) - 1) / Nk_ + 1
      k = idx(i) - (b - 1) * Nk_
      SortedEnergies_1Spin(i) = key(i)
      zSortedEigenvectors(:, i) = zEigenvectors(:, b, k)
      nSortedMomenta(i, :) = nMomentaComponents(k, :)
    end do
  end subroutine SortEnergies_1Spin

  ! Ascending heap sort of key with idx carried along.
  subroutine heap_sort(key, idx, N)
    integer(dp), intent(in) :: N
    real(dp), intent(inout) :: key(N)
    integer(dp), intent(inout) :: idx(N)
    integer(dp) :: i, last
    do i = N / 2, 1, -1
      call sift(i, N)
    end do
    do last = N, 2, -1
      call swap(1_dp, last)
      call sift(1_dp, last - 1)
    end do
  contains
    subroutine sift(root, hi)
      integer(dp), intent(in) :: root, hi
      integer(dp) :: r, c
      r = root
      do
        c = 2 * r
        if (c > hi) exit
        if (c < hi) then
          if (key(c + 1) > key(c)) c = c + 1
        end if
        if (key(r) >= key(c)) exit
        call swap(r, c)
        r = c
      end do
    end subroutine sift
    subroutine swap(a, b)
      integer(dp), intent(in) :: a, b
      real(dp) :: t
      integer(dp) :: ti
      t = key(a); key(a) = key(b); key(b) = t
      ti = idx(a); idx(a) = idx(b); idx(b) = ti
    end subroutine swap
  end subroutine heap_sort

  subroutine GetKineticEnergy(fsum, Coords, zFock, Delta_, nUnitCell_1, nUnitCell_2, ndim_, &
                              numNeighborCells, tn)
    integer(dp), intent(in) :: ndim_, numNeighborCells
    integer(dp), intent(in) :: nUnitCell_1(numNeighborCells), nUnitCell_2(numNeighborCells)
    real(dp), intent(in) :: tn(2, 3), Coords(ndim_, 3), Delta_
    complex(dp), intent(in) :: zFock(ndim_, ndim_, numNeighborCells)
    real(dp), intent(out) :: fsum
    real(dp) :: e(4), d(ndim_)
    integer(dp) :: i
    call need(ndim_, numNeighborCells, 'GetKineticEnergy')
    ! the layer-bias term reads the site densities of zFock itself
    do i = 1, ndim_
      d(i) = real(zFock(i, i, 1), dp) - 0.5_dp
    end do
    call forge_check(forge_energies(ctx, zFock, d, e), 'GetKineticEnergy')
    fsum = e(1)
  end subroutine GetKineticEnergy

  subroutine GetFockEnergy(fsum, Coords, zFock, nUnitCell_1, nUnitCell_2, ndim_, numNeighborCells, tn)
    integer(dp), intent(in) :: ndim_, numNeighborCells
    integer(dp), intent(in) :: nUnitCell_1(numNeighborCells), nUnitCell_2(numNeighborCells)
    real(dp), intent(in) :: tn(2, 3), Coords(ndim_, 3)
    complex(dp), intent(in) :: zFock(ndim_, ndim_, numNeighborCells)
    real(dp), intent(out) :: fsum
    real(dp) :: e(4), d(ndim_)
    call need(ndim_, numNeighborCells, 'GetFockEnergy')
    d = 0.0_dp
    call forge_check(forge_energies(ctx, zFock, d, e), 'GetFockEnergy')
    fsum = e(2)
  end subroutine GetFockEnergy

  subroutine GetHartreeHubbardEnergy(fsumH, fsumU, LongRange, dens1, dens2, ndim_)
    integer(dp), intent(in) :: ndim_
    real(dp), intent(in) :: LongRange(ndim_, ndim_)
    real(dp), intent(in) :: dens1(ndim_), dens2(ndim_)
    real(dp), intent(out) :: fsumH, fsumU
    real(dp) :: e(2)
    call need(ndim_, -1_dp, 'GetHartreeHubbardEnergy')
    call forge_check(forge_hartree_hubbard(ctx, dens1, dens2, e), 'GetHartreeHubbardEnergy')
    fsumH = e(1)
    fsumU = e(2)
  end subroutine GetHartreeHubbardEnergy

  ! Optimal damping: the total energy is exact-quadratic in the mixing weight;
  ! fit it through E(0) (the energies passed in), E(1) and E(1/2), take the
  ! minimiser when it lies in (0, 1), else 1 if E(1) < E(0), else
  ! StepAlternative; on return the energies are those of the mixed state.
  subroutine OptimalStep(Step, HartreeEnergy, FockEnergy, HubbardEnergy, KineticEnergy, zFock, &
                         zFockIn, Density, DensityIn, DensitySub, ndim_, numNeighborCells, Coords, &
                         LongRange, Delta_, nUnitCell_1, nUnitCell_2, tn)
    integer(dp), intent(in) :: ndim_, numNeighborCells
    integer(dp), intent(in) :: nUnitCell_1(numNeighborCells), nUnitCell_2(numNeighborCells)
    complex(dp), intent(in) :: zFock(ndim_, ndim_, numNeighborCells, numS)
    complex(dp), intent(in) :: zFockIn(ndim_, ndim_, numNeighborCells, numS)
    real(dp), intent(in) :: Coords(ndim_, 3), LongRange(ndim_, ndim_), tn(2, 3), Delta_
    real(dp), intent(in) :: Density(ndim_, numS), DensityIn(ndim_, numS), DensitySub(ndim_)
    real(dp), intent(inout) :: HartreeEnergy, FockEnergy(numS), HubbardEnergy, KineticEnergy(numS)
    real(dp), intent(out) :: Step
    complex(dp), allocatable :: zmix(:, :, :, :)
    real(dp) :: d(ndim_, numS), c0, c1, c2, E1, Eh
    call need(ndim_, numNeighborCells, 'OptimalStep')
    c0 = total()
    d = Density - spread(DensitySub, 2, numS)
    call energies_of(zFock, d)
    E1 = total()
    allocate(zmix(ndim_, ndim_, numNeighborCells, numS))
    zmix = 0.5_dp * (zFockIn + zFock)
    d = 0.5_dp * (Density + DensityIn) - spread(DensitySub, 2, numS)
    call energies_of(zmix, d)
    Eh = total()
    c1 = 4.0_dp * Eh - E1 - 3.0_dp * c0
    c2 = E1 - c0 - c1
    if (c2 > 0.0_dp .and. abs(-c1 / 2.0_dp / c2 - 0.5_dp) < 0.5_dp) then
      Step = -c1 / 2.0_dp / c2
    else if (E1 < c0) then
      Step = 1.0_dp
    else
      Step = StepAlternative
    end if
    zmix = (1.0_dp - Step) * zFockIn + Step * zFock
    d = (1.0_dp - Step) * DensityIn + Step * Density - spread(DensitySub, 2, numS)
    call energies_of(zmix, d)
  contains
    ! the energies of HartreeFock.f90:1162-1221 for one density per spin: per
    ! spin Ekin and EF; EH and EU of the pair (d1, d2), or (d, d) for numS = 1
    subroutine energies_of(z, dd)
      complex(dp), intent(in) :: z(ndim_, ndim_, numNeighborCells, numS)
      real(dp), intent(in) :: dd(ndim_, numS)
      real(dp) :: e(4), hh(2)
      integer(dp) :: s
      do s = 1, numS
        call forge_check(forge_energies(ctx, z(:, :, :, s), dd(:, s), e), 'OptimalStep')
        KineticEnergy(s) = e(1)
        FockEnergy(s) = e(2)
      end do
      if (numS == 1) then
        HartreeEnergy = e(3)
        HubbardEnergy = e(4)
      else
        call forge_check(forge_hartree_hubbard(ctx, dd(:, 1), dd(:, 2), hh), 'OptimalStep')
        HartreeEnergy = hh(1)
        HubbardEnergy = hh(2)
      end if
    end subroutine energies_of
    real(dp) function total()
      total = alpha * sum(FockEnergy) * real(3 - numS, dp) + sum(KineticEnergy) * real(3 - numS, dp) &
              + U * HubbardEnergy + alphaH * HartreeEnergy
    end function total
  end subroutine OptimalStep

  ! SortEnergies_2Spins, HartreeFock.f90:929-1004.  S1(:, s) holds spin s's
  ! sorted window energies; S2 receives (energy, spin label) rows: the
  ! lead-in rows of spin 1 and of spin 2 (the reference fills spin 2's lead-in
  ! with spin 2's energy at the spin-1 counter; kept), then both spins merged
  ! by energy while each counter stays within row Nk (W(2, s) - 1), then the
  ! remaining rows of spin 1 and of spin 2.
  subroutine SortEnergies_2Spins(numk_, numb_, W, S1, S2)
    integer(dp), intent(in) :: numk_, numb_, W(2, 2)
    real(dp), intent(in) :: S1(numb_ * numk_ * numk_, 2)
    real(dp), intent(inout) :: S2(2 * numb_ * numk_ * numk_, 2)
    integer(dp) :: nk2, last, c(2), r, q
    nk2 = numk_ * numk_
    last = numb_ * nk2
    r = 0
    c = 1
    do q = 1, (W(1, 1) - 1) * nk2
      call put(S1(c(1), 1), 1_dp)
      c(1) = c(1) + 1
    end do
    do q = 1, (W(1, 2) - 1) * nk2
      call put(S1(c(1), 2), 2_dp)
      c(2) = c(2) + 1
    end do
    do while (c(1) <= nk2 * (W(2, 1) - 1) .and. c(2) <= nk2 * (W(2, 2) - 1))
      if (S1(c(1), 1) <= S1(c(2), 2)) then
        call run(1_dp, 2_dp)
      else
        call run(2_dp, 1_dp)
      end if
    end do
    do q = 1, 2
      do while (c(q) <= last)
        call put(S1(c(q), q), int(q, dp))
        c(q) = c(q) + 1
      end do
    end do
  contains
    subroutine put(e, lab)
      real(dp), intent(in) :: e
      integer(dp), intent(in) :: lab
      r = r + 1
      S2(r, 1) = e
      S2(r, 2) = real(lab, dp)
    end subroutine put
    ! rows of spin a while they do not exceed the current row of spin o, then
    ! that row of spin o
    subroutine run(a, o)
      integer(dp), intent(in) :: a, o
      call put(S1(c(a), a), a)
      c(a) = c(a) + 1
      do while (c(a) <= last)
        if (S1(c(a), a) > S1(c(o), o)) exit
        call put(S1(c(a), a), a)
        c(a) = c(a) + 1
      end do
      call put(S1(c(o), o), o)
      c(o) = c(o) + 1
    end subroutine run
  end subroutine SortEnergies_2Spins

end module forge_hartreefock
