! forge_main.f90 -- Fortran host of the MI355X Hartree-Fock engine.
!
! Reproduces the call sequence of the reference driver (Main.f90:79-679) for
! the spin-singlet case: geometry, LongRange, initial state (first
! diagonalisation or the checkpoint read of Main.f90:187-235), energies and
! potential, then the SCF loop (Solve, SortEnergies_1Spin, Fermi level,
! GetFock, OptimalStep, mixing, potential) with the Energy and Convergence
! rows of Main.f90:616-639.
!
!   mode=reference  every hot-path routine is the drop-in shim of
!                   forge_hartreefock, called exactly where the reference calls
!                   its own routine; host arrays cross the C ABI on each call;
!   mode=device     the loop body runs device-resident (forge_scf_run); the
!                   host only sets options and collects the rows;
!   mode=geometry   writes the geometry and stops (no device needed).
!
! Arguments (key=value): ntheta nlayers numk nfilling itmax xi_nm epsilon U
! Delta mode start (first | path of a checkpoint in the reference's Fock
! format) occ mix kT beta etol precision switch_de out (trajectory file)
! refout (optional: a directory prefix such as output/, which must exist,
! receiving the reference's own Energy<parameters> and
! Convergence<parameters> files, Main.f90:352-358, 616-631) c3 (nenforceC3,
! Setup.f90:27: 1 = Solve_C3 / GetFock_C3 where Main.f90 switches on it)
! numS (1, or 2 for independent up and down spins with the reference's
! weights and optimal damping: mode=reference calls the per-spin shims,
! mode=device runs the two-spin loop forge_scf2_run) start2 (spin 2's checkpoint when
! numS = 2; start= is then spin 1's) nrelax (1: relaxed bilayer lattice,
! Main.f90:156-160; relaxtable= overrides forge_amd/data/relaxation_fourier.txt)
! T C2 C2T (nenforceT / nenforceC2 /
! nenforceC2T, Setup.f90:28-30: 1 = symmetrise each new density as
! Main.f90:514-534 does) V (nenforceValley, Setup.f90:31: 1 = the valley
! projection of Main.f90:469-494, also together with c3=1; numS = 1).
program forge_main
  use iso_c_binding
  use forge_c
  use forge_setup
  use forge_geometry
  use forge_hartreefock
  implicit none

  character(len=16) :: mode = 'reference'
  character(len=512) :: start = 'first', start2 = '', outfile = 'forge_traj.txt', refout = ''
  character(len=512) :: relaxtable = ''
  integer(c_int32_t) :: occ = 0, mix = 0, prec = 0
  integer(dp) :: nenforceC3 = 0, nenforceValley = 0
  integer(c_int64_t), allocatable, target :: nC3pairs(:), rotC3(:), rotV(:)
  integer(c_int64_t), allocatable :: nC2pairs(:)
  integer(c_int32_t) :: symflags = 0
  real(dp) :: kT = 0.001_dp, beta = 0.3_dp, etol = 0.0_dp, switch_de = 0.0_dp
  integer(dp) :: nnc, it, n, nFermiLevel, nOcc, nPart, icount, i, nU
  integer(dp), allocatable :: nUnitCell_1(:), nUnitCell_2(:), nMomentaComponents(:, :)
  integer(dp), allocatable :: nSortedMomenta(:, :, :)
  real(dp), allocatable :: Coords(:, :), MomentaValues(:, :, :), LongRange(:, :)
  real(dp), allocatable :: Potential(:, :), Density(:, :), DensityIn(:, :), DensitySub(:)
  real(dp), allocatable :: Bands(:, :, :), SortedEnergies(:, :), traj(:, :)
  complex(dp), allocatable, target :: zFock(:, :, :, :)
  complex(dp), allocatable :: zFockIn(:, :, :, :), zFockBulk(:, :, :, :)
  complex(dp), allocatable :: zEigenvectors(:, :, :, :), zSorted(:, :, :)
  real(dp) :: t1(2), t2(2), t3(2), g1(2), g12(2), tn(2, 3), RotMatrix(2, 2)
  real(dp) :: KineticEnergy(2), FockEnergy(2), HartreeEnergy, HubbardEnergy, TotalEnergy
  real(dp) :: KineticEnergyIn(2), FockEnergyIn(2), HartreeEnergyIn, HubbardEnergyIn
  real(dp) :: TotalEnergyIn, FermiEnergy, DegFactor, Step, DensityConvergence
  type(c_ptr) :: ctx, ctx2 = c_null_ptr, pair = c_null_ptr
  type(forge_scf_options_t) :: opt
  integer(c_int64_t) :: done
  integer :: fout

  call parse_args()
  if (len_trim(relaxtable) == 0) call default_relaxtable()
  call setup_init()
  n = ndim
  nU = nUpper - nLower + 1

  ! ---- geometry (Main.f90:119-160)
  nnc = NumberNeighborCells(numI)
  allocate(nUnitCell_1(nnc), nUnitCell_2(nnc))
  call OrderNeighborCells(nUnitCell_1, nUnitCell_2, numI, nnc)
  allocate(Coords(n, 3), MomentaValues(numk, numk, 2), nMomentaComponents(Nk, 2))
  call build_cell(Coords, t1, t2, t3, g1, g12)
  ! LatticeRelaxation for bilayers (Main.f90:156-160)
  if (nrelax == 1 .and. nlayers == 2) call lattice_relaxation(Coords, g1, g12, trim(relaxtable))
  call SampleBZ(nMomentaComponents, MomentaValues, numk, g1, g12)
  tn = reshape([t1, t2, t3], [2, 3])
  RotMatrix = rotation_matrix()
  ! C2_RelatedPoints (Main.f90:162)
  allocate(nC2pairs(n))
  nC2pairs = 0
  if (symflags /= 0) call C2_RelatedPoints(nC2pairs, Coords, n)
  if (nenforceC3 == 1) then
    ! C3_RelatedPoints (Main.f90:163)
    allocate(nC3pairs(n), rotC3(nlayers))
    rotC3 = RotateLayers
    call C3_RelatedPoints(nC3pairs, Coords, n)
  end if
  if (trim(mode) == 'geometry') then
    call write_geometry()
  else
    call run_engine()
  end if

contains

  subroutine run_engine()
    if (forge_device_count() < 1) then
      write(*, *) 'forge_main: no HIP device'
      error stop 3
    end if
    call forge_check(forge_ctx_create(setup_params(nnc), Coords, tn, MomentaValues, &
                                      nMomentaComponents, nUnitCell_1, nUnitCell_2, ctx), &
                     'forge_ctx_create')
    call forge_attach(ctx, n, nnc, alpha, Delta, Coords, tn)

    allocate(zFock(n, n, nnc, numS), zFockIn(n, n, nnc, numS))
    allocate(Density(n, numS), DensityIn(n, numS), DensitySub(n), Potential(n, numS))
    DensitySub = 0.5_dp
    Potential = 0.0_dp
    zFock = (0.0_dp, 0.0_dp)
    if (numS == 2 .and. (occ > 1 .or. mix /= 0)) then
      write(*, *) 'forge_main: numS = 2 runs with the T = 0 rule (occ=0 or 1) and mix=0'
      error stop 1
    end if
    if (numS == 2 .and. nenforceValley == 1) then
      write(*, *) 'forge_main: the valley projection runs with numS = 1'
      error stop 1
    end if
    if (trim(start) /= 'first') call read_fock(trim(start), 1_dp)
    if (numS == 2 .and. trim(start) /= 'first') then
      if (len_trim(start2) == 0) then
        write(*, *) 'forge_main: numS = 2 from checkpoints needs start2= (spin 2)'
        error stop 1
      end if
      call read_fock(trim(start2), 2_dp)
    end if

    open(newunit=fout, file=trim(outfile), status='replace')
    write(fout, '(A,I0,A,I0,A,I0,A,F4.1,A,A)') '# forge_main ntheta=', ntheta, ' nlayers=', nlayers, &
        ' numk=', numk, ' nfilling=', nfilling, ' mode=', trim(mode)

    if (trim(mode) == 'device') then
      opt%occ_mode = occ; opt%mix_mode = mix; opt%kT = kT; opt%beta = beta; opt%etol = etol
      opt%precision = prec; opt%switch_de = switch_de
      if (numS == 2) then
        ! spin 2's context: the same system; the pair keeps both spins on the device
        call forge_check(forge_ctx_create(setup_params(nnc), Coords, tn, MomentaValues, &
                                          nMomentaComponents, nUnitCell_1, nUnitCell_2, ctx2), &
                         'forge_ctx_create')
      end if
      call device_flags(ctx)
      if (numS == 2) call device_flags(ctx2)
      allocate(traj(12, itmax))
      if (numS == 2) then
        call forge_check(forge_scf2_create(ctx, ctx2, pair), 'forge_scf2_create')
        if (trim(start) == 'first') then
          call forge_check(forge_scf2_init(pair, c_null_ptr), 'forge_scf2_init')
        else
          call forge_check(forge_scf2_init(pair, c_loc(zFock)), 'forge_scf2_init')
        end if
        call forge_check(forge_scf2_run(pair, int(itmax, c_int64_t), traj, done), 'forge_scf2_run')
        call forge_check(forge_scf2_destroy(pair), 'forge_scf2_destroy')
        call forge_check(forge_ctx_destroy(ctx2), 'forge_ctx_destroy')
      else
        if (trim(start) == 'first') then
          call forge_check(forge_scf_init(ctx, c_null_ptr), 'forge_scf_init')
        else
          call forge_check(forge_scf_init(ctx, c_loc(zFock)), 'forge_scf_init')
        end if
        call forge_check(forge_scf_run(ctx, int(itmax, c_int64_t), traj, done), 'forge_scf_run')
      end if
      write(fout, '(A)') '# Energy: it Etot 2Ekin 2EF EH EU'
      do i = 1, done
        write(fout, '(F5.1,3X,5(F24.9,3X))') traj(1, i), traj(2:6, i)
      end do
      write(fout, '(A)') '# Convergence: it maxdDensity Step dE'
      do i = 1, done
        write(fout, '(F5.1,3X,F9.6,3X,F5.3,3X,F24.9,3X)') traj(1, i), traj(7, i), traj(8, i), traj(9, i)
      end do
      if (len_trim(refout) > 0) &
        call write_reference_files(traj(1:6, 1:done), &
                                   reshape([traj(1, 1:done), traj(7, 1:done), traj(8, 1:done), &
                                            traj(9, 1:done)], [int(done), 4], order=[1, 2]), &
                                   int(done, dp))
    else if (numS == 2) then
      call reference_sequence_2spins()
    else
      call reference_sequence()
    end if
    close(fout)
    call forge_check(forge_ctx_destroy(ctx), 'forge_ctx_destroy')
  end subroutine run_engine

  ! options, C3 set, symmetry and valley flags of a device-resident context
  subroutine device_flags(c)
    type(c_ptr), intent(in) :: c
    call forge_check(forge_scf_set_options(c, opt), 'forge_scf_set_options')
    if (nenforceC3 == 1) &
      call forge_check(forge_scf_set_c3(c, c_loc(rotC3), c_loc(nC3pairs)), 'forge_scf_set_c3')
    if (symflags /= 0) &
      call forge_check(forge_scf_set_symmetry(c, symflags, nC2pairs), 'forge_scf_set_symmetry')
    if (nenforceValley == 1) then
      if (.not. allocated(rotV)) then
        allocate(rotV(nlayers))
        rotV = RotateLayers
      end if
      call forge_check(forge_scf_set_valley(c, c_loc(rotV)), 'forge_scf_set_valley')
    end if
  end subroutine device_flags

  subroutine reference_sequence()
    real(dp), allocatable :: erow(:, :), crow(:, :)
    allocate(LongRange(n, n), zFockBulk(n, n, nnc, numS), Bands(numb, Nk, numS))
    allocate(zEigenvectors(n, numb, Nk, numS), zSorted(n, numb * Nk, numS))
    allocate(SortedEnergies(numb * Nk, numS), nSortedMomenta(numb * Nk, 2, numS))
    allocate(erow(6, itmax), crow(4, itmax))
    call LongRangeInteraction(LongRange, Coords, n, t1, t2)

    ! initial state: first diagonalisation (Main.f90:236-288) unless read
    if (trim(start) == 'first') then
      call solve_k(1_dp)
      call SortEnergies_1Spin(n, numk, numb, Nk, nMomentaComponents, zSorted(:, :, 1), &
                              SortedEnergies(:, 1), nSortedMomenta(:, :, 1), &
                              zEigenvectors(:, :, :, 1), Bands(:, :, 1))
      call fermi_level(SortedEnergies(:, 1), 1_dp)
      zFock(:, :, :, 1) = (0.0_dp, 0.0_dp)
      call getfock_k(1_dp, nOcc, nPart)
      zFock(:, :, :, 1) = (zFock(:, :, :, 1) + zFockBulk(:, :, :, 1)) / cmplx(real(Nk, dp), 0.0_dp, dp)
    end if
    do i = 1, n
      Density(i, 1) = real(zFock(i, i, 1, 1), dp)
    end do
    ! energies and potential of the initial state (Main.f90:290-327)
    call GetKineticEnergy(KineticEnergy(1), Coords, zFock(:, :, :, 1), Delta, nUnitCell_1, &
                          nUnitCell_2, n, nnc, tn)
    call GetFockEnergy(FockEnergy(1), Coords, zFock(:, :, :, 1), nUnitCell_1, nUnitCell_2, n, nnc, tn)
    call GetHartreeHubbardEnergy(HartreeEnergy, HubbardEnergy, LongRange, Density(:, 1) - DensitySub, &
                                 Density(:, 1) - DensitySub, n)
    TotalEnergy = total_energy()
    call update_potential()
    DensityIn = Density
    zFockIn = zFock

    it = 0
    do while (it < itmax)
      it = it + 1
      call solve_k(1_dp)
      call SortEnergies_1Spin(n, numk, numb, Nk, nMomentaComponents, zSorted(:, :, 1), &
                              SortedEnergies(:, 1), nSortedMomenta(:, :, 1), &
                              zEigenvectors(:, :, :, 1), Bands(:, :, 1))
      call fermi_level(SortedEnergies(:, 1), numS)
      zFock(:, :, :, 1) = (0.0_dp, 0.0_dp)
      call getfock_k(1_dp, nOcc, nPart)
      if (nenforceValley == 1) then
        ! Main.f90:469-494: GetFock of the states and of their valley
        ! transform (same energies, so the same sort), averaged
        call ValleyTransform(zEigenvectors(:, :, :, 1), Coords, MomentaValues, nMomentaComponents, &
                             numb, n, numk, Nk, nnc, nUnitCell_1, nUnitCell_2, tn, RotMatrix)
        call SortEnergies_1Spin(n, numk, numb, Nk, nMomentaComponents, zSorted(:, :, 1), &
                                SortedEnergies(:, 1), nSortedMomenta(:, :, 1), &
                                zEigenvectors(:, :, :, 1), Bands(:, :, 1))
        call getfock_k(1_dp, nOcc, nPart)
        zFock(:, :, :, 1) = 0.5_dp * zFock(:, :, :, 1) + zFockBulk(:, :, :, 1)
        zFock(:, :, :, 1) = zFock(:, :, :, 1) / cmplx(real(Nk, dp), 0.0_dp, dp)
      else
        zFock(:, :, :, 1) = (zFock(:, :, :, 1) + zFockBulk(:, :, :, 1)) / cmplx(real(Nk, dp), 0.0_dp, dp)
      end if
      call symmetrize_spin(1_dp)

      HartreeEnergyIn = HartreeEnergy
      FockEnergyIn = FockEnergy
      HubbardEnergyIn = HubbardEnergy
      KineticEnergyIn = KineticEnergy
      TotalEnergyIn = TotalEnergy
      do i = 1, n
        Density(i, 1) = real(zFock(i, i, 1, 1), dp)
      end do
      if (mix == 0) then
        call OptimalStep(Step, HartreeEnergy, FockEnergy, HubbardEnergy, KineticEnergy, zFock, &
                         zFockIn, Density, DensityIn, DensitySub, n, nnc, Coords, LongRange, Delta, &
                         nUnitCell_1, nUnitCell_2, tn)
        TotalEnergy = total_energy()
        zFock = Step * zFock + (1.0_dp - Step) * zFockIn
      else
        Step = beta
        zFock = Step * zFock + (1.0_dp - Step) * zFockIn
        do i = 1, n
          Density(i, 1) = real(zFock(i, i, 1, 1), dp)
        end do
        call GetKineticEnergy(KineticEnergy(1), Coords, zFock(:, :, :, 1), Delta, nUnitCell_1, &
                              nUnitCell_2, n, nnc, tn)
        call GetFockEnergy(FockEnergy(1), Coords, zFock(:, :, :, 1), nUnitCell_1, nUnitCell_2, n, &
                           nnc, tn)
        call GetHartreeHubbardEnergy(HartreeEnergy, HubbardEnergy, LongRange, &
                                     Density(:, 1) - DensitySub, Density(:, 1) - DensitySub, n)
        TotalEnergy = total_energy()
      end if
      do i = 1, n
        Density(i, 1) = real(zFock(i, i, 1, 1), dp)
      end do
      DensityConvergence = maxval(abs(Density(:, 1) - DensityIn(:, 1)))
      DensityIn = Density
      zFockIn = zFock
      call update_potential()
      erow(:, it) = [real(it, dp), TotalEnergyIn, 2.0_dp * KineticEnergyIn(1), &
                     2.0_dp * FockEnergyIn(1), HartreeEnergyIn, HubbardEnergyIn]
      crow(:, it) = [real(it, dp), DensityConvergence, Step, TotalEnergyIn - TotalEnergy]
      write(*, '(A,I0,A,F20.9,A,ES10.3)') 'it=', it, ' E=', TotalEnergy, ' dE=', &
          TotalEnergyIn - TotalEnergy
      if (etol > 0.0_dp .and. abs(TotalEnergyIn - TotalEnergy) < etol) exit
    end do
    write(fout, '(A)') '# Energy: it Etot 2Ekin 2EF EH EU'
    do i = 1, min(it, itmax)
      write(fout, '(F5.1,3X,5(F24.9,3X))') erow(:, i)
    end do
    write(fout, '(A)') '# Convergence: it maxdDensity Step dE'
    do i = 1, min(it, itmax)
      write(fout, '(F5.1,3X,F9.6,3X,F5.3,3X,F24.9,3X)') crow(:, i)
    end do
    if (len_trim(refout) > 0) &
      call write_reference_files(erow(:, 1:min(it, itmax)), transpose(crow(:, 1:min(it, itmax))), &
                                 min(it, itmax))
  end subroutine reference_sequence

  ! The run-parameter string of the reference's file names (Main.f90:180-183;
  ! no relaxation, state label SSS, dp = 8).
  function run_parameters() result(par)
    character(len=210) :: par
    character(len=1) :: dop
    dop = '+'
    if (nfilling < 0.0_dp) dop = '-'
    write(par, '(A1,A3,A8,A1,F3.1,A5,I0,A2,I0,A8,I0,A6,I0,A4,F0.1,A2,F4.2,A7,I0,A3,F0.1,A6,F0.3,&
               &A5,I0,A5,I0,A5,I0,A3,I0,A4)') &
        '-', 'SSS', '-filling', dop, abs(nfilling), '-numS', numS, '-i', ntheta, '-nlayers', nlayers, &
        '-relax', nrelax, '-eps', epsilon, '-U', U, '-screen', nscreen, '-xi', xi * 0.246_dp, &
        '-delta', Delta, '-numI', numI, '-numC', numC, '-numk', numk, '-dp', 8, '.dat'
  end function run_parameters

  ! refout//Energy<parameters> and refout//Convergence<parameters> in the
  ! reference's record formats (Main.f90:616-631).  e(6, nr): it, E_tot,
  ! 2 Ekin, 2 EF, EH, EU; c(nr, 4): it, max|dDensity|, Step, E_in - E_new.
  subroutine write_reference_files(e, c, nr)
    real(dp), intent(in) :: e(:, :), c(:, :)
    integer(dp), intent(in) :: nr
    integer :: ue, uc
    integer(dp) :: r
    character(len=210) :: par
    par = run_parameters()
    open(newunit=ue, file=trim(refout)//'Energy'//trim(par), status='replace')
    open(newunit=uc, file=trim(refout)//'Convergence'//trim(par), status='replace')
    do r = 1, nr
      write(ue, '(F5.1,3X,5(F24.9,3X))') e(:, r)
      write(uc, '(F5.1,3X,F9.6,3X,F5.3,3X,F24.9,3X)') c(r, :)
    end do
    close(ue)
    close(uc)
  end subroutine write_reference_files

  ! Fermi level and degenerate shell (Main.f90:396-414): states within
  ! EnergyTolerance of E_F share the remaining electrons equally.
  subroutine fermi_level(S, mult)
    real(dp), intent(in) :: S(:)
    integer(dp), intent(in) :: mult
    nFermiLevel = mult * nint(real(Nk * NeutralityPoint, dp) + real(Nk, dp) / 2.0_dp * nfilling, dp)
    FermiEnergy = (S(nFermiLevel) + S(nFermiLevel + 1)) / 2.0_dp
    nOcc = 0
    nPart = 0
    do icount = 1, Nk * numb
      if (S(icount) - FermiEnergy < -EnergyTolerance) then
        nOcc = nOcc + 1
      else if (abs(S(icount) - FermiEnergy) < EnergyTolerance) then
        nOcc = nOcc + 1
        nPart = nPart + 1
      end if
    end do
    DegFactor = 1.0_dp / real(nPart, dp) * real(nFermiLevel + nPart - nOcc, dp)
  end subroutine fermi_level

  real(dp) function total_energy()
    total_energy = alpha * sum(FockEnergy(1:numS)) * real(3 - numS, dp) + &
                   sum(KineticEnergy(1:numS)) * real(3 - numS, dp) + U * HubbardEnergy + &
                   alphaH * HartreeEnergy
  end function total_energy

  ! V = 2 alphaH L (D - 1/2) + U (D - 1/2)   (Main.f90:577-585)
  subroutine update_potential()
    Potential(:, 1) = 2.0_dp * alphaH * matmul(LongRange, Density(:, 1) - DensitySub) + &
                      U * (Density(:, 1) - DensitySub)
  end subroutine update_potential

  ! Density matrix in the reference's checkpoint layout (Main.f90:187-235):
  ! direct access, one complex(8) per record; cell 1 as its lower triangle by
  ! rows (the upper triangle is the conjugate), then cells 2..nnc in full.
  subroutine read_fock(path, sp)
    character(*), intent(in) :: path
    integer(dp), intent(in) :: sp
    integer :: fu, ios
    integer(dp) :: rec, a, b, m
    complex(dp) :: z
    open(newunit=fu, file=path, form='unformatted', status='old', access='direct', recl=16, &
         iostat=ios)
    if (ios /= 0) then
      write(*, *) 'forge_main: cannot open ', path
      error stop 4
    end if
    rec = 0
    do a = 1, n
      do b = 1, a - 1
        rec = rec + 1
        read(fu, rec=rec) z
        zFock(a, b, 1, sp) = z
        zFock(b, a, 1, sp) = conjg(z)
      end do
      rec = rec + 1
      read(fu, rec=rec) z
      zFock(a, a, 1, sp) = z
    end do
    do m = 2, nnc
      do a = 1, n
        do b = 1, n
          rec = rec + 1
          read(fu, rec=rec) z
          zFock(a, b, m, sp) = z
        end do
      end do
    end do
    close(fu)
  end subroutine read_fock

  ! mode=geometry: the host-side geometry as text (17 significant digits,
  ! exact on read-back), without touching the device
  subroutine write_geometry()
    open(newunit=fout, file=trim(outfile), status='replace')
    write(fout, '(A,2I8)') '# Coords', n, 3
    do i = 1, n
      write(fout, '(3ES26.17E3)') Coords(i, :)
    end do
    write(fout, '(A)') '# tn'
    write(fout, '(6ES26.17E3)') tn
    write(fout, '(A)') '# g1 g12'
    write(fout, '(4ES26.17E3)') g1, g12
    write(fout, '(A)') '# MomentaValues'
    write(fout, '(4ES26.17E3)') MomentaValues
    write(fout, '(A)') '# nMomentaComponents'
    write(fout, '(2I8)') (nMomentaComponents(i, :), i = 1, Nk)
    write(fout, '(A)') '# nUnitCell'
    write(fout, '(2I8)') (nUnitCell_1(i), nUnitCell_2(i), i = 1, nnc)
    write(fout, '(A,A)') '# parameters ', trim(run_parameters())
    close(fout)
  end subroutine write_geometry

  ! Solve / Solve_C3 and GetFock / GetFock_C3 as Main.f90:244-250, 273-276,
  ! 375-378 and 500-505 choose between them.
  subroutine solve_k(s)
    integer(dp), intent(in) :: s
    if (nenforceC3 == 1) then
      call Solve_C3(zFockBulk(:, :, :, s), zFock(:, :, :, s), Potential(:, s), alpha, Bands(:, :, s), &
                    zEigenvectors(:, :, :, s), Coords, MomentaValues, nMomentaComponents, nLower, &
                    nUpper, n, rotC3, numk, Nk, nnc, nUnitCell_1, nUnitCell_2, tn, g1, g12, nC3pairs)
    else
      call Solve(zFockBulk(:, :, :, s), zFock(:, :, :, s), Potential(:, s), alpha, Bands(:, :, s), &
                 zEigenvectors(:, :, :, s), Coords, MomentaValues, nMomentaComponents, nLower, &
                 nUpper, n, numk, Nk, nnc, nUnitCell_1, nUnitCell_2, tn)
    end if
  end subroutine solve_k

  subroutine getfock_k(s, no, np)
    integer(dp), intent(in) :: s, no, np
    if (nenforceC3 == 1) then
      call GetFock_C3(zFock(:, :, :, s), zSorted(:, :, s), n, rotC3, numb, numk, DegFactor, no, &
                      np, Coords, MomentaValues, nnc, nUnitCell_1, nUnitCell_2, &
                      nSortedMomenta(:, :, s), nC3pairs)
    else
      call GetFock(zFock(:, :, :, s), zSorted(:, :, s), n, numb, numk, DegFactor, no, np, &
                   Coords, MomentaValues, nnc, nUnitCell_1, nUnitCell_2, nSortedMomenta(:, :, s))
    end if
  end subroutine getfock_k

  ! The two-spin sequence (numS = 2, Main.f90:236-679): per spin Solve and
  ! SortEnergies_1Spin; SortEnergies_2Spins and the two-spin Fermi block
  ! (Main.f90:416-458); per spin GetFock with its own occupied counts and the
  ! shared DegFactor; OptimalStep over both spins; the spin-exchanged Hubbard
  ! potential (Main.f90:586-595).  The first diagonalisation fills spin 1 and
  ! copies it to spin 2 (Main.f90:238-288).
  subroutine reference_sequence_2spins()
    real(dp), allocatable :: erow(:, :), crow(:, :), S2(:, :)
    integer(dp) :: s, q, W(2, 2), nO(2), nP(2)
    real(dp) :: dconv
    allocate(LongRange(n, n), zFockBulk(n, n, nnc, 2), Bands(numb, Nk, 2))
    allocate(zEigenvectors(n, numb, Nk, 2), zSorted(n, numb * Nk, 2))
    allocate(SortedEnergies(numb * Nk, 2), nSortedMomenta(numb * Nk, 2, 2), S2(2 * numb * Nk, 2))
    allocate(erow(6, itmax), crow(4, itmax))
    call LongRangeInteraction(LongRange, Coords, n, t1, t2)
    if (trim(start) == 'first') then
      call solve_k(1_dp)
      call SortEnergies_1Spin(n, numk, numb, Nk, nMomentaComponents, zSorted(:, :, 1), &
                              SortedEnergies(:, 1), nSortedMomenta(:, :, 1), &
                              zEigenvectors(:, :, :, 1), Bands(:, :, 1))
      call fermi_level(SortedEnergies(:, 1), 1_dp)
      zFock(:, :, :, 1) = (0.0_dp, 0.0_dp)
      call getfock_k(1_dp, nOcc, nPart)
      zFock(:, :, :, 1) = (zFock(:, :, :, 1) + zFockBulk(:, :, :, 1)) / cmplx(real(Nk, dp), 0.0_dp, dp)
      zFock(:, :, :, 2) = zFock(:, :, :, 1)
    end if
    call spin_densities()
    do s = 1, 2
      call GetKineticEnergy(KineticEnergy(s), Coords, zFock(:, :, :, s), Delta, nUnitCell_1, &
                            nUnitCell_2, n, nnc, tn)
      call GetFockEnergy(FockEnergy(s), Coords, zFock(:, :, :, s), nUnitCell_1, nUnitCell_2, n, nnc, tn)
    end do
    call GetHartreeHubbardEnergy(HartreeEnergy, HubbardEnergy, LongRange, Density(:, 1) - DensitySub, &
                                 Density(:, 2) - DensitySub, n)
    TotalEnergy = total_energy()
    call update_potential_2spins()
    DensityIn = Density
    zFockIn = zFock
    W = reshape([NeutralityPoint - 1, NeutralityPoint + 3, NeutralityPoint - 1, NeutralityPoint + 3], &
                [2, 2])
    it = 0
    do while (it < itmax)
      it = it + 1
      do s = 1, 2
        call solve_k(s)
        call SortEnergies_1Spin(n, numk, numb, Nk, nMomentaComponents, zSorted(:, :, s), &
                                SortedEnergies(:, s), nSortedMomenta(:, :, s), &
                                zEigenvectors(:, :, :, s), Bands(:, :, s))
      end do
      call SortEnergies_2Spins(numk, numb, W, SortedEnergies, S2)
      nFermiLevel = 2 * nint(real(Nk * NeutralityPoint, dp) + real(Nk, dp) / 2.0_dp * nfilling, dp)
      FermiEnergy = (S2(nFermiLevel, 1) + S2(nFermiLevel + 1, 1)) / 2.0_dp
      nO = 0
      nP = 0
      do q = 1, 2 * numb * Nk
        s = nint(S2(q, 2), dp)
        if (S2(q, 1) - FermiEnergy < -EnergyTolerance) then
          nO(s) = nO(s) + 1
        else if (abs(S2(q, 1) - FermiEnergy) < EnergyTolerance) then
          nO(s) = nO(s) + 1
          nP(s) = nP(s) + 1
        end if
      end do
      if (nP(1) /= 0 .or. nP(2) /= 0) then
        DegFactor = 1.0_dp / real(nP(1) + nP(2), dp) * &
                    real(nFermiLevel + nP(1) + nP(2) - nO(1) - nO(2), dp)
      else
        DegFactor = 0.0_dp
      end if
      do s = 1, 2
        zFock(:, :, :, s) = (0.0_dp, 0.0_dp)
        call getfock_k(s, nO(s), nP(s))
        zFock(:, :, :, s) = (zFock(:, :, :, s) + zFockBulk(:, :, :, s)) / cmplx(real(Nk, dp), 0.0_dp, dp)
        call symmetrize_spin(s)
      end do
      HartreeEnergyIn = HartreeEnergy
      FockEnergyIn = FockEnergy
      HubbardEnergyIn = HubbardEnergy
      KineticEnergyIn = KineticEnergy
      TotalEnergyIn = TotalEnergy
      call spin_densities()
      call OptimalStep(Step, HartreeEnergy, FockEnergy, HubbardEnergy, KineticEnergy, zFock, &
                       zFockIn, Density, DensityIn, DensitySub, n, nnc, Coords, LongRange, Delta, &
                       nUnitCell_1, nUnitCell_2, tn)
      TotalEnergy = total_energy()
      zFock = Step * zFock + (1.0_dp - Step) * zFockIn
      call spin_densities()
      dconv = maxval(abs(Density - DensityIn))
      DensityIn = Density
      zFockIn = zFock
      call update_potential_2spins()
      erow(:, it) = [real(it, dp), TotalEnergyIn, KineticEnergyIn(1) + KineticEnergyIn(2), &
                     FockEnergyIn(1) + FockEnergyIn(2), HartreeEnergyIn, HubbardEnergyIn]
      crow(:, it) = [real(it, dp), dconv, Step, TotalEnergyIn - TotalEnergy]
      write(*, '(A,I0,A,F20.9,A,ES10.3)') 'it=', it, ' E=', TotalEnergy, ' dE=', &
          TotalEnergyIn - TotalEnergy
      if (etol > 0.0_dp .and. abs(TotalEnergyIn - TotalEnergy) < etol) exit
    end do
    write(fout, '(A)') '# Energy: it Etot Ekin EF EH EU'
    do i = 1, min(it, itmax)
      write(fout, '(F5.1,3X,5(F24.9,3X))') erow(:, i)
    end do
    write(fout, '(A)') '# Convergence: it maxdDensity Step dE'
    do i = 1, min(it, itmax)
      write(fout, '(F5.1,3X,F9.6,3X,F5.3,3X,F24.9,3X)') crow(:, i)
    end do
    if (len_trim(refout) > 0) &
      call write_reference_files(erow(:, 1:min(it, itmax)), transpose(crow(:, 1:min(it, itmax))), &
                                 min(it, itmax))
  end subroutine reference_sequence_2spins

  ! nenforceT / nenforceC2 / nenforceC2T on spin s's new density
  ! (Main.f90:514-534), on the device
  subroutine symmetrize_spin(s)
    integer(dp), intent(in) :: s
    if (symflags /= 0) &
      call forge_check(forge_symmetrize(ctx, zFock(:, :, :, s), symflags, nC2pairs), 'symmetrize')
  end subroutine symmetrize_spin

  subroutine spin_densities()
    integer(dp) :: s
    do s = 1, numS
      do i = 1, n
        Density(i, s) = real(zFock(i, i, 1, s), dp)
      end do
    end do
  end subroutine spin_densities

  ! V_1 = alphaH L (D_1 + D_2 - 1) + U (D_2 - 1/2), V_2 likewise with D_1
  ! (Main.f90:586-595)
  subroutine update_potential_2spins()
    real(dp) :: base(n)
    base = alphaH * matmul(LongRange, Density(:, 1) + Density(:, 2) - 2.0_dp * DensitySub)
    Potential(:, 2) = base + U * (Density(:, 1) - DensitySub)
    Potential(:, 1) = base + U * (Density(:, 2) - DensitySub)
  end subroutine update_potential_2spins

  ! forge_amd/data/relaxation_fourier.txt next to the executable's directory
  subroutine default_relaxtable()
    character(len=512) :: exe
    integer :: k
    call get_command_argument(0, exe)
    k = index(exe, '/', back=.true.)
    if (k > 0) then
      relaxtable = exe(1:k)//'../data/relaxation_fourier.txt'
    else
      relaxtable = '../data/relaxation_fourier.txt'
    end if
  end subroutine default_relaxtable

  subroutine parse_args()
    integer :: k, eq
    character(len=512) :: arg, key, val
    do k = 1, command_argument_count()
      call get_command_argument(k, arg)
      eq = index(arg, '=')
      if (eq == 0) then
        write(*, *) 'forge_main: expected key=value, got ', trim(arg)
        error stop 1
      end if
      key = arg(1:eq - 1)
      val = arg(eq + 1:)
      select case (trim(key))
      case ('ntheta'); read(val, *) ntheta
      case ('nlayers'); read(val, *) nlayers
      case ('numk'); read(val, *) numk
      case ('nfilling'); read(val, *) nfilling
      case ('itmax'); read(val, *) itmax
      case ('xi_nm'); read(val, *) xi; xi = xi / 0.246_dp
      case ('epsilon'); read(val, *) epsilon
      case ('U'); read(val, *) U
      case ('Delta'); read(val, *) Delta
      case ('mode'); mode = val
      case ('start'); start = val
      case ('occ'); read(val, *) occ
      case ('mix'); read(val, *) mix
      case ('kT'); read(val, *) kT
      case ('beta'); read(val, *) beta
      case ('etol'); read(val, *) etol
      case ('precision'); read(val, *) prec
      case ('switch_de'); read(val, *) switch_de
      case ('out'); outfile = val
      case ('refout'); refout = val
      case ('numS'); read(val, *) numS
      case ('start2'); start2 = val
      case ('nrelax'); read(val, *) nrelax
      case ('relaxtable'); relaxtable = val
      case ('T'); read(val, *) i; if (i == 1) symflags = ior(symflags, 1_c_int32_t)
      case ('C2'); read(val, *) i; if (i == 1) symflags = ior(symflags, 2_c_int32_t)
      case ('C2T'); read(val, *) i; if (i == 1) symflags = ior(symflags, 4_c_int32_t)
      case ('c3'); read(val, *) nenforceC3
      case ('V'); read(val, *) nenforceValley
      case default
        write(*, *) 'forge_main: unknown key ', trim(key)
        error stop 1
      end select
    end do
    if (numS /= 1 .and. numS /= 2) then
      write(*, *) 'forge_main: numS must be 1 or 2'
      error stop 1
    end if
    if (trim(mode) /= 'reference' .and. trim(mode) /= 'device' .and. trim(mode) /= 'geometry') then
      write(*, *) 'forge_main: mode must be reference, device or geometry'
      error stop 1
    end if
  end subroutine parse_args

end program forge_main
