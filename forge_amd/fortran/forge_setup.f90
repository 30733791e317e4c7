! forge_setup.f90 -- run-time model parameters of the Fortran host.
!
! Same names and meaning as the reference's Setup module (Setup.f90:6-74);
! here they are variables set once from the command line by setup_init, so a
! single executable covers every cell size, k grid and filling.  numS = 1
! (spin singlet) and numS = 2 (independent up and down spins) are covered;
! symmetry enforcement other than C3, and relaxation, are not.
module forge_setup
  use iso_c_binding, only: c_int64_t
  use forge_c, only: forge_params_t
  implicit none
  integer, parameter :: dp = 8
  integer(dp) :: numS = 1
  real(dp), parameter :: pi = 4.0_dp * atan(1.0_dp)
  real(dp), parameter :: a1(2) = [0.5_dp, sqrt(3.0_dp) * 0.5_dp]
  real(dp), parameter :: a2(2) = [-0.5_dp, sqrt(3.0_dp) * 0.5_dp]
  real(dp), parameter :: a0 = 1.0_dp / sqrt(3.0_dp)   ! carbon-carbon distance / a
  real(dp), parameter :: r0 = 0.184_dp                ! hopping decay constant
  real(dp), parameter :: d0 = 1.35772_dp              ! interlayer hopping distance / a
  real(dp), parameter :: tz = 1.35772_dp              ! layer separation / a
  real(dp), parameter :: EnergyTolerance = 0.0001_dp  ! degenerate-shell width (eV)
  real(dp), parameter :: StepAlternative = 1.0_dp     ! ODA step when no minimum exists
  integer(dp), parameter :: fphase = 0

  integer(dp) :: nlayers = 2, ntheta = 5, numk = 6, numI = 1, numC = 3, nscreen = 2
  integer(dp) :: itmax = 2, nRead = 0, nrelax = 0
  integer(dp), allocatable :: RotateLayers(:)
  real(dp) :: nfilling = 0.0_dp, U = 4.0_dp, epsilon = 10.0_dp, Delta = 0.0_dp
  real(dp) :: xi = 10.0_dp / 0.246_dp
  real(dp) :: alpha, alphaH
  integer(dp) :: ndim, numb, ncb, nLower, nUpper, NeutralityPoint, Nk

contains

  ! Derived parameters, as Setup.f90 defines them from the primary ones.
  subroutine setup_init()
    integer(dp) :: l
    ndim = nlayers * 2 * (ntheta**2 + (ntheta + 1)**2 + (ntheta + 1) * ntheta)
    numb = min(20_dp, ndim)
    ncb = numb / 2
    nLower = (ndim - 2 * numb + 2 * ncb) / 2 + 1
    nUpper = (ndim + 2 * ncb) / 2
    NeutralityPoint = (ndim / 2 - nLower) + 1
    Nk = numk * numk
    alpha = 5.853_dp / epsilon * 1.0_dp
    alphaH = 5.853_dp / epsilon * 1.0_dp
    if (allocated(RotateLayers)) deallocate(RotateLayers)
    allocate(RotateLayers(nlayers))
    ! alternating twist: -1, +1, -1, ...
    do l = 1, nlayers
      RotateLayers(l) = merge(-1_dp, 1_dp, mod(l, 2_dp) == 1)
    end do
  end subroutine setup_init

  ! The engine's copy of the parameters (struct forge_params).
  function setup_params(nnc) result(p)
    integer(dp), intent(in) :: nnc
    type(forge_params_t) :: p
    p%ndim = ndim; p%nlayers = nlayers; p%numk = numk; p%numI = numI; p%numC = numC
    p%nscreen = nscreen; p%fphase = fphase; p%nnc = nnc
    p%nLower = nLower; p%nUpper = nUpper; p%numb = numb
    p%nfilling = nfilling; p%U = U; p%alpha = alpha; p%alphaH = alphaH; p%Delta = Delta
    p%xi = xi; p%tz = tz; p%a0 = a0; p%r0 = r0; p%d0 = d0
    p%energy_tol = EnergyTolerance; p%step_alt = StepAlternative
  end function setup_params

end module forge_setup
