! forge_c.f90 -- ISO_C_BINDING interfaces of libforge_hip.so (include/forge_hf.h).
!
! This is the layer a Fortran host uses to call the MI355X engine.  Arrays are
! passed as the host's own column-major arrays (assumed size), int64 is
! integer(dp), complex(c_double_complex) is complex(dp).  Every function
! returns 0 on success, > 0 for an eigensolver failure count, < 0 for an
! argument or HIP error (text in forge_error_message()).
module forge_c
  use iso_c_binding
  implicit none
  private
  public :: forge_params_t, forge_scf_options_t, forge_error_message
  public :: forge_device_count, forge_ctx_create, forge_ctx_destroy, forge_longrange
  public :: forge_tb_hamiltonian, forge_hamiltonian_hf, forge_zdiagred, forge_solve
  public :: forge_getfock, forge_energies, forge_scf_set_options, forge_scf_init
  public :: forge_scf_run, forge_scf_get_density, forge_scf_timings
  public :: forge_sweep_create, forge_sweep_run, forge_sweep_destroy
  public :: forge_scf2_create, forge_scf2_init, forge_scf2_run, forge_scf2_get_density, forge_scf2_destroy
  public :: forge_c3_pairs, forge_solve_c3, forge_getfock_c3, forge_scf_set_c3
  public :: forge_c2_pairs, forge_symmetrize, forge_scf_set_symmetry, forge_hartree_hubbard
  public :: forge_scf_init_guess, forge_scf_init_guess_update, forge_scf_init_update
  public :: forge_mlp_param_count, forge_mlp_loss_grad, forge_mlp_train, forge_mlp_infer
  public :: forge_valley_phase, forge_valley_transform, forge_scf_set_valley
  public :: forge_z_vector_mul, forge_z_matrix_mul, forge_z_matrix_overlap, forge_c_matrix_mul
  public :: forge_zdiagred_noalloc

  ! struct forge_params (Setup.f90:6-74 at run time)
  type, bind(C) :: forge_params_t
    integer(c_int64_t) :: ndim, nlayers, numk, numI, numC, nscreen, fphase, nnc
    integer(c_int64_t) :: nLower, nUpper, numb
    real(c_double) :: nfilling, U, alpha, alphaH, Delta, xi, tz, a0, r0, d0
    real(c_double) :: energy_tol, step_alt
  end type forge_params_t

  ! struct forge_scf_options
  type, bind(C) :: forge_scf_options_t
    integer(c_int32_t) :: occ_mode = 0, mix_mode = 0
    real(c_double) :: kT = 0.0_c_double, beta = 0.3_c_double, etol = 0.0_c_double
    integer(c_int32_t) :: precision = 0, timing = 0
    real(c_double) :: switch_de = 0.0_c_double
    integer(c_int32_t) :: graph = 0
    integer(c_int32_t) :: reduction = 0
  end type forge_scf_options_t

  interface
    function forge_last_error() bind(C, name="forge_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function forge_last_error

    function forge_device_count() bind(C, name="forge_device_count") result(r)
      import :: c_int
      integer(c_int) :: r
    end function forge_device_count

    function forge_ctx_create(p, Coords, tn, MomentaValues, nMomentaComponents, nUnitCell_1, &
                              nUnitCell_2, ctx) bind(C, name="forge_ctx_create") result(r)
      import :: c_int, c_double, c_int64_t, c_ptr, forge_params_t
      type(forge_params_t), intent(in) :: p
      real(c_double), intent(in) :: Coords(*), tn(*), MomentaValues(*)
      integer(c_int64_t), intent(in) :: nMomentaComponents(*), nUnitCell_1(*), nUnitCell_2(*)
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: r
    end function forge_ctx_create

    function forge_ctx_destroy(ctx) bind(C, name="forge_ctx_destroy") result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: r
    end function forge_ctx_destroy

    function forge_longrange(ctx, LongRange) bind(C, name="forge_longrange") result(r)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: LongRange(*)
      integer(c_int) :: r
    end function forge_longrange

    function forge_tb_hamiltonian(ctx, zH, vk) bind(C, name="forge_tb_hamiltonian") result(r)
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: zH(*)
      real(c_double), intent(in) :: vk(2)
      integer(c_int) :: r
    end function forge_tb_hamiltonian

    function forge_hamiltonian_hf(ctx, zH, Potential, zFock, vk) &
        bind(C, name="forge_hamiltonian_hf") result(r)
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: zH(*)
      real(c_double), intent(in) :: Potential(*), vk(2)
      complex(c_double_complex), intent(in) :: zFock(*)
      integer(c_int) :: r
    end function forge_hamiltonian_hf

    function forge_zdiagred(A, n, eigenvalues, il, iu, info) bind(C, name="forge_zdiagred") result(r)
      import :: c_int, c_int32_t, c_int64_t, c_double, c_double_complex
      complex(c_double_complex), intent(inout) :: A(*)
      integer(c_int64_t), value :: n, il, iu
      real(c_double), intent(out) :: eigenvalues(*)
      integer(c_int32_t), intent(out) :: info
      integer(c_int) :: r
    end function forge_zdiagred

    ! forge_zdiagred without device allocation (n <= 512; no reference counterpart)
    function forge_zdiagred_noalloc(A, n, eigenvalues, il, iu, info) &
        bind(C, name="forge_zdiagred_noalloc") result(r)
      import :: c_int, c_int32_t, c_int64_t, c_double, c_double_complex
      complex(c_double_complex), intent(inout) :: A(*)
      integer(c_int64_t), value :: n, il, iu
      real(c_double), intent(out) :: eigenvalues(*)
      integer(c_int32_t), intent(out) :: info
      integer(c_int) :: r
    end function forge_zdiagred_noalloc

    ! products of the LAPACK shim (LapackRoutines.f90:119-236); z_factor may be
    ! c_null_ptr (1)
    function forge_z_vector_mul(C, A, n, k, z_factor) bind(C, name="forge_z_vector_mul") result(r)
      import :: c_int, c_int64_t, c_double_complex, c_ptr
      complex(c_double_complex), intent(inout) :: C(*)
      complex(c_double_complex), intent(in) :: A(*)
      integer(c_int64_t), value :: n, k
      type(c_ptr), value :: z_factor
      integer(c_int) :: r
    end function forge_z_vector_mul

    function forge_z_matrix_mul(C, A, B, m, k, n) bind(C, name="forge_z_matrix_mul") result(r)
      import :: c_int, c_int64_t, c_double_complex
      complex(c_double_complex), intent(out) :: C(*)
      complex(c_double_complex), intent(in) :: A(*), B(*)
      integer(c_int64_t), value :: m, k, n
      integer(c_int) :: r
    end function forge_z_matrix_mul

    function forge_z_matrix_overlap(C, A, B, m1, m0, z_factor) &
        bind(C, name="forge_z_matrix_overlap") result(r)
      import :: c_int, c_int64_t, c_double_complex, c_ptr
      complex(c_double_complex), intent(out) :: C(*)
      complex(c_double_complex), intent(in) :: A(*), B(*)
      integer(c_int64_t), value :: m1, m0
      type(c_ptr), value :: z_factor
      integer(c_int) :: r
    end function forge_z_matrix_overlap

    function forge_c_matrix_mul(C, A, B, m, k, n) bind(C, name="forge_c_matrix_mul") result(r)
      import :: c_int, c_int64_t, c_float_complex
      complex(c_float_complex), intent(out) :: C(*)
      complex(c_float_complex), intent(in) :: A(*), B(*)
      integer(c_int64_t), value :: m, k, n
      integer(c_int) :: r
    end function forge_c_matrix_mul

    function forge_solve(ctx, zFockBulk, zFock, Potential, Bands, zEigenvectors) &
        bind(C, name="forge_solve") result(r)
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: zFockBulk(*), zEigenvectors(*)
      complex(c_double_complex), intent(in) :: zFock(*)
      real(c_double), intent(in) :: Potential(*)
      real(c_double), intent(out) :: Bands(*)
      integer(c_int) :: r
    end function forge_solve

    function forge_getfock(ctx, zFock, zSortedEigenvectors, nSortedMomenta, DegFactor, OccStates, &
                           PartOccStates) bind(C, name="forge_getfock") result(r)
      import :: c_int, c_ptr, c_double, c_double_complex, c_int64_t
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(inout) :: zFock(*)
      complex(c_double_complex), intent(in) :: zSortedEigenvectors(*)
      integer(c_int64_t), intent(in) :: nSortedMomenta(*)
      real(c_double), value :: DegFactor
      integer(c_int64_t), value :: OccStates, PartOccStates
      integer(c_int) :: r
    end function forge_getfock

    function forge_energies(ctx, zFock, dens, out4) bind(C, name="forge_energies") result(r)
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: zFock(*)
      real(c_double), intent(in) :: dens(*)
      real(c_double), intent(out) :: out4(4)
      integer(c_int) :: r
    end function forge_energies

    function forge_c3_pairs(Coords, ndim, nlayers, nC3pairs) bind(C, name="forge_c3_pairs") &
        result(r)
      import :: c_double, c_int64_t
      real(c_double), intent(in) :: Coords(*)
      integer(c_int64_t), value :: ndim, nlayers
      integer(c_int64_t), intent(out) :: nC3pairs(*)
      integer(c_int64_t) :: r
    end function forge_c3_pairs
    function forge_solve_c3(ctx, zFockBulk, zFock, Potential, Bands, zEigenvectors, RotateLayers, &
                            nC3pairs) bind(C, name="forge_solve_c3") result(r)
      import :: c_int, c_ptr, c_double, c_double_complex, c_int64_t
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: zFockBulk(*), zEigenvectors(*)
      complex(c_double_complex), intent(in) :: zFock(*)
      real(c_double), intent(in) :: Potential(*)
      real(c_double), intent(out) :: Bands(*)
      integer(c_int64_t), intent(in) :: RotateLayers(*), nC3pairs(*)
      integer(c_int) :: r
    end function forge_solve_c3
    function forge_getfock_c3(ctx, zFock, zSortedEigenvectors, nSortedMomenta, DegFactor, &
                              OccStates, PartOccStates, RotateLayers, nC3pairs) &
        bind(C, name="forge_getfock_c3") result(r)
      import :: c_int, c_ptr, c_double, c_double_complex, c_int64_t
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(inout) :: zFock(*)
      complex(c_double_complex), intent(in) :: zSortedEigenvectors(*)
      integer(c_int64_t), intent(in) :: nSortedMomenta(*), RotateLayers(*), nC3pairs(*)
      real(c_double), value :: DegFactor
      integer(c_int64_t), value :: OccStates, PartOccStates
      integer(c_int) :: r
    end function forge_getfock_c3
    function forge_scf_set_c3(ctx, RotateLayers, nC3pairs) bind(C, name="forge_scf_set_c3") &
        result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: RotateLayers, nC3pairs
      integer(c_int) :: r
    end function forge_scf_set_c3
    function forge_c2_pairs(Coords, ndim, nlayers, nC2pairs) bind(C, name="forge_c2_pairs") &
        result(r)
      import :: c_double, c_int64_t
      real(c_double), intent(in) :: Coords(*)
      integer(c_int64_t), value :: ndim, nlayers
      integer(c_int64_t), intent(out) :: nC2pairs(*)
      integer(c_int64_t) :: r
    end function forge_c2_pairs
    function forge_symmetrize(ctx, zFock, flags, nC2pairs) bind(C, name="forge_symmetrize") &
        result(r)
      import :: c_int, c_ptr, c_double_complex, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(inout) :: zFock(*)
      integer(c_int32_t), value :: flags
      integer(c_int64_t), intent(in) :: nC2pairs(*)
      integer(c_int) :: r
    end function forge_symmetrize
    function forge_scf_set_symmetry(ctx, flags, nC2pairs) bind(C, name="forge_scf_set_symmetry") &
        result(r)
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: flags
      integer(c_int64_t), intent(in) :: nC2pairs(*)
      integer(c_int) :: r
    end function forge_scf_set_symmetry
    function forge_hartree_hubbard(ctx, dens1, dens2, out2) bind(C, name="forge_hartree_hubbard") &
        result(r)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: dens1(*), dens2(*)
      real(c_double), intent(out) :: out2(*)
      integer(c_int) :: r
    end function forge_hartree_hubbard
    function forge_scf_init_guess(ctx, rho_guess) bind(C, name="forge_scf_init_guess") result(r)
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: rho_guess(*)
      integer(c_int) :: r
    end function forge_scf_init_guess
    function forge_scf_init_guess_update(ctx, diag, nupd, row, col, cell, delta) &
        bind(C, name="forge_scf_init_guess_update") result(r)
      import :: c_int, c_ptr, c_double, c_double_complex, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: diag(*)
      integer(c_int64_t), value :: nupd
      integer(c_int64_t), intent(in) :: row(*), col(*), cell(*)
      complex(c_double_complex), intent(in) :: delta(*)
      integer(c_int) :: r
    end function forge_scf_init_guess_update
    function forge_scf_init_update(ctx, diag, nupd, row, col, cell, delta) &
        bind(C, name="forge_scf_init_update") result(r)
      import :: c_int, c_ptr, c_double, c_double_complex, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: diag(*)
      integer(c_int64_t), value :: nupd
      integer(c_int64_t), intent(in) :: row(*), col(*), cell(*)
      complex(c_double_complex), intent(in) :: delta(*)
      integer(c_int) :: r
    end function forge_scf_init_update
    function forge_mlp_param_count(din, dout) bind(C, name="forge_mlp_param_count") result(r)
      import :: c_int64_t
      integer(c_int64_t), value :: din, dout
      integer(c_int64_t) :: r
    end function forge_mlp_param_count
    function forge_mlp_loss_grad(params, din, dout, X, T, mask, N, loss, grad) &
        bind(C, name="forge_mlp_loss_grad") result(r)
      import :: c_int, c_float, c_int64_t
      real(c_float), intent(in) :: params(*), X(*), T(*), mask(*)
      integer(c_int64_t), value :: din, dout, N
      real(c_float), intent(out) :: loss, grad(*)
      integer(c_int) :: r
    end function forge_mlp_loss_grad
    function forge_mlp_train(params, din, dout, X, T, mask, N, steps, batch, lr, seed, loss_hist) &
        bind(C, name="forge_mlp_train") result(r)
      import :: c_int, c_float, c_double, c_int64_t
      real(c_float), intent(inout) :: params(*)
      integer(c_int64_t), value :: din, dout, N, steps, batch, seed
      real(c_float), intent(in) :: X(*), T(*), mask(*)
      real(c_double), value :: lr
      real(c_float), intent(out) :: loss_hist(*)
      integer(c_int) :: r
    end function forge_mlp_train
    function forge_mlp_infer(params, din, dout, X, N, bf16, Y) bind(C, name="forge_mlp_infer") &
        result(r)
      import :: c_int, c_float, c_int32_t, c_int64_t
      real(c_float), intent(in) :: params(*), X(*)
      integer(c_int64_t), value :: din, dout, N
      integer(c_int32_t), value :: bf16
      real(c_float), intent(out) :: Y(*)
      integer(c_int) :: r
    end function forge_mlp_infer
    function forge_valley_phase(ctx, vk, RotateLayers, zV) bind(C, name="forge_valley_phase") &
        result(r)
      import :: c_int, c_ptr, c_double, c_double_complex, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: vk(*)
      integer(c_int64_t), intent(in) :: RotateLayers(*)
      complex(c_double_complex), intent(out) :: zV(*)
      integer(c_int) :: r
    end function forge_valley_phase
    function forge_valley_transform(ctx, zEigenvectors, RotateLayers) &
        bind(C, name="forge_valley_transform") result(r)
      import :: c_int, c_ptr, c_double_complex, c_int64_t
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(inout) :: zEigenvectors(*)
      integer(c_int64_t), intent(in) :: RotateLayers(*)
      integer(c_int) :: r
    end function forge_valley_transform
    function forge_scf_set_valley(ctx, RotateLayers) bind(C, name="forge_scf_set_valley") result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: RotateLayers
      integer(c_int) :: r
    end function forge_scf_set_valley
    function forge_scf_set_options(ctx, opt) bind(C, name="forge_scf_set_options") result(r)
      import :: c_int, c_ptr, forge_scf_options_t
      type(c_ptr), value :: ctx
      type(forge_scf_options_t), intent(in) :: opt
      integer(c_int) :: r
    end function forge_scf_set_options

    ! rho_start: c_loc of the ndim x ndim x nnc start density, or c_null_ptr
    ! for the first diagonalisation
    function forge_scf_init(ctx, rho_start) bind(C, name="forge_scf_init") result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, rho_start
      integer(c_int) :: r
    end function forge_scf_init

    function forge_scf_run(ctx, max_iter, traj, iters_done) bind(C, name="forge_scf_run") result(r)
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: max_iter
      real(c_double), intent(out) :: traj(*)
      integer(c_int64_t), intent(out) :: iters_done
      integer(c_int) :: r
    end function forge_scf_run

    ! two spins (numS = 2) on the device: a pair of contexts, one per spin
    function forge_scf2_create(spin1, spin2, pair) bind(C, name="forge_scf2_create") result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: spin1, spin2
      type(c_ptr), intent(out) :: pair
      integer(c_int) :: r
    end function forge_scf2_create

    function forge_scf2_init(pair, rho_start) bind(C, name="forge_scf2_init") result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: pair, rho_start
      integer(c_int) :: r
    end function forge_scf2_init

    function forge_scf2_run(pair, max_iter, traj, iters_done) bind(C, name="forge_scf2_run") result(r)
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: pair
      integer(c_int64_t), value :: max_iter
      real(c_double), intent(out) :: traj(*)
      integer(c_int64_t), intent(out) :: iters_done
      integer(c_int) :: r
    end function forge_scf2_run

    function forge_scf2_get_density(pair, rho) bind(C, name="forge_scf2_get_density") result(r)
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: pair
      complex(c_double_complex), intent(out) :: rho(*)
      integer(c_int) :: r
    end function forge_scf2_get_density

    function forge_scf2_destroy(pair) bind(C, name="forge_scf2_destroy") result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: pair
      integer(c_int) :: r
    end function forge_scf2_destroy

    ! twist-angle sweeps: contexts advanced together, one uneven-size batch
    function forge_sweep_create(ctxs, nsys, sweep) bind(C, name="forge_sweep_create") result(r)
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), intent(in) :: ctxs(*)
      integer(c_int64_t), value :: nsys
      type(c_ptr), intent(out) :: sweep
      integer(c_int) :: r
    end function forge_sweep_create

    function forge_sweep_run(sweep, max_iter, traj, iters_done) bind(C, name="forge_sweep_run") result(r)
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: sweep
      integer(c_int64_t), value :: max_iter
      real(c_double), intent(out) :: traj(*)
      integer(c_int64_t), intent(out) :: iters_done(*)
      integer(c_int) :: r
    end function forge_sweep_run

    function forge_sweep_destroy(sweep) bind(C, name="forge_sweep_destroy") result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: sweep
      integer(c_int) :: r
    end function forge_sweep_destroy

    function forge_scf_get_density(ctx, rho) bind(C, name="forge_scf_get_density") result(r)
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: rho(*)
      integer(c_int) :: r
    end function forge_scf_get_density

    function forge_scf_timings(ctx, out24) bind(C, name="forge_scf_timings") result(r)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out24(24)
      integer(c_int) :: r
    end function forge_scf_timings
  end interface

contains

  ! forge_last_error() as a Fortran string
  function forge_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    interface
      function c_strlen(str) bind(C, name="strlen") result(len)
        import :: c_ptr, c_size_t
        type(c_ptr), value :: str
        integer(c_size_t) :: len
      end function c_strlen
    end interface
    p = forge_last_error()
    if (.not. c_associated(p)) then
      msg = ""
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function forge_error_message

end module forge_c
