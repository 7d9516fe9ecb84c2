! mgs_extradof.f90 -- drop-in replacements for the reference's multigrid
! scalar-field routines, with the reference's names and argument lists,
! whose bodies call libmgsolve through the mgsolve module.
!
!   scalar_field(i_step)                          PMP2extradof.f90:12-178
!   initialise_extradof                           PMP2extradof.f90:181-344
!   V_cycle(ilevel)                               PMP2extradof.f90:352-555
!   F_cycle(ilevel), W_cycle(ilevel)              PMP2extradof.f90:563-1157
!   restrict_density(ilevel)                      PMP2extradof.f90:1168
!   restrict_scalar_field(ilevel)                 PMP2extradof.f90:1262
!   correct_solution(ilevel)                      PMP2extradof.f90:1361
!   relaxation_iterations_fR(ilevel, redstep)     PMP2MGsolver_fR.f90:20
!   calculate_residual_fR(ilevel, res)            PMP2MGsolver_fR.f90:308
!   restrict_residual_fR(ilevel)                  PMP2MGsolver_fR.f90:569
!   calculate_physical_right_hand_side_fR(ilevel) PMP2MGsolver_fR.f90:789
!   relaxation_iterations_DGP(ilevel, redstep)    PMP2MGsolver_DGP.f90:8
!   calculate_residual_DGP(ilevel, res)           PMP2MGsolver_DGP.f90:240
!   restrict_residual_DGP(ilevel)                 PMP2MGsolver_DGP.f90:418
!   calculate_physical_right_hand_side_DGP(ilevel) PMP2MGsolver_DGP.f90:793
!   DENSIT (CIC branch)                           PMP2mod_density.f90:17-281
!   POTENTfft5                                    PMP2main.f90:705-867
!   POTENT_FOR_MG                                 PMP2main.f90:869-1015
!   MOVE(FactV) (f(R)/nDGP)                       PMP2main.f90:304-685
!
! The reference keeps its state in module Tools (FI = delta, FI2 = u or phi,
! FI3 = the packed hierarchy).  Here the hierarchy lives on the GPU, owned
! by one libmgsolve handle; the host arrays FI and FI2 are borrowed through
! mgs_attach_fields and synchronised by scalar_field (density in, field in
! and out), or explicitly by mgs_fields_to_device / mgs_fields_to_host for
! callers that drive the per-level routines themselves.  The PM routines
! take the particle arrays XPAR..VZ borrowed through mgs_attach_particles
! and, like the reference, read and write FI and those arrays on the host at
! each call (the particles, density and potential also stay on the GPU for
! callers of the C-ABI).  Errors print
! libmgsolve's message and STOP, as the reference does for an unsupported
! model (PMP2extradof.f90:408-411); non-convergence is a printed warning.
module mgs_extradof
  use iso_c_binding
  use mgsolve
  implicit none
  private

  public :: mgs_configure, mgs_attach_fields, mgs_release, mgs_fields_to_device, mgs_fields_to_host
  public :: mgs_handle_of
  public :: scalar_field, initialise_extradof, V_cycle, F_cycle, W_cycle
  public :: restrict_density, restrict_scalar_field, correct_solution
  public :: relaxation_iterations_fR, calculate_residual_fR, restrict_residual_fR
  public :: calculate_physical_right_hand_side_fR
  public :: relaxation_iterations_DGP, calculate_residual_DGP, restrict_residual_DGP
  public :: calculate_physical_right_hand_side_DGP
  public :: mgs_attach_particles, DENSIT, POTENTfft5, POTENT_FOR_MG, MOVE

  ! run control of the reference's scalar_field (PMP2mod_tools.f90 scalars)
  integer, public, save :: iter_count_max = 100
  real(8), public, save :: res_conv = 1.0d-8
  ! cycle type of scalar_field, as the reference's switches (PMP2extradof.f90:86-88)
  integer, public, save :: Vcycle = 1, Fcycle = 0, Wcycle = 0
  logical, public, save :: print_residuals = .true.
  ! PM step: the time step ASTEP that MOVE reads from module Tools, and the
  ! kinetic energy MOVE leaves there (ENKIN, PMP2main.f90:681)
  real(8), public, save :: ASTEP = 0.0d0, ENKIN = 0.0d0

  type(c_ptr), save :: h = c_null_ptr
  type(mgs_params), save :: prm
  logical, save :: configured = .false.
  real(4), pointer, save :: FI(:,:,:) => null(), FI2(:,:,:) => null()
  real(4), pointer, save :: XPAR(:) => null(), YPAR(:) => null(), ZPAR(:) => null()
  real(4), pointer, save :: VX(:) => null(), VY(:) => null(), VZ(:) => null()

contains

  ! ---- set-up ----------------------------------------------------------

  ! model and cosmology: NGRID, Box, Om, OmL, AEXPN, MG_model, fr_n, fR0,
  ! H0rc, N_branch (1 normal / 0 self-accelerating); precision 0 plain FP64,
  ! 1 mixed.  A later call with the same ngrid/model/precision updates the
  ! scalars of the existing hierarchy (a new time step).
  subroutine mgs_configure(ngrid, box, om, oml, aexpn, mg_model, fr_n, fr0, h0rc, n_branch, precision)
    integer, intent(in) :: ngrid, mg_model, fr_n, n_branch, precision
    real(8), intent(in) :: box, om, oml, aexpn, fr0, h0rc
    type(mgs_params) :: p
    p%ngrid = ngrid
    p%model = mg_model
    p%fr_n = fr_n
    p%branch = merge(1, -1, n_branch /= 0)
    p%box = box
    p%omega_m = om
    p%omega_l = oml
    p%aexpn = aexpn
    p%fr0 = fr0
    p%h0rc = h0rc
    p%precision = precision
    p%flags = 0
    if (c_associated(h)) then
      if (p%ngrid == prm%ngrid .and. p%model == prm%model .and. p%precision == prm%precision) then
        call check(mgs_set_params(h, p), 'mgs_set_params')
        prm = p
        return
      end if
      call mgs_release()
    end if
    prm = p
    configured = .true.
  end subroutine mgs_configure

  ! the host arrays scalar_field reads and writes (the reference's FI, FI2)
  subroutine mgs_attach_fields(density, field)
    real(4), target, intent(inout) :: density(:,:,:), field(:,:,:)
    FI => density
    FI2 => field
  end subroutine mgs_attach_fields

  subroutine mgs_release()
    integer(c_int) :: rc
    if (c_associated(h)) rc = mgs_destroy(h)
    h = c_null_ptr
  end subroutine mgs_release

  type(c_ptr) function mgs_handle_of()
    call ensure_handle()
    mgs_handle_of = h
  end function mgs_handle_of

  subroutine ensure_handle()
    if (c_associated(h)) return
    if (.not. configured) then
      write(*, '(A)') 'mgs_extradof: mgs_configure must be called first'
      stop 1
    end if
    call check(mgs_create(prm, h), 'mgs_create')
  end subroutine ensure_handle

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc == MGS_OK) return
    write(*, '(A,A,A,I3,A,A)') 'libmgsolve: ', what, ' failed (', rc, '): ', mgs_error_text(h)
    stop 1
  end subroutine check

  subroutine need_fields()
    if (.not. associated(FI) .or. .not. associated(FI2)) then
      write(*, '(A)') 'mgs_extradof: mgs_attach_fields(FI, FI2) must be called first'
      stop 1
    end if
  end subroutine need_fields

  ! FI -> device density, FI2 -> device field (raw u or phi)
  subroutine mgs_fields_to_device()
    call ensure_handle()
    call need_fields()
    call check(mgs_set_density(h, c_loc(FI), MGS_F32, MGS_HOST), 'mgs_set_density')
    call check(mgs_set_field(h, c_loc(FI2), MGS_F32, MGS_HOST), 'mgs_set_field')
  end subroutine mgs_fields_to_device

  ! device field (raw) -> FI2
  subroutine mgs_fields_to_host()
    call ensure_handle()
    call need_fields()
    call check(mgs_get_field(h, c_loc(FI2), MGS_F32, MGS_HOST, 1.0d0), 'mgs_get_field')
  end subroutine mgs_fields_to_host

  ! ---- the outer loop ----------------------------------------------------

  ! scalar_field(i_step), PMP2extradof.f90:12-178: step 1 starts from the
  ! background, later steps from FI2 (normalised by u_bg, :47-62); V-cycles
  ! until the reference's exits (:107-122); FI2 left normalised (:131-146)
  subroutine scalar_field(i_step)
    integer, intent(in) :: i_step
    type(mgs_solve_opts) :: o
    type(mgs_solve_report) :: r
    real(8) :: fr_bg, u_bg
    integer :: it
    call ensure_handle()
    call need_fields()
    if (i_step == 1) call initialise_extradof()
    call check(mgs_set_density(h, c_loc(FI), MGS_F32, MGS_HOST), 'mgs_set_density')
    u_bg = 1.0d0
    if (prm%model == MGS_MODEL_FR) call check(mgs_background(h, fr_bg, u_bg), 'mgs_background')
    if (i_step > 1) then
      if (prm%model == MGS_MODEL_FR) FI2 = FI2 * u_bg
      call check(mgs_set_field(h, c_loc(FI2), MGS_F32, MGS_HOST), 'mgs_set_field')
    end if
    o%mode = 0
    o%max_cycles = iter_count_max
    o%res_conv = res_conv
    o%rel_tol = 0.0d0
    o%init_guess = 0
    o%inner_cycles = 1
    o%cycle = 0
    if (Fcycle == 1) o%cycle = 1
    if (Wcycle == 1) o%cycle = 2
    o%reserved = 0
    call check(mgs_solve(h, o, r), 'mgs_solve')
    if (print_residuals) then
      do it = 1, min(int(r%outer), 128)
        write(*, '(A,F20.12,A,I5,A)') 'The residual is equal to ', r%history(it), ' after ', it, ' V-cycles.'
      end do
      write(*, '(A,F20.12)') 'Residual on the PM grid = ', r%res_final
      if (r%outer >= iter_count_max) &
        write(*, '(A,I5,A)') 'ATTENTION: not converged after a maximum of ', int(r%outer), ' iterations'
    end if
    if (prm%model == MGS_MODEL_FR) then
      call check(mgs_get_field(h, c_loc(FI2), MGS_F32, MGS_HOST, 1.0d0 / u_bg), 'mgs_get_field')
    else
      call check(mgs_get_field(h, c_loc(FI2), MGS_F32, MGS_HOST, 1.0d0), 'mgs_get_field')
    end if
  end subroutine scalar_field

  ! initialise_extradof, PMP2extradof.f90:181-344: u := u_bg (f(R)) or
  ! phi := 0 (DGP) on the device and in FI2
  subroutine initialise_extradof()
    call ensure_handle()
    call check(mgs_initialise(h), 'mgs_initialise')
    if (associated(FI2)) call check(mgs_get_field(h, c_loc(FI2), MGS_F32, MGS_HOST, 1.0d0), 'mgs_get_field')
  end subroutine initialise_extradof

  ! ---- per-level operators on the device hierarchy ----------------------

  subroutine V_cycle(ilevel)
    integer, intent(in) :: ilevel
    call ensure_handle()
    call check(mgs_vcycle(h, ilevel), 'V_cycle')
  end subroutine V_cycle

  subroutine F_cycle(ilevel)
    integer, intent(in) :: ilevel
    call ensure_handle()
    call check(mgs_fcycle(h, ilevel), 'F_cycle')
  end subroutine F_cycle

  subroutine W_cycle(ilevel)
    integer, intent(in) :: ilevel
    call ensure_handle()
    call check(mgs_wcycle(h, ilevel), 'W_cycle')
  end subroutine W_cycle

  subroutine restrict_density(ilevel)
    integer, intent(in) :: ilevel
    call ensure_handle()
    call check(mgs_restrict_density(h, ilevel), 'restrict_density')
  end subroutine restrict_density

  subroutine restrict_scalar_field(ilevel)
    integer, intent(in) :: ilevel
    call ensure_handle()
    call check(mgs_restrict_field(h, ilevel), 'restrict_scalar_field')
  end subroutine restrict_scalar_field

  subroutine correct_solution(ilevel)
    integer, intent(in) :: ilevel
    call ensure_handle()
    call check(mgs_correct(h, ilevel), 'correct_solution')
  end subroutine correct_solution

  subroutine relaxation_iterations_fR(ilevel, redstep)
    integer, intent(in) :: ilevel
    logical, intent(in) :: redstep
    call ensure_handle()
    call check(mgs_smooth(h, ilevel, merge(1, 0, redstep)), 'relaxation_iterations_fR')
  end subroutine relaxation_iterations_fR

  subroutine calculate_residual_fR(ilevel, res)
    integer, intent(in) :: ilevel
    real(8), intent(out) :: res
    call ensure_handle()
    call check(mgs_residual(h, ilevel, res), 'calculate_residual_fR')
  end subroutine calculate_residual_fR

  subroutine restrict_residual_fR(ilevel)
    integer, intent(in) :: ilevel
    call ensure_handle()
    call check(mgs_restrict_residual(h, ilevel), 'restrict_residual_fR')
  end subroutine restrict_residual_fR

  subroutine calculate_physical_right_hand_side_fR(ilevel)
    integer, intent(in) :: ilevel
    call ensure_handle()
    call check(mgs_rhs(h, ilevel), 'calculate_physical_right_hand_side_fR')
  end subroutine calculate_physical_right_hand_side_fR

  subroutine relaxation_iterations_DGP(ilevel, redstep)
    integer, intent(in) :: ilevel
    logical, intent(in) :: redstep
    call ensure_handle()
    call check(mgs_smooth(h, ilevel, merge(1, 0, redstep)), 'relaxation_iterations_DGP')
  end subroutine relaxation_iterations_DGP

  subroutine calculate_residual_DGP(ilevel, res)
    integer, intent(in) :: ilevel
    real(8), intent(out) :: res
    call ensure_handle()
    call check(mgs_residual(h, ilevel, res), 'calculate_residual_DGP')
  end subroutine calculate_residual_DGP

  subroutine restrict_residual_DGP(ilevel)
    integer, intent(in) :: ilevel
    call ensure_handle()
    call check(mgs_restrict_residual(h, ilevel), 'restrict_residual_DGP')
  end subroutine restrict_residual_DGP

  subroutine calculate_physical_right_hand_side_DGP(ilevel)
    integer, intent(in) :: ilevel
    call ensure_handle()
    call check(mgs_rhs(h, ilevel), 'calculate_physical_right_hand_side_DGP')
  end subroutine calculate_physical_right_hand_side_DGP

  ! ---- PM step around the solve -------------------------------------------

  ! the particle arrays of module Tools (PMP2mod_tools.f90:46), REAL*4,
  ! positions in grid units in [1, NGRID+1)
  subroutine mgs_attach_particles(x, y, z, vx_, vy_, vz_)
    real(4), target, intent(inout) :: x(:), y(:), z(:), vx_(:), vy_(:), vz_(:)
    XPAR => x
    YPAR => y
    ZPAR => z
    VX => vx_
    VY => vy_
    VZ => vz_
  end subroutine mgs_attach_particles

  subroutine need_particles()
    if (.not. associated(XPAR)) then
      write(*, '(A)') 'mgs_extradof: mgs_attach_particles(XPAR, YPAR, ZPAR, VX, VY, VZ) must be called first'
      stop 1
    end if
  end subroutine need_particles

  subroutine particles_to_device()
    call need_particles()
    call check(mgs_particles_set(h, c_loc(XPAR), c_loc(YPAR), c_loc(ZPAR), c_loc(VX), c_loc(VY), c_loc(VZ), &
                                 int(size(XPAR), c_int64_t), MGS_HOST), 'mgs_particles_set')
  end subroutine particles_to_device

  ! DENSIT's CIC branch, PMP2mod_density.f90:37-55, 235-281: FI = -1 plus
  ! WPAR times each particle's CIC weights
  subroutine DENSIT()
    call ensure_handle()
    call need_fields()
    call particles_to_device()
    call check(mgs_cic_deposit(h), 'DENSIT')
    call check(mgs_get_density(h, c_loc(FI), MGS_F32, MGS_HOST), 'DENSIT')
  end subroutine DENSIT

  ! POTENTfft5, PMP2main.f90:705-867: FI (delta) -> FI (potential), with
  ! trfi = 1.5 Om / AEXPN (the reference's branch without massive neutrinos)
  subroutine POTENTfft5()
    call ensure_handle()
    call need_fields()
    call check(mgs_set_density(h, c_loc(FI), MGS_F32, MGS_HOST), 'POTENTfft5')
    call check(mgs_potential(h, 1.5d0 * prm%omega_m / prm%aexpn), 'POTENTfft5')
    call check(mgs_get_potential(h, c_loc(FI), MGS_F32, MGS_HOST), 'POTENTfft5')
  end subroutine POTENTfft5

  ! POTENT_FOR_MG, PMP2main.f90:869-935 (f(R), nDGP): FI += the scalar
  ! field's potential, from the field of the last solve on the GPU
  subroutine POTENT_FOR_MG()
    call ensure_handle()
    call need_fields()
    call check(mgs_set_potential(h, c_loc(FI), MGS_F32, MGS_HOST), 'POTENT_FOR_MG')
    call check(mgs_potential_add_fifth_force(h), 'POTENT_FOR_MG')
    call check(mgs_get_potential(h, c_loc(FI), MGS_F32, MGS_HOST), 'POTENT_FOR_MG')
  end subroutine POTENT_FOR_MG

  ! MOVE(FactV), PMP2main.f90:304-685 for f(R) and nDGP: PCONST and XCONST
  ! from AEXPN, ASTEP, Om, OmL (:390-393), the push with the gradient of FI,
  ! ENKIN = SVEL / 2 / (AEXPN + ASTEP/2)^2 (:681)
  subroutine MOVE(FactV)
    real(8), intent(in) :: FactV
    real(8) :: pconst, xconst, ahalf, svel, a, om, oml
    call ensure_handle()
    call need_fields()
    a = prm%aexpn
    om = prm%omega_m
    oml = prm%omega_l
    pconst = -sqrt(a / (om + oml * a**3)) * ASTEP * 0.5d0 / FactV
    ahalf = a + ASTEP / 2.0d0
    xconst = ASTEP / sqrt(ahalf * (om + oml * ahalf**3)) / ahalf
    call particles_to_device()
    call check(mgs_set_potential(h, c_loc(FI), MGS_F32, MGS_HOST), 'MOVE')
    call check(mgs_move(h, pconst, xconst, svel), 'MOVE')
    call check(mgs_particles_get(h, c_loc(XPAR), c_loc(YPAR), c_loc(ZPAR), c_loc(VX), c_loc(VY), c_loc(VZ), &
                                 MGS_HOST), 'MOVE')
    ENKIN = svel / 2.0d0 / (a + ASTEP / 2.0d0)**2
  end subroutine MOVE

end module mgs_extradof
