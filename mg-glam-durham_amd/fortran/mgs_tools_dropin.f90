! mgs_tools_dropin.f90 -- the reference's multigrid routines as EXTERNAL
! procedures with the reference's names and argument types, compiled against
! the application's own module Tools (PMP2mod_tools.f90:9-110), whose bodies
! call libmgsolve through the mgsolve interface module.
!
! A build of the reference links this file (with mgsolve.f90 and
! mgs_tools_dropin_pm.f90) in place of PMP2extradof.f90, PMP2MGsolver_fR.f90
! and PMP2MGsolver_DGP.f90.  No line of the application changes: every input
! (FI, FI2, NGRID, Box, Om, OmL, AEXPN, MG_model, fr_n, fR0, H0rc,
! N_branch/S_branch, Vcycle/Fcycle/Wcycle, iter_count_max, res_conv) is read
! from Tools where the reference reads it, and every output (FI2, levelmax,
! levelmin) is written there.  The multigrid hierarchy (the reference's FI3)
! lives on the GPU, owned by one libmgsolve handle that follows NGRID and the
! model scalars from call to call.
!
!   scalar_field(i_step)                           PMP2extradof.f90:12-178
!   initialise_extradof                            PMP2extradof.f90:181-344
!   V_cycle(ilevel), F_cycle, W_cycle              PMP2extradof.f90:352, 563, 867
!   restrict_density(ilevel)                       PMP2extradof.f90:1168
!   restrict_scalar_field(ilevel)                  PMP2extradof.f90:1262
!   correct_solution(ilevel)                       PMP2extradof.f90:1361
!   relaxation_iterations_fR(ilevel, redstep)      PMP2MGsolver_fR.f90:20
!   calculate_residual_fR(ilevel, res)             PMP2MGsolver_fR.f90:308
!   restrict_residual_fR(ilevel)                   PMP2MGsolver_fR.f90:569
!   calculate_physical_right_hand_side_fR(ilevel)  PMP2MGsolver_fR.f90:789
!   relaxation_iterations_DGP(ilevel, redstep)     PMP2MGsolver_DGP.f90:8
!   calculate_residual_DGP(ilevel, res)            PMP2MGsolver_DGP.f90:240
!   restrict_residual_DGP(ilevel)                  PMP2MGsolver_DGP.f90:418
!   calculate_physical_right_hand_side_DGP(ilevel) PMP2MGsolver_DGP.f90:793
!
! Besides those, three routines of this library's own:
!   mgs_dropin_precision(p)   optional configure call: 0 the reference's
!                             method in FP64 (default), 1 the mixed-precision
!                             solve (FP32 inner cycles in an FP64 defect
!                             correction)
!   mgs_fields_to_device()    FI, FI2 -> the GPU hierarchy, and
!   mgs_fields_to_host()      the GPU field -> FI2 (raw), for callers that
!                             drive the per-level routines themselves
!
! The per-level routines act on the GPU hierarchy, which scalar_field keeps in
! step with FI and FI2.  Only f(R) (MG_model 1) and nDGP (2) are supported;
! another model STOPs with a message, as the reference does for a model it
! does not know (PMP2extradof.f90:408-411).  Any libmgsolve error (no GPU
! among them: the library has no CPU fallback) prints the library's message
! and STOPs.

module mgs_dropin_state
  use iso_c_binding
  use mgsolve
  implicit none
  type(c_ptr), save :: hd = c_null_ptr
  type(mgs_params), save :: pd
  integer, save :: prec_choice = 0
contains

  subroutine dropin_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc == MGS_OK) return
    write (*, '(A,A,A,I3,A,A)') ' libmgsolve: ', what, ' failed (', rc, '): ', mgs_error_text(hd)
    error stop 'libmgsolve error'
  end subroutine dropin_check

  ! the handle for the model and cosmology now in Tools: created on first use
  ! or when NGRID, the model or the precision change, else updated with the
  ! new scalars (a new time step)
  subroutine dropin_sync()
    use Tools, only: NGRID, Box, Om, OmL, AEXPN, MG_model, fr_n, fR0, H0rc, N_branch
    type(mgs_params) :: p
    integer(c_int) :: rc
    if (MG_model /= 1 .and. MG_model /= 2) then
      write (*, '(A,I3,A)') ' libmgsolve: MG_model ', MG_model, ' is not supported (1 f(R), 2 nDGP)'
      error stop 'Model currently unsupported.'
    end if
    p%ngrid = NGRID
    p%model = MG_model
    p%fr_n = fr_n
    p%branch = merge(1, -1, N_branch /= 0)
    p%box = dble(Box)
    p%omega_m = dble(Om)
    p%omega_l = dble(OmL)
    p%aexpn = dble(AEXPN)
    p%fr0 = fR0
    p%h0rc = H0rc
    p%precision = prec_choice
    p%flags = 0
    if (c_associated(hd)) then
      if (p%ngrid == pd%ngrid .and. p%model == pd%model .and. p%precision == pd%precision) then
        call dropin_check(mgs_set_params(hd, p), 'mgs_set_params')
        pd = p
        return
      end if
      rc = mgs_destroy(hd)
      hd = c_null_ptr
    end if
    pd = p
    call dropin_check(mgs_create(pd, hd), 'mgs_create')
  end subroutine dropin_sync

  ! u_bg = |fR_bg|^(e/(n+1)) with the reference's typing (PMP2extradof.f90:
  ! 48-52, 131-135): Om/AEXPN**3 is a REAL*4 quotient
  real(8) function dropin_ubg(e)
    use Tools, only: Om, OmL, AEXPN, fR0, fr_n
    real(8), intent(in) :: e
    real(8) :: R_bg, R0_bg, fR_bg
    R_bg = 3.0D0*(Om/AEXPN**3 + 4.0D0*OmL)
    R0_bg = 3.0D0*(Om + 4.0D0*OmL)
    fR_bg = fR0*(R0_bg/R_bg)**(fr_n + 1)
    dropin_ubg = (DABS(fR_bg))**(e/(1.0D0 + DBLE(fr_n)))
  end function dropin_ubg

end module mgs_dropin_state

subroutine mgs_dropin_precision(precision)
  use mgs_dropin_state
  implicit none
  integer, intent(in) :: precision
  if (precision /= 0 .and. precision /= 1) error stop 'mgs_dropin_precision: 0 (FP64) or 1 (mixed)'
  prec_choice = precision
end subroutine mgs_dropin_precision

subroutine mgs_fields_to_device()
  use Tools, only: FI, FI2
  use mgs_dropin_state
  implicit none
  call dropin_sync()
  call dropin_check(mgs_set_density(hd, c_loc(FI), MGS_F32, MGS_HOST), 'mgs_set_density')
  call dropin_check(mgs_set_field(hd, c_loc(FI2), MGS_F32, MGS_HOST), 'mgs_set_field')
end subroutine mgs_fields_to_device

subroutine mgs_fields_to_host()
  use Tools, only: FI2
  use mgs_dropin_state
  implicit none
  call dropin_sync()
  call dropin_check(mgs_get_field(hd, c_loc(FI2), MGS_F32, MGS_HOST, 1.0d0), 'mgs_get_field')
end subroutine mgs_fields_to_host

! scalar_field(i_step), PMP2extradof.f90:12-178: step 1 starts from the
! background, later steps from the normalised FI2 times u_bg (:44-60); cycles
! until the reference's exits (:107-122) with its printed lines; FI2 left
! normalised by u_bg (:127-141).
SUBROUTINE scalar_field(i_step)
  use Tools, only: FI, FI2, MG_model, MG_test, Vcycle, Fcycle, Wcycle, iter_count_max, res_conv, istep
  use mgs_dropin_state
  implicit none
  integer :: i_step
  type(mgs_solve_opts) :: o
  type(mgs_solve_report) :: r
  real(8) :: u_bg
  integer :: it, ncyc
  external :: initialise_extradof
  call dropin_sync()
  if (i_step == 1) call initialise_extradof
  call dropin_check(mgs_set_density(hd, c_loc(FI), MGS_F32, MGS_HOST), 'mgs_set_density')
  if (i_step > 1) then
    if (MG_model == 1) then
      u_bg = dropin_ubg(1.0D0)
      FI2 = FI2*u_bg
    end if
    call dropin_check(mgs_set_field(hd, c_loc(FI2), MGS_F32, MGS_HOST), 'mgs_set_field')
  end if
  ncyc = 0
  if (Vcycle == 1) ncyc = ncyc + 1
  if (Fcycle == 1) ncyc = ncyc + 1
  if (Wcycle == 1) ncyc = ncyc + 1
  if (ncyc /= 1) error stop 'libmgsolve: exactly one of Vcycle, Fcycle, Wcycle must be 1'
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
  call dropin_check(mgs_solve(hd, o, r), 'mgs_solve')
  do it = 1, min(int(r%outer), 128)
    WRITE (*, '(A,F20.12,A,I5,A)') 'The residual is equal to ', r%history(it), ' after ', it, ' V-cycles.'
  end do
  WRITE (*, '(A,F20.12)') 'Residual on the PM grid = ', r%res_final
  if (r%outer >= iter_count_max .and. r%converged == 0) &
    WRITE (*, '(A,I5,A)') 'ATTENTION: not converged after a maximum of ', int(r%outer), ' iterations'
  if (MG_model == 1) then
    call dropin_check(mgs_get_field(hd, c_loc(FI2), MGS_F32, MGS_HOST, dropin_ubg(-1.0D0)), 'mgs_get_field')
  else
    call dropin_check(mgs_get_field(hd, c_loc(FI2), MGS_F32, MGS_HOST, 1.0d0), 'mgs_get_field')
  end if
  if (MG_test == 1) WRITE (*, '(A,I2,A)') 'scalar field step', istep, ' done.'
END SUBROUTINE scalar_field

! initialise_extradof, PMP2extradof.f90:181-344: levelmax/levelmin into
! Tools, u := u_bg (f(R)) or phi := 0 (nDGP) on the GPU and in FI2
SUBROUTINE initialise_extradof
  use Tools, only: FI2, levelmax, levelmin
  use mgs_dropin_state
  implicit none
  integer(c_int32_t) :: lmax, lmin, lc
  call dropin_sync()
  call dropin_check(mgs_levels(hd, lmax, lmin, lc), 'mgs_levels')
  levelmax = lmax
  levelmin = lmin
  call dropin_check(mgs_initialise(hd), 'mgs_initialise')
  if (allocated(FI2)) call dropin_check(mgs_get_field(hd, c_loc(FI2), MGS_F32, MGS_HOST, 1.0d0), 'mgs_get_field')
END SUBROUTINE initialise_extradof

SUBROUTINE V_cycle(ilevel)
  use mgs_dropin_state
  implicit none
  integer :: ilevel
  call dropin_check(mgs_vcycle(hd, ilevel), 'V_cycle')
END SUBROUTINE V_cycle

SUBROUTINE F_cycle(ilevel)
  use mgs_dropin_state
  implicit none
  integer :: ilevel
  call dropin_check(mgs_fcycle(hd, ilevel), 'F_cycle')
END SUBROUTINE F_cycle

SUBROUTINE W_cycle(ilevel)
  use mgs_dropin_state
  implicit none
  integer :: ilevel
  call dropin_check(mgs_wcycle(hd, ilevel), 'W_cycle')
END SUBROUTINE W_cycle

SUBROUTINE restrict_density(ilevel)
  use mgs_dropin_state
  implicit none
  integer :: ilevel
  call dropin_check(mgs_restrict_density(hd, ilevel), 'restrict_density')
END SUBROUTINE restrict_density

SUBROUTINE restrict_scalar_field(ilevel)
  use mgs_dropin_state
  implicit none
  integer :: ilevel
  call dropin_check(mgs_restrict_field(hd, ilevel), 'restrict_scalar_field')
END SUBROUTINE restrict_scalar_field

SUBROUTINE correct_solution(ilevel)
  use mgs_dropin_state
  implicit none
  integer :: ilevel
  call dropin_check(mgs_correct(hd, ilevel), 'correct_solution')
END SUBROUTINE correct_solution

SUBROUTINE relaxation_iterations_fR(ilevel, redstep)
  use mgs_dropin_state
  implicit none
  integer :: ilevel
  logical :: redstep
  call dropin_check(mgs_smooth(hd, ilevel, merge(1, 0, redstep)), 'relaxation_iterations_fR')
END SUBROUTINE relaxation_iterations_fR

SUBROUTINE calculate_residual_fR(ilevel, res)
  use mgs_dropin_state
  implicit none
  integer :: ilevel
  real*8, intent(out) :: res
  call dropin_check(mgs_residual(hd, ilevel, res), 'calculate_residual_fR')
END SUBROUTINE calculate_residual_fR

SUBROUTINE restrict_residual_fR(ilevel)
  use mgs_dropin_state
  implicit none
  integer :: ilevel
  call dropin_check(mgs_restrict_residual(hd, ilevel), 'restrict_residual_fR')
END SUBROUTINE restrict_residual_fR

SUBROUTINE calculate_physical_right_hand_side_fR(ilevel)
  use mgs_dropin_state
  implicit none
  integer :: ilevel
  call dropin_check(mgs_rhs(hd, ilevel), 'calculate_physical_right_hand_side_fR')
END SUBROUTINE calculate_physical_right_hand_side_fR

SUBROUTINE relaxation_iterations_DGP(ilevel, redstep)
  use mgs_dropin_state
  implicit none
  integer :: ilevel
  logical :: redstep
  call dropin_check(mgs_smooth(hd, ilevel, merge(1, 0, redstep)), 'relaxation_iterations_DGP')
END SUBROUTINE relaxation_iterations_DGP

SUBROUTINE calculate_residual_DGP(ilevel, res)
  use mgs_dropin_state
  implicit none
  integer :: ilevel
  real*8, intent(out) :: res
  call dropin_check(mgs_residual(hd, ilevel, res), 'calculate_residual_DGP')
END SUBROUTINE calculate_residual_DGP

SUBROUTINE restrict_residual_DGP(ilevel)
  use mgs_dropin_state
  implicit none
  integer :: ilevel
  call dropin_check(mgs_restrict_residual(hd, ilevel), 'restrict_residual_DGP')
END SUBROUTINE restrict_residual_DGP

SUBROUTINE calculate_physical_right_hand_side_DGP(ilevel)
  use mgs_dropin_state
  implicit none
  integer :: ilevel
  call dropin_check(mgs_rhs(hd, ilevel), 'calculate_physical_right_hand_side_DGP')
END SUBROUTINE calculate_physical_right_hand_side_DGP
