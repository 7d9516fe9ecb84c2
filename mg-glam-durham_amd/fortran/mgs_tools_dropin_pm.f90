! mgs_tools_dropin_pm.f90 -- the PM-step routines around the solve as
! EXTERNAL procedures with the reference's names and argument types,
! compiled against the application's module Tools (PMP2mod_tools.f90).
!
!   POTENTfft5       PMP2main.f90:705-867   FI (density contrast) -> FI (potential)
!   POTENT_FOR_MG    PMP2main.f90:869-935   FI += the scalar field's potential
!   MOVE(FactV)      PMP2main.f90:304-685   push of XPAR..VZ with the gradient of FI
!   mgs_densit       PMP2mod_density.f90:17-281 (CIC branch): FI from XPAR..ZPAR
!
! The reference defines POTENTfft5, POTENT_FOR_MG and MOVE in its main
! program's file, so a build that uses these removes them there; this file is
! a separate object so that a build keeping them links without duplicates.
! DENSIT is a procedure of the reference's module Density, which other
! modules use; its GPU form is therefore named mgs_densit (a CALL DENSIT
! becomes CALL mgs_densit).  The constants are formed with the reference's
! typing (default REAL where it uses default REAL: trfi, PCONST, XCONST,
! Ahalf, ENKIN), the arithmetic is libmgsolve's (mgs_potential,
! mgs_potential_add_fifth_force, mgs_move, mgs_cic_deposit), and FI and the
! particle arrays are read and written on the host at each call, as the
! reference's are.  POTENT_FOR_MG adds the potential of the field of the last
! scalar_field solve, which the GPU still holds.

! the particle arrays of Tools (PMP2mod_tools.f90:46) to and from the GPU;
! they are passed as explicit-shape TARGET dummies, Tools declaring them
! without TARGET
subroutine dropin_particles_io(to_device, n, x, y, z, u, v, w)
  use mgs_dropin_state
  implicit none
  logical, intent(in) :: to_device
  integer(c_int64_t), intent(in) :: n
  real(4), target, intent(inout) :: x(n), y(n), z(n), u(n), v(n), w(n)
  if (to_device) then
    call dropin_check(mgs_particles_set(hd, c_loc(x), c_loc(y), c_loc(z), c_loc(u), c_loc(v), c_loc(w), n, MGS_HOST), &
                      'mgs_particles_set')
  else
    call dropin_check(mgs_particles_get(hd, c_loc(x), c_loc(y), c_loc(z), c_loc(u), c_loc(v), c_loc(w), MGS_HOST), &
                      'mgs_particles_get')
  end if
end subroutine dropin_particles_io

subroutine dropin_particles_to_device()
  use Tools, only: XPAR, YPAR, ZPAR, VX, VY, VZ, Nparticles
  use mgs_dropin_state, only: c_int64_t
  implicit none
  external :: dropin_particles_io
  call dropin_particles_io(.true., int(Nparticles, c_int64_t), XPAR, YPAR, ZPAR, VX, VY, VZ)
end subroutine dropin_particles_to_device

! DENSIT's CIC branch (PMP2mod_density.f90:37-55, 235-281): FI = -1 plus
! WPAR times each particle's CIC weights
subroutine mgs_densit()
  use Tools, only: FI
  use mgs_dropin_state
  implicit none
  external :: dropin_particles_to_device
  call dropin_sync()
  call dropin_particles_to_device()
  call dropin_check(mgs_cic_deposit(hd), 'DENSIT')
  call dropin_check(mgs_get_density(hd, c_loc(FI), MGS_F32, MGS_HOST), 'DENSIT')
end subroutine mgs_densit

SUBROUTINE POTENTfft5
  use Tools, only: FI, Om, Omb, AEXPN
  use mgs_dropin_state
  implicit none
  real :: trfi
  call dropin_sync()
  ! PMP2main.f90:727-731, in default REAL
  If (Omb < 0.005) Then
    trfi = 1.5*(Om - Omb)/AEXPN
  Else
    trfi = 1.5*Om/AEXPN
  end If
  call dropin_check(mgs_set_density(hd, c_loc(FI), MGS_F32, MGS_HOST), 'POTENTfft5')
  call dropin_check(mgs_potential(hd, dble(trfi)), 'POTENTfft5')
  call dropin_check(mgs_get_potential(hd, c_loc(FI), MGS_F32, MGS_HOST), 'POTENTfft5')
END SUBROUTINE POTENTfft5

SUBROUTINE POTENT_FOR_MG
  use Tools, only: FI
  use mgs_dropin_state
  implicit none
  call dropin_sync()
  call dropin_check(mgs_set_potential(hd, c_loc(FI), MGS_F32, MGS_HOST), 'POTENT_FOR_MG')
  call dropin_check(mgs_potential_add_fifth_force(hd), 'POTENT_FOR_MG')
  call dropin_check(mgs_get_potential(hd, c_loc(FI), MGS_F32, MGS_HOST), 'POTENT_FOR_MG')
END SUBROUTINE POTENT_FOR_MG

! MOVE(FactV), PMP2main.f90:304-685 for f(R) and nDGP
SUBROUTINE MOVE(FactV)
  use Tools, only: FI, XPAR, YPAR, ZPAR, VX, VY, VZ, Nparticles, AEXPN, ASTEP, Om, OmL, ENKIN
  use mgs_dropin_state
  implicit none
  real :: FactV
  real*8 :: PCONST, XCONST, svel
  real :: Ahalf
  external :: dropin_particles_to_device, dropin_particles_io
  call dropin_sync()
  ! :333-335, every operand default REAL
  PCONST = -SQRT(AEXPN/(Om + OmL*AEXPN**3))*ASTEP*0.5/FactV
  Ahalf = AEXPN + ASTEP/2.
  XCONST = ASTEP/SQRT(Ahalf*(Om + OmL*Ahalf**3))/Ahalf
  call dropin_particles_to_device()
  call dropin_check(mgs_set_potential(hd, c_loc(FI), MGS_F32, MGS_HOST), 'MOVE')
  call dropin_check(mgs_move(hd, PCONST, XCONST, svel), 'MOVE')
  call dropin_particles_io(.false., int(Nparticles, c_int64_t), XPAR, YPAR, ZPAR, VX, VY, VZ)
  ! :689
  ENKIN = svel/2./(AEXPN + ASTEP/2.)**2
END SUBROUTINE MOVE
