! ref_caller.f90 -- TEST DRIVER.  A main program in the reference's own
! terms: it uses the reference's module Tools (compiled from the reference's
! sources by oracle/Makefile), sets the Tools state as the reference's
! ReadSetup/Initialize would (PMP2mod_tools.f90:114-231, PMP2main.f90:91-195),
! and calls the drop-in's external routines the way the reference's time-step
! loop calls its own (PMP2main.f90:36-58), so that tests/test_fortran_dropin.py
! can compare every output with the same calls made through libmgsolve's
! Python binding.  The reference's own main program is linked against the
! same archive too (oracle/_ref/dropin/pmp2main_dropin.exe); this driver
! exists because that program reads its particles through direct-access
! records whose length the reference gives in ifx's 4-byte units, which
! flang counts in bytes.
!
!   ref_caller <dir>
!
! reads <dir>/params.nml (namelist /run/), <dir>/delta1.bin, <dir>/delta2.bin
! and <dir>/field0.bin (N^3 float32, x fastest) and, with npart > 0,
! <dir>/particles.bin (XPAR, YPAR, ZPAR, VX, VY, VZ); writes
!   solve1.bin, solve2.bin : FI2 after scalar_field(1) and scalar_field(2)
!   ops.bin, ops.txt       : FI2 and the residuals after per-level routines
!   pm_density.bin, pm_phi.bin, pm_phimg.bin, pm_part.bin, pm.txt :
!                            FI after mgs_densit, POTENTfft5, POTENT_FOR_MG,
!                            the particles after MOVE(FactV), and ENKIN
program ref_caller
  use Tools
  implicit none
  integer :: ngrid_in = 16, mg_model_in = 1, fr_n_in = 1, n_branch_in = 1, precision = 0, iter_max = 30, npart = 0
  real(8) :: box_in = 64d0, om_in = 0.3d0, oml_in = 0.7d0, aexpn_in = 1d0, fr0_in = -1d-5, h0rc_in = 1d0
  real(8) :: conv = 1d-10, astep_in = 0.02d0, factv_in = 1d0
  namelist /run/ ngrid_in, mg_model_in, fr_n_in, n_branch_in, precision, iter_max, box_in, om_in, oml_in, &
                 aexpn_in, fr0_in, h0rc_in, conv, npart, astep_in, factv_in
  character(len=512) :: dir
  real*8 :: r0, r1, r2
  real :: FactV
  integer :: u
  external :: scalar_field, mgs_dropin_precision, mgs_fields_to_device, mgs_fields_to_host
  external :: relaxation_iterations_fR, calculate_residual_fR, restrict_residual_fR
  external :: calculate_physical_right_hand_side_fR, relaxation_iterations_DGP, calculate_residual_DGP
  external :: restrict_residual_DGP, calculate_physical_right_hand_side_DGP
  external :: restrict_scalar_field, restrict_density, correct_solution, V_cycle
  external :: mgs_densit, POTENTfft5, POTENT_FOR_MG, MOVE

  call get_command_argument(1, dir)
  open (newunit=u, file=trim(dir)//'/params.nml', status='old')
  read (u, nml=run)
  close (u)

  ! what ReadSetup and Initialize leave in Tools
  NGRID = ngrid_in
  Box = real(box_in, 4)
  Om = real(om_in, 4)
  OmL = real(oml_in, 4)
  Omb = 0.045
  AEXPN = real(aexpn_in, 4)
  ASTEP = real(astep_in, 4)
  MG_flag = 1
  MG_test = 0
  MG_model = mg_model_in
  Vcycle = 1
  Fcycle = 0
  Wcycle = 0
  iter_count_max = iter_max
  res_conv = conv
  fr_n = fr_n_in
  fR0 = fr0_in
  H0rc = h0rc_in
  N_branch = merge(1, 0, n_branch_in /= 0)
  S_branch = merge(0, 1, n_branch_in /= 0)
  allocate (FI(NGRID, NGRID, NGRID), FI2(NGRID, NGRID, NGRID), FI3(NGRID, NGRID, NGRID))

  ! the drop-in's one optional configure call
  if (precision /= 0) call mgs_dropin_precision(precision)

  ! two time steps of the solve
  call read_grid('delta1.bin', FI)
  call scalar_field(1)
  call write_grid('solve1.bin', FI2)
  call read_grid('delta2.bin', FI)
  call scalar_field(2)
  call write_grid('solve2.bin', FI2)

  ! one V-cycle's worth of per-level routines, driven by hand
  call read_grid('field0.bin', FI2)
  call mgs_fields_to_device()
  if (MG_model == 1) then
    call relaxation_iterations_fR(0, .true.)
    call relaxation_iterations_fR(0, .false.)
    call calculate_residual_fR(0, r0)
    call restrict_residual_fR(1)
    call restrict_scalar_field(1)
    call restrict_density(1)
    call calculate_physical_right_hand_side_fR(1)
    call relaxation_iterations_fR(1, .true.)
    call relaxation_iterations_fR(1, .false.)
    call calculate_residual_fR(1, r1)
    call correct_solution(1)
    call V_cycle(0)
    call calculate_residual_fR(0, r2)
  else
    call relaxation_iterations_DGP(0, .true.)
    call relaxation_iterations_DGP(0, .false.)
    call calculate_residual_DGP(0, r0)
    call restrict_residual_DGP(1)
    call restrict_scalar_field(1)
    call restrict_density(1)
    call calculate_physical_right_hand_side_DGP(1)
    call relaxation_iterations_DGP(1, .true.)
    call relaxation_iterations_DGP(1, .false.)
    call calculate_residual_DGP(1, r1)
    call correct_solution(1)
    call V_cycle(0)
    call calculate_residual_DGP(0, r2)
  end if
  call mgs_fields_to_host()
  call write_grid('ops.bin', FI2)
  open (newunit=u, file=trim(dir)//'/ops.txt', status='replace')
  write (u, '(3ES25.17)') r0, r1, r2
  close (u)

  ! the PM step's routines with the field the cycles above left on the GPU
  if (npart > 0) then
    Nparticles = npart
    allocate (XPAR(npart), YPAR(npart), ZPAR(npart), VX(npart), VY(npart), VZ(npart))
    open (newunit=u, file=trim(dir)//'/particles.bin', access='stream', form='unformatted', status='old')
    read (u) XPAR, YPAR, ZPAR, VX, VY, VZ
    close (u)
    call mgs_densit()
    call write_grid('pm_density.bin', FI)
    call POTENTfft5()
    call write_grid('pm_phi.bin', FI)
    call POTENT_FOR_MG()
    call write_grid('pm_phimg.bin', FI)
    FactV = real(factv_in)
    call MOVE(FactV)
    open (newunit=u, file=trim(dir)//'/pm_part.bin', access='stream', form='unformatted', status='replace')
    write (u) XPAR, YPAR, ZPAR, VX, VY, VZ
    close (u)
    open (newunit=u, file=trim(dir)//'/pm.txt', status='replace')
    write (u, '(ES25.17)') ENKIN
    close (u)
  end if

contains

  subroutine read_grid(name, a)
    character(len=*), intent(in) :: name
    real(4), intent(out) :: a(:,:,:)
    integer :: v
    open (newunit=v, file=trim(dir)//'/'//name, access='stream', form='unformatted', status='old')
    read (v) a
    close (v)
  end subroutine read_grid

  subroutine write_grid(name, a)
    character(len=*), intent(in) :: name
    real(4), intent(in) :: a(:,:,:)
    integer :: v
    open (newunit=v, file=trim(dir)//'/'//name, access='stream', form='unformatted', status='replace')
    write (v) a
    close (v)
  end subroutine write_grid

end program ref_caller
