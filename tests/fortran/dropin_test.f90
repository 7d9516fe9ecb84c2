! dropin_test.f90 -- drives the Fortran drop-in routines (module
! mgs_extradof) the way the reference's time-step loop does, for the GPU test
! test_fortran_dropin.py, which repeats the same calls through the Python
! binding and compares the outputs bit for bit.
!
!   dropin_test <dir>
!
! reads <dir>/params.nml (namelist /run/), <dir>/delta1.bin, <dir>/delta2.bin
! and <dir>/field0.bin (N^3 float32, x fastest) and writes
!   <dir>/solve1.bin, <dir>/solve2.bin : FI2 after scalar_field(1) and (2)
!   <dir>/ops.bin  : FI2 after a hand-driven sequence of per-level routines
!   <dir>/ops.txt  : the residuals that sequence returned
program dropin_test
  use mgs_extradof
  implicit none
  integer :: ngrid = 16, mg_model = 1, fr_n = 1, n_branch = 1, precision = 0, iter_max = 30, npart = 0
  real(8) :: box = 64d0, om = 0.3d0, oml = 0.7d0, aexpn = 1d0, fr0 = -1d-5, h0rc = 1d0, conv = 1d-10, astep_in = 0.02d0
  namelist /run/ ngrid, mg_model, fr_n, n_branch, precision, iter_max, box, om, oml, aexpn, fr0, h0rc, conv, &
                 npart, astep_in
  character(len=512) :: dir
  real(4), allocatable, target :: FI(:,:,:), FI2(:,:,:)
  real(4), allocatable, target :: XP(:), YP(:), ZP(:), VXP(:), VYP(:), VZP(:)
  real(8) :: r0, r1, r2
  integer :: u

  call get_command_argument(1, dir)
  open(newunit=u, file=trim(dir)//'/params.nml', status='old')
  read(u, nml=run)
  close(u)
  allocate(FI(ngrid, ngrid, ngrid), FI2(ngrid, ngrid, ngrid))

  call mgs_configure(ngrid, box, om, oml, aexpn, mg_model, fr_n, fr0, h0rc, n_branch, precision)
  call mgs_attach_fields(FI, FI2)
  iter_count_max = iter_max
  res_conv = conv

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
  if (mg_model == 1) then
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
  open(newunit=u, file=trim(dir)//'/ops.txt', status='replace')
  write(u, '(3ES25.17)') r0, r1, r2
  close(u)

  ! the PM step's routines (DENSIT, POTENTfft5, POTENT_FOR_MG, MOVE) with the
  ! field the cycles above left on the GPU; particles.bin holds XPAR, YPAR,
  ! ZPAR, VX, VY, VZ (npart REAL*4 each)
  if (npart > 0) then
    allocate(XP(npart), YP(npart), ZP(npart), VXP(npart), VYP(npart), VZP(npart))
    open(newunit=u, file=trim(dir)//'/particles.bin', access='stream', form='unformatted', status='old')
    read(u) XP, YP, ZP, VXP, VYP, VZP
    close(u)
    call mgs_attach_particles(XP, YP, ZP, VXP, VYP, VZP)
    ASTEP = astep_in
    call DENSIT()
    call write_grid('pm_density.bin', FI)
    call POTENTfft5()
    call write_grid('pm_phi.bin', FI)
    call POTENT_FOR_MG()
    call write_grid('pm_phimg.bin', FI)
    call MOVE(1.0d0)
    open(newunit=u, file=trim(dir)//'/pm_part.bin', access='stream', form='unformatted', status='replace')
    write(u) XP, YP, ZP, VXP, VYP, VZP
    close(u)
    open(newunit=u, file=trim(dir)//'/pm.txt', status='replace')
    write(u, '(ES25.17)') ENKIN
    close(u)
  end if
  call mgs_release()

contains

  subroutine read_grid(name, a)
    character(len=*), intent(in) :: name
    real(4), intent(out) :: a(:,:,:)
    integer :: v
    open(newunit=v, file=trim(dir)//'/'//name, access='stream', form='unformatted', status='old')
    read(v) a
    close(v)
  end subroutine read_grid

  subroutine write_grid(name, a)
    character(len=*), intent(in) :: name
    real(4), intent(in) :: a(:,:,:)
    integer :: v
    open(newunit=v, file=trim(dir)//'/'//name, access='stream', form='unformatted', status='replace')
    write(v) a
    close(v)
  end subroutine write_grid

end program dropin_test
