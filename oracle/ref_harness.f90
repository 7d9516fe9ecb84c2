! ref_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
! A bind(C) shim around the reference's own multigrid routines, compiled
! together with the reference sources (unmodified, from where they lie) by
! oracle/Makefile into oracle/_ref/libmgref.so.  It lets the test suite and
! the golden-vector generator drive one reference routine at a time:
!   relaxation_iterations_fR/_DGP, calculate_residual_fR/_DGP,
!   restrict_residual_fR/_DGP, calculate_physical_right_hand_side_fR/_DGP
!   (PMP2MGsolver_fR.f90, PMP2MGsolver_DGP.f90),
!   restrict_density, restrict_scalar_field, correct_solution, V_cycle,
!   scalar_field (PMP2extradof.f90),
! on the module state of Tools (PMP2mod_tools.f90:9-104).  Nothing here
! computes any part of the solve itself.
module ref_harness
   use iso_c_binding
   use Tools
   implicit none
   external :: relaxation_iterations_fR, calculate_residual_fR, restrict_residual_fR
   external :: calculate_physical_right_hand_side_fR
   external :: relaxation_iterations_DGP, calculate_residual_DGP, restrict_residual_DGP
   external :: calculate_physical_right_hand_side_DGP
   external :: restrict_density, restrict_scalar_field, correct_solution
   external :: V_cycle, scalar_field, initialise_extradof
contains

   subroutine ref_init(ngrid_, model_, frn_, branch_, itmax_, box_, om_, oml_, a_, fr0_, h0rc_, resconv_) &
      bind(C, name='ref_init')
      integer(c_int), value :: ngrid_, model_, frn_, branch_, itmax_
      real(c_double), value :: box_, om_, oml_, a_, fr0_, h0rc_, resconv_
      NGRID = ngrid_
      Box = real(box_, 4)
      Om = real(om_, 4)
      OmL = real(oml_, 4)
      AEXPN = real(a_, 4)
      MG_flag = 1
      MG_test = 0
      MG_model = model_
      fr_n = frn_
      fR0 = fr0_
      H0rc = h0rc_
      N_branch = 0
      S_branch = 0
      if (branch_ >= 0) then
         N_branch = 1
      else
         S_branch = 1
      end if
      iter_count_max = itmax_
      res_conv = resconv_
      Vcycle = 1
      Fcycle = 0
      Wcycle = 0
      if (allocated(FI)) deallocate (FI)
      if (allocated(FI2)) deallocate (FI2)
      if (allocated(FI3)) deallocate (FI3)
      allocate (FI(NGRID, NGRID, NGRID), FI2(NGRID, NGRID, NGRID), FI3(NGRID, NGRID, NGRID))
      FI = 0.0
      call initialise_extradof
   end subroutine

   integer(c_int) function ref_levelmax() bind(C, name='ref_levelmax')
      ref_levelmax = levelmax
   end function

   integer(c_int) function ref_levelmin() bind(C, name='ref_levelmin')
      ref_levelmin = levelmin
   end function

   subroutine ref_put(which, src) bind(C, name='ref_put')
      integer(c_int), value :: which
      real(c_float), intent(in) :: src(*)
      integer :: n3
      n3 = NGRID*NGRID*NGRID
      select case (which)
      case (1)
         FI = reshape(src(1:n3), [NGRID, NGRID, NGRID])
      case (2)
         FI2 = reshape(src(1:n3), [NGRID, NGRID, NGRID])
      case (3)
         FI3 = reshape(src(1:n3), [NGRID, NGRID, NGRID])
      end select
   end subroutine

   subroutine ref_get(which, dst) bind(C, name='ref_get')
      integer(c_int), value :: which
      real(c_float), intent(out) :: dst(*)
      integer :: n3
      n3 = NGRID*NGRID*NGRID
      select case (which)
      case (1)
         dst(1:n3) = reshape(FI, [n3])
      case (2)
         dst(1:n3) = reshape(FI2, [n3])
      case (3)
         dst(1:n3) = reshape(FI3, [n3])
      end select
   end subroutine

   subroutine ref_smooth(ilevel, red) bind(C, name='ref_smooth')
      integer(c_int), value :: ilevel, red
      integer :: l
      logical :: rs
      l = ilevel
      rs = (red /= 0)
      if (MG_model == 1) then
         call relaxation_iterations_fR(l, rs)
      else
         call relaxation_iterations_DGP(l, rs)
      end if
   end subroutine

   real(c_double) function ref_residual(ilevel) bind(C, name='ref_residual')
      integer(c_int), value :: ilevel
      integer :: l
      real*8 :: res
      l = ilevel
      if (MG_model == 1) then
         call calculate_residual_fR(l, res)
      else
         call calculate_residual_DGP(l, res)
      end if
      ref_residual = res
   end function

   subroutine ref_restrict_residual(ilevel) bind(C, name='ref_restrict_residual')
      integer(c_int), value :: ilevel
      integer :: l
      l = ilevel
      if (MG_model == 1) then
         call restrict_residual_fR(l)
      else
         call restrict_residual_DGP(l)
      end if
   end subroutine

   subroutine ref_rhs(ilevel) bind(C, name='ref_rhs')
      integer(c_int), value :: ilevel
      integer :: l
      l = ilevel
      if (MG_model == 1) then
         call calculate_physical_right_hand_side_fR(l)
      else
         call calculate_physical_right_hand_side_DGP(l)
      end if
   end subroutine

   subroutine ref_restrict_density(ilevel) bind(C, name='ref_restrict_density')
      integer(c_int), value :: ilevel
      integer :: l
      l = ilevel
      call restrict_density(l)
   end subroutine

   subroutine ref_restrict_field(ilevel) bind(C, name='ref_restrict_field')
      integer(c_int), value :: ilevel
      integer :: l
      l = ilevel
      call restrict_scalar_field(l)
   end subroutine

   subroutine ref_correct(ilevel) bind(C, name='ref_correct')
      integer(c_int), value :: ilevel
      integer :: l
      l = ilevel
      call correct_solution(l)
   end subroutine

   ! ncycles V-cycles from the current FI2, recording the level-0 RMS
   ! residual after each one (the quantity scalar_field prints,
   ! PMP2extradof.f90:86-105, at full precision)
   subroutine ref_vcycles(ncycles, hist) bind(C, name='ref_vcycles')
      integer(c_int), value :: ncycles
      real(c_double), intent(out) :: hist(*)
      integer :: c, l0
      real*8 :: res
      l0 = 0
      do c = 1, ncycles
         call V_cycle(l0)
         if (MG_model == 1) then
            call calculate_residual_fR(l0, res)
         else
            call calculate_residual_DGP(l0, res)
         end if
         hist(c) = res
      end do
   end subroutine

   subroutine ref_cycles(ncycles, cycle, hist) bind(C, name='ref_cycles')
      integer(c_int), value :: ncycles, cycle
      real(c_double), intent(out) :: hist(*)
      integer :: c, l0
      real*8 :: res
      l0 = 0
      do c = 1, ncycles
         if (cycle == 1) then
            call F_cycle(l0)
         else if (cycle == 2) then
            call W_cycle(l0)
         else
            call V_cycle(l0)
         end if
         if (MG_model == 1) then
            call calculate_residual_fR(l0, res)
         else
            call calculate_residual_DGP(l0, res)
         end if
         hist(c) = res
      end do
   end subroutine

   subroutine ref_scalar_field(istep) bind(C, name='ref_scalar_field')
      integer(c_int), value :: istep
      integer :: s
      s = istep
      call scalar_field(s)
   end subroutine

end module ref_harness
