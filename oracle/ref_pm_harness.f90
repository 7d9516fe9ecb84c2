! ref_pm_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
! A bind(C) shim around the reference's own PM-step routines, compiled
! together with the whole reference object set (unmodified sources, from
! where they lie) by oracle/Makefile into oracle/_ref/libmgref_pm.so:
!   DENSIT (module Density, PMP2mod_density.f90:17, CIC branch :235-281),
!   POTENTfft5, POTENT_FOR_MG and MOVE(FactV) (external subroutines of
!   PMP2main.f90:705, :869, :304),
!   POWERfft5(iPol, iSwitch) (module Power, PMP2mod_power.f90:209),
! on the module state of Tools (PMP2mod_tools.f90:9-110).  It only sets that
! state and copies arrays in and out; it computes nothing itself.  The golden
! vector generator (tests/golden/make_golden.py) is its only user.
module ref_pm_harness
   use iso_c_binding
   use Tools
   use Density, only: DENSIT
   use Power, only: POWERfft5, Pk, dNharm, kharm, PkLog, dNharmLog, kharmLog, iCIC
   implicit none
   external :: POTENTfft5, POTENT_FOR_MG, MOVE
contains

   ! cosmology and model scalars, FI/FI2 of ngrid^3, npart particles
   subroutine ref_pm_init(ngrid_, model_, frn_, npart_, box_, om_, oml_, a_, astep_, fr0_, h0rc_) &
      bind(C, name='ref_pm_init')
      integer(c_int), value :: ngrid_, model_, frn_
      integer(c_int64_t), value :: npart_
      real(c_double), value :: box_, om_, oml_, a_, astep_, fr0_, h0rc_
      logical :: u17
      NGRID = ngrid_
      Box = real(box_, 4)
      Om = real(om_, 4)
      Omb = 0.045  ! >= 0.005: POTENTfft5's standard LCDM branch
      OmL = real(oml_, 4)
      AEXPN = real(a_, 4)
      ASTEP = real(astep_, 4)
      MG_flag = 1
      MG_test = 0
      MG_model = model_
      fr_n = frn_
      fR0 = fr0_
      H0rc = h0rc_
      N_branch = 1
      S_branch = 0
      Nparticles = npart_
      if (allocated(FI)) deallocate (FI)
      if (allocated(FI2)) deallocate (FI2)
      allocate (FI(NGRID, NGRID, NGRID), FI2(NGRID, NGRID, NGRID))
      FI = 0.0
      FI2 = 0.0
      if (allocated(XPAR)) deallocate (XPAR, YPAR, ZPAR, VX, VY, VZ)
      allocate (XPAR(Nparticles), YPAR(Nparticles), ZPAR(Nparticles), VX(Nparticles), VY(Nparticles), &
                VZ(Nparticles))
      XPAR = 1.0
      YPAR = 1.0
      ZPAR = 1.0
      VX = 0.0
      VY = 0.0
      VZ = 0.0
      if (allocated(Pk)) deallocate (Pk, dNharm, kharm, PkLog, dNharmLog, kharmLog)
      allocate (Pk(NGRID, 4), dNharm(NGRID, 4), kharm(NGRID, 4), PkLog(NGRID, 4), dNharmLog(NGRID, 4), &
                kharmLog(NGRID, 4))
      Pk = 0.0
      dNharm = 0.0
      kharm = 0.0
      PkLog = 0.0
      dNharmLog = 0.0
      kharmLog = 0.0
      ! POWERfft5 also writes its header lines to unit 17
      inquire (unit=17, opened=u17)
      if (.not. u17) open (unit=17, file='/dev/null', status='unknown')
   end subroutine

   subroutine ref_pm_put_particles(x, y, z, u, v, w) bind(C, name='ref_pm_put_particles')
      real(c_float), intent(in) :: x(*), y(*), z(*), u(*), v(*), w(*)
      XPAR(1:Nparticles) = x(1:Nparticles)
      YPAR(1:Nparticles) = y(1:Nparticles)
      ZPAR(1:Nparticles) = z(1:Nparticles)
      VX(1:Nparticles) = u(1:Nparticles)
      VY(1:Nparticles) = v(1:Nparticles)
      VZ(1:Nparticles) = w(1:Nparticles)
   end subroutine

   subroutine ref_pm_get_particles(x, y, z, u, v, w) bind(C, name='ref_pm_get_particles')
      real(c_float), intent(out) :: x(*), y(*), z(*), u(*), v(*), w(*)
      x(1:Nparticles) = XPAR(1:Nparticles)
      y(1:Nparticles) = YPAR(1:Nparticles)
      z(1:Nparticles) = ZPAR(1:Nparticles)
      u(1:Nparticles) = VX(1:Nparticles)
      v(1:Nparticles) = VY(1:Nparticles)
      w(1:Nparticles) = VZ(1:Nparticles)
   end subroutine

   ! which = 1: FI, 2: FI2 (ngrid^3 REAL*4 in the reference's order)
   subroutine ref_pm_put(which, src) bind(C, name='ref_pm_put')
      integer(c_int), value :: which
      real(c_float), intent(in) :: src(*)
      integer :: n3
      n3 = NGRID*NGRID*NGRID
      if (which == 1) then
         FI = reshape(src(1:n3), [NGRID, NGRID, NGRID])
      else
         FI2 = reshape(src(1:n3), [NGRID, NGRID, NGRID])
      end if
   end subroutine

   subroutine ref_pm_get(which, dst) bind(C, name='ref_pm_get')
      integer(c_int), value :: which
      real(c_float), intent(out) :: dst(*)
      integer :: n3
      n3 = NGRID*NGRID*NGRID
      if (which == 1) then
         dst(1:n3) = reshape(FI, [n3])
      else
         dst(1:n3) = reshape(FI2, [n3])
      end if
   end subroutine

   subroutine ref_pm_densit() bind(C, name='ref_pm_densit')
      call DENSIT
   end subroutine

   subroutine ref_pm_potentfft5() bind(C, name='ref_pm_potentfft5')
      call POTENTfft5
   end subroutine

   subroutine ref_pm_potent_for_mg() bind(C, name='ref_pm_potent_for_mg')
      call POTENT_FOR_MG
   end subroutine

   ! MOVE(FactV) with a default-REAL FactV as PMP2main.f90:48-58 passes it;
   ! returns ENKIN (:689)
   real(c_double) function ref_pm_move(factv_) bind(C, name='ref_pm_move')
      real(c_double), value :: factv_
      real :: fv
      fv = real(factv_)
      call MOVE(fv)
      ref_pm_move = ENKIN
   end function

   ! POWERfft5 of FI (destroyed) with iPol and the k-space binning icic;
   ! the NGRID/2 linear bins (k, P, harmonics) and the NGRID/2 log bins
   subroutine ref_pm_power(ipol, iswitch, icic_, k, p, cnt, klog, plog, cntlog) bind(C, name='ref_pm_power')
      integer(c_int), value :: ipol, iswitch, icic_
      real(c_double), intent(out) :: k(*), p(*), cnt(*), klog(*), plog(*), cntlog(*)
      integer :: ii, i, ip, isw
      ip = ipol
      isw = iswitch
      iCIC = icic_
      ii = 2*isw - ip
      Pk(:, ii) = 0.0
      dNharm(:, ii) = 0.0
      kharm(:, ii) = 0.0
      PkLog(:, ii) = 0.0
      dNharmLog(:, ii) = 0.0
      kharmLog(:, ii) = 0.0
      call POWERfft5(ip, isw)
      do i = 1, NGRID/2
         k(i) = kharm(i, ii)
         p(i) = Pk(i, ii)
         cnt(i) = dNharm(i, ii)
         klog(i) = kharmLog(i, ii)
         plog(i) = PkLog(i, ii)
         cntlog(i) = dNharmLog(i, ii)
      end do
   end subroutine

end module ref_pm_harness
