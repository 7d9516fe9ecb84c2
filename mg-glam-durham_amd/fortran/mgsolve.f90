! mgsolve.f90 -- Fortran (ISO_C_BINDING) interface to libmgsolve's C-ABI.
!
! One bind(C) interface per entry point of include/mgsolve.h, and the three
! parameter/report structures as interoperable derived types (same member
! order, so the same layout as the C structs).  Arrays are passed as
! type(c_ptr) (c_loc of a contiguous real(4) or real(8) array in the
! reference's FI/FI2 memory order, PMP2mod_tools.f90:40-42).
module mgsolve
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: MGS_OK = 0
  integer(c_int32_t), parameter :: MGS_MODEL_FR = 1, MGS_MODEL_DGP = 2
  integer(c_int32_t), parameter :: MGS_PREC_FP64 = 0, MGS_PREC_MIXED = 1
  integer(c_int32_t), parameter :: MGS_F32 = 0, MGS_F64 = 1
  integer(c_int32_t), parameter :: MGS_HOST = 0, MGS_DEVICE = 1
  integer(c_int32_t), parameter :: MGS_FLAG_REFERENCE_ROOTS = 1

  type, bind(C) :: mgs_params
    integer(c_int32_t) :: ngrid, model, fr_n, branch
    real(c_double) :: box, omega_m, omega_l, aexpn, fr0, h0rc
    integer(c_int32_t) :: precision, flags
  end type mgs_params

  type, bind(C) :: mgs_solve_opts
    integer(c_int32_t) :: mode, max_cycles
    real(c_double) :: res_conv, rel_tol
    integer(c_int32_t) :: init_guess, inner_cycles
    integer(c_int32_t) :: cycle, reserved
  end type mgs_solve_opts

  type, bind(C) :: mgs_solve_report
    integer(c_int32_t) :: cycles, outer, converged, reserved
    real(c_double) :: res0, res_final, seconds
    real(c_double) :: history(128)
  end type mgs_solve_report

  interface
    integer(c_int) function mgs_create(params, h) bind(C, name='mgs_create')
      import :: c_int, c_ptr, mgs_params
      type(mgs_params), intent(in) :: params
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function mgs_destroy(h) bind(C, name='mgs_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function mgs_last_error(h) bind(C, name='mgs_last_error')
      import :: c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mgs_levels(h, levelmax, levelmin, coarsest) bind(C, name='mgs_levels')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: levelmax, levelmin, coarsest
    end function
    integer(c_int) function mgs_set_params(h, params) bind(C, name='mgs_set_params')
      import :: c_int, c_ptr, mgs_params
      type(c_ptr), value :: h
      type(mgs_params), intent(in) :: params
    end function
    integer(c_int) function mgs_set_density(h, delta, dtype, where) bind(C, name='mgs_set_density')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h, delta
      integer(c_int32_t), value :: dtype, where
    end function
    integer(c_int) function mgs_set_field(h, u, dtype, where) bind(C, name='mgs_set_field')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h, u
      integer(c_int32_t), value :: dtype, where
    end function
    integer(c_int) function mgs_get_field(h, u, dtype, where, scale) bind(C, name='mgs_get_field')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: h, u
      integer(c_int32_t), value :: dtype, where
      real(c_double), value :: scale
    end function
    integer(c_int) function mgs_set_level_array(h, level, which, src, dtype, where) &
        bind(C, name='mgs_set_level_array')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h, src
      integer(c_int32_t), value :: level, which, dtype, where
    end function
    integer(c_int) function mgs_get_level_array(h, level, which, dst, dtype, where) &
        bind(C, name='mgs_get_level_array')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h, dst
      integer(c_int32_t), value :: level, which, dtype, where
    end function
    integer(c_int) function mgs_get_density(h, delta, dtype, where) bind(C, name='mgs_get_density')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h, delta
      integer(c_int32_t), value :: dtype, where
    end function
    integer(c_int) function mgs_background(h, fr_bg, u_bg) bind(C, name='mgs_background')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: fr_bg, u_bg
    end function
    integer(c_int) function mgs_initialise(h) bind(C, name='mgs_initialise')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mgs_smooth(h, level, red) bind(C, name='mgs_smooth')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: level, red
    end function
    integer(c_int) function mgs_residual(h, level, rms) bind(C, name='mgs_residual')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: level
      real(c_double), intent(out) :: rms
    end function
    integer(c_int) function mgs_restrict_residual(h, level) bind(C, name='mgs_restrict_residual')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: level
    end function
    integer(c_int) function mgs_restrict_field(h, level) bind(C, name='mgs_restrict_field')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: level
    end function
    integer(c_int) function mgs_restrict_density(h, level) bind(C, name='mgs_restrict_density')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: level
    end function
    integer(c_int) function mgs_rhs(h, level) bind(C, name='mgs_rhs')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: level
    end function
    integer(c_int) function mgs_correct(h, level) bind(C, name='mgs_correct')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: level
    end function
    integer(c_int) function mgs_vcycle(h, level) bind(C, name='mgs_vcycle')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: level
    end function
    integer(c_int) function mgs_fcycle(h, level) bind(C, name='mgs_fcycle')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: level
    end function
    integer(c_int) function mgs_wcycle(h, level) bind(C, name='mgs_wcycle')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: level
    end function
    integer(c_int) function mgs_solve(h, opts, report) bind(C, name='mgs_solve')
      import :: c_int, c_ptr, mgs_solve_opts, mgs_solve_report
      type(c_ptr), value :: h
      type(mgs_solve_opts), intent(in) :: opts
      type(mgs_solve_report), intent(out) :: report
    end function
    integer(c_int) function mgs_fifth_force(h, psi, gx, gy, gz, dtype, where) bind(C, name='mgs_fifth_force')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h, psi, gx, gy, gz
      integer(c_int32_t), value :: dtype, where
    end function
    integer(c_int) function mgs_synchronize(h) bind(C, name='mgs_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    ! PM step around the solve (particles, CIC deposit, Poisson FFT, push)
    integer(c_int) function mgs_particles_set(h, x, y, z, vx, vy, vz, npart, where) bind(C, name='mgs_particles_set')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: h, x, y, z, vx, vy, vz
      integer(c_int64_t), value :: npart
      integer(c_int32_t), value :: where
    end function
    integer(c_int) function mgs_particles_get(h, x, y, z, vx, vy, vz, where) bind(C, name='mgs_particles_get')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h, x, y, z, vx, vy, vz
      integer(c_int32_t), value :: where
    end function
    integer(c_int) function mgs_cic_deposit(h) bind(C, name='mgs_cic_deposit')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mgs_potential(h, trfi) bind(C, name='mgs_potential')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: trfi
    end function
    integer(c_int) function mgs_potential_add_fifth_force(h) bind(C, name='mgs_potential_add_fifth_force')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mgs_get_potential(h, dst, dtype, where) bind(C, name='mgs_get_potential')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h, dst
      integer(c_int32_t), value :: dtype, where
    end function
    integer(c_int) function mgs_set_potential(h, src, dtype, where) bind(C, name='mgs_set_potential')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h, src
      integer(c_int32_t), value :: dtype, where
    end function
    integer(c_int) function mgs_move(h, pconst, xconst, svel) bind(C, name='mgs_move')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: pconst, xconst
      real(c_double), intent(out) :: svel
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! mgs_last_error as a Fortran string
  function mgs_error_text(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    p = mgs_last_error(h)
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function mgs_error_text

end module mgsolve
