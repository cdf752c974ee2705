! ISO_C_BINDING interface to libaxgrav (include/axgrav.h).
!
! A Fortran host (the reference's main.f90/settings.f90 driver, or axg_driver.f90 here) uses
! this module to run the lattice step and diagnostics on an MI355X.  Every function returns
! the C status (0 = ok); axg_check() turns a failure into the message of axg_last_error()
! and a stop, the way the reference's MpiStop does.
module axgrav_capi
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: AXG_OK = 0, AXG_FP64 = 0, AXG_FP32C = 1, AXG_FP32 = 2

  ! struct axg_params
  type, bind(C) :: axg_params
     integer(c_int64_t) :: ngrid = 0
     integer(c_int32_t) :: precision = 0
     integer(c_int32_t) :: grav_spec = 0
     integer(c_int32_t) :: initial_loop = 0
     integer(c_int32_t) :: gw_fix_swap = 0
     integer(c_int32_t) :: num_bin_grav = 25
     integer(c_int32_t) :: num_bin = 50
     integer(c_int32_t) :: id_strings = 0
     integer(c_int32_t) :: id_walls = 0
     integer(c_int32_t) :: spec_est = 0
     integer(c_int32_t) :: device = 0
     real(c_double) :: dspa = 0, dwav = 0, dv_spa = 0, dv_wav = 0
     real(c_double) :: sigma_cr = 0, tau_cr = 0, a_now = 0
     real(c_double) :: kappa = 0
     real(c_double) :: eps_dtau = 0
     real(c_double) :: dtau_fixed = 0
     real(c_double) :: potential_kappa0 = 0
     integer(c_int32_t) :: degrade_spec = 0
     integer(c_int32_t) :: string_vel = 0
     integer(c_int32_t) :: reserved(6) = 0
  end type axg_params

  ! struct axg_keys
  type, bind(C) :: axg_keys
     real(c_double) :: sigma, g_rel, kappa, tau_fin, boxsize
     integer(c_int64_t) :: ngrid
  end type axg_keys

  ! struct axg_defect_out
  type, bind(C) :: axg_defect_out
     integer(c_int64_t) :: n_string = 0, n_wall = 0, n_mask = 0
  end type axg_defect_out

  ! struct axg_plaquette_out: per face orientation (x,y), (y,z), (z,x)
  type, bind(C) :: axg_plaquette_out
     integer(c_int64_t) :: n_pierced(3) = 0, n_plus(3) = 0, n_minus(3) = 0, n_quadrant(3) = 0, net(3) = 0
  end type axg_plaquette_out

  ! struct axg_string_out
  type, bind(C) :: axg_string_out
     integer(c_int64_t) :: n_string = 0, n_vel = 0
     real(c_double) :: length = 0
     real(c_double) :: vel(3) = 0
  end type axg_string_out

  ! struct axg_wall_out
  type, bind(C) :: axg_wall_out
     integer(c_int64_t) :: n_wall = 0
     real(c_double) :: area = 0
  end type axg_wall_out

  ! struct axg_group_out
  type, bind(C) :: axg_group_out
     integer(c_int64_t) :: n_group = 0
     real(c_double) :: th_dist = 0
     real(c_double) :: sum(3) = 0
  end type axg_group_out

  interface
     integer(c_int) function axg_derive(k, p, zeta, width_cr, thickness_cr) bind(C, name="axg_derive")
       import :: c_int, c_double, axg_keys, axg_params
       type(axg_keys), intent(in) :: k
       type(axg_params), intent(inout) :: p
       real(c_double), intent(out) :: zeta, width_cr, thickness_cr
     end function axg_derive

     integer(c_int) function axg_get_dtau(p, t1, t2, dtau, nsteps) bind(C, name="axg_get_dtau")
       import :: c_int, c_int32_t, c_double, axg_params
       type(axg_params), intent(in) :: p
       real(c_double), value :: t1, t2
       real(c_double), intent(out) :: dtau
       integer(c_int32_t), intent(out) :: nsteps
     end function axg_get_dtau

     integer(c_int) function axg_create(p, h) bind(C, name="axg_create")
       import :: c_int, c_ptr, axg_params
       type(axg_params), intent(in) :: p
       type(c_ptr), intent(out) :: h
     end function axg_create

     integer(c_int) function axg_destroy(h) bind(C, name="axg_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function axg_destroy

     type(c_ptr) function axg_last_error_c(h) bind(C, name="axg_last_error")
       import :: c_ptr
       type(c_ptr), value :: h
     end function axg_last_error_c

     integer(c_int) function axg_set_fields(h, phi, pi, nsite) bind(C, name="axg_set_fields")
       import :: c_int, c_ptr, c_int64_t, c_double_complex
       type(c_ptr), value :: h
       complex(c_double_complex), intent(in) :: phi(*), pi(*)
       integer(c_int64_t), value :: nsite
     end function axg_set_fields

     integer(c_int) function axg_get_fields(h, phi, pi, nsite) bind(C, name="axg_get_fields")
       import :: c_int, c_ptr, c_int64_t, c_double_complex
       type(c_ptr), value :: h
       complex(c_double_complex), intent(out) :: phi(*), pi(*)
       integer(c_int64_t), value :: nsite
     end function axg_get_fields

     integer(c_int) function axg_set_stream(h, seed1, seed2, position) bind(C, name="axg_set_stream")
       import :: c_int, c_ptr, c_int32_t, c_int64_t
       type(c_ptr), value :: h
       integer(c_int32_t), value :: seed1, seed2   ! uint32_t in C: the same bits
       integer(c_int64_t), value :: position
     end function axg_set_stream

     integer(c_int) function axg_get_stream(h, position) bind(C, name="axg_get_stream")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: h
       integer(c_int64_t), intent(out) :: position
     end function axg_get_stream

     integer(c_int) function axg_plaquettes(h, out) bind(C, name="axg_plaquettes")
       import :: c_int, c_ptr, axg_plaquette_out
       type(c_ptr), value :: h
       type(axg_plaquette_out), intent(out) :: out
     end function axg_plaquettes

     ! rec(5, cap): (ix, iy, iz, o, n) per pierced face; cap = 0 is a size query
     integer(c_int) function axg_get_plaquettes(h, rec, cap, count) bind(C, name="axg_get_plaquettes")
       import :: c_int, c_ptr, c_int32_t, c_int64_t
       type(c_ptr), value :: h
       integer(c_int32_t), intent(out) :: rec(*)
       integer(c_int64_t), value :: cap
       integer(c_int64_t), intent(out) :: count
     end function axg_get_plaquettes

     integer(c_int) function axg_init_flat(h, seed, kmax) bind(C, name="axg_init_flat")
       import :: c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: h
       integer(c_int64_t), value :: seed
       real(c_double), value :: kmax
     end function axg_init_flat

     integer(c_int) function axg_init_thermal(h, seed, tau_ini) bind(C, name="axg_init_thermal")
       import :: c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: h
       integer(c_int64_t), value :: seed
       real(c_double), value :: tau_ini
     end function axg_init_thermal

     integer(c_int) function axg_get_gw_state(h, a, count) bind(C, name="axg_get_gw_state")
       import :: c_int, c_ptr, c_int64_t, c_double_complex
       type(c_ptr), value :: h
       complex(c_double_complex), intent(out) :: a(*)
       integer(c_int64_t), value :: count
     end function axg_get_gw_state

     integer(c_int) function axg_set_gw_state(h, a, count) bind(C, name="axg_set_gw_state")
       import :: c_int, c_ptr, c_int64_t, c_double_complex
       type(c_ptr), value :: h
       complex(c_double_complex), intent(in) :: a(*)
       integer(c_int64_t), value :: count
     end function axg_set_gw_state

     integer(c_int) function axg_advance(h, t1, t2, nsteps) bind(C, name="axg_advance")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: h
       real(c_double), value :: t1, t2
       integer(c_int32_t), intent(out) :: nsteps
     end function axg_advance

     ! the step kernels' launch configuration, timed on the live lattice (results unchanged)
     integer(c_int) function axg_tune(h, reps, variant, ms, nms) bind(C, name="axg_tune")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: h
       integer(c_int32_t), value :: reps
       integer(c_int32_t), intent(out) :: variant
       type(c_ptr), value :: ms
       integer(c_int32_t), value :: nms
     end function axg_tune

     integer(c_int) function axg_advance_steps(h, t1, dt, nsteps) bind(C, name="axg_advance_steps")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: h
       real(c_double), value :: t1, dt
       integer(c_int32_t), value :: nsteps
     end function axg_advance_steps

     integer(c_int) function axg_gw_spectrum(h, qmid, power, numer, denom, nbins) bind(C, name="axg_gw_spectrum")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: h
       real(c_double), intent(out) :: qmid(*), power(*), numer(*), denom(*)
       integer(c_int32_t), value :: nbins
     end function axg_gw_spectrum

     integer(c_int) function axg_defects(h, out) bind(C, name="axg_defects")
       import :: c_int, c_ptr, axg_defect_out
       type(c_ptr), value :: h
       type(axg_defect_out), intent(out) :: out
     end function axg_defects

     integer(c_int) function axg_get_mask(h, ixyz, cap, count) bind(C, name="axg_get_mask")
       import :: c_int, c_ptr, c_int32_t, c_int64_t
       type(c_ptr), value :: h
       integer(c_int32_t), intent(out) :: ixyz(*)
       integer(c_int64_t), value :: cap
       integer(c_int64_t), intent(out) :: count
     end function axg_get_mask

     integer(c_int) function axg_get_flags(h, flags, nsite) bind(C, name="axg_get_flags")
       import :: c_int, c_ptr, c_int8_t, c_int64_t
       type(c_ptr), value :: h
       integer(c_int8_t), intent(out) :: flags(*)
       integer(c_int64_t), value :: nsite
     end function axg_get_flags

     integer(c_int) function axg_strings(h, seed0, seed1, out) bind(C, name="axg_strings")
       import :: c_int, c_ptr, c_int32_t, axg_string_out
       type(c_ptr), value :: h
       integer(c_int32_t), value :: seed0, seed1   ! uint32_t in C: the same bits
       type(axg_string_out), intent(out) :: out
     end function axg_strings

     integer(c_int) function axg_get_strings(h, rec, cap, count) bind(C, name="axg_get_strings")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: h
       real(c_double), intent(out) :: rec(*)
       integer(c_int64_t), value :: cap
       integer(c_int64_t), intent(out) :: count
     end function axg_get_strings

     integer(c_int) function axg_walls(h, out) bind(C, name="axg_walls")
       import :: c_int, c_ptr, axg_wall_out
       type(c_ptr), value :: h
       type(axg_wall_out), intent(out) :: out
     end function axg_walls

     integer(c_int) function axg_get_walls(h, rec, cap, count) bind(C, name="axg_get_walls")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: h
       real(c_double), intent(out) :: rec(*)
       integer(c_int64_t), value :: cap
       integer(c_int64_t), intent(out) :: count
     end function axg_get_walls

     integer(c_int) function axg_group_strings(h, width, hor_size, out) bind(C, name="axg_group_strings")
       import :: c_int, c_ptr, c_double, axg_group_out
       type(c_ptr), value :: h
       real(c_double), value :: width, hor_size
       type(axg_group_out), intent(out) :: out
     end function axg_group_strings

     integer(c_int) function axg_group_walls(h, thickness, hor_size, out) bind(C, name="axg_group_walls")
       import :: c_int, c_ptr, c_double, axg_group_out
       type(c_ptr), value :: h
       real(c_double), value :: thickness, hor_size
       type(axg_group_out), intent(out) :: out
     end function axg_group_walls

     integer(c_int) function axg_get_groups(h, kind, ids, cap, count) bind(C, name="axg_get_groups")
       import :: c_int, c_ptr, c_int32_t, c_int64_t
       type(c_ptr), value :: h
       integer(c_int32_t), value :: kind
       integer(c_int32_t), intent(out) :: ids(*)
       integer(c_int64_t), value :: cap
       integer(c_int64_t), intent(out) :: count
     end function axg_get_groups

     integer(c_int) function axg_field_spectrum(h, tau, qmid, power, conv, num, rho, nbins) &
          bind(C, name="axg_field_spectrum")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: h
       real(c_double), value :: tau
       real(c_double), intent(out) :: qmid(*), power(*), conv(*), num(*), rho
       integer(c_int32_t), value :: nbins
     end function axg_field_spectrum

     integer(c_int) function axg_field_covariance(h, seed, min_mc, max_mc, tol, cov, ncov, iters) &
          bind(C, name="axg_field_covariance")
       import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
       type(c_ptr), value :: h
       integer(c_int64_t), value :: seed
       integer(c_int32_t), value :: min_mc, max_mc
       real(c_double), value :: tol
       real(c_double), intent(out) :: cov(*)
       integer(c_int64_t), value :: ncov
       integer(c_int32_t), intent(out) :: iters
     end function axg_field_covariance

     integer(c_int) function axg_energy(h, tau, out) bind(C, name="axg_energy")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: tau
       real(c_double), intent(out) :: out(4)
     end function axg_energy
  end interface

contains

  ! Message of the last failure on h (or of the last axg_create when h is null).
  function axg_last_error(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=512) :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ''
    p = axg_last_error_c(h)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [512])
    do i = 1, 512
       if (s(i) == c_null_char) exit
       msg(i:i) = s(i)
    end do
  end function axg_last_error

  ! Stop with the library's message when rc is not AXG_OK (the reference's MpiStop).
  subroutine axg_check(rc, h, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: h
    character(*), intent(in) :: what
    if (rc /= AXG_OK) then
       write(*, '(A,I0,A)') ' error: '//trim(what)//' failed (', rc, '): '//trim(axg_last_error(h))
       stop 1
    end if
  end subroutine axg_check

end module axgrav_capi
