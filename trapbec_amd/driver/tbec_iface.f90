! iso_c_binding interface to libtbec (include/tbec.h).
!
! The reference's host program (src/run.f90) owns Psi, rhs and w as module arrays and
! calls get_rhs (src/finite_diff.f90:8) directly.  Here the device state lives behind a
! libtbec handle; this module declares the C entry points and the parameter and
! diagnostics structs with the same layout as the C header.
module tbec_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: TBEC_F32 = 32, TBEC_F64 = 64
  integer(c_int), parameter :: TBEC_SUM_REFERENCE = 0, TBEC_SUM_COMPENSATED = 1, &
                                TBEC_SUM_CENTRED = 2
  integer(c_int), parameter :: TBEC_BC_ZERO = 0, TBEC_BC_SOLID = 1, TBEC_BC_PERIODIC = 2
  integer(c_int), parameter :: TBEC_REAL_RK3 = 0, TBEC_IMAG_RK3 = 1, TBEC_IMAG_EULER_REF = 2
  ! the reference's `potential` names (src/finite_diff.f90:30-117), plus a caller's table
  integer(c_int), parameter :: TBEC_POT_ZERO = 0, TBEC_POT_HARMONIC = 1, &
    TBEC_POT_HARMONIC_RING = 2, TBEC_POT_HARMONIC_RING2 = 3, TBEC_POT_BARRIER = 4, &
    TBEC_POT_BOX_TRAP = 5, TBEC_POT_CHANNEL = 6, TBEC_POT_QUARTIC = 7, TBEC_POT_SPHERE = 8, &
    TBEC_POT_MOBIUS = 9, TBEC_POT_TABLE = 10
  integer(c_int), parameter :: TBEC_INSERT_PCURRENT = 0, TBEC_INSERT_RING = 1, &
    TBEC_INSERT_DIPOLE = 2, TBEC_INSERT_PHASE_IMPRINT = 3, TBEC_INSERT_NOISE = 4

  type, bind(C) :: tbec_params
    integer(c_int32_t) :: nx, ny, nz
    integer(c_int32_t) :: prec, sum_mode, tile
    real(c_double) :: Lx, Ly, Lz
    real(c_double) :: harm_wx, harm_wy, harm_wz
    real(c_double) :: omega_z
    real(c_double) :: c0, gnl
    real(c_double) :: dt_re, dt_im
    real(c_double) :: gamma
  end type

  type, bind(C) :: tbec_diag
    real(c_double) :: norm, energy, mu, residual
    real(c_double) :: max_abs, min_abs, mean_abs, int_psi2
  end type

  type, bind(C) :: tbec_physics
    integer(c_int32_t) :: potential, lbeam, two_dim, pad_
    real(c_double) :: wbeam, potential_off, t_spindown, advective_vx
  end type

  type, bind(C) :: tbec_reference_ic
    integer(c_int32_t) :: kind, winding_number, vor_ring_count
    integer(c_int32_t) :: seed
    real(c_double) :: ring_rad
  end type

  type, bind(C) :: tbec_solve_result
    integer(c_int64_t) :: steps
    integer(c_int32_t) :: converged, pad_
    real(c_double) :: energy, mu, residual, norm, rel_change
  end type

  interface
    integer(c_int) function tbec_create(p, h) bind(C, name="tbec_create")
      import :: c_int, c_ptr, tbec_params
      type(tbec_params), intent(in) :: p
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function tbec_destroy(h) bind(C, name="tbec_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function tbec_upload_psi(h, psi) bind(C, name="tbec_upload_psi")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, psi
    end function
    integer(c_int) function tbec_download_psi(h, psi) bind(C, name="tbec_download_psi")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, psi
    end function
    integer(c_int) function tbec_set_boundaries(h, bc, periodicz) &
        bind(C, name="tbec_set_boundaries")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: bc, periodicz
    end function
    integer(c_int) function tbec_set_physics(h, ph) bind(C, name="tbec_set_physics")
      import :: c_ptr, c_int, tbec_physics
      type(c_ptr), value :: h
      type(tbec_physics), intent(in) :: ph
    end function
    integer(c_int) function tbec_set_potential_table(h, trap) &
        bind(C, name="tbec_set_potential_table")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, trap
    end function
    integer(c_int) function tbec_imprint_vortex(h, x0, y0, dir) &
        bind(C, name="tbec_imprint_vortex")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), value :: x0, y0, dir
    end function
    integer(c_int) function tbec_reference_state(h, ic) bind(C, name="tbec_reference_state")
      import :: c_ptr, c_int, tbec_reference_ic
      type(c_ptr), value :: h
      type(tbec_reference_ic), intent(in) :: ic
    end function
    integer(c_int) function tbec_reference_insert(h, what, ic) &
        bind(C, name="tbec_reference_insert")
      import :: c_ptr, c_int, tbec_reference_ic
      type(c_ptr), value :: h
      integer(c_int), value :: what
      type(tbec_reference_ic), intent(in) :: ic
    end function
    integer(c_int) function tbec_velocity(h, vel) bind(C, name="tbec_velocity")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, vel
    end function
    integer(c_int) function tbec_qpressure(h, qp) bind(C, name="tbec_qpressure")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, qp
    end function
    integer(c_int) function tbec_step(h, nsteps, mode, t) bind(C, name="tbec_step")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int64_t), value :: nsteps
      integer(c_int), value :: mode
      real(c_double), intent(inout) :: t
    end function
    integer(c_int) function tbec_diagnose(h, d) bind(C, name="tbec_diagnose")
      import :: c_int, c_ptr, tbec_diag
      type(c_ptr), value :: h
      type(tbec_diag), intent(out) :: d
    end function
    integer(c_int) function tbec_imag_solve(h, max_steps, tol, norm, check_every, r) &
        bind(C, name="tbec_imag_solve")
      import :: c_int, c_ptr, c_int64_t, c_double, tbec_solve_result
      type(c_ptr), value :: h
      integer(c_int64_t), value :: max_steps
      real(c_double), value :: tol, norm
      integer(c_int), value :: check_every
      type(tbec_solve_result), intent(out) :: r
    end function
    integer(c_int) function tbec_sync(h) bind(C, name="tbec_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function tbec_last_error_c() bind(C, name="tbec_last_error")
      import :: c_ptr
    end function
  end interface

contains

  ! The library's last error message as a Fortran string.
  function tbec_last_error() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    p = tbec_last_error_c()
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function
end module
