! tbec_run: the Fortran host program for libtbec.
!
! It keeps the reference's driver role (src/run.f90:1-81): run parameters (here a
! runtime namelist instead of the compile-time src/cparam.local), the initial state
! (src/initial.f90), the imaginary-time stage (src/run.f90:19-35), the real-time loop
! with its output cadence (src/run.f90:42-76), the per-shot diagnostics line
! (src/run.f90:60-63), the var files and restart file (src/output.f90:71-108) and the
! timing log (src/output.f90:111-118).  Every step runs on the GPU through libtbec.
!
!   tbec_run [namelist file]       (default: tbec.nml)
!
! Errors follow the reference's emergency_stop (src/mpi_var.f90:45-62): message on
! stdout, the text in ./ERROR, non-zero exit.
program tbec_run
  use iso_c_binding
  use iso_fortran_env, only: real32, real64, int64
  use, intrinsic :: ieee_arithmetic, only: ieee_value, ieee_positive_inf
  use tbec_iface
  use tbec_ic
  implicit none

  ! ---- namelist (reference names where the reference has the parameter) ----
  integer :: Nx = 64, Ny = 64, Nz = 64
  real(dp) :: Lx = 16.0_dp, Ly = 16.0_dp, Lz = 16.0_dp
  real(dp) :: dt_re = 1.0e-3_dp, dt_im = 0.0_dp, gamma = 0.0_dp
  integer :: nsteps = 1000, shots = 1000
  real(dp) :: harm_wx = 256.0_dp, harm_wy = 256.0_dp, harm_wz = 256.0_dp
  real(dp) :: omega_z = 0.0_dp, c0 = 1.0_dp, gnl = 1.0_dp
  integer :: prec = 64, sum_mode = -1, tile = -1
  character(len=32) :: init_cond = 'tf_vortex'
  real(dp) :: trap_omega(3) = [1.0_dp, 1.0_dp, 1.0_dp]
  real(dp) :: mu = 10.0_dp, g_phys = 500.0_dp, tf_norm = 0.0_dp, psi_scale = 1.0_dp
  integer :: n_lines = 1
  real(dp) :: line_x(8) = 0.0_dp, line_y(8) = 0.0_dp
  integer :: line_q(8) = 1
  real(dp) :: xi = 0.0_dp, ring_radius = 0.0_dp
  integer(int64) :: seed = 20261203_int64
  real(dp) :: noise_amp = 0.0_dp, phase_noise_amp = 0.0_dp
  logical :: imaginary_time = .false., ic_only = .false., full_binary_output = .true.
  logical :: vel_print = .false., qpress_print = .false., periodicz = .false.
  character(len=16) :: boundaries = 'zero'
  ! physics of get_rhs (src/finite_diff.f90:30-130) and the forcing (src/run.f90:70-74)
  character(len=32) :: potential = 'harmonic'
  logical :: lbeam = .false., forcing = .false., two_dim = .false.
  real(dp) :: wbeam = 4.0_dp, advective_vx = 0.0_dp
  real(dp) :: potential_off = huge(1.0_dp), t_spindown = huge(1.0_dp)
  integer :: force_freq = 10
  ! the reference's own initial states and insertions (src/initial.f90:58-156,669-870)
  integer :: winding_number = 1, vor_ring_count = 1
  real(dp) :: ring_rad = 4.0_dp
  logical :: add_pcurrent = .false., add_ring = .false., add_dipole = .false.
  character(len=256) :: trap_file = ''
  integer :: imag_nsteps = 0, imag_check = 50
  real(dp) :: imag_tol = 1.0e-10_dp, imag_norm = 1.0_dp
  ! imaginary-time scheme: 'rk3' (renormalised RK3 to imag_tol, the default) or 'euler'
  ! (the reference's own unnormalised Euler for imag_nsteps steps, src/run.f90:19-33);
  ! imaginary_output writes imag_var%04d.dat snapshots (src/output.f90:62-69)
  character(len=8) :: imag_scheme = 'rk3'
  logical :: imaginary_output = .false.
  character(len=256) :: out_dir = 'data'
  namelist /tbec_run/ Nx, Ny, Nz, Lx, Ly, Lz, dt_re, dt_im, gamma, nsteps, shots, &
    harm_wx, harm_wy, harm_wz, omega_z, c0, gnl, prec, sum_mode, tile, init_cond, &
    trap_omega, mu, g_phys, tf_norm, psi_scale, n_lines, line_x, line_y, line_q, xi, &
    ring_radius, seed, noise_amp, phase_noise_amp, imaginary_time, ic_only, &
    full_binary_output, imag_nsteps, imag_check, imag_tol, imag_norm, out_dir, &
    vel_print, qpress_print, boundaries, periodicz, potential, lbeam, wbeam, &
    potential_off, t_spindown, advective_vx, forcing, force_freq, trap_file, two_dim, &
    winding_number, vor_ring_count, ring_rad, add_pcurrent, add_ring, add_dipole, &
    imag_scheme, imaginary_output

  character(len=512) :: nml_file
  integer :: u, ios, itime, nstart, l
  integer(int64) :: nforce = 0
  integer :: ref_kind = -1
  type(tbec_reference_ic) :: ric
  real(dp), allocatable :: xx(:), yy(:), zz(:)
  complex(dp), allocatable, target :: psi(:, :, :)
  complex(real32), allocatable, target :: psi4(:, :, :)
  real(dp), allocatable, target :: fld(:, :, :, :)
  real(real32), allocatable, target :: fld4(:, :, :, :)
  type(tbec_params) :: p
  type(tbec_diag) :: d
  type(tbec_solve_result) :: sr
  type(c_ptr) :: h
  real(c_double) :: t
  integer(int64) :: c_start, c_end, c_rate
  real(dp) :: time_per_step

  nml_file = 'tbec.nml'
  if (command_argument_count() >= 1) call get_command_argument(1, nml_file)
  open(newunit=u, file=trim(nml_file), status='old', action='read', iostat=ios)
  if (ios /= 0) call emergency_stop('cannot open namelist file '//trim(nml_file))
  read(u, nml=tbec_run, iostat=ios)
  if (ios /= 0) call emergency_stop('bad namelist in '//trim(nml_file))
  close(u)
  if (prec /= 32 .and. prec /= 64) call emergency_stop('prec must be 32 or 64')
  if (shots < 1) call emergency_stop('shots must be positive')
  if (forcing .and. force_freq < 1) call emergency_stop('force_freq must be positive')
  if (trim(imag_scheme) /= 'rk3' .and. trim(imag_scheme) /= 'euler') &
    call emergency_stop('imag_scheme must be rk3 or euler')
  l = pot_kind()   ! a wrong name stops here, before the device is touched
  if (sum_mode < 0) then
    sum_mode = TBEC_SUM_REFERENCE
    if (prec == 32) sum_mode = TBEC_SUM_CENTRED
  end if

  write(*, '(a)') ' TrapBEC (MI355X) driver initialised'
  write(*, '(a,i4.4,a,i4.4,a,i4.4,a)') ' Running a ', Nx, 'x', Ny, 'x', Nz, ' GPE simulation'
  write(*, '(a,i2,a,i1)') ' precision: fp', prec, '  stencil sum mode: ', sum_mode
  ! the reference's stability warning (src/initial.f90:29-33)
  if (dt_re > (1.0_dp / 8.0_dp) * min(Lx / Nx, Ly / Ny, Lz / Nz)**2) then
    write(*, '(a)') ' WARNING TIMESTEP IS PROBABLY TOO LARGE!'
  else
    write(*, '(a,es13.6)') ' dt is below upper-limit of', &
      (1.0_dp / 8.0_dp) * min(Lx / Nx, Ly / Ny, Lz / Nz)**2
  end if

  allocate(xx(Nx), yy(Ny), zz(Nz), psi(Nz, Ny, Nx))
  call coords(Nx, Lx, xx)
  call coords(Ny, Ly, yy)
  call coords(Nz, Lz, zz)
  t = 0.0_c_double
  nstart = 1

  select case (trim(init_cond))
  case ('tf_vortex')
    if (tf_norm > 0.0_dp) mu = tf_mu_for_norm(xx, yy, zz, trap_omega, g_phys, &
                                              (Lx / Nx) * (Ly / Ny) * (Lz / Nz), tf_norm)
    call tf_profile(psi, xx, yy, zz, trap_omega, g_phys, mu)
    if (xi <= 0.0_dp) xi = 1.0_dp / sqrt(mu)
    if (ring_radius > 0.0_dp) call imprint_ring(psi, xx, yy, zz, ring_radius, 0.0_dp, xi)
    do l = 1, min(n_lines, 8)
      call imprint_line(psi, xx, yy, line_x(l), line_y(l), line_q(l), xi)
    end do
    if (noise_amp > 0.0_dp) call add_noise(psi, seed, noise_amp, 0_int64)
    if (phase_noise_amp > 0.0_dp) call add_phase_noise(psi, seed, phase_noise_amp, 0_int64)
    psi = psi * psi_scale
  case ('restart')
    call read_restart()
  case default
    ref_kind = ref_ic_kind()   ! generated on the device after the handle exists
  end select
  ric%kind = max(ref_kind, 0)
  ric%winding_number = winding_number
  ric%vor_ring_count = vor_ring_count
  ric%seed = int(iand(seed, 2147483647_int64), c_int32_t)
  ric%ring_rad = ring_rad

  call execute_command_line('mkdir -p '//trim(out_dir)//'/proc000')
  ! dims.log as dims_print writes it (src/output.f90:48-58): one domain
  open(newunit=u, file=trim(out_dir)//'/dims.log', status='replace')
  write(u, *) Nx
  write(u, *) Ny
  write(u, *) Nz
  write(u, *) 1
  write(u, *) 1
  write(u, *) Nx
  write(u, *) Ny
  write(u, *) Nz
  close(u)
  if (ic_only .and. ref_kind < 0) then
    itime = 0
    call write_var(0)
    write(*, '(a)') ' initial state written; ic_only set, stopping'
    stop
  end if

  p%nx = Nx; p%ny = Ny; p%nz = Nz
  p%prec = prec; p%sum_mode = sum_mode; p%tile = tile
  p%Lx = Lx; p%Ly = Ly; p%Lz = Lz
  p%harm_wx = harm_wx; p%harm_wy = harm_wy; p%harm_wz = harm_wz
  p%omega_z = omega_z; p%c0 = c0; p%gnl = gnl
  p%dt_re = dt_re; p%dt_im = dt_im; p%gamma = gamma
  call check(tbec_create(p, h))
  call set_bc(h)
  call set_phys(h)
  if (ref_kind >= 0) then
    call check(tbec_reference_state(h, ric))
    call download()
    if (ic_only) then
      itime = 0
      call write_var(0)
      write(*, '(a)') ' initial state written; ic_only set, stopping'
      call check(tbec_destroy(h))
      stop
    end if
  else
    call upload()
  end if

  if (imaginary_time .and. nstart == 1) then
    if (trim(imag_scheme) == 'euler') then
      call imaginary_stage_euler()
    else
      ! renormalised RK3 needs dt = (0, -dtau): run it on a second handle
      call imaginary_stage()
    end if
  end if
  ! src/run.f90:37-40
  if (add_pcurrent) call check(tbec_reference_insert(h, TBEC_INSERT_PCURRENT, ric))
  if (add_ring) call check(tbec_reference_insert(h, TBEC_INSERT_RING, ric))
  if (add_dipole) call check(tbec_reference_insert(h, TBEC_INSERT_DIPOLE, ric))

  call system_clock(c_start, c_rate)
  itime = nstart - 1
  do while (itime < nsteps)
    l = min(shots - mod(itime, shots), nsteps - itime)
    if (forcing) l = min(l, force_freq - mod(itime, force_freq))
    call check(tbec_step(h, int(l, c_int64_t), TBEC_REAL_RK3, t))
    itime = itime + l
    if (mod(itime, shots) == 0) then
      call check(tbec_diagnose(h, d))
      print *, itime / shots, t, d%max_abs, d%min_abs, d%mean_abs, d%int_psi2
      call download()
      call write_var(itime / shots)
      if (vel_print) call write_field('vel', itime / shots)
      if (qpress_print) call write_field('qpress', itime / shots)
    end if
    if (forcing) then
      if (mod(itime, force_freq) == 0) call force_random_point()
    end if
  end do
  call check(tbec_sync(h))
  call system_clock(c_end)
  time_per_step = real(c_end - c_start, dp) / real(c_rate, dp) / max(1, nsteps - nstart + 1)
  open(newunit=u, file=trim(out_dir)//'/timing.log')
  write(u, *) int(Nx, int64) * Ny * Nz, 1, time_per_step
  close(u)
  call check(tbec_destroy(h))

contains

  subroutine emergency_stop(msg)
    character(*), intent(in) :: msg
    integer :: ue
    print *, '*****FATAL ERROR ABORTING CODE*****'
    print *, msg
    open(newunit=ue, file='ERROR')
    write(ue, *) msg
    close(ue)
    error stop 1
  end subroutine

  subroutine check(rc)
    integer(c_int), intent(in) :: rc
    if (rc /= 0) call emergency_stop('libtbec: '//tbec_last_error())
  end subroutine

  subroutine upload()
    if (prec == 64) then
      call check(tbec_upload_psi(h, c_loc(psi)))
    else
      psi4 = cmplx(psi, kind=real32)
      call check(tbec_upload_psi(h, c_loc(psi4)))
    end if
  end subroutine

  subroutine download()
    if (prec == 64) then
      call check(tbec_download_psi(h, c_loc(psi)))
    else
      if (.not. allocated(psi4)) allocate(psi4(Nz, Ny, Nx))
      call check(tbec_download_psi(h, c_loc(psi4)))
      psi = cmplx(psi4, kind=dp)
    end if
  end subroutine

  subroutine imaginary_stage()
    type(tbec_params) :: q
    type(c_ptr) :: hi
    q = p
    q%dt_re = 0.0_c_double
    q%dt_im = -abs(dt_im)
    if (dt_im == 0.0_dp) q%dt_im = -abs(dt_re)
    call check(tbec_create(q, hi))
    call set_bc(hi)
    call set_phys(hi)
    if (prec == 64) then
      call check(tbec_upload_psi(hi, c_loc(psi)))
    else
      call check(tbec_upload_psi(hi, c_loc(psi4)))
    end if
    call check(tbec_imag_solve(hi, int(imag_nsteps, c_int64_t), imag_tol, imag_norm, &
                               imag_check, sr))
    write(*, '(a,i8,a,i2,a,es22.14,a,es12.4)') ' imaginary time: steps ', sr%steps, &
      ' converged ', sr%converged, ' E ', sr%energy, ' residual ', sr%residual
    if (prec == 64) then
      call check(tbec_download_psi(hi, c_loc(psi)))
    else
      call check(tbec_download_psi(hi, c_loc(psi4)))
      psi = cmplx(psi4, kind=dp)
    end if
    call check(tbec_destroy(hi))
    call upload()
    ! one snapshot of the solve's final state, numbered by the shots it spans
    if (imaginary_output) call write_snapshot('imag_var', int((sr%steps + shots - 1) / shots))
    write(*, '(a)') ' finished imaginary time propagation'
  end subroutine

  ! The reference's imaginary stage (src/run.f90:19-33): imag_nsteps unnormalised steps
  ! psi += dt*0.5*rhs on the main handle (the run's own dt); every `shots` steps the
  ! norm ratio of the last step, intPsi2/intPsi2_old (Psi_norm's sum of |psi|^2, the old
  ! value 1 before the first step), and with imaginary_output the imag_var snapshot.
  subroutine imaginary_stage_euler()
    integer :: it, n
    real(dp) :: n_old, n_new
    it = 0
    n_old = 1.0_dp
    do while (it < imag_nsteps)
      n = min(shots - mod(it, shots), imag_nsteps - it)
      if (n > 1) then
        call check(tbec_step(h, int(n - 1, c_int64_t), TBEC_IMAG_EULER_REF, t))
        it = it + n - 1
        call check(tbec_diagnose(h, d))
        n_old = d%int_psi2
      end if
      call check(tbec_step(h, 1_c_int64_t, TBEC_IMAG_EULER_REF, t))
      it = it + 1
      call check(tbec_diagnose(h, d))
      n_new = d%int_psi2
      if (mod(it, shots) == 0) then
        print *, it / shots, n_new / n_old
        if (imaginary_output) then
          call download()
          call write_snapshot('imag_var', it / shots)
        end if
      end if
      n_old = n_new
    end do
    call download()
    write(*, '(a)') ' finished imaginary time propagation'
  end subroutine

  ! <prefix>%04d.dat: one record coords(2), Lx, Ly, Lz, t, Psi (src/output.f90:65-67,
  ! 75-77), in the run's precision.
  subroutine write_snapshot(prefix, shot)
    character(len=*), intent(in) :: prefix
    integer, intent(in) :: shot
    character(len=512) :: fname
    integer :: uv
    integer(4) :: crd(2)
    crd = 0
    write(fname, '(a,a,a,i4.4,a)') trim(out_dir), '/proc000/', prefix, shot, '.dat'
    open(newunit=uv, file=trim(fname), status='replace', form='unformatted')
    if (prec == 64) then
      write(uv) crd, Lx, Ly, Lz, real(t, dp), psi
    else
      write(uv) crd, real(Lx, real32), real(Ly, real32), real(Lz, real32), &
        real(t, real32), cmplx(psi, kind=real32)
    end if
    close(uv)
  end subroutine

  ! var%04d.dat (write_snapshot); var.dat: [itime+1] [t] [Psi] (src/output.f90:103-107),
  ! written in the run's precision.
  subroutine write_var(shot)
    integer, intent(in) :: shot
    character(len=512) :: fname
    integer :: uv
    if (full_binary_output) call write_snapshot('var', shot)
    write(fname, '(a,a)') trim(out_dir), '/proc000/var.dat'
    open(newunit=uv, file=trim(fname), status='replace', form='unformatted')
    write(uv) int(itime + 1, 4)
    if (prec == 64) then
      write(uv) real(t, dp)
      write(uv) psi
    else
      write(uv) real(t, real32)
      write(uv) cmplx(psi, kind=real32)
    end if
    close(uv)
  end subroutine

  ! potential / lbeam / wbeam / potential_off / t_spindown / advective_vx as in
  ! src/cparam.local; huge() switch-off times mean never
  subroutine set_phys(hh)
    type(c_ptr), intent(in) :: hh
    type(tbec_physics) :: ph
    real(dp), allocatable, target :: tab(:, :, :)
    real(real32), allocatable, target :: tab4(:, :, :)
    integer :: ut
    ph%potential = pot_kind()
    ph%lbeam = merge(1, 0, lbeam)
    ph%two_dim = merge(1, 0, two_dim)
    ph%pad_ = 0
    ph%wbeam = wbeam
    ph%potential_off = never_if_huge(potential_off)
    ph%t_spindown = never_if_huge(t_spindown)
    ph%advective_vx = advective_vx
    call check(tbec_set_physics(hh, ph))
    if (ph%potential /= TBEC_POT_TABLE) return
    ! Trap(k,j,i) as raw reals of the working precision, z fastest
    if (len_trim(trap_file) == 0) call emergency_stop('potential = table needs trap_file')
    open(newunit=ut, file=trim(trap_file), access='stream', form='unformatted', &
         status='old', action='read', iostat=ios)
    if (ios /= 0) call emergency_stop('cannot open trap_file '//trim(trap_file))
    if (prec == 64) then
      allocate(tab(Nz, Ny, Nx))
      read(ut, iostat=ios) tab
      if (ios == 0) call check(tbec_set_potential_table(hh, c_loc(tab)))
    else
      allocate(tab4(Nz, Ny, Nx))
      read(ut, iostat=ios) tab4
      if (ios == 0) call check(tbec_set_potential_table(hh, c_loc(tab4)))
    end if
    close(ut)
    if (ios /= 0) call emergency_stop('short trap_file '//trim(trap_file))
  end subroutine

  ! the reference's init_cond names (src/initial.f90:58-156); 'tf_vortex' and 'restart'
  ! are handled on the host
  integer function ref_ic_kind()
    select case (trim(init_cond))
    case ('constant'); ref_ic_kind = 0
    case ('soliton'); ref_ic_kind = 1
    case ('random_phase'); ref_ic_kind = 2
    case ('crow'); ref_ic_kind = 3
    case ('single_vortex'); ref_ic_kind = 4
    case ('sausage'); ref_ic_kind = 5
    case ('random_points'); ref_ic_kind = 6
    case ('leap_frog'); ref_ic_kind = 7
    case ('vortex_rings'); ref_ic_kind = 8
    case ('line_of_lines'); ref_ic_kind = 9
    case ('linked_rings'); ref_ic_kind = 10
    case ('lattice'); ref_ic_kind = 11
    case ('lattice2'); ref_ic_kind = 12
    case ('bundle'); ref_ic_kind = 13
    case default
      ref_ic_kind = -1
      call emergency_stop('wrong initial conditions set')
    end select
  end function

  ! the reference's `potential` names (src/finite_diff.f90:30-117) plus 'table'
  integer(c_int32_t) function pot_kind()
    select case (trim(potential))
    case ('zero'); pot_kind = TBEC_POT_ZERO
    case ('harmonic'); pot_kind = TBEC_POT_HARMONIC
    case ('harmonic_ring'); pot_kind = TBEC_POT_HARMONIC_RING
    case ('harmonic_ring2'); pot_kind = TBEC_POT_HARMONIC_RING2
    case ('barrier'); pot_kind = TBEC_POT_BARRIER
    case ('box_trap'); pot_kind = TBEC_POT_BOX_TRAP
    case ('channel'); pot_kind = TBEC_POT_CHANNEL
    case ('quartic'); pot_kind = TBEC_POT_QUARTIC
    case ('sphere'); pot_kind = TBEC_POT_SPHERE
    case ('mobius'); pot_kind = TBEC_POT_MOBIUS
    case ('table'); pot_kind = TBEC_POT_TABLE
    case default
      pot_kind = -1
      call emergency_stop('wrong potential set')
    end select
  end function

  real(c_double) function never_if_huge(x)
    real(dp), intent(in) :: x
    never_if_huge = x
    if (x >= huge(1.0_dp)) never_if_huge = ieee_value(1.0_c_double, ieee_positive_inf)
  end function

  ! force_random_points (src/force.f90:12-44): a vortex line at a uniform random
  ! position in [-L/4, L/4]^2 with a random sense, drawn here from the seeded
  ! counter-based generator (the reference's generator is clock-seeded).
  subroutine force_random_point()
    real(dp) :: x0, y0, dir
    x0 = -Lx / 4.0_dp + uniform(seed, 7_int64, 3 * nforce) * (Lx / 4.0_dp + Lx / 4.0_dp)
    y0 = -Ly / 4.0_dp + uniform(seed, 7_int64, 3 * nforce + 1) * (Ly / 4.0_dp + Ly / 4.0_dp)
    dir = 1.0_dp
    if (uniform(seed, 7_int64, 3 * nforce + 2) < 0.5_dp) dir = -1.0_dp
    nforce = nforce + 1
    call check(tbec_imprint_vortex(h, x0, y0, dir))
  end subroutine

  ! boundaries / periodicz as in src/cparam.local, applied by ghost_comm's rules
  subroutine set_bc(hh)
    type(c_ptr), intent(in) :: hh
    integer(c_int) :: bc, pz
    select case (trim(boundaries))
    case ('zero')
      bc = TBEC_BC_ZERO
    case ('solid')
      bc = TBEC_BC_SOLID
    case ('periodic')
      bc = TBEC_BC_PERIODIC
    case default
      call emergency_stop('wrong boundaries set')
    end select
    pz = 0
    if (periodicz) pz = 1
    call check(tbec_set_boundaries(hh, bc, pz))
  end subroutine

  ! vel%02d.dat / qpress%02d.dat: one record coords(2), field(nz, ny, nx, 3) in the run's
  ! precision (src/output.f90:80-91), computed on the device from the current psi.
  subroutine write_field(name, shot)
    character(len=*), intent(in) :: name
    integer, intent(in) :: shot
    character(len=512) :: fname
    integer :: uv
    integer(4) :: crd(2)
    crd = 0
    if (prec == 64) then
      if (.not. allocated(fld)) allocate(fld(Nz, Ny, Nx, 3))
      if (name == 'vel') then
        call check(tbec_velocity(h, c_loc(fld)))
      else
        call check(tbec_qpressure(h, c_loc(fld)))
      end if
    else
      if (.not. allocated(fld4)) allocate(fld4(Nz, Ny, Nx, 3))
      if (name == 'vel') then
        call check(tbec_velocity(h, c_loc(fld4)))
      else
        call check(tbec_qpressure(h, c_loc(fld4)))
      end if
    end if
    write(fname, '(a,a,a,i2.2,a)') trim(out_dir), '/proc000/', name, shot, '.dat'
    open(newunit=uv, file=trim(fname), status='replace', form='unformatted')
    if (prec == 64) then
      write(uv) crd, fld
    else
      write(uv) crd, fld4
    end if
    close(uv)
  end subroutine

  ! restart from out_dir/proc000/var.dat (src/initial.f90:796-818)
  subroutine read_restart()
    character(len=512) :: fname
    integer :: uv, it4
    real(dp) :: t8
    real(real32) :: t4
    write(fname, '(a,a)') trim(out_dir), '/proc000/var.dat'
    open(newunit=uv, file=trim(fname), status='old', form='unformatted', iostat=ios)
    if (ios /= 0) call emergency_stop('cannot open restart file '//trim(fname))
    read(uv) it4
    if (prec == 64) then
      read(uv) t8
      read(uv) psi
      t = t8
    else
      allocate(psi4(Nz, Ny, Nx))
      read(uv) t4
      read(uv) psi4
      psi = cmplx(psi4, kind=dp)
      t = t4
    end if
    close(uv)
    nstart = it4
    write(*, '(a,es14.6)') ' succsesful restart, t=', t
  end subroutine
end program
