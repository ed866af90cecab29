! Seeded initial conditions of the host driver (FP64).
!
! The reference sets up its states in src/initial.f90 (Thomas-Fermi setup_thomas_fermi
! :637-652, vortex imprint single_vortex :109-121, add_noise :741-757) with a
! clock-seeded RNG (src/mpi_var.f90:64-90).  Here every random number is a function of
! (seed, stream, flat point index) through a 32-bit integer mixer evaluated in
! integer(8) -- the same generator as trapbec_amd/initial.py, so the driver and the
! Python host build identical states.  Arrays are psi(k, j, i) = (z, y, x), z fastest.
module tbec_ic
  use iso_fortran_env, only: real64, int64
  implicit none
  integer, parameter :: dp = real64
  integer(int64), parameter :: M32 = 4294967295_int64
  real(dp), parameter :: PI_D = 3.141592653589793_dp

contains

  pure elemental function hash32(x0) result(x)
    integer(int64), intent(in) :: x0
    integer(int64) :: x
    x = iand(x0, M32)
    x = ieor(x, shiftr(x, 16))
    x = iand(x * 569420461_int64, M32)      ! 0x21f0aaad
    x = ieor(x, shiftr(x, 15))
    x = iand(x * 1935289751_int64, M32)     ! 0x735a2d97
    x = ieor(x, shiftr(x, 15))
  end function

  pure elemental real(dp) function uniform(seed, stream, idx)
    integer(int64), intent(in) :: seed, stream, idx
    integer(int64) :: key
    key = hash32(ieor(iand(seed, M32), hash32(stream)))
    uniform = (real(hash32(ieor(key, hash32(idx))), dp) + 0.5_dp) / 4294967296.0_dp
  end function

  ! x_i = ((2i - 1)/(2n)) L - L/2, i = 1..n  (src/initial.f90:34-42, one domain)
  subroutine coords(n, L, x)
    integer, intent(in) :: n
    real(dp), intent(in) :: L
    real(dp), intent(out) :: x(n)
    integer :: i
    do i = 1, n
      x(i) = (real(2 * i - 1, dp) / (2.0_dp * n)) * L - L / 2.0_dp
    end do
  end subroutine

  ! sqrt(max(mu - V, 0)/g) with V = (wx^2 x^2 + wy^2 y^2 + wz^2 z^2)/2
  subroutine tf_profile(psi, xx, yy, zz, om, g, mu)
    complex(dp), intent(out) :: psi(:, :, :)
    real(dp), intent(in) :: xx(:), yy(:), zz(:), om(3), g, mu
    integer :: i, j, k
    real(dp) :: v
    do i = 1, size(xx)
      do j = 1, size(yy)
        do k = 1, size(zz)
          v = 0.5_dp * (om(1) * om(1) * xx(i) * xx(i) + om(2) * om(2) * yy(j) * yy(j) + &
                        om(3) * om(3) * zz(k) * zz(k))
          psi(k, j, i) = cmplx(sqrt(max(mu - v, 0.0_dp) / g), 0.0_dp, dp)
        end do
      end do
    end do
  end subroutine

  ! mu such that sum |psi_TF|^2 dV = norm (bisection)
  real(dp) function tf_mu_for_norm(xx, yy, zz, om, g, dv, norm) result(mu)
    real(dp), intent(in) :: xx(:), yy(:), zz(:), om(3), g, dv, norm
    real(dp) :: lo, hi, mid
    integer :: it
    lo = 0.0_dp
    hi = 1.0_dp
    do while (n_of(hi) < norm)
      hi = 2.0_dp * hi
    end do
    do it = 1, 200
      mid = 0.5_dp * (lo + hi)
      if (n_of(mid) < norm) then
        lo = mid
      else
        hi = mid
      end if
      if (hi - lo <= 1.0e-15_dp * hi) exit
    end do
    mu = 0.5_dp * (lo + hi)
  contains
    real(dp) function n_of(m)
      real(dp), intent(in) :: m
      integer :: i, j, k
      real(dp) :: v
      n_of = 0.0_dp
      do i = 1, size(xx)
        do j = 1, size(yy)
          do k = 1, size(zz)
            v = 0.5_dp * (om(1) * om(1) * xx(i) * xx(i) + om(2) * om(2) * yy(j) * yy(j) + &
                          om(3) * om(3) * zz(k) * zz(k))
            n_of = n_of + max(m - v, 0.0_dp)
          end do
        end do
      end do
      n_of = n_of * dv / g
    end function
  end function

  ! multiply by tanh(r/(sqrt(2) xi)) exp(i q atan2(y-y0, x-x0)): a straight z-line
  subroutine imprint_line(psi, xx, yy, x0, y0, q, xi)
    complex(dp), intent(inout) :: psi(:, :, :)
    real(dp), intent(in) :: xx(:), yy(:), x0, y0, xi
    integer, intent(in) :: q
    integer :: i, j
    real(dp) :: dx, dy, r, th
    complex(dp) :: f
    do i = 1, size(xx)
      do j = 1, size(yy)
        dx = xx(i) - x0
        dy = yy(j) - y0
        r = sqrt(dx * dx + dy * dy)
        th = atan2(dy, dx)
        f = tanh(r / (sqrt(2.0_dp) * xi)) * exp(cmplx(0.0_dp, q * th, dp))
        psi(:, j, i) = psi(:, j, i) * f
      end do
    end do
  end subroutine

  ! a vortex ring of radius R in the plane z = z0 about the z axis
  subroutine imprint_ring(psi, xx, yy, zz, radius, z0, xi)
    complex(dp), intent(inout) :: psi(:, :, :)
    real(dp), intent(in) :: xx(:), yy(:), zz(:), radius, z0, xi
    integer :: i, j, k
    real(dp) :: rho, a, b, d
    do i = 1, size(xx)
      do j = 1, size(yy)
        rho = sqrt(xx(i) * xx(i) + yy(j) * yy(j))
        do k = 1, size(zz)
          a = rho - radius
          b = zz(k) - z0
          d = sqrt(a * a + b * b)
          psi(k, j, i) = psi(k, j, i) * tanh(d / (sqrt(2.0_dp) * xi)) * &
                         exp(cmplx(0.0_dp, atan2(b, a), dp))
        end do
      end do
    end do
  end subroutine

  ! psi += amp*(g1 + i g2), Box-Muller normals from streams (2s, 2s+1)
  subroutine add_noise(psi, seed, amp, stream)
    complex(dp), intent(inout) :: psi(:, :, :)
    integer(int64), intent(in) :: seed, stream
    real(dp), intent(in) :: amp
    integer :: i, j, k, nz, ny
    integer(int64) :: idx
    real(dp) :: u1, u2, r
    nz = size(psi, 1)
    ny = size(psi, 2)
    do i = 1, size(psi, 3)
      do j = 1, ny
        do k = 1, nz
          idx = (int(i - 1, int64) * ny + (j - 1)) * nz + (k - 1)
          u1 = uniform(seed, 2 * stream, idx)
          u2 = uniform(seed, 2 * stream + 1, idx)
          r = sqrt(-2.0_dp * log(u1))
          psi(k, j, i) = psi(k, j, i) + amp * cmplx(r * cos(2.0_dp * PI_D * u2), &
                                                    r * sin(2.0_dp * PI_D * u2), dp)
        end do
      end do
    end do
  end subroutine

  ! psi *= exp(i theta), theta uniform in [-amp, amp]
  subroutine add_phase_noise(psi, seed, amp, stream)
    complex(dp), intent(inout) :: psi(:, :, :)
    integer(int64), intent(in) :: seed, stream
    real(dp), intent(in) :: amp
    integer :: i, j, k, nz, ny
    integer(int64) :: idx
    real(dp) :: th
    nz = size(psi, 1)
    ny = size(psi, 2)
    do i = 1, size(psi, 3)
      do j = 1, ny
        do k = 1, nz
          idx = (int(i - 1, int64) * ny + (j - 1)) * nz + (k - 1)
          th = amp * (2.0_dp * uniform(seed, stream, idx) - 1.0_dp)
          psi(k, j, i) = psi(k, j, i) * exp(cmplx(0.0_dp, th, dp))
        end do
      end do
    end do
  end subroutine
end module
