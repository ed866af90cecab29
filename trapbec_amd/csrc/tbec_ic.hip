// Seeded initial states generated on the device, straight into the padded layout.
//
// The same states as trapbec_amd/initial.py and the Fortran driver's tbec_ic.f90 (the
// reference builds its states on the host in src/initial.f90): a Thomas-Fermi profile
// sqrt(max(mu - V, 0)/g), times a vortex ring, times straight z-line vortices, times a
// uniform phase noise, plus complex Gaussian noise, all times a final scale.  Random
// numbers come from a counter-based 32-bit mixer of (seed, stream, flat point index), so
// the device, numpy and Fortran generators agree to the rounding of their libm calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "gpe_stage.hpp"
#include "tbec_internal.hpp"

using namespace tbec;

namespace {

struct IcArgs {
  tbec_ic ic;
  int nx, ny, nz;
  double Lx, Ly, Lz;
  long long sy, sx, base;
};

template <typename T>
__global__ void k_init(const IcArgs A, cplx<T> *__restrict__ psi) {
  const tbec_ic &ic = A.ic;
  const long long n = (long long)A.nx * A.ny * A.nz;
  const double pi2 = 6.283185307179586;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const long long k = t % A.nz, r = t / A.nz, j = r % A.ny, i = r / A.ny;
    const double x = ((double)(2 * i + 1) / (2.0 * A.nx)) * A.Lx - A.Lx / 2.0;
    const double y = ((double)(2 * j + 1) / (2.0 * A.ny)) * A.Ly - A.Ly / 2.0;
    const double z = ((double)(2 * k + 1) / (2.0 * A.nz)) * A.Lz - A.Lz / 2.0;
    const double V = 0.5 * (ic.omega[0] * ic.omega[0] * x * x +
                            ic.omega[1] * ic.omega[1] * y * y +
                            ic.omega[2] * ic.omega[2] * z * z);
    double re = sqrt(fmax(ic.mu - V, 0.0) / ic.g), im = 0.0;
    const double s2xi = 1.4142135623730951 * ic.xi;
    auto mul = [&](double a, double b) {
      const double nr = re * a - im * b, ni = re * b + im * a;
      re = nr;
      im = ni;
    };
    if (ic.ring_radius > 0.0) {
      const double rho = sqrt(x * x + y * y);
      const double a = rho - ic.ring_radius, b = z - ic.ring_z0;
      const double amp = tanh(sqrt(a * a + b * b) / s2xi);
      const double th = atan2(b, a);
      mul(amp * cos(th), amp * sin(th));
    }
    for (int l = 0; l < ic.n_lines && l < 8; ++l) {
      const double dx = x - ic.line_x[l], dy = y - ic.line_y[l];
      const double amp = tanh(sqrt(dx * dx + dy * dy) / s2xi);
      const double th = ic.line_q[l] * atan2(dy, dx);
      mul(amp * cos(th), amp * sin(th));
    }
    const uint32_t idx = (uint32_t)t;
    if (ic.noise_amp > 0.0) {
      const double u1 = uniform01(ic.seed, 0, idx), u2 = uniform01(ic.seed, 1, idx);
      const double rr = sqrt(-2.0 * log(u1));
      re += ic.noise_amp * rr * cos(pi2 * u2);
      im += ic.noise_amp * rr * sin(pi2 * u2);
    }
    if (ic.phase_noise_amp > 0.0) {
      const double th = ic.phase_noise_amp * (2.0 * uniform01(ic.seed, 0, idx) - 1.0);
      mul(cos(th), sin(th));
    }
    cplx<T> v;
    v.re = T(re * ic.scale);
    v.im = T(im * ic.scale);
    psi[A.base + i * A.sx + j * A.sy + k] = v;
  }
}

}  // namespace

extern "C" int tbec_init_state(tbec_handle *h, const tbec_ic *ic) {
  if (!h || !ic) return set_error(TBEC_ERR_ARG, "null argument");
  if (!(ic->g > 0.0)) return set_error(TBEC_ERR_ARG, "initial state needs g > 0");
  if (ic->n_lines < 0 || ic->n_lines > 8) return set_error(TBEC_ERR_ARG, "n_lines in 0..8");
  if ((long long)h->p.nx * h->p.ny * h->p.nz > 0xFFFFFFFFLL)
    return set_error(TBEC_ERR_ARG, "grid too large for the 32-bit point index");
  IcArgs A;
  A.ic = *ic;
  A.nx = h->p.nx; A.ny = h->p.ny; A.nz = h->p.nz;
  A.Lx = h->p.Lx; A.Ly = h->p.Ly; A.Lz = h->p.Lz;
  A.sy = h->sy; A.sx = h->sx; A.base = h->base;
  const long long n = (long long)A.nx * A.ny * A.nz;
  const int grid = (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 16384));
  if (h->p.prec == TBEC_F32)
    hipLaunchKernelGGL(k_init<float>, dim3(grid), dim3(256), 0, h->stream, A,
                       (cplx<float> *)h->psi[h->cur]);
  else
    hipLaunchKernelGGL(k_init<double>, dim3(grid), dim3(256), 0, h->stream, A,
                       (cplx<double> *)h->psi[h->cur]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(TBEC_ERR_HIP, hipGetErrorString(e));
  return tbec_reset_state(h);
}
