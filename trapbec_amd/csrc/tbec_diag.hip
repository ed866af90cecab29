// Output-time diagnostics of the reference, on the device: the velocity field from the
// phase (get_phase, src/diagnostic.f90:27-46) and the quantum-pressure field
// (get_qpressure, src/diagnostic.f90:50-83), both written by var_print as vel%02d.dat
// and qpress%02d.dat (src/output.f90:80-91).
//
// The handle's w array (not carried across steps) is the scratch: its two real halves
// hold the phase or the density, and the Laplacian of the density, over the padded
// grid.  The ghost faces are filled first as ghost_comm fills them (zero, solid or
// periodic; with zero boundaries the ghost phase is atan2(0, 0) = 0 and the ghost density
// 0); corners stay zero, which is the reference's one-rank case.  Sums keep the
// reference's operand order and are not contracted into FMAs.
#include <hip/hip_runtime.h>

#include <string>

#include "gpe_stage.hpp"
#include "tbec_internal.hpp"

using namespace tbec;

namespace {

template <typename T>
__global__ void k_phase(const cplx<T> *__restrict__ psi, long long n, T *__restrict__ out,
                        int density) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const cplx<T> v = psi[q];
    out[q] = density ? hypot(v.re, v.im) : atan2(v.im, v.re);
  }
}

struct Geo {
  int nx, ny, nz;
  long long sy, sx, base;
  double Lx, Ly, Lz;
};

// 6th-order first derivative of the phase along each axis (first_deriv,
// src/deriv.f90:5-11) at every interior point; out is [axis][x][y][z].
template <typename T>
__global__ void k_velocity(const T *__restrict__ ph, const Geo g, T *__restrict__ out) {
#pragma clang fp contract(off)
  const long long n = (long long)g.nx * g.ny * g.nz;
  const T h[3] = {T(g.Lx) / T(g.nx), T(g.Ly) / T(g.ny), T(g.Lz) / T(g.nz)};
  const long long st[3] = {g.sx, g.sy, 1};
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const long long k = t % g.nz, r = t / g.nz, j = r % g.ny, i = r / g.ny;
    const T *c = ph + g.base + i * g.sx + j * g.sy + k;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const long long s = st[a];
      const T v = ((((-c[-3 * s] + T(9) * c[-2 * s]) - T(45) * c[-s]) + T(45) * c[s]) -
                   T(9) * c[2 * s]) + c[3 * s];
      out[a * n + t] = v / (T(60) * h[a]);
    }
  }
}

// 4th-order Laplacian of the density (real_second_deriv, src/deriv.f90:26-32, summed
// x + y + z) on the interior extended by one cell on every side.
template <typename T>
__global__ void k_laplace(const T *__restrict__ dn, const Geo g, T *__restrict__ lp) {
#pragma clang fp contract(off)
  const int ex = g.nx + 2, ey = g.ny + 2, ez = g.nz + 2;
  const long long n = (long long)ex * ey * ez;
  const T h[3] = {T(g.Lx) / T(g.nx), T(g.Ly) / T(g.ny), T(g.Lz) / T(g.nz)};
  const long long st[3] = {g.sx, g.sy, 1};
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const long long k = t % ez - 1, r = t / ez, j = r % ey - 1, i = r / ey - 1;
    const long long o = g.base + i * g.sx + j * g.sy + k;
    const T *c = dn + o;
    T d[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const long long s = st[a];
      const T v = (((-c[-2 * s] + T(16) * c[-s]) - T(30) * c[0]) + T(16) * c[s]) - c[2 * s];
      d[a] = v / (T(12) * (h[a] * h[a]));
    }
    lp[o] = (d[0] + d[1]) + d[2];
  }
}

// 2nd-order first derivative of the Laplacian (low_first_deriv, src/deriv.f90:33-39).
template <typename T>
__global__ void k_qpressure(const T *__restrict__ lp, const Geo g, T *__restrict__ out) {
#pragma clang fp contract(off)
  const long long n = (long long)g.nx * g.ny * g.nz;
  const T h[3] = {T(g.Lx) / T(g.nx), T(g.Ly) / T(g.ny), T(g.Lz) / T(g.nz)};
  const long long st[3] = {g.sx, g.sy, 1};
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const long long k = t % g.nz, r = t / g.nz, j = r % g.ny, i = r / g.ny;
    const T *c = lp + g.base + i * g.sx + j * g.sy + k;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const long long s = st[a];
      out[a * n + t] = (-c[-s] + c[s]) / (T(2) * h[a]);
    }
  }
}

int blocks_for(long long n) {
  return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 16384));
}

template <typename T>
int run_field(tbec_handle *h, bool velocity, void *out_host) {
  const long long n = (long long)h->p.nx * h->p.ny * h->p.nz;
  const long long elems = h->elems;
  Geo g{h->p.nx, h->p.ny, h->p.nz, h->sy, h->sx, h->base, h->p.Lx, h->p.Ly, h->p.Lz};
  const cplx<T> *psi = (const cplx<T> *)h->psi[h->cur];
  T *scr0 = (T *)h->w;            // phase, or the Laplacian
  T *scr1 = scr0 + elems;         // density
  T *dout = nullptr;
  hipError_t e = hipMalloc(&dout, sizeof(T) * 3 * (size_t)n);
  if (e != hipSuccess) return set_error(TBEC_ERR_OOM, "diagnostic output allocation failed");
  if (velocity) {
    hipLaunchKernelGGL(k_phase<T>, dim3(blocks_for(elems)), dim3(256), 0, h->stream, psi,
                       elems, scr0, 0);
    hipLaunchKernelGGL(k_velocity<T>, dim3(blocks_for(n)), dim3(256), 0, h->stream, scr0, g,
                       dout);
  } else {
    hipLaunchKernelGGL(k_phase<T>, dim3(blocks_for(elems)), dim3(256), 0, h->stream, psi,
                       elems, scr1, 1);
    const long long ne = (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
    hipLaunchKernelGGL(k_laplace<T>, dim3(blocks_for(ne)), dim3(256), 0, h->stream, scr1, g,
                       scr0);
    hipLaunchKernelGGL(k_qpressure<T>, dim3(blocks_for(n)), dim3(256), 0, h->stream, scr0, g,
                       dout);
  }
  e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out_host, dout, sizeof(T) * 3 * (size_t)n, hipMemcpyDeviceToHost,
                       h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(dout);
  if (e != hipSuccess) return set_error(TBEC_ERR_HIP, std::string("diagnostic field: ") +
                                                          hipGetErrorString(e));
  return 0;
}

int field(tbec_handle *h, bool velocity, void *out_host) {
  if (!h || !out_host) return set_error(TBEC_ERR_ARG, "null argument");
  int rc = tbec_materialize_scale(h);   // a pending imaginary-time scale
  if (!rc) rc = refresh_ghosts(h);      // ghost_comm before get_phase / get_qpressure
  if (rc) return rc;
  return h->p.prec == TBEC_F32 ? run_field<float>(h, velocity, out_host)
                               : run_field<double>(h, velocity, out_host);
}

}  // namespace

extern "C" int tbec_velocity(tbec_handle *h, void *vel_host) { return field(h, true, vel_host); }

extern "C" int tbec_qpressure(tbec_handle *h, void *qp_host) { return field(h, false, qp_host); }
