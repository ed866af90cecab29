// Prolongation between grids of one box: the current psi of a handle on a coarser (or
// any) grid, trilinearly interpolated at the cell centres of another handle's grid.
// This is the step between the levels of a coarse-to-fine imaginary-time solve: the slow
// modes of a ground-state search (a vortex lattice settling) relax on a time scale set
// by the physics, not by the grid, so they are cheapest to relax where a step is cheap
// and dtau is large.  No counterpart in the reference, which solves on one grid.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "gpe_stage.hpp"
#include "tbec_internal.hpp"

using namespace tbec;

namespace {

// Lower interpolation index and weight of the upper neighbour along one axis: a fine
// cell centre i maps to the coarse coordinate u = (i + 1/2) m/n - 1/2 (cell centres of
// both grids on the same interval), clamped to [0, m-1].
__device__ inline void axis_weight(long long i, int n, int m, int &i0, int &i1, double &f) {
  const double u = ((double)i + 0.5) * (double)m / (double)n - 0.5;
  const double c = fmin(fmax(u, 0.0), (double)(m - 1));
  i0 = m > 1 ? min((int)floor(c), m - 2) : 0;
  i1 = m > 1 ? i0 + 1 : 0;
  f = m > 1 ? c - (double)i0 : 0.0;
}

template <typename T>
__global__ void k_prolong(const cplx<T> *__restrict__ src, int mx, int my, int mz,
                          long long ssy, long long ssx, long long sbase,
                          const double *__restrict__ scale, cplx<T> *__restrict__ dst,
                          int nx, int ny, int nz, long long dsy, long long dsx,
                          long long dbase) {
  const long long n = (long long)nx * ny * nz;
  const T s = T(*scale);
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const long long k = t % nz, r = t / nz, j = r % ny, i = r / ny;
    int x0, x1, y0, y1, z0, z1;
    double fx, fy, fz;
    axis_weight(i, nx, mx, x0, x1, fx);
    axis_weight(j, ny, my, y0, y1, fy);
    axis_weight(k, nz, mz, z0, z1, fz);
    const int xi[2] = {x0, x1}, yi[2] = {y0, y1}, zi[2] = {z0, z1};
    const T wx[2] = {T(1.0 - fx), T(fx)}, wy[2] = {T(1.0 - fy), T(fy)},
            wz[2] = {T(1.0 - fz), T(fz)};
    T re = T(0), im = T(0);
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 2; ++c) {
          const T w = wx[a] * wy[b] * wz[c];
          const cplx<T> v = src[sbase + xi[a] * ssx + yi[b] * ssy + zi[c]];
          re += w * v.re;
          im += w * v.im;
        }
    cplx<T> o;
    o.re = re * s;
    o.im = im * s;
    dst[dbase + i * dsx + j * dsy + k] = o;
  }
}

}  // namespace

hipError_t tbec::launch_prolong(int prec, const void *src, const GridView &s,
                                const double *scale, void *dst, const GridView &d,
                                hipStream_t stream) {
  const long long n = (long long)d.nx * d.ny * d.nz;
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (prec == TBEC_F32)
    hipLaunchKernelGGL(k_prolong<float>, dim3((unsigned)blocks), dim3(256), 0, stream,
                       (const cplx<float> *)src, s.nx, s.ny, s.nz, s.sy, s.sx, s.base, scale,
                       (cplx<float> *)dst, d.nx, d.ny, d.nz, d.sy, d.sx, d.base);
  else
    hipLaunchKernelGGL(k_prolong<double>, dim3((unsigned)blocks), dim3(256), 0, stream,
                       (const cplx<double> *)src, s.nx, s.ny, s.nz, s.sy, s.sx, s.base, scale,
                       (cplx<double> *)dst, d.nx, d.ny, d.nz, d.sy, d.sx, d.base);
  return hipGetLastError();
}

bool tbec::same_box(const tbec_params &a, const tbec_params &b) {
  const double La[3] = {a.Lx, a.Ly, a.Lz}, Lb[3] = {b.Lx, b.Ly, b.Lz};
  for (int d = 0; d < 3; ++d)
    if (!(std::fabs(La[d] - Lb[d]) <= 1e-12 * std::fabs(La[d]))) return false;
  return true;
}

extern "C" int tbec_prolong(tbec_handle *dst, tbec_handle *src) {
  if (!dst || !src) return set_error(TBEC_ERR_ARG, "null handle");
  if (dst == src) return set_error(TBEC_ERR_ARG, "prolongation needs two handles");
  if (dst->p.prec != src->p.prec)
    return set_error(TBEC_ERR_ARG, "prolongation between handles of different precision");
  if (dst->device != src->device)
    return set_error(TBEC_ERR_ARG, "prolongation between handles on different devices");
  if (!same_box(dst->p, src->p))
    return set_error(TBEC_ERR_ARG, "prolongation between different boxes");
  // the source's pending steps finish before its psi is read
  hipError_t e = hipStreamSynchronize(src->stream);
  if (e != hipSuccess) return set_error(TBEC_ERR_HIP, hipGetErrorString(e));
  const GridView sv{src->p.nx, src->p.ny, src->p.nz, src->sy, src->sx, src->base};
  const GridView dv{dst->p.nx, dst->p.ny, dst->p.nz, dst->sy, dst->sx, dst->base};
  e = launch_prolong(dst->p.prec, src->psi[src->cur], sv, src->dstate + ST_SCALE,
                     dst->psi[dst->cur], dv, dst->stream);
  if (e != hipSuccess) return set_error(TBEC_ERR_HIP, hipGetErrorString(e));
  const int rc = tbec_reset_state(dst);
  if (rc) return rc;
  // the source may be destroyed as soon as this returns
  e = hipStreamSynchronize(dst->stream);
  if (e != hipSuccess) return set_error(TBEC_ERR_HIP, hipGetErrorString(e));
  return 0;
}
