// Vortex-core detection by phase winding around grid plaquettes (FP64 phases).
// The reference detects vortices only in MATLAB post-processing (bin/vortex_detect.m,
// bin/gpeget2dvort_homg.m); here the winding of psi's phase around every plaquette of
// one orientation is computed on the device from the current state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <string>
#include <vector>

#include "gpe_stage.hpp"
#include "tbec_internal.hpp"

using namespace tbec;

namespace {

template <typename T>
__global__ void k_vortex(const cplx<T> *__restrict__ p, int nx, int ny, int nz, long long sy,
                         long long sx, long long base, const double *scale, double dens_min,
                         int orient, int *out, long long cap, unsigned long long *count) {
  // plaquette domain and the two in-plane axis offsets (first, second)
  const int ei = orient == 1 ? 0 : 1, ej = orient == 2 ? 0 : 1, ek = orient == 0 ? 0 : 1;
  const long long da = orient == 0 ? sx : (orient == 1 ? sy : 1);
  const long long db = orient == 0 ? sy : (orient == 1 ? 1 : sx);
  const long long mi = nx - ei, mj = ny - ej, mk = nz - ek;
  const long long n = mi * mj * mk;
  const double s = *scale;
  const double two_pi = 6.283185307179586;
  const double pi = 3.141592653589793;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const long long k = t % mk, r = t / mk, j = r % mj, i = r / mj;
    const long long o0 = base + i * sx + j * sy + k;
    const long long o[4] = {o0, o0 + da, o0 + da + db, o0 + db};
    double ph[4];
    bool ok = true;
    for (int q = 0; q < 4; ++q) {
      const cplx<T> v = p[o[q]];
      const double a = (double)v.re * s, b = (double)v.im * s;
      ok = ok && (a * a + b * b >= dens_min);
      ph[q] = atan2(b, a);
    }
    if (!ok) continue;
    double w = 0.0;
    for (int q = 0; q < 4; ++q) {
      double d = ph[(q + 1) & 3] - ph[q];
      if (d > pi) d -= two_pi;
      if (d <= -pi) d += two_pi;
      w += d;
    }
    const int nw = (int)llround(w / two_pi);
    if (nw != 0) {
      const unsigned long long idx = atomicAdd(count, 1ULL);
      if ((long long)idx < cap) {
        out[4 * idx] = (int)i;
        out[4 * idx + 1] = (int)j;
        out[4 * idx + 2] = (int)k;
        out[4 * idx + 3] = nw;
      }
    }
  }
}

}  // namespace

extern "C" int tbec_vortex_cores(tbec_handle *h, double dens_min, int orientation,
                                 int32_t *out, int64_t cap, int64_t *count) {
  if (!h || !count || (cap > 0 && !out) || orientation < 0 || orientation > 2)
    return set_error(TBEC_ERR_ARG, "bad vortex-detection argument");
  int *d_out = nullptr;
  unsigned long long *d_cnt = nullptr;
  const long long dcap = std::max<long long>(cap, 1);
  auto fail = [&](const char *what, hipError_t e) {
    if (d_out) (void)hipFree(d_out);
    if (d_cnt) (void)hipFree(d_cnt);
    return set_error(TBEC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  hipError_t e = hipMalloc(&d_out, sizeof(int) * 4 * dcap);
  if (e != hipSuccess) return fail("hipMalloc", e);
  e = hipMalloc(&d_cnt, sizeof(unsigned long long));
  if (e != hipSuccess) return fail("hipMalloc", e);
  e = hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long), h->stream);
  if (e != hipSuccess) return fail("hipMemsetAsync", e);
  const long long n = (long long)h->p.nx * h->p.ny * h->p.nz;
  const int grid = (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 8192));
  if (h->p.prec == TBEC_F32)
    hipLaunchKernelGGL(k_vortex<float>, dim3(grid), dim3(256), 0, h->stream,
                       (const cplx<float> *)h->psi[h->cur], h->p.nx, h->p.ny, h->p.nz, h->sy,
                       h->sx, h->base, (const double *)(h->dstate + ST_SCALE), dens_min,
                       orientation, d_out, dcap, d_cnt);
  else
    hipLaunchKernelGGL(k_vortex<double>, dim3(grid), dim3(256), 0, h->stream,
                       (const cplx<double> *)h->psi[h->cur], h->p.nx, h->p.ny, h->p.nz, h->sy,
                       h->sx, h->base, (const double *)(h->dstate + ST_SCALE), dens_min,
                       orientation, d_out, dcap, d_cnt);
  e = hipGetLastError();
  if (e != hipSuccess) return fail("vortex kernel launch", e);
  unsigned long long cnt = 0;
  e = hipMemcpyAsync(&cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return fail("count copy", e);
  const long long nget = std::min<long long>((long long)cnt, cap);
  if (nget > 0) {
    std::vector<std::array<int, 4>> v(nget);
    e = hipMemcpy(v.data(), d_out, sizeof(int) * 4 * nget, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail("result copy", e);
    // atomics give an arbitrary order: sort to the (i, j, k) order of a serial scan
    std::sort(v.begin(), v.end());
    for (long long q = 0; q < nget; ++q)
      for (int c = 0; c < 4; ++c) out[4 * q + c] = v[q][c];
  }
  (void)hipFree(d_out);
  (void)hipFree(d_cnt);
  *count = (int64_t)cnt;
  return 0;
}
