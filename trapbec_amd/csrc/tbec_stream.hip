// Streaming rates of plain kernels on this GPU, for the stepper's streaming ceiling: a
// copy (one array read, one written: the first RK stage's mix with w rebuilt) and an add
// (two read, one written: the later stages' mix), 16-byte accesses with non-temporal
// writes; the copy reads non-temporally, the add is timed with plain and with
// non-temporal reads and the faster is kept.  A measurement utility; no counterpart in
// the reference.
#include <hip/hip_runtime.h>

#include <string>

#include "tbec_internal.hpp"

using namespace tbec;

namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_stream_copy(const u32x4 *__restrict__ a,
                                                     u32x4 *__restrict__ c, long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    __builtin_nontemporal_store(__builtin_nontemporal_load(&a[i]), &c[i]);
}

// four vectors per thread in flight
template <bool NT>
__global__ void __launch_bounds__(256) k_stream_add(const f32x4 *__restrict__ a,
                                                    const f32x4 *__restrict__ b,
                                                    f32x4 *__restrict__ c, long long n) {
  const long long per = 256 * 4;
  for (long long base = blockIdx.x * per; base < n; base += (long long)gridDim.x * per) {
    f32x4 x[4], y[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long long i = base + q * 256 + threadIdx.x;
      if (i < n) {
        if (NT) {
          x[q] = __builtin_nontemporal_load(&a[i]);
          y[q] = __builtin_nontemporal_load(&b[i]);
        } else {
          x[q] = a[i];
          y[q] = b[i];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long long i = base + q * 256 + threadIdx.x;
      if (i < n) __builtin_nontemporal_store(x[q] + y[q], &c[i]);
    }
  }
}

}  // namespace

extern "C" int tbec_stream_rates(int64_t bytes, int reps, double *copy_us, double *add_us) {
  if (bytes < 16 || (bytes & 15) || reps < 1 || !copy_us || !add_us)
    return set_error(TBEC_ERR_ARG, "bytes must be a positive multiple of 16, reps >= 1");
  const long long n = bytes / 16;
  void *a = nullptr, *b = nullptr, *c = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, bytes);
  if (e == hipSuccess) e = hipMalloc(&b, bytes);
  if (e == hipSuccess) e = hipMalloc(&c, bytes);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, bytes, s);
  if (e == hipSuccess) e = hipMemsetAsync(b, 0, bytes, s);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  double us[3] = {0.0, 0.0, 0.0};
  for (int kind = 0; kind < 3 && e == hipSuccess; ++kind) {
    const int grid = kind == 0 ? 65536 : 16384;
    auto launch = [&]() {
      if (kind == 0)
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, s, (const u32x4 *)a,
                           (u32x4 *)c, n);
      else if (kind == 1)
        hipLaunchKernelGGL(k_stream_add<false>, dim3(grid), dim3(256), 0, s,
                           (const f32x4 *)a, (const f32x4 *)b, (f32x4 *)c, n);
      else
        hipLaunchKernelGGL(k_stream_add<true>, dim3(grid), dim3(256), 0, s,
                           (const f32x4 *)a, (const f32x4 *)b, (f32x4 *)c, n);
    };
    launch();   // warm-up
    e = hipEventRecord(e0, s);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
      launch();
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    us[kind] = 1e3 * ms / reps;
  }
  *copy_us = us[0];
  *add_us = us[1] < us[2] ? us[1] : us[2];
  if (s) (void)hipStreamSynchronize(s);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  for (void *p : {a, b, c})
    if (p) (void)hipFree(p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(TBEC_ERR_HIP, std::string("stream rates: ") + hipGetErrorString(e));
  }
  return 0;
}
