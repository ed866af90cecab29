// Deterministic FP64 reductions of per-workgroup partials and the imaginary-time
// bookkeeping that follows them (energy, chemical potential, GP residual, convergence
// test, renormalisation scale).  Shared by the single-condensate and batched paths.
#pragma once
#include <hip/hip_runtime.h>

#include "tbec_internal.hpp"

namespace tbec {

// Fixed-order sum (or max/min for the first two slots) of partials [beg, end) x 4 with
// one 256-thread workgroup; every thread returns the result.
__device__ inline void sum_partials(const double *part, int beg, int end, double out[4],
                                    bool minmax) {
  __shared__ double red[4][256];
  double v[4];
  if (minmax) { v[0] = 0.0; v[1] = INFINITY; v[2] = 0.0; v[3] = 0.0; }
  else { v[0] = v[1] = v[2] = v[3] = 0.0; }
  for (int i = beg + threadIdx.x; i < end; i += 256) {
    if (minmax) {
      v[0] = fmax(v[0], part[4 * i]);
      v[1] = fmin(v[1], part[4 * i + 1]);
    } else {
      v[0] += part[4 * i];
      v[1] += part[4 * i + 1];
    }
    v[2] += part[4 * i + 2];
    v[3] += part[4 * i + 3];
  }
  for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      if (minmax) {
        red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + o]);
        red[1][threadIdx.x] = fmin(red[1][threadIdx.x], red[1][threadIdx.x + o]);
      } else {
        red[0][threadIdx.x] += red[0][threadIdx.x + o];
        red[1][threadIdx.x] += red[1][threadIdx.x + o];
      }
      red[2][threadIdx.x] += red[2][threadIdx.x + o];
      red[3][threadIdx.x] += red[3][threadIdx.x + o];
    }
    __syncthreads();
  }
  for (int q = 0; q < 4; ++q) out[q] = red[q][0];
  __syncthreads();
}

// After a stage-1 (or rhs-only) pass with reductions s = {sum|psi|^2, Re sum psi* rhs,
// sum|psi|^4, sum|rhs|^2} (without dV): energy, mu, residual; with check_conv also the
// convergence test and the c0 of the next step.  Thread 0 only.
//
// In imaginary time (check_conv) E and mu are those of the state scaled exactly to the
// target norm: with r = N_target/N the quadratic part of H scales by r and the
// interaction part by r^2.  The state read in stage 1 is (FP32 rounding of the scale)*psi,
// whose norm is off by ~1e-7; without this the energy sequence carries that rounding and
// its step-to-step change stalls near 1e-8 instead of reaching tolerances like 1e-10.
__device__ inline void fin_energy(const double s[4], double *st, int *flag, int check_conv) {
  const double dV = st[ST_DV];
  const double N = s[0] * dV, S1 = s[1] * dV, S4 = s[2] * dV, R = s[3] * dV;
  const double c0 = check_conv ? st[ST_C0EFF] : st[ST_C0BASE];
  const double H = -0.5 * S1 + 0.5 * c0 * N;
  double E = H - 0.25 * st[ST_GNL] * S4;
  double mu = H / N;
  const double am = 0.5 * c0 - mu;
  double r2 = 0.25 * R - am * S1 + am * am * N;
  if (r2 < 0) r2 = 0;
  st[ST_N] = N;
  st[ST_E] = E;
  st[ST_MU] = mu;
  st[ST_RES] = sqrt(r2 / N);
  if (!check_conv) return;
  if (!isfinite(E) || !isfinite(N) || N <= 0.0) {
    flag[0] = 2;
    flag[1] = (int)st[ST_STEP];
    return;
  }
  {
    const double r = st[ST_NORMT] / N, g4 = 0.25 * st[ST_GNL] * S4;
    const double Hlin = H - 2.0 * g4;              // the part quadratic in psi
    E = r * Hlin + r * r * g4;
    mu = (r * Hlin + r * r * 2.0 * g4) / st[ST_NORMT];
    st[ST_E] = E;
    st[ST_MU] = mu;
  }
  const double rel = st[ST_STEP] > 0.0 ? fabs(E - st[ST_EPREV]) / fabs(E) : INFINITY;
  st[ST_REL] = rel;
  st[ST_EPREV] = E;
  // the kernels read c0 in their own precision: store it rounded so that the energy
  // above (which adds c0*N/2 back) uses exactly the c0 the rhs was computed with
  const double c0n = st[ST_C0BASE] + 2.0 * mu;
  st[ST_C0NEXT] = st[ST_F32] != 0.0 ? (double)(float)c0n : c0n;
  // a step counts toward convergence when the energy change is below the tolerance and,
  // with a residual gate, the GP residual of the step's input is at or below the gate:
  // a slowly drifting state far from stationary (a vortex-free saddle of a rotating
  // trap) then cannot meet the rule
  const bool res_ok = st[ST_RGATE] <= 0.0 || st[ST_RES] <= st[ST_RGATE];
  st[ST_HITS] = (rel < st[ST_TOL] && res_ok) ? st[ST_HITS] + 1.0 : 0.0;
  if (st[ST_HITS] >= (double)kConvergeHits) {
    flag[0] = 1;
    flag[1] = (int)st[ST_STEP];
  }
}

// After stage 3 with reductions s[0] = sum |psi_new|^2: scale for the next read.
__device__ inline void fin_norm(const double s[4], double *st, int *flag) {
  const double N = s[0] * st[ST_DV];
  if (!isfinite(N) || N <= 0.0) {
    flag[0] = 2;
    flag[1] = (int)st[ST_STEP];
    return;
  }
  st[ST_SCALE] = sqrt(st[ST_NORMT] / N);
  st[ST_C0EFF] = st[ST_C0NEXT];
  st[ST_STEP] += 1.0;
}

}  // namespace tbec
