// Launchers of the stage kernel, one per (precision, tile shape).  Each is compiled in
// its own translation unit (stage_inst.hip, built once per pair) so the kernel
// instantiations compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "gpe_stage.hpp"

namespace tbec {

// Tile shapes (TZ, TY, VZ) per precision; the id is the index here.
template <int PREC, int ID> struct TileShape;
template <> struct TileShape<32, 0> { static constexpr int TZ = 64, TY = 16, VZ = 4; };
template <> struct TileShape<32, 1> { static constexpr int TZ = 64, TY = 32, VZ = 2; };
template <> struct TileShape<32, 2> { static constexpr int TZ = 64, TY = 16, VZ = 2; };
template <> struct TileShape<32, 3> { static constexpr int TZ = 32, TY = 32, VZ = 4; };
template <> struct TileShape<32, 4> { static constexpr int TZ = 64, TY = 16, VZ = 1; };
template <> struct TileShape<64, 0> { static constexpr int TZ = 32, TY = 16, VZ = 2; };
template <> struct TileShape<64, 1> { static constexpr int TZ = 64, TY = 8, VZ = 2; };
template <> struct TileShape<64, 2> { static constexpr int TZ = 32, TY = 16, VZ = 1; };
template <> struct TileShape<64, 3> { static constexpr int TZ = 64, TY = 16, VZ = 1; };
template <> struct TileShape<64, 4> { static constexpr int TZ = 32, TY = 8, VZ = 1; };

#define TBEC_DECL_LAUNCHER(P, I) \
  hipError_t launch_stage_p##P##_t##I(const StageArgs &a, int sum, int phys, bool imag, \
                                      hipStream_t s);
TBEC_DECL_LAUNCHER(32, 0)
TBEC_DECL_LAUNCHER(32, 1)
TBEC_DECL_LAUNCHER(32, 2)
TBEC_DECL_LAUNCHER(32, 3)
TBEC_DECL_LAUNCHER(32, 4)
TBEC_DECL_LAUNCHER(64, 0)
TBEC_DECL_LAUNCHER(64, 1)
TBEC_DECL_LAUNCHER(64, 2)
TBEC_DECL_LAUNCHER(64, 3)
TBEC_DECL_LAUNCHER(64, 4)
#undef TBEC_DECL_LAUNCHER

#define TBEC_DECL_BLAUNCHER(P, I)                                                      \
  hipError_t launch_batch_p##P##_t##I(const BatchLaunch &L, int sum, bool rot, int grid, \
                                      hipStream_t s);
TBEC_DECL_BLAUNCHER(32, 0)
TBEC_DECL_BLAUNCHER(32, 1)
TBEC_DECL_BLAUNCHER(32, 2)
TBEC_DECL_BLAUNCHER(32, 3)
TBEC_DECL_BLAUNCHER(32, 4)
TBEC_DECL_BLAUNCHER(64, 0)
TBEC_DECL_BLAUNCHER(64, 1)
TBEC_DECL_BLAUNCHER(64, 2)
TBEC_DECL_BLAUNCHER(64, 3)
TBEC_DECL_BLAUNCHER(64, 4)
#undef TBEC_DECL_BLAUNCHER

// Dispatch to the (precision, tile) launcher.
hipError_t launch_batch(int prec, int tile, const BatchLaunch &L, int sum, bool rot, int grid,
                        hipStream_t s);

// Stages in the default summation of each precision (centred FP32, reference FP64) also
// come as instances with the stage kind fixed at compile time: with w rebuilt in real and
// imaginary time and in the batch kernel, and with w stored in real time: the plane loop then carries no stage
// branches and the imaginary-time variants spill less (measured: real time at 512^3
// FP32 -8.5% per step, FP64 rotating -4%; imaginary time at 256^3 rotating FP32 -12%,
// FP64 -18%; a batch of 12 x 96^3 -21% / -16%; stored w at 512^3 FP32 -2%, FP64 0%).  The
// other FP32 sums and imaginary time with stored w use the run-time stage kind.
#ifndef TBEC_STAGE_SPECIALISE
#define TBEC_STAGE_SPECIALISE 1
#endif
#ifndef TBEC_STAGE_SPECIALISE_IMAG
#define TBEC_STAGE_SPECIALISE_IMAG 1
#endif
template <typename T, int S, int PH> constexpr bool stage_specialised() {
  return TBEC_STAGE_SPECIALISE &&
         ((sizeof(T) == 4 && S == SUM_CENTRED) || (sizeof(T) == 8 && S == SUM_REF));
}

template <typename T, int TZ, int TY, int VZ, int S, int PH>
static void launch_phys(const StageArgs &a, bool imag, hipStream_t s) {
  constexpr int NT = (TZ / VZ) * TY;
  dim3 grid(a.nblocks), block(NT);
  if (imag) {
    if constexpr (stage_specialised<T, S, PH>() && TBEC_STAGE_SPECIALISE_IMAG) {
      const bool rebuilt = !a.store_w && (a.stage == ST_RK1 || a.wrec != 0.0);
      if (rebuilt && a.stage == ST_RK1) {
        hipLaunchKernelGGL((gpe_stage<T, TZ, TY, VZ, S, PH, true, ST_RK1>), grid, block, 0, s, a);
        return;
      }
      if (rebuilt && a.stage == ST_RK2) {
        hipLaunchKernelGGL((gpe_stage<T, TZ, TY, VZ, S, PH, true, ST_RK2>), grid, block, 0, s, a);
        return;
      }
      if (rebuilt && a.stage == ST_RK3) {
        hipLaunchKernelGGL((gpe_stage<T, TZ, TY, VZ, S, PH, true, ST_RK3>), grid, block, 0, s, a);
        return;
      }
    }
    hipLaunchKernelGGL((gpe_stage<T, TZ, TY, VZ, S, PH, true>), grid, block, 0, s, a);
    return;
  }
  if constexpr (stage_specialised<T, S, PH>()) {
    const bool rebuilt = !a.store_w && (a.stage == ST_RK1 || a.wrec != 0.0);
    if (rebuilt && a.stage == ST_RK1) {
      hipLaunchKernelGGL((gpe_stage<T, TZ, TY, VZ, S, PH, false, ST_RK1>), grid, block, 0, s, a);
      return;
    }
    if (rebuilt && a.stage == ST_RK2) {
      hipLaunchKernelGGL((gpe_stage<T, TZ, TY, VZ, S, PH, false, ST_RK2>), grid, block, 0, s, a);
      return;
    }
    if (rebuilt && a.stage == ST_RK3) {
      hipLaunchKernelGGL((gpe_stage<T, TZ, TY, VZ, S, PH, false, ST_RK3>), grid, block, 0, s, a);
      return;
    }
    // the stored-w scheme: stages 1 and 2 write w, stages 2 and 3 read it
    if (a.stage == ST_RK1 && a.store_w && a.wrec == 0.0) {
      hipLaunchKernelGGL((gpe_stage<T, TZ, TY, VZ, S, PH, false, ST_RK1 | SK_STORED>), grid,
                         block, 0, s, a);
      return;
    }
    if (a.stage == ST_RK2 && a.store_w && a.wrec == 0.0) {
      hipLaunchKernelGGL((gpe_stage<T, TZ, TY, VZ, S, PH, false, ST_RK2 | SK_STORED>), grid,
                         block, 0, s, a);
      return;
    }
    if (a.stage == ST_RK3 && !a.store_w && a.wrec == 0.0) {
      hipLaunchKernelGGL((gpe_stage<T, TZ, TY, VZ, S, PH, false, ST_RK3 | SK_STORED>), grid,
                         block, 0, s, a);
      return;
    }
  }
  hipLaunchKernelGGL((gpe_stage<T, TZ, TY, VZ, S, PH, false>), grid, block, 0, s, a);
}

template <typename T, int TZ, int TY, int VZ, int S>
static void launch_sum(const StageArgs &a, int phys, bool imag, hipStream_t s) {
  if (phys == PH_GEN) launch_phys<T, TZ, TY, VZ, S, PH_GEN>(a, imag, s);
  else if (phys == PH_ROT) launch_phys<T, TZ, TY, VZ, S, PH_ROT>(a, imag, s);
  else launch_phys<T, TZ, TY, VZ, S, PH_HARM>(a, imag, s);
}

// FP64 uses the reference summation only (it is the parity path); FP32 has all three.
template <typename T, int TZ, int TY, int VZ>
static hipError_t launch_tile(const StageArgs &a, int sum, int phys, bool imag,
                              hipStream_t s) {
  if constexpr (sizeof(T) == 8) {
    launch_sum<T, TZ, TY, VZ, SUM_REF>(a, phys, imag, s);
  } else {
    if (sum == SUM_REF) launch_sum<T, TZ, TY, VZ, SUM_REF>(a, phys, imag, s);
    else if (sum == SUM_COMP) launch_sum<T, TZ, TY, VZ, SUM_COMP>(a, phys, imag, s);
    else launch_sum<T, TZ, TY, VZ, SUM_CENTRED>(a, phys, imag, s);
  }
  return hipGetLastError();
}

// Batched imaginary-time stages (reductions always on).
template <typename T, int TZ, int TY, int VZ, int S, int PH>
static void launch_batch_phys(const BatchLaunch &L, int grid, hipStream_t s) {
  constexpr int NT = (TZ / VZ) * TY;
  // batches always rebuild w: the three RK stage kinds have their own instances in the
  // default summations
  if constexpr (stage_specialised<T, S, PH>() && TBEC_STAGE_SPECIALISE_IMAG) {
    if (L.stage == ST_RK1) {
      hipLaunchKernelGGL((gpe_stage_batch<T, TZ, TY, VZ, S, PH, true, ST_RK1>), dim3(grid),
                         dim3(NT), 0, s, L);
      return;
    }
    if (L.stage == ST_RK2) {
      hipLaunchKernelGGL((gpe_stage_batch<T, TZ, TY, VZ, S, PH, true, ST_RK2>), dim3(grid),
                         dim3(NT), 0, s, L);
      return;
    }
    if (L.stage == ST_RK3) {
      hipLaunchKernelGGL((gpe_stage_batch<T, TZ, TY, VZ, S, PH, true, ST_RK3>), dim3(grid),
                         dim3(NT), 0, s, L);
      return;
    }
  }
  hipLaunchKernelGGL((gpe_stage_batch<T, TZ, TY, VZ, S, PH, true>), dim3(grid), dim3(NT), 0, s,
                     L);
}

template <typename T, int TZ, int TY, int VZ, int S>
static void launch_batch_sum(const BatchLaunch &L, bool rot, int grid, hipStream_t s) {
  if (rot) launch_batch_phys<T, TZ, TY, VZ, S, PH_ROT>(L, grid, s);
  else launch_batch_phys<T, TZ, TY, VZ, S, PH_HARM>(L, grid, s);
}

template <typename T, int TZ, int TY, int VZ>
static hipError_t launch_batch_tile(const BatchLaunch &L, int sum, bool rot, int grid,
                                    hipStream_t s) {
  if constexpr (sizeof(T) == 8) {
    launch_batch_sum<T, TZ, TY, VZ, SUM_REF>(L, rot, grid, s);
  } else {
    if (sum == SUM_REF) launch_batch_sum<T, TZ, TY, VZ, SUM_REF>(L, rot, grid, s);
    else if (sum == SUM_COMP) launch_batch_sum<T, TZ, TY, VZ, SUM_COMP>(L, rot, grid, s);
    else launch_batch_sum<T, TZ, TY, VZ, SUM_CENTRED>(L, rot, grid, s);
  }
  return hipGetLastError();
}

}  // namespace tbec
