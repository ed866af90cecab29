// One (precision, tile) group of stage-kernel instantiations; compiled once per pair
// with -DINST_PREC=32|64 -DINST_TILE=0..4.
#include "stage_launch.hpp"

#ifndef INST_PREC
#error "INST_PREC must be defined"
#endif
#ifndef INST_TILE
#error "INST_TILE must be defined"
#endif

#define TBEC_CAT4(a, b, c, d) a##b##c##d
#define TBEC_NAME(P, I) TBEC_CAT4(launch_stage_p, P, _t, I)
#define TBEC_BNAME(P, I) TBEC_CAT4(launch_batch_p, P, _t, I)

namespace tbec {
using InstT = std::conditional_t<INST_PREC == 32, float, double>;
using InstShape = TileShape<INST_PREC, INST_TILE>;

hipError_t TBEC_NAME(INST_PREC, INST_TILE)(const StageArgs &a, int sum, int phys, bool imag,
                                           hipStream_t s) {
  return launch_tile<InstT, InstShape::TZ, InstShape::TY, InstShape::VZ>(a, sum, phys, imag, s);
}

hipError_t TBEC_BNAME(INST_PREC, INST_TILE)(const BatchLaunch &L, int sum, bool rot, int grid,
                                            hipStream_t s) {
  return launch_batch_tile<InstT, InstShape::TZ, InstShape::TY, InstShape::VZ>(L, sum, rot,
                                                                               grid, s);
}
}  // namespace tbec
