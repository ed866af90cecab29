// Internal state of a libtbec handle (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <string>
#include <tuple>

#include "tbec.h"

namespace tbec {

// Device state slots (doubles) of a handle.
enum StateSlot {
  ST_SCALE = 0,   // psi is stored unnormalised; stage 1 reads scale*psi
  ST_C0EFF,       // c0 used by the imaginary-time stages of the current step
  ST_C0NEXT,      // c0 + 2*mu of the current step, for the next one
  ST_C0BASE,      // c0 of the parameters
  ST_GNL,
  ST_DV,
  ST_NORMT,       // renormalisation target
  ST_TOL,         // relative energy change that stops the solve
  ST_STEP,        // completed imaginary-time steps
  ST_EPREV,
  ST_N, ST_E, ST_MU, ST_RES, ST_REL,
  ST_MAXA, ST_MINA, ST_SUMA, ST_SUMA2,
  ST_HITS,        // consecutive steps below the tolerance
  ST_DC0,         // diagnostics: 2*mu, the c0 of the residual pass
  ST_F32,         // 1 when the kernels compute in FP32 (c0 values are stored rounded)
  ST_RGATE,       // a step counts toward convergence only with GP residual <= this (0: any)
  ST_COUNT
};

struct TileCfg { int tz, ty, vz; };
struct StageArgs;

// Padded device layout of one condensate (complex elements).
struct Layout {
  long long PZ = 0, PY = 0, PX = 0, sy = 0, sx = 0, base = 0, elems = 0;
  double hx = 0, hy = 0, hz = 0;
  int csize = 8;
};
void compute_layout(const tbec_params &p, Layout &L);
int fill_static_args(const tbec_params &p, const Layout &L, int tile, int seg, StageArgs &a);
bool rot_enabled(const tbec_params &p);
long long round_up(long long v, long long m);
const TileCfg &tile_cfg(int prec, int id);
extern const int kNumTiles;
long long tiles_per_plane(const tbec_params &p, int tile);
int auto_seg(long long tiles, int nx, long long want);
constexpr long long kAutoSegBlocks = 1024;

// Counter-based random numbers: a function of (seed, stream, index) only, identical on
// the host, the device and in trapbec_amd/initial.py (uniform in (0, 1)).
__host__ __device__ inline uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x21F0AAADu;
  x ^= x >> 15;
  x *= 0x735A2D97u;
  x ^= x >> 15;
  return x;
}

__host__ __device__ inline double uniform01(uint32_t seed, uint32_t stream, uint32_t idx) {
  const uint32_t key = hash32(seed ^ hash32(stream));
  return ((double)hash32(key ^ hash32(idx)) + 0.5) / 4294967296.0;
}

// A solve has converged after this many consecutive steps with a relative energy
// change below the tolerance (one step can dip below it by rounding noise in FP32).
constexpr int kConvergeHits = 8;

// Progress lines on stderr for long solves when TBEC_PROGRESS is set (seconds between
// lines, default 10).
double progress_interval();

}  // namespace tbec

struct tbec_handle {
  tbec_params p{};
  int device = 0;                       // HIP device current at tbec_create
  int tile = 0, seg = 64;
  int tile1 = -1, seg1 = 0;             // first RK stage in real time (tile1 < 0: as tile)
  int remap = 1;                        // XCD-aware tile order
  int w_stored = 0;                     // 1: RK3 stores w (10 transfers), 0: rebuilt (8)
  int bc = 0, periodicz = 0;            // boundary faces (TBEC_BC_*), z periodic
  bool resid_pass = false;              // next ST_RHS reduction runs with c0 = ST_DC0
  int csize = 8;
  double hx = 0, hy = 0, hz = 0;
  double dt_re = 0, dt_im = 0;          // dt*(1 - i*gamma)
  long long PZ = 0, PY = 0, PX = 0, sy = 0, sx = 0, base = 0, elems = 0;
  void *psi[2] = {nullptr, nullptr};
  void *w = nullptr;
  void *stage = nullptr;                // interior-sized staging buffer (lazy)
  int cur = 0;                          // psi[cur] holds the current state
  double *partials = nullptr;
  int npart = 0;
  double *dstate = nullptr, *hstate = nullptr;
  int *dflag = nullptr, *hflag = nullptr;
  double norm_target = 1.0;
  double tol = 0.0;
  double rgate = 0.0;                   // GP-residual gate of the stop rule (0: none)
  bool scale_pending = false;
  hipStream_t stream = nullptr;
  // captured step graphs: (mode, steps, parity) -> executable graph
  std::map<std::tuple<int, long long, int>, hipGraphExec_t> graphs;
  // physics beyond the harmonic trap (tbec_set_physics)
  tbec_physics phys{TBEC_POT_HARMONIC, 0, 0, 0, 4.0, HUGE_VAL, HUGE_VAL, 0.0};
  double t_now = 0.0;                   // the handle's time t (real-time steps advance it)
  void *vtab = nullptr;                 // tabulated potential: padded layout, one real per point
  void *trap = nullptr;                 // Trap table (mobius or the caller's): interior layout
  void *coords = nullptr;               // cell-centre coordinates x, y, z (working precision)
  int trap_owner = 0;                   // kind whose Trap the table holds (MOBIUS / TABLE), 0 none
  bool vtab_valid = false;
  double vkey[5] = {0, 0, 0, 0, 0};     // time key of the table held in vtab
  // stage configuration of the current step (set by tbec::prepare_physics)
  int ph_mode = 0;                      // PH_HARM / PH_ROT / PH_GEN
  int ph_drift = 0;                     // PH_GEN: 0 none, 1 frame term, 2 advective
  bool ph_wl_zero = false;              // analytic modes: the potential is zero
  int ph_sig = -1;                      // signature of the three above (graphs are dropped
                                        // when it changes)
};

namespace tbec {
// part: the reduction partials of the stage (4 doubles per workgroup); null = the
// handle's own buffer, sized (allocated) on demand
int enqueue_stage(tbec_handle *h, int kind, bool imag, bool reduce, double *part = nullptr);
// tbec_diagnose with its reduction partials in `part` (null = the handle's own buffer,
// allocated on demand); `part_slots` is the number of 4-double slots `part` holds
int diagnose_with(tbec_handle *h, tbec_diag *out, double *part, long long part_slots);
// 4-double partial slots tbec_diagnose needs for the handle's grid, tile and segment
long long diagnose_partial_slots(tbec_handle *h);
// preparing the physics of a step at time t would allocate a device table
bool physics_would_allocate(const tbec_handle *h, double t);
int tbec_reserve_partials(tbec_handle *h, int nb);
// Interior extent and padded strides of one grid in device memory.
struct GridView {
  int nx, ny, nz;
  long long sy, sx, base;
};
// Trilinear interpolation of src (scaled by *scale) at dst's cell centres, same box
// (tbec_prolong); enqueued on `stream`.
hipError_t launch_prolong(int prec, const void *src, const GridView &s, const double *scale,
                          void *dst, const GridView &d, hipStream_t stream);
bool same_box(const tbec_params &a, const tbec_params &b);
bool handle_rot(const tbec_handle *h);
int set_error(int code, const std::string &msg);
}  // namespace tbec

int tbec_reset_state(tbec_handle *h);
int tbec_materialize_scale(tbec_handle *h);
namespace tbec {
int refresh_ghosts(tbec_handle *h);   // ghost faces of the current psi (non-zero boundaries)
}  // namespace tbec
void tbec_drop_graphs(tbec_handle *h);
int fill_args_count(tbec_handle *h);
namespace tbec {
// Stage configuration and, for tabulated kinds, the potential table of a step that
// starts at time t (enqueued on the handle's stream when the table changes).
int prepare_physics(tbec_handle *h, double t);
// Physics of the handle changes with t (switch-off times, time-dependent kinds).
bool physics_time_dependent(const tbec_handle *h);
// The step starting at t2 has the same stage configuration and table as one starting at t1.
bool same_step_physics(const tbec_handle *h, double t1, double t2);
// t + Re(dt) in the working precision (src/run.f90:55).
double advance_time(const tbec_handle *h, double t);
void free_physics(tbec_handle *h);
}  // namespace tbec
