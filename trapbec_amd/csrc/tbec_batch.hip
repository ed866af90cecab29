// Batched ensembles: many independent condensates of uneven grids and parameters,
// advanced in imaginary time together.  Every RK stage of every instance is one launch
// of gpe_stage_batch with one workgroup per (instance, tile) work item; the per-instance
// finalisations (energy, convergence, renormalisation) are one
// launch with a workgroup per instance.  Runs of steps are captured into a graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "gpe_stage.hpp"
#include "stage_launch.hpp"
#include "tbec_fin.hpp"
#include "tbec_internal.hpp"

using namespace tbec;

#define BCHK(x)                                                                         \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess)                                                               \
      return set_error(TBEC_ERR_HIP, std::string(#x " failed: ") + hipGetErrorString(e_)); \
  } while (0)

struct tbec_batch {
  int n = 0;
  int prec = 64, sum = 0, tile = 0, seg = 64;
  bool rot = false;
  std::vector<tbec_params> p;
  std::vector<Layout> L;
  std::vector<void *> bufs;          // 2 per instance: the psi ping-pong pair
  std::vector<int> item_begin;       // n + 1
  std::vector<int> cur;              // per-instance current buffer after a solve
  int parity = 0;
  int total_items = 0;
  int active_items = 0;              // entries of d_items: the unfinished instances' items
  int active_done = 0;               // finished instances when d_items was last built
  int tail_seg_on = 0;               // shorten the segment for a tail of few instances
  double rgate = 0.0;                // GP-residual gate of the stop rule (0: none)
  BatchInst *d_inst = nullptr;
  int *d_item_inst = nullptr;
  int *d_item_begin = nullptr;
  int *d_items = nullptr;
  double *d_state = nullptr, *h_state = nullptr;
  int *d_flags = nullptr, *h_flags = nullptr;
  double *d_part = nullptr;
  void *stage = nullptr;
  size_t stage_bytes = 0;
  hipStream_t stream = nullptr;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  long long graph_steps = 0;
};

namespace tbec {
hipError_t launch_batch(int prec, int tile, const BatchLaunch &L, int sum, bool rot, int grid,
                        hipStream_t s) {
  if (prec == TBEC_F32) {
    switch (tile) {
      case 0: return launch_batch_p32_t0(L, sum, rot, grid, s);
      case 1: return launch_batch_p32_t1(L, sum, rot, grid, s);
      case 2: return launch_batch_p32_t2(L, sum, rot, grid, s);
      case 3: return launch_batch_p32_t3(L, sum, rot, grid, s);
      default: return launch_batch_p32_t4(L, sum, rot, grid, s);
    }
  }
  switch (tile) {
    case 0: return launch_batch_p64_t0(L, sum, rot, grid, s);
    case 1: return launch_batch_p64_t1(L, sum, rot, grid, s);
    case 2: return launch_batch_p64_t2(L, sum, rot, grid, s);
    case 3: return launch_batch_p64_t3(L, sum, rot, grid, s);
    default: return launch_batch_p64_t4(L, sum, rot, grid, s);
  }
}
}  // namespace tbec

namespace {

__global__ void __launch_bounds__(256)
k_batch_fin_energy(const double *part, const int *item_begin, double *state, int *flags) {
  const int i = blockIdx.x;
  if (flags[4 * i]) return;
  double s[4];
  sum_partials(part, item_begin[i], item_begin[i + 1], s, false);
  if (threadIdx.x == 0) fin_energy(s, state + (long long)i * ST_COUNT, flags + 4 * i, 1);
}

__global__ void __launch_bounds__(256)
k_batch_fin_norm(const double *part, const int *item_begin, double *state, int *flags) {
  const int i = blockIdx.x;
  if (flags[4 * i]) return;
  double s[4];
  sum_partials(part, item_begin[i], item_begin[i + 1], s, false);
  if (threadIdx.x == 0) fin_norm(s, state + (long long)i * ST_COUNT, flags + 4 * i);
}

template <typename T>
__global__ void k_batch_unpack(const cplx<T> *__restrict__ src, cplx<T> *__restrict__ dst,
                               int nx, int ny, int nz, long long sy, long long sx,
                               long long base) {
  const long long n = (long long)nx * ny * nz;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long z = i % nz, r = i / nz, y = r % ny, x = r / ny;
    dst[base + x * sx + y * sy + z] = src[i];
  }
}

template <typename T>
__global__ void k_batch_pack(const cplx<T> *__restrict__ src, cplx<T> *__restrict__ dst,
                             int nx, int ny, int nz, long long sy, long long sx, long long base,
                             const double *scale) {
  const long long n = (long long)nx * ny * nz;
  const T s = T(*scale);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long z = i % nz, r = i / nz, y = r % ny, x = r / ny;
    cplx<T> v = src[base + x * sx + y * sy + z];
    v.re *= s;
    v.im *= s;
    dst[i] = v;
  }
}

int grid_blocks(long long n) {
  long long b = (n + 255) / 256;
  return (int)std::max<long long>(1, std::min<long long>(b, 8192));
}

// Automatic segment length for the whole ensemble: the items of all instances together
// should number a few per CU, with the single-handle rule (auto_seg) applied to their sum:
// the longest segment that gives kAutoSegBlocks items, else 8, or 4 where 8 gives fewer
// than half of that.
int batch_auto_seg(const tbec_batch *b, int tile) {
  auto items = [&](int seg) {
    long long n = 0;
    for (int i = 0; i < b->n; ++i)
      n += tiles_per_plane(b->p[i], tile) * ((b->p[i].nx + seg - 1) / seg);
    return n;
  };
  for (int seg = 256; seg > 8; seg >>= 1)
    if (items(seg) >= kAutoSegBlocks) return seg;
  return 2 * items(8) < kAutoSegBlocks ? 4 : 8;
}

// Segment for the unfinished instances alone, once their items number under one per CU
// (kAutoSegBlocks / 4): the automatic rule applied to them, if shorter than the current
// one; otherwise the current segment.  A long tail of a few slow instances (C4: one 96^3
// instance alone for the last ~70% of its 1.5 M fine steps) then gets more workgroups.
int tail_seg(const tbec_batch *b, const int *flags) {
  auto items = [&](int seg) {
    long long n = 0;
    for (int i = 0; i < b->n; ++i)
      if (flags[4 * i] == 0)
        n += tiles_per_plane(b->p[i], b->tile) * ((b->p[i].nx + seg - 1) / seg);
    return n;
  };
  if (items(b->seg) * 4 >= kAutoSegBlocks) return b->seg;
  int seg = 2 * items(8) < kAutoSegBlocks ? 4 : 8;
  for (int s = 256; s > 8; s >>= 1)
    if (items(s) >= kAutoSegBlocks) { seg = s; break; }
  return std::min(seg, b->seg);
}

// Build the device instance table for the current tile / segment choice.
int build_table(tbec_batch *b) {
  std::vector<BatchInst> inst(b->n);
  b->item_begin.assign(b->n + 1, 0);
  for (int i = 0; i < b->n; ++i) {
    BatchInst &bi = inst[i];
    const int nt = fill_static_args(b->p[i], b->L[i], b->tile, b->seg, bi.a);
    bi.a.w = nullptr;               // w is rebuilt from the psi pair (stages 2, 3)
    bi.psi[0] = b->bufs[2 * i];
    bi.psi[1] = b->bufs[2 * i + 1];
    bi.dt_re = b->p[i].dt_re;
    bi.dt_im = b->p[i].dt_im;
    bi.item_begin = b->item_begin[i];
    bi.pad_ = 0;
    b->item_begin[i + 1] = b->item_begin[i] + nt;
  }
  b->total_items = b->item_begin[b->n];
  std::vector<int> item_inst(b->total_items);
  for (int i = 0; i < b->n; ++i)
    for (int k = b->item_begin[i]; k < b->item_begin[i + 1]; ++k) item_inst[k] = i;
  if (b->d_inst) (void)hipFree(b->d_inst);
  if (b->d_item_inst) (void)hipFree(b->d_item_inst);
  if (b->d_item_begin) (void)hipFree(b->d_item_begin);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_part) (void)hipFree(b->d_part);
  b->d_inst = nullptr;
  b->d_item_inst = b->d_item_begin = b->d_items = nullptr;
  b->d_part = nullptr;
  BCHK(hipMalloc(&b->d_inst, sizeof(BatchInst) * b->n));
  BCHK(hipMalloc(&b->d_item_inst, sizeof(int) * b->total_items));
  BCHK(hipMalloc(&b->d_items, sizeof(int) * b->total_items));
  BCHK(hipMalloc(&b->d_item_begin, sizeof(int) * (b->n + 1)));
  BCHK(hipMalloc(&b->d_part, sizeof(double) * 4 * (size_t)b->total_items));
  BCHK(hipMemcpy(b->d_inst, inst.data(), sizeof(BatchInst) * b->n, hipMemcpyHostToDevice));
  BCHK(hipMemcpy(b->d_item_inst, item_inst.data(), sizeof(int) * b->total_items,
                 hipMemcpyHostToDevice));
  BCHK(hipMemcpy(b->d_item_begin, b->item_begin.data(), sizeof(int) * (b->n + 1),
                 hipMemcpyHostToDevice));
  std::vector<int> all(b->total_items);
  for (int k = 0; k < b->total_items; ++k) all[k] = k;
  BCHK(hipMemcpy(b->d_items, all.data(), sizeof(int) * b->total_items, hipMemcpyHostToDevice));
  b->active_items = b->total_items;
  b->active_done = 0;
  return 0;
}

// A replay may still be in flight: drain the stream before destroying its graphs.
void drop_graphs(tbec_batch *b) {
  bool any = false;
  for (auto &g : b->graph) any = any || g;
  if (any && b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto &g : b->graph)
    if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
  b->graph_steps = 0;
}

// Rebuild the launch's item list from the instances whose flag is 0 (every instance when
// `flags` is null).  The captured graphs hold the grid size, so they are dropped; the
// stream is drained before the list is rewritten.  Partials keep their item indices, so
// every instance's arithmetic and reduction order are unchanged.
int set_active_items(tbec_batch *b, const int *flags) {
  std::vector<int> items;
  items.reserve(b->total_items);
  int done = 0;
  for (int i = 0; i < b->n; ++i) {
    if (flags && flags[4 * i] != 0) {
      ++done;
      continue;
    }
    for (int k = b->item_begin[i]; k < b->item_begin[i + 1]; ++k) items.push_back(k);
  }
  if (items.empty()) return 0;            // nothing left to launch: keep the last list
  if ((int)items.size() == b->active_items && done == b->active_done) return 0;
  drop_graphs(b);
  BCHK(hipStreamSynchronize(b->stream));
  BCHK(hipMemcpyAsync(b->d_items, items.data(), sizeof(int) * items.size(),
                      hipMemcpyHostToDevice, b->stream));
  BCHK(hipStreamSynchronize(b->stream));
  b->active_items = (int)items.size();
  b->active_done = done;
  return 0;
}

int enqueue_batch_stage(tbec_batch *b, int kind) {
  BatchLaunch L;
  L.inst = b->d_inst;
  L.item_inst = b->d_item_inst;
  L.items = b->d_items;
  L.total_items = b->active_items;
  L.parity = b->parity;
  L.stage = kind;
  L.A = kind == ST_RK1 ? 0.0 : (kind == ST_RK2 ? -2.0 / 3.0 : -1.0);
  L.B = kind == ST_RK1 ? 1.0 / 3.0 : (kind == ST_RK2 ? 1.0 : 0.5);
  L.partials = b->d_part;
  L.state = b->d_state;
  L.state_stride = ST_COUNT;
  L.flags = b->d_flags;
  const int grid = b->active_items;  // one workgroup per (instance, tile) work item
  hipError_t e = launch_batch(b->prec, b->tile, L, b->sum, b->rot, grid, b->stream);
  if (e != hipSuccess)
    return set_error(TBEC_ERR_HIP, std::string("batch stage launch: ") + hipGetErrorString(e));
  if (kind == ST_RK1)
    hipLaunchKernelGGL(k_batch_fin_energy, dim3(b->n), dim3(256), 0, b->stream, b->d_part,
                       b->d_item_begin, b->d_state, b->d_flags);
  else if (kind == ST_RK3)
    hipLaunchKernelGGL(k_batch_fin_norm, dim3(b->n), dim3(256), 0, b->stream, b->d_part,
                       b->d_item_begin, b->d_state, b->d_flags);
  BCHK(hipGetLastError());
  b->parity ^= 1;
  return 0;
}

int enqueue_batch_steps(tbec_batch *b, long long n) {
  for (long long s = 0; s < n; ++s) {
    int rc = enqueue_batch_stage(b, ST_RK1);
    if (!rc) rc = enqueue_batch_stage(b, ST_RK2);
    if (!rc) rc = enqueue_batch_stage(b, ST_RK3);
    if (rc) return rc;
  }
  return 0;
}

// Replay a captured graph of n (even) steps from the current parity.
int replay_batch(tbec_batch *b, long long n) {
  if (b->graph_steps != n) drop_graphs(b);
  hipGraphExec_t &g = b->graph[b->parity];
  if (!g) {
    const int p0 = b->parity;
    BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_batch_steps(b, n);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &gr);
    b->parity = p0;
    if (rc) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess) return set_error(TBEC_ERR_HIP, "batch graph capture failed");
    e = hipGraphInstantiate(&g, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) return set_error(TBEC_ERR_HIP, "batch graph instantiate failed");
    b->graph_steps = n;
  }
  BCHK(hipGraphLaunch(g, b->stream));
  b->parity ^= (int)((3 * n) & 1);
  return 0;
}

int reset_states(tbec_batch *b, double norm, double tol) {
  for (int i = 0; i < b->n; ++i) {
    double *s = b->h_state + (size_t)i * ST_COUNT;
    for (int k = 0; k < ST_COUNT; ++k) s[k] = 0.0;
    s[ST_SCALE] = 1.0;
    s[ST_C0EFF] = s[ST_C0NEXT] = s[ST_C0BASE] =
        b->prec == TBEC_F32 ? (double)(float)b->p[i].c0 : b->p[i].c0;
    s[ST_GNL] = b->p[i].gnl;
    s[ST_DV] = b->L[i].hx * b->L[i].hy * b->L[i].hz;
    s[ST_NORMT] = norm;
    s[ST_TOL] = tol;
    s[ST_RGATE] = b->rgate;
    s[ST_F32] = b->prec == TBEC_F32 ? 1.0 : 0.0;
    for (int k = 0; k < 4; ++k) b->h_flags[4 * i + k] = 0;
  }
  BCHK(hipMemcpyAsync(b->d_state, b->h_state, sizeof(double) * ST_COUNT * b->n,
                      hipMemcpyHostToDevice, b->stream));
  BCHK(hipMemcpyAsync(b->d_flags, b->h_flags, sizeof(int) * 4 * b->n, hipMemcpyHostToDevice,
                      b->stream));
  BCHK(hipStreamSynchronize(b->stream));
  return 0;
}

}  // namespace

extern "C" {

int tbec_batch_create(int32_t n, const tbec_params *params, tbec_batch **out) {
  if (!params || !out || n < 1) return set_error(TBEC_ERR_ARG, "bad batch arguments");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    const tbec_params &q = params[i];
    if (q.prec != params[0].prec || q.sum_mode != params[0].sum_mode)
      return set_error(TBEC_ERR_ARG, "batch instances must share prec and sum_mode");
    if (q.nx < 1 || q.ny < 1 || q.nz < 1 || !(q.Lx > 0 && q.Ly > 0 && q.Lz > 0))
      return set_error(TBEC_ERR_ARG, "batch instance " + std::to_string(i) + ": bad grid");
    if (q.dt_re != 0.0 || !(q.dt_im < 0.0))
      return set_error(TBEC_ERR_ARG, "batch instances are imaginary-time: dt = (0, -dtau)");
  }
  if (params[0].prec != TBEC_F32 && params[0].prec != TBEC_F64)
    return set_error(TBEC_ERR_ARG, "prec must be 32 or 64");
  if (params[0].prec == TBEC_F64 && params[0].sum_mode != TBEC_SUM_REFERENCE)
    return set_error(TBEC_ERR_ARG, "FP64 runs use TBEC_SUM_REFERENCE");
  tbec_batch *b = new tbec_batch();
  b->n = n;
  b->prec = params[0].prec;
  b->sum = params[0].sum_mode;
  b->tile = params[0].tile < 0 ? 2 : params[0].tile;
  b->p.assign(params, params + n);
  b->L.resize(n);
  b->cur.assign(n, 0);
  for (int i = 0; i < n; ++i) {
    compute_layout(b->p[i], b->L[i]);
    b->rot = b->rot || rot_enabled(b->p[i]);
    if (b->L[i].sx * b->L[i].csize > 0x7fffffffLL) {
      delete b;
      return set_error(TBEC_ERR_ARG, "an instance's x-plane exceeds 2^31 bytes");
    }
  }
  auto cleanup = [&](int rc) { tbec_batch_destroy(b); return rc; };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(set_error(TBEC_ERR_HIP, "stream creation failed"));
  for (int i = 0; i < n; ++i) {
    const size_t bytes = (size_t)b->L[i].elems * b->L[i].csize;
    for (int k = 0; k < 2; ++k) {
      void *ptr = nullptr;
      if (hipMalloc(&ptr, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return cleanup(set_error(TBEC_ERR_OOM, "batch buffer allocation failed"));
      }
      b->bufs.push_back(ptr);
      if (hipMemsetAsync(ptr, 0, bytes, b->stream) != hipSuccess)
        return cleanup(set_error(TBEC_ERR_HIP, "memset failed"));
    }
    b->stage_bytes = std::max(b->stage_bytes, (size_t)b->p[i].nx * b->p[i].ny * b->p[i].nz *
                                                  b->L[i].csize);
  }
  if (hipMalloc(&b->stage, b->stage_bytes) != hipSuccess ||
      hipMalloc(&b->d_state, sizeof(double) * ST_COUNT * n) != hipSuccess ||
      hipMalloc(&b->d_flags, sizeof(int) * 4 * n) != hipSuccess ||
      hipHostMalloc(&b->h_state, sizeof(double) * ST_COUNT * n) != hipSuccess ||
      hipHostMalloc(&b->h_flags, sizeof(int) * 4 * n) != hipSuccess) {
    (void)hipGetLastError();
    return cleanup(set_error(TBEC_ERR_OOM, "batch state allocation failed"));
  }
  b->seg = batch_auto_seg(b, b->tile);
  int rc = build_table(b);
  if (!rc) rc = reset_states(b, 1.0, 0.0);
  if (rc) return cleanup(rc);
  *out = b;
  return 0;
}

int tbec_batch_destroy(tbec_batch *b) {
  if (!b) return 0;
  drop_graphs(b);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (void *p : b->bufs) (void)hipFree(p);
  if (b->stage) (void)hipFree(b->stage);
  if (b->d_inst) (void)hipFree(b->d_inst);
  if (b->d_item_inst) (void)hipFree(b->d_item_inst);
  if (b->d_item_begin) (void)hipFree(b->d_item_begin);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_part) (void)hipFree(b->d_part);
  if (b->d_state) (void)hipFree(b->d_state);
  if (b->d_flags) (void)hipFree(b->d_flags);
  if (b->h_state) (void)hipHostFree(b->h_state);
  if (b->h_flags) (void)hipHostFree(b->h_flags);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int tbec_batch_set_tuning(tbec_batch *b, int tile, int seg) {
  if (!b) return set_error(TBEC_ERR_ARG, "null batch");
  if (tile < 0 || tile >= kNumTiles || seg < 0 || seg > 4096)
    return set_error(TBEC_ERR_ARG, "tile or segment out of range");
  BCHK(hipStreamSynchronize(b->stream));
  drop_graphs(b);
  b->tile = tile;
  b->seg = seg > 0 ? seg : batch_auto_seg(b, tile);
  return build_table(b);
}

int tbec_batch_set_tail_segment(tbec_batch *b, int on) {
  if (!b) return set_error(TBEC_ERR_ARG, "null batch");
  if (on != 0 && on != 1) return set_error(TBEC_ERR_ARG, "tail segment flag must be 0 or 1");
  b->tail_seg_on = on;
  return 0;
}

int tbec_batch_set_residual_gate(tbec_batch *b, double rmax) {
  if (!b) return set_error(TBEC_ERR_ARG, "null batch");
  if (!(rmax >= 0.0) || !std::isfinite(rmax))
    return set_error(TBEC_ERR_ARG, "residual gate must be finite and >= 0");
  b->rgate = rmax;
  return 0;
}

void *tbec_batch_stream(tbec_batch *b) { return b ? (void *)b->stream : nullptr; }

int tbec_batch_upload_psi(tbec_batch *b, int32_t i, const void *psi_host) {
  if (!b || !psi_host || i < 0 || i >= b->n) return set_error(TBEC_ERR_ARG, "bad argument");
  const tbec_params &q = b->p[i];
  const Layout &L = b->L[i];
  const long long n = (long long)q.nx * q.ny * q.nz;
  BCHK(hipMemcpyAsync(b->stage, psi_host, n * L.csize, hipMemcpyHostToDevice, b->stream));
  void *dst = b->bufs[2 * i + b->parity];
  if (b->prec == TBEC_F32)
    hipLaunchKernelGGL(k_batch_unpack<float>, dim3(grid_blocks(n)), dim3(256), 0, b->stream,
                       (const cplx<float> *)b->stage, (cplx<float> *)dst, q.nx, q.ny, q.nz,
                       L.sy, L.sx, L.base);
  else
    hipLaunchKernelGGL(k_batch_unpack<double>, dim3(grid_blocks(n)), dim3(256), 0, b->stream,
                       (const cplx<double> *)b->stage, (cplx<double> *)dst, q.nx, q.ny, q.nz,
                       L.sy, L.sx, L.base);
  BCHK(hipGetLastError());
  // the uploaded state is the instance's psi as given: no pending renormalisation
  // (tbec_batch_download_psi multiplies by the scale a solve leaves behind)
  const double one = 1.0;
  BCHK(hipMemcpyAsync(b->d_state + (size_t)i * ST_COUNT + ST_SCALE, &one, sizeof(double),
                      hipMemcpyHostToDevice, b->stream));
  BCHK(hipStreamSynchronize(b->stream));
  b->cur[i] = b->parity;
  return 0;
}

int tbec_batch_download_psi(tbec_batch *b, int32_t i, void *psi_host) {
  if (!b || !psi_host || i < 0 || i >= b->n) return set_error(TBEC_ERR_ARG, "bad argument");
  const tbec_params &q = b->p[i];
  const Layout &L = b->L[i];
  const long long n = (long long)q.nx * q.ny * q.nz;
  const void *src = b->bufs[2 * i + b->cur[i]];
  const double *scale = b->d_state + (size_t)i * ST_COUNT + ST_SCALE;
  if (b->prec == TBEC_F32)
    hipLaunchKernelGGL(k_batch_pack<float>, dim3(grid_blocks(n)), dim3(256), 0, b->stream,
                       (const cplx<float> *)src, (cplx<float> *)b->stage, q.nx, q.ny, q.nz,
                       L.sy, L.sx, L.base, scale);
  else
    hipLaunchKernelGGL(k_batch_pack<double>, dim3(grid_blocks(n)), dim3(256), 0, b->stream,
                       (const cplx<double> *)src, (cplx<double> *)b->stage, q.nx, q.ny, q.nz,
                       L.sy, L.sx, L.base, scale);
  BCHK(hipGetLastError());
  BCHK(hipMemcpyAsync(psi_host, b->stage, n * L.csize, hipMemcpyDeviceToHost, b->stream));
  BCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int tbec_batch_prolong(tbec_batch *dst, tbec_batch *src) {
  if (!dst || !src || dst == src) return set_error(TBEC_ERR_ARG, "prolongation needs two batches");
  if (dst->n != src->n) return set_error(TBEC_ERR_ARG, "batches of different instance counts");
  if (dst->prec != src->prec)
    return set_error(TBEC_ERR_ARG, "prolongation between batches of different precision");
  for (int i = 0; i < dst->n; ++i)
    if (!same_box(dst->p[i], src->p[i]))
      return set_error(TBEC_ERR_ARG, "instance " + std::to_string(i) + ": different boxes");
  // the source's pending steps finish before its states are read
  BCHK(hipStreamSynchronize(src->stream));
  const double one = 1.0;
  for (int i = 0; i < dst->n; ++i) {
    const Layout &Ls = src->L[i], &Ld = dst->L[i];
    const GridView sv{src->p[i].nx, src->p[i].ny, src->p[i].nz, Ls.sy, Ls.sx, Ls.base};
    const GridView dv{dst->p[i].nx, dst->p[i].ny, dst->p[i].nz, Ld.sy, Ld.sx, Ld.base};
    BCHK(launch_prolong(dst->prec, src->bufs[2 * i + src->cur[i]], sv,
                        src->d_state + (size_t)i * ST_COUNT + ST_SCALE,
                        dst->bufs[2 * i + dst->parity], dv, dst->stream));
    // the interpolated state carries its source's renormalisation: none is pending
    BCHK(hipMemcpyAsync(dst->d_state + (size_t)i * ST_COUNT + ST_SCALE, &one, sizeof(double),
                        hipMemcpyHostToDevice, dst->stream));
  }
  BCHK(hipStreamSynchronize(dst->stream));
  for (int i = 0; i < dst->n; ++i) dst->cur[i] = dst->parity;
  return 0;
}

int tbec_batch_imag_solve(tbec_batch *b, int64_t max_steps, double tol, double norm,
                          int check_every, tbec_solve_result *out) {
  if (!b || !out || max_steps < 1 || check_every < 1 || !(norm > 0))
    return set_error(TBEC_ERR_ARG, "bad argument");
  // every instance must start from the same buffer parity
  for (int i = 0; i < b->n; ++i)
    if (b->cur[i] != b->parity)
      return set_error(TBEC_ERR_STATE, "upload every instance before a batch solve");
  int rc = reset_states(b, norm, tol);
  if (!rc) rc = set_active_items(b, nullptr);   // every instance starts unfinished
  if (rc) return rc;
  const int p0 = b->parity;
  const int seg0 = b->seg;
  long long chunk = check_every + (check_every & 1);
  long long done = 0;
  bool all = false;
  const double every = progress_interval();
  auto last = std::chrono::steady_clock::now();
  while (done < max_steps && !all) {
    const long long nst = std::min<long long>(chunk, max_steps - done);
    rc = nst == chunk && chunk > 1 ? replay_batch(b, nst) : enqueue_batch_steps(b, nst);
    if (rc) return rc;
    done += nst;
    BCHK(hipMemcpyAsync(b->h_flags, b->d_flags, sizeof(int) * 4 * b->n,
                        hipMemcpyDeviceToHost, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    all = true;
    int nconv = 0;
    for (int i = 0; i < b->n; ++i) {
      all = all && b->h_flags[4 * i] != 0;
      nconv += b->h_flags[4 * i] != 0;
    }
    if (every > 0) {
      const auto now = std::chrono::steady_clock::now();
      if (std::chrono::duration<double>(now - last).count() >= every) {
        std::fprintf(stderr, "[tbec] batch solve: %lld steps, %d/%d finished\n", done, nconv,
                     b->n);
        last = now;
      }
    }
    // finished instances leave the launches (their workgroups would only exit); with the
    // tail option a tail of few unfinished instances also gets a shorter segment
    // (restored after the solve, so that every solve starts with the batch's own tuning)
    if (!all && nconv > b->active_done) {
      const int ts = b->tail_seg_on ? tail_seg(b, b->h_flags) : b->seg;
      if (ts < b->seg) {
        drop_graphs(b);
        b->seg = ts;
        rc = build_table(b);
        if (rc) return rc;
      }
      rc = set_active_items(b, b->h_flags);
      if (rc) return rc;
    }
  }
  if (b->seg != seg0) {
    drop_graphs(b);
    b->seg = seg0;
    rc = build_table(b);
    if (rc) return rc;
  }
  BCHK(hipMemcpyAsync(b->h_state, b->d_state, sizeof(double) * ST_COUNT * b->n,
                      hipMemcpyDeviceToHost, b->stream));
  BCHK(hipStreamSynchronize(b->stream));
  int bad = 0;
  for (int i = 0; i < b->n; ++i) {
    const double *s = b->h_state + (size_t)i * ST_COUNT;
    const int flag = b->h_flags[4 * i];
    out[i].steps = (int64_t)s[ST_STEP];
    out[i].converged = flag == 1;
    out[i].pad_ = 0;
    out[i].energy = s[ST_E];
    out[i].mu = s[ST_MU];
    out[i].residual = s[ST_RES];
    out[i].norm = s[ST_N];
    out[i].rel_change = s[ST_REL];
    // a finished instance holds psi_n in the input buffer of its step n
    b->cur[i] = flag ? (p0 ^ (int)((3 * (long long)b->h_flags[4 * i + 1]) & 1)) : b->parity;
    if (flag == 2) bad = 1;
  }
  if (bad) return set_error(TBEC_ERR_NONFINITE, "non-finite energy in a batch instance");
  return 0;
}

}  // extern "C"
