// libtbec: handle, device layout, launch plumbing and reductions for the fused GPE
// stage kernel (gpe_stage.hpp).  C ABI in include/tbec.h.
#include "tbec.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>

#include "gpe_stage.hpp"
#include "stage_launch.hpp"
#include "tbec_internal.hpp"
#include "tbec_fin.hpp"

using namespace tbec;

namespace {
thread_local std::string g_err;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
}  // namespace

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      return fail(TBEC_ERR_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                             \
  } while (0)

namespace tbec {

int set_error(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// ---------------------------------------------------------------- tile table ----
// Each precision has a small set of tile shapes (TZ x TY x VZ); the segment length SX
// is a runtime parameter.  The autotuner (trapbec_amd/autotune.py) picks (tile, SX).
const TileCfg kTilesF32[] = {{64, 16, 4}, {64, 32, 2}, {64, 16, 2}, {32, 32, 4},
                             {64, 16, 1}};
const TileCfg kTilesF64[] = {{32, 16, 2}, {64, 8, 2}, {32, 16, 1}, {64, 16, 1}, {32, 8, 1}};
const int kNumTiles = 5;

const TileCfg &tile_cfg(int prec, int id) {
  return prec == TBEC_F32 ? kTilesF32[id] : kTilesF64[id];
}

long long tiles_per_plane(const tbec_params &p, int tile) {
  const TileCfg &t = tile_cfg(p.prec, tile);
  return (long long)((p.nz + t.tz - 1) / t.tz) * ((p.ny + t.ty - 1) / t.ty);
}

// Segment length when none is given: the longest of 256, 128, ..., 8 planes that still
// gives `want` workgroups (a few per CU on 256 CUs).  Grids too small for that take 8, or
// 4 where 8 would leave fewer than want/2 workgroups: there a shorter march's extra
// prologue loads cost less than idle CUs (FP32, tile (64,16,2): 128^3 imaginary time
// 54.6 -> 52.1 us per step, real time 40.3 -> 36.7; 64^3 42.9 -> 31.2 and 29.1 -> 20.4;
// at 192^3 the 864 workgroups of segment 8 beat segment 4: 133.5 against 138.8 us).
int auto_seg(long long tiles, int nx, long long want) {
  for (int seg = 256; seg > 8; seg >>= 1)
    if (tiles * ((nx + seg - 1) / seg) >= want) return seg;
  return 2 * tiles * ((nx + 7) / 8) < want ? 4 : 8;
}

hipError_t launch_stage(int prec, int tile, const StageArgs &a, int sum, int phys,
                        bool imag, hipStream_t s) {
  if (prec == TBEC_F32) {
    switch (tile) {
      case 0: return launch_stage_p32_t0(a, sum, phys, imag, s);
      case 1: return launch_stage_p32_t1(a, sum, phys, imag, s);
      case 2: return launch_stage_p32_t2(a, sum, phys, imag, s);
      case 3: return launch_stage_p32_t3(a, sum, phys, imag, s);
      default: return launch_stage_p32_t4(a, sum, phys, imag, s);
    }
  } else {
    switch (tile) {
      case 0: return launch_stage_p64_t0(a, sum, phys, imag, s);
      case 1: return launch_stage_p64_t1(a, sum, phys, imag, s);
      case 2: return launch_stage_p64_t2(a, sum, phys, imag, s);
      case 3: return launch_stage_p64_t3(a, sum, phys, imag, s);
      default: return launch_stage_p64_t4(a, sum, phys, imag, s);
    }
  }
}

// ------------------------------------------------------------- ghost fill ----
// ghost_comm's boundary faces (src/ghost.f90:27-192) on the padded array: one thread per
// face row (two axes' interior point), writing the 3 ghosts on both sides of the third
// axis.  bc: 0 zero, 1 solid (ghost -g = psi g, 0-based), 2 periodic (ghost -g =
// psi n-g); z follows bc unless zperiodic.  Corners are not touched (they stay zero);
// the stencils never read them.
template <typename T>
__global__ void k_fill_ghosts(cplx<T> *__restrict__ psi, int nx, int ny, int nz, long long sy,
                              long long sx, long long base, int bc, int zperiodic) {
  const long long fx = (long long)ny * nz, fy = (long long)nx * nz, fz = (long long)nx * ny;
  const long long total = fx + fy + fz;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int a, n, b = bc;
    long long row, s;
    if (t < fx) {
      const long long j = t / nz, k = t % nz;
      a = 0; n = nx; s = sx; row = base + j * sy + k;
    } else if (t < fx + fy) {
      const long long q = t - fx, i = q / nz, k = q % nz;
      a = 1; n = ny; s = sy; row = base + i * sx + k;
    } else {
      const long long q = t - fx - fy, i = q / ny, j = q % ny;
      a = 2; n = nz; s = 1; row = base + i * sx + j * sy;
      if (zperiodic) b = 2;
    }
    (void)a;
    for (int g = 1; g <= 3; ++g) {
      cplx<T> lo{T(0), T(0)}, hi{T(0), T(0)};
      if (b == 1) {
        lo = psi[row + (long long)g * s];
        hi = psi[row + (long long)(n - 1 - g) * s];
      } else if (b == 2) {
        lo = psi[row + (long long)(n - g) * s];
        hi = psi[row + (long long)(g - 1) * s];
      }
      psi[row - (long long)g * s] = lo;
      psi[row + (long long)(n - 1 + g) * s] = hi;
    }
  }
}

hipError_t fill_ghosts(tbec_handle *h, void *buf, int bc, int zperiodic) {
  const long long total = (long long)h->p.ny * h->p.nz + (long long)h->p.nx * h->p.nz +
                          (long long)h->p.nx * h->p.ny;
  const int grid = (int)std::max<long long>(1, std::min<long long>((total + 255) / 256, 4096));
  if (h->p.prec == TBEC_F32)
    hipLaunchKernelGGL(k_fill_ghosts<float>, dim3(grid), dim3(256), 0, h->stream,
                       (cplx<float> *)buf, h->p.nx, h->p.ny, h->p.nz, h->sy, h->sx, h->base,
                       bc, zperiodic);
  else
    hipLaunchKernelGGL(k_fill_ghosts<double>, dim3(grid), dim3(256), 0, h->stream,
                       (cplx<double> *)buf, h->p.nx, h->p.ny, h->p.nz, h->sy, h->sx, h->base,
                       bc, zperiodic);
  return hipGetLastError();
}

// Fill the ghosts of the current psi when the boundaries are not all zero (the zero
// halo is permanent and needs no pass).
int refresh_ghosts(tbec_handle *h) {
  if (h->bc == TBEC_BC_ZERO && !h->periodicz) return 0;
  hipError_t e = fill_ghosts(h, h->psi[h->cur], h->bc, h->periodicz);
  if (e != hipSuccess) return fail(TBEC_ERR_HIP, "ghost fill failed: %s", hipGetErrorString(e));
  return 0;
}

// ------------------------------------------------------------- small kernels ----
template <typename T>
__global__ void k_unpack(const cplx<T> *__restrict__ src, cplx<T> *__restrict__ dst, int nx,
                         int ny, int nz, long long sy, long long sx, long long base) {
  const long long n = (long long)nx * ny * nz;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long z = i % nz, r = i / nz, y = r % ny, x = r / ny;
    dst[base + x * sx + y * sy + z] = src[i];
  }
}

template <typename T>
__global__ void k_pack(const cplx<T> *__restrict__ src, cplx<T> *__restrict__ dst, int nx,
                       int ny, int nz, long long sy, long long sx, long long base,
                       const double *scale) {
  const long long n = (long long)nx * ny * nz;
  const T s = scale ? T(*scale) : T(1);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long z = i % nz, r = i / nz, y = r % ny, x = r / ny;
    cplx<T> v = src[base + x * sx + y * sy + z];
    if (scale) { v.re *= s; v.im *= s; }
    dst[i] = v;
  }
}

// psi *= scale over the interior, then scale = 1.
template <typename T>
__global__ void k_apply_scale(cplx<T> *__restrict__ p, int nx, int ny, int nz, long long sy,
                              long long sx, long long base, const double *scale) {
  const long long n = (long long)nx * ny * nz;
  const T s = T(*scale);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long z = i % nz, r = i / nz, y = r % ny, x = r / ny;
    cplx<T> &v = p[base + x * sx + y * sy + z];
    v.re *= s;
    v.im *= s;
  }
}

__global__ void k_reset_scale(double *st) { st[ST_SCALE] = 1.0; }

// Psi_info partials: max|psi|, min|psi|, sum|psi|, sum|psi|^2 (src/diagnostic.f90:7-25).
template <typename T>
__global__ void __launch_bounds__(256)
k_info(const cplx<T> *__restrict__ p, int nx, int ny, int nz, long long sy, long long sx,
       long long base, const double *scale, double *part) {
  const long long n = (long long)nx * ny * nz;
  const double s = scale ? *scale : 1.0;
  double mx = 0.0, mn = INFINITY, s1 = 0.0, s2 = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long z = i % nz, r = i / nz, y = r % ny, x = r / ny;
    const cplx<T> v = p[base + x * sx + y * sy + z];
    const double re = (double)(T(s) * v.re), im = (double)(T(s) * v.im);
    const double a2 = re * re + im * im, a = sqrt(a2);
    mx = fmax(mx, a);
    mn = fmin(mn, a);
    s1 += a;
    s2 += a2;
  }
  __shared__ double red[4][256];
  red[0][threadIdx.x] = mx;
  red[1][threadIdx.x] = mn;
  red[2][threadIdx.x] = s1;
  red[3][threadIdx.x] = s2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + o]);
      red[1][threadIdx.x] = fmin(red[1][threadIdx.x], red[1][threadIdx.x + o]);
      red[2][threadIdx.x] += red[2][threadIdx.x + o];
      red[3][threadIdx.x] += red[3][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int q = 0; q < 4; ++q) part[4 * blockIdx.x + q] = red[q][0];
  }
}

__global__ void __launch_bounds__(256)
k_fin_energy(const double *part, int nb, double *st, int *flag, int check_conv) {
  if (check_conv && flag[0]) return;
  double s[4];
  sum_partials(part, 0, nb, s, false);
  if (threadIdx.x == 0) fin_energy(s, st, flag, check_conv);
}

// Diagnostics: c0 = 2*mu for the residual pass.  With rhs(c) = K(psi) + c*psi and
// H psi = -K/2, (H - mu) psi = -rhs(2 mu)/2, so the pass's sum of |rhs|^2 is the
// residual without the cancellation of (1/4) R - a S1 + a^2 N at a = c0/2 - mu.
__global__ void k_set_resid_c0(double *st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double c = 2.0 * st[ST_MU];
    st[ST_DC0] = st[ST_F32] != 0.0 ? (double)(float)c : c;   // as the kernel will read it
  }
}

__global__ void __launch_bounds__(256)
k_fin_resid(const double *part, int nb, double *st) {
  double s[4];
  sum_partials(part, 0, nb, s, false);
  if (threadIdx.x == 0) {
    // (H - mu) psi = -rhs(c)/2 + a psi with a = c/2 - mu, |a| ~ rounding of c only
    const double R = s[3] * st[ST_DV], S1 = s[1] * st[ST_DV], N = st[ST_N];
    const double a = 0.5 * st[ST_DC0] - st[ST_MU];
    double r2 = 0.25 * R - a * S1 + a * a * N;
    if (r2 < 0) r2 = 0;
    st[ST_RES] = N > 0.0 ? sqrt(r2 / N) : 0.0;
  }
}

__global__ void __launch_bounds__(256)
k_fin_norm(const double *part, int nb, double *st, int *flag) {
  if (flag[0]) return;
  double s[4];
  sum_partials(part, 0, nb, s, false);
  if (threadIdx.x == 0) fin_norm(s, st, flag);
}

__global__ void __launch_bounds__(256) k_fin_info(const double *part, int nb, double *st) {
  double s[4];
  sum_partials(part, 0, nb, s, true);
  if (threadIdx.x != 0) return;
  st[ST_MAXA] = s[0];
  st[ST_MINA] = s[1];
  st[ST_SUMA] = s[2];
  st[ST_SUMA2] = s[3];
}

}  // namespace tbec

// ------------------------------------------------------------------ handle ------
namespace tbec {
long long round_up(long long v, long long m) { return (v + m - 1) / m * m; }

// Padded device layout of one condensate: interior z starts at a 128-byte boundary;
// tiles of up to 128 in z and 64 in y may overrun the interior and read zeros; x has 3
// ghost planes below and 4 above (the register queue prefetches plane x+4; where the stage
// kernel prefetches plane x+5, the loads of planes past the segment's last used one,
// x+5 > xe+2, which covers every plane past nx+3, go out at an offset past the record
// count of an existing plane).
void compute_layout(const tbec_params &p, Layout &L) {
  L.csize = p.prec == TBEC_F32 ? 8 : 16;
  L.hx = p.Lx / p.nx;
  L.hy = p.Ly / p.ny;
  L.hz = p.Lz / p.nz;
  const long long zoff = 128 / L.csize;
  L.PZ = round_up(zoff + round_up(p.nz, 128) + 4, zoff);
  L.PY = 3 + round_up(p.ny, 64) + 3;
  L.PX = 3 + (long long)p.nx + 4;
  L.sy = L.PZ;
  L.sx = L.PZ * L.PY;
  L.base = 3 * L.sx + 3 * L.sy + zoff;
  L.elems = L.sx * L.PX;
}

// Static stage arguments of one condensate (geometry, physics, tiling); returns the
// number of tiles.  Buffers, stage fields and reduction pointers are set by the caller.
int fill_static_args(const tbec_params &p, const Layout &L, int tile, int seg, StageArgs &a) {
  const TileCfg &tc = tile_cfg(p.prec, tile);
  a.psi_in = nullptr;
  a.psi_out = nullptr;
  a.w = nullptr;
  a.nx = p.nx; a.ny = p.ny; a.nz = p.nz;
  a.seg_len = seg;
  a.tiles_z = (int)((p.nz + tc.tz - 1) / tc.tz);
  a.tiles_y = (int)((p.ny + tc.ty - 1) / tc.ty);
  a.nseg = (p.nx + seg - 1) / seg;
  a.nblocks = a.tiles_z * a.tiles_y * a.nseg;
  a.sy = L.sy; a.sx = L.sx; a.base = L.base; a.elems = L.elems;
  a.x0 = -0.5 * p.Lx + 0.5 * L.hx;
  a.y0 = -0.5 * p.Ly + 0.5 * L.hy;
  a.z0 = -0.5 * p.Lz + 0.5 * L.hz;
  a.hx = L.hx; a.hy = L.hy; a.hz = L.hz;
  a.k2[0] = 1.0 / (180.0 * L.hx * L.hx);
  a.k2[1] = 1.0 / (180.0 * L.hy * L.hy);
  a.k2[2] = 1.0 / (180.0 * L.hz * L.hz);
  a.k1[0] = 1.0 / (60.0 * L.hx);
  a.k1[1] = 1.0 / (60.0 * L.hy);
  a.wl[0] = p.harm_wx / (p.Lx * p.Lx);
  a.wl[1] = p.harm_wy / (p.Ly * p.Ly);
  a.wl[2] = p.harm_wz / (p.Lz * p.Lz);
  a.c0 = p.c0; a.gnl = p.gnl; a.om = p.omega_z;
  a.cre = a.cim = a.A = a.B = 0.0;
  a.wrec = 0.0;
  a.store_w = 0;
  a.stage = 0;
  a.remap = 1;
  a.scale = nullptr; a.prev_scale = nullptr; a.c0dev = nullptr; a.skip = nullptr;
  a.partials = nullptr;
  a.vtab = nullptr; a.drift = 0; a.vx2 = 0.0;
  return a.nblocks;
}

static Layout layout_of(const tbec_handle *h) {
  Layout L;
  L.csize = h->csize;
  L.hx = h->hx; L.hy = h->hy; L.hz = h->hz;
  L.PZ = h->PZ; L.PY = h->PY; L.PX = h->PX;
  L.sy = h->sy; L.sx = h->sx; L.base = h->base; L.elems = h->elems;
  return L;
}

static int fill_args(tbec_handle *h, StageArgs &a, int stage, int src, double A, double B,
                     double cre, double cim, int tile, int seg) {
  fill_static_args(h->p, layout_of(h), tile, seg, a);
  a.psi_in = h->psi[src];
  a.psi_out = h->psi[src ^ 1];
  a.w = h->w;
  a.cre = cre; a.cim = cim; a.A = A; a.B = B;
  a.stage = stage;
  // TBEC_REMAP_FIRST_STAGE: the XCD-contiguous order for the stages that read one array
  // (the first RK stage, Euler, rhs-only), the dispatch order for the two-array stages
  a.remap = h->remap == 1 ? 1 : (h->remap == 2 && stage != ST_RK2 && stage != ST_RK3) ? 1 : 0;
  return a.nblocks;
}

bool rot_enabled(const tbec_params &p) {
  // the reference enables the frame term when |omega_z| > epsilon(0.) of the working
  // precision (src/finite_diff.f90:128)
  const double eps = p.prec == TBEC_F32 ? 1.1920928955078125e-07 : 2.220446049250313e-16;
  return std::fabs(p.omega_z) > eps;
}

bool handle_rot(const tbec_handle *h) { return rot_enabled(h->p); }

int tbec_reserve_partials(tbec_handle *h, int nb) {
  if (nb > h->npart) {
    if (h->partials) (void)hipFree(h->partials);
    h->partials = nullptr;
    HIPCHK(hipMalloc(&h->partials, sizeof(double) * 4 * (size_t)nb));
    h->npart = nb;
  }
  return 0;
}

// Enqueue one stage.  kind: StageKind; imag: use the imaginary-time device state.
int enqueue_stage(tbec_handle *h, int kind, bool imag, bool reduce, double *part) {
  StageArgs a;
  const double dr = h->dt_re, di = h->dt_im;
  // eye*0.5*dt (src/run.f90:47)
  double cre = -0.5 * di, cim = 0.5 * dr;
  double A = 0.0, B = 1.0;
  if (kind == ST_RK1) { A = 0.0; B = 1.0 / 3.0; }
  else if (kind == ST_RK2) { A = -2.0 / 3.0; B = 1.0; }
  else if (kind == ST_RK3) { A = -1.0; B = 0.5; }
  else if (kind == ST_EULER) { cre = 0.5 * h->p.dt_re; cim = 0.5 * h->p.dt_im; }  // dt*0.5
  // the first RK stage of a real-time step may run with a tiling of its own
  // (tbec_set_first_stage_tuning); every other stage uses the handle's
  const bool own1 = h->tile1 >= 0 && kind == ST_RK1 && !imag && !reduce;
  const int tile = own1 ? h->tile1 : h->tile, seg = own1 ? h->seg1 : h->seg;
  int nb = fill_args(h, a, kind, h->cur, A, B, cre, cim, tile, seg);
  const bool want_red = imag || reduce;
  if (want_red) {
    if (!part) {
      int rc = tbec_reserve_partials(h, nb);
      if (rc) return rc;
      part = h->partials;
    }
    a.partials = part;
  }
  if (imag) {
    a.c0dev = h->dstate + ST_C0EFF;
    a.skip = h->dflag;
    if (kind == ST_RK1) a.scale = h->dstate + ST_SCALE;
    if (kind == ST_RK2) a.prev_scale = h->dstate + ST_SCALE;
  } else if (reduce) {
    a.scale = h->dstate + ST_SCALE;
    if (h->resid_pass) a.c0dev = h->dstate + ST_DC0;
  }
  if (h->w_stored) {
    a.store_w = kind == ST_RK1 || kind == ST_RK2;
  } else if (kind == ST_RK2 || kind == ST_RK3) {
    // w of the previous stage comes back from that stage's input, which is in the
    // buffer this stage overwrites (1/B_1 = 3, 1/B_2 = 1)
    a.w = a.psi_out;
    a.wrec = kind == ST_RK2 ? 3.0 : 1.0;
  }
  // physics of the current step (tbec::prepare_physics); two_dim drops D2z exactly
  // (x + 0*y == x for finite y)
  if (h->phys.two_dim) a.k2[2] = 0.0;
  if (h->ph_mode == PH_GEN) {
    a.vtab = h->vtab;
    a.drift = h->ph_drift;
    a.vx2 = 2.0 * h->phys.advective_vx;
  } else if (h->ph_wl_zero) {
    a.wl[0] = a.wl[1] = a.wl[2] = 0.0;
  }
  int rcg = refresh_ghosts(h);
  if (rcg) return rcg;
  hipError_t e = launch_stage(h->p.prec, tile, a, h->p.sum_mode, h->ph_mode, want_red,
                              h->stream);
  if (e != hipSuccess)
    return fail(TBEC_ERR_HIP, "stage launch failed: %s", hipGetErrorString(e));
  if (imag) {
    if (kind == ST_RK1) {
      hipLaunchKernelGGL(k_fin_energy, dim3(1), dim3(256), 0, h->stream, part, nb,
                         h->dstate, h->dflag, 1);
    } else if (kind == ST_RK3) {
      hipLaunchKernelGGL(k_fin_norm, dim3(1), dim3(256), 0, h->stream, part, nb,
                         h->dstate, h->dflag);
    }
    HIPCHK(hipGetLastError());
  }
  if (kind != ST_RHS) h->cur ^= 1;
  return 0;
}

static int grid_for(long long n) {
  long long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace tbec

extern "C" {

const char *tbec_last_error(void) { return g_err.c_str(); }

const char *tbec_version(void) { return "libtbec 0.1 (gfx950)"; }

int tbec_create(const tbec_params *p, tbec_handle **out) {
  if (!p || !out) return fail(TBEC_ERR_ARG, "null argument");
  *out = nullptr;
  if (p->nx < 1 || p->ny < 1 || p->nz < 1)
    return fail(TBEC_ERR_ARG, "grid must be positive (got %d x %d x %d)", p->nx, p->ny, p->nz);
  if (p->prec != TBEC_F32 && p->prec != TBEC_F64)
    return fail(TBEC_ERR_ARG, "prec must be 32 or 64 (got %d)", p->prec);
  if (p->sum_mode < TBEC_SUM_REFERENCE || p->sum_mode > TBEC_SUM_CENTRED)
    return fail(TBEC_ERR_ARG, "unknown sum_mode %d", p->sum_mode);
  if (p->prec == TBEC_F64 && p->sum_mode != TBEC_SUM_REFERENCE)
    return fail(TBEC_ERR_ARG, "FP64 runs use TBEC_SUM_REFERENCE");
  if (!(p->Lx > 0 && p->Ly > 0 && p->Lz > 0))
    return fail(TBEC_ERR_ARG, "box lengths must be positive");
  if (p->tile < -1 || p->tile >= kNumTiles)
    return fail(TBEC_ERR_ARG, "tile id %d out of range [-1, %d)", p->tile, kNumTiles);
  Layout L;
  compute_layout(*p, L);
  if (L.sx * L.csize > 0x7fffffffLL)     // one x-plane is one buffer record range
    return fail(TBEC_ERR_ARG, "x-plane of %lld bytes exceeds 2^31", L.sx * L.csize);
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  tbec_handle *h = new tbec_handle();
  h->p = *p;
  h->device = dev;
  h->tile = p->tile < 0 ? 2 : p->tile;
  h->seg = auto_seg(tiles_per_plane(*p, h->tile), p->nx, kAutoSegBlocks);
  h->csize = L.csize;
  h->hx = L.hx;
  h->hy = L.hy;
  h->hz = L.hz;
  // dt*(1 - i*gamma) for real time (src/run.f90:42)
  h->dt_re = p->dt_re + p->gamma * p->dt_im;
  h->dt_im = p->dt_im - p->gamma * p->dt_re;
  h->PZ = L.PZ; h->PY = L.PY; h->PX = L.PX;
  h->sy = L.sy; h->sx = L.sx; h->base = L.base; h->elems = L.elems;
  const size_t bytes = (size_t)h->elems * h->csize;
  hipError_t e1 = hipMalloc(&h->psi[0], bytes);
  hipError_t e2 = e1 == hipSuccess ? hipMalloc(&h->psi[1], bytes) : e1;
  hipError_t e3 = e2 == hipSuccess ? hipMalloc(&h->w, bytes) : e2;
  if (e3 != hipSuccess) {
    tbec_destroy(h);
    (void)hipGetLastError();
    return fail(TBEC_ERR_OOM, "device allocation of 3 x %.3f GB failed: %s", bytes / 1e9,
                hipGetErrorString(e3));
  }
  HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIPCHK(hipMemsetAsync(h->psi[0], 0, bytes, h->stream));
  HIPCHK(hipMemsetAsync(h->psi[1], 0, bytes, h->stream));
  HIPCHK(hipMemsetAsync(h->w, 0, bytes, h->stream));
  HIPCHK(hipMalloc(&h->dstate, sizeof(double) * ST_COUNT));
  HIPCHK(hipMalloc(&h->dflag, sizeof(int) * 4));
  HIPCHK(hipHostMalloc(&h->hstate, sizeof(double) * ST_COUNT));
  HIPCHK(hipHostMalloc(&h->hflag, sizeof(int) * 4));
  h->norm_target = 1.0;
  int rc = tbec_reset_state(h);
  if (rc) { tbec_destroy(h); return rc; }
  HIPCHK(hipStreamSynchronize(h->stream));
  *out = h;
  return 0;
}

int tbec_destroy(tbec_handle *h) {
  if (!h) return 0;
  tbec_drop_graphs(h);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_physics(h);
  if (h->psi[0]) (void)hipFree(h->psi[0]);
  if (h->psi[1]) (void)hipFree(h->psi[1]);
  if (h->w) (void)hipFree(h->w);
  if (h->stage) (void)hipFree(h->stage);
  if (h->partials) (void)hipFree(h->partials);
  if (h->dstate) (void)hipFree(h->dstate);
  if (h->dflag) (void)hipFree(h->dflag);
  if (h->hstate) (void)hipHostFree(h->hstate);
  if (h->hflag) (void)hipHostFree(h->hflag);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int tbec_set_residual_gate(tbec_handle *h, double rmax) {
  if (!h) return fail(TBEC_ERR_ARG, "null handle");
  if (!(rmax >= 0.0) || !std::isfinite(rmax))
    return fail(TBEC_ERR_ARG, "residual gate must be finite and >= 0");
  h->rgate = rmax;
  return 0;
}

int tbec_set_norm(tbec_handle *h, double norm) {
  if (!h || !(norm > 0)) return fail(TBEC_ERR_ARG, "norm must be positive");
  h->norm_target = norm;
  return tbec_reset_state(h);
}

static int ensure_stage_buf(tbec_handle *h) {
  if (!h->stage) {
    const size_t n = (size_t)h->p.nx * h->p.ny * h->p.nz * h->csize;
    hipError_t e = hipMalloc(&h->stage, n);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(TBEC_ERR_OOM, "staging buffer allocation failed: %s", hipGetErrorString(e));
    }
  }
  return 0;
}

static int unpack_from(tbec_handle *h, const void *src_dev) {
  const long long n = (long long)h->p.nx * h->p.ny * h->p.nz;
  if (h->p.prec == TBEC_F32)
    hipLaunchKernelGGL(k_unpack<float>, dim3(grid_for(n)), dim3(256), 0, h->stream,
                       (const cplx<float> *)src_dev, (cplx<float> *)h->psi[h->cur], h->p.nx,
                       h->p.ny, h->p.nz, h->sy, h->sx, h->base);
  else
    hipLaunchKernelGGL(k_unpack<double>, dim3(grid_for(n)), dim3(256), 0, h->stream,
                       (const cplx<double> *)src_dev, (cplx<double> *)h->psi[h->cur], h->p.nx,
                       h->p.ny, h->p.nz, h->sy, h->sx, h->base);
  HIPCHK(hipGetLastError());
  return tbec_reset_state(h);
}

static int pack_to(tbec_handle *h, const void *src_padded, void *dst_dev, const double *scale) {
  const long long n = (long long)h->p.nx * h->p.ny * h->p.nz;
  if (h->p.prec == TBEC_F32)
    hipLaunchKernelGGL(k_pack<float>, dim3(grid_for(n)), dim3(256), 0, h->stream,
                       (const cplx<float> *)src_padded, (cplx<float> *)dst_dev, h->p.nx, h->p.ny,
                       h->p.nz, h->sy, h->sx, h->base, scale);
  else
    hipLaunchKernelGGL(k_pack<double>, dim3(grid_for(n)), dim3(256), 0, h->stream,
                       (const cplx<double> *)src_padded, (cplx<double> *)dst_dev, h->p.nx,
                       h->p.ny, h->p.nz, h->sy, h->sx, h->base, scale);
  HIPCHK(hipGetLastError());
  return 0;
}

int tbec_upload_psi(tbec_handle *h, const void *psi_host) {
  if (!h || !psi_host) return fail(TBEC_ERR_ARG, "null argument");
  int rc = ensure_stage_buf(h);
  if (rc) return rc;
  const size_t n = (size_t)h->p.nx * h->p.ny * h->p.nz * h->csize;
  HIPCHK(hipMemcpyAsync(h->stage, psi_host, n, hipMemcpyHostToDevice, h->stream));
  rc = unpack_from(h, h->stage);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int tbec_upload_psi_device(tbec_handle *h, const void *psi_dev) {
  if (!h || !psi_dev) return fail(TBEC_ERR_ARG, "null argument");
  return unpack_from(h, psi_dev);
}

int tbec_download_psi(tbec_handle *h, void *psi_host) {
  if (!h || !psi_host) return fail(TBEC_ERR_ARG, "null argument");
  int rc = ensure_stage_buf(h);
  if (rc) return rc;
  rc = pack_to(h, h->psi[h->cur], h->stage, h->dstate + ST_SCALE);
  if (rc) return rc;
  const size_t n = (size_t)h->p.nx * h->p.ny * h->p.nz * h->csize;
  HIPCHK(hipMemcpyAsync(psi_host, h->stage, n, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int tbec_download_psi_device(tbec_handle *h, void *psi_dev) {
  if (!h || !psi_dev) return fail(TBEC_ERR_ARG, "null argument");
  return pack_to(h, h->psi[h->cur], psi_dev, h->dstate + ST_SCALE);
}

int tbec_rhs(tbec_handle *h, void *rhs_host) {
  if (!h || !rhs_host) return fail(TBEC_ERR_ARG, "null argument");
  int rc = tbec_materialize_scale(h);
  if (!rc) rc = prepare_physics(h, h->t_now);
  if (rc) return rc;
  rc = enqueue_stage(h, ST_RHS, false, false);
  if (rc) return rc;
  rc = ensure_stage_buf(h);
  if (rc) return rc;
  rc = pack_to(h, h->w, h->stage, nullptr);
  if (rc) return rc;
  const size_t n = (size_t)h->p.nx * h->p.ny * h->p.nz * h->csize;
  HIPCHK(hipMemcpyAsync(rhs_host, h->stage, n, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int tbec_diagnose(tbec_handle *h, tbec_diag *out) {
  if (!h || !out) return fail(TBEC_ERR_ARG, "null argument");
  return diagnose_with(h, out, nullptr, 0);
}

int tbec_set_tuning(tbec_handle *h, int tile, int seg) {
  if (!h) return fail(TBEC_ERR_ARG, "null handle");
  if (tile < 0 || tile >= kNumTiles) return fail(TBEC_ERR_ARG, "tile id %d out of range", tile);
  if (seg < 0 || seg > 4096) return fail(TBEC_ERR_ARG, "segment length %d out of range", seg);
  HIPCHK(hipStreamSynchronize(h->stream));
  tbec_drop_graphs(h);
  h->tile = tile;
  h->seg = seg > 0 ? seg : auto_seg(tiles_per_plane(h->p, tile), h->p.nx, kAutoSegBlocks);
  return 0;
}

int tbec_set_first_stage_tuning(tbec_handle *h, int tile, int seg) {
  if (!h) return fail(TBEC_ERR_ARG, "null handle");
  if (tile < -1 || tile >= kNumTiles) return fail(TBEC_ERR_ARG, "tile id %d out of range", tile);
  if (seg < 0 || seg > 4096) return fail(TBEC_ERR_ARG, "segment length %d out of range", seg);
  HIPCHK(hipStreamSynchronize(h->stream));
  tbec_drop_graphs(h);
  h->tile1 = tile;
  h->seg1 = tile < 0 ? 0
                     : (seg > 0 ? seg : auto_seg(tiles_per_plane(h->p, tile), h->p.nx,
                                                  kAutoSegBlocks));
  return 0;
}

int tbec_set_boundaries(tbec_handle *h, int boundaries, int periodicz) {
  if (!h) return fail(TBEC_ERR_ARG, "null handle");
  if (boundaries < TBEC_BC_ZERO || boundaries > TBEC_BC_PERIODIC)
    return fail(TBEC_ERR_ARG, "boundaries must be TBEC_BC_ZERO, _SOLID or _PERIODIC");
  if (periodicz != 0 && periodicz != 1) return fail(TBEC_ERR_ARG, "periodicz must be 0 or 1");
  HIPCHK(hipStreamSynchronize(h->stream));
  tbec_drop_graphs(h);
  h->bc = boundaries;
  h->periodicz = periodicz;
  // zero faces leave the permanent zero halo the zero boundary relies on
  for (int q = 0; q < 2; ++q) HIPCHK(fill_ghosts(h, h->psi[q], TBEC_BC_ZERO, 0));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int tbec_set_w_storage(tbec_handle *h, int stored) {
  if (!h) return fail(TBEC_ERR_ARG, "null handle");
  if (stored != 0 && stored != 1) return fail(TBEC_ERR_ARG, "stored must be 0 or 1");
  HIPCHK(hipStreamSynchronize(h->stream));
  tbec_drop_graphs(h);
  h->w_stored = stored;
  return 0;
}

int tbec_set_xcd_remap(tbec_handle *h, int on) {
  if (!h) return fail(TBEC_ERR_ARG, "null handle");
  if (on < 0 || on > 2) return fail(TBEC_ERR_ARG, "xcd remap must be 0, 1 or 2");
  HIPCHK(hipStreamSynchronize(h->stream));
  tbec_drop_graphs(h);
  h->remap = on;
  return 0;
}

int tbec_sync(tbec_handle *h) {
  if (!h) return fail(TBEC_ERR_ARG, "null handle");
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

void *tbec_stream(tbec_handle *h) { return h ? (void *)h->stream : nullptr; }

}  // extern "C"

// ------------------------------------------------------ state helpers -----------
int fill_args_count(tbec_handle *h) {
  StageArgs a;
  return fill_args(h, a, ST_RHS, h->cur, 0, 0, 0, 0, h->tile, h->seg);
}

namespace tbec {

long long diagnose_partial_slots(tbec_handle *h) {
  const long long n = (long long)h->p.nx * h->p.ny * h->p.nz;
  return std::max<long long>(fill_args_count(h), grid_for(n));
}

int diagnose_with(tbec_handle *h, tbec_diag *out, double *part, long long part_slots) {
  // rhs-only pass with reductions (psi read with the pending scale), then Psi_info
  const long long n = (long long)h->p.nx * h->p.ny * h->p.nz;
  const int nbi = grid_for(n);
  if (!part) {
    int rc = tbec_reserve_partials(h, (int)diagnose_partial_slots(h));
    if (rc) return rc;
    part = h->partials;
  } else if (diagnose_partial_slots(h) > part_slots) {
    return fail(TBEC_ERR_ARG, "diagnostics need %lld reduction slots, the buffer holds %lld",
                diagnose_partial_slots(h), part_slots);
  }
  int rc = prepare_physics(h, h->t_now);
  if (!rc) rc = enqueue_stage(h, ST_RHS, false, true, part);
  if (rc) return rc;
  const int nb_rhs = fill_args_count(h);
  hipLaunchKernelGGL(k_fin_energy, dim3(1), dim3(256), 0, h->stream, part, nb_rhs,
                     h->dstate, h->dflag, 0);
  HIPCHK(hipGetLastError());
  // second pass for the GP residual, with c0 = 2*mu of the first
  hipLaunchKernelGGL(k_set_resid_c0, dim3(1), dim3(64), 0, h->stream, h->dstate);
  h->resid_pass = true;
  rc = enqueue_stage(h, ST_RHS, false, true, part);
  h->resid_pass = false;
  if (rc) return rc;
  hipLaunchKernelGGL(k_fin_resid, dim3(1), dim3(256), 0, h->stream, part, nb_rhs,
                     h->dstate);
  HIPCHK(hipGetLastError());
  if (h->p.prec == TBEC_F32)
    hipLaunchKernelGGL(k_info<float>, dim3(nbi), dim3(256), 0, h->stream,
                       (const cplx<float> *)h->psi[h->cur], h->p.nx, h->p.ny, h->p.nz, h->sy,
                       h->sx, h->base, h->dstate + ST_SCALE, part);
  else
    hipLaunchKernelGGL(k_info<double>, dim3(nbi), dim3(256), 0, h->stream,
                       (const cplx<double> *)h->psi[h->cur], h->p.nx, h->p.ny, h->p.nz, h->sy,
                       h->sx, h->base, h->dstate + ST_SCALE, part);
  hipLaunchKernelGGL(k_fin_info, dim3(1), dim3(256), 0, h->stream, part, nbi, h->dstate);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(h->hstate, h->dstate, sizeof(double) * ST_COUNT,
                        hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  const double *s = h->hstate;
  out->norm = s[ST_N];
  out->energy = s[ST_E];
  out->mu = s[ST_MU];
  out->residual = s[ST_RES];
  out->max_abs = s[ST_MAXA];
  out->min_abs = s[ST_MINA];
  out->mean_abs = s[ST_SUMA] / ((double)h->p.nx * h->p.ny * h->p.nz);
  out->int_psi2 = s[ST_SUMA2];
  if (!std::isfinite(out->energy) || !std::isfinite(out->norm))
    return fail(TBEC_ERR_NONFINITE, "non-finite norm or energy in diagnostics");
  return 0;
}

}  // namespace tbec

int tbec_reset_state(tbec_handle *h) {
  double *s = h->hstate;
  for (int i = 0; i < ST_COUNT; ++i) s[i] = 0.0;
  s[ST_SCALE] = 1.0;
  // c0 as the kernels of this precision read it
  s[ST_C0EFF] = s[ST_C0NEXT] = s[ST_C0BASE] =
      h->p.prec == TBEC_F32 ? (double)(float)h->p.c0 : h->p.c0;
  s[ST_GNL] = h->p.gnl;
  s[ST_DV] = h->hx * h->hy * h->hz;
  s[ST_NORMT] = h->norm_target;
  s[ST_TOL] = h->tol;
  s[ST_RGATE] = h->rgate;
  s[ST_F32] = h->p.prec == TBEC_F32 ? 1.0 : 0.0;
  HIPCHK(hipMemcpyAsync(h->dstate, s, sizeof(double) * ST_COUNT, hipMemcpyHostToDevice,
                        h->stream));
  for (int i = 0; i < 4; ++i) h->hflag[i] = 0;
  HIPCHK(hipMemcpyAsync(h->dflag, h->hflag, sizeof(int) * 4, hipMemcpyHostToDevice,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->scale_pending = false;
  return 0;
}

// Fold the pending renormalisation scale into the stored psi (scale*psi rounded to the
// working precision, exactly what the next stage-1 read would compute), scale = 1.
int tbec_materialize_scale(tbec_handle *h) {
  if (!h->scale_pending) return 0;
  const long long n = (long long)h->p.nx * h->p.ny * h->p.nz;
  if (h->p.prec == TBEC_F32)
    hipLaunchKernelGGL(k_apply_scale<float>, dim3(grid_for(n)), dim3(256), 0, h->stream,
                       (cplx<float> *)h->psi[h->cur], h->p.nx, h->p.ny, h->p.nz, h->sy, h->sx,
                       h->base, (const double *)(h->dstate + ST_SCALE));
  else
    hipLaunchKernelGGL(k_apply_scale<double>, dim3(grid_for(n)), dim3(256), 0, h->stream,
                       (cplx<double> *)h->psi[h->cur], h->p.nx, h->p.ny, h->p.nz, h->sy,
                       h->sx, h->base, (const double *)(h->dstate + ST_SCALE));
  hipLaunchKernelGGL(k_reset_scale, dim3(1), dim3(1), 0, h->stream, h->dstate);
  HIPCHK(hipGetLastError());
  h->scale_pending = false;
  return 0;
}
