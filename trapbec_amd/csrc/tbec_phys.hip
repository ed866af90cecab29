// Physics beyond the harmonic trap: the reference's potential kinds, laser beam,
// switch-off times and advective term (get_rhs, src/finite_diff.f90:30-130), the mobius
// Trap table (src/initial.f90:160-196) and the vortex imprint of the forcing
// (force_random_points, src/force.f90:32-53).
//
// Design.  The analytic harmonic trap stays inside the stage kernel (PH_HARM /
// PH_ROT).  Every other kind is evaluated by k_potential into a tabulated potential V
// (one real per point, in the padded layout of psi so the stage kernel reads it at the
// same offsets) and the stage kernel's PH_GEN variant reads it: 4 (FP32) or 8 (FP64)
// more bytes per point and stage, one extra HBM stream next to the 8 complex transfers.
// The table is rebuilt only when the potential changes: at a switch-off time, while a
// ring or barrier opening max(0, 5 - 0.1 t) is still closing, and every step for the
// box trap's 5 sin(0.1 t) tilt.  The stepper groups consecutive steps with the same
// stage configuration and table and replays captured graphs within a group.
//
// Arithmetic.  V follows the reference's expressions operand for operand with
// contraction off; integer powers are multiplied out left to right as amdflang emits
// them (r**8 = r*r*r*r*r*r*r*r); sqrt and division are correctly rounded on both sides,
// so kinds built from them (harmonic, rings, barrier, box trap, channel, quartic, tables)
// give the reference's V bit for bit.  The scalars that depend on t alone (the opening,
// the tilt, the laser centre) are formed on the host in the working precision with the
// C library, as the reference forms them; pow, atan2, acos, sin, cos and exp evaluated
// per point (ring2, sphere, laser, mobius) come from the device math library and can
// differ from the host's in the last place.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gpe_stage.hpp"
#include "tbec_internal.hpp"

using namespace tbec;

namespace {

struct PotArgs {
  int kind, lbeam, pot_on;
  int nx, ny, nz;
  long long sy, sx, base;        // padded layout of the output table (elements)
  double Lx, Ly, Lz, wx, wy, wz; // converted to T in the kernel
  double s, b, las_x, las_y;     // T values formed on the host
  double wbeam;
  const void *trap;              // interior layout, T
  const void *coords;            // xx(1:nx), yy(1:ny), zz(1:nz) in T (src/initial.f90:34-42)
};

// Cell-centred coordinate of 0-based index i on an axis of n points and length L
// (src/initial.f90:34-42, one domain): ((2i+1) / (2n)) L - L/2 in the working precision.
template <typename T> __device__ __forceinline__ T coord(int i, int n, T L) {
#pragma clang fp contract(off)
  return ((T)(2 * i + 1) / ((T)2 * (T)n)) * L - L / (T)2;
}

template <typename T> __device__ T pot_point(const PotArgs &a, int i, int j, int k) {
#pragma clang fp contract(off)
  const T Lx = (T)a.Lx, Ly = (T)a.Ly, Lz = (T)a.Lz;
  const T *cc = (const T *)a.coords;
  const T x = cc[i], y = cc[a.nx + j], z = cc[a.nx + a.ny + k];
  const T s = (T)a.s;
  T pot = (T)0;
  switch (a.kind) {
    case TBEC_POT_ZERO:
      break;
    case TBEC_POT_HARMONIC: {
      const T ax = x / Lx, ay = y / Ly, az = z / Lz;
      pot = (T)a.wx * (ax * ax) + (T)a.wy * (ay * ay) + (T)a.wz * (az * az);
      break;
    }
    case TBEC_POT_HARMONIC_RING: {
      const T ax = x / Lx, ay = y / Lx;
      const T r = sqrt(ax * ax + ay * ay);
      const T d = r - (T)0.3;
      pot = (T)a.wx * (d * d);
      if (fabs(r - (T)0.3) < (T)0.01) pot = s;
      break;
    }
    case TBEC_POT_HARMONIC_RING2: {
      const T pw = (T)6;
      const T a1 = pow((T)0.25, pw);
      const T u1 = pow(x / Lx, pw);
      const T v1 = pow(y / Ly - (T)0.2, pw), v2 = pow(y / Ly + (T)0.2, pw);
      T r1 = pow(fabs((a1 - (T)0.15 * u1) - (T)4 * v1), (T)1 / pw);
      T r2 = pow(fabs((a1 - (T)0.15 * u1) - (T)4 * v2), (T)1 / pw);
      r1 = r1 < (T)0.2495 ? (T)0 : (T)5;
      r2 = r2 < (T)0.2495 ? (T)0 : (T)5;
      pot = r1 * r2;
      if (pot < (T)0) pot = (T)0;
      if (fabs(y / Ly) < (T)0.01 && pot < (T)0.01) pot = s;
      break;
    }
    case TBEC_POT_BARRIER:
      if (fabs(y / Ly) > (T)0.3) pot = (T)5;
      if (fabs(y / Ly) < (T)0.01) pot = s;
      break;
    case TBEC_POT_BOX_TRAP: {
      pot = (T)10;
      const T r = sqrt(x * x + y * y);
      if (r < Lx / (T)3 && fabs(z) < Lz / (T)2.5) pot = (T)0;
      pot = pot + ((T)a.b * z) / Lz;
      break;
    }
    case TBEC_POT_CHANNEL: {
      if (fabs(y) > ((T)0.75 * Ly) / (T)2 || fabs(z) > ((T)0.75 * Lz) / (T)2) pot = (T)5;
      const T hxg = Lx / (T)a.nx, hyg = Ly / (T)a.ny;
      if (fabs(x) < hxg) {
        for (int l1 = -1; l1 <= 1; ++l1) {
          if (fabs(y - ((T)l1 * Ly) / (T)5) < hyg) pot = (T)7.5;
        }
        for (int l1 = -1; l1 <= 1; ++l1) {
          if (fabs(z - ((T)l1 * Lz) / (T)5) < hyg) pot = (T)7.5;
        }
      }
      break;
    }
    case TBEC_POT_QUARTIC: {
      const T ax = x / Lx, ay = y / Ly;
      const T r = sqrt(ax * ax + ay * ay);
      pot = (T)a.wx * (r * r * r * r * r * r * r * r);
      break;
    }
    case TBEC_POT_SPHERE: {
      const T theta = atan2(y, x);
      const T ax = x / Lx, ay = y / Ly, az = z / Lz;
      const T r1 = sqrt((ax * ax + ay * ay) + az * az);
      const T phi = acos(az / r1);
      const T lim = ((T)0.2 + (T)0.01 * sin((T)20 * phi)) + (T)0.01 * cos((T)30 * theta + (T)0.03);
      pot = r1 < lim ? (T)10 : (T)0;
      break;
    }
    case TBEC_POT_MOBIUS:
    case TBEC_POT_TABLE: {
      pot = ((const T *)a.trap)[((long long)i * a.ny + j) * a.nz + k];
      if (a.kind == TBEC_POT_MOBIUS && fabs(y) < (T)1 && fabs(x - (T)25) < (T)1)
        pot = pot + (T)10;
      break;
    }
    default:
      break;
  }
  if (a.lbeam) {
    const T wb = (T)a.wbeam;
    const T u = (x - (T)a.las_x) / wb, v = (y - (T)a.las_y) / wb;
    pot = pot + (T)20 * exp(-(u * u) - v * v);
  }
  if (!a.pot_on) pot = (T)0;
  return pot;
}

// One x-row of z points per workgroup row: grid (ceil(nz/256), ny, nx), so the point's
// indices come from the launch coordinates (no 64-bit division per point).
template <typename T> __global__ void __launch_bounds__(256) k_potential(const PotArgs a, T *v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, i = blockIdx.z;
  if (k >= a.nz) return;
  v[a.base + i * a.sx + j * a.sy + k] = pot_point<T>(a, i, j, k);
}

// Pack a padded real table to interior order.
template <typename T>
__global__ void k_pack_real(const T *__restrict__ src, T *__restrict__ dst, int nx, int ny,
                            int nz, long long sy, long long sx, long long base) {
  const long long n = (long long)nx * ny * nz;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const long long k = q % nz, r = q / nz, j = r % ny, i = r / ny;
    dst[q] = src[base + i * sx + j * sy + k];
  }
}

// Mobius Trap, pass 1: for each interior point the sum over the 200 x 200 strip points
// of exp(-1.5 |x - m|^2) in (ii1, ii2) order, capped at 3 after every term
// (src/initial.f90:173-185).  The strip is staged through LDS in chunks; one thread per
// point keeps the reference's sequential order.  Block maxima go to part.
constexpr int kStrip = 200 * 200;
constexpr int kStripChunk = 512;

template <typename T>
__global__ void __launch_bounds__(256)
k_mobius_sum(const T *__restrict__ strip, int nx, int ny, int nz, double Lx, double Ly,
             double Lz, T *__restrict__ trap, double *part) {
#pragma clang fp contract(off)
  __shared__ T sm[kStripChunk * 3];
  __shared__ double bmax[256];
  const long long n = (long long)nx * ny * nz;
  const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const bool live = q < n;
  T x = 0, y = 0, z = 0;
  if (live) {
    const int k = (int)(q % nz);
    const long long r = q / nz;
    const int j = (int)(r % ny), i = (int)(r / ny);
    x = coord<T>(i, nx, (T)Lx);
    y = coord<T>(j, ny, (T)Ly);
    z = coord<T>(k, nz, (T)Lz);
  }
  T acc = (T)0;
  for (int c0 = 0; c0 < kStrip; c0 += kStripChunk) {
    const int m = min(kStripChunk, kStrip - c0);
    __syncthreads();
    for (int e = threadIdx.x; e < 3 * m; e += blockDim.x) sm[e] = strip[3 * (long long)c0 + e];
    __syncthreads();
    if (live) {
      for (int e = 0; e < m; ++e) {
        const T dx = x - sm[3 * e], dy = y - sm[3 * e + 1], dz = z - sm[3 * e + 2];
        acc = acc + exp(-((T)1.5 * (dx * dx)) - (T)1.5 * (dy * dy) - (T)1.5 * (dz * dz));
        if (acc > (T)3) acc = (T)3;
      }
    }
  }
  if (live) trap[q] = acc;
  bmax[threadIdx.x] = live ? (double)acc : -HUGE_VAL;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) bmax[threadIdx.x] = fmax(bmax[threadIdx.x], bmax[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = bmax[0];
}

// Pass 2: the largest block maximum (one workgroup), into part[nparts].
__global__ void __launch_bounds__(256) k_max_parts(double *part, int nparts) {
  __shared__ double red[256];
  double m = -HUGE_VAL;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) m = fmax(m, part[b]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[nparts] = red[0];
}

// Pass 3: Trap = max(Trap) - Trap (src/initial.f90:187-196).
template <typename T>
__global__ void __launch_bounds__(256)
k_mobius_finish(T *__restrict__ trap, long long n, const double *mx_) {
  const T mx = (T)*mx_;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x)
    trap[q] = mx - trap[q];
}

// force_random_points' imprint (src/force.f90:32-43): psi *= pade(r) exp(i dir theta)
// about (x0, y0); pade's constants c1, c2, c3 come from the host in T.
template <typename T>
__global__ void __launch_bounds__(256)
k_imprint(cplx<T> *__restrict__ psi, int nx, int ny, int nz, long long sy, long long sx,
          long long base, double Lx, double Ly, double x0, double y0, double dir, double c1,
          double c2, double c3) {
#pragma clang fp contract(off)
  const long long n = (long long)nx * ny * nz;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(q % nz);
    const long long r0 = q / nz;
    const int j = (int)(r0 % ny), i = (int)(r0 / ny);
    const T ddx = coord<T>(i, nx, (T)Lx) - (T)x0, ddy = coord<T>(j, ny, (T)Ly) - (T)y0;
    const T r = sqrt(ddx * ddx + ddy * ddy);
    const T theta = atan2(ddy, ddx);
    const T r2 = r * r, r4 = r * r * r * r;
    const T func = sqrt((r2 * ((T)c1 + (T)c2 * r2)) / (((T)1 + (T)c3 * r2) + (T)c2 * r4));
    const T ang = (T)dir * theta;
    const T fr = func * cos(ang), fi = func * sin(ang);
    cplx<T> &c = psi[base + i * sx + j * sy + k];
    const T a = c.re, b = c.im;
    c.re = a * fr - b * fi;
    c.im = a * fi + b * fr;
  }
}

int blocks_for(long long n) {
  return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 16384));
}

bool is_f32(const tbec_handle *h) { return h->p.prec == TBEC_F32; }

// T-precision helpers on the host: the working precision of the reference's `real`.
double to_t(const tbec_handle *h, double v) { return is_f32(h) ? (double)(float)v : v; }

double eps_of(const tbec_handle *h) {
  return is_f32(h) ? 1.1920928955078125e-07 : 2.220446049250313e-16;   // epsilon(0.)
}

bool analytic_kind(const tbec_handle *h) {
  const int k = h->phys.potential;
  return (k == TBEC_POT_HARMONIC || k == TBEC_POT_ZERO) && !h->phys.lbeam &&
         !(std::fabs(to_t(h, h->phys.advective_vx)) > eps_of(h));
}

bool kind_moves(int k) {   // the kind's own table depends on t
  return k == TBEC_POT_HARMONIC_RING || k == TBEC_POT_HARMONIC_RING2 ||
         k == TBEC_POT_BARRIER || k == TBEC_POT_BOX_TRAP;
}

// Everything of a step at time t that decides its stage launches and its table.
struct StepPhys {
  int mode, drift;
  bool wl_zero, pot_on;
  double key[5];   // table key: pot_on, s, b, las_x, las_y (tabulated kinds)
};

template <typename T> void time_scalars(T t, double out[4]) {
  // src/finite_diff.f90:39,63,69 (max(0., 5.-0.1*t)), :77 (5*sin(0.1*t)), :119-121
  const T m = (T)5 - (T)0.1 * t;
  out[0] = (double)((T)0 > m ? (T)0 : m);
  out[1] = (double)((T)5 * (T)std::sin((T)0.1 * t));
  const T lf = (T)0;
  out[2] = (double)(((T)70 * (T)std::cos(lf * t)) * ((T)1 - (T)std::sin(lf * t)));
  out[3] = (double)(((T)30 * (T)std::sin(lf * t)) * (T)std::cos(lf * t));
}

void time_scalars_of(const tbec_handle *h, double t, double out[4]) {
  if (is_f32(h)) {
    float ts = (float)t;
    float s = (float)5 - (float)0.1 * ts;
    out[0] = (double)((float)0 > s ? (float)0 : s);
    out[1] = (double)((float)5 * sinf((float)0.1 * ts));
    const float lf = 0.0f;
    out[2] = (double)(((float)70 * cosf(lf * ts)) * ((float)1 - sinf(lf * ts)));
    out[3] = (double)(((float)30 * sinf(lf * ts)) * cosf(lf * ts));
  } else {
    time_scalars<double>(t, out);
  }
}

StepPhys step_phys(const tbec_handle *h, double t) {
  StepPhys r{};
  const double tt = to_t(h, t);
  r.pot_on = !(tt > to_t(h, h->phys.potential_off));
  const bool adv = std::fabs(to_t(h, h->phys.advective_vx)) > eps_of(h);
  const bool rot = !adv && rot_enabled(h->p) && tt < to_t(h, h->phys.t_spindown);
  if (analytic_kind(h)) {
    r.mode = rot ? PH_ROT : PH_HARM;
    r.drift = 0;
    r.wl_zero = h->phys.potential == TBEC_POT_ZERO || !r.pot_on;
    for (double &k : r.key) k = 0.0;
    return r;
  }
  r.mode = PH_GEN;
  r.drift = adv ? 2 : (rot ? 1 : 0);
  r.wl_zero = false;
  double sc[4];
  time_scalars_of(h, tt, sc);
  const int k = h->phys.potential;
  r.key[0] = r.pot_on ? 1.0 : 0.0;
  const bool opening = k == TBEC_POT_HARMONIC_RING || k == TBEC_POT_HARMONIC_RING2 ||
                       k == TBEC_POT_BARRIER;
  r.key[1] = r.pot_on && opening ? sc[0] : 0.0;
  r.key[2] = r.pot_on && k == TBEC_POT_BOX_TRAP ? sc[1] : 0.0;
  r.key[3] = r.pot_on && h->phys.lbeam ? sc[2] : 0.0;
  r.key[4] = r.pot_on && h->phys.lbeam ? sc[3] : 0.0;
  return r;
}

bool same_key(const double a[5], const double b[5]) {
  for (int i = 0; i < 5; ++i)
    if (std::memcmp(&a[i], &b[i], sizeof(double)) != 0) return false;
  return true;
}

PotArgs pot_args(const tbec_handle *h, double t, bool pot_on) {
  PotArgs a{};
  a.kind = h->phys.potential;
  a.lbeam = h->phys.lbeam;
  a.pot_on = pot_on ? 1 : 0;
  a.nx = h->p.nx; a.ny = h->p.ny; a.nz = h->p.nz;
  a.sy = h->sy; a.sx = h->sx; a.base = h->base;
  a.Lx = h->p.Lx; a.Ly = h->p.Ly; a.Lz = h->p.Lz;
  a.wx = h->p.harm_wx; a.wy = h->p.harm_wy; a.wz = h->p.harm_wz;
  double sc[4];
  time_scalars_of(h, to_t(h, t), sc);
  a.s = sc[0]; a.b = sc[1]; a.las_x = sc[2]; a.las_y = sc[3];
  a.wbeam = h->phys.wbeam;
  a.trap = h->trap;
  return a;
}

// The cell-centre coordinates in the working precision, formed on the host with the
// same IEEE operations the reference uses and kept on the device (one table per handle).
template <typename T> std::vector<T> coord_table(const tbec_handle *h) {
  std::vector<T> c;
  const int n[3] = {h->p.nx, h->p.ny, h->p.nz};
  const T L[3] = {(T)h->p.Lx, (T)h->p.Ly, (T)h->p.Lz};
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < n[a]; ++i)
      c.push_back(((T)(2 * i + 1) / ((T)2 * (T)n[a])) * L[a] - L[a] / (T)2);
  return c;
}

int ensure_coords(tbec_handle *h) {
  if (h->coords) return 0;
  const size_t cnt = (size_t)h->p.nx + h->p.ny + h->p.nz;
  const size_t bytes = cnt * (h->csize / 2);
  hipError_t e = hipMalloc(&h->coords, bytes);
  if (e == hipSuccess) {
    if (is_f32(h)) {
      const std::vector<float> c = coord_table<float>(h);
      e = hipMemcpy(h->coords, c.data(), bytes, hipMemcpyHostToDevice);
    } else {
      const std::vector<double> c = coord_table<double>(h);
      e = hipMemcpy(h->coords, c.data(), bytes, hipMemcpyHostToDevice);
    }
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    if (h->coords) (void)hipFree(h->coords);
    h->coords = nullptr;
    return set_error(TBEC_ERR_OOM, std::string("coordinate table: ") + hipGetErrorString(e));
  }
  return 0;
}

// Fill `table` (padded layout, real T) with pot at time t.
int eval_table(tbec_handle *h, double t, bool pot_on, void *table) {
  const int k = h->phys.potential;
  if ((k == TBEC_POT_MOBIUS || k == TBEC_POT_TABLE) && h->trap_owner != k)
    return set_error(TBEC_ERR_STATE, k == TBEC_POT_TABLE
                                         ? "TBEC_POT_TABLE needs tbec_set_potential_table first"
                                         : "the mobius Trap was replaced; select it again");
  if (h->p.nx > 65535 || h->p.ny > 65535)
    return set_error(TBEC_ERR_ARG, "tabulated potentials need nx, ny <= 65535");
  int rc = ensure_coords(h);
  if (rc) return rc;
  PotArgs a = pot_args(h, t, pot_on);
  a.coords = h->coords;
  const dim3 grid((unsigned)((h->p.nz + 255) / 256), (unsigned)h->p.ny, (unsigned)h->p.nx);
  if (is_f32(h))
    hipLaunchKernelGGL(k_potential<float>, grid, dim3(256), 0, h->stream, a, (float *)table);
  else
    hipLaunchKernelGGL(k_potential<double>, grid, dim3(256), 0, h->stream, a,
                       (double *)table);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(TBEC_ERR_HIP, std::string("potential: ") +
                                                          hipGetErrorString(e));
  return 0;
}

int ensure_vtab(tbec_handle *h) {
  if (h->vtab) return 0;
  const size_t bytes = (size_t)h->elems * (h->csize / 2);
  hipError_t e = hipMalloc(&h->vtab, bytes);
  if (e == hipSuccess) e = hipMemsetAsync(h->vtab, 0, bytes, h->stream);   // padding stays 0
  if (e != hipSuccess) {
    (void)hipGetLastError();
    h->vtab = nullptr;
    return set_error(TBEC_ERR_OOM, std::string("potential table allocation failed: ") +
                                       hipGetErrorString(e));
  }
  h->vtab_valid = false;
  return 0;
}

int ensure_trap(tbec_handle *h) {
  if (h->trap) return 0;
  const size_t bytes = (size_t)h->p.nx * h->p.ny * h->p.nz * (h->csize / 2);
  hipError_t e = hipMalloc(&h->trap, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    h->trap = nullptr;
    return set_error(TBEC_ERR_OOM, std::string("Trap table allocation failed: ") +
                                       hipGetErrorString(e));
  }
  return 0;
}

// The mobius strip points (src/initial.f90:161-170) in the working precision, on the
// host with the C library's cos and sin, as the reference forms them.
template <typename T> std::vector<T> mobius_strip() {
  std::vector<T> m((size_t)kStrip * 3);
  const T pi = (T)3.14159265359;   // src/cdata.f90:13
  for (int i = 1; i <= 200; ++i)
    for (int j = 1; j <= 200; ++j) {
      const T r = ((T)(i - 1) * (T)2) / (T)200 - (T)1;
      const T th = (((T)(j - 1) * (T)2) * pi) / (T)200;
      const T ch = (T)std::cos(th / (T)2), sh = (T)std::sin(th / (T)2);
      const T c = (T)std::cos(th), sn = (T)std::sin(th);
      T *e = m.data() + ((size_t)(i - 1) * 200 + (j - 1)) * 3;
      e[0] = ((T)25 + ((T)5 * r) * ch) * c;
      e[1] = ((T)25 + ((T)5 * r) * ch) * sn;
      e[2] = ((T)5 * r) * sh;
    }
  return m;
}

template <> std::vector<float> mobius_strip<float>() {
  std::vector<float> m((size_t)kStrip * 3);
  const float pi = (float)3.14159265359;
  for (int i = 1; i <= 200; ++i)
    for (int j = 1; j <= 200; ++j) {
      const float r = ((float)(i - 1) * 2.0f) / 200.0f - 1.0f;
      const float th = (((float)(j - 1) * 2.0f) * pi) / 200.0f;
      const float ch = cosf(th / 2.0f), sh = sinf(th / 2.0f);
      const float c = cosf(th), sn = sinf(th);
      float *e = m.data() + ((size_t)(i - 1) * 200 + (j - 1)) * 3;
      e[0] = (25.0f + (5.0f * r) * ch) * c;
      e[1] = (25.0f + (5.0f * r) * ch) * sn;
      e[2] = (5.0f * r) * sh;
    }
  return m;
}

template <typename T> int build_mobius(tbec_handle *h) {
  int rc = ensure_trap(h);
  if (rc) return rc;
  h->trap_owner = 0;   // the buffer is rewritten below: no table is valid until it ends
  const std::vector<T> strip = mobius_strip<T>();
  T *dstrip = nullptr;
  hipError_t e = hipMalloc(&dstrip, strip.size() * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(TBEC_ERR_OOM, "mobius strip allocation failed");
  }
  const long long n = (long long)h->p.nx * h->p.ny * h->p.nz;
  const int nb = (int)((n + 255) / 256);
  double *part = nullptr;
  e = hipMalloc(&part, sizeof(double) * ((size_t)nb + 1));
  if (e == hipSuccess)
    e = hipMemcpyAsync(dstrip, strip.data(), strip.size() * sizeof(T), hipMemcpyHostToDevice,
                       h->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mobius_sum<T>, dim3(nb), dim3(256), 0, h->stream, dstrip, h->p.nx,
                       h->p.ny, h->p.nz, h->p.Lx, h->p.Ly, h->p.Lz, (T *)h->trap, part);
    hipLaunchKernelGGL(k_max_parts, dim3(1), dim3(256), 0, h->stream, part, nb);
    hipLaunchKernelGGL(k_mobius_finish<T>, dim3(blocks_for(n)), dim3(256), 0, h->stream,
                       (T *)h->trap, n, (const double *)(part + nb));
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(dstrip);
  if (part) (void)hipFree(part);
  if (e != hipSuccess)
    return set_error(TBEC_ERR_HIP, std::string("mobius Trap: ") + hipGetErrorString(e));
  h->trap_owner = TBEC_POT_MOBIUS;
  return 0;
}

}  // namespace

namespace tbec {

double advance_time(const tbec_handle *h, double t) {
  return is_f32(h) ? (double)((float)t + (float)h->dt_re) : t + h->dt_re;
}

bool physics_time_dependent(const tbec_handle *h) {
  if (std::isfinite(h->phys.potential_off) || std::isfinite(h->phys.t_spindown)) return true;
  return !analytic_kind(h) && kind_moves(h->phys.potential);
}

bool same_step_physics(const tbec_handle *h, double t1, double t2) {
  const StepPhys a = step_phys(h, t1), b = step_phys(h, t2);
  return a.mode == b.mode && a.drift == b.drift && a.wl_zero == b.wl_zero &&
         same_key(a.key, b.key);
}

int prepare_physics(tbec_handle *h, double t) {
  const StepPhys sp = step_phys(h, t);
  const int sig = sp.mode * 8 + sp.drift * 2 + (sp.wl_zero ? 1 : 0);
  if (sig != h->ph_sig) {
    tbec_drop_graphs(h);   // the captured stage arguments no longer apply
    h->ph_sig = sig;
  }
  h->ph_mode = sp.mode;
  h->ph_drift = sp.drift;
  h->ph_wl_zero = sp.wl_zero;
  if (sp.mode != PH_GEN) return 0;
  int rc = ensure_vtab(h);
  if (rc) return rc;
  if (h->vtab_valid && same_key(h->vkey, sp.key)) return 0;
  rc = eval_table(h, t, sp.pot_on, h->vtab);
  if (rc) return rc;
  std::memcpy(h->vkey, sp.key, sizeof h->vkey);
  h->vtab_valid = true;
  return 0;
}

bool physics_would_allocate(const tbec_handle *h, double t) {
  const StepPhys sp = step_phys(h, t);
  if (sp.mode != PH_GEN) return false;
  if (!h->vtab) return true;
  return !(h->vtab_valid && same_key(h->vkey, sp.key)) && !h->coords;
}

void free_physics(tbec_handle *h) {
  if (h->vtab) (void)hipFree(h->vtab);
  if (h->trap) (void)hipFree(h->trap);
  if (h->coords) (void)hipFree(h->coords);
  h->vtab = h->trap = h->coords = nullptr;
}

}  // namespace tbec

extern "C" {

int tbec_set_physics(tbec_handle *h, const tbec_physics *ph) {
  if (!h || !ph) return set_error(TBEC_ERR_ARG, "null argument");
  if (ph->potential < TBEC_POT_ZERO || ph->potential > TBEC_POT_TABLE)
    return set_error(TBEC_ERR_ARG, "unknown potential kind " + std::to_string(ph->potential));
  if (ph->lbeam && !(ph->wbeam != 0.0))
    return set_error(TBEC_ERR_ARG, "the laser beam needs a non-zero width");
  if (std::isnan(ph->potential_off) || std::isnan(ph->t_spindown) ||
      !std::isfinite(ph->advective_vx))
    return set_error(TBEC_ERR_ARG, "switch-off times must not be NaN; vx must be finite");
  int rc = tbec_materialize_scale(h);
  if (rc) return rc;
  const tbec_physics before = h->phys;
  h->phys = *ph;
  h->vtab_valid = false;
  h->ph_sig = -1;
  tbec_drop_graphs(h);
  if (ph->potential == TBEC_POT_MOBIUS && h->trap_owner != TBEC_POT_MOBIUS) {
    rc = is_f32(h) ? build_mobius<float>(h) : build_mobius<double>(h);
    if (rc) h->phys = before;   // the handle keeps its previous physics
    return rc;
  }
  return 0;
}

int tbec_set_potential_table(tbec_handle *h, const void *trap_host) {
  if (!h || !trap_host) return set_error(TBEC_ERR_ARG, "null argument");
  int rc = ensure_trap(h);
  if (rc) return rc;
  const size_t bytes = (size_t)h->p.nx * h->p.ny * h->p.nz * (h->csize / 2);
  hipError_t e = hipMemcpyAsync(h->trap, trap_host, bytes, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess)
    return set_error(TBEC_ERR_HIP, std::string("table upload: ") + hipGetErrorString(e));
  h->trap_owner = TBEC_POT_TABLE;   // replaces a mobius Trap, if one was built
  h->vtab_valid = false;
  return 0;
}

int tbec_potential(tbec_handle *h, double t, void *pot_host) {
  if (!h || !pot_host) return set_error(TBEC_ERR_ARG, "null argument");
  // evaluate into the w array (scratch between steps), then pack
  const double tt = to_t(h, t);
  const bool pot_on = !(tt > to_t(h, h->phys.potential_off));
  void *scr = h->w;
  const size_t rbytes = (size_t)h->elems * (h->csize / 2);
  hipError_t e = hipMemsetAsync(scr, 0, rbytes, h->stream);
  if (e != hipSuccess) return set_error(TBEC_ERR_HIP, hipGetErrorString(e));
  int rc = eval_table(h, tt, pot_on, scr);
  if (rc) return rc;
  const long long n = (long long)h->p.nx * h->p.ny * h->p.nz;
  void *dout = nullptr;
  e = hipMalloc(&dout, (size_t)n * (h->csize / 2));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(TBEC_ERR_OOM, "potential output allocation failed");
  }
  if (is_f32(h))
    hipLaunchKernelGGL(k_pack_real<float>, dim3(blocks_for(n)), dim3(256), 0, h->stream,
                       (const float *)scr, (float *)dout, h->p.nx, h->p.ny, h->p.nz, h->sy,
                       h->sx, h->base);
  else
    hipLaunchKernelGGL(k_pack_real<double>, dim3(blocks_for(n)), dim3(256), 0, h->stream,
                       (const double *)scr, (double *)dout, h->p.nx, h->p.ny, h->p.nz, h->sy,
                       h->sx, h->base);
  e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(pot_host, dout, (size_t)n * (h->csize / 2), hipMemcpyDeviceToHost,
                       h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(dout);
  if (e != hipSuccess) return set_error(TBEC_ERR_HIP, std::string("potential: ") +
                                                          hipGetErrorString(e));
  return 0;
}

int tbec_set_time(tbec_handle *h, double t) {
  if (!h) return set_error(TBEC_ERR_ARG, "null handle");
  if (!std::isfinite(t)) return set_error(TBEC_ERR_ARG, "t must be finite");
  h->t_now = to_t(h, t);
  return 0;
}

int tbec_imprint_vortex(tbec_handle *h, double x0, double y0, double dir) {
  if (!h) return set_error(TBEC_ERR_ARG, "null handle");
  if (!std::isfinite(x0) || !std::isfinite(y0) || !(dir == 1.0 || dir == -1.0))
    return set_error(TBEC_ERR_ARG, "imprint needs a finite centre and dir = +1 or -1");
  int rc = tbec_materialize_scale(h);
  if (rc) return rc;
  // pade_density's constants (src/force.f90:49-51) in the working precision
  double c1, c2, c3;
  if (is_f32(h)) {
    const float a = 11.0f / 32.0f, b = (5.0f - 32.0f * a) / (48.0f - 192.0f * a);
    c1 = a; c3 = b; c2 = (double)(a * (b - 0.25f));
  } else {
    c1 = 11.0 / 32.0;
    c3 = (5.0 - 32.0 * c1) / (48.0 - 192.0 * c1);
    c2 = c1 * (c3 - 0.25);
  }
  const long long n = (long long)h->p.nx * h->p.ny * h->p.nz;
  if (is_f32(h))
    hipLaunchKernelGGL(k_imprint<float>, dim3(blocks_for(n)), dim3(256), 0, h->stream,
                       (cplx<float> *)h->psi[h->cur], h->p.nx, h->p.ny, h->p.nz, h->sy, h->sx,
                       h->base, h->p.Lx, h->p.Ly, (double)(float)x0, (double)(float)y0, dir, c1,
                       c2, c3);
  else
    hipLaunchKernelGGL(k_imprint<double>, dim3(blocks_for(n)), dim3(256), 0, h->stream,
                       (cplx<double> *)h->psi[h->cur], h->p.nx, h->p.ny, h->p.nz, h->sy,
                       h->sx, h->base, h->p.Lx, h->p.Ly, x0, y0, dir, c1, c2, c3);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(TBEC_ERR_HIP, std::string("imprint: ") +
                                                          hipGetErrorString(e));
  return 0;
}

}  // extern "C"
