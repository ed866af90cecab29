// The reference's own initial states (mesh_init's init_cond select,
// src/initial.f90:58-156, with setup_* at :202-652) and the insertions applied before the
// real-time loop (setup_pcurrent, setup_add_ring, setup_add_dipole, phase_imprint,
// add_noise: src/initial.f90:669-870, called at src/run.f90:26,37-40), on the device.
//
// One thread per interior point (grid (ceil(nz/256), ny, nx), indices from the launch
// coordinates).  Each point's value is the reference's product of vortex factors, in the
// reference's order, with contraction off: real x complex is the full complex product
// (d, 0)*(c, s) as amdflang forms it.  Values formed once per state (trig of fixed
// angles, the bundle centres, exp(eye*pi)) and the values the reference draws at random
// come from the host in the working precision; per-point transcendentals (atan2, pow for
// r**1.15, exp, cos, sin) come from the device library and can differ from the host's in
// the last place.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "gpe_stage.hpp"
#include "tbec_internal.hpp"

using namespace tbec;

namespace {

constexpr int kMaxPrm = 128;
constexpr int kMaxRings = 8;

struct RefArgs {
  int kind, what, winding, potential, two_dim;
  int nx, ny, nz;
  long long sy, sx, base;
  uint32_t seed;
  double Lx, Ly, Lz, wx, wy, wz, ring_rad;
  int nprm;
  double prm[kMaxPrm];
};

template <typename T> struct Cx { T re, im; };

template <typename T> __device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b) {
#pragma clang fp contract(off)
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// real x complex as amdflang forms it: (d, 0) * (c, s)
template <typename T> __device__ __forceinline__ Cx<T> rmul(T d, T c, T s) {
  return cmul<T>({d, (T)0}, {c, s});
}

template <typename T> __device__ __forceinline__ T coord(int i, int n, T L) {
#pragma clang fp contract(off)
  return ((T)(2 * i + 1) / ((T)2 * (T)n)) * L - L / (T)2;
}

// func*exp(s*eye*theta) of a straight line (src/initial.f90:96-104 and its copies):
// func = 1 - exp(-c*r**1.15), theta = atan2(d2, d1).
template <typename T> __device__ Cx<T> line(T d1, T d2, T c, T s) {
#pragma clang fp contract(off)
  const T r = sqrt(d1 * d1 + d2 * d2);
  const T theta = atan2(d2, d1);
  const T func = (T)1 - exp(-(c * pow(r, (T)1.15)));
  const T a = s * theta;
  return rmul<T>(func, cos(a), sin(a));
}

// ring_density (src/initial.f90:628-635)
template <typename T> __device__ T ring_density(T r) {
#pragma clang fp contract(off)
  const T c1 = (T)11 / (T)32;
  const T c3 = ((T)5 - (T)32 * c1) / ((T)48 - (T)192 * c1);
  const T c2 = c1 * (c3 - (T)0.25);
  return sqrt(((r * r) * (c1 + c2 * (r * r))) / (((T)1 + c3 * (r * r)) + c2 * (r * r * r * r)));
}

// ring_density(r)*exp(eye*theta) with theta = atan2(a, q +- R), r = |(a, q +- R)|
template <typename T> __device__ Cx<T> ring_term(T a, T q, T R, bool plus) {
#pragma clang fp contract(off)
  const T qq = plus ? q + R : q - R;
  const T theta = atan2(a, qq);
  const T r = sqrt(a * a + qq * qq);
  return rmul<T>(ring_density<T>(r), cos(theta), sin(theta));
}

template <typename T> __device__ __forceinline__ Cx<T> conjg(Cx<T> a) { return {a.re, -a.im}; }

template <typename T> __device__ T tf_dense(const RefArgs &a, T x, T y, T z) {
#pragma clang fp contract(off)
  const T ax = x / (T)a.Lx, ay = y / (T)a.Ly, az = z / (T)a.Lz;
  return (((T)1 - (T)a.wx * (ax * ax)) - (T)a.wy * (ay * ay)) - (T)a.wz * (az * az);
}

template <typename T> __device__ Cx<T> state(const RefArgs &a, T x, T y, T z) {
#pragma clang fp contract(off)
  const T Lx = (T)a.Lx, Ly = (T)a.Ly, Lz = (T)a.Lz, R = (T)a.ring_rad;
  const T pi = (T)3.14159265359;   // src/cdata.f90:13
  const double *prm = a.prm;
  Cx<T> psi{(T)0, (T)0};
  switch (a.kind) {
    case TBEC_IC_CONSTANT: {
      const T c = (T)1 / sqrt((T)2);
      psi = {c, c};
      break;
    }
    case TBEC_IC_SOLITON: {
      T tf = tf_dense<T>(a, x, y, z) / sqrt((T)2);
      if (tf < (T)0) tf = (T)0;
      const T ld = hypot(tf, tf);
      psi = x < (T)0 ? rmul<T>(ld, (T)1, (T)0) : rmul<T>(ld, (T)prm[0], (T)prm[1]);
      break;
    }
    case TBEC_IC_RANDOM_PHASE: {
      for (int q = 0; q < 20; ++q) {
        const double *w = prm + 5 * q;
        T s = x * (T)w[0] + y * (T)w[1];
        if (!a.two_dim) s = s + z * (T)w[2];
        const T arg = (((s * (T)2) * pi) * Lx) / (Lx * Lx);   // eye*s*2*pi/Lx
        const Cx<T> t = cmul<T>({(T)w[3], (T)w[4]}, {cos(arg), sin(arg)});
        psi = {psi.re + t.re, psi.im + t.im};
      }
      T tf = tf_dense<T>(a, x, y, z);
      if (tf < (T)0) tf = (T)0;
      psi = {psi.re * tf, psi.im * tf};
      break;
    }
    case TBEC_IC_CROW: {
      const T v2 = -Ly / (T)4;
      const T v1 = (T)(-4) + cos((pi * z) / Lz), w1 = (T)4 - cos((pi * z) / Lz);
      psi = line<T>(x - v1, y - v2, (T)0.7, (T)1);
      psi = cmul<T>(psi, line<T>(x - w1, y - v2, (T)0.7, (T)(-1)));
      break;
    }
    case TBEC_IC_SINGLE_VORTEX:
      psi = line<T>(x - (T)0, y - (T)0, (T)0.7, (T)a.winding);
      break;
    case TBEC_IC_SAUSAGE:
      psi = line<T>(x - (T)0, y - (T)0, fabs(z) < (T)4 ? (T)0.1 : (T)0.7, (T)1);
      break;
    case TBEC_IC_RANDOM_POINTS:
      psi = {(T)0.7071, (T)0.7071};
      for (int q = 0; q < 20; ++q)
        psi = cmul<T>(psi, line<T>(x - (T)prm[3 * q], y - (T)prm[3 * q + 1], (T)0.7,
                                   (T)prm[3 * q + 2]));
      break;
    case TBEC_IC_LEAP_FROG: {
      psi = {(T)0.7071, (T)0.7071};
      const T px[4] = {(T)(-3), (T)3, (T)(-3), (T)3}, py[4] = {(T)3, (T)12, (T)(-3), (T)(-12)};
      const T dir[4] = {(T)1, (T)1, (T)(-1), (T)(-1)};
      for (int q = 0; q < 4; ++q) psi = cmul<T>(psi, line<T>(x - px[q], y - py[q], (T)0.7, dir[q]));
      break;
    }
    case TBEC_IC_VORTEX_RINGS: {
      psi = {(T)1, (T)0};
      for (int q = 0; q < a.nprm / 12; ++q) {
        const double *c = prm + 12 * q;
        const T tx = ((T)c[0] * x + (T)c[1] * y) + (T)c[2] * z;
        const T ty = ((T)c[3] * x + (T)c[4] * y) + (T)c[5] * z;
        const T tz = ((T)c[6] * x - (T)c[7] * y) + (T)c[8] * z;
        const T dx = tx - (T)c[9], dy = ty - (T)c[10], dz = tz - (T)c[11];
        const T qd = sqrt(dy * dy + dz * dz);
        const T theta = atan2(dz, dy);
        const T ct = cos(theta), st = sin(theta);
        Cx<T> p12[2];
        for (int sg = 0; sg < 2; ++sg) {
          const T qq = sg == 0 ? qd + R : qd - R;
          const T r = sqrt(dx * dx + qq * qq);
          const T d = sqrt(((r * r) * ((T)0.347 + (T)0.0286 * (r * r))) /
                           (((T)1 + (T)0.3333 * (r * r)) + (T)0.0286 * (r * r * r * r)));
          p12[sg] = cmul<T>({d, (T)0}, {ct, st});
        }
        psi = cmul<T>(psi, cmul<T>(p12[0], conjg<T>(p12[1])));
      }
      break;
    }
    case TBEC_IC_LINE_OF_LINES: {
      const T c = (T)1 / sqrt((T)2);
      psi = {c, c};
      const T ax = x - (T)0, qd = sqrt(y * y + z * z);
      psi = cmul<T>(psi, ring_term<T>(ax, qd, R, true));
      psi = cmul<T>(psi, conjg<T>(ring_term<T>(ax, qd, R, false)));
      break;
    }
    case TBEC_IC_LINKED_RINGS: {
      const T t1 = R * (T)0;
      const T ax = x - (T)0, q1 = sqrt(y * y + (z - t1) * (z - t1));
      psi = cmul<T>(ring_term<T>(ax, q1, R, true), conjg<T>(ring_term<T>(ax, q1, R, false)));
      const T t2 = R * (T)1.2;
      const T by = y - (T)0, q2 = sqrt(x * x + (z - t2) * (z - t2));
      psi = cmul<T>(psi, ring_term<T>(by, q2, R, true));
      psi = cmul<T>(psi, conjg<T>(ring_term<T>(by, q2, R, false)));
      break;
    }
    case TBEC_IC_LATTICE:
      psi = {(T)1, (T)1};
      for (int v1 = 1; v1 <= 10; ++v1)
        for (int v2 = 1; v2 <= 10; ++v2) {
          const T p1 = (T)v1 * (Lx / (T)16) - Lx / (T)3;
          const T p2 = (T)v2 * (Ly / (T)16) - Ly / (T)3;
          psi = cmul<T>(psi, line<T>(x - p1, y - p2, (T)0.7, (T)1));
        }
      break;
    case TBEC_IC_LATTICE2: {
      T tf = tf_dense<T>(a, x, y, z) / sqrt((T)2);
      if (tf < (T)0) tf = (T)0;
      psi = {tf, tf};
      for (int v1 = 1; v1 <= 3; ++v1)
        for (int v2 = 1; v2 <= 3; ++v2) {
          const T p1 = (T)v1 * (Lx / (T)4) - Lx / (T)2;
          const T p2 = (T)v2 * (Ly / (T)4) - Ly / (T)2;
          psi = cmul<T>(psi, line<T>(x - p1, y - p2, (T)0.7, (T)1));
        }
      for (int v1 = 1; v1 <= 3; ++v1)
        for (int v2 = 1; v2 <= 3; ++v2) {
          const T p1 = ((T)v1 * (Lx / (T)4) + (T)10) - Lx / (T)2;
          const T p2 = ((T)v2 * (Lz / (T)4) + (T)10) - Lz / (T)2;
          psi = cmul<T>(psi, line<T>(x - p1, z - p2, (T)0.7, (T)1));
        }
      break;
    }
    case TBEC_IC_BUNDLE: {
      T tf = tf_dense<T>(a, x, y, z) / sqrt((T)2);
      if (tf < (T)0) tf = (T)0;
      psi = {tf, tf};
      for (int q = 0; q < 7; ++q)
        psi = cmul<T>(psi, line<T>(x - (T)prm[2 * q], y - (T)prm[2 * q + 1], (T)0.7, (T)1));
      for (int q = 7; q < 14; ++q)
        psi = cmul<T>(psi, line<T>(x - (T)prm[2 * q], z - (T)prm[2 * q + 1], (T)0.7, (T)1));
      for (int q = 14; q < 21; ++q)
        psi = cmul<T>(psi, line<T>(y - (T)prm[2 * q], z - (T)prm[2 * q + 1], (T)0.7, (T)1));
      break;
    }
    default:
      break;
  }
  return psi;
}

template <typename T> __device__ Cx<T> insert(const RefArgs &a, Cx<T> c, T x, T y, T z,
                                              long long flat) {
#pragma clang fp contract(off)
  const T Lx = (T)a.Lx, Ly = (T)a.Ly, R = (T)a.ring_rad;
  const T pi = (T)3.14159265359;
  switch (a.what) {
    case TBEC_INSERT_PCURRENT:
    case TBEC_INSERT_PHASE_IMPRINT: {
      T theta = (T)0, n = (T)1;
      if (a.what == TBEC_INSERT_PHASE_IMPRINT) {
        theta = atan2(y - (T)0, x - (T)0);
        n = (T)a.winding;
      } else {
        switch (a.potential) {
          case TBEC_POT_BARRIER:
            theta = y < (T)0 ? ((x * (T)50) * pi) / Lx : (T)20 * pi - ((x * (T)50) * pi) / Lx;
            break;
          case TBEC_POT_MOBIUS:
            theta = (T)20 * atan2(y, x);
            break;
          case TBEC_POT_HARMONIC_RING: {
            const T ax = x / Lx, ay = y / Lx;
            theta = sqrt(ax * ax + ay * ay) < (T)0.3 ? atan2(y, x) : atan2(-y, x);
            n = (T)10;
            break;
          }
          case TBEC_POT_HARMONIC_RING2:
            theta = y < (T)0 ? atan2((T)0.5 * (y + Ly / (T)5), x)
                             : atan2((T)0.5 * (y - Ly / (T)5), x) - pi;
            n = (T)20;
            break;
          case TBEC_POT_CHANNEL:
            theta = (((T)20 * pi) * (x + (T)0.5 * Lx)) / Lx;
            n = (T)2;
            break;
          default:
            return c;   // no current for this potential (the select has no case)
        }
      }
      const T ld = hypot(c.re, c.im);
      const T ang = n * theta;
      return rmul<T>(ld, cos(ang), sin(ang));
    }
    case TBEC_INSERT_RING: {
      const T ax = x - (T)(-0.25) * Lx, qd = sqrt(y * y + (z - (T)0) * (z - (T)0));
      c = cmul<T>(c, ring_term<T>(ax, qd, R, true));
      return cmul<T>(c, conjg<T>(ring_term<T>(ax, qd, R, false)));
    }
    case TBEC_INSERT_DIPOLE:
      c = cmul<T>(c, line<T>(x - (T)(-20), y - (T)5, (T)0.7, (T)1));
      return cmul<T>(c, line<T>(x - (T)(-20), y - (T)(-5), (T)0.7, (T)(-1)));
    case TBEC_INSERT_NOISE: {
      const T e1 = (T)uniform01(a.seed, 204, (uint32_t)(2 * flat)) * (T)2 - (T)1;
      const T e2 = (T)uniform01(a.seed, 204, (uint32_t)(2 * flat + 1)) * (T)2 - (T)1;
      return {c.re + (T)0.01 * e1, c.im + (T)0.01 * e2};
    }
    default:
      return c;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_reference(const RefArgs a, cplx<T> *__restrict__ psi) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, i = blockIdx.z;
  if (k >= a.nz) return;
  const T x = coord<T>(i, a.nx, (T)a.Lx), y = coord<T>(j, a.ny, (T)a.Ly),
          z = coord<T>(k, a.nz, (T)a.Lz);
  cplx<T> &p = psi[a.base + i * a.sx + j * a.sy + k];
  Cx<T> v;
  if (a.what < 0) {
    v = state<T>(a, x, y, z);
  } else {
    const long long flat = ((long long)i * a.ny + j) * a.nz + k;
    v = insert<T>(a, Cx<T>{p.re, p.im}, x, y, z, flat);
  }
  p.re = v.re;
  p.im = v.im;
}

// ---- host: the values formed once per state, in the working precision --------------
template <typename T> T runif(uint32_t seed, uint32_t stream, uint32_t idx, T lo, T hi) {
  const T u = (T)uniform01(seed, stream, idx);
  return lo + u * (hi - lo);   // src/statistics.f90:29-35
}

template <typename T> T tcos(T v);
template <typename T> T tsin(T v);
template <> float tcos<float>(float v) { return cosf(v); }
template <> float tsin<float>(float v) { return sinf(v); }
template <> double tcos<double>(double v) { return std::cos(v); }
template <> double tsin<double>(double v) { return std::sin(v); }

template <typename T> int fill_prm(const tbec_handle *h, const tbec_reference_ic *ic, RefArgs &a) {
  const T pi = (T)3.14159265359;
  const T Lx = (T)h->p.Lx, Ly = (T)h->p.Ly;
  double *prm = a.prm;
  int n = 0;
  switch (ic->kind) {
    case TBEC_IC_SOLITON:   // exp(eye*pi) (src/initial.f90:76)
      prm[n++] = (double)tcos<T>(pi);
      prm[n++] = (double)tsin<T>(pi);
      break;
    case TBEC_IC_RANDOM_PHASE:   // setup_quench's draws (src/initial.f90:351-357)
      for (int q = 0; q < 20; ++q) {
        for (int w = 0; w < 3; ++w)
          prm[n++] = (double)std::floor(runif<T>(ic->seed, 201, 5 * q + w, (T)(-20), (T)20));
        for (int w = 3; w < 5; ++w)
          prm[n++] = (double)runif<T>(ic->seed, 201, 5 * q + w, (T)(-0.1), (T)0.1);
      }
      break;
    case TBEC_IC_RANDOM_POINTS:   // src/initial.f90:276-285
      for (int q = 0; q < 20; ++q) {
        prm[n++] = (double)runif<T>(ic->seed, 202, 3 * q, -Lx / (T)5, Lx / (T)5);
        prm[n++] = (double)runif<T>(ic->seed, 202, 3 * q + 1, -Ly / (T)5, Ly / (T)5);
        const T dir = runif<T>(ic->seed, 202, 3 * q + 2, (T)0.9, (T)1);
        prm[n++] = dir < (T)0.5 ? -1.0 : 1.0;
      }
      break;
    case TBEC_IC_VORTEX_RINGS: {   // src/initial.f90:216-253
      if (ic->vor_ring_count < 1 || ic->vor_ring_count > kMaxRings)
        return set_error(TBEC_ERR_ARG, "vor_ring_count must be 1..8");
      for (int v = 0; v < ic->vor_ring_count; ++v) {
        const uint32_t b = 6 * (uint32_t)v;
        const T ax = runif<T>(ic->seed, 203, b, (T)0, (T)0);
        const T ay = runif<T>(ic->seed, 203, b + 1, (T)0, (T)2 * pi);
        const T az = runif<T>(ic->seed, 203, b + 2, (T)0, (T)2 * pi);
        const T tu = runif<T>(ic->seed, 203, b + 3, (T)(-1), (T)1);
        const T tt = runif<T>(ic->seed, 203, b + 4, (T)0, (T)2 * pi);
        const T tr = runif<T>(ic->seed, 203, b + 5, (T)0, Lx / (T)6);
        const T cax = tcos<T>(ax), sax = tsin<T>(ax), cay = tcos<T>(ay), say = tsin<T>(ay);
        const T caz = tcos<T>(az), saz = tsin<T>(az);
        const T c[12] = {cay * caz,
                         cax * saz + (sax * say) * caz,
                         sax * saz - (cax * say) * caz,
                         -(cay * saz),
                         cax * caz - (sax * say) * saz,
                         sax * cay + (cax * say) * saz,
                         say,
                         sax * cay,
                         cax * cay,
                         (tr * std::sqrt((T)1 - tu * tu)) * tcos<T>(tt),
                         (tr * std::sqrt((T)1 - tu * tu)) * tsin<T>(tt),
                         tr * tu};
        for (const T v12 : c) prm[n++] = (double)v12;
      }
      break;
    }
    case TBEC_IC_BUNDLE:   // the three bundles' centres (src/initial.f90:476-535)
      for (int bdl = 0; bdl < 3; ++bdl)
        for (int v1 = 1; v1 <= 7; ++v1) {
          const T rpos = v1 == 1 ? (T)0 : (T)10;
          const T tp = v1 == 1 ? (T)0 : (((T)(v1 - 2) * (T)2) * pi) / (T)6;
          T p1 = rpos * tcos<T>(tp), p2 = rpos * tsin<T>(tp);
          if (bdl == 1) p2 = p2 - (T)30;
          else p1 = p1 - (T)30;
          prm[n++] = (double)p1;
          prm[n++] = (double)p2;
        }
      break;
    default:
      break;
  }
  a.nprm = n;
  return 0;
}

RefArgs ref_args(const tbec_handle *h, const tbec_reference_ic *ic) {
  RefArgs a{};
  a.kind = ic->kind;
  a.what = -1;
  a.winding = ic->winding_number;
  a.potential = h->phys.potential;
  a.two_dim = h->phys.two_dim;
  a.nx = h->p.nx; a.ny = h->p.ny; a.nz = h->p.nz;
  a.sy = h->sy; a.sx = h->sx; a.base = h->base;
  a.seed = ic->seed;
  a.Lx = h->p.Lx; a.Ly = h->p.Ly; a.Lz = h->p.Lz;
  a.wx = h->p.harm_wx; a.wy = h->p.harm_wy; a.wz = h->p.harm_wz;
  a.ring_rad = ic->ring_rad;
  return a;
}

int launch(tbec_handle *h, const RefArgs &a) {
  if (h->p.nx > 65535 || h->p.ny > 65535)
    return set_error(TBEC_ERR_ARG, "reference states need nx, ny <= 65535");
  const dim3 grid((unsigned)((h->p.nz + 255) / 256), (unsigned)h->p.ny, (unsigned)h->p.nx);
  if (h->p.prec == TBEC_F32)
    hipLaunchKernelGGL(k_reference<float>, grid, dim3(256), 0, h->stream, a,
                       (cplx<float> *)h->psi[h->cur]);
  else
    hipLaunchKernelGGL(k_reference<double>, grid, dim3(256), 0, h->stream, a,
                       (cplx<double> *)h->psi[h->cur]);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess)
    return set_error(TBEC_ERR_HIP, std::string("reference state: ") + hipGetErrorString(e));
  return 0;
}

}  // namespace

extern "C" {

int tbec_reference_state(tbec_handle *h, const tbec_reference_ic *ic) {
  if (!h || !ic) return set_error(TBEC_ERR_ARG, "null argument");
  if (ic->kind < TBEC_IC_CONSTANT || ic->kind > TBEC_IC_BUNDLE)
    return set_error(TBEC_ERR_ARG, "unknown reference state " + std::to_string(ic->kind));
  int rc = tbec_materialize_scale(h);
  if (rc) return rc;
  RefArgs a = ref_args(h, ic);
  rc = h->p.prec == TBEC_F32 ? fill_prm<float>(h, ic, a) : fill_prm<double>(h, ic, a);
  if (rc) return rc;
  rc = launch(h, a);
  if (rc) return rc;
  return tbec_reset_state(h);   // a new state: the imaginary-time bookkeeping restarts
}

int tbec_reference_insert(tbec_handle *h, int what, const tbec_reference_ic *ic) {
  if (!h || !ic) return set_error(TBEC_ERR_ARG, "null argument");
  if (what < TBEC_INSERT_PCURRENT || what > TBEC_INSERT_NOISE)
    return set_error(TBEC_ERR_ARG, "unknown insertion " + std::to_string(what));
  int rc = tbec_materialize_scale(h);
  if (rc) return rc;
  RefArgs a = ref_args(h, ic);
  a.what = what;
  rc = launch(h, a);
  if (rc) return rc;
  return tbec_reset_state(h);
}

}  // extern "C"
