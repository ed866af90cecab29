// Fused GPE right-hand side + low-storage RK3 stage for gfx950 (CDNA4).
//
// One launch = one RK stage over the whole grid: it reads psi once from HBM, evaluates
//   rhs = D2x psi + D2y psi + D2z psi + om*i*(x D1y psi - y D1x psi)
//         + (c0 - gnl*|psi|^2 - pot) psi,     pot = sum_a harm_w[a]*(x_a/L_a)^2
// (src/finite_diff.f90:19-135, src/deriv.f90:12-25) and applies the stage update
//   w <- A*w + c*rhs,  psi_out <- psi + B*w     (c = i*dt/2)
// (src/run.f90:46-54), writing psi_out and w.  Variants: the reference imaginary
// Euler update psi + (dt/2)*rhs (src/run.f90:22) and a rhs-only pass (test hook).
//
// Two ways to carry w between stages.  Stored (the reference's scheme): stages 1 and 2
// write w, stages 2 and 3 read it: 10 HBM transfers of a complex value per point and
// step.  Rebuilt: w is not stored.  The stage input psi_s and the previous stage's input
// psi_{s-1} (still held in the other ping-pong buffer) give it back pointwise:
// w_{s-1} = (psi_s - psi_{s-1}) / B_{s-1}, since psi_s = psi_{s-1} + B_{s-1} w_{s-1}.
// Stage 2 reads psi_1 (stencil) and psi_0 and writes psi_2 over psi_0; stage 3 reads
// psi_2 and psi_1 and writes psi_3 over psi_1 (each point is read and rewritten by the
// same thread only).  That is 8 HBM transfers of a complex value per point and step
// instead of the 10 of a stored w.  In imaginary time stage 1 reads scale*psi_0; stage
// 2 forms the same product (same rounding) from psi_0 before the difference.
//
// Layout (per condensate): psi, w are padded arrays, z fastest.  Interior (0,0,0) sits
// at `base`; strides sy (y) and sx (x) in complex elements.  The halo (3 cells, more at
// row ends) is zero and never written, which is the reference's zero boundary
// (src/ghost.f90:55,88,121,154,179-181).
//
// Work decomposition: a workgroup owns a TZ x TY tile of (z, y) and marches along x
// over a segment of SX planes.  Each thread owns VZ consecutive z points of one y row
// and keeps their x-neighbourhood (7 planes) in a register queue, so D2x/D1x need no
// memory traffic.  The centre plane of the tile plus a 3-deep y halo and a 4-wide
// (aligned) z halo is staged in LDS (double buffered, one barrier per plane) for the y
// and z stencils.  Plane x+4 (own points) and plane x+1 (halo) are prefetched one
// iteration ahead (own points two ahead, plane x+5, in the first RK stage of the
// real-time analytic-trap instances).  Consecutive workgroup ids are spread over the 8 XCDs by the
// dispatcher; the tile index is remapped so each XCD gets a contiguous block of tiles
// (shared halos stay in that XCD's L2).
//
// Summation modes for the 7-tap D2 sums:
//   SUM_REF  : the reference's left-to-right sum (FP64 parity path)
//   SUM_COMP : compensated: each symmetric pair f[-k]+f[k] is formed by an error-free
//              TwoSum (s, e); d_k = (s - 2 f0) + e (the subtraction is exact by
//              Sterbenz for smooth data), D2 = (270 d1 - 27 d2 + 2 d3)/(180 h^2).
//              The rounding is then relative to h^2 f'' instead of to f.
//   SUM_CENTRED: d_k = (f[-k] - f0) + (f[k] - f0): both differences are exact (Sterbenz)
//              for smooth data and the one rounding of the sum is relative to d_k, the
//              same error as SUM_COMP at less than half the instructions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace tbec {

enum SumMode { SUM_REF = 0, SUM_COMP = 1, SUM_CENTRED = 2 };
// Physics variants of the stage kernel (a template parameter):
//   PH_HARM: the analytic harmonic trap, no first-derivative term;
//   PH_ROT : the same plus the rotating-frame term om*i*(x D1y - y D1x);
//   PH_GEN : pot read from a tabulated potential V (same padded layout as psi, one real
//            per point; any of the reference's potential kinds), and a run-time drift:
//            0 none, 1 the frame term, 2 the advective term 2 vx i D1x
//            (src/finite_diff.f90:126-130).
enum PhysKind { PH_HARM = 0, PH_ROT = 1, PH_GEN = 2 };
enum StageKind { ST_RK1 = 0, ST_RK2 = 1, ST_RK3 = 2, ST_EULER = 3, ST_RHS = 4 };
// Stage-kind template argument of the kernel instances with the kind fixed at compile
// time: an RK stage kind, plus SK_STORED for the stored-w scheme (w rebuilt otherwise).
constexpr int SK_STORED = 8;

template <typename T> struct alignas(2 * sizeof(T)) cplx { T re, im; };

struct StageArgs {
  const void *psi_in;
  void *psi_out;
  void *w;
  int nx, ny, nz;
  int seg_len;                 // SX
  int tiles_z, tiles_y, nseg;  // tile grid
  int nblocks;                 // tiles_z * tiles_y * nseg
  long long sy, sx, base;      // strides / interior offset (complex elements)
  long long elems;             // allocated complex elements per buffer
  double x0, y0, z0;           // coordinate of interior index 0 on each axis
  double hx, hy, hz;
  double k2[3];                // 1 / (180 h_a^2)
  double k1[2];                // 1 / (60 h_x), 1 / (60 h_y)
  double wl[3];                // harm_w[a] / L_a^2
  double c0, gnl, om;
  double cre, cim;             // complex update coefficient
  double A, B;                 // stage coefficients
  double wrec;                 // 0: read w from w_buf.  Else (stages 2, 3, rebuilt w):
                               // w = (psi - pscale*w_buf) * wrec, where w_buf holds the
                               // previous stage's input
  int store_w;                 // write w (stored scheme, stages 1 and 2)
  int stage;                   // StageKind
  int remap;                   // 1: spread tiles over XCDs in contiguous chunks
  const double *scale;         // imaginary time: psi is read as scale*psi (or null)
  const double *prev_scale;    // imaginary time, stage 2: scale of w_buf (or null)
  const double *c0dev;         // imaginary time: c0 from device memory (or null)
  const int *skip;             // imaginary time: converged flag; skip the pass when set
  double *partials;            // reductions: 4 doubles per workgroup
  const void *vtab;            // PH_GEN: tabulated potential (padded layout, real T)
  int drift;                   // PH_GEN: 0 none, 1 frame term (om), 2 advective (vx2)
  double vx2;                  // PH_GEN: 2 * advective_vx
};

// ---- small helpers -------------------------------------------------------------
__device__ __forceinline__ bool stage_reads_w(int st) { return st == ST_RK2 || st == ST_RK3; }

template <typename T> __device__ __forceinline__ void two_sum(T a, T b, T &s, T &e) {
  s = a + b;
  T bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Explicit FMAs for the stage arithmetic: fmaT(a, b, c) = a*b + c rounded once, and
// fma2(a, b, c, d) = a*b + c*d as fma(a, b, c*d).  Left to the compiler, the contraction
// would not be the same in every unrolled copy of the plane loop: for a sum of two
// products it may fuse either one, and a product that also occurs in another copy (the
// x stencil's 270*q[x+1] of plane x is plane x+2's 270*q[x+1]) is computed once and then
// not fused at all.  A plane's rounding would then depend on its position in the unrolled
// loop, and so on the segment length.  Sums of two products and the stencil chains of
// plane values go through these.
template <typename T> __device__ __forceinline__ T fmaT(T a, T b, T c) {
  if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}
template <typename T> __device__ __forceinline__ T fma2(T a, T b, T c, T d) {
  return fmaT(a, b, c * d);
}

template <typename T, int VZ> struct VecIO;
template <> struct VecIO<float, 4> {   // 4 complex floats = 2 x 16 B
  static __device__ __forceinline__ void ld(const cplx<float> *p, cplx<float> *o) {
    const float4 *q = reinterpret_cast<const float4 *>(p);
    float4 a = q[0], b = q[1];
    o[0] = {a.x, a.y}; o[1] = {a.z, a.w}; o[2] = {b.x, b.y}; o[3] = {b.z, b.w};
  }
  static __device__ __forceinline__ void st(cplx<float> *p, const cplx<float> *o) {
    float4 *q = reinterpret_cast<float4 *>(p);
    q[0] = make_float4(o[0].re, o[0].im, o[1].re, o[1].im);
    q[1] = make_float4(o[2].re, o[2].im, o[3].re, o[3].im);
  }
};
template <> struct VecIO<float, 2> {
  static __device__ __forceinline__ void ld(const cplx<float> *p, cplx<float> *o) {
    float4 a = *reinterpret_cast<const float4 *>(p);
    o[0] = {a.x, a.y}; o[1] = {a.z, a.w};
  }
  static __device__ __forceinline__ void st(cplx<float> *p, const cplx<float> *o) {
    *reinterpret_cast<float4 *>(p) = make_float4(o[0].re, o[0].im, o[1].re, o[1].im);
  }
};
template <> struct VecIO<double, 2> {  // 2 complex doubles = 2 x 16 B
  static __device__ __forceinline__ void ld(const cplx<double> *p, cplx<double> *o) {
    const double2 *q = reinterpret_cast<const double2 *>(p);
    double2 a = q[0], b = q[1];
    o[0] = {a.x, a.y}; o[1] = {b.x, b.y};
  }
  static __device__ __forceinline__ void st(cplx<double> *p, const cplx<double> *o) {
    double2 *q = reinterpret_cast<double2 *>(p);
    q[0] = make_double2(o[0].re, o[0].im);
    q[1] = make_double2(o[1].re, o[1].im);
  }
};
template <> struct VecIO<float, 1> {
  static __device__ __forceinline__ void ld(const cplx<float> *p, cplx<float> *o) {
    float2 a = *reinterpret_cast<const float2 *>(p);
    o[0] = {a.x, a.y};
  }
  static __device__ __forceinline__ void st(cplx<float> *p, const cplx<float> *o) {
    *reinterpret_cast<float2 *>(p) = make_float2(o[0].re, o[0].im);
  }
};
template <> struct VecIO<double, 1> {
  static __device__ __forceinline__ void ld(const cplx<double> *p, cplx<double> *o) {
    double2 a = *reinterpret_cast<const double2 *>(p);
    o[0] = {a.x, a.y};
  }
  static __device__ __forceinline__ void st(cplx<double> *p, const cplx<double> *o) {
    *reinterpret_cast<double2 *>(p) = make_double2(o[0].re, o[0].im);
  }
};

// Raw buffer access for the stage kernel: a uniform 64-bit plane base in a buffer
// resource (SGPRs) plus a 32-bit per-lane byte offset (one VGPR), instead of a 64-bit
// address per lane.  The resource's record count is one x-plane, so every access is
// range-checked against the plane it belongs to; the march only touches planes
// -3 .. nx+3, which the padded layout holds.
typedef unsigned int tbec_u32x4 __attribute__((ext_vector_type(4)));
constexpr int kRsrcWord3 = 0x00020000;   // gfx9-family raw buffer, no swizzle

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void *p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, bytes, kRsrcWord3);
}

typedef unsigned int tbec_u32x2 __attribute__((ext_vector_type(2)));

// Cache policy (the buffer instructions' cpol operand; bit 1 is NT, non-temporal, on
// gfx950).  Each point's output and its pointwise inputs (w or psi_{s-1}, the tabulated
// potential) are touched once per stage: they are streamed with NT so that they do not
// push out of L2 the stencil input's lines that neighbouring tiles still read as halo.
// The stencil input keeps the default policy.  Measured at 512^3 on MI355X: FP32 -3.4%,
// FP64 -9.3% per step; NT on the halo loads (+2%) or on the own-point loads (+9%) is worse.
#ifndef TBEC_STORE_CPOL
#define TBEC_STORE_CPOL 2
#endif
#ifndef TBEC_POINTWISE_LOAD_CPOL
#define TBEC_POINTWISE_LOAD_CPOL 2
#endif
#ifndef TBEC_OWN_LOAD_CPOL
#define TBEC_OWN_LOAD_CPOL 0
#endif
#ifndef TBEC_HALO_LOAD_CPOL
#define TBEC_HALO_LOAD_CPOL 0
#endif
constexpr int kStoreCpol = TBEC_STORE_CPOL;
constexpr int kPointwiseLoadCpol = TBEC_POINTWISE_LOAD_CPOL;
constexpr int kOwnLoadCpol = TBEC_OWN_LOAD_CPOL;
constexpr int kHaloLoadCpol = TBEC_HALO_LOAD_CPOL;
#ifndef TBEC_UNIFORM_VMEM
#define TBEC_UNIFORM_VMEM 1
#endif
#ifndef TBEC_AHEAD2
#define TBEC_AHEAD2 1
#endif
// A buffer offset past every plane's record count (a plane is < 2^31 bytes): loads there
// return zeros without a memory access, stores there are dropped.
constexpr uint32_t kOffNone = 0x80000000u;

template <typename T, int VZ> struct BufIO {
  // VZ complex values per thread: 16-byte chunks, or one 8-byte access for a single
  // complex<float>
  static constexpr int BYTES = VZ * 2 * (int)sizeof(T);
  static constexpr int NB = BYTES / 16;
  static_assert(BYTES == 8 || BYTES % 16 == 0, "unsupported vector width");
  template <int CP = 0>
  static __device__ __forceinline__ void ld(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                            cplx<T> *o) {
    if constexpr (BYTES == 8) {
      const tbec_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, CP);
      o[0] = {__uint_as_float(v.x), __uint_as_float(v.y)};
    } else {
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        const tbec_u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16 * c, 0, CP);
        if constexpr (sizeof(T) == 4) {
          o[2 * c] = {__uint_as_float(v.x), __uint_as_float(v.y)};
          o[2 * c + 1] = {__uint_as_float(v.z), __uint_as_float(v.w)};
        } else {
          o[c] = {__hiloint2double(v.y, v.x), __hiloint2double(v.w, v.z)};
        }
      }
    }
  }
  template <int CP = 0>
  static __device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                            const cplx<T> *o) {
    if constexpr (BYTES == 8) {
      const tbec_u32x2 v = tbec_u32x2{__float_as_uint(o[0].re), __float_as_uint(o[0].im)};
      __builtin_amdgcn_raw_buffer_store_b64(v, r, off, 0, CP);
    } else {
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        tbec_u32x4 v;
        if constexpr (sizeof(T) == 4) {
          v = tbec_u32x4{__float_as_uint(o[2 * c].re), __float_as_uint(o[2 * c].im),
                         __float_as_uint(o[2 * c + 1].re), __float_as_uint(o[2 * c + 1].im)};
        } else {
          v = tbec_u32x4{(unsigned)__double2loint(o[c].re), (unsigned)__double2hiint(o[c].re),
                         (unsigned)__double2loint(o[c].im), (unsigned)__double2hiint(o[c].im)};
        }
        __builtin_amdgcn_raw_buffer_store_b128(v, r, off + 16 * c, 0, CP);
      }
    }
  }
};

// VZ consecutive reals (the tabulated potential) at a byte offset of a plane resource.
template <typename T, int VZ> struct BufIOR {
  static constexpr int BYTES = VZ * (int)sizeof(T);
  static_assert(BYTES == 4 || BYTES == 8 || BYTES == 16, "unsupported vector width");
  template <int CP = 0>
  static __device__ __forceinline__ void ld(__amdgpu_buffer_rsrc_t r, uint32_t off, T *o) {
    if constexpr (BYTES == 4) {
      o[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, CP));
    } else if constexpr (BYTES == 8) {
      const tbec_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, CP);
      if constexpr (sizeof(T) == 4) {
        o[0] = __uint_as_float(v.x);
        o[1] = __uint_as_float(v.y);
      } else {
        o[0] = __hiloint2double(v.y, v.x);
      }
    } else {
      const tbec_u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, CP);
      if constexpr (sizeof(T) == 4) {
        o[0] = __uint_as_float(v.x); o[1] = __uint_as_float(v.y);
        o[2] = __uint_as_float(v.z); o[3] = __uint_as_float(v.w);
      } else {
        o[0] = __hiloint2double(v.y, v.x);
        o[1] = __hiloint2double(v.w, v.z);
      }
    }
  }
};

// 7-tap second derivative sum (without the 1/(180 h^2) factor) of one component.
template <typename T, int SUM>
__device__ __forceinline__ T d2sum(T m3, T m2, T m1, T c, T p1, T p2, T p3) {
  if constexpr (SUM == SUM_REF) {
    T r = fma2(T(2), m3, T(-27), m2);
    r = fmaT(T(270), m1, r);
    r = fmaT(T(-490), c, r);
    r = fmaT(T(270), p1, r);
    r = fmaT(T(-27), p2, r);
    return fmaT(T(2), p3, r);
  } else if constexpr (SUM == SUM_COMP) {
    T s1, e1, s2, e2, s3, e3;
    two_sum(m1, p1, s1, e1);
    two_sum(m2, p2, s2, e2);
    two_sum(m3, p3, s3, e3);
    const T c2 = c + c;
    const T d1 = (s1 - c2) + e1, d2 = (s2 - c2) + e2, d3 = (s3 - c2) + e3;
    return fma2(T(270), d1, T(-27), d2) + T(2) * d3;
  } else {
    const T d1 = (m1 - c) + (p1 - c);
    const T d2 = (m2 - c) + (p2 - c);
    const T d3 = (m3 - c) + (p3 - c);
    return fma2(T(270), d1, T(-27), d2) + T(2) * d3;
  }
}

// The centred 7-tap sum from its six differences dm_k = f[-k] - f0, dp_k = f[k] - f0: the
// operations of d2sum<SUM_CENTRED> after its subtractions, in the same order.
template <typename T>
__device__ __forceinline__ T d2diff(T dm1, T dp1, T dm2, T dp2, T dm3, T dp3) {
  const T d1 = dm1 + dp1, d2 = dm2 + dp2, d3 = dm3 + dp3;
  return fma2(T(270), d1, T(-27), d2) + T(2) * d3;
}

// 7-tap first derivative sum (without 1/(60 h)).
template <typename T>
__device__ __forceinline__ T d1sum(T m3, T m2, T m1, T p1, T p2, T p3) {
  T r = fmaT(T(9), m2, -m3);
  r = fmaT(T(-45), m1, r);
  r = fmaT(T(45), p1, r);
  r = fmaT(T(-9), p2, r);
  return r + p3;
}

// Workgroup-wide sum of 4 doubles; thread 0 writes them.
template <int NT>
__device__ __forceinline__ void block_sum4(double v[4], double *out) {
  __shared__ double red[NT / 64][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) red[wid][q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double s = 0.0;
      for (int k = 0; k < NT / 64; ++k) s += red[k][q];
      out[q] = s;
    }
  }
}

// ---- the stage kernel ----------------------------------------------------------
// Process one tile (TZ x TY in z, y; one x segment) of one condensate: logical tile id
// lb in [0, tiles_z*tiles_y*nseg).  Accumulates this thread's reduction terms in red.
template <typename T, int TZ, int TY, int VZ, int SUM, int PHYS, bool IMAG, int SK = -1>
__device__ __forceinline__ void march_tile(const StageArgs &a, const int lb, double red[4]) {
  using C = cplx<T>;
  constexpr int NTZ = TZ / VZ;
  constexpr int NT = NTZ * TY;
  constexpr int HZV = 4 / VZ;           // vectors per z-halo side (4 complex wide)
  constexpr int NHALO = 6 * NTZ + 2 * TY * HZV;
  static_assert(NHALO <= NT, "one halo vector per thread at most");
  constexpr int LZ = TZ + 8;            // 4 halo columns each side
  constexpr int LY = TY + 6;
  __shared__ C lds[2][LY][LZ];

  const int tiles = a.tiles_z * a.tiles_y;
  const int seg = lb / tiles;
  const int t2 = lb - seg * tiles;
  const int tyi = t2 / a.tiles_z;
  const int tzi = t2 - tyi * a.tiles_z;
  const int zb = tzi * TZ, yb = tyi * TY;
  const int xs = seg * a.seg_len;
  const int xe = min(xs + a.seg_len, a.nx);

  const int tid = threadIdx.x;
  const int tz = tid % NTZ, ty = tid / NTZ;
  const int z = zb + tz * VZ, y = yb + ty;

  // Plane bases point at (row -3, column -4) of the interior so every per-thread offset
  // within an x-plane is a non-negative 32-bit value; with a uniform 64-bit base that
  // lets the loads and stores use the scalar-base + 32-bit-offset addressing form.
  const long long sx = a.sx, sy = a.sy;
  const long long rb = a.base - 3 * sy - 4;
  const C *__restrict__ pin = reinterpret_cast<const C *>(a.psi_in) + rb;
  // in stages 2, 3 psi_out and w are the same buffer (read, then rewritten, pointwise)
  C *pout = reinterpret_cast<C *>(a.psi_out) + rb;
  C *pw = reinterpret_cast<C *>(a.w) + rb;

  // Imaginary time: stage 1 reads psi as scale*psi, stage 2 forms the same product of
  // psi_{s-1}; the host sets scale / prev_scale only for those stages.  Instances with the
  // stage kind fixed carry only the one they use (a multiply by 1 is exact, so the other
  // instances give the same results).
  // SK >= 0: the stage kind is fixed at compile time (KIND; with SK_STORED the stored-w
  // scheme, else w rebuilt), which removes the run-time stage branches from the plane loop
  constexpr bool FIX = SK >= 0;
  constexpr bool STO = FIX && (SK & SK_STORED) != 0;
  constexpr int KIND = FIX ? (SK & 7) : -1;
  constexpr bool USE_SC = IMAG && (!FIX || KIND == ST_RK1);
  constexpr bool USE_PSC = IMAG && (!FIX || KIND == ST_RK2);
  T sc = T(1), psc = T(1);
  T c0 = T(a.c0);
  if constexpr (IMAG) {
    if constexpr (USE_SC) {
      if (a.scale) sc = T(*a.scale);
    }
    if constexpr (USE_PSC) {
      if (a.prev_scale) psc = T(*a.prev_scale);
    }
    if (a.c0dev) c0 = T(*a.c0dev);
  }

  // Own column offset within a plane.  Rows/columns beyond the interior read zeros from
  // the padding; their results are never stored.
  const uint32_t own = (uint32_t)(y + 3) * (uint32_t)sy + (uint32_t)(z + 4);

  // Halo assignment: vectors [0, 6*NTZ) are y-halo rows, the rest z-halo columns.
  uint32_t hoff = 0;
  int hrow = 0, hcol = 0;
  const bool hl = tid < NHALO;
  if (hl) {
    if (tid < 6 * NTZ) {
      const int r = tid / NTZ, cz = tid - r * NTZ;
      const int dy = r < 3 ? r - 3 : TY + r - 3;
      hrow = 3 + dy;
      hcol = 4 + cz * VZ;
      hoff = (uint32_t)(yb + dy + 3) * (uint32_t)sy + (uint32_t)(zb + cz * VZ + 4);
    } else {
      const int q = tid - 6 * NTZ;            // [0, 2*TY*HZV)
      const int side = q / (TY * HZV);
      const int r2 = q - side * TY * HZV;
      const int row = r2 / HZV, v = r2 - row * HZV;
      hrow = 3 + row;
      const int zc = side == 0 ? (-4 + v * VZ) : (TZ + v * VZ);
      hcol = 4 + zc;
      hoff = (uint32_t)(yb + row + 3) * (uint32_t)sy + (uint32_t)(zb + zc + 4);
    }
  }

  constexpr bool UNI = TBEC_UNIFORM_VMEM && sizeof(T) == 4 && FIX && !STO && KIND == ST_RK1;
  // Prefetch distance of the own points: plane x+3+AH is loaded in the iteration of plane
  // x (the halo of plane x+1 always one ahead).  AH2: two planes ahead, in the first RK
  // stage of the real-time analytic-trap instances, which have the registers for it
  // (512^3 FP32 on (64,32,2): 111 -> 120 VGPRs; the step 1.5351 -> 1.5236 ms, three
  // alternating runs each on one box; three ahead, at 128 VGPRs, gains nothing more, and
  // the halo two ahead as well spills).  The other FP32 instances would spill; the FP64
  // first stage fits (98 -> 110 VGPRs) but its step was 0.5% slower (3.052 -> 3.067 ms),
  // and the rotating one (127 VGPRs) measured the same as one ahead (1.693 / 1.698 ms).
  // Stages 2 and 3 two ahead fit only without their shared first differences (125
  // VGPRs) and were slower: 1.5417 -> 1.5632 ms per step (without the shared
  // differences alone, 1.5557).
  constexpr bool AH2 = TBEC_AHEAD2 && UNI && !IMAG && PHYS == PH_HARM;
  constexpr int AH = AH2 ? 2 : 1;
  // Register queue of QN plane slots: plane p lives in slot (p - xs + 3) % QN, and the
  // halo of plane p in hb[(p - xs) & 1].  The loop is unrolled by QN so every slot index
  // is a compile-time constant (no register moves when the queue advances; only the live
  // slots take registers) and the LDS buffer / halo parity is static too (QN even).
  constexpr int QN = AH2 ? 10 : 8;
  C q[QN][VZ];
  C hb[2][VZ];
  // buffer resources for x-plane `x` of the three arrays: base = the plane's row -3,
  // column -4; record count = one plane (sx complex values, < 2^31 bytes: checked by
  // the host)
  // Each stream keeps a pointer to the plane it touches next and advances it by one
  // plane (sx) per plane: a 64-bit scalar add, no multiply by the loop-varying x.
  const int pbytes = (int)(sx * (long long)sizeof(C));
  auto rs = [&](const C *plane) { return plane_rsrc(plane, pbytes); };
  const C *p_own = pin + (long long)(xs + 3 + AH) * sx;   // own points of plane x+3+AH
  const C *p_halo = pin + (long long)(xs + 1) * sx;       // halo of plane x+1
  C *p_wnext = pw + (long long)(xs + 1) * sx;          // w (or psi_{s-1}) of plane x+1
  C *p_out = pout + (long long)xs * sx;                // outputs of plane x
  C *p_w = pw + (long long)xs * sx;
  // UNI: every vector memory instruction of the plane loop is issued unconditionally.
  // Lanes with no halo vector issue the halo load at an offset past the plane's record
  // count (it returns zeros and touches no memory); the w prefetch runs on the last
  // plane too (plane x+1 <= nx exists); lanes that store nothing store past the record
  // count.  Each iteration then issues the same number of them on every path, and the
  // compiler's in-order wait counts let the loads of the next planes stay in flight
  // across the barrier; with a branch around a load or store, the wait before the
  // barrier also waits for the prefetch issued just before it.  It pays in the first RK
  // stage, the one with the most arithmetic per byte: 512^3 FP32 stage 1 472 -> 437 us,
  // rotating 525 -> 469 us (A/B on one box, kernel traces).  Stages 2 and 3, at the HBM
  // rate for their mix, are 1.8% slower with it (the extra address work of the dummy
  // halo loads), and the stored-w scheme and FP64 slower still, so they keep the branches.
  const uint32_t ownb = own * (uint32_t)sizeof(C);
  const uint32_t hoffb = (UNI && !hl) ? kOffNone : hoff * (uint32_t)sizeof(C);
  auto scale_in = [&](C *o) {
    if constexpr (USE_SC) {
#pragma unroll
      for (int v = 0; v < VZ; ++v) { o[v].re *= sc; o[v].im *= sc; }
    }
  };
  auto ldo = [&](__amdgpu_buffer_rsrc_t r, uint32_t off, C *o) {   // own points
    BufIO<T, VZ>::template ld<kOwnLoadCpol>(r, off, o);
    scale_in(o);
  };
  auto ldh = [&](__amdgpu_buffer_rsrc_t r, uint32_t off, C *o) {   // halo points
    BufIO<T, VZ>::template ld<kHaloLoadCpol>(r, off, o);
    scale_in(o);
  };
#pragma unroll
  for (int k = 0; k < 6 + AH; ++k) {
    // planes past xe+2 are not used; more than one ahead they can lie past plane nx+3
    const bool use = AH == 1 || xs - 3 + k <= xe + 2;
    ldo(rs(pin + (long long)(use ? xs - 3 + k : 0) * sx), use ? ownb : kOffNone, q[k]);
  }
  if (UNI || hl) ldh(rs(pin + (long long)xs * sx), hoffb, hb[0]);
  // Centred sums share first differences between neighbouring points (below), in the
  // instances with the stage kind fixed, and there in the analytic trap, in the tabulated
  // potential in real time, and in the first RK stage of every physics: stages 2 and 3 of
  // the rotating and of the imaginary-time tabulated instances spill 12-20 B with the
  // extra registers, and the run-time-kind instances more.  The first stages of the
  // rotating frame joined once the explicit FMAs had lowered their registers (real-time
  // Omega = 0.7 step 1.6699 -> 1.6533 ms, 256^3 rotating imaginary time 233.0 -> 228.9
  // us per step; real-time results bitwise unchanged).  xd1: x_i - x_{i-1} of the previous
  // plane, per point (plane xs: slots 3 and 2).
  constexpr bool SHARE_D1 = SUM == SUM_CENTRED && FIX &&
      (PHYS == PH_HARM || (PHYS == PH_GEN && !IMAG) || KIND == ST_RK1);
  C xd1[VZ];
  if constexpr (SHARE_D1) {
#pragma unroll
    for (int v = 0; v < VZ; ++v) xd1[v] = {q[3][v].re - q[2][v].re, q[3][v].im - q[2][v].im};
  }
  // w of the first plane (stages that read w); later planes are prefetched one ahead
  C wq[2][VZ];
  if (stage_reads_w(FIX ? KIND : a.stage))
    BufIO<T, VZ>::template ld<kPointwiseLoadCpol>(rs(p_w), ownb, wq[0]);
  // tabulated potential (PH_GEN): own points of plane x, see vq below
  constexpr bool VT = PHYS == PH_GEN;
  const int vbytes = (int)(sx * (long long)sizeof(T));
  const T *p_v = VT ? reinterpret_cast<const T *>(a.vtab) + rb + (long long)xs * sx : nullptr;
  const uint32_t ownv = own * (uint32_t)sizeof(T);
  // planes x .. x+VA of it in a ring of VS slots (plane x in slot R % VS, QN % VS == 0:
  // the instances that read it keep QN = 8).  In real
  // time the potential, a fourth stream, is prefetched two planes ahead (512^3 FP32: -1%
  // per step); the imaginary-time instances keep one plane (two would spill)
  constexpr int VA = IMAG ? 1 : 2, VS = VA == 1 ? 2 : 4;
  static_assert(!VT || QN % VS == 0, "potential ring must divide the unroll");
  T vq[VS][VZ];
  if constexpr (VT) {
#pragma unroll
    for (int k = 0; k < VA; ++k)
      if (xs + k < xe)
        BufIOR<T, VZ>::template ld<kPointwiseLoadCpol>(plane_rsrc(p_v + k * sx, vbytes), ownv,
                                                       vq[k]);
  }
  const int drift = VT ? a.drift : 0;
  const T vx2 = T(a.vx2);

  const T k2x = T(a.k2[0]), k2y = T(a.k2[1]), k2z = T(a.k2[2]);
  const T k1x = T(a.k1[0]), k1y = T(a.k1[1]);
  const T wlx = T(a.wl[0]), wly = T(a.wl[1]), wlz = T(a.wl[2]);
  const T gnl = T(a.gnl), om = T(a.om);
  const T cre = T(a.cre), cim = T(a.cim), A = T(a.A), B = T(a.B);
  const T wrec = T(a.wrec);
  const bool rebuild = FIX ? (!STO && (KIND == ST_RK2 || KIND == ST_RK3)) : a.wrec != 0.0;
  const bool store_w = FIX ? (STO && (KIND == ST_RK1 || KIND == ST_RK2)) : a.store_w != 0;
  const int stage = FIX ? KIND : a.stage;
  const T yc = T(a.y0 + a.hy * y);
  T zc2[VZ];
#pragma unroll
  for (int v = 0; v < VZ; ++v) {
    const T zc = T(a.z0 + a.hz * (z + v));
    zc2[v] = wlz * (zc * zc);
  }
  const T poty = wly * (yc * yc);
  const T omyk = om * yc * k1x;     // FP32 frame term: om * y / (60 hx)
  bool zin[VZ];
  bool full = true;
#pragma unroll
  for (int v = 0; v < VZ; ++v) {
    zin[v] = (z + v) < a.nz;
    full = full && zin[v];
  }
  const bool yin = y < a.ny;
  // store offset: this thread's vector where all of it is interior, else past the record
  // count (dropped); the partial vectors at the z edge are stored element by element
  const uint32_t ownst = (yin && full) ? ownb : kOffNone;

  auto plane = [&](auto Rc, const int x) {
    constexpr int R = decltype(Rc)::value;   // (x - xs) % QN
    constexpr int buf = R & 1;
    // prefetch plane x+3+AH (own) and the halo of plane x+1.  Two ahead, the own plane
    // x+5 is past the padded layout's last plane (nx+3) at x = nx-1, and past what the
    // segment uses (plane xe+2) from x = xe-2: those loads are issued at an offset past the
    // record count on a plane that exists (zeros, no memory access).
    if constexpr (AH2) {
      const bool use = x + 3 + AH <= xe + 2;
      ldo(rs(use ? p_own : pin), use ? ownb : kOffNone, q[(R + 6 + AH) % QN]);
    } else {
      ldo(rs(p_own), ownb, q[(R + 6 + AH) % QN]);
    }
    if (UNI || hl) ldh(rs(p_halo), hoffb, hb[buf ^ 1]);
    // (plane x+1 <= nx and plane x+VA <= nx+1 exist: 4 padding planes above the interior)
    if (stage_reads_w(stage) && (UNI || x + 1 < xe))
      BufIO<T, VZ>::template ld<kPointwiseLoadCpol>(rs(p_wnext), ownb, wq[buf ^ 1]);
    if constexpr (VT) {
      if (UNI || x + VA < xe)
        BufIOR<T, VZ>::template ld<kPointwiseLoadCpol>(plane_rsrc(p_v + VA * sx, vbytes),
                                                       ownv, vq[(R + VA) % VS]);
    }
    const C *wv = wq[buf];

    // stage centre plane + halo into LDS
    {
      C *row = &lds[buf][3 + ty][4 + tz * VZ];
#pragma unroll
      for (int v = 0; v < VZ; ++v) row[v] = q[(R + 3) % QN][v];
      if (hl) {
        C *hr = &lds[buf][hrow][hcol];
#pragma unroll
        for (int v = 0; v < VZ; ++v) hr[v] = hb[buf][v];
      }
    }
    __syncthreads();

    const T xc = T(a.x0 + a.hx * x);
    const T potxy = wlx * (xc * xc) + poty;
    const T omxk = om * xc * k1y;   // FP32 frame term: om * x / (60 hy)
    // z row segment: columns tz*VZ .. tz*VZ+VZ+7  (z-4 .. z+VZ+3)
    C zr[VZ + 8];
    {
      const C *row = &lds[buf][3 + ty][tz * VZ];
#pragma unroll
      for (int k = 0; k < VZ + 8; ++k) zr[k] = row[k];
    }
    C out[VZ], wo[VZ];
    C zd_prev;   // centred sums: point v-1's (z_v - z_{v-1}) (see below)
    // FP32 instances of the last imaginary-time stage sum the VZ values of |psi|^2 of a
    // plane in FP32 and add that to the FP64 accumulator once per plane: the norm only
    // sets the renormalisation scale of an FP32 psi (the energy sums stay FP64 per point)
    constexpr bool PN32 = IMAG && sizeof(T) == 4 && KIND == ST_RK3;
    T pn = T(0);
#pragma unroll
    for (int v = 0; v < VZ; ++v) {
      const int col = 4 + tz * VZ + v;
      const C ym3 = lds[buf][ty][col], ym2 = lds[buf][ty + 1][col],
              ym1 = lds[buf][ty + 2][col], yp1 = lds[buf][ty + 4][col],
              yp2 = lds[buf][ty + 5][col], yp3 = lds[buf][ty + 6][col];
      const C xm3 = q[R % QN][v], xm2 = q[(R + 1) % QN][v], xm1 = q[(R + 2) % QN][v];
      const C f = q[(R + 3) % QN][v];
      const C xp1 = q[(R + 4) % QN][v], xp2 = q[(R + 5) % QN][v], xp3 = q[(R + 6) % QN][v];
      const C zm3 = zr[v + 1], zm2 = zr[v + 2], zm1 = zr[v + 3], zp1 = zr[v + 5],
              zp2 = zr[v + 6], zp3 = zr[v + 7];
      T rr, ri;
      if constexpr (SHARE_D1) {
        // First differences shared with a neighbour instead of recomputed: x_{i-1} - x_i
        // is minus the previous plane's x_i - x_{i-1} (carried in xd1), and within a
        // thread's z run z_{v-1} - z_v is minus point v-1's z_v - z_{v-1}.  IEEE gives
        // a - b = -(b - a) exactly (up to the sign of a zero difference), so every
        // difference and pair sum is the one d2sum<SUM_CENTRED> forms; the weighted sums
        // after them are the same expressions, though the compiler may contract them into
        // FMAs differently (FP32 error against FP64 at 128^3 after 200 steps: 4.1387e-7
        // before, 4.1335e-7 with the shared differences).  Two subtractions fewer per
        // point (x) and one per pair of points (z): 512^3 FP32 1.5765 -> 1.5640 ms/step.
        const C xdp1 = {xp1.re - f.re, xp1.im - f.im};
        const C xdm1 = {-xd1[v].re, -xd1[v].im};
        xd1[v] = xdp1;
        C zdp1 = {zp1.re - f.re, zp1.im - f.im};
        C zdm1 = {zm1.re - f.re, zm1.im - f.im};
        if (v + 1 < VZ) {
          const C fn = q[(R + 3) % QN][(v + 1) % VZ];
          zdp1 = {fn.re - f.re, fn.im - f.im};
        }
        if (v > 0) zdm1 = {-zd_prev.re, -zd_prev.im};
        zd_prev = zdp1;
        const T dxr = d2diff<T>(xdm1.re, xdp1.re, xm2.re - f.re, xp2.re - f.re, xm3.re - f.re,
                                xp3.re - f.re);
        const T dxi = d2diff<T>(xdm1.im, xdp1.im, xm2.im - f.im, xp2.im - f.im, xm3.im - f.im,
                                xp3.im - f.im);
        const T dyr = d2diff<T>(ym1.re - f.re, yp1.re - f.re, ym2.re - f.re, yp2.re - f.re,
                                ym3.re - f.re, yp3.re - f.re);
        const T dyi = d2diff<T>(ym1.im - f.im, yp1.im - f.im, ym2.im - f.im, yp2.im - f.im,
                                ym3.im - f.im, yp3.im - f.im);
        const T dzr = d2diff<T>(zdm1.re, zdp1.re, zm2.re - f.re, zp2.re - f.re, zm3.re - f.re,
                                zp3.re - f.re);
        const T dzi = d2diff<T>(zdm1.im, zdp1.im, zm2.im - f.im, zp2.im - f.im, zm3.im - f.im,
                                zp3.im - f.im);
        rr = fma2(dxr, k2x, dyr, k2y) + dzr * k2z;
        ri = fma2(dxi, k2x, dyi, k2y) + dzi * k2z;
      } else {
        const T dxr = d2sum<T, SUM>(xm3.re, xm2.re, xm1.re, f.re, xp1.re, xp2.re, xp3.re);
        const T dxi = d2sum<T, SUM>(xm3.im, xm2.im, xm1.im, f.im, xp1.im, xp2.im, xp3.im);
        const T dyr = d2sum<T, SUM>(ym3.re, ym2.re, ym1.re, f.re, yp1.re, yp2.re, yp3.re);
        const T dyi = d2sum<T, SUM>(ym3.im, ym2.im, ym1.im, f.im, yp1.im, yp2.im, yp3.im);
        const T dzr = d2sum<T, SUM>(zm3.re, zm2.re, zm1.re, f.re, zp1.re, zp2.re, zp3.re);
        const T dzi = d2sum<T, SUM>(zm3.im, zm2.im, zm1.im, f.im, zp1.im, zp2.im, zp3.im);
        rr = fma2(dxr, k2x, dyr, k2y) + dzr * k2z;
        ri = fma2(dxi, k2x, dyi, k2y) + dzi * k2z;
      }
      if (PHYS == PH_ROT || drift == 1) {
        if constexpr (sizeof(T) == 4) {
          // FP32: om*x/(60 hy) and om*y/(60 hx) folded into one factor each, so the
          // frame term is two FMAs per component (fewer live registers)
          const T s1yr = d1sum<T>(ym3.re, ym2.re, ym1.re, yp1.re, yp2.re, yp3.re);
          const T s1yi = d1sum<T>(ym3.im, ym2.im, ym1.im, yp1.im, yp2.im, yp3.im);
          const T s1xr = d1sum<T>(xm3.re, xm2.re, xm1.re, xp1.re, xp2.re, xp3.re);
          const T s1xi = d1sum<T>(xm3.im, xm2.im, xm1.im, xp1.im, xp2.im, xp3.im);
          rr = fmaT(omyk, s1xi, fmaT(-omxk, s1yi, rr));   // (0, om) * (x D1y - y D1x)
          ri = fmaT(-omyk, s1xr, fmaT(omxk, s1yr, ri));
        } else {
          const T d1yr = d1sum<T>(ym3.re, ym2.re, ym1.re, yp1.re, yp2.re, yp3.re) * k1y;
          const T d1yi = d1sum<T>(ym3.im, ym2.im, ym1.im, yp1.im, yp2.im, yp3.im) * k1y;
          const T d1xr = d1sum<T>(xm3.re, xm2.re, xm1.re, xp1.re, xp2.re, xp3.re) * k1x;
          const T d1xi = d1sum<T>(xm3.im, xm2.im, xm1.im, xp1.im, xp2.im, xp3.im) * k1x;
          const T pr = fma2(xc, d1yr, -yc, d1xr);
          const T pim = fma2(xc, d1yi, -yc, d1xi);
          rr -= om * pim;   // (0, om) * (pr, pim)
          ri += om * pr;
        }
      } else if (VT && drift == 2) {
        const T d1xr = d1sum<T>(xm3.re, xm2.re, xm1.re, xp1.re, xp2.re, xp3.re) * k1x;
        const T d1xi = d1sum<T>(xm3.im, xm2.im, xm1.im, xp1.im, xp2.im, xp3.im) * k1x;
        rr -= vx2 * d1xi;  // (0, 2 vx) * (d1xr, d1xi)
        ri += vx2 * d1xr;
      }
      const T n2 = fma2(f.re, f.re, f.im, f.im);
      T pot;
      if constexpr (VT) pot = vq[R % VS][v];
      else pot = potxy + zc2[v];
      const T fac = (c0 - gnl * n2) - pot;
      rr += fac * f.re;
      ri += fac * f.im;

      const T ur = fma2(cre, rr, -cim, ri), ui = fma2(cre, ri, cim, rr);
      if (stage == ST_RHS) {
        wo[v] = {rr, ri};
      } else if (stage == ST_RK1 || stage == ST_EULER) {
        wo[v] = {ur, ui};
        out[v] = {f.re + B * ur, f.im + B * ui};
      } else {
        C wp = wv[v];
        if (rebuild) {
          if constexpr (USE_PSC) { wp.re *= psc; wp.im *= psc; }
          wp = {(f.re - wp.re) * wrec, (f.im - wp.im) * wrec};
        }
        wo[v] = {A * wp.re + ur, A * wp.im + ui};
        out[v] = {f.re + B * wo[v].re, f.im + B * wo[v].im};
      }
      if constexpr (IMAG) {
        if (yin && zin[v]) {
          if (stage == ST_RK1 || stage == ST_RHS) {
            red[0] += (double)n2;
            red[1] += fma2((double)f.re, (double)rr, (double)f.im, (double)ri);
            red[2] += (double)n2 * (double)n2;
            red[3] += fma2((double)rr, (double)rr, (double)ri, (double)ri);
          } else if (stage == ST_RK3 || stage == ST_EULER) {
            if constexpr (PN32)
              pn += fma2(out[v].re, out[v].re, out[v].im, out[v].im);
            else
              red[0] += fma2((double)out[v].re, (double)out[v].re, (double)out[v].im,
                             (double)out[v].im);
          }
        }
      }
    }
    if constexpr (PN32) red[0] += (double)pn;
    // stores (interior points only; padding stays zero)
    if constexpr (UNI) {
      if (stage == ST_RHS) {
        BufIO<T, VZ>::template st<kStoreCpol>(rs(p_w), ownst, wo);
      } else {
        BufIO<T, VZ>::template st<kStoreCpol>(rs(p_out), ownst, out);
        if (store_w) BufIO<T, VZ>::template st<kStoreCpol>(rs(p_w), ownst, wo);
      }
    }
    if (yin && !(UNI && full)) {
      C *po = p_out;   // uniform plane bases
      C *pwx = p_w;
      if (!UNI && full) {
        if (stage == ST_RHS) {
          BufIO<T, VZ>::template st<kStoreCpol>(rs(p_w), ownb, wo);
        } else {
          BufIO<T, VZ>::template st<kStoreCpol>(rs(p_out), ownb, out);
          if (store_w) BufIO<T, VZ>::template st<kStoreCpol>(rs(p_w), ownb, wo);
        }
      } else {
#pragma unroll
        for (int v = 0; v < VZ; ++v) {
          if (!zin[v]) continue;
          if (stage == ST_RHS) {
            pwx[own + v] = wo[v];
          } else {
            po[own + v] = out[v];
            if (store_w) pwx[own + v] = wo[v];
          }
        }
      }
    }
    p_own += sx;
    p_halo += sx;
    p_wnext += sx;
    p_out += sx;
    p_w += sx;
    if constexpr (VT) p_v += sx;
  };

  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  using I5 = std::integral_constant<int, 5>;
  using I6 = std::integral_constant<int, 6>;
  using I7 = std::integral_constant<int, 7>;
  using I8 = std::integral_constant<int, 8>;
  using I9 = std::integral_constant<int, 9>;
  for (int x = xs; x < xe; x += QN) {
    plane(I0{}, x);
    if (x + 1 < xe) plane(I1{}, x + 1);
    if (x + 2 < xe) plane(I2{}, x + 2);
    if (x + 3 < xe) plane(I3{}, x + 3);
    if (x + 4 < xe) plane(I4{}, x + 4);
    if (x + 5 < xe) plane(I5{}, x + 5);
    if (x + 6 < xe) plane(I6{}, x + 6);
    if (x + 7 < xe) plane(I7{}, x + 7);
    if constexpr (QN == 10) {
      if (x + 8 < xe) plane(I8{}, x + 8);
      if (x + 9 < xe) plane(I9{}, x + 9);
    }
  }
}

// Minimum waves per SIMD asked of the compiler.  512-thread FP32 kernels are held to
// 128 VGPRs so that two workgroups share a CU: the rotating and the imaginary-time
// variants need 130-145 otherwise, which halves the resident waves (measured: 20-30%
// slower even where the floor makes the compiler spill a few registers).  FP64 kernels
// would spill heavily and keep the compiler's choice.
template <typename T, int NT, bool IMAG> constexpr int stage_min_waves() {
  return (sizeof(T) == 4 && NT == 512) ? 4 : 1;
}

template <typename T, int TZ, int TY, int VZ, int SUM, int PHYS, bool IMAG, int SK = -1>
__global__ void __launch_bounds__((TZ / VZ) * TY, (stage_min_waves<T, (TZ / VZ) * TY, IMAG>()))
gpe_stage(const StageArgs a) {
  constexpr int NT = (TZ / VZ) * TY;
  if constexpr (IMAG) {
    if (a.skip && *a.skip) return;
  }
  // XCD-aware tile order: hardware id b lands on XCD b % 8; give XCD k the logical
  // range [k*per, (k+1)*per) so neighbouring tiles share that XCD's L2.
  const int nb = a.nblocks;
  const int b = blockIdx.x;
  int lb = b;
  if (a.remap && (nb & 7) == 0) lb = (b & 7) * (nb >> 3) + (b >> 3);
  double red[4] = {0.0, 0.0, 0.0, 0.0};
  march_tile<T, TZ, TY, VZ, SUM, PHYS, IMAG, SK>(a, lb, red);
  if constexpr (IMAG) {
    if (a.partials) block_sum4<NT>(red, a.partials + 4LL * lb);
  }
}

// Many condensates in one launch.  Work items are (instance, tile) pairs, one per
// workgroup.  Per-instance static arguments live in a device table; the launch supplies
// the stage and the buffer parity.
struct BatchInst {
  StageArgs a;          // psi_in/psi_out/w, stage fields are filled per item
  void *psi[2];
  double dt_re, dt_im;  // per-instance complex step
  int item_begin;       // first work item of this instance
  int pad_;
};

struct BatchLaunch {
  const BatchInst *inst;
  const int *item_inst; // work item -> instance
  const int *items;     // workgroup -> work item: the items of the unfinished instances
  int total_items;      // workgroups in the launch (entries of items)
  int parity;           // psi[parity] is read, psi[parity ^ 1] written
  int stage;            // StageKind
  double A, B;
  double *partials;     // 4 doubles per work item
  const double *state;  // per-instance state rows (stride state_stride)
  int state_stride;
  const int *flags;     // per-instance flags (stride 4); skip when flags[4*i] != 0
};

template <typename T, int TZ, int TY, int VZ, int SUM, int PHYS, bool IMAG, int SK = -1>
__global__ void __launch_bounds__((TZ / VZ) * TY, (stage_min_waves<T, (TZ / VZ) * TY, IMAG>()))
gpe_stage_batch(const BatchLaunch L) {
  constexpr int NT = (TZ / VZ) * TY;
  // one work item per workgroup (the grid is total_items): no loop state stays live
  // across the plane march.  The item list holds the unfinished instances' items only
  // (the host compacts it as instances finish), so a tail of a few slow instances does
  // not launch the finished ones' workgroups; partials stay at the item's own index.
  if ((int)blockIdx.x >= L.total_items) return;
  const int item = L.items[blockIdx.x];
  {
    const int i = L.item_inst[item];
    if (L.flags[4 * i] != 0) return;            // converged or failed: skip (uniform)
    const BatchInst &bi = L.inst[i];
    StageArgs a = bi.a;
    a.psi_in = bi.psi[L.parity];
    a.psi_out = bi.psi[L.parity ^ 1];
    a.stage = L.stage;
    a.A = L.A;
    a.B = L.B;
    if (L.stage == ST_EULER) { a.cre = 0.5 * bi.dt_re; a.cim = 0.5 * bi.dt_im; }
    else { a.cre = -0.5 * bi.dt_im; a.cim = 0.5 * bi.dt_re; }
    const double *st = L.state + (long long)i * L.state_stride;
    a.scale = (IMAG && L.stage == ST_RK1) ? st + 0 : nullptr;   // ST_SCALE
    a.prev_scale = (IMAG && L.stage == ST_RK2) ? st + 0 : nullptr;
    a.w = a.psi_out;                     // w rebuilt from the psi pair (stages 2, 3)
    a.wrec = L.stage == ST_RK2 ? 3.0 : (L.stage == ST_RK3 ? 1.0 : 0.0);
    a.store_w = 0;
    a.c0dev = IMAG ? st + 1 : nullptr;                            // ST_C0EFF
    a.skip = nullptr;
    double red[4] = {0.0, 0.0, 0.0, 0.0};
    march_tile<T, TZ, TY, VZ, SUM, PHYS, IMAG, SK>(a, item - bi.item_begin, red);
    if constexpr (IMAG) block_sum4<NT>(red, L.partials + 4LL * item);
  }
}

}  // namespace tbec
