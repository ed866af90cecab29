// Learned warm start: a small MLP (inputs -> depth x width GELU -> outputs) trained and
// evaluated on the GPU with hand-written BF16 MFMA GEMMs that accumulate in FP32, plus
// the trilinear upsampling that turns its coarse prediction into a starting state.
// No reference counterpart (the reference has no learned component).
//
// GEMMs on v_mfma_f32_16x16x32_bf16, BF16 operands, FP32 accumulation:
// - k_gemm_bt, for every product whose operands are both contiguous in the reduction
//   index (forward C = H W, and dH = dZ W^T): the weight operand is a BF16 copy of the
//   weights kept in the layout that makes it so (W^T for the forward, W itself for dH),
//   refreshed once after every parameter update; activations are FP32 rounded to BF16
//   (nearest-even) when staged.  A 64x128 output tile per 256-thread workgroup (4 waves of
//   32x64 = 2x4 fragments), K in steps of 32 through double-buffered LDS (27.6 KB, so
//   four workgroups share a CU) with 16-byte global loads, the next step's loads in
//   flight while the current step's 8 MFMAs per wave run, one barrier per step.
// - k_gemm, the generic strided form (dW = H^T dZ, where the reduction runs over the
//   batch, and the 5-wide input layer): a 64x64 tile, operands rounded to BF16 when
//   staged.
// Long-K products are split over workgroups and the partials summed in a fixed order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "gpe_stage.hpp"
#include "tbec.h"
#include "tbec_internal.hpp"

using namespace tbec;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define MCHK(x)                                                                         \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess)                                                               \
      return set_error(TBEC_ERR_HIP, std::string(#x " failed: ") + hipGetErrorString(e_)); \
  } while (0)

namespace {

constexpr int BM = 64, BN = 64, BK = 32, KPAD = 8;

// FP32 -> BF16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float gelu(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

enum Epi { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_PARTIAL = 3 };

struct GemmArgs {
  const float *A;
  long long sam, sak;       // A[m][k] = A[m*sam + k*sak]
  const float *B;
  long long sbk, sbn;       // B[k][n] = B[k*sbk + n*sbn]
  float *C;                 // row-major M x N (ldc = N); EPI_PARTIAL: [split][M][N]
  float *Z;                 // EPI_BIAS_GELU: pre-activation output (may be null)
  const float *bias;
  int M, N, K;
  int ksplit;               // number of K ranges (gridDim.z)
  int epi;
};

__global__ void __launch_bounds__(256) k_gemm(const GemmArgs g) {
  __shared__ uint16_t As[BM][BK + KPAD];
  __shared__ uint16_t Bs[BN][BK + KPAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int kper = ((g.K + g.ksplit - 1) / g.ksplit + BK - 1) / BK * BK;
  const int kbeg = blockIdx.z * kper;
  const int kend = min(g.K, kbeg + kper);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int lr = tid >> 2, lk = (tid & 3) * 8;   // staging: row (m or n), k chunk
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    {
      const int m = m0 + lr;
      uint16_t v[8];
      for (int q = 0; q < 8; ++q) {
        const int k = k0 + lk + q;
        const float f = (m < g.M && k < kend) ? g.A[m * g.sam + k * g.sak] : 0.f;
        v[q] = f2bf(f);
      }
      for (int q = 0; q < 8; ++q) As[lr][lk + q] = v[q];
      const int n = n0 + lr;
      for (int q = 0; q < 8; ++q) {
        const int k = k0 + lk + q;
        const float f = (n < g.N && k < kend) ? g.B[k * g.sbk + n * g.sbn] : 0.f;
        v[q] = f2bf(f);
      }
      for (int q = 0; q < 8; ++q) Bs[lr][lk + q] = v[q];
    }
    __syncthreads();
    // fragments: lane l holds row/col (l % 16), k = (l / 16)*8 .. +7
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    bf16x8 af[2], bfr[2];
    for (int i = 0; i < 2; ++i)
      af[i] = *reinterpret_cast<const bf16x8 *>(&As[wm + i * 16 + fr][fk]);
    for (int j = 0; j < 2; ++j)
      bfr[j] = *reinterpret_cast<const bf16x8 *>(&Bs[wn + j * 16 + fr][fk]);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // D layout: lane l holds rows (l / 16)*4 + r, column l % 16 of each 16x16 fragment
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        const int n = n0 + wn + j * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        float v = acc[i][j][r];
        if (g.epi == EPI_PARTIAL) {
          g.C[((long long)blockIdx.z * g.M + m) * g.N + n] = v;
          continue;
        }
        if (g.epi == EPI_BIAS || g.epi == EPI_BIAS_GELU) v += g.bias[n];
        if (g.epi == EPI_BIAS_GELU) {
          if (g.Z) g.Z[(long long)m * g.N + n] = v;
          v = gelu(v);
        }
        g.C[(long long)m * g.N + n] = v;
      }
}

constexpr int TBM = 64, TBN = 128, TBK = 32, TPAD = 8;

struct GemmBtArgs {
  const float *A;           // A[m][k] = A[m*lda + k], FP32
  long long lda;
  const uint16_t *Bt;       // B[k][n] = Bt[n*ldb + k], BF16 bits
  long long ldb;
  float *C;                 // row-major M x N; EPI_PARTIAL: [split][M][N]
  float *Z;
  const float *bias;
  int M, N, K;
  int ksplit;
  int epi;
};

__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
  return (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
}

__global__ void __launch_bounds__(256) k_gemm_bt(const GemmBtArgs g) {
  __shared__ __align__(16) uint16_t As[2][TBM][TBK + TPAD];
  __shared__ __align__(16) uint16_t Bs[2][TBN][TBK + TPAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * TBM, n0 = blockIdx.x * TBN;
  const int kper = ((g.K + g.ksplit - 1) / g.ksplit + TBK - 1) / TBK * TBK;
  const int kbeg = blockIdx.z * kper;
  const int kend = min(g.K, kbeg + kper);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
  f32x4 acc[2][4];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // staging: A as 2 float4 per thread (row idx/8, k (idx%8)*4), B as 2 x 16 bytes per
  // thread (row idx/4, k (idx%4)*8); elements outside [M] x [N] x [kbeg, kend) are zero
  float4 ra[2];
  uint4 rb[2];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int idx = q * 256 + tid, r = idx >> 3, k = k0 + (idx & 7) * 4, m = m0 + r;
      if (m < g.M && k + 3 < kend) {
        ra[q] = *reinterpret_cast<const float4 *>(g.A + (long long)m * g.lda + k);
      } else {
        float v[4];
        for (int e = 0; e < 4; ++e)
          v[e] = (m < g.M && k + e < kend) ? g.A[(long long)m * g.lda + k + e] : 0.f;
        ra[q] = make_float4(v[0], v[1], v[2], v[3]);
      }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int idx = q * 256 + tid, r = idx >> 2, k = k0 + (idx & 3) * 8, n = n0 + r;
      if (n < g.N && k + 7 < kend) {
        rb[q] = *reinterpret_cast<const uint4 *>(g.Bt + (long long)n * g.ldb + k);
      } else {
        uint32_t w[4];
        for (int e = 0; e < 4; ++e) {
          const uint32_t lo = (n < g.N && k + 2 * e < kend)
                                  ? g.Bt[(long long)n * g.ldb + k + 2 * e] : 0u;
          const uint32_t hi = (n < g.N && k + 2 * e + 1 < kend)
                                  ? g.Bt[(long long)n * g.ldb + k + 2 * e + 1] : 0u;
          w[e] = lo | (hi << 16);
        }
        rb[q] = make_uint4(w[0], w[1], w[2], w[3]);
      }
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int idx = q * 256 + tid, r = idx >> 3, kq = (idx & 7) * 4;
      *reinterpret_cast<uint2 *>(&As[buf][r][kq]) =
          make_uint2(pack_bf16x2(ra[q].x, ra[q].y), pack_bf16x2(ra[q].z, ra[q].w));
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int idx = q * 256 + tid, r = idx >> 2, k8 = (idx & 3) * 8;
      *reinterpret_cast<uint4 *>(&Bs[buf][r][k8]) = rb[q];
    }
  };
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  int buf = 0;
  if (kbeg < kend) {
    gload(kbeg);
    lstore(0);
  }
  __syncthreads();
  for (int k0 = kbeg; k0 < kend; k0 += TBK) {
    const bool more = k0 + TBK < kend;
    if (more) gload(k0 + TBK);
#pragma unroll
    for (int kk = 0; kk < TBK; kk += 32) {
      bf16x8 af[2], bfr[4];
#pragma unroll
      for (int i = 0; i < 2; ++i)
        af[i] = *reinterpret_cast<const bf16x8 *>(&As[buf][wm + i * 16 + fr][kk + fk]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        bfr[j] = *reinterpret_cast<const bf16x8 *>(&Bs[buf][wn + j * 16 + fr][kk + fk]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
    if (more) lstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 4; ++j)
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        const int n = n0 + wn + j * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        float v = acc[i][j][r];
        if (g.epi == EPI_PARTIAL) {
          g.C[((long long)blockIdx.z * g.M + m) * g.N + n] = v;
          continue;
        }
        if (g.epi == EPI_BIAS || g.epi == EPI_BIAS_GELU) v += g.bias[n];
        if (g.epi == EPI_BIAS_GELU) {
          if (g.Z) g.Z[(long long)m * g.N + n] = v;
          v = gelu(v);
        }
        g.C[(long long)m * g.N + n] = v;
      }
}

// BF16 copies of one K x N weight matrix W (row-major): Wb (same layout) and WbT (N x K),
// through 64 x 64 LDS tiles so both writes are coalesced.
__global__ void __launch_bounds__(256) k_weights_bf16(const float *__restrict__ W,
                                                       uint16_t *__restrict__ Wb,
                                                       uint16_t *__restrict__ WbT, int K,
                                                       int N) {
  __shared__ uint16_t t[64][66];
  const int k0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  for (int i = threadIdx.x; i < 64 * 64; i += 256) {
    const int r = i >> 6, c = i & 63, k = k0 + r, n = n0 + c;
    if (k < K && n < N) {
      const uint16_t v = f2bf(W[(long long)k * N + n]);
      Wb[(long long)k * N + n] = v;
      t[r][c] = v;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 64 * 64; i += 256) {
    const int r = i >> 6, c = i & 63, n = n0 + r, k = k0 + c;
    if (k < K && n < N) WbT[(long long)n * K + k] = t[c][r];
  }
}

// C[m][n] = sum over s of P[s][m][n] in order s = 0, 1, ...
__global__ void k_sum_splits(const float *P, float *C, long long mn, int splits) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < mn;
       i += (long long)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int q = 0; q < splits; ++q) s += P[q * mn + i];
    C[i] = s;
  }
}

// dZ = dH * gelu'(Z)
__global__ void k_gelu_bwd(const float *dH, const float *Z, float *dZ, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dZ[i] = dH[i] * gelu_grad(Z[i]);
}

// db[n] = sum over rows of dZ[m][n] (rows summed in order)
__global__ void k_colsum(const float *dZ, float *db, int M, int N) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += dZ[(long long)m * N + n];
  db[n] = s;
}

// Mean-squared error: dY = 2 (Y - T) / count, per-block partial sums of (Y - T)^2.
__global__ void __launch_bounds__(256)
k_mse(const float *Y, const float *T, float *dY, long long n, double *part) {
  __shared__ double red[256];
  double s = 0.0;
  const float inv = 2.0f / (float)n;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float d = Y[i] - T[i];
    dY[i] = inv * d;
    s += (double)d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam with bias correction (FP32 master weights).
__global__ void k_adam(float *w, const float *g, float *m, float *v, long long n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// Gather rows idx[0..B) of a row-major [n x d] matrix into [B x d].
__global__ void k_gather_rows(const float *src, const int *idx, float *dst, int B, int d) {
  const long long n = (long long)B * d;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long r = i / d, c = i % d;
    dst[i] = src[(long long)idx[r] * d + c];
  }
}

int nblk(long long n) {
  long long b = (n + 255) / 256;
  return (int)std::max<long long>(1, std::min<long long>(b, 8192));
}

}  // namespace

struct tbec_mlp {
  int n_in = 0, width = 0, depth = 0, n_out = 0;
  std::vector<int> dims;              // layer sizes: n_in, width x depth, n_out
  std::vector<long long> woff, boff;  // parameter offsets per layer
  long long nparam = 0;
  float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  uint16_t *wb = nullptr, *wbt = nullptr;   // BF16 weights: W's layout and transposed
  int cap_batch = 0;
  std::vector<float *> H, Z;          // activations per layer (H[0] = input)
  float *dA = nullptr, *dB = nullptr; // gradient scratch (max layer width x batch)
  float *Y = nullptr, *dY = nullptr, *T = nullptr, *split = nullptr;
  long long split_elems = 0;
  double *part = nullptr, *h_part = nullptr;
  float *X_all = nullptr, *T_all = nullptr;
  int n_data = 0;
  int *d_idx = nullptr;
  long long adam_t = 0;
  hipStream_t stream = nullptr;
};

namespace {

int gemm(tbec_mlp *m, const float *A, long long sam, long long sak, const float *B,
         long long sbk, long long sbn, float *C, int M, int N, int K, int epi,
         const float *bias = nullptr, float *Z = nullptr) {
  GemmArgs g{A, sam, sak, B, sbk, sbn, C, Z, bias, M, N, K, 1, epi};
  const long long tiles = (long long)((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  int splits = 1;
  if (epi == EPI_STORE && tiles < 128 && K >= 4 * BK) {
    splits = (int)std::min<long long>(K / (2 * BK), 256 / std::max<long long>(tiles, 1));
    splits = std::max(1, splits);
  }
  if (splits > 1) {
    const long long need = (long long)splits * M * N;
    if (need > m->split_elems) {
      if (m->split) (void)hipFree(m->split);
      m->split = nullptr;
      MCHK(hipMalloc(&m->split, sizeof(float) * need));
      m->split_elems = need;
    }
    g.C = m->split;
    g.epi = EPI_PARTIAL;
    g.ksplit = splits;
  }
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, splits);
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, m->stream, g);
  MCHK(hipGetLastError());
  if (splits > 1) {
    hipLaunchKernelGGL(k_sum_splits, dim3(nblk((long long)M * N)), dim3(256), 0, m->stream,
                       m->split, C, (long long)M * N, splits);
    MCHK(hipGetLastError());
  }
  return 0;
}

// C = A * B with A FP32 row-major (lda) and B given as BF16 Bt[n][k] (ldb); split over K
// as gemm() does when the output has few tiles.
int gemm_bt(tbec_mlp *m, const float *A, long long lda, const uint16_t *Bt, long long ldb,
            float *C, int M, int N, int K, int epi, const float *bias = nullptr,
            float *Z = nullptr) {
  GemmBtArgs g{A, lda, Bt, ldb, C, Z, bias, M, N, K, 1, epi};
  const long long tiles = (long long)((M + TBM - 1) / TBM) * ((N + TBN - 1) / TBN);
  int splits = 1;
  // few output tiles over a long K (dH of the wide output layer): split K until about
  // 512 workgroups stream the weights, each over at least 4 K-steps
  if (epi == EPI_STORE && tiles < 128 && K >= 8 * TBK) {
    splits = (int)std::min<long long>(K / (4 * TBK), 512 / std::max<long long>(tiles, 1));
    splits = std::max(1, splits);
  }
  if (splits > 1) {
    const long long need = (long long)splits * M * N;
    if (need > m->split_elems) {
      if (m->split) (void)hipFree(m->split);
      m->split = nullptr;
      MCHK(hipMalloc(&m->split, sizeof(float) * need));
      m->split_elems = need;
    }
    g.C = m->split;
    g.epi = EPI_PARTIAL;
    g.ksplit = splits;
  }
  dim3 grid((N + TBN - 1) / TBN, (M + TBM - 1) / TBM, splits);
  hipLaunchKernelGGL(k_gemm_bt, grid, dim3(256), 0, m->stream, g);
  MCHK(hipGetLastError());
  if (splits > 1) {
    hipLaunchKernelGGL(k_sum_splits, dim3(nblk((long long)M * N)), dim3(256), 0, m->stream,
                       m->split, C, (long long)M * N, splits);
    MCHK(hipGetLastError());
  }
  return 0;
}

// k_gemm_bt needs both operands contiguous in k with 16-byte aligned rows.
bool bt_ok(const float *A, long long lda, long long ldb) {
  return ((uintptr_t)A & 15) == 0 && (lda & 3) == 0 && (ldb & 7) == 0;
}

// Refresh the BF16 weight copies from the FP32 master weights (after every update).
int refresh_bf16(tbec_mlp *m) {
  for (size_t l = 0; l + 1 < m->dims.size(); ++l) {
    const int K = m->dims[l], N = m->dims[l + 1];
    dim3 grid((N + 63) / 64, (K + 63) / 64);
    hipLaunchKernelGGL(k_weights_bf16, grid, dim3(256), 0, m->stream, m->params + m->woff[l],
                       m->wb + m->woff[l], m->wbt + m->woff[l], K, N);
  }
  MCHK(hipGetLastError());
  return 0;
}

int ensure_batch(tbec_mlp *m, int B) {
  if (B <= m->cap_batch) return 0;
  for (float *p : m->H) if (p) (void)hipFree(p);
  for (float *p : m->Z) if (p) (void)hipFree(p);
  for (float *p : {m->dA, m->dB, m->Y, m->dY, m->T}) if (p) (void)hipFree(p);
  if (m->part) (void)hipFree(m->part);
  if (m->h_part) (void)hipHostFree(m->h_part);
  if (m->d_idx) (void)hipFree(m->d_idx);
  const int L = (int)m->dims.size() - 1;
  m->H.assign(L, nullptr);
  m->Z.assign(L, nullptr);
  int maxw = 0;
  for (int d : m->dims) maxw = std::max(maxw, d);
  for (int l = 0; l < L; ++l) {
    MCHK(hipMalloc(&m->H[l], sizeof(float) * (size_t)B * m->dims[l]));
    MCHK(hipMalloc(&m->Z[l], sizeof(float) * (size_t)B * m->dims[l + 1]));
  }
  MCHK(hipMalloc(&m->dA, sizeof(float) * (size_t)B * maxw));
  MCHK(hipMalloc(&m->dB, sizeof(float) * (size_t)B * maxw));
  MCHK(hipMalloc(&m->Y, sizeof(float) * (size_t)B * m->n_out));
  MCHK(hipMalloc(&m->dY, sizeof(float) * (size_t)B * m->n_out));
  MCHK(hipMalloc(&m->T, sizeof(float) * (size_t)B * m->n_out));
  MCHK(hipMalloc(&m->part, sizeof(double) * 8192));
  MCHK(hipHostMalloc(&m->h_part, sizeof(double) * 8192));
  MCHK(hipMalloc(&m->d_idx, sizeof(int) * (size_t)B));
  m->cap_batch = B;
  return 0;
}

// Forward pass of a batch whose input is already in H[0]; output in m->Y.
int forward(tbec_mlp *m, int B) {
  const int L = (int)m->dims.size() - 1;
  for (int l = 0; l < L; ++l) {
    const int K = m->dims[l], N = m->dims[l + 1];
    const float *W = m->params + m->woff[l], *b = m->params + m->boff[l];
    const bool last = l == L - 1;
    float *out = last ? m->Y : m->H[l + 1];
    const int epi = last ? EPI_BIAS : EPI_BIAS_GELU;
    int rc;
    if (K >= TBK && bt_ok(m->H[l], K, K))   // B[k][n] = W[k][n] = WbT[n][k]
      rc = gemm_bt(m, m->H[l], K, m->wbt + m->woff[l], K, out, B, N, K, epi, b,
                   last ? nullptr : m->Z[l]);
    else
      rc = gemm(m, m->H[l], K, 1, W, N, 1, out, B, N, K, epi, b, last ? nullptr : m->Z[l]);
    if (rc) return rc;
  }
  return 0;
}

// Loss and parameter gradients for targets in m->T; returns the mean squared error.
int backward(tbec_mlp *m, int B, double *loss) {
  const int L = (int)m->dims.size() - 1;
  const long long ny = (long long)B * m->n_out;
  const int nb = std::min(nblk(ny), 8192);
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, m->stream, m->Y, m->T, m->dY, ny, m->part);
  MCHK(hipGetLastError());
  MCHK(hipMemcpyAsync(m->h_part, m->part, sizeof(double) * nb, hipMemcpyDeviceToHost,
                      m->stream));
  const float *dZ = m->dY;
  for (int l = L - 1; l >= 0; --l) {
    const int K = m->dims[l], N = m->dims[l + 1];
    float *gW = m->grads + m->woff[l], *gb = m->grads + m->boff[l];
    // dW = H^T dZ  (A = H^T: A[k][b] = H[b][k])
    int rc = gemm(m, m->H[l], 1, K, dZ, N, 1, gW, K, N, B, EPI_STORE);
    if (rc) return rc;
    hipLaunchKernelGGL(k_colsum, dim3((N + 255) / 256), dim3(256), 0, m->stream, dZ, gb, B, N);
    MCHK(hipGetLastError());
    if (l == 0) break;
    // dH = dZ W^T  (B = W^T: B[n][k] = W[k][n]); then dZ_prev = dH * gelu'(Z_prev)
    float *dH = (dZ == m->dA) ? m->dB : m->dA;
    if (N >= TBK && bt_ok(dZ, N, N))   // B[n][k] = W[k][n] = Wb[k][n]
      rc = gemm_bt(m, dZ, N, m->wb + m->woff[l], N, dH, B, K, N, EPI_STORE);
    else
      rc = gemm(m, dZ, N, 1, m->params + m->woff[l], 1, N, dH, B, K, N, EPI_STORE);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk((long long)B * K)), dim3(256), 0, m->stream, dH,
                       m->Z[l - 1], dH, (long long)B * K);
    MCHK(hipGetLastError());
    dZ = dH;
  }
  MCHK(hipStreamSynchronize(m->stream));
  double s = 0.0;
  for (int i = 0; i < nb; ++i) s += m->h_part[i];
  *loss = s / (double)ny;
  return 0;
}

// splitmix64 for deterministic initialisation and shuffles
uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int tbec_mlp_create(int32_t n_in, int32_t width, int32_t depth, int32_t n_out, uint64_t seed,
                    tbec_mlp **out) {
  if (!out || n_in < 1 || width < 1 || depth < 1 || n_out < 1)
    return set_error(TBEC_ERR_ARG, "bad MLP shape");
  *out = nullptr;
  tbec_mlp *m = new tbec_mlp();
  m->n_in = n_in; m->width = width; m->depth = depth; m->n_out = n_out;
  m->dims.push_back(n_in);
  for (int i = 0; i < depth; ++i) m->dims.push_back(width);
  m->dims.push_back(n_out);
  long long off = 0;
  for (size_t l = 0; l + 1 < m->dims.size(); ++l) {
    m->woff.push_back(off);
    off += (long long)m->dims[l] * m->dims[l + 1];
    m->boff.push_back(off);
    off += m->dims[l + 1];
  }
  m->nparam = off;
  auto bail = [&](int rc) { tbec_mlp_destroy(m); return rc; };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(set_error(TBEC_ERR_HIP, "stream creation failed"));
  const size_t bytes = sizeof(float) * (size_t)m->nparam;
  const size_t hbytes = sizeof(uint16_t) * (size_t)m->nparam;
  if (hipMalloc(&m->params, bytes) != hipSuccess || hipMalloc(&m->grads, bytes) != hipSuccess ||
      hipMalloc(&m->adam_m, bytes) != hipSuccess || hipMalloc(&m->adam_v, bytes) != hipSuccess ||
      hipMalloc(&m->wb, hbytes) != hipSuccess || hipMalloc(&m->wbt, hbytes) != hipSuccess)
    return bail(set_error(TBEC_ERR_OOM, "MLP parameter allocation failed"));
  // He-style uniform init: W ~ U(-a, a), a = sqrt(6 / fan_in); b = 0.
  std::vector<float> h(m->nparam, 0.f);
  uint64_t s = seed;
  for (size_t l = 0; l + 1 < m->dims.size(); ++l) {
    const double a = std::sqrt(6.0 / m->dims[l]);
    const long long nw = (long long)m->dims[l] * m->dims[l + 1];
    for (long long i = 0; i < nw; ++i) {
      const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
      h[m->woff[l] + i] = (float)((2.0 * u - 1.0) * a);
    }
  }
  if (hipMemcpy(m->params, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->adam_m, 0, bytes) != hipSuccess || hipMemset(m->adam_v, 0, bytes) != hipSuccess)
    return bail(set_error(TBEC_ERR_HIP, "MLP initialisation failed"));
  if (int rc = refresh_bf16(m)) return bail(rc);
  if (hipStreamSynchronize(m->stream) != hipSuccess)
    return bail(set_error(TBEC_ERR_HIP, "MLP initialisation failed"));
  *out = m;
  return 0;
}

int tbec_mlp_destroy(tbec_mlp *m) {
  if (!m) return 0;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (float *p : m->H) if (p) (void)hipFree(p);
  for (float *p : m->Z) if (p) (void)hipFree(p);
  for (float *p : {m->params, m->grads, m->adam_m, m->adam_v, m->dA, m->dB, m->Y, m->dY, m->T,
                   m->split, m->X_all, m->T_all})
    if (p) (void)hipFree(p);
  for (uint16_t *p : {m->wb, m->wbt})
    if (p) (void)hipFree(p);
  if (m->part) (void)hipFree(m->part);
  if (m->h_part) (void)hipHostFree(m->h_part);
  if (m->d_idx) (void)hipFree(m->d_idx);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

int64_t tbec_mlp_num_params(tbec_mlp *m) { return m ? m->nparam : -1; }

int tbec_mlp_get_params(tbec_mlp *m, float *out) {
  if (!m || !out) return set_error(TBEC_ERR_ARG, "null argument");
  MCHK(hipMemcpy(out, m->params, sizeof(float) * m->nparam, hipMemcpyDeviceToHost));
  return 0;
}

int tbec_mlp_set_params(tbec_mlp *m, const float *in) {
  if (!m || !in) return set_error(TBEC_ERR_ARG, "null argument");
  MCHK(hipMemcpy(m->params, in, sizeof(float) * m->nparam, hipMemcpyHostToDevice));
  int rc = refresh_bf16(m);
  if (rc) return rc;
  MCHK(hipStreamSynchronize(m->stream));
  m->adam_t = 0;
  MCHK(hipMemset(m->adam_m, 0, sizeof(float) * m->nparam));
  MCHK(hipMemset(m->adam_v, 0, sizeof(float) * m->nparam));
  return 0;
}

int tbec_mlp_forward(tbec_mlp *m, const float *x, int32_t batch, float *y) {
  if (!m || !x || !y || batch < 1) return set_error(TBEC_ERR_ARG, "bad argument");
  int rc = ensure_batch(m, batch);
  if (rc) return rc;
  MCHK(hipMemcpyAsync(m->H[0], x, sizeof(float) * (size_t)batch * m->n_in,
                      hipMemcpyHostToDevice, m->stream));
  rc = forward(m, batch);
  if (rc) return rc;
  MCHK(hipMemcpyAsync(y, m->Y, sizeof(float) * (size_t)batch * m->n_out, hipMemcpyDeviceToHost,
                      m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  return 0;
}

int tbec_mlp_loss_grad(tbec_mlp *m, const float *x, const float *t, int32_t batch,
                       double *loss, float *grads) {
  if (!m || !x || !t || !loss || batch < 1) return set_error(TBEC_ERR_ARG, "bad argument");
  int rc = ensure_batch(m, batch);
  if (rc) return rc;
  MCHK(hipMemcpyAsync(m->H[0], x, sizeof(float) * (size_t)batch * m->n_in,
                      hipMemcpyHostToDevice, m->stream));
  MCHK(hipMemcpyAsync(m->T, t, sizeof(float) * (size_t)batch * m->n_out, hipMemcpyHostToDevice,
                      m->stream));
  rc = forward(m, batch);
  if (!rc) rc = backward(m, batch, loss);
  if (rc) return rc;
  if (grads) {
    MCHK(hipMemcpy(grads, m->grads, sizeof(float) * m->nparam, hipMemcpyDeviceToHost));
  }
  return 0;
}

int tbec_mlp_set_data(tbec_mlp *m, const float *x, const float *t, int32_t n) {
  if (!m || !x || !t || n < 1) return set_error(TBEC_ERR_ARG, "bad argument");
  if (m->X_all) (void)hipFree(m->X_all);
  if (m->T_all) (void)hipFree(m->T_all);
  m->X_all = m->T_all = nullptr;
  MCHK(hipMalloc(&m->X_all, sizeof(float) * (size_t)n * m->n_in));
  MCHK(hipMalloc(&m->T_all, sizeof(float) * (size_t)n * m->n_out));
  MCHK(hipMemcpy(m->X_all, x, sizeof(float) * (size_t)n * m->n_in, hipMemcpyHostToDevice));
  MCHK(hipMemcpy(m->T_all, t, sizeof(float) * (size_t)n * m->n_out, hipMemcpyHostToDevice));
  m->n_data = n;
  return 0;
}

// Minibatch Adam over the stored data set: `epochs` passes in seeded shuffled order.
// loss_hist (optional) receives the mean training loss of each epoch.
int tbec_mlp_train(tbec_mlp *m, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                   double *loss_hist) {
  if (!m || epochs < 1 || batch < 1 || !(lr > 0)) return set_error(TBEC_ERR_ARG, "bad argument");
  if (m->n_data < 1) return set_error(TBEC_ERR_STATE, "no training data set");
  const int B = std::min(batch, m->n_data);
  int rc = ensure_batch(m, B);
  if (rc) return rc;
  std::vector<int> perm(m->n_data);
  uint64_t s = seed;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int i = 0; i < m->n_data; ++i) perm[i] = i;
    for (int i = m->n_data - 1; i > 0; --i) std::swap(perm[i], perm[splitmix(s) % (i + 1)]);
    double lsum = 0.0;
    int nbat = 0;
    for (int b0 = 0; b0 + B <= m->n_data; b0 += B) {
      MCHK(hipMemcpyAsync(m->d_idx, perm.data() + b0, sizeof(int) * B, hipMemcpyHostToDevice,
                          m->stream));
      hipLaunchKernelGGL(k_gather_rows, dim3(nblk((long long)B * m->n_in)), dim3(256), 0,
                         m->stream, m->X_all, m->d_idx, m->H[0], B, m->n_in);
      hipLaunchKernelGGL(k_gather_rows, dim3(nblk((long long)B * m->n_out)), dim3(256), 0,
                         m->stream, m->T_all, m->d_idx, m->T, B, m->n_out);
      MCHK(hipGetLastError());
      double loss = 0.0;
      rc = forward(m, B);
      if (!rc) rc = backward(m, B, &loss);
      if (rc) return rc;
      ++m->adam_t;
      const float c1 = 1.f - std::pow(b1, (float)m->adam_t);
      const float c2 = 1.f - std::pow(b2, (float)m->adam_t);
      hipLaunchKernelGGL(k_adam, dim3(nblk(m->nparam)), dim3(256), 0, m->stream, m->params,
                         m->grads, m->adam_m, m->adam_v, m->nparam, lr, b1, b2, eps, c1, c2);
      MCHK(hipGetLastError());
      rc = refresh_bf16(m);
      if (rc) return rc;
      lsum += loss;
      ++nbat;
    }
    if (loss_hist) loss_hist[ep] = lsum / std::max(1, nbat);
  }
  MCHK(hipStreamSynchronize(m->stream));
  return 0;
}

}  // extern "C"

namespace {

// Trilinear interpolation of a coarse (3, cn, cn, cn) field -- sqrt-density a, cos(phi),
// sin(phi), C order [x][y][z], cell-centred on the same box -- at the fine cell centres,
// combined into psi = a (c + i s)/|c + i s|.  Coarse indices are clamped at the edges.
template <typename T>
__global__ void k_upsample(const float *__restrict__ coarse, int cn, cplx<T> *__restrict__ psi,
                           int nx, int ny, int nz, long long sy, long long sx, long long base) {
  const long long n = (long long)nx * ny * nz;
  const long long cn3 = (long long)cn * cn * cn;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const long long k = t % nz, r = t / nz, j = r % ny, i = r / ny;
    const float u[3] = {(i + 0.5f) * cn / nx - 0.5f, (j + 0.5f) * cn / ny - 0.5f,
                        (k + 0.5f) * cn / nz - 0.5f};
    int i0[3];
    float f[3];
    for (int d = 0; d < 3; ++d) {
      const float c = fminf(fmaxf(u[d], 0.f), (float)(cn - 1));
      i0[d] = min((int)floorf(c), cn - 2);
      f[d] = c - i0[d];
    }
    float ch[3];
    for (int q = 0; q < 3; ++q) {
      const float *g = coarse + q * cn3;
      float acc = 0.f;
      for (int dx = 0; dx < 2; ++dx)
        for (int dy = 0; dy < 2; ++dy)
          for (int dz = 0; dz < 2; ++dz) {
            const float w = (dx ? f[0] : 1.f - f[0]) * (dy ? f[1] : 1.f - f[1]) *
                            (dz ? f[2] : 1.f - f[2]);
            acc += w * g[((long long)(i0[0] + dx) * cn + (i0[1] + dy)) * cn + (i0[2] + dz)];
          }
      ch[q] = acc;
    }
    const float a = fmaxf(ch[0], 0.f);
    const float m = sqrtf(ch[1] * ch[1] + ch[2] * ch[2]);
    cplx<T> v;
    if (m > 1e-12f) { v.re = T(a * ch[1] / m); v.im = T(a * ch[2] / m); }
    else { v.re = T(a); v.im = T(0); }
    psi[base + i * sx + j * sy + k] = v;
  }
}

}  // namespace

extern "C" int tbec_upload_coarse(tbec_handle *h, const float *coarse, int32_t cn) {
  if (!h || !coarse || cn < 2) return set_error(TBEC_ERR_ARG, "bad argument");
  float *d = nullptr;
  const size_t bytes = sizeof(float) * 3 * (size_t)cn * cn * cn;
  MCHK(hipMalloc(&d, bytes));
  hipError_t e = hipMemcpyAsync(d, coarse, bytes, hipMemcpyHostToDevice, h->stream);
  const long long n = (long long)h->p.nx * h->p.ny * h->p.nz;
  if (e == hipSuccess) {
    if (h->p.prec == TBEC_F32)
      hipLaunchKernelGGL(k_upsample<float>, dim3(nblk(n)), dim3(256), 0, h->stream, d, cn,
                         (cplx<float> *)h->psi[h->cur], h->p.nx, h->p.ny, h->p.nz, h->sy, h->sx,
                         h->base);
    else
      hipLaunchKernelGGL(k_upsample<double>, dim3(nblk(n)), dim3(256), 0, h->stream, d, cn,
                         (cplx<double> *)h->psi[h->cur], h->p.nx, h->p.ny, h->p.nz, h->sy,
                         h->sx, h->base);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return set_error(TBEC_ERR_HIP, hipGetErrorString(e));
  return tbec_reset_state(h);
}
