// Surrogate network kernels: an ensemble of E fully connected GELU networks
// of identical shape, trained and evaluated together (blockIdx.z = member).
//
// * gemm_bf16_kernel: C[m][n] = sum_k A(m,k) B(n,k) on the matrix cores
//   (v_mfma_f32_32x32x8 bf16, FP32 accumulation), 64x64 output tiles of four
//   waves, 32-deep K slices staged through LDS; either operand may be stored
//   k-contiguous or row-contiguous, so forward (X W^T), input gradient
//   (dZ W) and weight gradient (dZ^T X) are one kernel.  Epilogues fuse the
//   bias, GELU (pre-activation kept in FP32 for the backward pass) and the
//   GELU derivative.
// * mse_grad_kernel, bias_grad_kernel, adam_kernel, gather_kernel: loss
//   gradient, bias gradients, the Adam update (FP32 master weights, BF16
//   copies for the GEMMs) and minibatch assembly.
#include <cmath>
#include <cstring>
#include <memory>

#include "common.h"
#include "mlp.h"

namespace imped {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// round-to-nearest-even float -> bfloat16 (NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU(x) = x Phi(x) and its derivative Phi(x) + x phi(x), Phi(x) = (1 +
// erf(x / sqrt 2)) / 2 (training: full-precision erf)
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

// erf by the rational approximation of Abramowitz and Stegun, Handbook of
// Mathematical Functions, eq. 7.1.26 (after Hastings): erf(x) = 1 - (a1 t +
// a2 t^2 + a3 t^3 + a4 t^4 + a5 t^5) exp(-x^2), t = 1 / (1 + p x) for x >= 0,
// odd in x, absolute error at most 1.5e-7 (5.3e-7 in FP32 arithmetic) -- far
// below the BF16 rounding (2^-8 relative) of the activations it feeds, at a
// fraction of the instructions of the full-precision erf.  Inference only
// (the epilogue of a hidden layer evaluates sixteen per lane); training keeps
// gelu above, so a trained network does not depend on it.
__device__ __forceinline__ float erf_as(float x) {
  const float ax = fabsf(x);
  const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * ax);
  const float poly =
      t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
  return copysignf(1.0f - poly * __expf(-ax * ax), x);
}
__device__ __forceinline__ float gelu_inference(float x) {
  return 0.5f * x * (1.0f + erf_as(x * 0.70710678118654752f));
}

enum { EPI_BIAS_GELU = 0, EPI_BIAS = 1, EPI_DGELU = 2, EPI_F32 = 3 };

struct GemmArgs {
  const uint16_t* A;  // AK: A(m,k) = A[m*lda + k]; else A[k*lda + m]
  int64_t lda, a_z;
  const uint16_t* B;  // BK: B(n,k) = B[n*ldb + k]; else B[k*ldb + n]
  int64_t ldb, b_z;
  int M, N, K;
  float* Cf;  // row-major [M][ldc]
  uint16_t* Cb;
  int64_t ldc, c_z;
  const float* bias;  // [N] per member
  int64_t bias_z;
  const float* aux;  // EPI_DGELU: pre-activation, layout of C
  int vec;           // gemm_fwd_kernel: K, lda, ldb, a_z, b_z multiples of 8, A and B 16-byte aligned
  const float* Af;   // gemm_fwd_kernel<EPI, true>: A in FP32 (shared by the members), rounded to BF16 when staged
};

constexpr int TM = 64, TN = 64, TK = 32, LK = TK + 8;

// Stage a 64 x 32 slice (rows r0.., k0..) of an operand into tile[row][k],
// zero outside [0, R) x [0, K).
template <bool KC>
__device__ __forceinline__ void stage(const uint16_t* src, int64_t ld, int R, int K, int r0, int k0,
                                      uint16_t (*tile)[LK], int t) {
  if constexpr (KC) {
    const int r = t >> 2, kc = (t & 3) * 8;
    const int gr = r0 + r, gk = k0 + kc;
    uint16_t v[8];
    const uint16_t* p = src + (int64_t)gr * ld + gk;
    if (gr < R && gk + 8 <= K && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
      *reinterpret_cast<uint4*>(v) = *reinterpret_cast<const uint4*>(p);
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = (gr < R && gk + i < K) ? src[(int64_t)gr * ld + gk + i] : (uint16_t)0;
    }
    *reinterpret_cast<uint4*>(&tile[r][kc]) = *reinterpret_cast<const uint4*>(v);
  } else {
    const int k = t >> 3, rc = (t & 7) * 8;
    const int gk = k0 + k, gr = r0 + rc;
    uint16_t v[8];
    const uint16_t* p = src + (int64_t)gk * ld + gr;
    if (gk < K && gr + 8 <= R && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
      *reinterpret_cast<uint4*>(v) = *reinterpret_cast<const uint4*>(p);
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = (gk < K && gr + i < R) ? src[(int64_t)gk * ld + gr + i] : (uint16_t)0;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) tile[rc + i][k] = v[i];
  }
}

template <bool AK, bool BK, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
  __shared__ __align__(16) uint16_t As[TM][LK];
  __shared__ __align__(16) uint16_t Bs[TN][LK];
  const int z = blockIdx.z;
  const uint16_t* A = g.A + (int64_t)z * g.a_z;
  const uint16_t* B = g.B + (int64_t)z * g.b_z;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  for (int k0 = 0; k0 < g.K; k0 += TK) {
    stage<AK>(A, g.lda, g.M, g.K, m0, k0, As, t);
    stage<BK>(B, g.ldb, g.N, g.K, n0, k0, Bs, t);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; kk += 16) {
      // 16-deep steps (gfx950): lane l supplies row l%32 and k = kk + 8*(l/32) .. +7
      // of each operand (rows of LK = 40 elements keep every read 16-byte aligned)
      const int kq = kk + 8 * (lane >> 5);
      const uint4 a = *reinterpret_cast<const uint4*>(&As[wm + (lane & 31)][kq]);
      const uint4 b = *reinterpret_cast<const uint4*>(&Bs[wn + (lane & 31)][kq]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                    0, 0, 0);
    }
    __syncthreads();
  }
  // accumulator r of lane l holds row 8*(r/4) + 4*(l/32) + r%4, column l%32
  // (the column's bias read once and added before the first store)
  const int64_t cz = (int64_t)z * g.c_z;
  const int n = n0 + wn + (lane & 31);
  constexpr bool with_bias = EPI == EPI_BIAS_GELU || EPI == EPI_BIAS;
  const float bn = with_bias && n < g.N ? g.bias[(int64_t)z * g.bias_z + n] : 0.0f;
  float zr[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    zr[r] = with_bias ? acc[r] + bn : acc[r];
    asm volatile("" : "+v"(zr[r]));  // no instruction: keeps the sum here, out of the branches below
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int m = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (m >= g.M || n >= g.N) continue;
    const int64_t o = cz + (int64_t)m * g.ldc + n;
    const float v = zr[r];
    if constexpr (EPI == EPI_BIAS_GELU) {
      g.Cf[o] = v;
      g.Cb[o] = f2bf(gelu(v));
    } else if constexpr (EPI == EPI_BIAS) {
      g.Cf[o] = v;
    } else if constexpr (EPI == EPI_DGELU) {
      g.Cb[o] = f2bf(v * gelu_grad(g.aux[o]));
    } else {
      g.Cf[o] = v;
    }
  }
}

// Forward GEMM C = X W^T (+ bias, GELU) with K staged FK at a time for both
// 64-row operand tiles: every global load of a stage is in flight together
// and one barrier precedes its MFMAs, so a 512-deep layer pays two load
// latencies instead of sixteen.  2 x 64 x (FK + 8) bf16 = 66 KiB of LDS, so
// two workgroups share a CU (the 1000-wide output layer has 320 of them).
constexpr int FK = 256, FLK = FK + 8;
constexpr size_t FWD_LDS = (size_t)2 * TM * FLK * sizeof(uint16_t);

// 8 consecutive bf16 of row gr from column gk (zeros outside [0, R) x [0, K)).
__device__ __forceinline__ uint4 load8(const uint16_t* src, int64_t ld, int R, int K, int gr, int gk) {
  const uint16_t* p = src + (int64_t)gr * ld + gk;
  if (gr < R && gk + 8 <= K && (reinterpret_cast<uintptr_t>(p) & 15) == 0) return *reinterpret_cast<const uint4*>(p);
  uint16_t v[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = (gr < R && gk + i < K) ? src[(int64_t)gr * ld + gk + i] : (uint16_t)0;
  return *reinterpret_cast<const uint4*>(v);
}

// Eight FP32 values rounded to BF16, packed as the BF16 row chunk.
__device__ __forceinline__ uint4 pack_bf16x8(const float4& a, const float4& b) {
  uint4 r;
  r.x = (uint32_t)f2bf(a.x) | ((uint32_t)f2bf(a.y) << 16);
  r.y = (uint32_t)f2bf(a.z) | ((uint32_t)f2bf(a.w) << 16);
  r.z = (uint32_t)f2bf(b.x) | ((uint32_t)f2bf(b.y) << 16);
  r.w = (uint32_t)f2bf(b.z) | ((uint32_t)f2bf(b.w) << 16);
  return r;
}

// AF: the A operand is FP32 (g.Af, vec path only; the predictions' first
// layer reads the batch inputs straight from mapped host memory, so no
// conversion launch precedes it).
template <int EPI, bool AF>
__global__ __launch_bounds__(256) void gemm_fwd_kernel(GemmArgs g) {
  extern __shared__ __align__(16) uint16_t fsm[];
  uint16_t(*As)[FLK] = reinterpret_cast<uint16_t(*)[FLK]>(fsm);
  uint16_t(*Bs)[FLK] = reinterpret_cast<uint16_t(*)[FLK]>(fsm + TM * FLK);
  // Tile order for the eight XCDs, which take the workgroups round-robin in
  // launch order (x fastest) and each have their own L2: with the launch
  // order, XCD i got output-column tile i of every row tile and member, so
  // each XCD fetched every activation tile (eight fetches per tile).
  // Renumbered so that each XCD takes one contiguous run of tiles, ordered
  // member, column tile, row tile (rows fastest): an XCD's run covers whole
  // columns of one or two members, every weight tile is fetched by one XCD
  // and every activation tile by about two.  Same arithmetic per tile.
  int z = blockIdx.z, bx = blockIdx.x, by = blockIdx.y;
  {
    const int gx = gridDim.x, gy = gridDim.y, tiles = gx * gy * gridDim.z;
    if ((tiles & 7) == 0) {
      const int l = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
      const int q = (l & 7) * (tiles >> 3) + (l >> 3);  // a bijection of [0, tiles) when 8 | tiles
      z = q / (gx * gy);
      const int r = q - z * gx * gy;
      bx = r / gy;
      by = r - bx * gy;
    }
  }
  const uint16_t* A = AF ? nullptr : g.A + (int64_t)z * g.a_z;
  const uint16_t* B = g.B + (int64_t)z * g.b_z;
  const int m0 = by * TM, n0 = bx * TN;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  for (int k0 = 0; k0 < g.K; k0 += FK) {
    const int kc8 = ((g.K - k0 < FK ? g.K - k0 : FK) + 7) & ~7;  // MFMA steps of 8, zero padded
    const int per_row = kc8 / 8;
    const int chunks = TM * per_row;
    if (g.vec) {
      // every load of the stage issued before the first LDS write; rows past
      // M or N read the last row instead (their outputs are not stored), so
      // no load waits behind a branch
      // (chunk indices past the stage repeat its last chunk: those threads
      // write the same values to the same LDS slot, so nothing branches)
      uint4 ra[8], rb[8];
      float4 fa[AF ? 8 : 1][2];
      int slot[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int c = t + 256 * i < chunks ? t + 256 * i : chunks - 1;
        const int r = c / per_row, k8 = (c - r * per_row) * 8;
        slot[i] = r * FLK + k8;
        const int ga = m0 + r < g.M ? m0 + r : g.M - 1, gb = n0 + r < g.N ? n0 + r : g.N - 1;
        if constexpr (AF) {
          const float* pa = g.Af + (int64_t)ga * g.lda + k0 + k8;
          fa[i][0] = *reinterpret_cast<const float4*>(pa);
          fa[i][1] = *reinterpret_cast<const float4*>(pa + 4);
        } else {
          ra[i] = *reinterpret_cast<const uint4*>(A + (int64_t)ga * g.lda + k0 + k8);
        }
        rb[i] = *reinterpret_cast<const uint4*>(B + (int64_t)gb * g.ldb + k0 + k8);
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        if constexpr (AF) ra[i] = pack_bf16x8(fa[i][0], fa[i][1]);
        *reinterpret_cast<uint4*>(&As[0][0] + slot[i]) = ra[i];
        *reinterpret_cast<uint4*>(&Bs[0][0] + slot[i]) = rb[i];
      }
    } else if constexpr (!AF) {
#pragma unroll 8
      for (int c = t; c < chunks; c += 256) {
        const int r = c / per_row, k8 = (c - r * per_row) * 8;
        const uint4 va = load8(A, g.lda, g.M, g.K, m0 + r, k0 + k8);
        const uint4 vb = load8(B, g.ldb, g.N, g.K, n0 + r, k0 + k8);
        *reinterpret_cast<uint4*>(&As[r][k8]) = va;
        *reinterpret_cast<uint4*>(&Bs[r][k8]) = vb;
      }
    }
    __syncthreads();
    int kk = 0;
    for (; kk + 16 <= kc8; kk += 16) {
      // 16-deep steps (gfx950): lane l supplies row l%32 and k = kk + 8*(l/32) .. +7
      // of each operand; both operands use the same k assignment
      const int kq = kk + 8 * (lane >> 5);
      const uint4 a = *reinterpret_cast<const uint4*>(&As[wm + (lane & 31)][kq]);
      const uint4 b = *reinterpret_cast<const uint4*>(&Bs[wn + (lane & 31)][kq]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                    0, 0, 0);
    }
    if (kk < kc8) {
      // an odd 8-deep step: lane l supplies row l%32 and k = kk + 4*(l/32) .. +3
      const int kq = kk + 4 * (lane >> 5);
      const bf16x4 a = *reinterpret_cast<const bf16x4*>(&As[wm + (lane & 31)][kq]);
      const bf16x4 b = *reinterpret_cast<const bf16x4*>(&Bs[wn + (lane & 31)][kq]);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // accumulator r of lane l holds row 8*(r/4) + 4*(l/32) + r%4, column l%32.
  // The column's bias is read once and added to every accumulator before the
  // first store: read after each store (as a pointer that may alias the
  // output would be), or first used inside the per-row branches, it puts a
  // wait for every earlier store (stores count in vmcnt) between the stores.
  const int64_t cz = (int64_t)z * g.c_z;
  const int n = n0 + wn + (lane & 31);
  const float bn = n < g.N ? g.bias[(int64_t)z * g.bias_z + n] : 0.0f;
  float zr[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    zr[r] = acc[r] + bn;
    asm volatile("" : "+v"(zr[r]));  // no instruction: keeps the sum here, out of the branches below
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int m = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (m >= g.M || n >= g.N) continue;
    const int64_t o = cz + (int64_t)m * g.ldc + n;
    const float zv = zr[r];
    if (EPI != EPI_BIAS_GELU || g.Cf) g.Cf[o] = zv;  // hidden layers of a prediction keep no pre-activation
    if constexpr (EPI == EPI_BIAS_GELU) g.Cb[o] = f2bf(g.Cf ? gelu(zv) : gelu_inference(zv));
  }
}

__global__ void to_bf16_kernel(const float* x, int64_t n, uint16_t* y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// minibatch rows idx[b] of the data set: inputs to BF16, targets as FP32
__global__ void gather_kernel(const float* X, int dx, const float* Y, int dy, const int32_t* idx, int B,
                              uint16_t* xb, float* yb) {
  const int b = blockIdx.x;
  if (b >= B) return;
  const int64_t r = idx[b];
  for (int j = threadIdx.x; j < dx; j += blockDim.x) xb[(int64_t)b * dx + j] = f2bf(X[r * dx + j]);
  for (int j = threadIdx.x; j < dy; j += blockDim.x) yb[(int64_t)b * dy + j] = Y[r * dy + j];
}

// dY = 2 (Y - T) / (B N) in BF16 per member; loss[z] += sum (Y - T)^2 / (B N)
__global__ __launch_bounds__(256) void mse_grad_kernel(const float* Y, const float* T, int64_t BN, uint16_t* dY,
                                                       double* loss) {
  __shared__ double red[4];
  const int z = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double s = 0.0;
  if (i < BN) {
    const float d = Y[(int64_t)z * BN + i] - T[i];
    dY[(int64_t)z * BN + i] = f2bf(2.0f * d / (float)BN);
    s = (double)d * (double)d;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&loss[z], (red[0] + red[1] + red[2] + red[3]) / (double)BN);
}

// Sum over a 256-thread workgroup, broadcast to every thread.
__device__ __forceinline__ double sum256(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  __syncthreads();  // red may still be read by the previous call
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// Ensemble statistics of one sample per workgroup, from the outputs
// y [E][B][D]: member spectra m_e = max(y_e / scale, 0) / area_e (area =
// sum_d w_d (.), the trapezoid rule when w holds its weights), their mean
// renormalized to unit area, the disagreement
// dis = (1/E) sum_e sum_d w_d |m_e,d - mean_d| and, when target is given,
// the misfit l1 = sum_d w_d |mean_d - target_d|.  E <= 64; the general
// path (ensemble_stats_reg_kernel covers E <= 8, D <= 1024).
__global__ __launch_bounds__(256) void ensemble_stats_kernel(const float* y, int E, int B, int D, const double* w,
                                                             double scale, const double* target, double* mean,
                                                             double* dis, double* l1) {
  __shared__ double red[4];
  __shared__ double area[64];
  const int b = blockIdx.x;
  for (int e = 0; e < E; ++e) {
    const float* ye = y + ((int64_t)e * B + b) * D;
    double p = 0.0;
    for (int d = threadIdx.x; d < D; d += 256) p += w[d] * fmax((double)ye[d] / scale, 0.0);
    p = sum256(p, red);
    if (threadIdx.x == 0) area[e] = p;
  }
  __syncthreads();
  double* mb = mean + (int64_t)b * D;
  double pm = 0.0;
  for (int d = threadIdx.x; d < D; d += 256) {
    double acc = 0.0;
    for (int e = 0; e < E; ++e) acc += fmax((double)y[((int64_t)e * B + b) * D + d] / scale, 0.0) / area[e];
    acc /= (double)E;
    mb[d] = acc;
    pm += w[d] * acc;
  }
  const double marea = sum256(pm, red);
  double pd = 0.0, pl = 0.0;
  for (int d = threadIdx.x; d < D; d += 256) {
    const double mn = mb[d] / marea;
    mb[d] = mn;
    for (int e = 0; e < E; ++e)
      pd += w[d] * fabs(fmax((double)y[((int64_t)e * B + b) * D + d] / scale, 0.0) / area[e] - mn);
    if (target) pl += w[d] * fabs(mn - target[d]);
  }
  pd = sum256(pd, red);
  if (target) pl = sum256(pl, red);
  if (threadIdx.x == 0) {
    dis[b] = pd / (double)E;
    if (target) l1[b] = pl;
  }
}

__device__ __forceinline__ double wave_sum64(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Sums of NV values over a 256-thread workgroup, broadcast to every thread;
// red [4][NV] is this round's own LDS, so one barrier suffices.
template <int NV>
__device__ __forceinline__ void sums256(double (&v)[NV], double (*red)[NV]) {
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) red[threadIdx.x >> 6][i] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = red[0][i] + red[1][i] + red[2][i] + red[3][i];
}

// The same statistics with the sample's member outputs loaded once into
// registers (EM members, D <= 256 K; thread t holds d = t + 256 k): each
// statistic is register arithmetic and one workgroup reduction (DPP wave
// sums, then LDS), three rounds in all, and every division is by a
// reciprocal formed once.
template <int K, int EM>
__global__ __launch_bounds__(256) void ensemble_stats_reg_kernel(const float* y, int E, int B, int D,
                                                                 const double* w, double scale,
                                                                 const double* target, double* mean, double* dis,
                                                                 double* l1) {
  __shared__ double red_a[4][EM], red_m[4][1], red_d[4][2];
  const int t = threadIdx.x;
  const int b = blockIdx.x;
  const double inv_scale = 1.0 / scale;
  double v[EM][K], wd[K];
#pragma unroll
  for (int e = 0; e < EM; ++e)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int d = t + 256 * k;
      v[e][k] = (e < E && d < D) ? fmax((double)y[((int64_t)e * B + b) * D + d] * inv_scale, 0.0) : 0.0;
    }
  double tg[K];  // the target's values, loaded with the rest (mean may be NULL: not written)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int d = t + 256 * k;
    wd[k] = d < D ? w[d] : 0.0;
    tg[k] = target && d < D ? target[d] : 0.0;
  }
  double inv_area[EM];
#pragma unroll
  for (int e = 0; e < EM; ++e) {
    double p = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) p += wd[k] * v[e][k];
    inv_area[e] = p;
  }
  sums256<EM>(inv_area, red_a);
#pragma unroll
  for (int e = 0; e < EM; ++e) inv_area[e] = e < E ? 1.0 / inv_area[e] : 0.0;
  double mn[K], pm[1] = {0.0};
  const double inv_e = 1.0 / (double)E;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double acc = 0.0;
#pragma unroll
    for (int e = 0; e < EM; ++e) acc += v[e][k] * inv_area[e];
    mn[k] = acc * inv_e;
    pm[0] += wd[k] * mn[k];
  }
  sums256<1>(pm, red_m);
  const double inv_marea = 1.0 / pm[0];
  double pd[2] = {0.0, 0.0};
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int d = t + 256 * k;
    const double m = mn[k] * inv_marea;
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < EM; ++e)
      if (e < E) s += fabs(v[e][k] * inv_area[e] - m);
    pd[0] += wd[k] * s;
    if (d < D) {
      if (mean) mean[(int64_t)b * D + d] = m;
      if (target) pd[1] += wd[k] * fabs(m - tg[k]);
    }
  }
  sums256<2>(pd, red_d);
  if (t == 0) {
    dis[b] = pd[0] * inv_e;
    if (target) l1[b] = pd[1];
  }
}

// db[z][n] = sum_b dZ[z][b][n]
__global__ void bias_grad_kernel(const uint16_t* dZ, int B, int N, int64_t dz_z, float* db, int64_t db_z) {
  const int z = blockIdx.y;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.0f;
  for (int b = 0; b < B; ++b) s += bf2f(dZ[(int64_t)z * dz_z + (int64_t)b * N + n]);
  db[(int64_t)z * db_z + n] = s;
}

__global__ void adam_kernel(float* w, const float* g, float* m, float* v, uint16_t* wb, int64_t n, float lr, float b1,
                            float b2, float eps, float bc1, float bc2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  w[i] = wi;
  wb[i] = f2bf(wi);
}

template <bool AK, bool BK, int EPI>
void gemm(const GemmArgs& g, int E, hipStream_t s) {
  const dim3 grid(blocks_for(g.N, TN), blocks_for(g.M, TM), E);
  hipLaunchKernelGGL((gemm_bf16_kernel<AK, BK, EPI>), grid, dim3(256), 0, s, g);
  HIP_CHECK(hipGetLastError());
}

// forward layer (X W^T + bias, GELU or not) by the whole-K kernel
template <int EPI, bool AF = false>
void fwd_attr() {
  static bool attr = false;
  if (!attr) {
    HIP_CHECK(hipFuncSetAttribute((const void*)gemm_fwd_kernel<EPI, AF>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)FWD_LDS));
    attr = true;
  }
}

template <int EPI>
void gemm_fwd(const GemmArgs& g0, int E, hipStream_t s) {
  GemmArgs g = g0;
  const void* a = g.Af ? (const void*)g.Af : (const void*)g.A;
  g.vec = g.K % 8 == 0 && g.lda % 8 == 0 && g.ldb % 8 == 0 && g.a_z % 8 == 0 && g.b_z % 8 == 0 &&
          (reinterpret_cast<uintptr_t>(a) & 15) == 0 && (reinterpret_cast<uintptr_t>(g.B) & 15) == 0;
  const dim3 grid(blocks_for(g.N, TN), blocks_for(g.M, TM), E);
  if (g.Af) {
    if (!g.vec) throw std::runtime_error("gemm_fwd: an FP32 operand needs K, strides and alignment in steps of 8");
    fwd_attr<EPI, true>();
    hipLaunchKernelGGL((gemm_fwd_kernel<EPI, true>), grid, dim3(256), FWD_LDS, s, g);
  } else {
    fwd_attr<EPI>();
    hipLaunchKernelGGL((gemm_fwd_kernel<EPI, false>), grid, dim3(256), FWD_LDS, s, g);
  }
  HIP_CHECK(hipGetLastError());
}

}  // namespace

// ------------------------------------------------------------------ network state

// Page-locked host memory mapped into the device's address space: kernels
// read small per-call inputs from it and write small results into it, so
// they move without a copy dispatch.  Every call that uses it synchronizes
// its stream before returning, so the next call may overwrite it.
struct HostIO {
  void* h = nullptr;  // host address
  void* d = nullptr;  // device address of the same memory
  size_t bytes = 0;
  void reserve(size_t n) {
    if (n <= bytes) return;
    release();
    HIP_CHECK(hipHostMalloc(&h, n, hipHostMallocDefault));
    HIP_CHECK(hipHostGetDevicePointer(&d, h, 0));
    bytes = n;
  }
  void release() {
    if (h) (void)hipHostFree(h);
    h = d = nullptr;
    bytes = 0;
  }
  HostIO() = default;
  HostIO(const HostIO&) = delete;
  HostIO& operator=(const HostIO&) = delete;
  ~HostIO() { release(); }
};

struct Mlp {
  int E = 0, L = 0, max_batch = 0;
  std::vector<int> dims;            // L + 1
  std::vector<int64_t> w_off, b_off;  // per layer offsets in a member's flat parameter vector
  int64_t P = 0;                    // parameters per member
  int64_t step = 0;                 // Adam steps taken
  DevBuf w, wb, g, m, v;            // [E][P] fp32 master, bf16 copy, gradient, Adam moments
  static constexpr int MAXL = 16;
  DevBuf zf[MAXL], hb[MAXL], dz[MAXL];  // per layer: pre-activation fp32, activation bf16, dZ bf16
  bool acts = false;
  DevBuf xb, yb, out, loss, idx;    // minibatch input bf16, targets fp32, output fp32 [E][B][dL]
  DevBuf dX, dY;                    // data set (fp32)
  DevBuf sw, smean, st;             // ensemble statistics: area weights, mean [B][dL], misfit target [dL]
  HostIO xin, sout;                 // batch input x [B][d0] (fp32); disagreement [B] and misfit [B]
  std::vector<double> sw_host, st_host;  // the area weights and target last uploaded
  // Prediction graphs (input conversion, forward GEMMs, statistics): one
  // launch instead of six, so the GPU does not wait for the host between
  // kernels.  Keyed by everything baked into the captured launches.
  struct PredictGraph {
    std::vector<uintptr_t> key;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<PredictGraph> graphs;
  hipStream_t cap = nullptr;  // capture stream
  ~Mlp() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    if (cap) (void)hipStreamDestroy(cap);
  }
  int64_t n_data = 0;
};

namespace {

void alloc_acts(Mlp& net, int B) {
  if (B <= net.max_batch && net.acts) return;
  net.max_batch = std::max(B, net.max_batch);
  net.acts = true;
  const int64_t Bm = net.max_batch;
  for (int l = 0; l < net.L; ++l) {
    const int64_t out = net.dims[l + 1];
    net.zf[l].reserve(net.E * Bm * out * 4);
    net.hb[l].reserve(net.E * Bm * out * 2);
    net.dz[l].reserve(net.E * Bm * out * 2);
  }
  net.xb.reserve(Bm * net.dims[0] * 2);
  net.yb.reserve(Bm * net.dims[net.L] * 4);
  net.out.reserve(net.E * Bm * net.dims[net.L] * 4);
  net.idx.reserve(Bm * 4);
  net.loss.reserve(net.E * 8);
}

// forward pass of every member on the BF16 minibatch in xb; output in out
// x_f32: the first layer reads FP32 inputs there (else the BF16 copy in xb);
// inference: the hidden layers' FP32 pre-activations (kept for the backward
// pass) are not written
void forward(Mlp& net, int B, hipStream_t s, const float* x_f32 = nullptr, bool inference = false) {
  const uint16_t* x = net.xb.as<uint16_t>();
  int64_t x_z = 0;  // the input is shared by all members
  for (int l = 0; l < net.L; ++l) {
    const int in = net.dims[l], outd = net.dims[l + 1];
    GemmArgs g{};
    g.A = x;
    if (l == 0) g.Af = x_f32;
    g.lda = in;
    g.a_z = x_z;
    g.B = net.wb.as<uint16_t>() + net.w_off[l];
    g.ldb = in;
    g.b_z = net.P;
    g.M = B;
    g.N = outd;
    g.K = in;
    g.ldc = outd;
    g.c_z = (int64_t)B * outd;
    g.bias = net.w.as<float>() + net.b_off[l];
    g.bias_z = net.P;
    TimedLaunch tl("mlp_gemm", s);
    if (l + 1 < net.L) {
      g.Cf = inference ? nullptr : net.zf[l].as<float>();
      g.Cb = net.hb[l].as<uint16_t>();
      gemm_fwd<EPI_BIAS_GELU>(g, net.E, s);
      x = net.hb[l].as<uint16_t>();
      x_z = (int64_t)B * outd;
    } else {
      g.Cf = net.out.as<float>();
      gemm_fwd<EPI_BIAS>(g, net.E, s);
    }
  }
}

// gradients of the mean-squared error against yb into g; loss per member
void backward(Mlp& net, int B, hipStream_t s) {
  const int L = net.L;
  const int64_t BN = (int64_t)B * net.dims[L];
  HIP_CHECK(hipMemsetAsync(net.loss.p, 0, net.E * 8, s));
  hipLaunchKernelGGL(mse_grad_kernel, dim3(blocks_for(BN, 256), net.E), dim3(256), 0, s, net.out.as<float>(),
                     net.yb.as<float>(), BN, net.dz[L - 1].as<uint16_t>(), net.loss.as<double>());
  HIP_CHECK(hipGetLastError());
  for (int l = L - 1; l >= 0; --l) {
    const int in = net.dims[l], outd = net.dims[l + 1];
    const uint16_t* dZ = net.dz[l].as<uint16_t>();
    const uint16_t* X = l > 0 ? net.hb[l - 1].as<uint16_t>() : net.xb.as<uint16_t>();
    const int64_t x_z = l > 0 ? (int64_t)B * in : 0;
    {
      // dW[o][i] = sum_b dZ[b][o] X[b][i]
      GemmArgs g{};
      g.A = dZ;
      g.lda = outd;
      g.a_z = (int64_t)B * outd;
      g.B = X;
      g.ldb = in;
      g.b_z = x_z;
      g.M = outd;
      g.N = in;
      g.K = B;
      g.Cf = net.g.as<float>() + net.w_off[l];
      g.ldc = in;
      g.c_z = net.P;
      TimedLaunch tl("mlp_gemm", s);
      gemm<false, false, EPI_F32>(g, net.E, s);
    }
    hipLaunchKernelGGL(bias_grad_kernel, dim3(blocks_for(outd, 256), net.E), dim3(256), 0, s, dZ, B, outd,
                       (int64_t)B * outd, net.g.as<float>() + net.b_off[l], net.P);
    HIP_CHECK(hipGetLastError());
    if (l > 0) {
      // dZ_{l-1}[b][i] = (sum_o dZ[b][o] W[o][i]) * gelu'(Z_{l-1}[b][i])
      GemmArgs g{};
      g.A = dZ;
      g.lda = outd;
      g.a_z = (int64_t)B * outd;
      g.B = net.wb.as<uint16_t>() + net.w_off[l];
      g.ldb = in;
      g.b_z = net.P;
      g.M = B;
      g.N = in;
      g.K = outd;
      g.Cb = net.dz[l - 1].as<uint16_t>();
      g.ldc = in;
      g.c_z = (int64_t)B * in;
      g.aux = net.zf[l - 1].as<float>();
      TimedLaunch tl("mlp_gemm", s);
      gemm<true, false, EPI_DGELU>(g, net.E, s);
    }
  }
}

// x into the mapped input buffer; convert_batch_input launches the
// conversion, which reads it there (no copy dispatch).
void stage_batch_input(Mlp& net, const float* x_host, int B) {
  const int64_t n = (int64_t)B * net.dims[0];
  net.xin.reserve(n * sizeof(float));
  std::memcpy(net.xin.h, x_host, n * sizeof(float));
}

void convert_batch_input(Mlp& net, int B, hipStream_t s) {
  const int64_t n = (int64_t)B * net.dims[0];
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, static_cast<const float*>(net.xin.d),
                     n, net.xb.as<uint16_t>());
  HIP_CHECK(hipGetLastError());
}

void upload_batch_input(Mlp& net, const float* x_host, int B, hipStream_t s) {
  stage_batch_input(net, x_host, B);
  convert_batch_input(net, B, s);
}

}  // namespace

Mlp* mlp_create(int E, int L, const int32_t* dims, int max_batch) {
  auto net = std::make_unique<Mlp>();
  net->E = E;
  net->L = L;
  net->dims.assign(dims, dims + L + 1);
  int64_t off = 0;
  for (int l = 0; l < L; ++l) {
    net->w_off.push_back(off);
    off += (int64_t)dims[l + 1] * dims[l];
    net->b_off.push_back(off);
    off += dims[l + 1];
  }
  // members start 16-element aligned so every BF16 row load stays aligned
  net->P = (off + 15) / 16 * 16;
  const int64_t tot = (int64_t)E * net->P;
  net->w.reserve(tot * 4);
  net->wb.reserve(tot * 2);
  net->g.reserve(tot * 4);
  net->m.reserve(tot * 4);
  net->v.reserve(tot * 4);
  HIP_CHECK(hipMemset(net->w.p, 0, tot * 4));
  HIP_CHECK(hipMemset(net->wb.p, 0, tot * 2));
  HIP_CHECK(hipMemset(net->g.p, 0, tot * 4));
  HIP_CHECK(hipMemset(net->m.p, 0, tot * 4));
  HIP_CHECK(hipMemset(net->v.p, 0, tot * 4));
  alloc_acts(*net, std::max(max_batch, 1));
  return net.release();
}

void mlp_destroy(Mlp* net) { delete net; }
int64_t mlp_param_count(const Mlp& net) { return net.P; }
int64_t mlp_data_count(const Mlp& net) { return net.n_data; }
int mlp_layers(const Mlp& net, int32_t* dims) {
  if (dims)
    for (int l = 0; l <= net.L; ++l) dims[l] = net.dims[l];
  return net.L;
}

void mlp_get_params(Mlp& net, float* params, hipStream_t s) {
  d2h(params, net.w.p, (size_t)net.E * net.P, s);
  stream_sync(s);
}

void mlp_set_params(Mlp& net, const float* params, hipStream_t s) {
  const int64_t tot = (int64_t)net.E * net.P;
  HIP_CHECK(hipMemcpyAsync(net.w.p, params, tot * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(tot, 256)), dim3(256), 0, s, net.w.as<float>(), tot,
                     net.wb.as<uint16_t>());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemsetAsync(net.m.p, 0, tot * 4, s));
  HIP_CHECK(hipMemsetAsync(net.v.p, 0, tot * 4, s));
  stream_sync(s);
  net.step = 0;
}

void mlp_forward(Mlp& net, const float* x_host, int B, float* y_host, hipStream_t s) {
  alloc_acts(net, B);
  upload_batch_input(net, x_host, B, s);
  forward(net, B, s, nullptr, true);  // the network's outputs as predictions form them
  d2h(y_host, net.out.p, (size_t)net.E * B * net.dims[net.L], s);
  stream_sync(s);
}

namespace {

// Uploads a host vector of D doubles to dev when it differs from the copy
// last uploaded (the grid's weights and a fit's target stay the same from
// call to call).
void upload_on_change(DevBuf& dev, std::vector<double>& last, const double* host, int D, hipStream_t s) {
  if (last.size() != (size_t)D || std::memcmp(last.data(), host, D * sizeof(double)) != 0) {
    last.assign(host, host + D);
    h2d(dev, last.data(), D, s);
  }
}

}  // namespace

namespace {

template <int K>
bool launch_stats_reg_k(int E, int B, const float* y, int D, const double* w, double scale, const double* tgt,
                        double* m, double* dis, double* l1, hipStream_t s) {
  const dim3 g(B), t(256);
  switch (E) {
    case 1: hipLaunchKernelGGL((ensemble_stats_reg_kernel<K, 1>), g, t, 0, s, y, E, B, D, w, scale, tgt, m, dis, l1); break;
    case 2: hipLaunchKernelGGL((ensemble_stats_reg_kernel<K, 2>), g, t, 0, s, y, E, B, D, w, scale, tgt, m, dis, l1); break;
    case 3: hipLaunchKernelGGL((ensemble_stats_reg_kernel<K, 3>), g, t, 0, s, y, E, B, D, w, scale, tgt, m, dis, l1); break;
    case 4: hipLaunchKernelGGL((ensemble_stats_reg_kernel<K, 4>), g, t, 0, s, y, E, B, D, w, scale, tgt, m, dis, l1); break;
    case 5: hipLaunchKernelGGL((ensemble_stats_reg_kernel<K, 5>), g, t, 0, s, y, E, B, D, w, scale, tgt, m, dis, l1); break;
    case 6: hipLaunchKernelGGL((ensemble_stats_reg_kernel<K, 6>), g, t, 0, s, y, E, B, D, w, scale, tgt, m, dis, l1); break;
    case 7: hipLaunchKernelGGL((ensemble_stats_reg_kernel<K, 7>), g, t, 0, s, y, E, B, D, w, scale, tgt, m, dis, l1); break;
    case 8: hipLaunchKernelGGL((ensemble_stats_reg_kernel<K, 8>), g, t, 0, s, y, E, B, D, w, scale, tgt, m, dis, l1); break;
    default: return false;
  }
  return true;
}

// The register kernel for E <= 8 members and K values per thread; false if none fits.
bool launch_stats_reg(int K, int E, int B, const float* y, int D, const double* w, double scale, const double* tgt,
                      double* m, double* dis, double* l1, hipStream_t s) {
  if (K == 1) return launch_stats_reg_k<1>(E, B, y, D, w, scale, tgt, m, dis, l1, s);
  if (K == 2) return launch_stats_reg_k<2>(E, B, y, D, w, scale, tgt, m, dis, l1, s);
  return launch_stats_reg_k<4>(E, B, y, D, w, scale, tgt, m, dis, l1, s);
}

// Conversion, forward pass and ensemble statistics of B staged samples.
void enqueue_predict(Mlp& net, int B, double scale, bool with_target, bool with_mean, hipStream_t s) {
  if (net.dims[0] % 8 == 0) {
    forward(net, B, s, static_cast<const float*>(net.xin.d), true);  // inputs read from the mapped buffer
  } else {
    convert_batch_input(net, B, s);
    forward(net, B, s, nullptr, true);
  }
  const int D = net.dims[net.L];
  const double* tgt = with_target ? net.st.as<double>() : nullptr;
  const float* y = net.out.as<float>();
  // disagreement and misfit are written straight into the mapped result buffer
  double *m = net.smean.as<double>(), *dis = static_cast<double*>(net.sout.d), *l1 = dis + B;
  const double* w = net.sw.as<double>();
  if (net.E <= 8 && D <= 1024 && !with_mean) m = nullptr;  // the register kernel writes no means then
  if (net.E > 8 || D > 1024 ||
      !launch_stats_reg(D <= 256 ? 1 : D <= 512 ? 2 : 4, net.E, B, y, D, w, scale, tgt, m, dis, l1, s))
    hipLaunchKernelGGL(ensemble_stats_kernel, dim3(B), dim3(256), 0, s, y, net.E, B, D, w, scale, tgt, m, dis, l1);
  HIP_CHECK(hipGetLastError());
}

// The captured enqueue_predict for this call's sizes and buffers.
hipGraphExec_t predict_graph(Mlp& net, int B, double scale, bool with_target, bool with_mean) {
  uint64_t sbits;
  std::memcpy(&sbits, &scale, sizeof(sbits));
  std::vector<uintptr_t> key = {(uintptr_t)B, (uintptr_t)sbits, (uintptr_t)with_target, (uintptr_t)with_mean,
                                (uintptr_t)net.xin.d, (uintptr_t)net.sout.d, (uintptr_t)net.smean.p,
                                (uintptr_t)net.sw.p, (uintptr_t)net.st.p, (uintptr_t)net.xb.p,
                                (uintptr_t)net.wb.p, (uintptr_t)net.w.p, (uintptr_t)net.out.p};
  for (int l = 0; l < net.L; ++l) {
    key.push_back((uintptr_t)net.zf[l].p);
    key.push_back((uintptr_t)net.hb[l].p);
  }
  for (const auto& g : net.graphs)
    if (g.key == key) return g.exec;
  if (!net.cap) HIP_CHECK(hipStreamCreateWithFlags(&net.cap, hipStreamNonBlocking));
  fwd_attr<EPI_BIAS_GELU>();  // kernel attributes are set outside the capture
  fwd_attr<EPI_BIAS>();
  fwd_attr<EPI_BIAS_GELU, true>();
  fwd_attr<EPI_BIAS, true>();
  HIP_CHECK(hipStreamBeginCapture(net.cap, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    enqueue_predict(net, B, scale, with_target, with_mean, net.cap);
  } catch (...) {
    (void)hipStreamEndCapture(net.cap, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  HIP_CHECK(hipStreamEndCapture(net.cap, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIP_CHECK(ie);
  if (net.graphs.size() >= 8) {
    // every call synchronizes before returning: no launch of the oldest graph is in flight
    HIP_CHECK(hipGraphExecDestroy(net.graphs.front().exec));
    net.graphs.erase(net.graphs.begin());
  }
  net.graphs.push_back(Mlp::PredictGraph{key, exec});
  return exec;
}

}  // namespace

void mlp_predict(Mlp& net, const float* x_host, int B, const double* w_host, double scale, const double* target_host,
                 double* mean_host, double* dis_host, double* l1_host, hipStream_t s) {
  mlp_predict_begin(net, x_host, B, w_host, scale, target_host, mean_host, s);
  mlp_predict_end(net, B, target_host != nullptr, dis_host, l1_host, s);
}

void mlp_predict_begin(Mlp& net, const float* x_host, int B, const double* w_host, double scale,
                       const double* target_host, double* mean_host, hipStream_t s) {
  alloc_acts(net, B);
  stage_batch_input(net, x_host, B);
  const int D = net.dims[net.L];
  upload_on_change(net.sw, net.sw_host, w_host, D, s);
  if (target_host) upload_on_change(net.st, net.st_host, target_host, D, s);
  net.smean.reserve((size_t)B * D * sizeof(double));
  net.sout.reserve((size_t)2 * B * sizeof(double));
  // per-kernel timing needs direct launches
  const bool with_target = target_host != nullptr, with_mean = mean_host != nullptr;
  if (timing_enabled() || !g_tuning.mlp_graph)
    enqueue_predict(net, B, scale, with_target, with_mean, s);
  else
    HIP_CHECK(hipGraphLaunch(predict_graph(net, B, scale, with_target, with_mean), s));
  if (mean_host) d2h(mean_host, net.smean.p, (size_t)B * D, s);
}

void mlp_predict_end(Mlp& net, int B, bool with_target, double* dis_host, double* l1_host, hipStream_t s) {
  stream_sync(s);
  const double* res = static_cast<const double*>(net.sout.h);
  std::memcpy(dis_host, res, B * sizeof(double));
  if (with_target && l1_host) std::memcpy(l1_host, res + B, B * sizeof(double));
}

void mlp_gradients(Mlp& net, const float* x_host, const float* y_host, int B, float* g_host, double* loss,
                   hipStream_t s) {
  alloc_acts(net, B);
  upload_batch_input(net, x_host, B, s);
  h2d(net.yb, y_host, (size_t)B * net.dims[net.L], s);
  HIP_CHECK(hipMemsetAsync(net.g.p, 0, (size_t)net.E * net.P * 4, s));
  forward(net, B, s);
  backward(net, B, s);
  if (g_host) d2h(g_host, net.g.p, (size_t)net.E * net.P, s);
  if (loss) d2h(loss, net.loss.p, net.E, s);
  stream_sync(s);
}

void mlp_set_data(Mlp& net, const float* X, const float* Y, int64_t n, hipStream_t s) {
  h2d(net.dX, X, (size_t)n * net.dims[0], s);
  h2d(net.dY, Y, (size_t)n * net.dims[net.L], s);
  stream_sync(s);
  net.n_data = n;
}

void mlp_train(Mlp& net, const int32_t* batches, int steps, int B, float lr, float b1, float b2, float eps,
               double* losses, hipStream_t s) {
  alloc_acts(net, B);
  DevBuf didx;
  h2d(didx, batches, (size_t)steps * B, s);
  const int64_t tot = (int64_t)net.E * net.P;
  DevBuf lbuf;
  lbuf.reserve((size_t)std::max(steps, 1) * net.E * 8);
  for (int it = 0; it < steps; ++it) {
    hipLaunchKernelGGL(gather_kernel, dim3(B), dim3(256), 0, s, net.dX.as<float>(), net.dims[0], net.dY.as<float>(),
                       net.dims[net.L], didx.as<int32_t>() + (size_t)it * B, B, net.xb.as<uint16_t>(),
                       net.yb.as<float>());
    HIP_CHECK(hipGetLastError());
    forward(net, B, s);
    backward(net, B, s);
    HIP_CHECK(hipMemcpyAsync(lbuf.as<double>() + (size_t)it * net.E, net.loss.p, net.E * 8, hipMemcpyDeviceToDevice,
                             s));
    ++net.step;
    const float bc1 = 1.0f - powf(b1, (float)net.step), bc2 = 1.0f - powf(b2, (float)net.step);
    TimedLaunch tl("mlp_adam", s);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(tot, 256)), dim3(256), 0, s, net.w.as<float>(), net.g.as<float>(),
                       net.m.as<float>(), net.v.as<float>(), net.wb.as<uint16_t>(), tot, lr, b1, b2, eps, bc1, bc2);
    HIP_CHECK(hipGetLastError());
  }
  if (losses) d2h(losses, lbuf.p, (size_t)steps * net.E, s);
  stream_sync(s);
}

}  // namespace imped
