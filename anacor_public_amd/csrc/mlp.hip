// Learned surrogate of the absorption factor: an MLP 60 -> 256 -> 256 -> 256
// -> 256 -> 1 (SiLU) from features of the incident and diffracted directions
// in the crystal frame to log A_h, trained and evaluated on the GPU by the kernels
// in this file.  The ray tracer (trace.hip) generates its training targets,
// judges its error and takes every reflection outside its envelope.
//
// Precision: the three 256x256 layers run on v_mfma_f32_16x16x32_bf16 with
// BF16 weights and FP32 accumulation.  Activations and back-propagated
// gradients stay FP32 and enter the MFMA as a BF16 hi + lo pair (two
// products), so their rounding puts no input-dependent noise on the output;
// weight gradients use hi/lo on both operands (three products).  In training
// the 60 -> 256 input layer and the 256 -> 1 output layer are FP32 VALU
// kernels; inference runs the input layer on MFMA as hi/lo pairs of both
// operands.  FP32 master weights, Adam.
//
// Parameter layout (FP32, ANACOR_MLP_PARAMS values):
//   W1[256][60] b1[256] W2[256][256] b2[256] W3[256][256] b3[256]
//   W4[256][256] b4[256] w5[256] b5[1]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

#include "errors.h"
#include "../../include/anacor_mi355x.h"

namespace anacor { namespace mlp {

using anacor::set_error;

constexpr int kIn = ANACOR_MLP_INPUTS;
constexpr int kH = ANACOR_MLP_WIDTH;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kIn;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH * kH;
constexpr int kOffW4 = kOffB3 + kH;
constexpr int kOffB4 = kOffW4 + kH * kH;
constexpr int kOffW5 = kOffB4 + kH;
constexpr int kOffB5 = kOffW5 + kH;
constexpr int kParams = kOffB5 + 1;
static_assert(kParams == ANACOR_MLP_PARAMS, "parameter count");
static_assert(kIn == 60 && kH == 256, "kernels assume 60 inputs and width 256");
constexpr int kInK = 64;          // the input layer's K on MFMA (inputs padded with zeros)

constexpr int kChunk = 1 << 16;   // largest training batch / gradient call

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// sigmoid with the hardware exp and reciprocal (v_exp_f32, v_rcp_f32: about
// 1 ulp each), the same function in training and inference.
__device__ __forceinline__ float sigmoidf_(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoidf_(z); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigmoidf_(z);
  return s * (1.0f + z * (1.0f - s));
}

// ------------------------------------------------------------------ GEMM
//
// C[M][N] = sum_k A(m,k) B(k,n).  A is stored K-contiguous ([M][K], A_KC) or
// M-contiguous ([K][M]); B K-contiguous ([N][K], B_KC) or N-contiguous
// ([K][N]).  FP32 operands (A_F32 / B_F32) are split into BF16 hi + lo in
// LDS; BF16 operands are used as they are.  Workgroup tile 64x64, 4 waves of
// 32x32 (2x2 MFMA 16x16x32 blocks), K step 32.
//
// MFMA 16x16x32 operand layout: lane l supplies row (A) / column (B) l % 16
// and the 8 k values 8 (l / 16) .. 8 (l / 16) + 7; the accumulator of lane l
// holds column l % 16, rows 4 (l / 16) .. 4 (l / 16) + 3.

constexpr int BM = 64, BN = 64, BK = 32, KP = 40;  // KP: padded LDS row (bf16)
constexpr int kSplitMax = 32;                      // K slices of the weight-gradient GEMMs

enum Epi { kEpiStore = 0, kEpiBiasSilu = 1, kEpiSiluBwd = 2, kEpiSplit = 3 };

struct GemmArgs {
  int M, N, K;
  const void* A;
  int lda;
  const void* B;
  int ldb;
  float* C;      // kEpiStore: C; kEpiBiasSilu: SiLU(z); kEpiSiluBwd: dZ; kEpiSplit: slices [z][M][N]
  int ldc;
  float* Z;      // kEpiBiasSilu: pre-activation out (nullable); kEpiSiluBwd: pre-activation in
  const float* bias;
  float* colsum; // kEpiSplit, A M-contiguous: per-slice sums of A over k, [z][M] (nullable)
  int k_chunk;   // kEpiSplit: k per slice (multiple of BK)
  int vec;       // 1: leading dimensions and bases allow 16-byte loads
};

// 8 consecutive elements of an FP32 or BF16 matrix as floats (`valid` of
// them in range; all 8 loaded as 16-byte vectors when valid == 8 and vec).
template <bool F32>
__device__ __forceinline__ void load8(const void* base, size_t idx, int valid, bool vec,
                                      float* v) {
  if (F32) {
    const float* p = (const float*)base + idx;
    if (valid == 8 && vec) {
      const float4 a = *(const float4*)p, b = *(const float4*)(p + 4);
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
      v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = e < valid ? p[e] : 0.f;
    }
  } else {
    const __bf16* p = (const __bf16*)base + idx;
    if (valid == 8 && vec) {
      const bf16x8 a = *(const bf16x8*)p;
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = (float)a[e];
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = e < valid ? (float)p[e] : 0.f;
    }
  }
}

// One operand tile (64 rows of the output dimension x 32 k, k < K) into LDS.
// KC: the operand is K-contiguous in global memory.  With SUM (M-contiguous
// operands only) the thread adds its 8 values to csum (its 8 rows, one k).
template <bool KC, bool F32, bool SUM>
__device__ __forceinline__ void stage(const void* base, int ld, int rows, int K, int r0, int k0,
                                      bool vec, __bf16 (*hi)[KP], __bf16 (*lo)[KP],
                                      float* csum) {
  const int t = (int)threadIdx.x;
  float v[8];
  if (KC) {  // thread: one row, 8 consecutive k
    const int r = t >> 2, kc = (t & 3) * 8;
    const int gr = r0 + r, gk = k0 + kc;
    const int valid = (gr < rows) ? min(8, max(0, K - gk)) : 0;
    if (valid > 0) {
      load8<F32>(base, (size_t)gr * ld + gk, valid, vec, v);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = 0.f;
    }
    bf16x8 h, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      h[e] = (__bf16)v[e];
      l[e] = (__bf16)(v[e] - (float)h[e]);
    }
    *(bf16x8*)&hi[r][kc] = h;
    if (F32) *(bf16x8*)&lo[r][kc] = l;
  } else {   // thread: one k, 8 consecutive rows
    const int k = t >> 3, rc = (t & 7) * 8;
    const int gk = k0 + k, gr = r0 + rc;
    const int valid = (gk < K) ? min(8, max(0, rows - gr)) : 0;
    if (valid > 0) {
      load8<F32>(base, (size_t)gk * ld + gr, valid, vec, v);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = 0.f;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const __bf16 h = (__bf16)v[e];
      hi[rc + e][k] = h;
      if (F32) lo[rc + e][k] = (__bf16)(v[e] - (float)h);
      if (SUM) csum[e] += v[e];
    }
  }
}

template <bool A_KC, bool B_KC, bool A_F32, bool B_F32, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
  constexpr bool SUM = (EPI == kEpiSplit) && !A_KC;
  __shared__ __attribute__((aligned(16))) __bf16 Ah[BM][KP];
  __shared__ __attribute__((aligned(16))) __bf16 Al[A_F32 ? BM : 1][KP];
  __shared__ __attribute__((aligned(16))) __bf16 Bh[BN][KP];
  __shared__ __attribute__((aligned(16))) __bf16 Bl[B_F32 ? BN : 1][KP];
  const int m0 = (int)blockIdx.y * BM, n0 = (int)blockIdx.x * BN;
  const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const bool vec = g.vec != 0;
  int kb = 0, ke = g.K;
  if (EPI == kEpiSplit) {
    kb = (int)blockIdx.z * g.k_chunk;
    ke = min(g.K, kb + g.k_chunk);
  }
  const bool do_sum = SUM && g.colsum != nullptr && blockIdx.x == 0;
  float csum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  const int r16 = lane & 15, kg = (lane >> 4) * 8;
  for (int k0 = kb; k0 < ke; k0 += BK) {
    if (do_sum)
      stage<A_KC, A_F32, true>(g.A, g.lda, g.M, ke, m0, k0, vec, Ah, Al, csum);
    else
      stage<A_KC, A_F32, false>(g.A, g.lda, g.M, ke, m0, k0, vec, Ah, Al, csum);
    stage<B_KC, B_F32, false>(g.B, g.ldb, g.N, ke, n0, k0, vec, Bh, Bl, nullptr);
    __syncthreads();
    bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      ah[i] = *(const bf16x8*)&Ah[wm + i * 16 + r16][kg];
      if (A_F32) al[i] = *(const bf16x8*)&Al[wm + i * 16 + r16][kg];
      bh[i] = *(const bf16x8*)&Bh[wn + i * 16 + r16][kg];
      if (B_F32) bl[i] = *(const bf16x8*)&Bl[wn + i * 16 + r16][kg];
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        if (A_F32)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        if (B_F32)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
      }
    __syncthreads();
  }
  if (SUM) {   // column sums of A over this slice: 32 k-threads per row group
    if (do_sum) {
      __shared__ float red[32][BM + 1];
      const int k = (int)threadIdx.x >> 3, rc = ((int)threadIdx.x & 7) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e) red[k][rc + e] = csum[e];
      __syncthreads();
      if (threadIdx.x < BM && m0 + (int)threadIdx.x < g.M) {
        float sum = 0.f;
        for (int q = 0; q < 32; ++q) sum += red[q][threadIdx.x];
        g.colsum[(size_t)blockIdx.z * g.M + m0 + threadIdx.x] = sum;
      }
    }
  }
  float* C = g.C;
  if (EPI == kEpiSplit) C += (size_t)blockIdx.z * g.M * g.ldc;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        const int n = n0 + wn + j * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        const float v = acc[i][j][r];
        const size_t o = (size_t)m * g.ldc + n;
        if (EPI == kEpiStore || EPI == kEpiSplit) {
          C[o] = v;
        } else if (EPI == kEpiBiasSilu) {
          const float z = v + g.bias[n];
          if (g.Z) g.Z[o] = z;
          C[o] = silu(z);
        } else {
          C[o] = v * silu_grad(g.Z[o]);
        }
      }
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

template <bool A_KC, bool B_KC, bool A_F32, bool B_F32, int EPI>
static void gemm(GemmArgs g, hipStream_t s, int slices = 1) {
  const int ea = A_F32 ? 4 : 8, eb = B_F32 ? 4 : 8;   // elements per 16 bytes
  g.vec = (g.lda % ea == 0 && g.ldb % eb == 0 && aligned16(g.A) && aligned16(g.B)) ? 1 : 0;
  const dim3 grid((unsigned)((g.N + BN - 1) / BN), (unsigned)((g.M + BM - 1) / BM),
                  (unsigned)slices);
  hipLaunchKernelGGL((gemm_kernel<A_KC, B_KC, A_F32, B_F32, EPI>), grid, dim3(256), 0, s, g);
}

// Sum of the K slices of a split GEMM (fixed order): out[m][n] (row stride
// ldo) and, when colsum is set, bias_out[m] from the per-slice column sums.
__global__ void split_reduce(const float* __restrict__ ws, int slices, int M, int N,
                             float* __restrict__ out, int ldo, const float* __restrict__ colsum,
                             float* __restrict__ bias_out) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < M * N) {
    const int m = i / N, n = i % N;
    float sum = 0.f;
    for (int z = 0; z < slices; ++z) sum += ws[(size_t)z * M * N + i];
    out[(size_t)m * ldo + n] = sum;
  }
  if (colsum && i < M) {
    float sum = 0.f;
    for (int z = 0; z < slices; ++z) sum += colsum[(size_t)z * M + i];
    bias_out[i] = sum;
  }
}

// ------------------------------------------------- fused inference kernel
//
// The whole network for a tile of 64 rows in one workgroup of 4 waves, on
// MFMA with the units as the M dimension (D[unit][row] = W . H^T): wave w
// owns units 64 (w % 4) .. + 63 of every layer for all 64 rows.  Weights come
// from L2 as 16-byte A fragments (BF16 copies; the 60 -> 256 layer as hi/lo
// pairs padded to K = 64); activations stay in LDS as BF16 hi/lo [row][unit],
// so a lane's 4 consecutive units of one row leave as one 8-byte store.  The
// inputs are staged as hi/lo in the first 64 columns of the same arrays and
// overwritten by the first layer's activations once every wave has read
// them.  The
// output layer is reduced in registers.  With `flag`/`A` the kernel also
// writes A = exp(y std + mean) inside the envelope and NaN outside.

constexpr int kHP = kH + 8;   // padded LDS row of activations (bf16)
// BF16 weight block: W2, W3, W4 [256][256], then W1 hi and lo [256][kInK].
constexpr size_t kWbW1h = (size_t)3 * kH * kH;
constexpr size_t kWbW1l = kWbW1h + (size_t)kH * kInK;
constexpr size_t kWbSize = kWbW1l + (size_t)kH * kInK;

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

template <int ROWS>
__global__ __launch_bounds__(256) void mlp_infer_kernel(
    const float* __restrict__ X, int64_t n, const float* __restrict__ P,
    const __bf16* __restrict__ Wb, const uint8_t* __restrict__ flag, double mean, double stdv,
    float* __restrict__ y, double* __restrict__ A) {
  constexpr int RT = ROWS / 16;   // 16-row MFMA tiles
  static_assert(ROWS % 16 == 0 && ROWS <= 256, "row tile");
  __shared__ __attribute__((aligned(16))) __bf16 Hh[ROWS][kHP];
  __shared__ __attribute__((aligned(16))) __bf16 Hl[ROWS][kHP];
  __shared__ float part[4][ROWS];
  static_assert(kInK <= kHP && kInK % 32 == 0, "input staging");
  const int t = (int)threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * ROWS;
  for (int i = t; i < ROWS * kInK; i += 256) {   // inputs as hi/lo, k >= kIn zero
    const int r = i / kInK, k = i % kInK;
    const float v = (k < kIn && r0 + r < n) ? X[(r0 + r) * kIn + k] : 0.f;
    const __bf16 h = (__bf16)v;
    Hh[r][k] = h;
    Hl[r][k] = (__bf16)(v - (float)h);
  }
  __syncthreads();
  const int lane = t & 63, wave = t >> 6;
  const int ub = wave * 64;                     // this wave's units
  const int r16 = lane & 15, kg = (lane >> 4) * 8, q4 = (lane >> 4) * 4;
  floatx4 acc[4][RT];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < RT; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  // layer 1: W1 (hi, lo) x X (hi, lo), three products, K = kInK in steps of 32
  for (int k0 = 0; k0 < kInK; k0 += 32) {
    bf16x8 wh[4], wl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      wh[i] = *(const bf16x8*)(Wb + kWbW1h + (size_t)(ub + i * 16 + r16) * kInK + k0 + kg);
      wl[i] = *(const bf16x8*)(Wb + kWbW1l + (size_t)(ub + i * 16 + r16) * kInK + k0 + kg);
    }
#pragma unroll
    for (int j = 0; j < RT; ++j) {
      const bf16x8 xh = *(const bf16x8*)&Hh[j * 16 + r16][k0 + kg];
      const bf16x8 xl = *(const bf16x8*)&Hl[j * 16 + r16][k0 + kg];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl[i], xh, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[i], xl, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[i], xh, acc[i][j], 0, 0, 0);
      }
    }
  }
  __syncthreads();   // every wave has read the inputs: the arrays take layer 1's output
  const int offb[4] = {kOffB1, kOffB2, kOffB3, kOffB4};
  for (int l = 0; l < 4; ++l) {
    if (l > 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < RT; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
      const __bf16* W = Wb + (size_t)(l - 1) * kH * kH;
      for (int k0 = 0; k0 < kH; k0 += 32) {
        bf16x8 w[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
          w[i] = *(const bf16x8*)(W + (size_t)(ub + i * 16 + r16) * kH + k0 + kg);
#pragma unroll
        for (int j = 0; j < RT; ++j) {
          const bf16x8 hh = *(const bf16x8*)&Hh[j * 16 + r16][k0 + kg];
          const bf16x8 hl = *(const bf16x8*)&Hl[j * 16 + r16][k0 + kg];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[i], hl, acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[i], hh, acc[i][j], 0, 0, 0);
          }
        }
      }
      __syncthreads();   // every wave has read this layer's input
    }
    if (l == 3) break;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int u = ub + i * 16 + q4;
      const float4 bias = *(const float4*)(P + offb[l] + u);
#pragma unroll
      for (int j = 0; j < RT; ++j) {
        const int row = j * 16 + r16;
        bf16x4 hh, hl;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float h = silu(acc[i][j][r] + (r == 0 ? bias.x : r == 1 ? bias.y
                                                : r == 2 ? bias.z : bias.w));
          hh[r] = (__bf16)h;
          hl[r] = (__bf16)(h - (float)hh[r]);
        }
        *(bf16x4*)&Hh[row][u] = hh;
        *(bf16x4*)&Hl[row][u] = hl;
      }
    }
    __syncthreads();
  }
  // output layer: per row, this wave's 64 units of SiLU(z4) . w5
#pragma unroll
  for (int j = 0; j < RT; ++j) {
    float v = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int u = ub + i * 16 + q4;
      const float4 bias = *(const float4*)(P + kOffB4 + u);
      const float4 w5 = *(const float4*)(P + kOffW5 + u);
      v = fmaf(silu(acc[i][j][0] + bias.x), w5.x, v);
      v = fmaf(silu(acc[i][j][1] + bias.y), w5.y, v);
      v = fmaf(silu(acc[i][j][2] + bias.z), w5.z, v);
      v = fmaf(silu(acc[i][j][3] + bias.w), w5.w, v);
    }
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    if (lane < 16) part[wave][j * 16 + lane] = v;
  }
  __syncthreads();
  if (t < ROWS && r0 + t < n) {
    const float yv = ((part[0][t] + part[1][t]) + (part[2][t] + part[3][t])) + P[kOffB5];
    y[r0 + t] = yv;
    if (A) A[r0 + t] = flag[r0 + t] ? exp((double)yv * stdv + mean) : __builtin_nan("");
  }
}

// ------------------------------------------------------- small kernels

// H1 = SiLU(X W1^T + b1), FP32; one thread per (row, unit).  With perm
// (training) row b reads X[perm[b]] and T[perm[b]] and copies them to
// Xb[b] / Tb[b] for the backward pass.
__global__ void layer1_fwd(const float* __restrict__ X, const int* __restrict__ perm,
                           const float* __restrict__ T, int n, const float* __restrict__ P,
                           float* __restrict__ Xb, float* __restrict__ Tb,
                           float* __restrict__ Z1, float* __restrict__ H1) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * kH) return;
  const int b = (int)(t / kH), u = (int)(t % kH);
  const int src = perm ? perm[b] : b;
  const float* w = P + kOffW1 + u * kIn;
  const float* x = X + (size_t)src * kIn;
  float z = P[kOffB1 + u];
#pragma unroll
  for (int k = 0; k < kIn; ++k) z = fmaf(x[k], w[k], z);
  if (Z1) Z1[t] = z;
  H1[t] = silu(z);
  if (perm) {
    if (u < kIn) Xb[(size_t)b * kIn + u] = x[u];
    if (u == kIn) Tb[b] = T[src];
  }
}

// y = H4 w5 + b5, FP32; one wave per row.
__device__ __forceinline__ float head_dot(const float* __restrict__ h, const float* __restrict__ P,
                                          int lane) {
  float s = 0.f;
#pragma unroll
  for (int k = lane; k < kH; k += 64) s = fmaf(h[k], P[kOffW5 + k], s);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  return s + P[kOffB5];
}

__global__ void out_fwd(const float* __restrict__ H4, int n, const float* __restrict__ P,
                        float* __restrict__ y) {
  const int row = (int)blockIdx.x * 4 + ((int)threadIdx.x >> 6);
  const int lane = (int)threadIdx.x & 63;
  if (row >= n) return;
  const float s = head_dot(H4 + (size_t)row * kH, P, lane);
  if (lane == 0) y[row] = s;
}

// Training head, one wave per row: y, the mean squared error term, dy =
// 2 (y - t) / n, dZ4 = dy w5 silu'(Z4); per-block partial sums of the loss
// and of dy (db5) in part[block][2].
__global__ void head_train(const float* __restrict__ H4, const float* __restrict__ Z4,
                           const float* __restrict__ T, int n, const float* __restrict__ P,
                           float* __restrict__ y, float* __restrict__ dy,
                           float* __restrict__ dZ4, float* __restrict__ part) {
  __shared__ float red[4][2];
  const int w = (int)threadIdx.x >> 6;
  const int row = (int)blockIdx.x * 4 + w;
  const int lane = (int)threadIdx.x & 63;
  float l = 0.f, g = 0.f;
  if (row < n) {
    const float yv = head_dot(H4 + (size_t)row * kH, P, lane);
    const float d = yv - T[row];
    g = 2.f * d / (float)n;
    l = d * d / (float)n;
    if (lane == 0) {
      y[row] = yv;
      dy[row] = g;
    }
    const float* z = Z4 + (size_t)row * kH;
    float* o = dZ4 + (size_t)row * kH;
#pragma unroll
    for (int k = lane; k < kH; k += 64) o[k] = g * P[kOffW5 + k] * silu_grad(z[k]);
  }
  if (lane == 0) {
    red[w][0] = l;
    red[w][1] = g;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
    part[2 * blockIdx.x + 1] = (red[0][1] + red[1][1]) + (red[2][1] + red[3][1]);
  }
}

// loss[0] += sum of the loss partials; G[b5] = sum of the dy partials.
__global__ void head_finish(const float* __restrict__ part, int nb, float* __restrict__ loss,
                            float* __restrict__ gb5) {
  float l = 0.f, g = 0.f;
  for (int i = (int)threadIdx.x; i < nb; i += 64) {
    l += part[2 * i];
    g += part[2 * i + 1];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    l += __shfl_xor(l, off);
    g += __shfl_xor(g, off);
  }
  if (threadIdx.x == 0) {
    loss[0] += l;
    gb5[0] = g;
  }
}

// BF16 copies of W2, W3, W4 (round to nearest even) and W1 as hi/lo pairs
// padded to K = kInK (the inference kernel's layout).
__device__ __forceinline__ void store_bf16_copy(int i, float p, __bf16* __restrict__ Wb) {
  int l = -1, j = 0;
  if (i >= kOffW2 && i < kOffB2) { l = 0; j = i - kOffW2; }
  else if (i >= kOffW3 && i < kOffB3) { l = 1; j = i - kOffW3; }
  else if (i >= kOffW4 && i < kOffB4) { l = 2; j = i - kOffW4; }
  if (l >= 0) {
    Wb[(size_t)l * kH * kH + j] = (__bf16)p;
  } else if (i >= kOffW1 && i < kOffB1) {
    const int u = (i - kOffW1) / kIn, k = (i - kOffW1) % kIn;
    const __bf16 h = (__bf16)p;
    Wb[kWbW1h + (size_t)u * kInK + k] = h;
    Wb[kWbW1l + (size_t)u * kInK + k] = (__bf16)(p - (float)h);
  }
}

// Adam on every parameter, refreshing the BF16 copies of the weights.
__global__ void adam_step(float* __restrict__ P, const float* __restrict__ G,
                          float* __restrict__ m, float* __restrict__ v, __bf16* __restrict__ Wb,
                          float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= kParams) return;
  const float g = G[i];
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float p = P[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  P[i] = p;
  store_bf16_copy(i, p, Wb);
}

__global__ void weights_to_bf16(const float* __restrict__ P, __bf16* __restrict__ Wb) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < kParams) store_bf16_copy(i, P[i], Wb);
}

__global__ void zero_floats(float* __restrict__ p, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) p[i] = 0.f;
}

// ------------------------------------------------------------- features
//
// The crystal frame of the tracer's setup: M = Rodrigues(omega_axis, omega) F,
// incident x = M^T raw_xray, diffracted s = M^T s1 (unit vectors here).
// Network inputs (ANACOR_MLP_INPUTS per row): x (3), s (3), then six more by
// feature set:
//   kFeaturesHarmonic: per vector v, 2 v0 v1, v0^2 - v1^2, 2 v2^2 - 1 (the
//     sin/cos of twice its azimuth and polar angle, scaled);
//   kFeaturesChamber: the side of the planes on which the tracer's discrete
//     rules switch (+-1): sign x2, sign(|x1| - |x2|), sign(|x0| - |x1|),
//     sign s2, sign(|s0| - |s1|), sign(|s1| - |s2|).  The floor/round choice
//     of cal_coord follows the sign of theta, i.e. of v2
//     (ray_tracing_basic.c:13-50, :428-438), and the absorption factor jumps
//     across those planes and some of the diagonal ones;
//   kFeaturesResonance: the chamber signs, and in [12:60] for each ray and
//     component pair the ratio t of the smaller to the larger magnitude and
//     2 |q t - rint(q t)| for q = 2..8: V-shaped kinks at every ratio p / q,
//     where the voxel steps of all crystal voxels fall into step together
//     and the traced absorption factor moves faster than elsewhere (the other
//     sets leave [12:60] zero).
// Envelope invariants: axial = x . u (u: crystal-frame rotation axis),
// two_theta = angle(x, s), alpha = azimuth of x about u (basis e1, e2),
// beta = azimuth of s about x (basis u x x, x (u x x)).  The trust grid
// splits (alpha, two_theta, beta) into cells; a reflection is inside the
// envelope when axial and two_theta are in range and its cell is trusted.

enum { kFeaturesHarmonic = 0, kFeaturesChamber = 1, kFeaturesResonance = 2 };
static_assert(kFeaturesHarmonic == ANACOR_MLP_FEATURES_HARMONIC &&
              kFeaturesChamber == ANACOR_MLP_FEATURES_CHAMBER &&
              kFeaturesResonance == ANACOR_MLP_FEATURES_RESONANCE, "feature set ids");

// Inputs [12:60] of kFeaturesResonance from the FP32 components v[6] (x, s):
// ray r, pair p (0, 1), (0, 2), (1, 2), k = 3 r + p: o[k] = t, o[6 + 7 k +
// q - 2] = 2 |q t - rint(q t)| (one rounding per operation, the last two
// exact).
__device__ __forceinline__ void resonance_inputs(const float* v, float* o) {
  constexpr int pi[3] = {0, 0, 1}, pj[3] = {1, 2, 2};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int r = k / 3, p = k % 3;
    const float a = fabsf(v[3 * r + pi[p]]), b = fabsf(v[3 * r + pj[p]]);
    const float hi = fmaxf(a, b);
    const float t = hi > 0.0f ? fminf(a, b) / hi : 0.0f;
    o[k] = t;
#pragma unroll
    for (int q = 2; q <= 8; ++q) {
      const float u = (float)q * t;
      o[6 + 7 * k + q - 2] = 2.0f * fabsf(u - rintf(u));
    }
  }
}

// Resonance bands (anacor_band_t): near a small rational ratio of two
// components of a ray's direction the voxel steps of all crystal voxels fall
// into step together, and the traced absorption factor oscillates (moire)
// faster than the training set samples it.  Reflections inside a band go to
// the tracer.
struct FeatureGeo {
  double raw[3], axis[3], F[9], u[3], e1[3], e2[3];
  double axial_lo, axial_hi, tth_lo, tth_hi;
  int cells[3];
  int feature_set;
  int use_env;
  int n_bands;
  const uint8_t* trust;   // device, cells[0] * cells[1] * cells[2]; null: all trusted
  anacor_band_t bands[ANACOR_MLP_MAX_BANDS];
};

// |min(|v_i|, |v_j|) / max(|v_i|, |v_j|) - ratio| < width for some band.
__device__ __forceinline__ bool in_band(const FeatureGeo& g, const double* x, const double* s) {
  for (int k = 0; k < g.n_bands; ++k) {
    const anacor_band_t& b = g.bands[k];
    const double* v = b.ray ? s : x;
    const double p = fabs(v[b.i]), q = fabs(v[b.j]);
    const double hi = fmax(p, q);
    const double t = hi > 0.0 ? fmin(p, q) / hi : 0.0;
    if (fabs(t - b.ratio) < b.width) return true;
  }
  return false;
}

__device__ void rodrigues_dev(const double* a, double t, double* r) {
  const double x = a[0], y = a[1], z = a[2];
  const double c = cos(t), s = sin(t), k = 1.0 - c;
  r[0] = c + x * x * k;      r[1] = x * y * k - z * s;  r[2] = y * s + x * z * k;
  r[3] = z * s + x * y * k;  r[4] = c + y * y * k;      r[5] = -x * s + y * z * k;
  r[6] = -y * s + x * z * k; r[7] = x * s + y * z * k;  r[8] = c + z * z * k;
}

__device__ __forceinline__ double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ float sgn(double v) { return v > 0.0 ? 1.f : (v < 0.0 ? -1.f : 0.f); }

__device__ __forceinline__ int bin(double v, double lo, double hi, int n) {
  if (!(hi > lo)) return 0;
  const double f = (v - lo) / (hi - lo) * (double)n;
  const int k = (int)floor(f);
  return k < 0 ? 0 : (k >= n ? n - 1 : k);
}

// inv[R][4] = (axial, two_theta, alpha, beta); cell[R]; flag[R] (any nullable).
__global__ void features_kernel(const double* __restrict__ s1, const double* __restrict__ omega,
                                int64_t R, FeatureGeo g, float* __restrict__ X,
                                double* __restrict__ inv, int32_t* __restrict__ cell,
                                uint8_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  double rw[9], m[9];
  rodrigues_dev(g.axis, omega[i], rw);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      m[a * 3 + b] = rw[a * 3 + 0] * g.F[0 * 3 + b] + rw[a * 3 + 1] * g.F[1 * 3 + b] +
                     rw[a * 3 + 2] * g.F[2 * 3 + b];
  double x[3], s[3];
  const double* sv = s1 + 3 * i;
  for (int b = 0; b < 3; ++b) {   // M^T v
    x[b] = m[0 * 3 + b] * g.raw[0] + m[1 * 3 + b] * g.raw[1] + m[2 * 3 + b] * g.raw[2];
    s[b] = m[0 * 3 + b] * sv[0] + m[1 * 3 + b] * sv[1] + m[2 * 3 + b] * sv[2];
  }
  const double nx = 1.0 / sqrt(dot3(x, x));
  const double ns = 1.0 / sqrt(dot3(s, s));
  for (int b = 0; b < 3; ++b) {
    x[b] *= nx;
    s[b] *= ns;
  }
  if (X) {
    float* o = X + (int64_t)kIn * i;
    o[0] = (float)x[0]; o[1] = (float)x[1]; o[2] = (float)x[2];
    o[3] = (float)s[0]; o[4] = (float)s[1]; o[5] = (float)s[2];
    if (g.feature_set != kFeaturesHarmonic) {
      o[6] = sgn(x[2]);
      o[7] = sgn(fabs(x[1]) - fabs(x[2]));
      o[8] = sgn(fabs(x[0]) - fabs(x[1]));
      o[9] = sgn(s[2]);
      o[10] = sgn(fabs(s[0]) - fabs(s[1]));
      o[11] = sgn(fabs(s[1]) - fabs(s[2]));
    } else {
      o[6] = (float)(2.0 * x[0] * x[1]);
      o[7] = (float)(x[0] * x[0] - x[1] * x[1]);
      o[8] = (float)(2.0 * x[2] * x[2] - 1.0);
      o[9] = (float)(2.0 * s[0] * s[1]);
      o[10] = (float)(s[0] * s[0] - s[1] * s[1]);
      o[11] = (float)(2.0 * s[2] * s[2] - 1.0);
    }
    if (g.feature_set == kFeaturesResonance) {
      resonance_inputs(o, o + 12);
    } else {
      for (int k = 12; k < kIn; ++k) o[k] = 0.f;
    }
  }
  const double axial = dot3(x, g.u);
  const double c2t = dot3(x, s);
  const double tth = acos(c2t > 1.0 ? 1.0 : (c2t < -1.0 ? -1.0 : c2t));
  const double alpha = atan2(dot3(x, g.e2), dot3(x, g.e1));
  double f1[3] = {g.u[1] * x[2] - g.u[2] * x[1], g.u[2] * x[0] - g.u[0] * x[2],
                  g.u[0] * x[1] - g.u[1] * x[0]};
  const double nf = sqrt(dot3(f1, f1));
  double beta = 0.0;
  if (nf > 1e-12) {
    for (int b = 0; b < 3; ++b) f1[b] /= nf;
    const double f2[3] = {x[1] * f1[2] - x[2] * f1[1], x[2] * f1[0] - x[0] * f1[2],
                          x[0] * f1[1] - x[1] * f1[0]};
    double p[3];
    for (int b = 0; b < 3; ++b) p[b] = s[b] - c2t * x[b];
    beta = atan2(dot3(p, f2), dot3(p, f1));
  }
  if (inv) {
    inv[4 * i + 0] = axial;
    inv[4 * i + 1] = tth;
    inv[4 * i + 2] = alpha;
    inv[4 * i + 3] = beta;
  }
  if (g.use_env) {
    const int ia = bin(alpha, -M_PI, M_PI, g.cells[0]);
    const int it = bin(tth, g.tth_lo, g.tth_hi, g.cells[1]);
    const int ib = bin(beta, -M_PI, M_PI, g.cells[2]);
    const int c = (ia * g.cells[1] + it) * g.cells[2] + ib;
    if (cell) cell[i] = c;
    if (flag) {
      const bool in = axial >= g.axial_lo && axial <= g.axial_hi && tth >= g.tth_lo &&
                      tth <= g.tth_hi && (!g.trust || g.trust[c] != 0) &&   // NaN compares false
                      !in_band(g, x, s);
      flag[i] = in ? 1 : 0;
    }
  }
}

// ------------------------------------------------------------- risk trees
//
// A boosted ensemble of complete binary trees over 60 resonance features of
// the two unit vectors (layout in anacor_mi355x.h): the ratios t of two
// components of a ray and t's distance to the nearest p / q, q <= 8.  Near
// such ratios the voxel steps of all crystal voxels fall into step together
// and the traced A_h carries a moire the network does not resolve; the trees
// score where its error is likely to exceed the bound, and rows scoring at
// least `cut` leave the envelope (the tracer takes them).  The trees are
// fitted on the host (risk_fit.cpp) to the training rows' errors.
//
// Every feature is one correctly rounded FP32 operation chain (|.|, min, max,
// one division, q * t, rint, an exact subtraction, one multiplication by the
// rounded 1/q), so a host restatement in FP32 gives the same bits.

constexpr int kRiskF = ANACOR_RISK_FEATURES;
static_assert(kRiskF == 60, "risk feature layout");

struct RiskNode {
  float thr;
  int feat;
};

struct RiskArgs {
  const RiskNode* nodes;   // [n_trees][n_int]
  const float* leaf;       // [n_trees][n_int + 1]
  int n_trees, depth, n_int;
  float base, cut;
};

__device__ __forceinline__ void risk_features(const float* v, float* o, int stride) {
  constexpr float inv_q[8] = {1.0f, 1.0f / 2.0f, 1.0f / 3.0f, 1.0f / 4.0f,
                              1.0f / 5.0f, 1.0f / 6.0f, 1.0f / 7.0f, 1.0f / 8.0f};
  constexpr int pi[3] = {0, 0, 1}, pj[3] = {1, 2, 2};
  for (int k = 0; k < 6; ++k) o[k * stride] = v[k];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int r = k / 3, p = k % 3;
    const float a = fabsf(v[3 * r + pi[p]]), b = fabsf(v[3 * r + pj[p]]);
    const float hi = fmaxf(a, b);
    const float t = hi > 0.0f ? fminf(a, b) / hi : 0.0f;
    o[(6 + k) * stride] = t;
#pragma unroll
    for (int q = 1; q <= 8; ++q) {
      const float u = (float)q * t;
      o[(12 + 8 * k + q - 1) * stride] = fabsf(u - rintf(u)) * inv_q[q - 1];
    }
  }
}

// One thread per row of network inputs X[n][kIn] (x, s in the first six): the features into LDS
// (column per thread, so the 64 lanes of a wave read 64 consecutive banks),
// optionally out to feat_out[n][60]; with trees the score (score[n],
// nullable) and flag[i] cleared when score >= cut (flag nullable).  No thread
// reads another's column, so the kernel has no barrier.
__global__ __launch_bounds__(256) void risk_kernel(const float* __restrict__ X, int64_t n,
                                                   RiskArgs r, float* __restrict__ feat_out,
                                                   float* __restrict__ score,
                                                   uint8_t* __restrict__ flag) {
  __shared__ float lds[kRiskF * 256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v[6];
  for (int k = 0; k < 6; ++k) v[k] = X[i * kIn + k];
  float* col = lds + threadIdx.x;
  risk_features(v, col, 256);
  if (feat_out)
    for (int k = 0; k < kRiskF; ++k) feat_out[i * kRiskF + k] = col[k * 256];
  if (r.n_trees <= 0) return;
  // eight trees walked side by side (eight independent chains of node loads
  // in flight; 1M rows of 200 depth-6 trees: 0.79 ms against 1.62 ms one
  // tree at a time, 1.00 ms with four, 0.77 ms with sixteen), their leaves
  // added in tree order as one at a time would
  constexpr int kT = 8;
  float s = r.base;
  int t = 0;
  for (; t + kT <= r.n_trees; t += kT) {
    int j[kT] = {};
    for (int d = 0; d < r.depth; ++d) {
      RiskNode e[kT];
#pragma unroll
      for (int q = 0; q < kT; ++q) e[q] = r.nodes[(size_t)(t + q) * r.n_int + j[q]];
#pragma unroll
      for (int q = 0; q < kT; ++q)
        j[q] = 2 * j[q] + 1 + (col[e[q].feat * 256] > e[q].thr ? 1 : 0);
    }
#pragma unroll
    for (int q = 0; q < kT; ++q) s += r.leaf[(size_t)(t + q) * (r.n_int + 1) + (j[q] - r.n_int)];
  }
  for (; t < r.n_trees; ++t) {
    const RiskNode* nd = r.nodes + (size_t)t * r.n_int;
    int j = 0;
    for (int d = 0; d < r.depth; ++d) {
      const RiskNode e = nd[j];
      j = 2 * j + 1 + (col[e.feat * 256] > e.thr ? 1 : 0);
    }
    s += r.leaf[(size_t)t * (r.n_int + 1) + (j - r.n_int)];
  }
  if (score) score[i] = s;
  if (flag && s >= r.cut) flag[i] = 0;
}

// ---------------------------------------------------------------- model

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  int grow(size_t want) {
    if (want <= n) return ANACOR_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, want * sizeof(T)) != hipSuccess) {
      set_error(ANACOR_ERR_HIP, "hipMalloc of %zu bytes failed", want * sizeof(T));
      return ANACOR_ERR_HIP;
    }
    n = want;
    return ANACOR_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

struct Model {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  DevBuf<float> P, G, m, v;       // FP32 master parameters, gradients, Adam moments
  DevBuf<__bf16> Wb;              // BF16 copies of W2, W3, W4
  int cap = 0;                    // rows of the activation buffers
  DevBuf<float> Z[4], H[4], dA, dB, dy, xb, tb, part, loss;
  DevBuf<float> ws, colws;        // split-K slices [kSplitMax][256][256], [kSplitMax][256]
  // feature / inference / training-set buffers (grown on demand)
  DevBuf<double> s1, omega, inv, A;
  DevBuf<float> X, y, T;
  DevBuf<uint8_t> flag, trust;
  DevBuf<int32_t> cell;
  int feature_set = kFeaturesHarmonic;
  bool has_env = false, trust_set = false;
  anacor_envelope_t env{};
  int n_bands = 0;
  anacor_band_t bands[ANACOR_MLP_MAX_BANDS] = {};
  DevBuf<RiskNode> r_nodes;       // risk trees (none while r_trees == 0)
  DevBuf<float> r_leaf, r_feat, r_score;
  int r_trees = 0, r_depth = 0;
  float r_base = 0.f, r_cut = 0.f;
  double y_mean = 0.0, y_std = 1.0;
  long long adam_t = 0;
  double last_mlp_ms = 0.0, last_total_ms = 0.0;
  int64_t last_rows = 0, last_in = 0;
  std::mutex mu;
};

#define MLP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      set_error(ANACOR_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));       \
      return ANACOR_ERR_HIP;                                                          \
    }                                                                                 \
  } while (0)

#define MLP_CHECK(expr)                  \
  do {                                   \
    int rc_ = (expr);                    \
    if (rc_ != ANACOR_OK) return rc_;    \
  } while (0)

static unsigned nblocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

static int ensure_rows(Model* M, int n) {
  if (n <= M->cap) return ANACOR_OK;
  for (int l = 0; l < 4; ++l) {
    MLP_CHECK(M->Z[l].grow((size_t)n * kH));
    MLP_CHECK(M->H[l].grow((size_t)n * kH));
  }
  MLP_CHECK(M->dA.grow((size_t)n * kH));
  MLP_CHECK(M->dB.grow((size_t)n * kH));
  MLP_CHECK(M->dy.grow((size_t)n));
  MLP_CHECK(M->xb.grow((size_t)n * kIn));
  MLP_CHECK(M->tb.grow((size_t)n));
  MLP_CHECK(M->part.grow((size_t)2 * nblocks(n, 4)));
  MLP_CHECK(M->ws.grow((size_t)kSplitMax * kH * kH));
  MLP_CHECK(M->colws.grow((size_t)kSplitMax * kH));
  M->cap = n;
  return ANACOR_OK;
}

static void refresh_bf16(Model* M) {
  hipLaunchKernelGGL(weights_to_bf16, dim3(nblocks(kParams, 256)), dim3(256), 0, M->stream,
                     M->P.p, M->Wb.p);
}

// Rows per inference workgroup (LDS: 2 x ROWS x 264 BF16 + 4 x ROWS floats).
constexpr int kInferRows = 64;

// Inference over n rows of device features X: y[n] and, with flag/A, the
// absorption factors (fused inference kernel, one launch).
static void infer(Model* M, const float* X, int64_t n, const uint8_t* flag, float* y,
                  double* A) {
  if (n <= 0) return;
  hipLaunchKernelGGL(mlp_infer_kernel<kInferRows>, dim3(nblocks(n, kInferRows)), dim3(256), 0,
                     M->stream, X, n, M->P.p, M->Wb.p, flag, M->y_mean, M->y_std, y, A);
}

// K slices of a weight-gradient GEMM with K = n rows.
static void split_shape(int n, int* slices, int* chunk) {
  int s = std::max(1, std::min(kSplitMax, (n + 63) / 64));
  int c = ((n + s - 1) / s + BK - 1) / BK * BK;
  *chunk = c;
  *slices = (n + c - 1) / c;
}

// dW[M][N] (row stride ldo) = A^T B over n rows, A [n][M] and B [n][N] FP32,
// plus bias[M] = column sums of A when bias is set.
static void weight_grad(Model* M, const float* A, int Mdim, const float* B, int N, int n,
                        float* dW, int ldo, float* bias) {
  int slices, chunk;
  split_shape(n, &slices, &chunk);
  GemmArgs g{};
  g.M = Mdim; g.N = N; g.K = n;
  g.A = A; g.lda = Mdim;
  g.B = B; g.ldb = N;
  g.C = M->ws.p; g.ldc = N;
  g.colsum = bias ? M->colws.p : nullptr;
  g.k_chunk = chunk;
  gemm<false, false, true, true, kEpiSplit>(g, M->stream, slices);
  hipLaunchKernelGGL(split_reduce, dim3(nblocks((long long)Mdim * N, 256)), dim3(256), 0,
                     M->stream, M->ws.p, slices, Mdim, N, dW, ldo,
                     bias ? M->colws.p : (const float*)nullptr, bias);
}

// Forward and backward pass of one batch of n rows: gradients of the mean
// squared error into G, the batch loss added to loss[0].  With perm the rows
// are X[perm[b]] / T[perm[b]] (gathered on the way), else X / T themselves.
static void fwd_bwd(Model* M, const float* X, const int* perm, const float* T, int n) {
  hipStream_t s = M->stream;
  float* G = M->G.p;
  hipLaunchKernelGGL(layer1_fwd, dim3(nblocks((long long)n * kH, 256)), dim3(256), 0, s, X, perm,
                     T, n, M->P.p, M->xb.p, M->tb.p, M->Z[0].p, M->H[0].p);
  const float* Xr = perm ? M->xb.p : X;
  const float* Tr = perm ? M->tb.p : T;
  const int offw[3] = {kOffW2, kOffW3, kOffW4};
  const int offb[3] = {kOffB2, kOffB3, kOffB4};
  for (int l = 0; l < 3; ++l) {
    GemmArgs g{};
    g.M = n; g.N = kH; g.K = kH;
    g.A = M->H[l].p; g.lda = kH;
    g.B = M->Wb.p + (size_t)l * kH * kH; g.ldb = kH;
    g.C = M->H[l + 1].p; g.ldc = kH;
    g.Z = M->Z[l + 1].p;
    g.bias = M->P.p + offb[l];
    gemm<true, true, true, false, kEpiBiasSilu>(g, s);
  }
  // head: y, loss, dy, dZ4; then db5 and the loss
  float* dz = M->dA.p;
  float* dnext = M->dB.p;
  const unsigned hb = nblocks(n, 4);
  hipLaunchKernelGGL(head_train, dim3(hb), dim3(256), 0, s, M->H[3].p, M->Z[3].p, Tr, n, M->P.p,
                     M->y.p, M->dy.p, dz, M->part.p);
  hipLaunchKernelGGL(head_finish, dim3(1), dim3(64), 0, s, M->part.p, (int)hb, M->loss.p,
                     G + kOffB5);
  weight_grad(M, M->H[3].p, kH, M->dy.p, 1, n, G + kOffW5, 1, nullptr);   // dw5 = H4^T dy
  for (int l = 2; l >= 0; --l) {
    // dW[out][in] = sum_b dZ[b][out] H[b][in], db = sum_b dZ[b]
    weight_grad(M, dz, kH, M->H[l].p, kH, n, G + offw[l], kH, G + offb[l]);
    // dZ_prev = (dZ W) silu'(Z_prev): M = batch, N = in, K = out
    GemmArgs gd{};
    gd.M = n; gd.N = kH; gd.K = kH;
    gd.A = dz; gd.lda = kH;                             // [M][K], K-contiguous
    gd.B = M->Wb.p + (size_t)l * kH * kH; gd.ldb = kH;  // W[out][in] = [K][N]
    gd.C = dnext; gd.ldc = kH;
    gd.Z = M->Z[l].p;
    gemm<true, false, true, false, kEpiSiluBwd>(gd, s);
    std::swap(dz, dnext);
  }
  weight_grad(M, dz, kH, Xr, kIn, n, G + kOffW1, kIn, G + kOffB1);   // input layer
}

static void adam(Model* M, float lr) {
  const double b1 = 0.9, b2 = 0.999;
  ++M->adam_t;
  const float c1 = (float)(1.0 - std::pow(b1, (double)M->adam_t));
  const float c2 = (float)(1.0 - std::pow(b2, (double)M->adam_t));
  hipLaunchKernelGGL(adam_step, dim3(nblocks(kParams, 256)), dim3(256), 0, M->stream, M->P.p,
                     M->G.p, M->m.p, M->v.p, M->Wb.p, lr, (float)b1, (float)b2, 1e-8f, c1, c2);
}

// The goniometer of a call and the model's envelope (if set) for the
// features kernel.  Without an envelope the axis is the unit F^T omega_axis.
static int geo_from(const Model* M, const double* raw, const double* axis, const double* F,
                    FeatureGeo* g) {
  if (!raw || !axis || !F) {
    set_error(ANACOR_ERR_ARG, "null goniometer arguments");
    return ANACOR_ERR_ARG;
  }
  std::memset(g, 0, sizeof(*g));
  std::memcpy(g->raw, raw, sizeof(g->raw));
  std::memcpy(g->axis, axis, sizeof(g->axis));
  std::memcpy(g->F, F, sizeof(g->F));
  g->feature_set = M->feature_set;
  if (M->has_env) {
    const anacor_envelope_t& e = M->env;
    std::memcpy(g->u, e.axis, sizeof(g->u));
    std::memcpy(g->e1, e.e1, sizeof(g->e1));
    std::memcpy(g->e2, e.e2, sizeof(g->e2));
    g->axial_lo = e.axial[0];
    g->axial_hi = e.axial[1];
    g->tth_lo = e.two_theta[0];
    g->tth_hi = e.two_theta[1];
    for (int k = 0; k < 3; ++k) g->cells[k] = e.cells[k];
    g->trust = M->trust_set ? M->trust.p : nullptr;
    g->use_env = 1;
    g->n_bands = M->n_bands;
    std::memcpy(g->bands, M->bands, sizeof(g->bands));
  } else {
    double u[3], nn = 0.0;
    for (int b = 0; b < 3; ++b) {
      u[b] = F[0 * 3 + b] * axis[0] + F[1 * 3 + b] * axis[1] + F[2 * 3 + b] * axis[2];
      nn += u[b] * u[b];
    }
    nn = std::sqrt(nn);
    for (int b = 0; b < 3; ++b) g->u[b] = nn > 0 ? u[b] / nn : 0.0;
    // any orthonormal pair normal to u
    const int k = std::fabs(g->u[0]) <= std::fabs(g->u[1])
                      ? (std::fabs(g->u[0]) <= std::fabs(g->u[2]) ? 0 : 2)
                      : (std::fabs(g->u[1]) <= std::fabs(g->u[2]) ? 1 : 2);
    double r[3] = {0, 0, 0};
    r[k] = 1.0;
    double e1[3] = {g->u[1] * r[2] - g->u[2] * r[1], g->u[2] * r[0] - g->u[0] * r[2],
                    g->u[0] * r[1] - g->u[1] * r[0]};
    const double n1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    for (int b = 0; b < 3; ++b) g->e1[b] = n1 > 0 ? e1[b] / n1 : 0.0;
    g->e2[0] = g->u[1] * g->e1[2] - g->u[2] * g->e1[1];
    g->e2[1] = g->u[2] * g->e1[0] - g->u[0] * g->e1[2];
    g->e2[2] = g->u[0] * g->e1[1] - g->u[1] * g->e1[0];
    g->cells[0] = g->cells[1] = g->cells[2] = 1;
    g->use_env = 0;
  }
  return ANACOR_OK;
}

static RiskArgs risk_args(const Model* M) {
  RiskArgs r{};
  r.nodes = M->r_nodes.p;
  r.leaf = M->r_leaf.p;
  r.n_trees = M->r_trees;
  r.depth = M->r_depth;
  r.n_int = (1 << M->r_depth) - 1;
  r.base = M->r_base;
  r.cut = M->r_cut;
  return r;
}

// s1/omega of R reflections to the device and the features kernel over them
// (inv and cell only when requested).
static int run_features(Model* M, const double* s1, const double* omega, int64_t R,
                        const FeatureGeo& g, bool want_inv, bool want_cell) {
  MLP_CHECK(M->s1.grow((size_t)R * 3));
  MLP_CHECK(M->omega.grow((size_t)R));
  MLP_CHECK(M->X.grow((size_t)R * kIn));
  MLP_CHECK(M->flag.grow((size_t)R));
  if (want_inv) MLP_CHECK(M->inv.grow((size_t)R * 4));
  if (want_cell) MLP_CHECK(M->cell.grow((size_t)R));
  MLP_TRY(hipMemcpyAsync(M->s1.p, s1, (size_t)R * 3 * sizeof(double), hipMemcpyHostToDevice,
                         M->stream));
  MLP_TRY(hipMemcpyAsync(M->omega.p, omega, (size_t)R * sizeof(double), hipMemcpyHostToDevice,
                         M->stream));
  hipLaunchKernelGGL(features_kernel, dim3(nblocks(R, 256)), dim3(256), 0, M->stream, M->s1.p,
                     M->omega.p, R, g, M->X.p, want_inv ? M->inv.p : nullptr,
                     want_cell ? M->cell.p : nullptr, M->flag.p);
  if (g.use_env && M->r_trees > 0)   // rows the trees score at or above the cut leave
    hipLaunchKernelGGL(risk_kernel, dim3(nblocks(R, 256)), dim3(256), 0, M->stream, M->X.p, R,
                       risk_args(M), (float*)nullptr, (float*)nullptr, M->flag.p);
  MLP_TRY(hipGetLastError());
  return ANACOR_OK;
}

static double ms_since(std::chrono::steady_clock::time_point a) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
}

}  // namespace mlp
}  // namespace anacor

struct anacor_mlp {
  anacor::mlp::Model* m;
};

using namespace anacor::mlp;

static Model* model_of(anacor_mlp* h) { return h ? h->m : nullptr; }

extern "C" {

size_t anacor_mlp_num_params(void) { return (size_t)kParams; }

anacor_mlp* anacor_mlp_create(uint64_t seed, int device) {
  anacor::clear_error();
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error(ANACOR_ERR_NO_DEVICE, "no HIP device visible");
    return nullptr;
  }
  if (device < 0 || device >= ndev) {
    set_error(ANACOR_ERR_ARG, "device %d of %d", device, ndev);
    return nullptr;
  }
  anacor_mlp* h = new anacor_mlp{new Model()};
  Model* M = h->m;
  M->device = device;
  auto fail = [&](const char* what) -> anacor_mlp* {
    set_error(ANACOR_ERR_HIP, "mlp create: %s failed", what);
    anacor_mlp_destroy(h);
    return nullptr;
  };
  if (hipSetDevice(device) != hipSuccess) return fail("hipSetDevice");
  if (hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking) != hipSuccess)
    return fail("hipStreamCreate");
  if (hipEventCreate(&M->ev[0]) != hipSuccess || hipEventCreate(&M->ev[1]) != hipSuccess)
    return fail("hipEventCreate");
  if (M->P.grow(kParams) || M->G.grow(kParams) || M->m.grow(kParams) || M->v.grow(kParams) ||
      M->Wb.grow(kWbSize) || M->loss.grow(1))
    return fail("hipMalloc");
  // uniform He-style initialisation, zero biases
  std::vector<float> w(kParams, 0.f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int fan_in, int count) {
    const float a = std::sqrt(6.f / (float)fan_in);
    std::uniform_real_distribution<float> u(-a, a);
    for (int i = 0; i < count; ++i) w[(size_t)off + i] = u(rng);
  };
  fill(kOffW1, kIn, kH * kIn);
  fill(kOffW2, kH, kH * kH);
  fill(kOffW3, kH, kH * kH);
  fill(kOffW4, kH, kH * kH);
  fill(kOffW5, kH, kH);
  if (hipMemcpy(M->P.p, w.data(), kParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return fail("hipMemcpy");
  hipLaunchKernelGGL(zero_floats, dim3(nblocks(kParams, 256)), dim3(256), 0, M->stream, M->m.p,
                     kParams);
  hipLaunchKernelGGL(zero_floats, dim3(nblocks(kParams, 256)), dim3(256), 0, M->stream, M->v.p,
                     kParams);
  // the K padding of the W1 hi/lo block stays zero for the model's lifetime
  if (hipMemsetAsync(M->Wb.p, 0, kWbSize * sizeof(__bf16), M->stream) != hipSuccess)
    return fail("hipMemsetAsync");
  refresh_bf16(M);
  if (hipStreamSynchronize(M->stream) != hipSuccess) return fail("initialisation");
  return h;
}

void anacor_mlp_destroy(anacor_mlp* mlp) {
  if (!mlp) return;
  Model* M = mlp->m;
  if (M) {
    (void)hipSetDevice(M->device);
    if (M->stream) (void)hipStreamSynchronize(M->stream);
    M->P.release(); M->G.release(); M->m.release(); M->v.release(); M->Wb.release();
    for (int l = 0; l < 4; ++l) {
      M->Z[l].release();
      M->H[l].release();
    }
    M->dA.release(); M->dB.release(); M->dy.release(); M->xb.release(); M->tb.release();
    M->part.release(); M->loss.release(); M->ws.release(); M->colws.release();
    M->s1.release(); M->omega.release(); M->inv.release(); M->A.release();
    M->X.release(); M->y.release(); M->T.release(); M->flag.release(); M->trust.release();
    M->cell.release();
    M->r_nodes.release(); M->r_leaf.release(); M->r_feat.release(); M->r_score.release();
    for (hipEvent_t e : M->ev)
      if (e) (void)hipEventDestroy(e);
    if (M->stream) (void)hipStreamDestroy(M->stream);
    delete M;
  }
  delete mlp;
}

int anacor_mlp_get_params(anacor_mlp* mlp, float* out, size_t n) {
  Model* M = model_of(mlp);
  if (!M || !out || n < (size_t)kParams) {
    set_error(ANACOR_ERR_ARG, "mlp get_params: bad arguments");
    return ANACOR_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(M->mu);
  MLP_TRY(hipSetDevice(M->device));
  MLP_TRY(hipMemcpyAsync(out, M->P.p, kParams * sizeof(float), hipMemcpyDeviceToHost, M->stream));
  MLP_TRY(hipStreamSynchronize(M->stream));
  return ANACOR_OK;
}

int anacor_mlp_set_params(anacor_mlp* mlp, const float* in, size_t n) {
  Model* M = model_of(mlp);
  if (!M || !in || n < (size_t)kParams) {
    set_error(ANACOR_ERR_ARG, "mlp set_params: bad arguments");
    return ANACOR_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(M->mu);
  MLP_TRY(hipSetDevice(M->device));
  MLP_TRY(hipMemcpyAsync(M->P.p, in, kParams * sizeof(float), hipMemcpyHostToDevice, M->stream));
  refresh_bf16(M);
  hipLaunchKernelGGL(zero_floats, dim3(nblocks(kParams, 256)), dim3(256), 0, M->stream, M->m.p,
                     kParams);
  hipLaunchKernelGGL(zero_floats, dim3(nblocks(kParams, 256)), dim3(256), 0, M->stream, M->v.p,
                     kParams);
  M->adam_t = 0;
  MLP_TRY(hipStreamSynchronize(M->stream));
  return ANACOR_OK;
}

int anacor_mlp_set_scale(anacor_mlp* mlp, double y_mean, double y_std) {
  Model* M = model_of(mlp);
  if (!M || !(y_std > 0.0) || !std::isfinite(y_mean)) {
    set_error(ANACOR_ERR_ARG, "mlp set_scale: bad arguments");
    return ANACOR_ERR_ARG;
  }
  M->y_mean = y_mean;
  M->y_std = y_std;
  return ANACOR_OK;
}

int anacor_mlp_set_features(anacor_mlp* mlp, int feature_set) {
  Model* M = model_of(mlp);
  if (!M || (feature_set != ANACOR_MLP_FEATURES_HARMONIC &&
             feature_set != ANACOR_MLP_FEATURES_CHAMBER &&
             feature_set != ANACOR_MLP_FEATURES_RESONANCE)) {
    set_error(ANACOR_ERR_ARG, "mlp set_features: bad arguments");
    return ANACOR_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(M->mu);
  M->feature_set = feature_set;
  return ANACOR_OK;
}

int anacor_mlp_set_envelope(anacor_mlp* mlp, const anacor_envelope_t* env, const uint8_t* trust) {
  Model* M = model_of(mlp);
  if (!M) {
    set_error(ANACOR_ERR_ARG, "mlp set_envelope: null model");
    return ANACOR_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(M->mu);
  if (!env) {
    M->has_env = false;
    M->trust_set = false;
    return ANACOR_OK;
  }
  int64_t n = 1;
  for (int k = 0; k < 3; ++k) {
    if (env->cells[k] < 1 || env->cells[k] > 4096) {
      set_error(ANACOR_ERR_ARG, "mlp set_envelope: cells[%d] = %d outside 1..4096", k,
                (int)env->cells[k]);
      return ANACOR_ERR_ARG;
    }
    n *= env->cells[k];
  }
  if (n > ((int64_t)1 << 26) || !(env->two_theta[1] >= env->two_theta[0]) ||
      !(env->axial[1] >= env->axial[0])) {
    set_error(ANACOR_ERR_ARG, "mlp set_envelope: bad ranges or too many cells");
    return ANACOR_ERR_ARG;
  }
  MLP_TRY(hipSetDevice(M->device));
  if (trust) {
    MLP_CHECK(M->trust.grow((size_t)n));
    MLP_TRY(hipMemcpyAsync(M->trust.p, trust, (size_t)n, hipMemcpyHostToDevice, M->stream));
    MLP_TRY(hipStreamSynchronize(M->stream));
  }
  M->env = *env;
  M->has_env = true;
  M->trust_set = trust != nullptr;
  return ANACOR_OK;
}

int anacor_mlp_set_bands(anacor_mlp* mlp, const anacor_band_t* bands, int n) {
  Model* M = model_of(mlp);
  if (!M || n < 0 || n > ANACOR_MLP_MAX_BANDS || (n > 0 && !bands)) {
    set_error(ANACOR_ERR_ARG, "mlp set_bands: null model or %d bands (at most %d)", n,
              ANACOR_MLP_MAX_BANDS);
    return ANACOR_ERR_ARG;
  }
  for (int k = 0; k < n; ++k) {
    const anacor_band_t& b = bands[k];
    if ((b.ray != 0 && b.ray != 1) || b.i < 0 || b.i > 2 || b.j < 0 || b.j > 2 || b.i == b.j ||
        !(b.ratio >= 0.0 && b.ratio <= 1.0) || !(b.width >= 0.0)) {
      set_error(ANACOR_ERR_ARG, "mlp set_bands: band %d is malformed", k);
      return ANACOR_ERR_ARG;
    }
  }
  std::lock_guard<std::mutex> lk(M->mu);
  M->n_bands = n;
  std::memset(M->bands, 0, sizeof(M->bands));
  if (n > 0) std::memcpy(M->bands, bands, (size_t)n * sizeof(anacor_band_t));
  return ANACOR_OK;
}

int anacor_mlp_set_risk(anacor_mlp* mlp, int n_trees, int depth, const uint8_t* feat,
                        const float* thr, const float* leaf, float base, float cut) {
  Model* M = model_of(mlp);
  if (!M || n_trees < 0 || n_trees > ANACOR_RISK_MAX_TREES ||
      (n_trees > 0 && (depth < 1 || depth > ANACOR_RISK_MAX_DEPTH || !feat || !thr || !leaf ||
                       !std::isfinite(base) || std::isnan(cut)))) {
    set_error(ANACOR_ERR_ARG, "mlp set_risk: null model or bad trees (%d trees, depth %d)",
              n_trees, depth);
    return ANACOR_ERR_ARG;
  }
  const int n_int = n_trees > 0 ? (1 << depth) - 1 : 0;
  std::vector<RiskNode> nodes((size_t)n_trees * n_int);
  for (size_t k = 0; k < nodes.size(); ++k) {
    if (feat[k] >= kRiskF || std::isnan(thr[k])) {
      set_error(ANACOR_ERR_ARG, "mlp set_risk: node %zu: feature %d (of %d) or NaN threshold",
                k, (int)feat[k], kRiskF);
      return ANACOR_ERR_ARG;
    }
    nodes[k].thr = thr[k];
    nodes[k].feat = feat[k];
  }
  for (size_t k = 0; k < (size_t)n_trees * (n_int + 1); ++k)
    if (!std::isfinite(leaf[k])) {
      set_error(ANACOR_ERR_ARG, "mlp set_risk: leaf %zu is not finite", k);
      return ANACOR_ERR_ARG;
    }
  std::lock_guard<std::mutex> lk(M->mu);
  MLP_TRY(hipSetDevice(M->device));
  if (n_trees > 0) {
    MLP_CHECK(M->r_nodes.grow(nodes.size()));
    MLP_CHECK(M->r_leaf.grow((size_t)n_trees * (n_int + 1)));
    MLP_TRY(hipMemcpyAsync(M->r_nodes.p, nodes.data(), nodes.size() * sizeof(RiskNode),
                           hipMemcpyHostToDevice, M->stream));
    MLP_TRY(hipMemcpyAsync(M->r_leaf.p, leaf, (size_t)n_trees * (n_int + 1) * sizeof(float),
                           hipMemcpyHostToDevice, M->stream));
    MLP_TRY(hipStreamSynchronize(M->stream));
  }
  M->r_trees = n_trees;
  M->r_depth = n_trees > 0 ? depth : 0;
  M->r_base = base;
  M->r_cut = cut;
  return ANACOR_OK;
}

int anacor_mlp_risk(anacor_mlp* mlp, const float* x, int64_t n, float* features,
                    float* scores) {
  Model* M = model_of(mlp);
  if (!M || n < 0 || n > ((int64_t)1 << 30) || (n > 0 && !x)) {
    set_error(ANACOR_ERR_ARG, "mlp risk: bad arguments");
    return ANACOR_ERR_ARG;
  }
  if (n == 0) return ANACOR_OK;
  std::lock_guard<std::mutex> lk(M->mu);
  if (scores && M->r_trees <= 0) {
    set_error(ANACOR_ERR_ARG, "mlp risk: scores asked for but no trees are set");
    return ANACOR_ERR_ARG;
  }
  MLP_TRY(hipSetDevice(M->device));
  MLP_CHECK(M->X.grow((size_t)n * kIn));
  if (features) MLP_CHECK(M->r_feat.grow((size_t)n * kRiskF));
  if (scores) MLP_CHECK(M->r_score.grow((size_t)n));
  MLP_TRY(hipMemcpyAsync(M->X.p, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice,
                         M->stream));
  RiskArgs r = risk_args(M);
  if (!scores) r.n_trees = 0;
  hipLaunchKernelGGL(risk_kernel, dim3(nblocks(n, 256)), dim3(256), 0, M->stream, M->X.p, n, r,
                     features ? M->r_feat.p : nullptr, scores ? M->r_score.p : nullptr,
                     (uint8_t*)nullptr);
  MLP_TRY(hipGetLastError());
  if (features)
    MLP_TRY(hipMemcpyAsync(features, M->r_feat.p, (size_t)n * kRiskF * sizeof(float),
                           hipMemcpyDeviceToHost, M->stream));
  if (scores)
    MLP_TRY(hipMemcpyAsync(scores, M->r_score.p, (size_t)n * sizeof(float),
                           hipMemcpyDeviceToHost, M->stream));
  MLP_TRY(hipStreamSynchronize(M->stream));
  return ANACOR_OK;
}

int anacor_mlp_features(anacor_mlp* mlp, const double* s1, const double* omega, int64_t R,
                        const double raw_xray[3], const double omega_axis[3], const double F[9],
                        float* features, double* invariants, int32_t* cells,
                        uint8_t* in_envelope) {
  Model* M = model_of(mlp);
  if (!M || R < 0 || (R > 0 && (!s1 || !omega))) {
    set_error(ANACOR_ERR_ARG, "mlp features: bad arguments");
    return ANACOR_ERR_ARG;
  }
  if (R == 0) return ANACOR_OK;
  std::lock_guard<std::mutex> lk(M->mu);
  if ((cells || in_envelope) && !M->has_env) {
    set_error(ANACOR_ERR_ARG, "mlp features: cells / in_envelope need an envelope");
    return ANACOR_ERR_ARG;
  }
  FeatureGeo g;
  MLP_CHECK(geo_from(M, raw_xray, omega_axis, F, &g));
  MLP_TRY(hipSetDevice(M->device));
  MLP_CHECK(run_features(M, s1, omega, R, g, invariants != nullptr, cells != nullptr));
  if (features)
    MLP_TRY(hipMemcpyAsync(features, M->X.p, (size_t)R * kIn * sizeof(float),
                           hipMemcpyDeviceToHost, M->stream));
  if (invariants)
    MLP_TRY(hipMemcpyAsync(invariants, M->inv.p, (size_t)R * 4 * sizeof(double),
                           hipMemcpyDeviceToHost, M->stream));
  if (cells)
    MLP_TRY(hipMemcpyAsync(cells, M->cell.p, (size_t)R * sizeof(int32_t),
                           hipMemcpyDeviceToHost, M->stream));
  if (in_envelope)
    MLP_TRY(hipMemcpyAsync(in_envelope, M->flag.p, (size_t)R, hipMemcpyDeviceToHost,
                           M->stream));
  MLP_TRY(hipStreamSynchronize(M->stream));
  return ANACOR_OK;
}

int anacor_mlp_forward(anacor_mlp* mlp, const float* x, int64_t n, float* y) {
  Model* M = model_of(mlp);
  if (!M || n < 0 || (n > 0 && (!x || !y))) {
    set_error(ANACOR_ERR_ARG, "mlp forward: bad arguments");
    return ANACOR_ERR_ARG;
  }
  if (n == 0) return ANACOR_OK;
  std::lock_guard<std::mutex> lk(M->mu);
  MLP_TRY(hipSetDevice(M->device));
  MLP_CHECK(M->X.grow((size_t)n * kIn));
  MLP_CHECK(M->y.grow((size_t)n));
  MLP_TRY(hipMemcpyAsync(M->X.p, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice,
                         M->stream));
  infer(M, M->X.p, n, nullptr, M->y.p, nullptr);
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipMemcpyAsync(y, M->y.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost,
                         M->stream));
  MLP_TRY(hipStreamSynchronize(M->stream));
  return ANACOR_OK;
}

int anacor_mlp_gradients(anacor_mlp* mlp, const float* x, const float* t, int64_t n,
                         float* grad, float* loss) {
  Model* M = model_of(mlp);
  if (!M || n <= 0 || n > kChunk || !x || !t || !grad) {
    set_error(ANACOR_ERR_ARG, "mlp gradients: bad arguments (1 <= n <= %d)", kChunk);
    return ANACOR_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(M->mu);
  MLP_TRY(hipSetDevice(M->device));
  MLP_CHECK(ensure_rows(M, (int)n));
  MLP_CHECK(M->y.grow((size_t)n));
  MLP_TRY(hipMemcpyAsync(M->xb.p, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice,
                         M->stream));
  MLP_TRY(hipMemcpyAsync(M->tb.p, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice,
                         M->stream));
  hipLaunchKernelGGL(zero_floats, dim3(1), dim3(64), 0, M->stream, M->loss.p, 1);
  fwd_bwd(M, M->xb.p, nullptr, M->tb.p, (int)n);
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipMemcpyAsync(grad, M->G.p, kParams * sizeof(float), hipMemcpyDeviceToHost,
                         M->stream));
  float l = 0.f;
  MLP_TRY(hipMemcpyAsync(&l, M->loss.p, sizeof(float), hipMemcpyDeviceToHost, M->stream));
  MLP_TRY(hipStreamSynchronize(M->stream));
  if (loss) *loss = l;
  return ANACOR_OK;
}

int anacor_mlp_train(anacor_mlp* mlp, const float* x, const float* t, int64_t n, int epochs,
                     int batch, float lr, float lr_min, uint64_t seed, float* epoch_loss) {
  Model* M = model_of(mlp);
  if (!M || n <= 0 || n > ((int64_t)1 << 30) || !x || !t || epochs <= 0 || batch <= 0 ||
      batch > kChunk || !(lr > 0.f) || !(lr_min >= 0.f)) {
    set_error(ANACOR_ERR_ARG, "mlp train: bad arguments");
    return ANACOR_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(M->mu);
  MLP_TRY(hipSetDevice(M->device));
  MLP_CHECK(ensure_rows(M, batch));
  MLP_CHECK(M->y.grow((size_t)batch));
  MLP_CHECK(M->X.grow((size_t)n * kIn));
  MLP_CHECK(M->T.grow((size_t)n));
  DevBuf<int> perm_d;
  MLP_CHECK(perm_d.grow((size_t)n));
  int rc = ANACOR_OK;
  auto hip_ok = [&](hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    set_error(ANACOR_ERR_HIP, "mlp train: %s failed: %s", what, hipGetErrorString(e));
    rc = ANACOR_ERR_HIP;
    return false;
  };
  std::vector<int> perm((size_t)n);
  for (int64_t i = 0; i < n; ++i) perm[(size_t)i] = (int)i;
  std::mt19937_64 rng(seed);
  const int64_t per_epoch = (n + batch - 1) / batch;
  const int64_t total = per_epoch * epochs;
  int64_t step = 0;
  if (hip_ok(hipMemcpyAsync(M->X.p, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice,
                            M->stream), "upload") &&
      hip_ok(hipMemcpyAsync(M->T.p, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice,
                            M->stream), "upload")) {
    for (int e = 0; e < epochs; ++e) {
      std::shuffle(perm.begin(), perm.end(), rng);
      if (!hip_ok(hipMemcpyAsync(perm_d.p, perm.data(), (size_t)n * sizeof(int),
                                 hipMemcpyHostToDevice, M->stream), "permutation upload"))
        break;
      hipLaunchKernelGGL(zero_floats, dim3(1), dim3(64), 0, M->stream, M->loss.p, 1);
      for (int64_t s0 = 0; s0 < n; s0 += batch, ++step) {
        const int c = (int)std::min<int64_t>(batch, n - s0);
        fwd_bwd(M, M->X.p, perm_d.p + s0, M->T.p, c);
        const double frac = total > 1 ? (double)step / (double)(total - 1) : 1.0;
        const double cur = lr_min + 0.5 * (lr - lr_min) * (1.0 + std::cos(M_PI * frac));
        adam(M, (float)cur);
      }
      float l = 0.f;
      if (!hip_ok(hipGetLastError(), "kernel launch") ||
          !hip_ok(hipMemcpyAsync(&l, M->loss.p, sizeof(float), hipMemcpyDeviceToHost,
                                 M->stream), "loss download") ||
          !hip_ok(hipStreamSynchronize(M->stream), "epoch"))
        break;
      if (epoch_loss) epoch_loss[e] = l / (float)per_epoch;
    }
  }
  (void)hipStreamSynchronize(M->stream);   // perm staging must outlive the copies
  perm_d.release();
  return rc;
}

int anacor_mlp_predict(anacor_mlp* mlp, const double* s1, const double* omega, int64_t R,
                       const double raw_xray[3], const double omega_axis[3], const double F[9],
                       double* A, uint8_t* in_envelope) {
  Model* M = model_of(mlp);
  if (!M || R < 0 || (R > 0 && (!s1 || !omega || !A))) {
    set_error(ANACOR_ERR_ARG, "mlp predict: bad arguments");
    return ANACOR_ERR_ARG;
  }
  if (R == 0) return ANACOR_OK;
  const auto t0 = std::chrono::steady_clock::now();
  std::lock_guard<std::mutex> lk(M->mu);
  if (!M->has_env) {
    set_error(ANACOR_ERR_ARG, "mlp predict: no envelope set");
    return ANACOR_ERR_ARG;
  }
  FeatureGeo g;
  MLP_CHECK(geo_from(M, raw_xray, omega_axis, F, &g));
  MLP_TRY(hipSetDevice(M->device));
  MLP_CHECK(run_features(M, s1, omega, R, g, false, false));
  MLP_CHECK(M->y.grow((size_t)R));
  MLP_CHECK(M->A.grow((size_t)R));
  MLP_TRY(hipEventRecord(M->ev[0], M->stream));
  infer(M, M->X.p, R, M->flag.p, M->y.p, M->A.p);
  MLP_TRY(hipEventRecord(M->ev[1], M->stream));
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipMemcpyAsync(A, M->A.p, (size_t)R * sizeof(double), hipMemcpyDeviceToHost,
                         M->stream));
  std::vector<uint8_t> fl;
  uint8_t* fdst = in_envelope;
  if (!fdst) {
    fl.resize((size_t)R);
    fdst = fl.data();
  }
  MLP_TRY(hipMemcpyAsync(fdst, M->flag.p, (size_t)R, hipMemcpyDeviceToHost, M->stream));
  MLP_TRY(hipStreamSynchronize(M->stream));
  float ms = 0.f;
  MLP_TRY(hipEventElapsedTime(&ms, M->ev[0], M->ev[1]));
  int64_t in = 0;
  for (int64_t i = 0; i < R; ++i) in += fdst[i];
  M->last_mlp_ms = ms;
  M->last_rows = R;
  M->last_in = in;
  M->last_total_ms = ms_since(t0);
  return ANACOR_OK;
}

int anacor_mlp_stats(const anacor_mlp* mlp, anacor_mlp_stats_t* st) {
  if (!mlp || !mlp->m || !st) {
    set_error(ANACOR_ERR_ARG, "mlp stats: bad arguments");
    return ANACOR_ERR_ARG;
  }
  const Model* M = mlp->m;
  st->mlp_ms = M->last_mlp_ms;
  st->total_ms = M->last_total_ms;
  st->rows = M->last_rows;
  st->in_envelope = M->last_in;
  return ANACOR_OK;
}

}  // extern "C"
