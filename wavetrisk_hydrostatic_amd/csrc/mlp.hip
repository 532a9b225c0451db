// MI355X kernels of the learned free-surface correction: a 3-layer MLP
// (192 -> 128 -> 128 -> 64, GELU) on 8x8 node patches, with BF16 weights
// and activations and FP32 accumulation on the matrix cores, its backward
// pass and an Adam optimiser, all on the device.
//
// Every product of the training passes is one tiled kernel, C[M x N] =
// op(A)[M x K] op(B)[K x N]: BF16 operands staged through LDS with k
// contiguous, 64 x 64 output tiles per 256-thread block, each wave a 32 x 32
// quarter of 2 x 2 MFMA tiles of gfx950's v_mfma_f32_16x16x32_bf16 (FP32
// accumulate, one instruction per 32 k).  op() transposes an operand by
// the way it is loaded, so the forward pass (X W), the weight gradients
// (X^T dZ, a reduction over the batch) and the input gradients (dZ W^T) share
// the kernel.  The epilogue adds the bias and applies GELU, writing the FP32
// pre-activation (kept for the backward pass) and the BF16 activation that
// feeds the next layer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "ctx.h"
#include "mlp.h"

namespace {

typedef float float4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

// 16 x 16 x 32 product on the matrix cores.  Lane l holds row l % 16 of A
// and column l % 16 of B with the 8 k values 8 (l / 16) .. +7 (any k order
// shared by A and B gives the same dot products), and rows 4 (l / 16) .. +3
// of column l % 16 of the FP32 accumulator.
__device__ __forceinline__ float4_t mfma32(const uint16_t* a, const uint16_t* b, float4_t c) {
  const bf16x8_t av = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const uint4*>(a));
  const bf16x8_t bv = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const uint4*>(b));
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c, 0, 0, 0);
}
__device__ __forceinline__ float4_t mfma32b(const uint16_t* a, uint4 b, float4_t c) {
  const bf16x8_t av = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const uint4*>(a));
  const bf16x8_t bv = __builtin_bit_cast(bf16x8_t, b);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c, 0, 0, 0);
}

constexpr int GT = 256;            // threads per GEMM block
constexpr int BM = 64, BN = 64;    // output tile
constexpr int BK = 32;             // k per LDS stage
constexpr int LP = BK + 8;         // LDS row pitch (bf16), breaks bank conflicts

// bf16 <-> fp32: round to nearest even (gfx950's v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint2 f2bf4(float a, float b, float c, float d) {
  const bf16x4_t v = {(__bf16)a, (__bf16)b, (__bf16)c, (__bf16)d};
  return __builtin_bit_cast(uint2, v);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// erf by Abramowitz and Stegun, Handbook of Mathematical Functions, 7.1.28:
// erf(x) = 1 - (1 + a1 x + ... + a6 x^6)^-16 for x >= 0, |error| <= 3e-7 in
// exact arithmetic; evaluated in FP32 (the powering amplifies the rounding)
// and its reciprocal by v_rcp_f32 it stays within 2e-6 of erf on [-6, 6], far
// below the 4e-3 relative step of
// the BF16 activations it feeds; odd in x.  Branch-free: the
// GELU of every activation of the forward and backward passes evaluates it,
// and a branchy erf made the inference epilogue the kernel's bottleneck.
__device__ __forceinline__ float erf_as(float x) {
  const float a = fabsf(x);
  float p = 0.0000430638f;
  p = p * a + 0.0002765672f;
  p = p * a + 0.0001520143f;
  p = p * a + 0.0092705272f;
  p = p * a + 0.0422820123f;
  p = p * a + 0.0705230784f;
  p = p * a + 1.0f;
  p = p * p;  // ^2
  p = p * p;  // ^4
  p = p * p;  // ^8
  p = p * p;  // ^16 (inf for large a: the quotient below is then 0)
  return copysignf(1.0f - __builtin_amdgcn_rcpf(p), x);  // v_rcp_f32 (1 ulp): rcp(inf) = 0
}
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erf_as(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  const float cdf = 0.5f * (1.0f + erf_as(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

enum Epi { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2 };

struct GemmArgs {
  const uint16_t* A;  // bf16
  const uint16_t* B;  // bf16
  int M, N, K, lda, ldb;
  float* C;           // fp32 [M][ldc]: the product (EPI_STORE), +bias (EPI_BIAS), pre-activation (EPI_BIAS_GELU)
  int ldc;
  const float* bias;  // [N]
  uint16_t* H;        // bf16 [M][ldc]: GELU activation (EPI_BIAS_GELU)
};

// op(A)(m, k) = TA ? A[k lda + m] : A[m lda + k];  op(B)(k, n) = TB ? B[n ldb + k] : B[k ldb + n]
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(GT) void gemm_bf16_kernel(GemmArgs a) {
  __shared__ __align__(16) uint16_t As[BM][LP];  // As[m][k]
  __shared__ __align__(16) uint16_t Bs[BN][LP];  // Bs[n][k]
  const int t = threadIdx.x, lane = t % 64, wave = t / 64;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;
  float4_t acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < a.K; k0 += BK) {
    // --- stage A: 64 x 32 bf16 = 256 runs of 8 ------------------------------
    {
      uint16_t v[8];
      if (!TA) {  // run of 8 consecutive k for one m
        const int m = t / 4, kk = (t % 4) * 8, gm = m0 + m;
        if (gm < a.M && k0 + kk + 8 <= a.K) {
          const uint4 q = *reinterpret_cast<const uint4*>(a.A + (int64_t)gm * a.lda + k0 + kk);
          std::memcpy(v, &q, 16);
        } else {
          for (int i = 0; i < 8; ++i) v[i] = (gm < a.M && k0 + kk + i < a.K) ? a.A[(int64_t)gm * a.lda + k0 + kk + i] : 0;
        }
        std::memcpy(&As[m][kk], v, 16);  // one 16-byte LDS store (row pitch 80 B, kk a multiple of 8)
      } else {  // run of 8 consecutive m for one k
        const int kk = t / 8, m = (t % 8) * 8, gk = k0 + kk;
        if (gk < a.K && m0 + m + 8 <= a.M) {
          const uint4 q = *reinterpret_cast<const uint4*>(a.A + (int64_t)gk * a.lda + m0 + m);
          std::memcpy(v, &q, 16);
        } else {
          for (int i = 0; i < 8; ++i) v[i] = (gk < a.K && m0 + m + i < a.M) ? a.A[(int64_t)gk * a.lda + m0 + m + i] : 0;
        }
        for (int i = 0; i < 8; ++i) As[m + i][kk] = v[i];
      }
    }
    // --- stage B: 32 x 64 ------------------------------------------------------
    {
      uint16_t v[8];
      if (TB) {  // run of 8 consecutive k for one n
        const int n = t / 4, kk = (t % 4) * 8, gn = n0 + n;
        if (gn < a.N && k0 + kk + 8 <= a.K) {
          const uint4 q = *reinterpret_cast<const uint4*>(a.B + (int64_t)gn * a.ldb + k0 + kk);
          std::memcpy(v, &q, 16);
        } else {
          for (int i = 0; i < 8; ++i) v[i] = (gn < a.N && k0 + kk + i < a.K) ? a.B[(int64_t)gn * a.ldb + k0 + kk + i] : 0;
        }
        std::memcpy(&Bs[n][kk], v, 16);
      } else {  // run of 8 consecutive n for one k
        const int kk = t / 8, n = (t % 8) * 8, gk = k0 + kk;
        if (gk < a.K && n0 + n + 8 <= a.N) {
          const uint4 q = *reinterpret_cast<const uint4*>(a.B + (int64_t)gk * a.ldb + n0 + n);
          std::memcpy(v, &q, 16);
        } else {
          for (int i = 0; i < 8; ++i) v[i] = (gk < a.K && n0 + n + i < a.N) ? a.B[(int64_t)gk * a.ldb + n0 + n + i] : 0;
        }
        for (int i = 0; i < 8; ++i) Bs[n + i][kk] = v[i];
      }
    }
    __syncthreads();
    // --- 2 x 2 tiles of 16 x 16 x 32: one 16-byte LDS read per fragment -----------
    {
      const int kq = 8 * (lane / 16);
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          acc[i][j] = mfma32(&As[wm + i * 16 + lane % 16][kq], &Bs[wn + j * 16 + lane % 16][kq], acc[i][j]);
    }
    __syncthreads();
  }
  // --- epilogue: lane holds rows 4 (lane / 16) + r, column lane % 16 of each tile
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn + j * 16 + lane % 16;
      if (n >= a.N) continue;
      const float b = EPI == EPI_STORE ? 0.f : a.bias[n];
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + 4 * (lane / 16) + r;
        if (m >= a.M) continue;
        const float z = acc[i][j][r] + b;
        a.C[(int64_t)m * a.ldc + n] = z;
        if (EPI == EPI_BIAS_GELU) a.H[(int64_t)m * a.ldc + n] = f2bf(gelu(z));
      }
    }
}

// ---------------------------------------------------------------------------
// Inference in one launch: a persistent block per CU holds the three weight
// matrices transposed in LDS (n rows of k contiguous BF16, 101 KB with
// padding) and the biases for the whole launch, and walks tiles of 128 rows;
// each of its 8 waves takes 16 rows through the three layers, its hidden
// activations in its own LDS rows, and loads the next tile's input rows while
// it finishes the current one.  Each product is computed transposed, Z^T =
// W^T X^T: the weights are the matrix cores' A operand and the 16 rows the
// columns of B, so a lane's accumulator holds 4 consecutive outputs of one row
// -- stored as one 8-byte BF16 activation or one 16-byte FP32 output.  Every
// output is the sum over k of the products of gemm_bf16_kernel's forward pass
// in the same k chunks and order, with the same epilogue arithmetic.
// ---------------------------------------------------------------------------
constexpr int IT = 512;               // threads per inference block (8 waves, 2 per SIMD)
constexpr int IW = 16;                // rows per wave
constexpr int IR = IT / 64 * IW;      // rows per block tile
constexpr int P1 = WTG_MLP_IN + 8;    // LDS pitches (bf16) of W1^T, W2^T / W3^T and the activations
constexpr int P2 = WTG_MLP_HIDDEN + 8;

struct InferSmem {
  uint16_t w1[WTG_MLP_HIDDEN][P1];
  uint16_t w2[WTG_MLP_HIDDEN][P2];
  uint16_t w3[WTG_MLP_OUT][P2];
  uint16_t h[IT / 64][IW][P2];
  float b[2 * WTG_MLP_HIDDEN + WTG_MLP_OUT];
};

// rows (n) of a transposed weight matrix [N][K] into LDS with pitch P, 16 bytes per copy
template <int N, int K, int P>
__device__ __forceinline__ void stage_wt(const uint16_t* __restrict__ wt, uint16_t (*dst)[P]) {
  for (int i = threadIdx.x; i < N * (K / 8); i += IT) {
    const int n = i / (K / 8), c = (i % (K / 8)) * 8;
    *reinterpret_cast<uint4*>(&dst[n][c]) = *reinterpret_cast<const uint4*>(wt + (int64_t)n * K + c);
  }
}

__global__ __launch_bounds__(IT) void mlp_infer_kernel(int B, const uint16_t* __restrict__ x,
                                                       const uint16_t* __restrict__ wt, const float* __restrict__ p,
                                                       float* __restrict__ y) {
  constexpr int NI_ = WTG_MLP_IN, NH_ = WTG_MLP_HIDDEN, NO_ = WTG_MLP_OUT;
  constexpr int OB1_ = NI_ * NH_, OW2_ = OB1_ + NH_, OB2_ = OW2_ + NH_ * NH_, OW3_ = OB2_ + NH_,
                OB3_ = OW3_ + NH_ * NO_;
  constexpr int K1 = NI_ / 32, K2 = NH_ / 32, J1 = NH_ / 16, J3 = NO_ / 16;
  __shared__ __align__(16) InferSmem s;
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int r16 = lane % 16, q8 = 8 * (lane / 16), q4 = 4 * (lane / 16);
  stage_wt<NH_, NI_, P1>(wt, s.w1);
  stage_wt<NH_, NH_, P2>(wt + NI_ * NH_, s.w2);
  stage_wt<NO_, NH_, P2>(wt + NI_ * NH_ + NH_ * NH_, s.w3);
  for (int i = threadIdx.x; i < NH_; i += IT) s.b[i] = p[OB1_ + i], s.b[NH_ + i] = p[OB2_ + i];
  for (int i = threadIdx.x; i < NO_; i += IT) s.b[2 * NH_ + i] = p[OB3_ + i];
  __syncthreads();
  uint16_t(*h)[P2] = s.h[wave];
  // this lane's input row fragments of a tile (rows past B read as zero)
  auto load_x = [&](int m0, uint4 v[K1]) {
    const int row = m0 + wave * IW + r16;
    for (int ks = 0; ks < K1; ++ks)
      v[ks] = row < B ? *reinterpret_cast<const uint4*>(x + (int64_t)row * NI_ + 32 * ks + q8)
                      : make_uint4(0u, 0u, 0u, 0u);
  };
  // bias + GELU of 4 consecutive hidden units n0 .. n0 + 3 of row r16, as BF16 into h
  auto act = [&](const float4_t& z, const float* bias, int n0) {
    const float4 bv = *reinterpret_cast<const float4*>(bias + n0);
    *reinterpret_cast<uint2*>(&h[r16][n0]) =
        f2bf4(gelu(z[0] + bv.x), gelu(z[1] + bv.y), gelu(z[2] + bv.z), gelu(z[3] + bv.w));
  };
  uint4 xa[K1];
  load_x(blockIdx.x * IR, xa);
  for (int m0 = blockIdx.x * IR; m0 < B; m0 += gridDim.x * IR) {
    // layer 1: Z1 = X W1 + b1, H1 = bf16(GELU(Z1))
    float4_t acc[J1];
    for (int j = 0; j < J1; ++j) acc[j] = float4_t{0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < K1; ++ks)
      for (int j = 0; j < J1; ++j) acc[j] = mfma32b(&s.w1[16 * j + r16][32 * ks + q8], xa[ks], acc[j]);
    load_x(m0 + gridDim.x * IR, xa);  // the next tile's rows, in flight through layers 2 and 3
    for (int j = 0; j < J1; ++j) act(acc[j], s.b, 16 * j + q4);
    // layer 2: the wave's activations in place (its reads of H1 precede its writes of H2)
    for (int j = 0; j < J1; ++j) acc[j] = float4_t{0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < K2; ++ks) {
      const uint4 hb = *reinterpret_cast<const uint4*>(&h[r16][32 * ks + q8]);
      for (int j = 0; j < J1; ++j) acc[j] = mfma32b(&s.w2[16 * j + r16][32 * ks + q8], hb, acc[j]);
    }
    for (int j = 0; j < J1; ++j) act(acc[j], s.b + NH_, 16 * j + q4);
    // layer 3: y = H2 W3 + b3 for the valid rows, 4 consecutive outputs per lane
    float4_t ac3[J3];
    for (int j = 0; j < J3; ++j) ac3[j] = float4_t{0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < K2; ++ks) {
      const uint4 hb = *reinterpret_cast<const uint4*>(&h[r16][32 * ks + q8]);
      for (int j = 0; j < J3; ++j) ac3[j] = mfma32b(&s.w3[16 * j + r16][32 * ks + q8], hb, ac3[j]);
    }
    const int row = m0 + wave * IW + r16;
    if (row < B)
      for (int j = 0; j < J3; ++j) {
        const float4 bv = *reinterpret_cast<const float4*>(s.b + 2 * NH_ + 16 * j + q4);
        *reinterpret_cast<float4*>(y + (int64_t)row * NO_ + 16 * j + q4) =
            make_float4(ac3[j][0] + bv.x, ac3[j][1] + bv.y, ac3[j][2] + bv.z, ac3[j][3] + bv.w);
      }
  }
}

// bf16 W^T of a [K][N] FP32 weight matrix: wt[n][k] = bf16(w[k][n])
__global__ void to_bf16_t_kernel(int K, int N, const float* __restrict__ w, uint16_t* __restrict__ wt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K * N) return;
  const int n = i / K, k = i % K;
  wt[i] = f2bf(w[(int64_t)k * N + n]);
}

__global__ void to_bf16_kernel(int64_t n, const float* __restrict__ x, uint16_t* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// rows idx[r] of a float dataset [*, w] -> bf16 [B, w] (rows >= nvalid zero)
__global__ void gather_rows_kernel(int B, int nvalid, int w, const int* __restrict__ idx, const float* __restrict__ src,
                                   float* __restrict__ dst_f, uint16_t* __restrict__ dst_b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * w) return;
  const int r = (int)(i / w), c = (int)(i % w);
  const float v = r < nvalid ? src[(int64_t)idx[r] * w + c] : 0.f;
  if (dst_f) dst_f[i] = v;
  if (dst_b) dst_b[i] = f2bf(v);
}

// loss = sum (y - t)^2 / (2 nvalid); dY = (y - t) / nvalid (bf16), rows >= nvalid zero
__global__ void mse_kernel(int B, int nvalid, int w, const float* __restrict__ y, const float* __restrict__ tgt,
                           uint16_t* __restrict__ dyb, float* __restrict__ partial) {
  __shared__ float red[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < (int64_t)B * w) {
    const int r = (int)(i / w);
    float d = 0.f;
    if (r < nvalid) d = y[i] - tgt[i];
    l = d * d;
    dyb[i] = f2bf(d / (float)nvalid);
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// dZ = dH * gelu'(Z) -> bf16
__global__ void gelu_back_kernel(int64_t n, const float* __restrict__ dh, const float* __restrict__ z,
                                 uint16_t* __restrict__ dzb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dzb[i] = f2bf(dh[i] * gelu_grad(z[i]));
}

// db[c] = sum over rows of dZ (bf16 values, fp32 sum), one block per column
__global__ void colsum_kernel(int B, int w, const uint16_t* __restrict__ dzb, float* __restrict__ db) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  float s = 0.f;
  for (int r = threadIdx.x; r < B; r += blockDim.x) s += bf2f(dzb[(int64_t)r * w + c]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] = red[0];
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters, BF16 copies of the weights refreshed
__global__ void adam_kernel(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                            float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2,
                            const int* __restrict__ wmap, uint16_t* __restrict__ pb, const int* __restrict__ wmapt,
                            uint16_t* __restrict__ pbt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi, b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float v = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  p[i] = v;
  if (wmap[i] >= 0) {
    pb[wmap[i]] = f2bf(v);
    pbt[wmapt[i]] = f2bf(v);
  }
}

inline unsigned blocks(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct wtg_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  int max_batch = 0;  // rows of the activation buffers (multiple of 64)
  std::string err;
  // FP32 master parameters, flattened W1 b1 W2 b2 W3 b3 (W row-major [in][out])
  float* p = nullptr;
  float* g = nullptr;
  float *m1 = nullptr, *m2 = nullptr;
  uint16_t* wb = nullptr;  // bf16 W1 W2 W3
  uint16_t* wbt = nullptr; // bf16 W1^T W2^T W3^T (inference)
  int infer_grid = 0;      // persistent inference blocks (one per CU)
  int* wmap = nullptr;     // parameter -> index in wb (-1 for biases)
  int* wmapt = nullptr;    // parameter -> index in wbt
  int adam_t = 0;
  // activations
  uint16_t *x0 = nullptr, *h1 = nullptr, *h2 = nullptr, *dyb = nullptr, *dz2 = nullptr, *dz1 = nullptr;
  float *x0f = nullptr, *z1 = nullptr, *z2 = nullptr, *y = nullptr, *tgt = nullptr, *dh = nullptr, *part = nullptr;
  int* idx = nullptr;
  // resident training set
  float *ds_x = nullptr, *ds_y = nullptr;
  int64_t ds_n = 0;
  std::vector<void*> allocs;
};

namespace {

constexpr int NI = WTG_MLP_IN, NH = WTG_MLP_HIDDEN, NO = WTG_MLP_OUT;
constexpr int OW1 = 0, OB1 = OW1 + NI * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NH * NO,
              NPAR = OB3 + NO;
constexpr int BW1 = 0, BW2 = NI * NH, BW3 = BW2 + NH * NH, NWB = BW3 + NH * NO;

int mfail(wtg_mlp* m, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (m)
    m->err = buf;
  else
    g_create_error = buf;
  return code;
}

#define MCHK(m, call)                                                                                 \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess) return mfail(m, WTG_E_DEVICE, "%s: %s", #call, hipGetErrorString(e_));      \
  } while (0)
#define MRC(x)                     \
  do {                             \
    if (int rc_ = (x)) return rc_; \
  } while (0)

template <class Tp>
int malloc_dev(wtg_mlp* m, Tp** p, size_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(count * sizeof(Tp), 1)) != hipSuccess)
    return mfail(m, WTG_E_DEVICE, "hipMalloc of %zu bytes failed", count * sizeof(Tp));
  m->allocs.push_back(q);
  *p = (Tp*)q;
  return 0;
}

template <bool TA, bool TB, int EPI>
int gemm(wtg_mlp* m, const uint16_t* A, int lda, const uint16_t* B, int ldb, int M, int N, int K, float* C, int ldc,
         const float* bias = nullptr, uint16_t* H = nullptr) {
  if (N % 16 || K % 8) return mfail(m, WTG_E_INTERNAL, "gemm shape %d x %d x %d", M, N, K);
  GemmArgs a{A, B, M, N, K, lda, ldb, C, ldc, bias, H};
  dim3 grid((unsigned)((M + BM - 1) / BM), (unsigned)((N + BN - 1) / BN));
  hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB, EPI>), grid, dim3(GT), 0, m->stream, a);
  MCHK(m, hipGetLastError());
  return 0;
}

// forward of B rows already in x0 (bf16): z1, h1, z2, h2, y
int forward(wtg_mlp* m, int B) {
  MRC((gemm<false, false, EPI_BIAS_GELU>(m, m->x0, NI, m->wb + BW1, NH, B, NH, NI, m->z1, NH, m->p + OB1, m->h1)));
  MRC((gemm<false, false, EPI_BIAS_GELU>(m, m->h1, NH, m->wb + BW2, NH, B, NH, NH, m->z2, NH, m->p + OB2, m->h2)));
  MRC((gemm<false, false, EPI_BIAS>(m, m->h2, NH, m->wb + BW3, NO, B, NO, NH, m->y, NO, m->p + OB3)));
  return 0;
}

// loss and gradients (into m->g) of rows in x0 / tgt, nvalid of them; B padded
int backward(wtg_mlp* m, int B, int nvalid, float* loss_out) {
  const unsigned nb = blocks((int64_t)B * NO);
  hipLaunchKernelGGL(mse_kernel, dim3(nb), dim3(256), 0, m->stream, B, nvalid, NO, m->y, m->tgt, m->dyb, m->part);
  MCHK(m, hipGetLastError());
  // layer 3: dW3 = h2^T dY, db3 = colsum dY, dh2 = dY W3^T
  MRC((gemm<true, false, EPI_STORE>(m, m->h2, NH, m->dyb, NO, NH, NO, B, m->g + OW3, NO)));
  hipLaunchKernelGGL(colsum_kernel, dim3(NO), dim3(256), 0, m->stream, B, NO, m->dyb, m->g + OB3);
  MRC((gemm<false, true, EPI_STORE>(m, m->dyb, NO, m->wb + BW3, NO, B, NH, NO, m->dh, NH)));
  // layer 2
  hipLaunchKernelGGL(gelu_back_kernel, dim3(blocks((int64_t)B * NH)), dim3(256), 0, m->stream, (int64_t)B * NH, m->dh,
                     m->z2, m->dz2);
  MRC((gemm<true, false, EPI_STORE>(m, m->h1, NH, m->dz2, NH, NH, NH, B, m->g + OW2, NH)));
  hipLaunchKernelGGL(colsum_kernel, dim3(NH), dim3(256), 0, m->stream, B, NH, m->dz2, m->g + OB2);
  MRC((gemm<false, true, EPI_STORE>(m, m->dz2, NH, m->wb + BW2, NH, B, NH, NH, m->dh, NH)));
  // layer 1
  hipLaunchKernelGGL(gelu_back_kernel, dim3(blocks((int64_t)B * NH)), dim3(256), 0, m->stream, (int64_t)B * NH, m->dh,
                     m->z1, m->dz1);
  MRC((gemm<true, false, EPI_STORE>(m, m->x0, NI, m->dz1, NH, NI, NH, B, m->g + OW1, NH)));
  hipLaunchKernelGGL(colsum_kernel, dim3(NH), dim3(256), 0, m->stream, B, NH, m->dz1, m->g + OB1);
  MCHK(m, hipGetLastError());
  if (loss_out) {
    std::vector<float> part(nb);
    MCHK(m, hipMemcpyAsync(part.data(), m->part, nb * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    double s = 0.0;
    for (float v : part) s += v;
    *loss_out = (float)(s / (2.0 * nvalid));
  }
  return 0;
}

int refresh_bf16(wtg_mlp* m) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(NI * NH)), dim3(256), 0, m->stream, (int64_t)NI * NH, m->p + OW1,
                     m->wb + BW1);
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(NH * NH)), dim3(256), 0, m->stream, (int64_t)NH * NH, m->p + OW2,
                     m->wb + BW2);
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(NH * NO)), dim3(256), 0, m->stream, (int64_t)NH * NO, m->p + OW3,
                     m->wb + BW3);
  hipLaunchKernelGGL(to_bf16_t_kernel, dim3(blocks(NI * NH)), dim3(256), 0, m->stream, NI, NH, m->p + OW1,
                     m->wbt + BW1);
  hipLaunchKernelGGL(to_bf16_t_kernel, dim3(blocks(NH * NH)), dim3(256), 0, m->stream, NH, NH, m->p + OW2,
                     m->wbt + BW2);
  hipLaunchKernelGGL(to_bf16_t_kernel, dim3(blocks(NH * NO)), dim3(256), 0, m->stream, NH, NO, m->p + OW3,
                     m->wbt + BW3);
  MCHK(m, hipGetLastError());
  return 0;
}

int upload_rows(wtg_mlp* m, const float* x, const float* y, int B) {
  // host rows -> x0f/x0 and tgt, padded with zero rows to a multiple of 64
  const int Bp = (B + 63) / 64 * 64;
  if (Bp > m->max_batch) return mfail(m, WTG_E_ARG, "batch %d exceeds the maximum %d", B, m->max_batch);
  MCHK(m, hipMemsetAsync(m->x0f, 0, (size_t)Bp * NI * sizeof(float), m->stream));
  MCHK(m, hipMemcpyAsync(m->x0f, x, (size_t)B * NI * sizeof(float), hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks((int64_t)Bp * NI)), dim3(256), 0, m->stream, (int64_t)Bp * NI, m->x0f,
                     m->x0);
  MCHK(m, hipGetLastError());
  if (y) {
    MCHK(m, hipMemsetAsync(m->tgt, 0, (size_t)Bp * NO * sizeof(float), m->stream));
    MCHK(m, hipMemcpyAsync(m->tgt, y, (size_t)B * NO * sizeof(float), hipMemcpyHostToDevice, m->stream));
  }
  return 0;
}

}  // namespace

// for the solver (elliptic.hip): input / output buffers and a forward pass
// of B rows on the caller's stream
void wtg_mlp_buffers(wtg_mlp* m, uint16_t** x0, float** y, int* max_batch) {
  *x0 = m->x0;
  *y = m->y;
  *max_batch = m->max_batch;
}

int wtg_mlp_infer_on(wtg_mlp* m, hipStream_t stream, int B) {
  if (B < 0 || B > m->max_batch) return mfail(m, WTG_E_ARG, "batch %d exceeds the maximum %d", B, m->max_batch);
  if (B == 0) return 0;
  const unsigned grid = (unsigned)std::min<int64_t>((B + IR - 1) / IR, m->infer_grid);
  hipLaunchKernelGGL(mlp_infer_kernel, dim3(grid), dim3(IT), 0, stream, B, m->x0, m->wbt, m->p, m->y);
  MCHK(m, hipGetLastError());
  return 0;
}


const char* wtg_mlp_error_text(const wtg_mlp* m) { return m->err.c_str(); }

extern "C" {

int wtg_mlp_create(int max_batch, uint64_t seed, wtg_mlp** out) {
  if (!out) return mfail(nullptr, WTG_E_ARG, "null output");
  *out = nullptr;
  if (max_batch < 1 || max_batch > (1 << 22)) return mfail(nullptr, WTG_E_ARG, "max_batch %d out of range", max_batch);
  auto m = std::make_unique<wtg_mlp>();
  if (hipGetDevice(&m->device) != hipSuccess) return mfail(nullptr, WTG_E_DEVICE, "no HIP device");
  if (hipStreamCreate(&m->stream) != hipSuccess) return mfail(nullptr, WTG_E_DEVICE, "hipStreamCreate failed");
  m->max_batch = (max_batch + 63) / 64 * 64;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, m->device) != hipSuccess) return mfail(nullptr, WTG_E_DEVICE, "no device properties");
    m->infer_grid = std::max(1, prop.multiProcessorCount);
  }
  const int Bp = m->max_batch;
  auto mk = [&](auto** p, size_t n) { return malloc_dev(m.get(), p, n); };
  int rc = 0;
  rc = rc ? rc : mk(&m->p, NPAR);
  rc = rc ? rc : mk(&m->g, NPAR);
  rc = rc ? rc : mk(&m->m1, NPAR);
  rc = rc ? rc : mk(&m->m2, NPAR);
  rc = rc ? rc : mk(&m->wb, NWB);
  rc = rc ? rc : mk(&m->wbt, NWB);
  rc = rc ? rc : mk(&m->wmap, NPAR);
  rc = rc ? rc : mk(&m->wmapt, NPAR);
  rc = rc ? rc : mk(&m->x0, (size_t)Bp * NI);
  rc = rc ? rc : mk(&m->x0f, (size_t)Bp * NI);
  rc = rc ? rc : mk(&m->h1, (size_t)Bp * NH);
  rc = rc ? rc : mk(&m->h2, (size_t)Bp * NH);
  rc = rc ? rc : mk(&m->z1, (size_t)Bp * NH);
  rc = rc ? rc : mk(&m->z2, (size_t)Bp * NH);
  rc = rc ? rc : mk(&m->dh, (size_t)Bp * NH);
  rc = rc ? rc : mk(&m->dz1, (size_t)Bp * NH);
  rc = rc ? rc : mk(&m->dz2, (size_t)Bp * NH);
  rc = rc ? rc : mk(&m->y, (size_t)Bp * NO);
  rc = rc ? rc : mk(&m->tgt, (size_t)Bp * NO);
  rc = rc ? rc : mk(&m->dyb, (size_t)Bp * NO);
  rc = rc ? rc : mk(&m->part, (size_t)Bp * NO / 256 + 1);
  rc = rc ? rc : mk(&m->idx, (size_t)Bp);
  if (rc) {
    g_create_error = m->err;
    wtg_mlp_destroy(m.release());
    return rc;
  }
  // He-normal weights (fan in), the output layer scaled by 0.1 so that an
  // untrained network proposes small corrections; zero biases
  std::vector<float> p(NPAR, 0.f);
  std::mt19937_64 gen(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  auto fill = [&](int off, int fin, int fout, double scale) {
    const double s = scale * std::sqrt(2.0 / fin);
    for (int i = 0; i < fin * fout; ++i) p[off + i] = (float)(s * nd(gen));
  };
  fill(OW1, NI, NH, 1.0);
  fill(OW2, NH, NH, 1.0);
  fill(OW3, NH, NO, 0.1);
  std::vector<int> wmap(NPAR, -1), wmapt(NPAR, -1);
  for (int i = 0; i < NI * NH; ++i) wmap[OW1 + i] = BW1 + i, wmapt[OW1 + i] = BW1 + (i % NH) * NI + i / NH;
  for (int i = 0; i < NH * NH; ++i) wmap[OW2 + i] = BW2 + i, wmapt[OW2 + i] = BW2 + (i % NH) * NH + i / NH;
  for (int i = 0; i < NH * NO; ++i) wmap[OW3 + i] = BW3 + i, wmapt[OW3 + i] = BW3 + (i % NO) * NH + i / NO;
  wtg_mlp* mm = m.get();
  if (hipMemcpy(mm->p, p.data(), NPAR * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(mm->wmap, wmap.data(), NPAR * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(mm->wmapt, wmapt.data(), NPAR * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(mm->m1, 0, NPAR * sizeof(float)) != hipSuccess ||
      hipMemset(mm->m2, 0, NPAR * sizeof(float)) != hipSuccess) {
    wtg_mlp_destroy(m.release());
    return mfail(nullptr, WTG_E_DEVICE, "parameter upload failed");
  }
  if (int r = refresh_bf16(mm)) {
    g_create_error = mm->err;
    wtg_mlp_destroy(m.release());
    return r;
  }
  if (hipStreamSynchronize(mm->stream) != hipSuccess) {
    wtg_mlp_destroy(m.release());
    return mfail(nullptr, WTG_E_DEVICE, "stream synchronisation failed");
  }
  *out = m.release();
  return 0;
}

void wtg_mlp_destroy(wtg_mlp* m) {
  if (!m) return;
  if (m->stream) hipStreamSynchronize(m->stream);
  for (void* q : m->allocs) hipFree(q);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
}

int wtg_mlp_last_error(const wtg_mlp* m, char* buf, size_t n) {
  if (!buf || !n) return WTG_E_ARG;
  const std::string& e = m ? m->err : g_create_error;
  std::snprintf(buf, n, "%s", e.c_str());
  return 0;
}

int wtg_mlp_get_params(wtg_mlp* m, float* host, size_t n) {
  if (!m || !host) return WTG_E_ARG;
  if (n != (size_t)NPAR) return mfail(m, WTG_E_ARG, "the MLP has %d parameters, %zu given", NPAR, n);
  MCHK(m, hipMemcpyAsync(host, m->p, NPAR * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return 0;
}

int wtg_mlp_set_params(wtg_mlp* m, const float* host, size_t n) {
  if (!m || !host) return WTG_E_ARG;
  if (n != (size_t)NPAR) return mfail(m, WTG_E_ARG, "the MLP has %d parameters, %zu given", NPAR, n);
  MCHK(m, hipMemcpyAsync(m->p, host, NPAR * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipMemsetAsync(m->m1, 0, NPAR * sizeof(float), m->stream));
  MCHK(m, hipMemsetAsync(m->m2, 0, NPAR * sizeof(float), m->stream));
  m->adam_t = 0;
  MRC(refresh_bf16(m));
  MCHK(m, hipStreamSynchronize(m->stream));
  return 0;
}

int wtg_mlp_forward(wtg_mlp* m, const float* x, int batch, float* y) {
  if (!m || !x || !y) return WTG_E_ARG;
  if (batch < 1) return mfail(m, WTG_E_ARG, "batch < 1");
  MRC(upload_rows(m, x, nullptr, batch));
  MRC(wtg_mlp_infer_on(m, m->stream, batch));
  MCHK(m, hipMemcpyAsync(y, m->y, (size_t)batch * NO * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return 0;
}

int wtg_mlp_forward_timed(wtg_mlp* m, int batch, int nrep, double* ms_per_pass) {
  if (!m || !ms_per_pass) return WTG_E_ARG;
  if (batch < 1 || nrep < 1) return mfail(m, WTG_E_ARG, "batch and nrep must be positive");
  const int Bp = (batch + 63) / 64 * 64;
  if (Bp > m->max_batch) return mfail(m, WTG_E_ARG, "batch %d exceeds the maximum %d", batch, m->max_batch);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return mfail(m, WTG_E_DEVICE, "hipEventCreate failed");
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return mfail(m, WTG_E_DEVICE, "hipEventCreate failed");
  }
  int rc = 0;
  if (hipEventRecord(e0, m->stream) != hipSuccess) rc = mfail(m, WTG_E_DEVICE, "hipEventRecord failed");
  for (int i = 0; i < nrep && !rc; ++i) rc = wtg_mlp_infer_on(m, m->stream, batch);
  if (!rc && (hipEventRecord(e1, m->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
    rc = mfail(m, WTG_E_DEVICE, "event synchronisation failed");
  if (!rc) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    *ms_per_pass = ms / nrep;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return rc;
}

int wtg_mlp_gradients(wtg_mlp* m, const float* x, const float* y_true, int batch, float* loss, float* grads) {
  if (!m || !x || !y_true || !loss || !grads) return WTG_E_ARG;
  if (batch < 1) return mfail(m, WTG_E_ARG, "batch < 1");
  MRC(upload_rows(m, x, y_true, batch));
  const int Bp = (batch + 63) / 64 * 64;
  MRC(forward(m, Bp));
  MRC(backward(m, Bp, batch, loss));
  MCHK(m, hipMemcpyAsync(grads, m->g, NPAR * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return 0;
}

int wtg_mlp_set_dataset(wtg_mlp* m, const float* x, const float* y, int64_t n) {
  if (!m || !x || !y) return WTG_E_ARG;
  if (n < 1) return mfail(m, WTG_E_ARG, "empty dataset");
  if (m->ds_x) {
    MCHK(m, hipStreamSynchronize(m->stream));
    hipFree(m->ds_x);
    hipFree(m->ds_y);
    m->ds_x = m->ds_y = nullptr;
  }
  if (hipMalloc((void**)&m->ds_x, (size_t)n * NI * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&m->ds_y, (size_t)n * NO * sizeof(float)) != hipSuccess)
    return mfail(m, WTG_E_DEVICE, "no memory for %lld training pairs", (long long)n);
  MCHK(m, hipMemcpyAsync(m->ds_x, x, (size_t)n * NI * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipMemcpyAsync(m->ds_y, y, (size_t)n * NO * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  m->ds_n = n;
  return 0;
}

int wtg_mlp_train(wtg_mlp* m, int nsteps, int batch, float lr, uint64_t seed, float* losses) {
  if (!m) return WTG_E_ARG;
  if (!m->ds_x) return mfail(m, WTG_E_STATE, "wtg_mlp_set_dataset must be called first");
  if (nsteps < 0 || batch < 1) return mfail(m, WTG_E_ARG, "bad step count or batch");
  const int Bp = (batch + 63) / 64 * 64;
  if (Bp > m->max_batch) return mfail(m, WTG_E_ARG, "batch %d exceeds the maximum %d", batch, m->max_batch);
  std::mt19937_64 gen(seed);
  std::vector<int> perm(m->ds_n);
  for (int64_t i = 0; i < m->ds_n; ++i) perm[i] = (int)i;
  int64_t pos = m->ds_n;
  std::vector<int> idx(Bp, 0);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int s = 0; s < nsteps; ++s) {
    for (int r = 0; r < batch; ++r) {  // epochs of a shuffled permutation
      if (pos >= m->ds_n) {
        std::shuffle(perm.begin(), perm.end(), gen);
        pos = 0;
      }
      idx[r] = perm[pos++];
    }
    MCHK(m, hipStreamSynchronize(m->stream));  // idx is reused
    MCHK(m, hipMemcpyAsync(m->idx, idx.data(), Bp * sizeof(int), hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(gather_rows_kernel, dim3(blocks((int64_t)Bp * NI)), dim3(256), 0, m->stream, Bp, batch, NI,
                       m->idx, m->ds_x, (float*)nullptr, m->x0);
    hipLaunchKernelGGL(gather_rows_kernel, dim3(blocks((int64_t)Bp * NO)), dim3(256), 0, m->stream, Bp, batch, NO,
                       m->idx, m->ds_y, m->tgt, (uint16_t*)nullptr);
    MCHK(m, hipGetLastError());
    MRC(forward(m, Bp));
    float loss = 0.f;
    MRC(backward(m, Bp, batch, losses ? &loss : nullptr));
    if (losses) losses[s] = loss;
    ++m->adam_t;
    const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks(NPAR)), dim3(256), 0, m->stream, NPAR, m->p, m->g, m->m1, m->m2, lr,
                       b1, b2, eps, c1, c2, m->wmap, m->wb, m->wmapt, m->wbt);
    MCHK(m, hipGetLastError());
  }
  MCHK(m, hipStreamSynchronize(m->stream));
  return 0;
}

}  // extern "C"
