// ocfd_mlp_tile.h -- device code of the fused MLP forward used inside the
// chemistry surrogate kernel: one 256-thread workgroup (4 waves) carries a
// tile of 128 cells through every layer, the activations staying in LDS as
// bfloat16 and each wave streaming its slice of the weights from L2.
//
// Layer shapes this code takes: input padded to 64, hidden widths 256,
// output padded; v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  MFMA
// fragment layouts (16x16x32): lane l holds A[l%16][8(l/16) .. +8] and
// B[8(l/16) .. +8][l%16]; the accumulator holds D[4(l/16) + r][l%16], r < 4.
#pragma once
#include <hip/hip_runtime.h>

typedef __bf16 tbf16;
typedef __bf16 tbf16x8 __attribute__((ext_vector_type(8)));
typedef float tfloatx4 __attribute__((ext_vector_type(4)));

constexpr int MT_ROWS = 128;  // cells per workgroup
constexpr int MT_LDH = 264;   // LDS row pitch (bf16): 256 + 8 against bank conflicts

// GELU (tanh form) as z * sigmoid(2u), the same evaluation as training
__device__ __forceinline__ float mt_gelu(float z) {
  const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
  return z / (1.0f + __expf(-2.0f * u));
}

// H[128][0:K] (LDS) -> act(H . W^T + b) written back into H[128][0:256].
// W: [256][ldw] bf16 row-major (out x in).  Wave w computes output columns
// [64w, 64w + 64) for all 128 rows: 8 x 4 accumulator tiles.
template <bool GELU>
__device__ __forceinline__ void mt_layer256(tbf16* H, const tbf16* __restrict__ W, const float* __restrict__ b, int K,
                                            int ldw, int lane, int w) {
  tfloatx4 acc[8][4];
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = tfloatx4{0.f, 0.f, 0.f, 0.f};
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  tbf16x8 bn[4];
#pragma unroll
  for (int j = 0; j < 4; j++) bn[j] = *(const tbf16x8*)&W[(long)(64 * w + 16 * j + lr) * ldw + lk];
  for (int k0 = 0; k0 < K; k0 += 32) {
    tbf16x8 bc[4];
#pragma unroll
    for (int j = 0; j < 4; j++) bc[j] = bn[j];
    if (k0 + 32 < K) {  // prefetch the next weight fragments
#pragma unroll
      for (int j = 0; j < 4; j++) bn[j] = *(const tbf16x8*)&W[(long)(64 * w + 16 * j + lr) * ldw + k0 + 32 + lk];
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const tbf16x8 a = *(const tbf16x8*)&H[(16 * i + lr) * MT_LDH + k0 + lk];
#pragma unroll
      for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bc[j], acc[i][j], 0, 0, 0);
    }
  }
  __syncthreads();  // every wave has read its input rows
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int col = 64 * w + 16 * j + lr;
    const float bias = b[col];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const float v = acc[i][j][r] + bias;
        H[(16 * i + (lane >> 4) * 4 + r) * MT_LDH + col] = (tbf16)(GELU ? mt_gelu(v) : v);
      }
  }
  __syncthreads();
}

// Output layer, first 16 output columns: O[128][16] (fp32, may alias H) =
// H[128][0:K] . W[0:16]^T + b.  Wave w takes rows [32w, 32w + 32).
__device__ __forceinline__ void mt_out16(const tbf16* H, const tbf16* __restrict__ W, const float* __restrict__ b,
                                         int K, int ldw, float* O, int lane, int w) {
  tfloatx4 acc[2] = {tfloatx4{0.f, 0.f, 0.f, 0.f}, tfloatx4{0.f, 0.f, 0.f, 0.f}};
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  for (int k0 = 0; k0 < K; k0 += 32) {
    const tbf16x8 bf = *(const tbf16x8*)&W[(long)lr * ldw + k0 + lk];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const tbf16x8 a = *(const tbf16x8*)&H[(32 * w + 16 * i + lr) * MT_LDH + k0 + lk];
      acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bf, acc[i], 0, 0, 0);
    }
  }
  __syncthreads();  // O may alias H
  const float bias = b[lr];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int r = 0; r < 4; r++) O[(32 * w + 16 * i + (lane >> 4) * 4 + r) * 16 + lr] = acc[i][r] + bias;
  __syncthreads();
}
