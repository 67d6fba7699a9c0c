// ocfd_mlp_tile.h -- device code of the fused MLP forward used inside the
// chemistry surrogate kernel: one workgroup of NW waves (4 or 8) carries a
// tile of cells (ROWS = 64, 96 or 128) through every layer, the activations
// staying in LDS as bfloat16 and each wave streaming its slice of the
// weights from L2.
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

constexpr int MT_ROWS = 128;  // largest number of cells per workgroup
constexpr int MT_LDH = 264;   // LDS row pitch (bf16): 256 + 8 against bank conflicts

// GELU (tanh form) as z * sigmoid(2u) with the hardware exp and reciprocal,
// the same evaluation as training
__device__ __forceinline__ float mt_gelu(float z) {
  const float w = fmaf(0.044715f * z * z, z, z);
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f((-2.0f * 0.7978845608028654f * 1.4426950408889634f) * w));
}

// H[ROWS][0:K] (LDS) -> act(b + H . W^T) written back into H[ROWS][0:256].
// W: [256][ldw] bf16 row-major (out x in).  NW waves; wave w computes output
// columns [256/NW w, 256/NW (w+1)) for all ROWS rows.  Every output element
// accumulates its K-chunks of 32 in increasing order whatever ROWS and NW
// are, so all tile shapes give bit-identical results.
template <bool GELU, int ROWS, int NW>
__device__ __forceinline__ void mt_layer256(tbf16* H, const tbf16* __restrict__ W, const float* __restrict__ b, int K,
                                            int ldw, int lane, int w) {
  constexpr int IT = ROWS / 16, JT = 16 / NW;
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  const int cb = 16 * JT * w;
  // the accumulators start from the bias of their column
  tfloatx4 acc[IT][JT];
#pragma unroll
  for (int j = 0; j < JT; j++) {
    const float bias = b[cb + 16 * j + lr];
#pragma unroll
    for (int i = 0; i < IT; i++) acc[i][j] = tfloatx4{bias, bias, bias, bias};
  }
  tbf16x8 bn[JT];
#pragma unroll
  for (int j = 0; j < JT; j++) bn[j] = *(const tbf16x8*)&W[(long)(cb + 16 * j + lr) * ldw + lk];
  for (int k0 = 0; k0 < K; k0 += 32) {
    tbf16x8 bc[JT];
#pragma unroll
    for (int j = 0; j < JT; j++) bc[j] = bn[j];
    if (k0 + 32 < K) {  // prefetch the next weight fragments
#pragma unroll
      for (int j = 0; j < JT; j++) bn[j] = *(const tbf16x8*)&W[(long)(cb + 16 * j + lr) * ldw + k0 + 32 + lk];
    }
#pragma unroll
    for (int i = 0; i < IT; i++) {
      const tbf16x8 a = *(const tbf16x8*)&H[(16 * i + lr) * MT_LDH + k0 + lk];
#pragma unroll
      for (int j = 0; j < JT; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bc[j], acc[i][j], 0, 0, 0);
    }
  }
  __syncthreads();  // every wave has read its input rows
#pragma unroll
  for (int j = 0; j < JT; j++) {
    const int col = cb + 16 * j + lr;
#pragma unroll
    for (int i = 0; i < IT; i++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const float v = acc[i][j][r];
        H[(16 * i + (lane >> 4) * 4 + r) * MT_LDH + col] = (tbf16)(GELU ? mt_gelu(v) : v);
      }
  }
  __syncthreads();
}

// Output layer, first 16 output columns: O[ROWS][16] (fp32, may alias H) =
// b + H[ROWS][0:K] . W[0:16]^T.  The 16-row tiles are dealt to the waves in
// turn (wave w takes tiles w, w + NW, ...), so ROWS need only be a multiple
// of 16; each element accumulates its K-chunks in increasing order.
template <int ROWS, int NW>
__device__ __forceinline__ void mt_out16(const tbf16* H, const tbf16* __restrict__ W, const float* __restrict__ b,
                                         int K, int ldw, float* O, int lane, int w) {
  static_assert(ROWS % 16 == 0, "whole 16-row tiles");
  constexpr int NT = ROWS / 16, TPW = (NT + NW - 1) / NW;
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  const float bias = b[lr];
  tfloatx4 acc[TPW];
#pragma unroll
  for (int it = 0; it < TPW; it++) acc[it] = tfloatx4{bias, bias, bias, bias};
  for (int k0 = 0; k0 < K; k0 += 32) {
    const tbf16x8 bf = *(const tbf16x8*)&W[(long)lr * ldw + k0 + lk];
#pragma unroll
    for (int it = 0; it < TPW; it++) {
      const int tt = w + it * NW;
      if (tt < NT) {
        const tbf16x8 a = *(const tbf16x8*)&H[(16 * tt + lr) * MT_LDH + k0 + lk];
        acc[it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bf, acc[it], 0, 0, 0);
      }
    }
  }
  __syncthreads();  // O may alias H
#pragma unroll
  for (int it = 0; it < TPW; it++) {
    const int tt = w + it * NW;
    if (tt < NT) {
#pragma unroll
      for (int r = 0; r < 4; r++) O[(16 * tt + (lane >> 4) * 4 + r) * 16 + lr] = acc[it][r];
    }
  }
  __syncthreads();
}
