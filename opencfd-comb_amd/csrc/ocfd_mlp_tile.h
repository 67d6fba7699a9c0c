// ocfd_mlp_tile.h -- device code of the MLP forward in the chemistry
// surrogate's network kernel: one workgroup of NW waves carries a tile of
// cells (ROWS = 64, 96 or 128) through every layer, the activations staying
// in LDS as bfloat16 and each wave streaming its slice of the weights, kept
// in MFMA fragment order (mt_pack_src), from L2.
//
// Layer shapes this code takes: input of at most 32 features (zero-padded),
// hidden widths 256, at most 16 outputs; v_mfma_f32_16x16x32_bf16 with FP32
// accumulation.  MFMA fragment layouts (16x16x32): lane l holds
// A[l%16][8(l/16) .. +8] and B[8(l/16) .. +8][l%16]; the accumulator holds
// D[4(l/16) + r][l%16], r < 4.
#pragma once
#include <hip/hip_runtime.h>

typedef __bf16 tbf16;
typedef __bf16 tbf16x8 __attribute__((ext_vector_type(8)));
typedef float tfloatx4 __attribute__((ext_vector_type(4)));

constexpr int MT_LDH = 264;   // LDS row pitch (bf16): 256 + 8 against bank conflicts

// GELU in its tanh form, z * sigmoid(2u), u = c (z + 0.044715 z^3), with
// the hardware exp2 and reciprocal; 2u log2(e) is formed as
// z (C0 + C1 z^2), the same evaluation as training
constexpr float MT_G0 = -2.0f * 0.7978845608028654f * 1.4426950408889634f;
constexpr float MT_G1 = MT_G0 * 0.044715f;
__device__ __forceinline__ float mt_gelu(float z) {
  const float t = z * fmaf(z * z, MT_G1, MT_G0);
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(t));
}

typedef __bf16 tbf16x4 __attribute__((ext_vector_type(4)));

// Fragment order of a weight matrix W[out][ldw] (bf16, row-major) for the
// MFMA A operand over K = 32 KS: element (block jb of 16 outputs, K-chunk
// s, lane l, q < 8) = W[16 jb + l % 16][32 s + 8 (l / 16) + q] sits at
// ((jb KS + s) 64 + l) 8 + q, so the 64 lanes' 16-byte fragment loads of
// one (jb, s) cover 1 KB contiguously.
__host__ __device__ inline long mt_pack_src(long e, int KS, int ldw) {
  const long jb = e / (KS * 512);
  const int rem = (int)(e % (KS * 512)), s = rem / 512, l = (rem % 512) / 8, q = rem % 8;
  return (16 * jb + l % 16) * ldw + 32 * s + 8 * (l / 16) + q;
}

// H[ROWS][0:K] (LDS) -> act(b + H . W^T) written back into H[ROWS][0:256].
// W: [256][ldw] bf16 row-major (out x in).  NW waves; wave w computes output
// columns [256/NW w, 256/NW (w+1)) for all ROWS rows.  The MFMA takes the
// weights as its A operand and the activations as B, so a lane's
// accumulator holds 4 consecutive output columns of one cell, converted
// and stored to LDS as one 8-byte write.  Every output element accumulates
// its K-chunks of 32 in increasing order whatever ROWS and NW are, so all
// tile shapes give bit-identical results.  K (32 or 256) is a template
// argument: the K loop unrolls and the weight fragments are loaded one or
// two chunks ahead (from L2, whose latency exceeds one chunk of MFMA work)
// into a rotating set of registers.  Wp holds the weights in fragment order
// (mt_pack_index): a wave's fragment load is 1 KB contiguous.
template <bool GELU, int ROWS, int NW, int K, int NWR = 1>
__device__ __forceinline__ void mt_layer256(tbf16* H, const tbf16* __restrict__ Wp, const float* __restrict__ b,
                                            int lane, int w) {
  constexpr int NWC = NW / NWR, IT = ROWS / NWR / 16, JT = 16 / NWC, KS = K / 32;
  static_assert(K % 32 == 0 && KS >= 1, "K in chunks of 32");
  static_assert(NW % NWR == 0 && ROWS % (16 * NWR) == 0, "whole row groups");
  const int lr = lane & 15, lk = (lane >> 4) * 8, lq = (lane >> 4) * 4;
  const int cb = 16 * JT * (w % NWC), r0 = (ROWS / NWR) * (w / NWC);
  // the accumulators start from the bias of their columns (the first
  // chunk's MFMA takes it as its C operand)
  tfloatx4 acc[IT][JT], bias[JT];
#pragma unroll
  for (int j = 0; j < JT; j++) bias[j] = *(const tfloatx4*)&b[cb + 16 * j + lq];
  // weight prefetch distance in K-chunks: 2, or 1 where the accumulators
  // of 128 x 64 outputs leave no room
  constexpr int PD = IT * JT >= 32 ? 1 : 2;
  const tbf16* Wl = Wp + (long)(cb / 16) * KS * 512 + lane * 8;
  // MFMA phase at raised wave priority, epilogue at the normal one, so that a
  // wave of the other workgroup in its epilogue fills the issue slots
  // between this wave's MFMAs
  __builtin_amdgcn_s_setprio(1);
  tbf16x8 wf[PD + 1][JT];
#pragma unroll
  for (int s = 0; s < PD && s < KS; s++)
#pragma unroll
    for (int j = 0; j < JT; j++) wf[s][j] = *(const tbf16x8*)&Wl[(j * KS + s) * 512];
#pragma unroll
  for (int s = 0; s < KS; s++) {
    if (s + PD < KS) {
#pragma unroll
      for (int j = 0; j < JT; j++)
        wf[(s + PD) % (PD + 1)][j] = *(const tbf16x8*)&Wl[(j * KS + s + PD) * 512];
    }
    tbf16x8 a[IT];
#pragma unroll
    for (int i = 0; i < IT; i++) a[i] = *(const tbf16x8*)&H[(r0 + 16 * i + lr) * MT_LDH + 32 * s + lk];
#pragma unroll
    for (int i = 0; i < IT; i++)
#pragma unroll
      for (int j = 0; j < JT; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[s % (PD + 1)][j], a[i], s == 0 ? bias[j] : acc[i][j],
                                                            0, 0, 0);
  }
  __builtin_amdgcn_s_setprio(0);
  __syncthreads();  // every wave has read its input rows
#pragma unroll
  for (int i = 0; i < IT; i++)
#pragma unroll
    for (int j = 0; j < JT; j++) {
      tfloatx4 v = acc[i][j];
      if (GELU) {
#pragma unroll
        for (int r = 0; r < 4; r++) v[r] = mt_gelu(v[r]);
      }
      *(tbf16x4*)&H[(r0 + 16 * i + lr) * MT_LDH + cb + 16 * j + lq] = __builtin_convertvector(v, tbf16x4);
    }
  __syncthreads();
}

// Output layer: Oout[col][n] = b + H[ROWS][0:K] . W[col]^T for the first
// NOUT (<= 16) columns, written from the accumulators (weights as the MFMA A
// operand, so a lane holds 4 consecutive columns of one cell; 16 cells per
// store).  The 16-row tiles are dealt to the waves in turn (wave w takes
// tiles w, w + NW, ...), so ROWS need only be a multiple of 16; each element
// accumulates its K-chunks in increasing order.  w must be wave-uniform.
template <int ROWS, int NW, int K, int NOUT>
__device__ __forceinline__ void mt_out_cols(const tbf16* H, const tbf16* __restrict__ Wp,
                                            const float* __restrict__ b, float* __restrict__ Oout, long n, long c0,
                                            int lane, int w) {
  static_assert(ROWS % 16 == 0 && K % 32 == 0 && NOUT <= 16, "whole tiles");
  constexpr int NT = ROWS / 16, TPW = (NT + NW - 1) / NW, KS = K / 32;
  const int lr = lane & 15, lk = (lane >> 4) * 8, lq = (lane >> 4) * 4;
  const tfloatx4 bias = *(const tfloatx4*)&b[lq];
  tbf16x8 wf[KS];
#pragma unroll
  for (int s = 0; s < KS; s++) wf[s] = *(const tbf16x8*)&Wp[s * 512 + lane * 8];
#pragma unroll
  for (int it = 0; it < TPW; it++) {
    const int tt = w + it * NW;
    if (NT % NW == 0 || tt < NT) {
      tfloatx4 acc = bias;
#pragma unroll
      for (int s = 0; s < KS; s++)
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[s], *(const tbf16x8*)&H[(16 * tt + lr) * MT_LDH + 32 * s + lk],
                                                      acc, 0, 0, 0);
      const long c = c0 + 16 * tt + lr;
      if (c < n) {
#pragma unroll
        for (int r = 0; r < 4; r++)
          if (lq + r < NOUT) Oout[(long)(lq + r) * n + c] = acc[r];
      }
    }
  }
}
