// ocfd_mlp.hip -- multilayer perceptron for the learned chemistry surrogate
// (no reference counterpart): BF16 MFMA GEMMs with FP32 accumulation for the
// forward and backward passes, FP32 master weights and Adam on the device.
//
// Network: dims[0] -> dims[1] -> ... -> dims[L]; GELU (tanh form) after every
// layer but the last.  Weights W_l are [dims[l+1]][dims[l]] (out x in,
// row-major) and biases b_l [dims[l+1]].  Internally every dimension is
// padded with zeros to a multiple of 64 (inputs, hidden and output widths),
// so the GEMM tiles need no bounds checks; padded weights
// receive zero gradients and stay zero.
//
// GEMM kernel: 64x64 output tile per 256-thread workgroup, K in steps of 32
// staged through LDS in a canonical [row][k] layout, each wave computing a
// 32x32 sub-tile with v_mfma_f32_16x16x32_bf16.  Operand layouts:
//   A(m,k) = TA ? A[k*lda + m] : A[m*lda + k]
//   B(k,n) = TB ? B[n*ldb + k] : B[k*ldb + n]
// Split-K writes per-split partial tiles that a second kernel sums in a fixed
// order, so training is deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ocfd_hip.h"
#include "ocfd_mlp.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int TM = 64, TN = 64, TK = 32, LDT = TK + 8;  // LDS row: 40 bf16 = 80 B

enum { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_DGELU = 2 };

// GELU in its tanh form, 0.5 z (1 + tanh u) with u = c (z + 0.044715 z^3),
// evaluated as z * sigmoid(2u) = z * rcp(1 + exp2(z (C0 + C1 z^2))), C0 =
// -2 c log2(e), C1 = 0.044715 C0, with the hardware exp2 and reciprocal
// (v_exp_f32, v_rcp_f32: 1 ulp, far below the bfloat16 rounding of the
// activations; a correctly rounded division cost a sixth of the fused
// inference kernel's time).  The inference tile evaluates the same
// expression.
__device__ __forceinline__ float gelu_sig(float z) {
  constexpr float C0 = -2.0f * 0.7978845608028654f * 1.4426950408889634f, C1 = C0 * 0.044715f;
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * fmaf(z * z, C1, C0)));
}
__device__ __forceinline__ float gelu_f(float z) { return z * gelu_sig(z); }
__device__ __forceinline__ float dgelu_f(float z) {
  const float s = gelu_sig(z);
  return s + z * s * (1.0f - s) * 2.0f * 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * z * z);
}

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, int kchunk, const bf16* __restrict__ A, long lda,
                                              const bf16* __restrict__ B, long ldb, float* __restrict__ C, long ldc,
                                              long split_stride, int epi, const float* __restrict__ bias,
                                              const float* __restrict__ Zin, long ldz, bf16* __restrict__ Cb,
                                              long ldcb) {
  __shared__ __attribute__((aligned(16))) bf16 As[TM][LDT];
  __shared__ __attribute__((aligned(16))) bf16 Bs[TN][LDT];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int n0 = blockIdx.x * TN, m0 = blockIdx.y * TM;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(K, kb + kchunk);
  const int wr = w >> 1, wc = w & 1;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = kb; k0 < ke; k0 += TK) {
    if (!TA) {
      const int r = t >> 2, c8 = (t & 3) * 8;
      *(bf16x8*)&As[r][c8] = *(const bf16x8*)&A[(long)(m0 + r) * lda + k0 + c8];
    } else {
      const int kk = t >> 3, m8 = (t & 7) * 8;
      const bf16x8 v = *(const bf16x8*)&A[(long)(k0 + kk) * lda + m0 + m8];
#pragma unroll
      for (int j = 0; j < 8; j++) As[m8 + j][kk] = v[j];
    }
    if (TB) {
      const int r = t >> 2, c8 = (t & 3) * 8;
      *(bf16x8*)&Bs[r][c8] = *(const bf16x8*)&B[(long)(n0 + r) * ldb + k0 + c8];
    } else {
      const int kk = t >> 3, n8 = (t & 7) * 8;
      const bf16x8 v = *(const bf16x8*)&B[(long)(k0 + kk) * ldb + n0 + n8];
#pragma unroll
      for (int j = 0; j < 8; j++) Bs[n8 + j][kk] = v[j];
    }
    __syncthreads();
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; i++) a[i] = *(const bf16x8*)&As[wr * 32 + i * 16 + (lane & 15)][(lane >> 4) * 8];
#pragma unroll
    for (int j = 0; j < 2; j++) b[j] = *(const bf16x8*)&Bs[wc * 32 + j * 16 + (lane & 15)][(lane >> 4) * 8];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }

  float* Cs = C + blockIdx.z * split_stride;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int n = n0 + wc * 32 + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wr * 32 + i * 16 + (lane >> 4) * 4 + r;
        float v = acc[i][j][r];
        if (epi == EPI_BIAS_GELU) {
          v = v + bias[n];
          Cs[(long)m * ldc + n] = v;
          Cb[(long)m * ldcb + n] = (bf16)gelu_f(v);
        } else if (epi == EPI_DGELU) {
          v = v * dgelu_f(Zin[(long)m * ldz + n]);
          Cs[(long)m * ldc + n] = v;
          Cb[(long)m * ldcb + n] = (bf16)v;
        } else {
          if (bias) v = v + bias[n];
          Cs[(long)m * ldc + n] = v;
          if (Cb) Cb[(long)m * ldcb + n] = (bf16)v;
        }
      }
    }
}

// out[i] = sum_s part[s*stride + i], fixed order
__global__ void k_sum_splits(long n, int S, long stride, const float* __restrict__ part, float* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int k = 0; k < S; k++) s += part[k * stride + i];
  out[i] = s;
}

// column sums of a [rows][cols] fp32 matrix (bias gradient) in a fixed
// order: block (x, y) sums rows [y*CS_ROWS, (y+1)*CS_ROWS) of 64 columns with 4
// row lanes, the partials are then summed over y by k_sum_splits
constexpr int CS_ROWS = 256;
__global__ __launch_bounds__(256) void k_colsum_part(int rows, int cols, const float* __restrict__ a,
                                                     float* __restrict__ part) {
  __shared__ float red[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), q = threadIdx.x >> 6;
  const int r0 = blockIdx.y * CS_ROWS, r1 = min(rows, r0 + CS_ROWS);
  float s = 0.f;
  if (c < cols)
    for (int r = r0 + q; r < r1; r += 4) s += a[(long)r * cols + c];
  red[q][threadIdx.x & 63] = s;
  __syncthreads();
  if (q == 0 && c < cols) part[(long)blockIdx.y * cols + c] = ((red[0][threadIdx.x] + red[1][threadIdx.x]) +
                                                               (red[2][threadIdx.x] + red[3][threadIdx.x]));
}

// [B][din] fp32 -> [Bp][dinp] bf16 with zero padding
__global__ void k_pack_in(long B, long Bp, int din, int dinp, const float* __restrict__ x, bf16* __restrict__ xb) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Bp * dinp) return;
  const long r = i / dinp;
  const int c = (int)(i % dinp);
  xb[i] = (bf16)((r < B && c < din) ? x[r * din + c] : 0.f);
}

// [Bp][doutp] fp32 -> [B][dout] fp32
__global__ void k_unpack_out(long B, int dout, int doutp, const float* __restrict__ z, float* __restrict__ y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * dout) return;
  const long r = i / dout;
  const int c = (int)(i % dout);
  y[i] = z[r * doutp + c];
}

// MSE loss L = mean over (B x dout) of (z - y)^2; dz = dL/dz (zero on padding);
// per-block partial sums of the squared error, summed in order by k_loss_final
__global__ __launch_bounds__(256) void k_mse(long B, long Bp, int dout, int doutp, const float* __restrict__ z,
                                             const float* __restrict__ y, float* __restrict__ dz, bf16* __restrict__ dzb,
                                             float* __restrict__ part) {
  __shared__ float red[256];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.f;
  if (i < Bp * doutp) {
    const long r = i / doutp;
    const int c = (int)(i % doutp);
    float g = 0.f;
    if (r < B && c < dout) {
      const float e = z[i] - y[r * dout + c];
      e2 = e * e;
      g = 2.0f * e / (float)(B * dout);
    }
    dz[i] = g;
    dzb[i] = (bf16)g;
  }
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// mean of the squared-error partials: a fixed-shape tree in one workgroup
__global__ __launch_bounds__(256) void k_loss_final(int nb, long count, const float* __restrict__ part,
                                                    float* __restrict__ loss) {
  __shared__ double red[256];
  double s = 0.0;
  for (int k = threadIdx.x; k < nb; k += 256) s += part[k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = (float)(red[0] / (double)count);
}

__global__ void k_adam(long n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, bf16* __restrict__ pb, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  if (pb) pb[i] = (bf16)pi;
}

__global__ void k_to_bf16(long n, const float* __restrict__ a, bf16* __restrict__ b) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (bf16)a[i];
}

int pad_to(int x, int m) { return (x + m - 1) / m * m; }

struct Layer {
  int din = 0, dout = 0, dinp = 0, doutp = 0;
  float *W = nullptr, *b = nullptr, *gW = nullptr, *gb = nullptr;
  float *mW = nullptr, *vW = nullptr, *mb = nullptr, *vb = nullptr;
  bf16* Wb = nullptr;
  float* Z = nullptr;   // [Bp][doutp] pre-activation (output of the last layer)
  bf16* Ab = nullptr;   // [Bp][doutp] GELU(Z) (hidden layers)
  float* dZ = nullptr;  // [Bp][doutp]
  bf16* dZb = nullptr;
};

struct MlpCtx {
  bool ready = false;
  int L = 0;
  std::vector<Layer> lay;
  long Bcap = 0;
  bf16* Xb = nullptr;
  float* part = nullptr;  // split-K partials
  long part_cap = 0;
  float* lpart = nullptr;  // loss partials
  long lpart_cap = 0;
  float* loss = nullptr;  // device scalar
  long step = 0;
  unsigned long long version = 0;  // bumped whenever the bf16 weights change
  hipStream_t stream = nullptr;
  std::string err;
  MlpDevice dev{};
};
MlpCtx X;

int mfail(const std::string& m) {
  X.err = m;
  return 1;
}

template <typename T>
bool dalloc(T** p, size_t n) {
  return hipMalloc((void**)p, n * sizeof(T) + 16) == hipSuccess;
}

void free_batch() {
  for (auto& l : X.lay) {
    hipFree(l.Z);
    hipFree(l.Ab);
    hipFree(l.dZ);
    hipFree(l.dZb);
    l.Z = nullptr;
    l.Ab = nullptr;
    l.dZ = nullptr;
    l.dZb = nullptr;
  }
  hipFree(X.Xb);
  X.Xb = nullptr;
  X.Bcap = 0;
}

void free_all() {
  free_batch();
  for (auto& l : X.lay) {
    hipFree(l.W);
    hipFree(l.b);
    hipFree(l.gW);
    hipFree(l.gb);
    hipFree(l.mW);
    hipFree(l.vW);
    hipFree(l.mb);
    hipFree(l.vb);
    hipFree(l.Wb);
  }
  X.lay.clear();
  hipFree(X.part);
  hipFree(X.lpart);
  hipFree(X.loss);
  X.part = X.lpart = X.loss = nullptr;
  X.part_cap = X.lpart_cap = 0;
  X.ready = false;
}

int reserve(long B) {
  const long Bp = pad_to((int)B, 64);
  if (Bp <= X.Bcap) return 0;
  free_batch();
  for (auto& l : X.lay) {
    if (!dalloc(&l.Z, (size_t)Bp * l.doutp) || !dalloc(&l.Ab, (size_t)Bp * l.doutp) ||
        !dalloc(&l.dZ, (size_t)Bp * l.doutp) || !dalloc(&l.dZb, (size_t)Bp * l.doutp))
      return mfail("ocfd_mlp: out of device memory for the batch");
  }
  if (!dalloc(&X.Xb, (size_t)Bp * X.lay[0].dinp)) return mfail("ocfd_mlp: out of device memory for the batch");
  X.Bcap = Bp;
  return 0;
}

unsigned blocks(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

template <bool TA, bool TB>
int gemm(int M, int N, int K, int splits, const bf16* A, long lda, const bf16* B, long ldb, float* C, long ldc,
         long split_stride, int epi, const float* bias, const float* Zin, long ldz, bf16* Cb, long ldcb) {
  if (M % TM || N % TN || K % TK || splits < 1) return mfail("ocfd_mlp: GEMM shape is not a multiple of the tile");
  int kchunk = pad_to((K + splits - 1) / splits, TK);
  const int S = (K + kchunk - 1) / kchunk;
  hipLaunchKernelGGL((k_gemm<TA, TB>), dim3(N / TN, M / TM, S), dim3(256), 0, X.stream, M, N, K, kchunk, A, lda, B,
                     ldb, C, ldc, split_stride, epi, bias, Zin, ldz, Cb, ldcb);
  return hipGetLastError() == hipSuccess ? 0 : mfail("ocfd_mlp: GEMM launch failed");
}

// forward over Bp rows; input already packed in X.Xb
int forward(long Bp) {
  const bf16* in = X.Xb;
  for (int l = 0; l < X.L; l++) {
    Layer& y = X.lay[l];
    const bool last = (l == X.L - 1);
    if (gemm<false, true>((int)Bp, y.doutp, y.dinp, 1, in, y.dinp, y.Wb, y.dinp, y.Z, y.doutp, 0,
                          last ? EPI_STORE : EPI_BIAS_GELU, y.b, nullptr, 0, last ? nullptr : y.Ab, y.doutp))
      return 1;
    in = y.Ab;
  }
  return 0;
}

int split_count(long Bp) {
  long s = Bp / 512;
  if (s < 1) s = 1;
  if (s > 64) s = 64;
  return (int)s;
}

// loss and gradients (dL/dW, dL/db into gW, gb) for targets y [B][dout]
int backward(long B, long Bp, const float* y) {
  Layer& o = X.lay[X.L - 1];
  const long cnt = Bp * o.doutp;
  const int nb = (int)blocks(cnt);
  if (nb > X.lpart_cap) {
    hipFree(X.lpart);
    if (!dalloc(&X.lpart, (size_t)nb)) return mfail("ocfd_mlp: out of device memory");
    X.lpart_cap = nb;
  }
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, X.stream, B, Bp, o.dout, o.doutp, o.Z, y, o.dZ, o.dZb, X.lpart);
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(256), 0, X.stream, nb, B * (long)o.dout, X.lpart, X.loss);
  for (int l = X.L - 1; l >= 0; l--) {
    Layer& y_ = X.lay[l];
    const bf16* ain = (l == 0) ? X.Xb : X.lay[l - 1].Ab;
    const int S = split_count(Bp);
    const long need = std::max((long)S * y_.doutp * y_.dinp, (Bp + CS_ROWS - 1) / CS_ROWS * (long)y_.doutp);
    if (need > X.part_cap) {
      hipFree(X.part);
      if (!dalloc(&X.part, (size_t)need)) return mfail("ocfd_mlp: out of device memory");
      X.part_cap = need;
    }
    // dW = dZ^T . A_in : M = doutp, N = dinp, K = Bp
    const int kchunk = pad_to((int)((Bp + S - 1) / S), TK);
    const int Sx = (int)((Bp + kchunk - 1) / kchunk);
    if (gemm<true, false>(y_.doutp, y_.dinp, (int)Bp, S, y_.dZb, y_.doutp, ain, y_.dinp, X.part, y_.dinp,
                          (long)y_.doutp * y_.dinp, EPI_STORE, nullptr, nullptr, 0, nullptr, 0))
      return 1;
    const long nw = (long)y_.doutp * y_.dinp;
    hipLaunchKernelGGL(k_sum_splits, dim3(blocks(nw)), dim3(256), 0, X.stream, nw, Sx, nw, X.part, y_.gW);
    const int ry = (int)((Bp + CS_ROWS - 1) / CS_ROWS);
    hipLaunchKernelGGL(k_colsum_part, dim3((y_.doutp + 63) / 64, ry), dim3(256), 0, X.stream, (int)Bp, y_.doutp, y_.dZ,
                       X.part);
    hipLaunchKernelGGL(k_sum_splits, dim3(blocks(y_.doutp)), dim3(256), 0, X.stream, (long)y_.doutp, ry,
                       (long)y_.doutp, X.part, y_.gb);
    if (l > 0) {
      Layer& p = X.lay[l - 1];
      // dA_in = dZ . W : M = Bp, N = dinp, K = doutp; times GELU'(Z_in)
      if (gemm<false, false>((int)Bp, y_.dinp, y_.doutp, 1, y_.dZb, y_.doutp, y_.Wb, y_.dinp, p.dZ, p.doutp, 0,
                             EPI_DGELU, nullptr, p.Z, p.doutp, p.dZb, p.doutp))
        return 1;
    }
  }
  return hipGetLastError() == hipSuccess ? 0 : mfail("ocfd_mlp: launch failed");
}

int pack_input(long B, const float* x) {
  if (reserve(B)) return 1;
  const long Bp = pad_to((int)B, 64);
  const Layer& f = X.lay[0];
  hipLaunchKernelGGL(k_pack_in, dim3(blocks(Bp * f.dinp)), dim3(256), 0, X.stream, B, Bp, f.din, f.dinp, x, X.Xb);
  return 0;
}

void refresh_device_view() {
  MlpDevice& d = X.dev;
  d.n_layers = X.L;
  for (int l = 0; l < X.L && l < MLP_MAX_LAYERS; l++) {
    d.din[l] = X.lay[l].din;
    d.dout[l] = X.lay[l].dout;
    d.dinp[l] = X.lay[l].dinp;
    d.doutp[l] = X.lay[l].doutp;
    d.W[l] = X.lay[l].Wb;
    d.b[l] = X.lay[l].b;
  }
}

}  // namespace

const MlpDevice* ocfd_mlp_device_view() { return X.ready ? &X.dev : nullptr; }
unsigned long long ocfd_mlp_weights_version() { return X.version; }
hipStream_t ocfd_mlp_stream_handle() { return X.stream; }

extern "C" {

const char* ocfd_mlp_last_error(void) { return X.err.c_str(); }

int32_t ocfd_mlp_create(int32_t n_layers, const int32_t* dims) {
  if (n_layers < 1 || n_layers > MLP_MAX_LAYERS) return mfail("ocfd_mlp_create: 1..8 layers");
  for (int l = 0; l <= n_layers; l++)
    if (dims[l] < 1 || dims[l] > 4096) return mfail("ocfd_mlp_create: bad layer width");
  free_all();
  X.L = n_layers;
  X.lay.resize(n_layers);
  for (int l = 0; l < n_layers; l++) {
    Layer& y = X.lay[l];
    y.din = dims[l];
    y.dout = dims[l + 1];
    y.dinp = (l == 0) ? pad_to(y.din, 64) : X.lay[l - 1].doutp;
    y.doutp = pad_to(y.dout, 64);
    const size_t nw = (size_t)y.doutp * y.dinp;
    if (!dalloc(&y.W, nw) || !dalloc(&y.gW, nw) || !dalloc(&y.mW, nw) || !dalloc(&y.vW, nw) || !dalloc(&y.Wb, nw) ||
        !dalloc(&y.b, y.doutp) || !dalloc(&y.gb, y.doutp) || !dalloc(&y.mb, y.doutp) || !dalloc(&y.vb, y.doutp))
      return mfail("ocfd_mlp_create: out of device memory");
    if (hipMemset(y.W, 0, nw * 4) || hipMemset(y.mW, 0, nw * 4) || hipMemset(y.vW, 0, nw * 4) ||
        hipMemset(y.Wb, 0, nw * 2) || hipMemset(y.b, 0, y.doutp * 4) || hipMemset(y.mb, 0, y.doutp * 4) ||
        hipMemset(y.vb, 0, y.doutp * 4) || hipMemset(y.gW, 0, nw * 4) || hipMemset(y.gb, 0, y.doutp * 4))
      return mfail("ocfd_mlp_create: memset failed");
  }
  if (!X.stream && hipStreamCreateWithFlags(&X.stream, hipStreamNonBlocking) != hipSuccess)
    return mfail("ocfd_mlp_create: no HIP stream");
  if (!dalloc(&X.loss, 1)) return mfail("ocfd_mlp_create: out of device memory");
  X.step = 0;
  X.ready = true;
  refresh_device_view();
  X.version++;
  return 0;
}

int32_t ocfd_mlp_destroy(void) {
  if (X.stream) hipStreamSynchronize(X.stream);
  free_all();
  return 0;
}

int64_t ocfd_mlp_param_count(void) {
  if (!X.ready) return -1;
  int64_t n = 0;
  for (auto& l : X.lay) n += (int64_t)l.dout * l.din + l.dout;
  return n;
}

// flat host layout: per layer W [dout][din] row-major then b [dout]
int32_t ocfd_mlp_set_params(const float* flat) {
  if (!X.ready) return mfail("ocfd_mlp_set_params: no network");
  if (hipStreamSynchronize(X.stream) != hipSuccess) return mfail("ocfd_mlp_set_params: stream error");
  const float* p = flat;
  for (auto& l : X.lay) {
    std::vector<float> w((size_t)l.doutp * l.dinp, 0.f), b(l.doutp, 0.f);
    for (int o = 0; o < l.dout; o++)
      for (int i = 0; i < l.din; i++) w[(size_t)o * l.dinp + i] = *p++;
    for (int o = 0; o < l.dout; o++) b[o] = *p++;
    if (hipMemcpy(l.W, w.data(), w.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(l.b, b.data(), b.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
      return mfail("ocfd_mlp_set_params: copy failed");
    const long nw = (long)w.size();
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(nw)), dim3(256), 0, X.stream, nw, l.W, l.Wb);
    if (hipMemsetAsync(l.mW, 0, nw * 4, X.stream) || hipMemsetAsync(l.vW, 0, nw * 4, X.stream) ||
        hipMemsetAsync(l.mb, 0, l.doutp * 4, X.stream) || hipMemsetAsync(l.vb, 0, l.doutp * 4, X.stream))
      return mfail("ocfd_mlp_set_params: memset failed");
  }
  X.step = 0;
  X.version++;
  return hipStreamSynchronize(X.stream) == hipSuccess ? 0 : mfail("ocfd_mlp_set_params: stream error");
}

int32_t ocfd_mlp_get_params(float* flat) {
  if (!X.ready) return mfail("ocfd_mlp_get_params: no network");
  if (hipStreamSynchronize(X.stream) != hipSuccess) return mfail("ocfd_mlp_get_params: stream error");
  float* p = flat;
  for (auto& l : X.lay) {
    std::vector<float> w((size_t)l.doutp * l.dinp), b(l.doutp);
    if (hipMemcpy(w.data(), l.W, w.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(b.data(), l.b, b.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
      return mfail("ocfd_mlp_get_params: copy failed");
    for (int o = 0; o < l.dout; o++)
      for (int i = 0; i < l.din; i++) *p++ = w[(size_t)o * l.dinp + i];
    for (int o = 0; o < l.dout; o++) *p++ = b[o];
  }
  return 0;
}

int32_t ocfd_mlp_set_stream(void* stream) {
  if (!X.ready) return mfail("ocfd_mlp_set_stream: no network");
  X.stream = (hipStream_t)stream;
  return 0;
}

// y [B][dout] = net(x [B][din]); device pointers, fp32
int32_t ocfd_mlp_forward(int64_t B, const float* x, float* y) {
  if (!X.ready) return mfail("ocfd_mlp_forward: no network");
  if (B <= 0) return 0;
  if (pack_input(B, x)) return 1;
  const long Bp = pad_to((int)B, 64);
  if (forward(Bp)) return 1;
  const Layer& o = X.lay[X.L - 1];
  hipLaunchKernelGGL(k_unpack_out, dim3(blocks(B * o.dout)), dim3(256), 0, X.stream, (long)B, o.dout, o.doutp, o.Z, y);
  return hipGetLastError() == hipSuccess ? 0 : mfail("ocfd_mlp_forward: launch failed");
}

// loss (mean squared error) and its gradient for one batch; grad (device,
// may be NULL) receives the gradient in the flat parameter layout; loss
// (device, may be NULL) the loss.
int32_t ocfd_mlp_gradients(int64_t B, const float* x, const float* y, float* grad, float* loss) {
  if (!X.ready) return mfail("ocfd_mlp_gradients: no network");
  if (B <= 0) return mfail("ocfd_mlp_gradients: empty batch");
  if (pack_input(B, x)) return 1;
  const long Bp = pad_to((int)B, 64);
  if (forward(Bp) || backward(B, Bp, y)) return 1;
  if (grad) {
    float* p = grad;
    for (auto& l : X.lay) {
      if (hipMemcpy2DAsync(p, (size_t)l.din * 4, l.gW, (size_t)l.dinp * 4, (size_t)l.din * 4, l.dout,
                           hipMemcpyDeviceToDevice, X.stream) != hipSuccess ||
          hipMemcpyAsync(p + (size_t)l.dout * l.din, l.gb, (size_t)l.dout * 4, hipMemcpyDeviceToDevice, X.stream) !=
              hipSuccess)
        return mfail("ocfd_mlp_gradients: copy failed");
      p += (size_t)l.dout * l.din + l.dout;
    }
  }
  if (loss && hipMemcpyAsync(loss, X.loss, 4, hipMemcpyDeviceToDevice, X.stream) != hipSuccess)
    return mfail("ocfd_mlp_gradients: copy failed");
  return 0;
}

// one Adam step (bias-corrected) on the batch; loss (device, may be NULL)
int32_t ocfd_mlp_train_step(int64_t B, const float* x, const float* y, float lr, float beta1, float beta2, float eps,
                            float* loss) {
  if (!X.ready) return mfail("ocfd_mlp_train_step: no network");
  if (B <= 0) return mfail("ocfd_mlp_train_step: empty batch");
  if (pack_input(B, x)) return 1;
  const long Bp = pad_to((int)B, 64);
  if (forward(Bp) || backward(B, Bp, y)) return 1;
  X.step++;
  X.version++;
  const float c1 = 1.f - powf(beta1, (float)X.step), c2 = 1.f - powf(beta2, (float)X.step);
  for (auto& l : X.lay) {
    const long nw = (long)l.doutp * l.dinp;
    hipLaunchKernelGGL(k_adam, dim3(blocks(nw)), dim3(256), 0, X.stream, nw, l.W, l.gW, l.mW, l.vW, l.Wb, lr, beta1,
                       beta2, eps, c1, c2);
    hipLaunchKernelGGL(k_adam, dim3(blocks(l.doutp)), dim3(256), 0, X.stream, (long)l.doutp, l.b, l.gb, l.mb, l.vb,
                       (bf16*)nullptr, lr, beta1, beta2, eps, c1, c2);
  }
  if (loss && hipMemcpyAsync(loss, X.loss, 4, hipMemcpyDeviceToDevice, X.stream) != hipSuccess)
    return mfail("ocfd_mlp_train_step: copy failed");
  return hipGetLastError() == hipSuccess ? 0 : mfail("ocfd_mlp_train_step: launch failed");
}

int32_t ocfd_mlp_synchronize(void) {
  if (!X.stream) return 0;
  return hipStreamSynchronize(X.stream) == hipSuccess ? 0 : mfail("ocfd_mlp_synchronize: stream error");
}

// generic GEMM for tests: C[M][N] fp32 = op(A) op(B), bf16 device operands
int32_t ocfd_mlp_gemm_bf16(int32_t M, int32_t N, int32_t K, int32_t ta, int32_t tb, const void* A, int64_t lda,
                           const void* B, int64_t ldb, float* C, int64_t ldc, int32_t splits) {
  if (!X.stream && hipStreamCreateWithFlags(&X.stream, hipStreamNonBlocking) != hipSuccess)
    return mfail("ocfd_mlp_gemm_bf16: no HIP stream");
  if (M <= 0 || N <= 0 || K <= 0 || lda % 8 || ldb % 8 || ldc < N || lda < (ta ? M : K) || ldb < (tb ? K : N))
    return mfail("ocfd_mlp_gemm_bf16: bad shape or leading dimension (multiples of 8 required)");
  const bf16* a = (const bf16*)A;
  const bf16* b = (const bf16*)B;
  if (splits <= 1) {
    if (ta && tb) return gemm<true, true>(M, N, K, 1, a, lda, b, ldb, C, ldc, 0, EPI_STORE, nullptr, nullptr, 0, nullptr, 0);
    if (ta) return gemm<true, false>(M, N, K, 1, a, lda, b, ldb, C, ldc, 0, EPI_STORE, nullptr, nullptr, 0, nullptr, 0);
    if (tb) return gemm<false, true>(M, N, K, 1, a, lda, b, ldb, C, ldc, 0, EPI_STORE, nullptr, nullptr, 0, nullptr, 0);
    return gemm<false, false>(M, N, K, 1, a, lda, b, ldb, C, ldc, 0, EPI_STORE, nullptr, nullptr, 0, nullptr, 0);
  }
  // split-K into a scratch buffer, then the ordered sum
  const int kchunk = pad_to((K + splits - 1) / splits, TK);
  const int S = (K + kchunk - 1) / kchunk;
  const long need = (long)S * M * ldc;
  if (need > X.part_cap) {
    hipFree(X.part);
    if (!dalloc(&X.part, (size_t)need)) return mfail("ocfd_mlp_gemm_bf16: out of device memory");
    X.part_cap = need;
  }
  int rc;
  if (ta && tb) rc = gemm<true, true>(M, N, K, S, a, lda, b, ldb, X.part, ldc, (long)M * ldc, EPI_STORE, nullptr, nullptr, 0, nullptr, 0);
  else if (ta) rc = gemm<true, false>(M, N, K, S, a, lda, b, ldb, X.part, ldc, (long)M * ldc, EPI_STORE, nullptr, nullptr, 0, nullptr, 0);
  else if (tb) rc = gemm<false, true>(M, N, K, S, a, lda, b, ldb, X.part, ldc, (long)M * ldc, EPI_STORE, nullptr, nullptr, 0, nullptr, 0);
  else rc = gemm<false, false>(M, N, K, S, a, lda, b, ldb, X.part, ldc, (long)M * ldc, EPI_STORE, nullptr, nullptr, 0, nullptr, 0);
  if (rc) return rc;
  const long n = (long)M * ldc;
  hipLaunchKernelGGL(k_sum_splits, dim3(blocks(n)), dim3(256), 0, X.stream, n, S, n, X.part, C);
  return hipGetLastError() == hipSuccess ? 0 : mfail("ocfd_mlp_gemm_bf16: launch failed");
}

}  // extern "C"
