// ocfd_diag.h -- diagnostics reductions of the flow state on the device:
// show_flow_msg (code/flow-solver/OCFD_Time_Adv.f90:200-239) and ana_box
// (code/others/OCFD_ANA.f90:197-303).
//
// The per-cell terms are the reference's, in its evaluation order.  The sums
// are deterministic trees (per-thread strided runs, 64-lane shuffles, the
// four waves of a block in fixed order, then one wave over the block
// partials), so a result does not depend on the run, only on the grid; it
// differs from the reference's sequential sum by the summation order alone
// (relative ~ log2(N) * 2^-53 for sums of positive terms).
#pragma once

#define DIAG_SLOTS 16

// slot layout of one reduction pass:
//   pass 0 (show_flow_msg): 0 f5/Ajac, 1 d |u|^2 / 2 / Ajac
//   pass 1 (ana_box):       0 u^2, 1 v^2, 2 w^2, 3 ux^2, 4 ux^3, 5 ux^4, 6 d,
//                           7 c, 8 Amu, 9 |u|^2 d / 2, 10 T, 11 u^3, 12 u^4
//   pass 2 (ana_box):       0 (d - d_aver)^2
template <int PASS, int CD>
__device__ __forceinline__ void diag_terms(const DevState& S, long c, double d_aver, double* v) {
  int i, j, k;
  cell_ijk(S, c, i, j, k);
  const long q = hq(S, i, j, k);
  if (PASS == 0) {
    const double tmp0 = 1.0 / S.Ajac[q];
    v[0] = S.f[c + 4 * S.ni] * tmp0;
    v[1] = S.d[q] * (S.u[q] * S.u[q] + S.v[q] * S.v[q] + S.w[q] * S.w[q]) * 0.5 * tmp0;
  } else if (PASS == 1) {
    // exchange_boundary_xyz(u) + OCFD_dx0/dy0/dz0(u, ., Scheme_Vis)
    auto du = [&](int dir, int s) {
      return deriv1<CD>(S, dir, s, [&](int o) {
        const int t = wrap(s + o, S.n[dir], S.per[dir]);
        return S.u[q + (long)(t - s) * S.sh[dir]];
      });
    };
    const double ux = du(0, i) * S.Akx[q] + du(1, j) * S.Aix[q] + du(2, k) * S.Asx[q];
    const double gamma = S.p[q] / S.Et[q] + 1.0;
    const double c1 = sqrt(gamma * S.p[q] / S.d[q]);
    const double uu = S.u[q], vv = S.v[q], ww = S.w[q], dd = S.d[q];
    v[0] = uu * uu;
    v[1] = vv * vv;
    v[2] = ww * ww;
    v[3] = ux * ux;
    v[4] = (ux * ux) * ux;
    v[5] = (ux * ux) * (ux * ux);
    v[6] = dd;
    v[7] = c1;
    v[8] = S.Amu[c];
    v[9] = (uu * uu + vv * vv + ww * ww) * dd * 0.5;
    v[10] = S.T[q];
    v[11] = (uu * uu) * uu;
    v[12] = (uu * uu) * (uu * uu);
  } else {
    const double x = S.d[q] - d_aver;
    v[0] = x * x;
  }
}

template <int PASS> struct DiagN;
template <> struct DiagN<0> { static constexpr int N = 2; };
template <> struct DiagN<1> { static constexpr int N = 13; };
template <> struct DiagN<2> { static constexpr int N = 1; };

// block partial sums part[block * DIAG_SLOTS + slot]
template <int PASS, int CD>
__global__ __launch_bounds__(256) void k_diag_partial(DevState S, double* part, double d_aver) {
  constexpr int N = DiagN<PASS>::N;
  __shared__ double sm[4][N];
  double acc[N];
#pragma unroll
  for (int a = 0; a < N; a++) acc[a] = 0.0;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < S.ni; c += (long)gridDim.x * blockDim.x) {
    double v[N];
    diag_terms<PASS, CD>(S, c, d_aver, v);
#pragma unroll
    for (int a = 0; a < N; a++) acc[a] = acc[a] + v[a];
  }
#pragma unroll
  for (int a = 0; a < N; a++) {
    double x = acc[a];
    for (int o = 32; o > 0; o >>= 1) x = x + __shfl_xor(x, o, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6][a] = x;
  }
  __syncthreads();
  if (threadIdx.x < N) {
    const int a = threadIdx.x;
    part[(long)blockIdx.x * DIAG_SLOTS + a] = ((sm[0][a] + sm[1][a]) + sm[2][a]) + sm[3][a];
  }
}

// out[slot] = (sum over the block partials) / (nx ny nz), one wave per slot
__global__ __launch_bounds__(64) void k_diag_final(DevState S, const double* part, int nparts, int nslots, double* out) {
  const int a = blockIdx.x;
  if (a >= nslots) return;
  double x = 0.0;
  for (int b = threadIdx.x; b < nparts; b += 64) x = x + part[(long)b * DIAG_SLOTS + a];
  for (int o = 32; o > 0; o >>= 1) x = x + __shfl_xor(x, o, 64);
  if (threadIdx.x == 0) out[a] = x / (1.0 * S.nx * S.ny * S.nz);
}
