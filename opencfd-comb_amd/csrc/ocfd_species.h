// ocfd_species.h -- kernels of the multi-species flow path (N_SPEC > 1):
// mixture thermodynamics (perfect or NASA-fit real gas), transport
// coefficients (Sutherland/Pr/Sc or Chemkin fits with Wilke mixing), species
// convection and diffusion, the species RK3 update, primitive recovery with
// Rescaling_di, face re-synthesis and the jet boundary condition.
//
// Same conventions as ocfd_device.h: FP64, no FMA contraction, the
// reference's operation order.  Included once, by ocfd_flow.hip.
#pragma once
#include <type_traits>

#include "ocfd_device.h"

#define GAS_NT_MAX 4
#define GAS_R0 8.314       // CHEM_modules.f90:20
#define GAS_ATM 1.013e5    // CHEM_modules.f90:22

// Species data of the flow path (read_Spec, CHEM_thermdyn.f90:16-51; the
// transport fits of TRANSPORT_modules.f90:80-253).
struct FlowGas {
  int ns, n_tct, real_gas, flag_vis, flag_cond, flag_diff;
  double tct[GAS_NT_MAX], tct_max;
  double Mi[OCFD_NS_MAX], Ri[OCFD_NS_MAX];
  double coef[OCFD_NS_MAX][GAS_NT_MAX][7], ect[OCFD_NS_MAX][GAS_NT_MAX];
  double mu_an[OCFD_NS_MAX][4], k_an[OCFD_NS_MAX][4], D_an[OCFD_NS_MAX][OCFD_NS_MAX][4];
  // Wilke factors that depend only on the molar masses, evaluated on the
  // host: wq = (M_j/M_k)^(1/4) and wrec = 1/sqrt(8 (1 + M_k/M_j))
  double wq[OCFD_NS_MAX][OCFD_NS_MAX], wrec[OCFD_NS_MAX][OCFD_NS_MAX];
};
__constant__ FlowGas fg;

__device__ __forceinline__ double gp3(double x) { double x2 = x * x; return x2 * x; }
__device__ __forceinline__ double gp4(double x) { double x2 = x * x; return x2 * x2; }
__device__ __forceinline__ double gp5(double x) { double x2 = x * x; return (x2 * x2) * x; }

__device__ __forceinline__ int gas_range(double T) {
  int jc = 0;
  for (int j = 0; j < fg.n_tct - 1; j++)
    if (T > fg.tct[j]) jc = j + 1;
  return jc;
}

// comput_E_pgas (CHEM_thermdyn.f90:114-122) / comput_E_rgas (:203-231)
__device__ double gas_E(const double* di, double T) {
  const int ns = fg.ns;
  if (!fg.real_gas) {
    double E = 0.0;
    for (int k = 0; k < ns; k++) E = E + 2.5 * fg.Ri[k] * di[k] * T;
    return E;
  }
  double A1 = 0, B1 = 0, C1 = 0, D1 = 0, E1 = 0, Ect = 0;
  const int jc = gas_range(T);
  for (int k = 0; k < ns; k++) {
    const double R1 = fg.Ri[k];
    A1 = A1 + (fg.coef[k][jc][0] - 1.0) * R1 * di[k];
    B1 = B1 + fg.coef[k][jc][1] * R1 * di[k];
    C1 = C1 + fg.coef[k][jc][2] * R1 * di[k];
    D1 = D1 + fg.coef[k][jc][3] * R1 * di[k];
    E1 = E1 + fg.coef[k][jc][4] * R1 * di[k];
    Ect = Ect + fg.ect[k][jc] * R1 * di[k];
  }
  if (T <= fg.tct_max) return A1 * T + B1 * (T * T) / 2.0 + C1 * gp3(T) / 3.0 + D1 * gp4(T) / 4.0 + E1 * gp5(T) / 5.0 + Ect;
  const double tm = fg.tct_max;
  const double Cv2 = A1 + B1 * tm + C1 * (tm * tm) + D1 * gp3(tm) + E1 * gp4(tm);
  return A1 * tm + B1 * (tm * tm) / 2.0 + C1 * gp3(tm) / 3.0 + D1 * gp4(tm) / 4.0 + E1 * gp5(tm) / 5.0 + Ect +
         Cv2 * (T - tm);
}

// comput_T_pgas (CHEM_thermdyn.f90:157-168) / comput_T_rgas (:326-382)
__device__ double gas_T(const double* di, double E) {
  const int ns = fg.ns;
  if (!fg.real_gas) {
    double rCv = 0.0;
    for (int k = 0; k < ns; k++) rCv = rCv + 2.5 * fg.Ri[k] * di[k];
    return E / rCv;
  }
  double A1[GAS_NT_MAX], B1[GAS_NT_MAX], C1[GAS_NT_MAX], D1[GAS_NT_MAX], E1[GAS_NT_MAX], Ect[GAS_NT_MAX];
#pragma unroll
  for (int j = 0; j < GAS_NT_MAX; j++) {
    A1[j] = B1[j] = C1[j] = D1[j] = E1[j] = Ect[j] = 0.0;
    if (j < fg.n_tct) {
      for (int k = 0; k < ns; k++) {
        const double R1 = fg.Ri[k];
        A1[j] = A1[j] + (fg.coef[k][j][0] - 1.0) * R1 * di[k];
        B1[j] = B1[j] + fg.coef[k][j][1] * R1 * di[k];
        C1[j] = C1[j] + fg.coef[k][j][2] * R1 * di[k];
        D1[j] = D1[j] + fg.coef[k][j][3] * R1 * di[k];
        E1[j] = E1[j] + fg.coef[k][j][4] * R1 * di[k];
        Ect[j] = Ect[j] + fg.ect[k][j] * R1 * di[k];
      }
    }
  }
  double T1 = E / A1[0], T2 = T1;
  for (int ks = 0; ks <= 30; ks++) {
    const int jc = gas_range(T1);
    double a = A1[0], b = B1[0], c = C1[0], d = D1[0], e = E1[0], ect = Ect[0];
#pragma unroll
    for (int j = 1; j < GAS_NT_MAX; j++)
      if (jc == j) { a = A1[j]; b = B1[j]; c = C1[j]; d = D1[j]; e = E1[j]; ect = Ect[j]; }
    double Et1, Ex;
    if (T1 <= fg.tct_max) {
      Et1 = a * T1 + b * (T1 * T1) / 2.0 + c * gp3(T1) / 3.0 + d * gp4(T1) / 4.0 + e * gp5(T1) / 5.0 + ect;
      Ex = a + b * T1 + c * (T1 * T1) + d * gp3(T1) + e * gp4(T1);
    } else {
      const double tm = fg.tct_max;
      Ex = a + b * tm + c * (tm * tm) + d * gp3(tm) + e * gp4(tm);
      Et1 = a * tm + b * (tm * tm) / 2.0 + c * gp3(tm) / 3.0 + d * gp4(tm) / 4.0 + e * gp5(tm) / 5.0 + ect +
            Ex * (T1 - tm);
    }
    T2 = T1 - (Et1 - E) / Ex;
    const double error = fabs(T2 - T1);
    T1 = T2;
    if (error < 1.0e-3) break;
  }
  return T2;
}

// comput_P, CHEM_thermdyn.f90:385-395
__device__ __forceinline__ double gas_P(const double* di, double T) {
  double p = 0.0;
  for (int k = 0; k < fg.ns; k++) p = p + fg.Ri[k] * di[k] * T;
  return p;
}

// comput_d, CHEM_thermdyn.f90:399-409 (ai = mass fractions)
__device__ __forceinline__ double gas_d(const double* ai, double T, double p) {
  double R1 = 0.0;
  for (int k = 0; k < fg.ns; k++) R1 = R1 + fg.Ri[k] * ai[k];
  return p / (T * R1);
}

// comput_Cp_pgas (CHEM_thermdyn.f90:125-136) / comput_Cp_rgas (:234-266)
__device__ double gas_Cp(double d, const double* di, double T) {
  const int ns = fg.ns;
  if (!fg.real_gas) {
    double Cp = 0.0;
    for (int k = 0; k < ns; k++) Cp = Cp + 3.5 * fg.Ri[k] * di[k];
    return Cp / d;
  }
  double A1 = 0, B1 = 0, C1 = 0, D1 = 0, E1 = 0;
  const int jc = gas_range(T);
  for (int k = 0; k < ns; k++) {
    const double R1 = fg.Ri[k];
    A1 = A1 + fg.coef[k][jc][0] * R1 * di[k];
    B1 = B1 + fg.coef[k][jc][1] * R1 * di[k];
    C1 = C1 + fg.coef[k][jc][2] * R1 * di[k];
    D1 = D1 + fg.coef[k][jc][3] * R1 * di[k];
    E1 = E1 + fg.coef[k][jc][4] * R1 * di[k];
  }
  const double t = T <= fg.tct_max ? T : fg.tct_max;
  return (A1 + B1 * t + C1 * (t * t) + D1 * gp3(t) + E1 * gp4(t)) / d;
}

// comput_hi_pgas (CHEM_thermdyn.f90:176-182) / comput_hi_rgas (:271-293)
__device__ double gas_hi(int k, double T) {
  if (!fg.real_gas) return 3.5 * fg.Ri[k] * T;
  const int jc = gas_range(T);
  const double c0 = fg.coef[k][jc][0], c1 = fg.coef[k][jc][1], c2 = fg.coef[k][jc][2], c3 = fg.coef[k][jc][3],
               c4 = fg.coef[k][jc][4], ect = fg.ect[k][jc];
  if (T <= fg.tct_max)
    return (c0 * T + c1 * (T * T) / 2.0 + c2 * gp3(T) / 3.0 + c3 * gp4(T) / 4.0 + c4 * gp5(T) / 5.0 + ect) * fg.Ri[k];
  const double tm = fg.tct_max;
  const double Cp_max = c0 + c1 * tm + c2 * (tm * tm) + c3 * gp3(tm) + c4 * gp4(tm);
  return (c0 * tm + c1 * (tm * tm) / 2.0 + c2 * gp3(tm) / 3.0 + c3 * gp4(tm) / 4.0 + c4 * gp5(tm) / 5.0 + ect +
          Cp_max * (T - tm)) * fg.Ri[k];
}

// Chemkin fit exp(a1 + a2 lnT + a3 lnT^2 + a4 lnT^3), TRANSPORT_modules.f90:276-277
__device__ __forceinline__ double chemkin_fit(const double* a, double lT) {
  return exp(a[0] + a[1] * lT + a[2] * (lT * lT) + a[3] * gp3(lT));
}

__device__ __forceinline__ void cell_ijk(const DevState& S, long c, int& i, int& j, int& k) {
  i = (int)(c % S.nx) + 1;
  j = (int)((c / S.nx) % S.ny) + 1;
  k = (int)(c / ((long)S.nx * S.ny)) + 1;
}

// ===========================================================================
// comput_Amu_Amk_AMD (Comput_flow_variables.f90:69-98) for N_SPEC > 1:
// comput_vis_coeff (TRANSPORT_modules.f90:259-306), comput_Cp,
// comput_thermal_conductivity (:312-372), comput_diffusion_coeff (:378-424).
// ===========================================================================
__global__ __launch_bounds__(256, 5) void k_transport_ms(DevState S) {
  const long c = cell_index();
  if (c >= S.ni) return;
  int i, j, k;
  cell_ijk(S, c, i, j, k);
  const long q = hq(S, i, j, k);
  const int ns = fg.ns;
  // species loops run to the fixed bound with guards, fully unrolled, so the
  // per-species arrays stay in registers
  double di0[OCFD_NS_MAX], Xi[OCFD_NS_MAX];
  double X0 = 0.0;
#pragma unroll
  for (int m = 0; m < OCFD_NS_MAX; m++) {
    di0[m] = m < ns ? S.di[q + (long)m * S.nh] : 0.0;
    Xi[m] = m < ns ? di0[m] / fg.Mi[m] : 0.0;
    if (m < ns) X0 = X0 + Xi[m];
  }
#pragma unroll
  for (int m = 0; m < OCFD_NS_MAX; m++)
    if (m < ns) Xi[m] = Xi[m] / X0;
  const double T = S.T[q];
  double amu;
  if (fg.flag_vis == 0) {
    const double TA = 288.15, TC = 110.4;
    const double tr = T / TA;
    amu = S.amu0 * sqrt(gp3(tr)) * (TA + TC) / (TC + T);
  } else {
    const double lT = log(T);
    double mui[OCFD_NS_MAX];
#pragma unroll
    for (int m = 0; m < OCFD_NS_MAX; m++) mui[m] = m < ns ? chemkin_fit(fg.mu_an[m], lT) * 0.1 : 1.0;
    // Wilke's Phi_kj = (1 + sqrt(mu_k/mu_j) (M_j/M_k)^(1/4))^2 / sqrt(8 (1 +
    // M_k/M_j)) (comput_vis_coeff, TRANSPORT_modules.f90:296-305) with
    // sqrt(mu_k/mu_j) = sqrt(mu_k) / sqrt(mu_j): 9 square roots and 9
    // divisions per cell instead of 81 and 162 (a few ulps from the
    // reference's evaluation)
    double smu[OCFD_NS_MAX], rmu[OCFD_NS_MAX];
#pragma unroll
    for (int m = 0; m < OCFD_NS_MAX; m++) {
      smu[m] = sqrt(mui[m]);
      rmu[m] = 1.0 / smu[m];
    }
    double mu = 0.0;
#pragma unroll
    for (int kk = 0; kk < OCFD_NS_MAX; kk++) {
      if (kk < ns) {
        double Pk0 = 0.0;
#pragma unroll
        for (int jj = 0; jj < OCFD_NS_MAX; jj++) {
          if (jj < ns) {
            const double a = 1.0 + smu[kk] * rmu[jj] * fg.wq[kk][jj];
            const double Pkj = (a * a) * fg.wrec[kk][jj];
            Pk0 = Pk0 + Xi[jj] * Pkj;
          }
        }
        mu = mu + Xi[kk] * mui[kk] / Pk0;
      }
    }
    amu = mu;
  }
  const double Cp = gas_Cp(S.d[q], di0, T);
  double amk;
  if (fg.flag_cond == 0) {
    amk = amu * Cp / S.pr;
  } else {
    const double lT = log(T);
    double Amk1 = 0.0, Amk2 = 0.0;
#pragma unroll
    for (int m = 0; m < OCFD_NS_MAX; m++)
      if (m < ns) {
        const double ki = chemkin_fit(fg.k_an[m], lT) * 1.0e-5;
        Amk1 = Amk1 + Xi[m] * ki;
        Amk2 = Amk2 + Xi[m] / ki;
      }
    amk = 0.5 * (Amk1 + 1.0 / Amk2);
  }
  S.Amu[c] = amu;
  S.Amk[c] = amk;
  if (fg.flag_diff == 0) {
    double d0 = 0.0;
#pragma unroll
    for (int m = 0; m < OCFD_NS_MAX; m++)
      if (m < ns) d0 = d0 + Xi[m] * fg.Mi[m];
#pragma unroll
    for (int m = 0; m < OCFD_NS_MAX; m++)
      if (m < ns)
        S.AmD[c + (long)m * S.ni] = amu * (1.0 - Xi[m] * fg.Mi[m] / d0) / ((1.0 - Xi[m]) * S.sc * S.d[q] + 1.0e-20);
  } else {
    // binary coefficients D(k,j) = D(j,k) from the fit of the pair k < j (the
    // reference fills the lower triangle from the upper one), each evaluated
    // once.  Visiting the pairs (a, b), a < b, with a outer and b inner adds
    // the terms of every ds_k in increasing j, the reference's order.
    const double lT = log(T);
    const double pq = S.p[q];
    double ds[OCFD_NS_MAX];
#pragma unroll
    for (int m = 0; m < OCFD_NS_MAX; m++) ds[m] = 0.0;
#pragma unroll
    for (int a = 0; a < OCFD_NS_MAX; a++)
#pragma unroll
      for (int b = a + 1; b < OCFD_NS_MAX; b++)
        if (b < ns) {
          const double rD = 1.0 / (chemkin_fit(fg.D_an[a][b], lT) * 1.0e-4);
          ds[a] = ds[a] + Xi[b] * rD;
          ds[b] = ds[b] + Xi[a] * rD;
        }
#pragma unroll
    for (int kk = 0; kk < OCFD_NS_MAX; kk++) {
      if (kk < ns)
        S.AmD[c + (long)kk * S.ni] = (1.0 - Xi[kk]) / (ds[kk] + 1.0e-20) * GAS_ATM / pq;
    }
  }
}

// ===========================================================================
// du_scaler_invis (OCFD_scaler_inviscous.f90:1-84) for all species at once:
// scalar Steger-Warming split F+- = (vs +- |vs|)/2 * di_m with vs = A1 u +
// A2 v + A3 w, the flow's reconstructions and closures, the same hybrid masks.
// Each species only reads its own di, so doing them together equals the
// reference's one-species-at-a-time order.
// ===========================================================================
template <int DIR> struct SpTile;
template <> struct SpTile<0> { static constexpr int NP = 4, TS = 64; };
template <> struct SpTile<1> { static constexpr int NP = 8, TS = 32; };
template <> struct SpTile<2> { static constexpr int NP = 8, TS = 32; };

template <int DIR>
__global__ __launch_bounds__(256) void k_species_sweep(DevState S, int m0, int m1) {
  constexpr int NP = SpTile<DIR>::NP, TS = SpTile<DIR>::TS;
  constexpr int NPT = TS + 8, NIF = TS + 1;
  // one species at a time through small LDS buffers: the split velocity,
  // the point index and the reconstruction kind are formed once
  __shared__ double sV[NP * NPT];
  __shared__ long sQ[NP * NPT];
  __shared__ double sF[2 * NP * NPT];
  __shared__ double sH[2 * NP * NIF];
  __shared__ uint8_t sK[2 * NP * NIF];
  __shared__ PencilBase<DIR, NP> PB;
  const int ns = S.ns;
  const int n = S.n[DIR];
  const long nb = (long)gridDim.x * gridDim.y;
  const long lb = xcd_block((long)blockIdx.y * gridDim.x + blockIdx.x, nb);
  const int bx = (int)(lb % gridDim.x);
  const long by = lb / gridDim.x;
  const int c0 = bx * TS + 1;
  const int ncell = min(TS, n - c0 + 1);
  const long npen = DIR == 0 ? (long)S.ny * S.nz : DIR == 1 ? (long)S.nx * S.nz : (long)S.nx * S.ny;
  const long pen0 = by * NP;
  auto lp = [](int pl, int s) { return DIR == 0 ? pl * NPT + s : s * NP + pl; };
  auto lh = [](int pl, int l) { return DIR == 0 ? pl * NIF + l : l * NP + pl; };
  const double* M1 = DIR == 0 ? S.Akx1 : DIR == 1 ? S.Aix1 : S.Asx1;
  const double* M2 = DIR == 0 ? S.Aky1 : DIR == 1 ? S.Aiy1 : S.Asy1;
  const double* M3 = DIR == 0 ? S.Akz1 : DIR == 1 ? S.Aiz1 : S.Asz1;
  const int tid = threadIdx.x;
  pencil_bases<DIR, NP>(S, pen0, npen, PB);
  __syncthreads();
  for (int t = tid; t < NP * NPT; t += blockDim.x) {
    const int pl = DIR == 0 ? t / NPT : t % NP;
    const int sl = DIR == 0 ? t % NPT : t / NP;
    const long pen = pen0 + pl;
    if (pen >= npen || sl > ncell + 7) continue;
    const long q = PB.hq1[pl] + (long)(wrap(c0 - 4 + sl, n, S.per[DIR]) - 1) * S.sh[DIR];
    sQ[lp(pl, sl)] = q;
    sV[lp(pl, sl)] = M1[q] * S.u[q] + M2[q] * S.v[q] + M3[q] * S.w[q];
  }
  const bool hyb = S.scheme == OCFD_SCHEME_HYBRID;
  for (int t = tid; t < 2 * NP * NIF; t += blockDim.x) {
    const int side = t / (NP * NIF);
    const int r = t % (NP * NIF);
    const int pl = DIR == 0 ? r / NIF : r % NP;
    const int l = DIR == 0 ? r % NIF : r / NP;
    const long pen = pen0 + pl;
    if (pen >= npen || l > ncell) continue;
    const int s_if = c0 - 1 + l;
    const int mask = hyb ? (int)S.mask[DIR][PB.mb[pl] + (long)s_if * mask_stride<DIR>(S)] : 0;
    sK[t] = (uint8_t)recon_kind(S, DIR, s_if, side, mask);
  }
  const double h1 = S.h_1[DIR];
  (void)ns;
  for (int m = m0; m < m1; m++) {
    const double* Di = S.di + (long)m * S.nh;
    for (int t = tid; t < NP * NPT; t += blockDim.x) {
      const int pl = DIR == 0 ? t / NPT : t % NP;
      const int sl = DIR == 0 ? t % NPT : t / NP;
      if (pen0 + pl >= npen || sl > ncell + 7) continue;
      const int a = lp(pl, sl);
      const double vs = sV[a], fs = Di[sQ[a]];
      sF[a] = 0.5 * (vs + fabs(vs)) * fs;
      sF[NP * NPT + a] = 0.5 * (vs - fabs(vs)) * fs;
    }
    __syncthreads();  // sF complete; the previous species' outputs have read sH
    for (int t = tid; t < 2 * NP * NIF; t += blockDim.x) {
      const int side = t / (NP * NIF);
      const int r = t % (NP * NIF);
      const int pl = DIR == 0 ? r / NIF : r % NP;
      const int l = DIR == 0 ? r % NIF : r / NP;
      if (pen0 + pl >= npen || l > ncell) continue;
      const double* F = sF + side * NP * NPT;
      double w[8];
#pragma unroll
      for (int qq = 0; qq < 8; qq++) w[qq] = F[lp(pl, side ? l + 7 - qq : l + qq)];
      sH[t] = recon(sK[t], w, S.ud7l);
    }
    __syncthreads();  // sH complete; sF may be overwritten
    for (int t = tid; t < NP * TS; t += blockDim.x) {
      const int pl = DIR == 0 ? t / TS : t % NP;
      const int c = DIR == 0 ? t % TS : t / NP;
      const long pen = pen0 + pl;
      if (pen >= npen || c >= ncell) continue;
      const double hp1 = sH[lh(pl, c + 1)], hp0 = sH[lh(pl, c)];
      const double hm1 = sH[NP * NIF + lh(pl, c + 1)], hm0 = sH[NP * NIF + lh(pl, c)];
      const double val = (hp1 - hp0 + hm1 - hm0) * h1;
      const long ci = PB.iq1[pl] + (long)(c0 + c - 1) * S.si[DIR] + (long)m * S.ni;
      if (DIR == 0) {
        S.dus[ci] = val;
      } else if (DIR == 1) {
        S.dus[ci] = S.dus[ci] + val;
      } else {
        const long q = sQ[lp(pl, c + 4)];
        S.dus[ci] = -(S.dus[ci] + val) * S.Ajac[q];
      }
    }
  }
}

// ===========================================================================
// du_scaler_viscous (OCFD_scaler_viscous.f90:4-79), part 1 for all species:
// grad(c_m) with c_m = di_m/d, the flux rho D_m grad(c_m) in the three
// computational directions, and Qp = sum_m rho D_m h_m grad(c_m) in species
// order (zeroed at the first species).  c_m is formed once per cell by
// k_species_conc (the reference's Ev1 = fs/d, :25-31).
// ===========================================================================
__global__ __launch_bounds__(256) void k_species_conc(DevState S, int m0, int m1) {
  const long c = cell_index();
  if (c >= S.ni) return;
  int i, j, k;
  cell_ijk(S, c, i, j, k);
  const long q = hq(S, i, j, k);
  const double d = S.d[q];
  for (int m = m0; m < m1; m++) S.cs[c + (long)m * S.ni] = S.di[q + (long)m * S.nh] / d;
}

// INT: every stencil of the cell is central (no wrap, no closure), so each
// species' derivatives are straight-line code whose loads issue together.
// Species m0 .. m1-1; Qp starts from zero at the first species (m0 = 0) and
// from its stored value otherwise, so species taken one at a time add up in
// the same order as all at once.
template <int CD, bool IX, bool IY, bool IZ>
__device__ __forceinline__ void species_visc_prep_cell(const DevState& S, long c, long q, int i, int j, int k, int m0,
                                                       int m1) {
  const double T0 = S.T[q], d0 = S.d[q];
  const double Akx = S.Akx[q], Aky = S.Aky[q], Akz = S.Akz[q];
  const double Aix = S.Aix[q], Aiy = S.Aiy[q], Aiz = S.Aiz[q];
  const double Asx = S.Asx[q], Asy = S.Asy[q], Asz = S.Asz[q];
  const double Akx1 = S.Akx1[q], Aky1 = S.Aky1[q], Akz1 = S.Akz1[q];
  const double Aix1 = S.Aix1[q], Aiy1 = S.Aiy1[q], Aiz1 = S.Aiz1[q];
  const double Asx1 = S.Asx1[q], Asy1 = S.Asy1[q], Asz1 = S.Asz1[q];
  const long ni = S.ni;
  double Qpx = 0.0, Qpy = 0.0, Qpz = 0.0;
  if (m0 > 0) {
    Qpx = S.Qp[c];
    Qpy = S.Qp[c + ni];
    Qpz = S.Qp[c + 2 * ni];
  }
  for (int m = m0; m < m1; m++) {
    const double* C = S.cs + (long)m * ni + c;
    auto dc = [&](auto central, int dir, int s) {
      const long st = S.si[dir];
      if constexpr (decltype(central)::value) {
        return deriv1<CD, true>(S, dir, s, [&](int o) { return C[o * st]; });
      } else {
        return deriv1<CD>(S, dir, s, [&](int o) {
          const int t = wrap(s + o, S.n[dir], S.per[dir]);
          return C[(t - s) * st];
        });
      }
    };
    const double uk = dc(std::bool_constant<IX>{}, 0, i), ui = dc(std::bool_constant<IY>{}, 1, j),
                 us = dc(std::bool_constant<IZ>{}, 2, k);
    const double dD = d0 * S.AmD[c + (long)m * ni];
    const double hi = gas_hi(m, T0);
    const double cix = uk * Akx + ui * Aix + us * Asx;
    const double ciy = uk * Aky + ui * Aiy + us * Asy;
    const double ciz = uk * Akz + ui * Aiz + us * Asz;
    const double E1 = dD * cix, E2 = dD * ciy, E3 = dD * ciz;
    Qpx = Qpx + E1 * hi;
    Qpy = Qpy + E2 * hi;
    Qpz = Qpz + E3 * hi;
    S.Evs[c + (long)(3 * m) * ni] = E1 * Akx1 + E2 * Aky1 + E3 * Akz1;
    S.Evs[c + (long)(3 * m + 1) * ni] = E1 * Aix1 + E2 * Aiy1 + E3 * Aiz1;
    S.Evs[c + (long)(3 * m + 2) * ni] = E1 * Asx1 + E2 * Asy1 + E3 * Asz1;
  }
  S.Qp[c] = Qpx;
  S.Qp[c + ni] = Qpy;
  S.Qp[c + 2 * ni] = Qpz;
}

template <int CD>
__global__ __launch_bounds__(256) void k_species_visc_prep(DevState S, int m0, int m1) {
  constexpr int R = CD == 1 ? 4 : 3;
  const long c = cell_index();
  if (c >= S.ni) return;
  int i, j, k;
  cell_ijk(S, c, i, j, k);
  const long q = hq(S, i, j, k);
  // central forms per direction: a wavefront at the xi faces keeps the
  // central eta and zeta forms
  const bool ix = wave_all(central_cell(i, S.nx, R));
  const bool iyz = wave_all(central_cell(j, S.ny, R) && central_cell(k, S.nz, R));
  if (ix && iyz) species_visc_prep_cell<CD, true, true, true>(S, c, q, i, j, k, m0, m1);
  else if (iyz) species_visc_prep_cell<CD, false, true, true>(S, c, q, i, j, k, m0, m1);
  else species_visc_prep_cell<CD, false, false, false>(S, c, q, i, j, k, m0, m1);
}

// Species update: Dus += div(diffusive flux) Ajac (OCFD_scaler_viscous.f90:67-77)
// and Scalar_time_adv_RK3 (OCFD_Time_Adv.f90:134-166) for every species;
// krk = 0 stores the complete right-hand side in dus instead.
template <int CD, bool IX, bool IY, bool IZ>
__device__ __forceinline__ void species_rk_cell(const DevState& S, long c, long q, int i, int j, int k, int krk,
                                                double dt, int visc, int m0, int m1) {
  const long ni = S.ni;
  const double ajac = S.Ajac[q];
  auto dv = [&](auto central, const double* F, int dir, int s) {
    const double* Fc = F + c;
    const long st = S.si[dir];
    if constexpr (decltype(central)::value) {
      return deriv1<CD, true>(S, dir, s, [&](int o) { return Fc[o * st]; });
    } else {
      return deriv1<CD>(S, dir, s, [&](int o) {
        const int t = wrap(s + o, S.n[dir], S.per[dir]);
        return Fc[(t - s) * st];
      });
    }
  };
  const double ra = krk == 2 ? 3.0 / 4.0 : 1.0 / 3.0;
  const double rb = krk == 2 ? 1.0 / 4.0 : 2.0 / 3.0;
  for (int m = m0; m < m1; m++) {
    const long qm = q + (long)m * S.nh, cm = c + (long)m * ni;
    // the stage's old values, read with the stencil loads
    const double dn = krk >= 1 ? S.din[cm] : 0.0;
    const double dold = krk >= 2 ? S.di[qm] : 0.0;
    double val = S.dus[cm];
    if (visc) {
      const double uk = dv(std::bool_constant<IX>{}, S.Evs + (long)(3 * m) * ni, 0, i);
      const double ui = dv(std::bool_constant<IY>{}, S.Evs + (long)(3 * m + 1) * ni, 1, j);
      const double us = dv(std::bool_constant<IZ>{}, S.Evs + (long)(3 * m + 2) * ni, 2, k);
      val = val + (uk + ui + us) * ajac;
    }
    if (krk == 0) {  // right-hand side only
      S.dus[cm] = val;
      continue;
    }
    if (krk == 1) S.di[qm] = dn + dt * val;
    else S.di[qm] = ra * dn + rb * (dold + dt * val);
  }
}

template <int CD>
__global__ __launch_bounds__(256) void k_species_rk(DevState S, int krk, double dt, int visc, int m0, int m1) {
  constexpr int R = CD == 1 ? 4 : 3;
  const long c = cell_index();
  if (c >= S.ni) return;
  int i, j, k;
  cell_ijk(S, c, i, j, k);
  const long q = hq(S, i, j, k);
  const bool ix = wave_all(central_cell(i, S.nx, R));
  const bool iyz = wave_all(central_cell(j, S.ny, R) && central_cell(k, S.nz, R));
  if (ix && iyz) species_rk_cell<CD, true, true, true>(S, c, q, i, j, k, krk, dt, visc, m0, m1);
  else if (iyz) species_rk_cell<CD, false, true, true>(S, c, q, i, j, k, krk, dt, visc, m0, m1);
  else species_rk_cell<CD, false, false, false>(S, c, q, i, j, k, krk, dt, visc, m0, m1);
}

// Rescaling_di (Comput_flow_variables.f90:101-121) and comput_T_P (:32-46)
// of one cell, given the new density and internal energy.
__device__ __forceinline__ void ms_prims(const DevState& S, long q, double dd, double Et, double& T, double& p) {
  double di[OCFD_NS_MAX];
  double d0 = 0.0;
  for (int m = 0; m < S.ns; m++) {
    double x = S.di[q + (long)m * S.nh];
    if (x < 0.0) x = 0.0;
    di[m] = x;
    d0 = d0 + x;
  }
  const double tmp = dd / d0;
  for (int m = 0; m < S.ns; m++) {
    di[m] = di[m] * tmp;
    S.di[q + (long)m * S.nh] = di[m];
  }
  T = gas_T(di, Et);
  p = gas_P(di, T);
}

// comput_T_P over the interior (after the chemistry substeps, OCFD_NS_Solver.f90:34)
__global__ __launch_bounds__(256) void k_comput_T_P(DevState S) {
  const long c = cell_index();
  if (c >= S.ni) return;
  int i, j, k;
  cell_ijk(S, c, i, j, k);
  const long q = hq(S, i, j, k);
  double di[OCFD_NS_MAX];
  for (int m = 0; m < S.ns; m++) di[m] = S.di[q + (long)m * S.nh];
  const double T = gas_T(di, S.Et[q]);
  S.T[q] = T;
  S.p[q] = gas_P(di, T);
}

// Face re-synthesis (OCFD_BC.f90:33-56) for N_SPEC > 1
__device__ __forceinline__ void face_update_ms(const DevState& S, long q, long c) {
  double di[OCFD_NS_MAX];
  double d0 = 0.0;
  for (int m = 0; m < S.ns; m++) {
    di[m] = S.di[q + (long)m * S.nh];
    d0 = d0 + di[m];
  }
  const double dtmp = S.d[q] / d0;
  for (int m = 0; m < S.ns; m++) {
    di[m] = di[m] * dtmp;
    S.di[q + (long)m * S.nh] = di[m];
  }
  const double T = S.T[q];
  const double Et = gas_E(di, T);
  const double p = gas_P(di, T);
  S.Et[q] = Et;
  S.p[q] = p;
  const double dd = S.d[q], uu = S.u[q], vv = S.v[q], ww = S.w[q];
  S.f[c] = dd;
  S.f[c + S.ni] = dd * uu;
  S.f[c + 2 * S.ni] = dd * vv;
  S.f[c + 3 * S.ni] = dd * ww;
  S.f[c + 4 * S.ni] = Et + dd * (uu * uu + vv * vv + ww * ww) * 0.5;
}

// ocfd_bc_jet inflow (OCFD_BC_jet.f90:22-73): Dirichlet data, or (BC_Para(1)=1)
// the pressure of i=2 where the inflow is subsonic
__global__ void k_jet_inlet(DevState S) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)S.ny * S.nz) return;
  const int j = (int)(t % S.ny) + 1, k = (int)(t / S.ny) + 1;
  const long q = hq(S, 1, j, k), pl = (long)S.ny * S.nz, r = t;
  const double* in = S.jet_in;
  const int ns = S.ns;
  if (S.jet_iflag == 0) {
    S.d[q] = in[r]; S.u[q] = in[r + pl]; S.v[q] = in[r + 2 * pl]; S.w[q] = in[r + 3 * pl]; S.T[q] = in[r + 4 * pl];
    for (int m = 0; m < ns; m++) S.di[q + (long)m * S.nh] = in[r + (5L + m) * pl];
    return;
  }
  S.u[q] = in[r + pl]; S.v[q] = in[r + 2 * pl]; S.w[q] = in[r + 3 * pl]; S.T[q] = in[r + 4 * pl];
  const double d1 = in[r];
  double di1[OCFD_NS_MAX], ai1[OCFD_NS_MAX];
  double d0 = 0.0;
  for (int m = 0; m < ns; m++) {
    di1[m] = in[r + (5L + m) * pl];
    d0 = d0 + di1[m];
  }
  for (int m = 0; m < ns; m++) ai1[m] = di1[m] / d0;
  const double T1 = S.T[q];
  const double p1 = gas_P(di1, T1);
  const double Et1 = gas_E(di1, T1);
  const double gamma = p1 / Et1 + 1.0;  // comput_soundspeed, CHEM_thermdyn.f90:412-418
  const double c1 = sqrt(gamma * p1 / d1);
  if (S.u[q] < c1) {
    S.p[q] = S.p[hq(S, 2, j, k)];
    const double d1new = gas_d(ai1, T1, S.p[q]);
    S.d[q] = d1new;
    for (int m = 0; m < ns; m++) S.di[q + (long)m * S.nh] = d1new * ai1[m];
  } else {
    S.d[q] = d1;
    for (int m = 0; m < ns; m++) S.di[q + (long)m * S.nh] = di1[m];
  }
}

// ocfd_bc_jet outflow (OCFD_BC_jet.f90:76-93): zeroth-order extrapolation
__global__ void k_jet_outlet(DevState S) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)S.ny * S.nz) return;
  const int j = (int)(t % S.ny) + 1, k = (int)(t / S.ny) + 1;
  const long q = hq(S, S.nx, j, k), qa = hq(S, S.nx - 1, j, k);
  S.d[q] = S.d[qa]; S.u[q] = S.u[qa]; S.v[q] = S.v[qa]; S.w[q] = S.w[qa];
  S.T[q] = S.T[qa]; S.p[q] = S.p[qa]; S.Et[q] = S.Et[qa];
  for (int m = 0; m < S.ns; m++) S.di[q + (long)m * S.nh] = S.di[qa + (long)m * S.nh];
}

// Chemistry coupling (Chemical_time_advance, OCFD_Time_Adv.f90:169-196):
// interior di, Et -> batch [ns][interior], E, dt ...
__global__ __launch_bounds__(256) void k_chem_pack(DevState S, double* rho, double* E, double* dtb, double dt2) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ni) return;
  int i, j, k;
  cell_ijk(S, c, i, j, k);
  const long q = hq(S, i, j, k);
  for (int m = 0; m < S.ns; m++) rho[(long)m * S.ni + c] = S.di[q + (long)m * S.nh];
  E[c] = S.Et[q];
  dtb[c] = dt2;
}

// ... and back: di = batch, f5 += Et1 - Et0, Et = Et1
__global__ __launch_bounds__(256) void k_chem_unpack(DevState S, const double* rho, const double* E) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ni) return;
  int i, j, k;
  cell_ijk(S, c, i, j, k);
  const long q = hq(S, i, j, k);
  for (int m = 0; m < S.ns; m++) S.di[q + (long)m * S.nh] = rho[(long)m * S.ni + c];
  const double Et1 = E[c], Et0 = S.Et[q];
  S.f[c + 4 * S.ni] = S.f[c + 4 * S.ni] + (Et1 - Et0);
  S.Et[q] = Et1;
}
