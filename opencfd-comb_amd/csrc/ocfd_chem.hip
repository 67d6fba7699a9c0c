// ocfd_chem.hip -- MI355X kernels for OpenCFD-Comb's finite-rate chemistry:
// the per-cell 0-D integrators of chemical_timeAdv (CHEM_reactions.f90:374-602)
// over a batch of independent cells, FP64 throughout.
//
//   * k_chem_fixed<MODE>: the reference integrator, nsub equal substeps per
//     cell, one cell per thread (MODE = CHEM_TimeAdv: 1 Euler, 2 RK2, 3 RK3,
//     -1 point-implicit Euler).
//   * k_chem_adaptive: point-implicit Euler with the adaptive per-cell substep
//     controller, one cell per thread.
//   * k_chem_queue: the same integrator in a persistent kernel whose lanes
//     pull cells from a device work queue, so lanes that finish early take
//     new cells instead of idling while the slowest lane of the wavefront
//     works through hundreds of substeps.
//
// Layout: rho is species-major [ns][n] (coalesced per species), E and dt [n].
// The mechanism lives in constant memory.  Arithmetic keeps the reference's
// operation order (-ffp-contract=off); exp/log/log10/pow are the device
// library's, so results agree with the CPU restatement to a few ulps per call.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/ocfd_hip.h"
#include "ocfd_mlp.h"
#include "ocfd_mlp_tile.h"

#define CH_NS 9
#define CH_NR_MAX 32
#define CH_NT_MAX 4
#define CH_R0 8.314
#define CH_ATM 1.013e5

struct ChemMech {
  int ns, nr, n_tct;
  int rereac;                      // 0 Gibbs Kc, 1 Arrhenius reverse rates, 2 Park fit of Kc
  double rev[CH_NR_MAX][5];        // ReReac 1: A_r, b_r, E_r; ReReac 2: A1r..A5r
  int af_type[CH_NR_MAX], tb[CH_NR_MAX], sgm[CH_NR_MAX];
  double fc[CH_NR_MAX], arr[CH_NR_MAX][2][3], eff[CH_NR_MAX][CH_NS];
  int vf[CH_NR_MAX][CH_NS], vr[CH_NR_MAX][CH_NS];
  double tct[CH_NT_MAX], tct_max;
  double mi[CH_NS], ri[CH_NS], coef[CH_NS][CH_NT_MAX][7], ect[CH_NS][CH_NT_MAX];
};

__constant__ ChemMech cm;

// The mechanism as seen by the device functions: a pointer into the constant
// bank (address space 4, so loads through it are scalar loads).  The kernels
// pass it through an empty asm statement at the top of every substep; the
// compiler then cannot hoist the hundreds of constant loads of the rate and
// thermo code out of the substep loop into registers, which otherwise
// exhausts the register file and spills.
typedef const __attribute__((address_space(4))) ChemMech* Mech;
__device__ __forceinline__ Mech mech_opaque(Mech m) {
  asm volatile("" : "+s"(m));
  return m;
}

// Integration state of a cell handed from the per-lane queue to the
// cooperative kernel (struct of arrays, capacity cap records).
struct HeavyBuf {
  double* di;  // [CH_NS][cap]
  double *E, *t, *h_cap;
  int *nsub, *nrej, *cell;
  long cap;
};

namespace {
struct ChemCtx {
  bool ready = false;
  bool rev_ready = false;  // reverse-rate data present (always for ReReac = 0)
  ChemMech h{};
  hipStream_t stream = nullptr, own = nullptr;
  unsigned long long* counter = nullptr;  // [0] queue head, [1] unfinished, [2] surrogate fallbacks
  int* fb_list = nullptr;                  // cells handed to the fallback integrator
  unsigned char* surr_buf = nullptr;       // surrogate phases: X[cap][16] (bf16), O[9][cap] (float), ok[cap]
  tbf16* surr_wpack = nullptr;             // network weights in MFMA fragment order
  long surr_wpack_cap = 0;
  const tbf16* surr_wp[MLP_MAX_LAYERS] = {};  // per layer, into surr_wpack
  unsigned long long surr_wp_version = ~0ull;  // network weights version packed
  long fb_cap = 0;
  HeavyBuf heavy{};                        // cells handed to the cooperative kernel
  int fallback_scheduler = 2;              // scheduler of the surrogate's fallback
  int coop_direct_max = 16384;             // fallback lists up to this size go to the cooperative kernel
  int coop_budget = 128;                   // attempts on one lane before the cooperative kernel
  unsigned char* heavy_mem = nullptr;
  // longest-first order of a batch: cell indices sorted by decreasing dt
  int order = 1;                           // 0: index order, 1: longest dt first
  int *ord_in = nullptr, *ord_out = nullptr;
  unsigned int *ord_keys = nullptr, *ord_keys_in = nullptr;
  unsigned long long* ord_count = nullptr;
  void* ord_temp = nullptr;
  size_t ord_temp_bytes = 0;
  long ord_cap = 0;
  unsigned char* ord_mem = nullptr;
  bool surr_ready = false;
  int surr_variant = 0;                    // tile shape of k_surr_mlp (SurrTile)
  float surr_tune_ms[8] = {};
  std::string err;
};
ChemCtx CX;
int cfail(const std::string& m) {
  CX.err = m;
  return 1;
}
}  // namespace

__device__ __forceinline__ double q2(double x) { return x * x; }
__device__ __forceinline__ double q3(double x) { double x2 = x * x; return x2 * x; }
__device__ __forceinline__ double q4(double x) { double x2 = x * x; return x2 * x2; }
__device__ __forceinline__ double q5(double x) { double x2 = x * x; return (x2 * x2) * x; }
// Fortran x**n for an integer n (repeated squaring; negative n -> 1/x**|n|)
__device__ __forceinline__ double ipow(double x, int n) {
  if (n == 0) return 1.0;
  int m = n < 0 ? -n : n;
  double r = 1.0, b = x;
  bool first = true;
  while (m) {
    if (m & 1) { r = first ? b : r * b; first = false; }
    m >>= 1;
    if (m) b = b * b;
  }
  return n < 0 ? 1.0 / r : r;
}

__device__ __forceinline__ int trange(Mech m, double T) {
  int jc = 0;
  for (int j = 0; j < m->n_tct - 1; j++)
    if (T > m->tct[j]) jc = j + 1;
  return jc;
}

// comput_T_rgas, CHEM_thermdyn.f90:326-382 (Newton, 1e-3 K, <= 31 iterations)
__device__ double chem_T(Mech m, const double* di, double E) {
  double A1[CH_NT_MAX], B1[CH_NT_MAX], C1[CH_NT_MAX], D1[CH_NT_MAX], E1[CH_NT_MAX], Ect[CH_NT_MAX];
#pragma unroll
  for (int j = 0; j < CH_NT_MAX; j++) {
    A1[j] = B1[j] = C1[j] = D1[j] = E1[j] = Ect[j] = 0.0;
    if (j < m->n_tct) {
#pragma unroll
      for (int k = 0; k < CH_NS; k++) {
        const double R1 = m->ri[k];
        A1[j] = A1[j] + (m->coef[k][j][0] - 1.0) * R1 * di[k];
        B1[j] = B1[j] + m->coef[k][j][1] * R1 * di[k];
        C1[j] = C1[j] + m->coef[k][j][2] * R1 * di[k];
        D1[j] = D1[j] + m->coef[k][j][3] * R1 * di[k];
        E1[j] = E1[j] + m->coef[k][j][4] * R1 * di[k];
        Ect[j] = Ect[j] + m->ect[k][j] * R1 * di[k];
      }
    }
  }
  double T1 = E / A1[0], T2 = T1;
  for (int ks = 0; ks <= 30; ks++) {
    const int jc = trange(m, T1);
    double a = A1[0], b = B1[0], c = C1[0], d = D1[0], e = E1[0], ect = Ect[0];
#pragma unroll
    for (int j = 1; j < CH_NT_MAX; j++)
      if (jc == j) { a = A1[j]; b = B1[j]; c = C1[j]; d = D1[j]; e = E1[j]; ect = Ect[j]; }
    double Et1, Ex;
    if (T1 <= m->tct_max) {
      Et1 = a * T1 + b * q2(T1) / 2.0 + c * q3(T1) / 3.0 + d * q4(T1) / 4.0 + e * q5(T1) / 5.0 + ect;
      Ex = a + b * T1 + c * q2(T1) + d * q3(T1) + e * q4(T1);
    } else {
      const double tm = m->tct_max;
      Ex = a + b * tm + c * q2(tm) + d * q3(tm) + e * q4(tm);
      Et1 = a * tm + b * q2(tm) / 2.0 + c * q3(tm) / 3.0 + d * q4(tm) / 4.0 + e * q5(tm) / 5.0 + ect + Ex * (T1 - tm);
    }
    T2 = T1 - (Et1 - E) / Ex;
    const double error = fabs(T2 - T1);
    T1 = T2;
    if (error < 1.0e-3) break;
  }
  return T2;
}

// comput_gibbs, CHEM_thermdyn.f90:296-319
__device__ __forceinline__ double chem_gibbs(Mech m, int k, double T, double logT) {
  const int jc = trange(m, T);
  const auto c = m->coef[k][jc];  // pointer into the constant bank
  if (T <= m->tct_max)
    return c[0] * (1.0 - logT) * T - c[1] * q2(T) / 2.0 - c[2] * q3(T) / 6.0 - c[3] * q4(T) / 12.0 - c[4] * q5(T) / 20.0 +
           c[5] - c[6] * T;
  const double tm = m->tct_max;
  const double Cp_max = c[0] + c[1] * tm + c[2] * q2(tm) + c[3] * q3(tm) + c[4] * q4(tm);
  const double h_max = c[0] * tm + c[1] * q2(tm) / 2.0 + c[2] * q3(tm) / 3.0 + c[3] * q4(tm) / 4.0 + c[4] * q5(tm) / 5.0 + c[5];
  const double S_max = c[0] * log(tm) + c[1] * tm + c[2] * q2(tm) / 2.0 + c[3] * q3(tm) / 3.0 + c[4] * q4(tm) / 4.0 + c[6];
  return h_max + Cp_max * (T - tm) - T * (S_max + Cp_max * log(T / tm));
}

// Reaction_rate_coef (CHEM_reactions.f90:130-197) + Reaction_rate_and_Matrix
// (:230-272).  ri in mol/cm3/s; A[i][j] = d ri / d cj if JAC.
// Constants of one reaction (read_Reac, CHEM_reactions.f90:15-71).
struct RxnConst {
  int af_type, tb, sgm, rereac;
  double fc, a0, b0, e0, a1, b1, e1;
  double r[5];                     // reverse-rate data (ReReac 1 or 2)
  double eff[CH_NS];
  int vf[CH_NS], vr[CH_NS];
};

__device__ __forceinline__ RxnConst load_rxn(Mech m, int k) {
  RxnConst R;
  R.af_type = m->af_type[k];
  R.tb = m->tb[k];
  R.sgm = m->sgm[k];
  R.fc = m->fc[k];
  R.a0 = m->arr[k][0][0];
  R.b0 = m->arr[k][0][1];
  R.e0 = m->arr[k][0][2];
  R.a1 = m->arr[k][1][0];
  R.b1 = m->arr[k][1][1];
  R.e1 = m->arr[k][1][2];
  R.rereac = m->rereac;
#pragma unroll
  for (int a = 0; a < 5; a++) R.r[a] = m->rev[k][a];
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    R.eff[i] = m->eff[k][i];
    R.vf[i] = m->vf[k][i];
    R.vr[i] = m->vr[k][i];
  }
  return R;
}

// Rate of one reaction with concentrations: Reaction_rate_coef
// (CHEM_reactions.f90:130-197) and the mass-action products of
// Reaction_rate_and_Matrix (:246-250).  gi = G_i/R_i of every species.
__device__ __forceinline__ void chem_reaction_c(const RxnConst& R, double T, const double* gi, const double* ci,
                                                double& wf1, double& wr1) {
  const double Rs = CH_R0 / CH_ATM * 1.0e6;
  // comput_KjX (CHEM_reactions.f90:76-103), used by ReReac = 0 only
  double kjx = 0.0;
  if (R.rereac == 0) {
    double dgi = 0.0;
#pragma unroll
    for (int i = 0; i < CH_NS; i++) dgi = dgi + gi[i] * (R.vr[i] - R.vf[i]);
    kjx = exp(-dgi / T) * ipow(Rs * T, -R.sgm);
  }
  double Ctrb = 0.0;
  if (R.tb == 1) {
#pragma unroll
    for (int i = 0; i < CH_NS; i++) Ctrb = Ctrb + ci[i] * R.eff[i];
  }
  const double wf1a = R.a0 * pow(T, R.b0) * exp(-R.e0 / T);
  double wf, wr;
  if (R.af_type == 0) {
    wf = wf1a;
  } else if (R.af_type == 2) {
    const double wf2 = R.a1 * pow(T, R.b1) * exp(-R.e1 / T);
    wf = wf1a + wf2;
  } else {
    const double wf2 = R.a1 * pow(T, R.b1) * exp(-R.e1 / T);
    const double Prs = wf1a / wf2 * Ctrb;
    const double Fc = R.fc;
    const double lfc = log10(Fc);
    const double ac = -0.4 - (double)0.67f * lfc;   // REAL(4) literals of the reference
    const double an = 0.75 - (double)1.27f * lfc;
    const double lp = log10(Prs);
    const double F = pow(10.0, lfc / (1.0 + q2((lp + ac) / (an - 0.14 * (lp + ac)))));
    wf = wf2 * Prs / (1.0 + Prs) * F;
  }
  if (R.rereac == 0) {
    wr = wf / kjx;
  } else if (R.rereac == 1) {
    wr = R.r[0] * pow(T, R.r[1]) * exp(-R.r[2] / T);      // Arrhenius reverse rate (:180-181)
  } else {
    // ParkCurveFit (:113-118) with the 1e-30 floor (:182-186)
    const double z = 10000.0 / T;
    double kp = exp(R.r[0] + R.r[1] * z + R.r[2] * q2(z) + R.r[3] * q3(z) + R.r[4] * q4(z));
    if (kp < 1.0e-30) kp = 1.0e-30;
    wr = wf / kp;
  }
  if (R.tb == 1 && R.af_type != 1) {
    wf = wf * Ctrb;
    wr = wr * Ctrb;
  }
  wf1 = wf;
  wr1 = wr;
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    if (R.vf[i] != 0) wf1 = wf1 * ipow(ci[i], R.vf[i]);
    if (R.vr[i] != 0) wr1 = wr1 * ipow(ci[i], R.vr[i]);
  }
}

__device__ __forceinline__ void chem_reaction(Mech m, int k, double T, const double* gi, const double* ci, double& wf1,
                                              double& wr1) {
  chem_reaction_c(load_rxn(m, k), T, gi, ci, wf1, wr1);
}

// d(ri)/d(cj) contribution of reaction k (CHEM_reactions.f90:262-268)
__device__ __forceinline__ double chem_dA(Mech m, int k, int j, double wf1, double wr1, const double* ci) {
  const double Awf = m->vf[k][j] * wf1 / (ci[j] + 1.0e-20);
  const double Awr = m->vr[k][j] * wr1 / (ci[j] + 1.0e-20);
  return Awf - Awr;
}

// Reaction_rate_coef (CHEM_reactions.f90:130-197) + Reaction_rate_and_Matrix
// (:230-272).  ri in mol/cm3/s; A[i][j] = d ri / d cj if JAC.  k is uniform
// across the wavefront, so the stoichiometry tests are scalar branches and
// the fully unrolled i/j loops keep ri and A in registers.
template <bool JAC>
__device__ void chem_rates(Mech m, const double* ci, double T, double* ri, double (*A)[CH_NS]) {
  double gi[CH_NS];
  if (m->rereac == 0) {
    const double logT = log(T);
#pragma unroll
    for (int i = 0; i < CH_NS; i++) gi[i] = chem_gibbs(m, i, T, logT);
  }
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    ri[i] = 0.0;
    if (JAC) {
#pragma unroll
      for (int j = 0; j < CH_NS; j++) A[i][j] = 0.0;
    }
  }
  for (int k = 0; k < m->nr; k++) {
    double wf1, wr1;
    chem_reaction(m, k, T, gi, ci, wf1, wr1);
#pragma unroll
    for (int i = 0; i < CH_NS; i++) {
      const int nu = m->vr[k][i] - m->vf[k][i];
      if (nu != 0) ri[i] = ri[i] + (wf1 - wr1) * nu;
    }
    if (JAC) {
#pragma unroll
      for (int j = 0; j < CH_NS; j++) {
        if (m->vf[k][j] == 0 && m->vr[k][j] == 0) continue;
        const double dA = chem_dA(m, k, j, wf1, wr1, ci);
#pragma unroll
        for (int i = 0; i < CH_NS; i++) {
          const int nu = m->vr[k][i] - m->vf[k][i];
          if (nu != 0) A[i][j] = A[i][j] + dA * nu;
        }
      }
    }
  }
}

// Gauss elimination with partial pivoting (CHEM_reactions.f90:523-567),
// fully unrolled so the 9x9 system stays in registers; row swaps are selects.
__device__ void chem_gauss(double (*a)[CH_NS], double* b, double* x) {
#pragma unroll
  for (int k = 0; k < CH_NS - 1; k++) {
    double fmax = fabs(a[k][k]);
    int kk = k;
#pragma unroll
    for (int i = k + 1; i < CH_NS; i++)
      if (fabs(a[i][k]) > fmax) { fmax = fabs(a[i][k]); kk = i; }
#pragma unroll
    for (int i = k + 1; i < CH_NS; i++) {
      const bool sw = (i == kk);
#pragma unroll
      for (int j = k; j < CH_NS; j++) {
        const double t = a[k][j];
        a[k][j] = sw ? a[i][j] : a[k][j];
        a[i][j] = sw ? t : a[i][j];
      }
      const double t = b[k];
      b[k] = sw ? b[i] : b[k];
      b[i] = sw ? t : b[i];
    }
#pragma unroll
    for (int i = k + 1; i < CH_NS; i++) {
      const double f = -a[i][k] / a[k][k];
#pragma unroll
      for (int j = 0; j < CH_NS; j++) a[i][j] = a[i][j] + f * a[k][j];
      b[i] = b[i] + f * b[k];
    }
  }
  x[CH_NS - 1] = b[CH_NS - 1] / a[CH_NS - 1][CH_NS - 1];
#pragma unroll
  for (int i = CH_NS - 2; i >= 0; i--) {
    double xi = b[i];
#pragma unroll
    for (int j = i + 1; j < CH_NS; j++) xi = xi - a[i][j] * x[j];
    x[i] = xi / a[i][i];
  }
}

// update_Et, CHEM_thermdyn.f90:184-196
__device__ __forceinline__ void chem_update_et(Mech m, const double* di, const double* dinew, double& E) {
#pragma unroll
  for (int i = 0; i < CH_NS; i++) E = E - (dinew[i] - di[i]) * m->ri[i] * m->coef[i][0][5];
}

// The linear system of chemical_Euler_implicit (CHEM_reactions.f90:491-505)
// with step h, solved in place: Ar becomes I - Ar (M_i/M_j) h and is
// eliminated; dd = the increment of the partial densities.
__device__ __forceinline__ void chem_implicit_solve(Mech m, const double* ri, double (*Ar)[CH_NS], double h, double* dd) {
  double rhs[CH_NS];
#pragma unroll
  for (int j = 0; j < CH_NS; j++)
#pragma unroll
    for (int i = 0; i < CH_NS; i++) Ar[i][j] = (i == j ? 1.0 : 0.0) - Ar[i][j] * m->mi[i] / m->mi[j] * h;
#pragma unroll
  for (int i = 0; i < CH_NS; i++) rhs[i] = ri[i] * m->mi[i] * 1.0e6 * h;
  chem_gauss(Ar, rhs, dd);
}

__device__ __forceinline__ void chem_apply(Mech m, double* di, double& E, const double* dd) {
  double dinew[CH_NS];
#pragma unroll
  for (int i = 0; i < CH_NS; i++) dinew[i] = di[i] + dd[i];
  chem_update_et(m, di, dinew, E);
#pragma unroll
  for (int i = 0; i < CH_NS; i++) di[i] = dinew[i];
}

// chemical_Euler_implicit, CHEM_reactions.f90:473-516
__device__ void chem_step_implicit(Mech m, double* di, double& E, double dt) {
  double ci[CH_NS], ri[CH_NS], Ar[CH_NS][CH_NS];
  const double T = chem_T(m, di, E);
#pragma unroll
  for (int i = 0; i < CH_NS; i++) ci[i] = di[i] / m->mi[i] * 1.0e-6;
  chem_rates<true>(m, ci, T, ri, Ar);
  double dd[CH_NS];
  chem_implicit_solve(m, ri, Ar, dt, dd);
  chem_apply(m, di, E, dd);
}

__device__ void chem_explicit_rates(Mech m, const double* di_c, double T, double* ri) {
  double ci[CH_NS];
#pragma unroll
  for (int i = 0; i < CH_NS; i++) ci[i] = di_c[i] / m->mi[i] * 1.0e-6;
  chem_rates<false>(m, ci, T, ri, nullptr);
}

// chemical_Euler (:395-425), chemical_RK2 (:429-469), chemical_RK3 (:572-602)
template <int MODE>
__device__ void chem_step(Mech m, double* di, double& E, double dt) {
  if (MODE == -1) {
    chem_step_implicit(m, di, E, dt);
  } else if (MODE == 1) {
    double ri[CH_NS], dinew[CH_NS];
    const double T = chem_T(m, di, E);
    chem_explicit_rates(m, di, T, ri);
#pragma unroll
    for (int i = 0; i < CH_NS; i++) dinew[i] = di[i] + ri[i] * m->mi[i] * 1.0e6 * dt;
    chem_update_et(m, di, dinew, E);
#pragma unroll
    for (int i = 0; i < CH_NS; i++) di[i] = dinew[i];
  } else if (MODE == 2) {
    double ri[CH_NS], di1[CH_NS], dinew[CH_NS];
    double T = chem_T(m, di, E);
    chem_explicit_rates(m, di, T, ri);
#pragma unroll
    for (int i = 0; i < CH_NS; i++) di1[i] = di[i] + ri[i] * m->mi[i] * 1.0e6 * dt;
    chem_update_et(m, di, di1, E);
    T = chem_T(m, di, E);  // the reference evaluates T from the old di here
    chem_explicit_rates(m, di1, T, ri);
#pragma unroll
    for (int i = 0; i < CH_NS; i++) dinew[i] = 0.5 * di[i] + 0.5 * di1[i] + 0.5 * ri[i] * m->mi[i] * 1.0e6 * dt;
    chem_update_et(m, di1, dinew, E);
#pragma unroll
    for (int i = 0; i < CH_NS; i++) di[i] = dinew[i];
  } else {
    const double Ra[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0}, Rb[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
    double din[CH_NS], ri[CH_NS], dinew[CH_NS];
#pragma unroll
    for (int i = 0; i < CH_NS; i++) din[i] = di[i];
    for (int krk = 0; krk < 3; krk++) {
      const double T = chem_T(m, di, E);
      chem_explicit_rates(m, di, T, ri);
#pragma unroll
      for (int i = 0; i < CH_NS; i++) dinew[i] = Ra[krk] * din[i] + Rb[krk] * (di[i] + ri[i] * m->mi[i] * 1.0e6 * dt);
      chem_update_et(m, di, dinew, E);
#pragma unroll
      for (int i = 0; i < CH_NS; i++) di[i] = dinew[i];
    }
  }
}

template <int MODE>
__global__ __launch_bounds__(128) void k_chem_fixed(long n, double* rho, double* Ecell, const double* dtcell, int nsub) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double di[CH_NS];
#pragma unroll
  for (int i = 0; i < CH_NS; i++) di[i] = rho[(long)i * n + c];
  double E = Ecell[c];
  const double h = dtcell[c] / nsub;
  Mech m = (Mech)&cm;
  for (int s = 0; s < nsub; s++) {
    m = mech_opaque(m);
    chem_step<MODE>(m, di, E, h);
  }
#pragma unroll
  for (int i = 0; i < CH_NS; i++) rho[(long)i * n + c] = di[i];
  Ecell[c] = E;
}

// Adaptive point-implicit substepping (no reference counterpart).  One
// attempt: T, rates and Jacobian at the current state (negative partial
// densities enter the rates as zero: the reference's regularised Jacobian
// d(ri)/d(cj) ~ 1/(cj + 1e-20) makes cj = -1e-20 mol/cm3 a fixed point of the
// point-implicit update of a consumed species, where it divides by zero);
// proposed step
//   h = min(dt - t, h_cap, min_k s_k/|w_k|, s_T/|q|)
// with s_k = rtol_y (max(rho_k,0) + y_floor rho), s_T = rtol_t T rcv,
// rcv = sum_k 2.5 R_k rho_k, w_k = ri_k M_k 1e6 and q = sum_k w_k R_k F_k;
// the linear solve gives dd, accepted when |dd_k| <= 2 s_k for all k and
// |sum_k dd_k R_k F_k| <= 2 s_T.  Accepted: the update is applied and
// h_cap = 2h.  Rejected: the state is unchanged and h_cap = h/2.  Returns
// whether the attempt was accepted; h_taken is the step on acceptance.
struct ChemCtl {
  double rtol_y, y_floor, rtol_t;
  int max_sub;
};

__device__ __forceinline__ bool chem_adaptive_attempt(Mech m, double* di, double& E, double dt, double t, double& h_cap,
                                                      const ChemCtl& ctl, double& h_taken) {
  double ci[CH_NS], ri[CH_NS], Ar[CH_NS][CH_NS], sk[CH_NS], dd[CH_NS];
  const double T = chem_T(m, di, E);
  double rho = 0.0;
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    ci[i] = (di[i] > 0.0 ? di[i] : 0.0) / m->mi[i] * 1.0e-6;
    rho = rho + di[i];
  }
  chem_rates<true>(m, ci, T, ri, Ar);
  double h = dt - t;
  if (h_cap < h) h = h_cap;
  double q = 0.0, rcv = 0.0;
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    const double w = ri[i] * m->mi[i] * 1.0e6;
    sk[i] = ctl.rtol_y * ((di[i] > 0.0 ? di[i] : 0.0) + ctl.y_floor * rho);
    if (fabs(w) * h > sk[i]) h = sk[i] / fabs(w);
    q = q + w * m->ri[i] * m->coef[i][0][5];
    rcv = rcv + 2.5 * m->ri[i] * di[i];
  }
  const double sT = ctl.rtol_t * T * rcv;
  if (fabs(q) * h > sT) h = sT / fabs(q);
  chem_implicit_solve(m, ri, Ar, h, dd);
  bool ok = true;
  double de = 0.0;
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    ok = ok && fabs(dd[i]) <= 2.0 * sk[i];
    de = de + dd[i] * m->ri[i] * m->coef[i][0][5];
  }
  ok = ok && fabs(de) <= 2.0 * sT;
  if (!ok) {
    h_cap = 0.5 * h;
    return false;
  }
  chem_apply(m, di, E, dd);
  h_cap = 2.0 * h;
  h_taken = h;
  return true;
}

// queue[0]: next cell of the work queue; queue[1]: cells left unfinished
// (budget of max_sub accepted and max_sub rejected attempts spent before dt)
__device__ __forceinline__ void chem_finish(long c, long n, double* rho, double* Ecell, const double* di, double E,
                                            int nsub, bool finished, int* nsub_out, unsigned long long* queue) {
#pragma unroll
  for (int i = 0; i < CH_NS; i++) rho[(long)i * n + c] = di[i];
  Ecell[c] = E;
  if (nsub_out) nsub_out[c] = nsub;
  if (!finished) atomicAdd(queue + 1, 1ull);
}

__global__ __launch_bounds__(128) void k_chem_adaptive(long n, double* rho, double* Ecell, const double* dtcell,
                                                        ChemCtl ctl, int* nsub_out, unsigned long long* queue) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double di[CH_NS];
#pragma unroll
  for (int i = 0; i < CH_NS; i++) di[i] = rho[(long)i * n + c];
  double E = Ecell[c];
  const double dt = dtcell[c];
  double t = 0.0, h_cap = dt, h = 0.0;
  int nsub = 0, nrej = 0;
  Mech m = (Mech)&cm;
  while (t < dt && nsub < ctl.max_sub && nrej < ctl.max_sub) {
    m = mech_opaque(m);
    if (!chem_adaptive_attempt(m, di, E, dt, t, h_cap, ctl, h)) {
      nrej++;
      continue;
    }
    t = t + h;
    nsub++;
    if (dt - t <= 1e-12 * dt) break;
  }
  chem_finish(c, n, rho, Ecell, di, E, nsub, t >= dt || dt - t <= 1e-12 * dt, nsub_out, queue);
}

// Persistent kernel with a lane-level work queue.  A lane whose cell is done
// takes the next one; the wavefront asks the queue once for all its idle
// lanes (one atomic per refill).  Each pass runs one attempt per busy lane,
// with the same per-cell loop as k_chem_adaptive.  A cell ends after at most
// 2 max_sub attempts, and a lane leaves once the queue is empty and its cell
// is done, so every wave reaches the exit.
// With a list, the queue runs over the cells list[0 .. *list_count) instead
// of 0 .. n (n stays the array stride).
// With budget > 0, a cell still running after `budget` attempts is handed
// to the cooperative kernel (k_chem_coop) with its full controller state, so
// the few cells that need hundreds of substeps do not hold a lane for the
// whole kernel.  A list of at most direct_max cells (the surrogate's
// rejects: few, and mostly stiff) hands every cell over after its first
// attempt: each then has a wavefront to itself.
__global__ __launch_bounds__(128) void k_chem_queue(long n, double* rho, double* Ecell, const double* dtcell,
                                                     ChemCtl ctl, int* nsub_out, unsigned long long* queue,
                                                     const int* list, const unsigned long long* list_count, int budget,
                                                     int direct_max, HeavyBuf hb) {
  const unsigned long long count =
      list ? (*list_count < (unsigned long long)n ? *list_count : (unsigned long long)n) : (unsigned long long)n;
  if (list && budget > 0 && count <= (unsigned long long)direct_max) budget = 1;
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = (1ull << lane) - 1ull;
  long cell = -1;  // -1: needs work, -2: queue exhausted
  double di[CH_NS];
  double E = 0.0, dt = 0.0, t = 0.0, h_cap = 0.0;
  int nsub = 0, nrej = 0;
  Mech m = (Mech)&cm;
  while (true) {
    m = mech_opaque(m);
    const bool need = (cell == -1);
    const unsigned long long mask = __ballot(need);
    if (mask) {
      const int leader = __ffsll((long long)mask) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(queue, (unsigned long long)__popcll(mask));
      base = __shfl(base, leader);
      if (need) {
        const unsigned long long my = base + (unsigned long long)__popcll(mask & lt);
        if (my < count) {
          cell = list ? (long)list[my] : (long)my;
          if (cell < 0 || cell >= n) cell = -1;  // a list entry outside the batch is skipped
        } else {
          cell = -2;
        }
        if (cell >= 0) {
#pragma unroll
          for (int i = 0; i < CH_NS; i++) di[i] = rho[(long)i * n + cell];
          E = Ecell[cell];
          dt = dtcell[cell];
          t = 0.0;
          h_cap = dt;
          nsub = 0;
          nrej = 0;
        }
      }
    }
    if (__ballot(cell != -2) == 0) break;  // every lane has found the queue empty
    if (cell >= 0) {
      bool done = !(t < dt && nsub < ctl.max_sub && nrej < ctl.max_sub);
      double h = 0.0;
      if (!done) {
        if (chem_adaptive_attempt(m, di, E, dt, t, h_cap, ctl, h)) {
          t = t + h;
          nsub++;
          done = dt - t <= 1e-12 * dt;
        } else {
          nrej++;
        }
        done = done || !(t < dt && nsub < ctl.max_sub && nrej < ctl.max_sub);
      }
      if (done) {
        chem_finish(cell, n, rho, Ecell, di, E, nsub, t >= dt || dt - t <= 1e-12 * dt, nsub_out, queue);
        cell = -1;
      } else if (budget > 0 && nsub + nrej >= budget) {
        const unsigned long long r = atomicAdd(queue + 3, 1ull);
#pragma unroll
        for (int i = 0; i < CH_NS; i++) hb.di[(long)i * hb.cap + r] = di[i];
        hb.E[r] = E;
        hb.t[r] = t;
        hb.h_cap[r] = h_cap;
        hb.nsub[r] = nsub;
        hb.nrej[r] = nrej;
        hb.cell[r] = (int)cell;
        cell = -1;
      }
    }
  }
}

// Cooperative integration of one cell per wavefront (64-thread workgroup).
// Every lane holds the same copy of the cell state; the parallel parts (the
// species' Gibbs functions, the reactions, ri, the Jacobian entries and the
// Gauss elimination) exchange values through LDS, and every value is computed
// with the same operations in the same order as chem_adaptive_attempt, so
// the results are bit-identical to the per-lane integrator.
struct CoopLds {
  double gi[CH_NS];
  double wf1[CH_NR_MAX], wr1[CH_NR_MAX];
  double ri[CH_NS];
  double A[CH_NS][CH_NS + 1];  // Jacobian, then the system; column CH_NS = right-hand side
  double f[CH_NS];
  double dA[CH_NR_MAX][CH_NS];  // d(rate of reaction k)/d(c_j), shared by every row i
  // mechanism tables read with lane-dependent indices (staged once)
  double coef[CH_NS][CH_NT_MAX][7];
  double mi[CH_NS];
  signed char vf[CH_NR_MAX][CH_NS], vr[CH_NR_MAX][CH_NS];
};

// comput_gibbs (CHEM_thermdyn.f90:296-319) with the coefficients from LDS
__device__ __forceinline__ double coop_gibbs(Mech m, const double (*coefk)[7], double T, double logT) {
  const int jc = trange(m, T);
  const double* c = coefk[jc];
  if (T <= m->tct_max)
    return c[0] * (1.0 - logT) * T - c[1] * q2(T) / 2.0 - c[2] * q3(T) / 6.0 - c[3] * q4(T) / 12.0 - c[4] * q5(T) / 20.0 +
           c[5] - c[6] * T;
  const double tm = m->tct_max;
  const double Cp_max = c[0] + c[1] * tm + c[2] * q2(tm) + c[3] * q3(tm) + c[4] * q4(tm);
  const double h_max = c[0] * tm + c[1] * q2(tm) / 2.0 + c[2] * q3(tm) / 3.0 + c[3] * q4(tm) / 4.0 + c[4] * q5(tm) / 5.0 + c[5];
  const double S_max = c[0] * log(tm) + c[1] * tm + c[2] * q2(tm) / 2.0 + c[3] * q3(tm) / 3.0 + c[4] * q4(tm) / 4.0 + c[6];
  return h_max + Cp_max * (T - tm) - T * (S_max + Cp_max * log(T / tm));
}

__device__ bool coop_attempt(Mech m, CoopLds& L, const RxnConst& R, int lane, double* di, double& E, double dt,
                             double t, double& h_cap, const ChemCtl& ctl, double& h_taken) {
  double ci[CH_NS], sk[CH_NS], dd[CH_NS];
  __syncthreads();
  const double T = chem_T(m, di, E);
  double rho = 0.0;
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    ci[i] = (di[i] > 0.0 ? di[i] : 0.0) / m->mi[i] * 1.0e-6;
    rho = rho + di[i];
  }
  if (lane < CH_NS) L.gi[lane] = coop_gibbs(m, L.coef[lane], T, log(T));
  __syncthreads();
  const int nr = m->nr;
  if (lane < nr) {
    double gi[CH_NS];
#pragma unroll
    for (int i = 0; i < CH_NS; i++) gi[i] = L.gi[i];
    double wf1, wr1;
    chem_reaction_c(R, T, gi, ci, wf1, wr1);
    L.wf1[lane] = wf1;
    L.wr1[lane] = wr1;
  }
  __syncthreads();
  if (lane < CH_NS) {
    double r = 0.0;
    for (int k = 0; k < nr; k++) {
      const int nu = L.vr[k][lane] - L.vf[k][lane];
      if (nu != 0) r = r + (L.wf1[k] - L.wr1[k]) * nu;
    }
    L.ri[lane] = r;
  }
  // the factor (Awf - Awr) of chem_dA once per (reaction, species) ...
  for (int e = lane; e < nr * CH_NS; e += 64) {
    const int k = e / CH_NS, j = e % CH_NS;
    const int vfj = L.vf[k][j], vrj = L.vr[k][j];
    double cjv = ci[0];  // ci[j] for the lane-dependent j, without dynamic register indexing
#pragma unroll
    for (int q = 1; q < CH_NS; q++)
      if (j == q) cjv = ci[q];
    const double Awf = vfj * L.wf1[k] / (cjv + 1.0e-20);
    const double Awr = vrj * L.wr1[k] / (cjv + 1.0e-20);
    L.dA[k][j] = Awf - Awr;
  }
  __syncthreads();
  // ... then every Jacobian entry sums its reactions in order
  for (int e = lane; e < CH_NS * CH_NS; e += 64) {
    const int i = e / CH_NS, j = e % CH_NS;
    double a = 0.0;
    for (int k = 0; k < nr; k++) {
      if (L.vf[k][j] == 0 && L.vr[k][j] == 0) continue;
      const int nu = L.vr[k][i] - L.vf[k][i];
      if (nu != 0) a = a + L.dA[k][j] * nu;
    }
    L.A[i][j] = a;
  }
  __syncthreads();
  double h = dt - t;
  if (h_cap < h) h = h_cap;
  double q = 0.0, rcv = 0.0;
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    const double w = L.ri[i] * m->mi[i] * 1.0e6;
    sk[i] = ctl.rtol_y * ((di[i] > 0.0 ? di[i] : 0.0) + ctl.y_floor * rho);
    if (fabs(w) * h > sk[i]) h = sk[i] / fabs(w);
    q = q + w * m->ri[i] * m->coef[i][0][5];
    rcv = rcv + 2.5 * m->ri[i] * di[i];
  }
  const double sT = ctl.rtol_t * T * rcv;
  if (fabs(q) * h > sT) h = sT / fabs(q);
  for (int e = lane; e < CH_NS * CH_NS; e += 64) {
    const int i = e / CH_NS, j = e % CH_NS;
    L.A[i][j] = (i == j ? 1.0 : 0.0) - L.A[i][j] * L.mi[i] / L.mi[j] * h;
  }
  if (lane < CH_NS) L.A[lane][CH_NS] = L.ri[lane] * L.mi[lane] * 1.0e6 * h;
  __syncthreads();
  // Gauss elimination with partial pivoting (CHEM_reactions.f90:523-567)
  for (int k = 0; k < CH_NS - 1; k++) {
    double fmax = fabs(L.A[k][k]);
    int kk = k;
    for (int i = k + 1; i < CH_NS; i++)
      if (fabs(L.A[i][k]) > fmax) {
        fmax = fabs(L.A[i][k]);
        kk = i;
      }
    if (kk != k && lane >= k && lane <= CH_NS) {
      const double tmp = L.A[k][lane];
      L.A[k][lane] = L.A[kk][lane];
      L.A[kk][lane] = tmp;
    }
    __syncthreads();
    if (lane > k && lane < CH_NS) L.f[lane] = -L.A[lane][k] / L.A[k][k];
    __syncthreads();
    const int w = CH_NS + 1 - k;  // columns k .. CH_NS (right-hand side included)
    for (int e = lane; e < (CH_NS - 1 - k) * w; e += 64) {
      const int i = k + 1 + e / w, j = k + e % w;
      L.A[i][j] = L.A[i][j] + L.f[i] * L.A[k][j];
    }
    __syncthreads();
  }
  dd[CH_NS - 1] = L.A[CH_NS - 1][CH_NS] / L.A[CH_NS - 1][CH_NS - 1];
#pragma unroll
  for (int i = CH_NS - 2; i >= 0; i--) {
    double xi = L.A[i][CH_NS];
#pragma unroll
    for (int j = i + 1; j < CH_NS; j++) xi = xi - L.A[i][j] * dd[j];
    dd[i] = xi / L.A[i][i];
  }
  bool ok = true;
  double de = 0.0;
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    ok = ok && fabs(dd[i]) <= 2.0 * sk[i];
    de = de + dd[i] * m->ri[i] * m->coef[i][0][5];
  }
  ok = ok && fabs(de) <= 2.0 * sT;
  if (!ok) {
    h_cap = 0.5 * h;
    return false;
  }
  chem_apply(m, di, E, dd);
  h_cap = 2.0 * h;
  h_taken = h;
  return true;
}

// One wavefront per handed-off cell, taken from a queue over the records.
__global__ __launch_bounds__(64) void k_chem_coop(long n, double* rho, double* Ecell, const double* dtcell, ChemCtl ctl,
                                                  int* nsub_out, unsigned long long* queue, HeavyBuf hb) {
  __shared__ CoopLds L;
  const int lane = threadIdx.x;
  Mech m = (Mech)&cm;
  const unsigned long long count = queue[3];
  // stage the lane-indexed mechanism data: this lane's reaction in
  // registers, the stoichiometry, NASA coefficients and molar masses in LDS
  const int nr = m->nr;
  RxnConst R = load_rxn(m, lane < nr ? lane : 0);
  for (int e = lane; e < CH_NR_MAX * CH_NS; e += 64) {
    const int k = e / CH_NS, i = e % CH_NS;
    L.vf[k][i] = (signed char)(k < nr ? m->vf[k][i] : 0);
    L.vr[k][i] = (signed char)(k < nr ? m->vr[k][i] : 0);
  }
  for (int e = lane; e < CH_NS * CH_NT_MAX * 7; e += 64) {
    const int i = e / (CH_NT_MAX * 7), j = (e / 7) % CH_NT_MAX, c = e % 7;
    L.coef[i][j][c] = m->coef[i][j][c];
  }
  if (lane < CH_NS) L.mi[lane] = m->mi[lane];
  __syncthreads();
  while (true) {
    unsigned long long r = 0;
    if (lane == 0) r = atomicAdd(queue + 4, 1ull);
    r = __shfl(r, 0);
    if (r >= count) break;
    const long c = hb.cell[r];
    double di[CH_NS];
#pragma unroll
    for (int i = 0; i < CH_NS; i++) di[i] = hb.di[(long)i * hb.cap + r];
    double E = hb.E[r], t = hb.t[r], h_cap = hb.h_cap[r];
    int nsub = hb.nsub[r], nrej = hb.nrej[r];
    const double dt = dtcell[c];
    // every lane holds the same state: readfirstlane states that the loop
    // decisions are uniform, which keeps the control flow scalar
    while (__builtin_amdgcn_readfirstlane((t < dt && nsub < ctl.max_sub && nrej < ctl.max_sub) ? 1 : 0)) {
      m = mech_opaque(m);
      double h = 0.0;
      const bool acc = coop_attempt(m, L, R, lane, di, E, dt, t, h_cap, ctl, h);
      if (!__builtin_amdgcn_readfirstlane(acc ? 1 : 0)) {
        nrej++;
        continue;
      }
      t = t + h;
      nsub++;
      if (__builtin_amdgcn_readfirstlane((dt - t <= 1e-12 * dt) ? 1 : 0)) break;
    }
    if (lane == 0) chem_finish(c, n, rho, Ecell, di, E, nsub, t >= dt || dt - t <= 1e-12 * dt, nsub_out, queue);
  }
}

__global__ void k_chem_reset(unsigned long long* queue) {
  if (threadIdx.x < 8 && blockIdx.x == 0) queue[threadIdx.x] = 0;
}

// ===========================================================================
// Learned surrogate (no reference counterpart).  Per cell: T from E by the
// reference's Newton iteration, features
//   f = [T, ln p, cbrt(Y_0..Y_8), ln dt],   x = (f - in_shift) * in_scale,
// the MLP (BF16 MFMA, FP32 accumulation) gives z, and in FP64
//   dY = P (z * out_scale)          (P projects onto element conservation)
// is accepted when the cell lies in the training domain, the H/O/N balance of
// dY is within elem_tol of the cell's total element content (moles of H, O
// and N per kg), Y + dY >= -y_neg_tol, every
// |dY_k| <= dy_max (the increment size up to which the network's error was
// measured to stay within tolerance) and the temperature after the update
// (Newton on the new state, update_Et energy) is finite and within
// [T_out_lo, T_out_hi].  Accepted cells are written
// (rho_k += rho dY_k, E by update_Et); the others go to a list for the plain
// FP64 integrator.
struct SurrCfg {
  double in_shift[12], in_scale[12];
  double out_scale[CH_NS];
  double T_lo, T_hi, p_lo, p_hi, dt_lo, dt_hi;
  double y_neg_tol, elem_tol, T_out_lo, T_out_hi, dy_max;
  double P[CH_NS][CH_NS];
  double Cel[3][CH_NS];
};
__constant__ SurrCfg scfg;
__constant__ MlpDevice cnet;
// the network's weights in fragment order (mt_pack_src): layer 0 over K = 32,
// hidden layers over 256, the output layer's first 16 rows
struct SurrPacked {
  const tbf16* W[MLP_MAX_LAYERS];
};
__constant__ SurrPacked cwp;

__global__ void k_pack_frag(long count, int KS, int ldw, const tbf16* __restrict__ src, tbf16* __restrict__ dst) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < count) dst[e] = src[mt_pack_src(e, KS, ldw)];
}

// raw surrogate features of one cell: [T, ln p, cbrt(Y_0..Y_8), ln dt]
__device__ __forceinline__ void surr_raw_features(Mech m, const double* di, double r, double E, double dt, double* f,
                                                  double& T, double& p) {
  T = chem_T(m, di, E);
  p = 0.0;
#pragma unroll
  for (int i = 0; i < CH_NS; i++) p = p + di[i] * m->ri[i] * T;
  f[0] = T;
  f[1] = log(p);
#pragma unroll
  for (int i = 0; i < CH_NS; i++) f[2 + i] = cbrt(di[i] / r);
  f[11] = log(dt);
}

// mode 0: T [n]; 1: raw features [n][12] (f64); 2: normalised inputs [n][12]
// (f32, (f - in_shift) * in_scale)
__global__ __launch_bounds__(128) void k_cell_features(long n, const double* __restrict__ rho,
                                                        const double* __restrict__ Ecell, const double* __restrict__ dtcell,
                                                        int mode, void* out) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  Mech m = (Mech)&cm;
  double di[CH_NS], r = 0.0;
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    di[i] = rho[(long)i * n + c];
    r = r + di[i];
  }
  if (mode == 0) {
    ((double*)out)[c] = chem_T(m, di, Ecell[c]);
    return;
  }
  double f[12], T, p;
  surr_raw_features(m, di, r, Ecell[c], dtcell[c], f, T, p);
#pragma unroll
  for (int j = 0; j < 12; j++) {
    if (mode == 1) ((double*)out)[c * 12 + j] = f[j];
    else ((float*)out)[c * 12 + j] = (float)((f[j] - scfg.in_shift[j]) * scfg.in_scale[j]);
  }
}

// normalised targets: y [n][9] = ((rho1_k - rho0_k) / rho0) / out_scale_k
__global__ void k_surr_targets(long n, const double* __restrict__ rho0, const double* __restrict__ rho1, float* y) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < CH_NS; i++) r = r + rho0[(long)i * n + c];
#pragma unroll
  for (int i = 0; i < CH_NS; i++)
    y[c * CH_NS + i] = (float)((rho1[(long)i * n + c] - rho0[(long)i * n + c]) / r / scfg.out_scale[i]);
}

__global__ void k_fill_u64(unsigned long long* p, int count) {
  if ((int)threadIdx.x < count && blockIdx.x == 0) p[threadIdx.x] = 0;
}

// Tile shapes of the surrogate network kernel (cells per workgroup, waves,
// workgroups per CU the register budget aims at, row groups of waves: with
// NWR = 2 the waves form 2 x 2 blocks of ROWS/2 rows x 128 columns, which
// halves the LDS reads per MFMA and doubles the weight loads).  All give
// bit-identical results; ocfd_chem_surrogate_autotune picks the fastest.
template <int V> struct SurrTile;
template <> struct SurrTile<0> { static constexpr int ROWS = 128, NW = 4, OCC = 2, NWR = 1; };
template <> struct SurrTile<1> { static constexpr int ROWS = 128, NW = 4, OCC = 2, NWR = 2; };
template <> struct SurrTile<2> { static constexpr int ROWS = 64, NW = 4, OCC = 3, NWR = 1; };
template <> struct SurrTile<3> { static constexpr int ROWS = 96, NW = 4, OCC = 2, NWR = 1; };
#define SURR_VARIANTS 4

// The surrogate step runs as three launches, so that every phase keeps all
// of its threads busy: the per-cell FP64 work (features with the Newton
// temperature; projection, checks, update_Et and a second Newton
// temperature) takes one cell per thread, and the network takes whole
// workgroups of MFMA waves.  (Fused into one kernel, the per-cell phases ran
// on the ROWS threads of a workgroup of NW * 64 while the other waves waited
// at its barriers.)
//
// Phase 1: normalised inputs X[n][16] (bfloat16, rounded to nearest even;
// features 12..15 zero), the network's first-layer operand rows, and the
// domain check ok[n].
__global__ __launch_bounds__(256) void k_surr_features(long n, const double* __restrict__ rho,
                                                       const double* __restrict__ Ecell,
                                                       const double* __restrict__ dtcell, tbf16* __restrict__ X,
                                                       unsigned char* __restrict__ okv) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  Mech m = (Mech)&cm;
  float x[12];
  double di[CH_NS];
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    di[i] = rho[(long)i * n + c];
    r = r + di[i];
  }
  const double E = Ecell[c], dt = dtcell[c];
  double f[12], T, p;
  surr_raw_features(m, di, r, E, dt, f, T, p);
  int ok = (T >= scfg.T_lo && T <= scfg.T_hi && p >= scfg.p_lo && p <= scfg.p_hi && dt >= scfg.dt_lo &&
            dt <= scfg.dt_hi && r > 0.0) ? 1 : 0;
#pragma unroll
  for (int j = 0; j < 12; j++) {
    x[j] = (float)((f[j] - scfg.in_shift[j]) * scfg.in_scale[j]);
    if (!(fabsf(x[j]) <= 1e6f)) ok = 0;  // also catches NaN
  }
  tbf16x8 lo, hi;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    lo[j] = (tbf16)(ok ? x[j] : 0.f);
    hi[j] = (tbf16)(ok && j < 4 ? x[8 + j] : 0.f);
  }
  *(tbf16x8*)&X[c * 16] = lo;
  *(tbf16x8*)&X[c * 16 + 8] = hi;
  okv[c] = (unsigned char)ok;
}

// Phase 2: the network on tiles of ROWS cells, O[9][n] (float) from X.  The
// input rows are staged into LDS columns 0..15 with columns 16..31 zeroed:
// the first layer runs over K = 32 (its weights beyond input 12 are zero).
template <int V>
__global__ __launch_bounds__(SurrTile<V>::NW * 64, SurrTile<V>::OCC) void k_surr_mlp(long n,
                                                                                       const tbf16* __restrict__ X,
                                                                                       float* __restrict__ Oout) {
  constexpr int ROWS = SurrTile<V>::ROWS, NW = SurrTile<V>::NW, NT = NW * 64;
  __shared__ __attribute__((aligned(16))) tbf16 H[ROWS * MT_LDH];
  const int t = threadIdx.x, lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
  const long c0 = (long)blockIdx.x * ROWS;
  const tbf16x8 zero8 = {};
  for (int e = t; e < ROWS * 2; e += NT) {
    const int row = e >> 1, h = (e & 1) * 8;
    const long c = c0 + row;
    *(tbf16x8*)&H[row * MT_LDH + h] = c < n ? *(const tbf16x8*)&X[c * 16 + h] : zero8;
    *(tbf16x8*)&H[row * MT_LDH + 16 + h] = zero8;
  }
  __syncthreads();
  const int L = cnet.n_layers;
  constexpr int NWR = SurrTile<V>::NWR;
  mt_layer256<true, ROWS, NW, 32, NWR>(H, cwp.W[0], cnet.b[0], lane, w);
  for (int l = 1; l < L - 1; l++) mt_layer256<true, ROWS, NW, 256, NWR>(H, cwp.W[l], cnet.b[l], lane, w);
  mt_out_cols<ROWS, NW, 256, CH_NS>(H, cwp.W[L - 1], cnet.b[L - 1], Oout, n, c0, lane, w);
}

// Phase 3: in FP64, dY = P (O * out_scale) (P projects onto element
// conservation); the cell is accepted when it lies in the training domain,
// the H/O/N balance of dY is within elem_tol of the cell's total element
// content (moles of H, O and N per kg), Y + dY >= -y_neg_tol, every
// |dY_k| <= dy_max (the increment size up to which the network's error was
// measured to stay within tolerance) and the temperature after the update
// (Newton on the new state, update_Et energy) is finite and within
// [T_out_lo, T_out_hi].  Accepted cells are written (rho_k += rho dY_k, E by
// update_Et); the others go to a list for the plain FP64 integrator.
__global__ __launch_bounds__(256) void k_surr_finish(long n, double* rho, double* Ecell,
                                                     const float* __restrict__ Oin,
                                                     const unsigned char* __restrict__ okv, int* flag_out,
                                                     int* fb_list, unsigned long long* counters) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  Mech m = (Mech)&cm;
  int why = okv[c] ? 0 : 1;
  double di[CH_NS], dY[CH_NS], dYp[CH_NS], dinew[CH_NS];
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    di[i] = rho[(long)i * n + c];
    r = r + di[i];
    dY[i] = (double)Oin[(long)i * n + c] * scfg.out_scale[i];
  }
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < CH_NS; j++) a = a + scfg.P[i][j] * dY[j];
    dYp[i] = a;
  }
  // each element's balance to elem_tol of the cell's total element content
  // (mol/kg): a cell without one element (pure air has no H) is held to the
  // same absolute bound as the others, not to zero
  double content = 0.0;
#pragma unroll
  for (int e = 0; e < 3; e++)
#pragma unroll
    for (int i = 0; i < CH_NS; i++) content = content + scfg.Cel[e][i] * fmax(di[i] / r, 0.0);
#pragma unroll
  for (int e = 0; e < 3; e++) {
    double res = 0.0;
#pragma unroll
    for (int i = 0; i < CH_NS; i++) res = res + scfg.Cel[e][i] * dYp[i];
    if (!(fabs(res) <= scfg.elem_tol * content + 1e-300)) why |= 2;
  }
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    if (!(di[i] / r + dYp[i] >= -scfg.y_neg_tol)) why |= 4;
    if (!(fabs(dYp[i]) <= scfg.dy_max)) why |= 16;  // also NaN
    dinew[i] = di[i] + r * dYp[i];
  }
  double E = Ecell[c];
  chem_update_et(m, di, dinew, E);
  const double Tn = chem_T(m, dinew, E);
  if (!(Tn >= scfg.T_out_lo && Tn <= scfg.T_out_hi)) why |= 8;  // also NaN
  if (why == 0) {
#pragma unroll
    for (int i = 0; i < CH_NS; i++) rho[(long)i * n + c] = dinew[i];
    Ecell[c] = E;
  } else {
    const unsigned long long k = atomicAdd(counters + 2, 1ull);
    fb_list[k] = (int)c;
  }
  if (flag_out) flag_out[c] = why;
}

// checks of the surrogate network, fallback list capacity, network view into
// the constant bank, counters cleared (on CX.stream)
int surr_prepare(long n, const char* who) {
  if (!CX.ready || !CX.surr_ready) return cfail(std::string(who) + ": mechanism or surrogate not set");
  if (!CX.rev_ready) return cfail(std::string(who) + ": ReReac 1/2 needs ocfd_chem_set_reverse_rates");
  const MlpDevice* net = ocfd_mlp_device_view();
  if (!net) return cfail(std::string(who) + ": no network (ocfd_mlp_create)");
  const int L = net->n_layers;
  bool shape_ok = L >= 2 && net->din[0] == 12 && net->dinp[0] == 64 && net->dout[L - 1] == CH_NS;
  for (int l = 0; l < L - 1; l++) shape_ok = shape_ok && net->dout[l] == 256;
  if (!shape_ok) return cfail(std::string(who) + ": the network must be 12 -> 256 x k -> 9");
  if (n > 0x7fffffffL) return cfail(std::string(who) + ": at most 2^31-1 cells");
  if (n > CX.fb_cap) {
    if (CX.fb_list) hipFree(CX.fb_list);
    if (CX.surr_buf) hipFree(CX.surr_buf);
    CX.fb_list = nullptr;
    CX.surr_buf = nullptr;
    CX.fb_cap = 0;
    // the phases' buffers: inputs X[n][16] (bf16), outputs O[9][n] (float), ok[n]
    if (hipMalloc(&CX.fb_list, (size_t)n * sizeof(int)) != hipSuccess ||
        hipMalloc(&CX.surr_buf, (size_t)n * (16 * sizeof(tbf16) + 9 * sizeof(float) + 1)) != hipSuccess)
      return cfail(std::string(who) + ": hipMalloc");
    CX.fb_cap = n;
  }
  // the weights in fragment order, repacked whenever the network's weights
  // changed: 16 output blocks x 1 chunk for layer 0, 16 x 8 for the hidden
  // layers, 1 x 8 for the output layer
  long off[MLP_MAX_LAYERS], need = 0;
  for (int l = 0; l < L; l++) {
    off[l] = need;
    need += (l == L - 1 ? 1 : 16) * (l == 0 ? 1 : 8) * 512L;
  }
  if (need > CX.surr_wpack_cap) {
    CX.surr_wp_version = ~0ull;
    if (CX.surr_wpack) hipFree(CX.surr_wpack);
    CX.surr_wpack = nullptr;
    CX.surr_wpack_cap = 0;
    if (hipMalloc(&CX.surr_wpack, need * sizeof(tbf16)) != hipSuccess)
      return cfail(std::string(who) + ": hipMalloc");
    CX.surr_wpack_cap = need;
  }
  if (CX.surr_wp_version != ocfd_mlp_weights_version()) {
    for (int l = 0; l < L; l++) {
      const long cnt = (l == L - 1 ? 1 : 16) * (l == 0 ? 1 : 8) * 512L;
      hipLaunchKernelGGL(k_pack_frag, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, CX.stream, cnt,
                         l == 0 ? 1 : 8, net->dinp[l], (const tbf16*)net->W[l], CX.surr_wpack + off[l]);
      CX.surr_wp[l] = CX.surr_wpack + off[l];
    }
    CX.surr_wp_version = ocfd_mlp_weights_version();
  }
  if (hipMemcpyToSymbolAsync(HIP_SYMBOL(cnet), net, sizeof(MlpDevice), 0, hipMemcpyHostToDevice, CX.stream) !=
          hipSuccess ||
      hipMemcpyToSymbolAsync(HIP_SYMBOL(cwp), CX.surr_wp, sizeof(SurrPacked), 0, hipMemcpyHostToDevice,
                             CX.stream) != hipSuccess)
    return cfail(std::string(who) + ": copy failed");
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, CX.stream, CX.counter, 8);
  return 0;
}

template <int V>
void launch_surrogate_v(long n, double* rho, double* E, const double* dt, int* flag_out) {
  constexpr int R = SurrTile<V>::ROWS, NW = SurrTile<V>::NW;
  tbf16* X = (tbf16*)CX.surr_buf;
  float* O = (float*)(X + 16 * (size_t)CX.fb_cap);
  unsigned char* ok = (unsigned char*)(O + 9 * (size_t)CX.fb_cap);
  const dim3 g1((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL(k_surr_features, g1, dim3(256), 0, CX.stream, n, (const double*)rho, (const double*)E, dt, X, ok);
  hipLaunchKernelGGL(k_surr_mlp<V>, dim3((unsigned)((n + R - 1) / R)), dim3(NW * 64), 0, CX.stream, n,
                     (const tbf16*)X, O);
  hipLaunchKernelGGL(k_surr_finish, g1, dim3(256), 0, CX.stream, n, rho, E, (const float*)O,
                     (const unsigned char*)ok, flag_out, CX.fb_list, CX.counter);
}
void launch_surrogate(int v, long n, double* rho, double* E, const double* dt, int* flag_out) {
  switch (v) {
    case 1: launch_surrogate_v<1>(n, rho, E, dt, flag_out); break;
    case 2: launch_surrogate_v<2>(n, rho, E, dt, flag_out); break;
    case 3: launch_surrogate_v<3>(n, rho, E, dt, flag_out); break;
    default: launch_surrogate_v<0>(n, rho, E, dt, flag_out); break;
  }
}

int ensure_heavy(long n) {
  if (n <= CX.heavy.cap) return 0;
  if (CX.heavy_mem) hipFree(CX.heavy_mem);
  CX.heavy_mem = nullptr;
  CX.heavy = HeavyBuf{};
  const size_t bytes = (size_t)n * ((CH_NS + 3) * sizeof(double) + 3 * sizeof(int)) + 256;
  if (hipMalloc(&CX.heavy_mem, bytes) != hipSuccess) return cfail("chemistry: out of device memory (work records)");
  HeavyBuf& h = CX.heavy;
  double* d = (double*)CX.heavy_mem;
  h.di = d;
  h.E = d + (size_t)CH_NS * n;
  h.t = h.E + n;
  h.h_cap = h.t + n;
  int* q = (int*)(h.h_cap + n);
  h.nsub = q;
  h.nrej = q + n;
  h.cell = q + 2 * n;
  h.cap = n;
  return 0;
}

// Sort key of a cell, an estimate of its work: dt T^2 (T by comput_T_rgas).
// On the 2^21-cell jet-flame batch this order gives 14.4 ms against 16.0 ms
// for dt alone and 11.4 ms for the true substep counts (dt exp(-Ta/T) is no
// better, at the cost of an exp).  The key is the top 32 bits of the value
// (sign, exponent and 20 mantissa bits, which order like the values for a
// positive finite number: to within 1e-6); 0 otherwise (such cells go last).
__device__ __forceinline__ unsigned int cell_key(Mech m, const double* __restrict__ rho,
                                                 const double* __restrict__ Ecell, const double* __restrict__ dtcell,
                                                 long n, long c) {
  double di[CH_NS];
#pragma unroll
  for (int i = 0; i < CH_NS; i++) di[i] = rho[(long)i * n + c];
  const double T = chem_T(m, di, Ecell[c]);
  const double w = dtcell[c] * T * T;
  return (w > 0.0 && w <= 1.7976931348623157e308) ? (unsigned int)((unsigned long long)__double_as_longlong(w) >> 32)
                                                  : 0u;
}

// keys[c] = cell_key of cell c, idx[c] = c for c < n, and *count = n
__global__ void k_order_keys(long n, const double* __restrict__ rho, const double* __restrict__ Ecell,
                             const double* __restrict__ dt, unsigned int* __restrict__ keys, int* idx,
                             unsigned long long* count) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) {
    keys[c] = cell_key((Mech)&cm, rho, Ecell, dt, n, c);
    idx[c] = (int)c;
  }
  if (c == 0) *count = (unsigned long long)n;
}

int ensure_order(long n) {
  if (n <= CX.ord_cap) return 0;
  if (CX.ord_mem) hipFree(CX.ord_mem);
  CX.ord_mem = nullptr;
  CX.ord_cap = 0;
  size_t temp = 0;
  if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, temp, (const unsigned int*)nullptr, (unsigned int*)nullptr,
                                                   (const int*)nullptr, (int*)nullptr, (int)n) != hipSuccess)
    return cfail("chemistry: radix sort query failed");
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_keys = up(sizeof(unsigned int) * (size_t)n), b_idx = up(sizeof(int) * (size_t)n), b_cnt = 256;
  if (hipMalloc(&CX.ord_mem, 2 * b_keys + 2 * b_idx + b_cnt + up(temp)) != hipSuccess)
    return cfail("chemistry: out of device memory (cell order)");
  unsigned char* p = CX.ord_mem;
  CX.ord_keys = (unsigned int*)p;
  CX.ord_keys_in = (unsigned int*)(p + b_keys);
  p += 2 * b_keys;
  CX.ord_in = (int*)p;
  CX.ord_out = (int*)(p + b_idx);
  CX.ord_count = (unsigned long long*)(p + 2 * b_idx);
  CX.ord_temp = p + 2 * b_idx + b_cnt;
  CX.ord_temp_bytes = temp;
  CX.ord_cap = n;
  return 0;
}

// Cell indices in order of decreasing cell_key, ties in index order (a
// stable radix sort): the cells with the longest integrations start first,
// so the batch does not end waiting on a stiff cell that a lane took up
// late.  The result is CX.ord_out[0 .. n) with *CX.ord_count = n.
int order_longest_first(long n, const double* rho, const double* E, const double* dt) {
  if (n > 0x7fffffffL) return cfail("chemistry: the cell order needs n < 2^31");
  if (ensure_order(n)) return 1;
  hipLaunchKernelGGL(k_order_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, CX.stream, n, rho, E, dt,
                     CX.ord_keys_in, CX.ord_in, CX.ord_count);
  size_t temp = CX.ord_temp_bytes;
  const hipError_t e = hipcub::DeviceRadixSort::SortPairsDescending(CX.ord_temp, temp, CX.ord_keys_in, CX.ord_keys,
                                                                    CX.ord_in, CX.ord_out, (int)n, 0, 32, CX.stream);
  if (e != hipSuccess) return cfail(std::string("chemistry: radix sort failed: ") + hipGetErrorString(e));
  return 0;
}

// Sort keys of a cell list whose length is on the device: cell_key of each
// listed cell, 0 past the end.  A stable descending sort keeps every listed
// cell ahead of the slots past the end (their keys are the smallest, and
// among equal keys the listed cells come first in index order).
__global__ void k_list_keys(long n, long cap, const int* __restrict__ list, const unsigned long long* __restrict__ count,
                            const double* __restrict__ rho, const double* __restrict__ Ecell,
                            const double* __restrict__ dt, unsigned int* __restrict__ keys) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cap) return;
  unsigned int k = 0u;
  if ((unsigned long long)i < *count) {
    const long c = list[i];
    if (c >= 0 && c < n) k = cell_key((Mech)&cm, rho, Ecell, dt, n, c);
  }
  keys[i] = k;
}

// The cells of list[0 .. *count) (count <= cap <= n) in order of decreasing
// cell_key, into CX.ord_out[0 .. *count).
int order_list_longest_first(long n, long cap, const int* list, const unsigned long long* count, const double* rho,
                             const double* E, const double* dt) {
  if (n > 0x7fffffffL || cap > n) return cfail("chemistry: the cell order needs cap <= n < 2^31");
  if (ensure_order(n)) return 1;
  hipLaunchKernelGGL(k_list_keys, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, CX.stream, n, cap, list, count,
                     rho, E, dt, CX.ord_keys_in);
  size_t temp = CX.ord_temp_bytes;
  const hipError_t e = hipcub::DeviceRadixSort::SortPairsDescending(CX.ord_temp, temp, CX.ord_keys_in, CX.ord_keys,
                                                                    list, CX.ord_out, (int)cap, 0, 32, CX.stream);
  if (e != hipSuccess) return cfail(std::string("chemistry: radix sort failed: ") + hipGetErrorString(e));
  return 0;
}

// the adaptive integrator over cells 0..n-1, or over list[0 .. *list_count)
// when list is given; counters [0,1,3,4] must be zero.  scheduler 0: one
// cell per thread; 1: per-lane work queue; 2: work queue with the cooperative
// kernel taking the cells that outlast CX.coop_budget attempts.  direct: a
// list of at most CX.coop_direct_max cells goes to the cooperative kernel
// after one attempt per cell (the surrogate's rejects).
int launch_adaptive(long n, double* rho, double* E, const double* dt, const ChemCtl& ctl, int* nsub_out, int scheduler,
                    const int* list, const unsigned long long* list_count, bool direct = true,
                    unsigned long long* counter = nullptr) {
  if (!counter) counter = CX.counter;
  if (scheduler == 0 && !list) {
    hipLaunchKernelGGL(k_chem_adaptive, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, CX.stream, n, rho, E, dt, ctl,
                       nsub_out, counter);
    return hipGetLastError() == hipSuccess ? 0 : cfail("chemistry: launch failed");
  }
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_chem_queue, 128, 0) != hipSuccess)
    return cfail("chemistry: device query failed");
  if (per < 1) per = 1;
  long blocks = (long)ncu * per;
  const long need = (n + 127) / 128;
  if (blocks > need) blocks = need;
  const bool coop = scheduler == 2;
  if (coop && ensure_heavy(n)) return 1;
  hipLaunchKernelGGL(k_chem_queue, dim3((unsigned)blocks), dim3(128), 0, CX.stream, n, rho, E, dt, ctl, nsub_out,
                     counter, list, list_count, coop ? CX.coop_budget : 0, direct ? CX.coop_direct_max : 0,
                     coop ? CX.heavy : HeavyBuf{});
  if (coop) {
    int per2 = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per2, k_chem_coop, 64, 0) != hipSuccess)
      return cfail("chemistry: device query failed");
    long b2 = (long)ncu * (per2 < 1 ? 1 : per2);
    if (b2 > n) b2 = n;
    hipLaunchKernelGGL(k_chem_coop, dim3((unsigned)b2), dim3(64), 0, CX.stream, n, rho, E, dt, ctl, nsub_out,
                       counter, CX.heavy);
  }
  return hipGetLastError() == hipSuccess ? 0 : cfail("chemistry: launch failed");
}

// ===========================================================================
// The batch entry behind ocfd_chem_adaptive_batch; the flow solver calls it
// with order = false (its cells share one dt, so the order is the index order).
int chem_adaptive_batch(long n, double* rho, double* E, const double* dt, double rtol_y, double y_floor, double rtol_t,
                        int max_sub, int* nsub_out, int scheduler, bool order) {
  if (!CX.ready) return cfail("ocfd_chem_adaptive_batch: call ocfd_chem_set_mechanism first");
  if (!CX.rev_ready) return cfail("ocfd_chem_adaptive_batch: ReReac 1/2 needs ocfd_chem_set_reverse_rates");
  if (max_sub < 1 || !(rtol_y > 0.0) || !(y_floor >= 0.0) || !(rtol_t > 0.0))
    return cfail("ocfd_chem_adaptive_batch: bad controller parameters");
  if (scheduler < 0 || scheduler > 2) return cfail("ocfd_chem_adaptive_batch: scheduler 0, 1 or 2");
  // the cooperative hand-off records a cell index as int
  if (scheduler == 2 && n > 0x7fffffffL) return cfail("ocfd_chem_adaptive_batch: scheduler 2 takes at most 2^31-1 cells");
  const ChemCtl ctl{rtol_y, y_floor, rtol_t, max_sub};
  hipLaunchKernelGGL(k_chem_reset, dim3(1), dim3(64), 0, CX.stream, CX.counter);
  if (n <= 0) return hipGetLastError() == hipSuccess ? 0 : cfail("ocfd_chem_adaptive_batch: launch failed");
  // the order lists cells as int: a batch of 2^31 cells or more runs in index order
  if (order && scheduler != 0 && n > 1 && n <= 0x7fffffffL) {
    if (order_longest_first(n, rho, E, dt)) return 1;
    return launch_adaptive(n, rho, E, dt, ctl, nsub_out, scheduler, CX.ord_out, CX.ord_count, false);
  }
  return launch_adaptive(n, rho, E, dt, ctl, nsub_out, scheduler, nullptr, nullptr);
}

// The batch entry behind ocfd_chem_surrogate_batch; with order, the fallback
// list is taken by decreasing dt (the flow solver passes false: one dt).
int chem_surrogate_batch(long n, double* rho, double* E, const double* dt, bool fallback, double rtol_y, double y_floor,
                         double rtol_t, int max_sub, int* flag_out, bool order) {
  if (fallback && (max_sub < 1 || !(rtol_y > 0.0) || !(y_floor >= 0.0) || !(rtol_t > 0.0)))
    return cfail("ocfd_chem_surrogate_batch: bad controller parameters");
  if (surr_prepare(n, "ocfd_chem_surrogate_batch")) return 1;
  if (n <= 0) return hipGetLastError() == hipSuccess ? 0 : cfail("ocfd_chem_surrogate_batch: launch failed");
  launch_surrogate(CX.surr_variant, n, rho, E, dt, flag_out);
  if (fallback) {
    const ChemCtl ctl{rtol_y, y_floor, rtol_t, max_sub};
    const int* list = CX.fb_list;
    if (order && CX.fallback_scheduler != 0 && n > 1) {
      // the list's length decides the work: none (no launch), or a sort of
      // that many keys only (sorting all n slots cost 0.3 ms on 2^21 cells)
      unsigned long long cnt = 0;
      if (hipMemcpyAsync(&cnt, CX.counter + 2, sizeof(cnt), hipMemcpyDeviceToHost, CX.stream) != hipSuccess ||
          hipStreamSynchronize(CX.stream) != hipSuccess)
        return cfail("ocfd_chem_surrogate_batch: reading the fallback count failed");
      if (cnt == 0) return hipGetLastError() == hipSuccess ? 0 : cfail("ocfd_chem_surrogate_batch: launch failed");
      const long cap = (long)std::min<unsigned long long>(cnt, (unsigned long long)n);
      if (cap > 1) {
        if (order_list_longest_first(n, cap, CX.fb_list, CX.counter + 2, rho, E, dt)) return 1;
        list = CX.ord_out;
      }
    }
    if (launch_adaptive(n, rho, E, dt, ctl, nullptr, CX.fallback_scheduler, list, CX.counter + 2)) return 1;
  }
  if (hipGetLastError() != hipSuccess) return cfail("ocfd_chem_surrogate_batch: launch failed");
  return 0;
}

// ===========================================================================
// Workspace of ocfd_chem_adaptive_batch_noalloc (no reference counterpart):
// the work counters, the cell-order keys and index lists and the radix-sort
// workspace of one batch of at most NA_MAX_CELLS cells, in one file-scope
// device array.  Each GPU holds its own copy (about 32.5 MiB) from module
// load until the process ends; host code reaches it only through the
// per-device address from hipGetSymbolAddress, never through the symbol.
constexpr long NA_MAX_CELLS = 1L << 20;
constexpr size_t NA_COUNTERS = 256;  // 8 u64 work counters (k_chem_reset), then padding
constexpr size_t NA_COUNT = 256;     // length of the ordered list (u64)
constexpr size_t NA_KEYS = sizeof(unsigned int) * NA_MAX_CELLS;
constexpr size_t NA_LIST = sizeof(int) * NA_MAX_CELLS;
constexpr size_t NA_TEMP = size_t(16) << 20;  // radix-sort workspace
constexpr size_t NA_BYTES = NA_COUNTERS + NA_COUNT + 2 * NA_KEYS + 2 * NA_LIST + NA_TEMP;
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
unsigned char* na_base[NA_MAX_DEVICES] = {};

// The current device's copy of scratch, resolved once per device (with that
// device current) on the first call there.
int na_workspace(unsigned char** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES)
    return cfail("ocfd_chem_adaptive_batch_noalloc: no usable current device");
  if (!na_base[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return cfail(std::string("ocfd_chem_adaptive_batch_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    na_base[dev] = (unsigned char*)p;
  }
  *out = na_base[dev];
  return 0;
}

extern "C" {

const char* ocfd_chem_last_error(void) { return CX.err.c_str(); }

int32_t ocfd_chem_set_mechanism(int32_t ns, int32_t nr, int32_t rereac, const int32_t* af_type,
                                const int32_t* third_body, const double* fc_troe, const double* arrhenius,
                                const int32_t* vf, const int32_t* vr, const double* eff, int32_t n_tct,
                                const double* tct, double tct_max, const double* mi, const double* coef) {
  if (ns != CH_NS) return cfail("ocfd_chem_set_mechanism: this build is compiled for 9 species");
  if (nr > CH_NR_MAX || n_tct > CH_NT_MAX || n_tct < 1) return cfail("ocfd_chem_set_mechanism: too many reactions or ranges");
  if (rereac < 0 || rereac > 2) return cfail("ocfd_chem_set_mechanism: ReReac must be 0, 1 or 2");
  ChemMech& h = CX.h;
  h = ChemMech{};
  h.ns = ns;
  h.rereac = rereac;
  CX.rev_ready = rereac == 0;
  h.nr = nr;
  h.n_tct = n_tct;
  for (int j = 0; j < nr; j++) {
    h.af_type[j] = af_type[j];
    h.tb[j] = third_body[j];
    h.fc[j] = fc_troe[j];
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 3; b++) h.arr[j][a][b] = arrhenius[(j * 2 + a) * 3 + b];
    h.sgm[j] = 0;
    for (int i = 0; i < ns; i++) {
      h.vf[j][i] = vf[j * ns + i];
      h.vr[j][i] = vr[j * ns + i];
      h.eff[j][i] = eff[j * ns + i];
      h.sgm[j] = h.sgm[j] + h.vr[j][i] - h.vf[j][i];
    }
  }
  for (int j = 0; j < n_tct - 1; j++) h.tct[j] = tct[j];
  h.tct_max = tct_max;
  auto P2 = [](double x) { return x * x; };
  auto P3 = [](double x) { double x2 = x * x; return x2 * x; };
  auto P4 = [](double x) { double x2 = x * x; return x2 * x2; };
  auto P5 = [](double x) { double x2 = x * x; return (x2 * x2) * x; };
  for (int k = 0; k < ns; k++) {
    h.mi[k] = mi[k];
    h.ri[k] = CH_R0 / mi[k];
    for (int j = 0; j < n_tct; j++)
      for (int c = 0; c < 7; c++) h.coef[k][j][c] = coef[(k * n_tct + j) * 7 + c];
    h.ect[k][0] = 0.0;  // read_Spec, CHEM_thermdyn.f90:43-47
    for (int j = 0; j < n_tct - 1; j++) {
      const double t = h.tct[j];
      h.ect[k][j + 1] = -((h.coef[k][j + 1][0] - h.coef[k][j][0]) * t + (h.coef[k][j + 1][1] - h.coef[k][j][1]) * P2(t) / 2.0 +
                          (h.coef[k][j + 1][2] - h.coef[k][j][2]) * P3(t) / 3.0 +
                          (h.coef[k][j + 1][3] - h.coef[k][j][3]) * P4(t) / 4.0 +
                          (h.coef[k][j + 1][4] - h.coef[k][j][4]) * P5(t) / 5.0) +
                        h.ect[k][j];
    }
  }
  if (!CX.own) {
    if (hipStreamCreateWithFlags(&CX.own, hipStreamNonBlocking) != hipSuccess) return cfail("no HIP device/stream");
    if (hipMalloc(&CX.counter, 8 * sizeof(unsigned long long)) != hipSuccess) return cfail("hipMalloc failed");
  }
  CX.stream = CX.own;
  // kernels still in flight read the constant bank being replaced
  if (hipDeviceSynchronize() != hipSuccess) return cfail("ocfd_chem_set_mechanism: device error");
  if (hipMemcpyToSymbol(HIP_SYMBOL(cm), &h, sizeof(ChemMech)) != hipSuccess) return cfail("hipMemcpyToSymbol failed");
  CX.ready = true;
  return 0;
}

// Reverse-rate data of ReReac = 1 or 2, rev[j*5 .. j*5+4] per reaction (the
// line read_Reac reads after the Arrhenius lines, CHEM_reactions.f90:46-47).
int32_t ocfd_chem_set_reverse_rates(const double* rev) {
  if (!CX.ready) return cfail("ocfd_chem_set_reverse_rates: call ocfd_chem_set_mechanism first");
  if (!rev) return cfail("ocfd_chem_set_reverse_rates: null data");
  ChemMech& h = CX.h;
  for (int j = 0; j < h.nr; j++)
    for (int a = 0; a < 5; a++) h.rev[j][a] = rev[j * 5 + a];
  if (hipDeviceSynchronize() != hipSuccess) return cfail("ocfd_chem_set_reverse_rates: device error");
  if (hipMemcpyToSymbol(HIP_SYMBOL(cm), &h, sizeof(ChemMech)) != hipSuccess) return cfail("hipMemcpyToSymbol failed");
  CX.rev_ready = true;
  return 0;
}

// Chemical_time_advance over a batch of independent cells (OCFD_Time_Adv.f90:169-196
// calls chemical_timeAdv per cell); device pointers; nsub substeps of dt/nsub.
int32_t ocfd_chem_timeadv_batch(int64_t n, double* rho, double* E, const double* dt, int32_t mode, int32_t nsub) {
  if (!CX.ready) return cfail("ocfd_chem_timeadv_batch: call ocfd_chem_set_mechanism first");
  if (!CX.rev_ready) return cfail("ocfd_chem_timeadv_batch: ReReac 1/2 needs ocfd_chem_set_reverse_rates");
  if (n <= 0) return 0;
  if (nsub < 1) return cfail("ocfd_chem_timeadv_batch: nsub must be >= 1");
  dim3 g((unsigned)((n + 127) / 128)), b(128);
  switch (mode) {
    case 1: hipLaunchKernelGGL(k_chem_fixed<1>, g, b, 0, CX.stream, (long)n, rho, E, dt, nsub); break;
    case 2: hipLaunchKernelGGL(k_chem_fixed<2>, g, b, 0, CX.stream, (long)n, rho, E, dt, nsub); break;
    case 3: hipLaunchKernelGGL(k_chem_fixed<3>, g, b, 0, CX.stream, (long)n, rho, E, dt, nsub); break;
    case -1: hipLaunchKernelGGL(k_chem_fixed<-1>, g, b, 0, CX.stream, (long)n, rho, E, dt, nsub); break;
    default: return cfail("ocfd_chem_timeadv_batch: CHEM_TimeAdv must be 1, 2, 3 or -1");
  }
  if (hipGetLastError() != hipSuccess) return cfail("ocfd_chem_timeadv_batch: launch failed");
  return 0;
}

int32_t ocfd_chem_adaptive_batch(int64_t n, double* rho, double* E, const double* dt, double rtol_y, double y_floor,
                                 double rtol_t, int32_t max_sub, int32_t* nsub_out, int32_t scheduler) {
  return chem_adaptive_batch((long)n, rho, E, dt, rtol_y, y_floor, rtol_t, (int)max_sub, (int*)nsub_out,
                             (int)scheduler, CX.order != 0);
}

int32_t ocfd_chem_adaptive_batch_noalloc(int64_t n, double* rho, double* E, const double* dt, double rtol_y,
                                         double y_floor, double rtol_t, int32_t max_sub, int32_t* nsub_out) {
  const char* who = "ocfd_chem_adaptive_batch_noalloc";
  if (!CX.ready) return cfail(std::string(who) + ": call ocfd_chem_set_mechanism first");
  if (!CX.rev_ready) return cfail(std::string(who) + ": ReReac 1/2 needs ocfd_chem_set_reverse_rates");
  if (max_sub < 1 || !(rtol_y > 0.0) || !(y_floor >= 0.0) || !(rtol_t > 0.0))
    return cfail(std::string(who) + ": bad controller parameters");
  if (n < 0 || n > NA_MAX_CELLS)
    return cfail(std::string(who) + ": n must be 0 .. 2^20 (" + std::to_string(NA_MAX_CELLS) +
                 ", the capacity of its static workspace), got " + std::to_string((long long)n));
  unsigned char* base = nullptr;
  if (na_workspace(&base)) return 1;
  unsigned long long* counter = (unsigned long long*)base;
  unsigned long long* count = (unsigned long long*)(base + NA_COUNTERS);
  unsigned char* p = base + NA_COUNTERS + NA_COUNT;
  unsigned int* keys_in = (unsigned int*)p;
  unsigned int* keys_out = (unsigned int*)(p + NA_KEYS);
  int* idx_in = (int*)(p + 2 * NA_KEYS);
  int* idx_out = (int*)(p + 2 * NA_KEYS + NA_LIST);
  void* temp = p + 2 * NA_KEYS + 2 * NA_LIST;
  const ChemCtl ctl{rtol_y, y_floor, rtol_t, max_sub};
  // the previous call's counters are still there: clear them (a kernel)
  hipLaunchKernelGGL(k_chem_reset, dim3(1), dim3(64), 0, CX.stream, counter);
  if (n == 0) return hipGetLastError() == hipSuccess ? 0 : cfail(std::string(who) + ": launch failed");
  const int* list = nullptr;
  const unsigned long long* list_count = nullptr;
  if (n > 1) {
    // the longest-first order of ocfd_chem_adaptive_batch, on the workspace
    size_t need = 0;
    if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, need, (const unsigned int*)nullptr, (unsigned int*)nullptr,
                                                     (const int*)nullptr, (int*)nullptr, (int)n) != hipSuccess)
      return cfail(std::string(who) + ": radix sort query failed");
    if (need > NA_TEMP)
      return cfail(std::string(who) + ": the radix sort needs " + std::to_string(need) + " bytes of workspace, more than " +
                   std::to_string(NA_TEMP));
    hipLaunchKernelGGL(k_order_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, CX.stream, (long)n, rho, E, dt,
                       keys_in, idx_in, count);
    size_t tb = need;
    const hipError_t e = hipcub::DeviceRadixSort::SortPairsDescending(temp, tb, keys_in, keys_out, idx_in, idx_out,
                                                                      (int)n, 0, 32, CX.stream);
    if (e != hipSuccess) return cfail(std::string(who) + ": radix sort failed: " + hipGetErrorString(e));
    list = idx_out;
    list_count = count;
  }
  return launch_adaptive((long)n, rho, E, dt, ctl, (int*)nsub_out, 1, list, list_count, false, counter);
}

int32_t ocfd_chem_surrogate_set(const double* cfg, int32_t count) {
  if (!CX.ready) return cfail("ocfd_chem_surrogate_set: call ocfd_chem_set_mechanism first");
  if (count != (int32_t)(sizeof(SurrCfg) / sizeof(double)))
    return cfail("ocfd_chem_surrogate_set: expected " + std::to_string(sizeof(SurrCfg) / sizeof(double)) + " values");
  SurrCfg h;
  std::memcpy(&h, cfg, sizeof(SurrCfg));
  if (hipStreamSynchronize(CX.stream) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(scfg), &h, sizeof(SurrCfg)) != hipSuccess)
    return cfail("ocfd_chem_surrogate_set: copy failed");
  CX.surr_ready = true;
  return 0;
}

int32_t ocfd_chem_surrogate_batch(int64_t n, double* rho, double* E, const double* dt, int32_t fallback,
                                  double rtol_y, double y_floor, double rtol_t, int32_t max_sub, int32_t* flag_out) {
  return chem_surrogate_batch((long)n, rho, E, dt, fallback != 0, rtol_y, y_floor, rtol_t, (int)max_sub,
                              (int*)flag_out, CX.order != 0);
}

// Autotuner of the surrogate network kernel: every tile shape runs `reps`
// times on scratch copies of the batch (the inputs are not modified), timed
// with HIP events on the chemistry stream; the fastest is kept.  Results are
// bit-identical for every shape.  ms_out (may be NULL) receives the mean ms
// per launch of each shape (SURR_VARIANTS floats).  Synchronises.
int32_t ocfd_chem_surrogate_autotune(int64_t n, const double* rho, const double* E, const double* dt, int32_t reps,
                                     float* ms_out) {
  if (reps < 1 || n <= 0) return cfail("ocfd_chem_surrogate_autotune: need reps >= 1 and cells");
  if (surr_prepare((long)n, "ocfd_chem_surrogate_autotune")) return 1;
  double *r2 = nullptr, *e2 = nullptr;
  if (hipMalloc(&r2, sizeof(double) * (size_t)n * CH_NS) != hipSuccess ||
      hipMalloc(&e2, sizeof(double) * (size_t)n) != hipSuccess) {
    if (r2) hipFree(r2);
    return cfail("ocfd_chem_surrogate_autotune: out of device memory");
  }
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  int best = 0;
  for (int v = 0; v < SURR_VARIANTS; v++) {
    float tot = 0.f;
    for (int r = 0; r <= reps; r++) {  // the first run warms up and is not counted
      hipMemcpyAsync(r2, rho, sizeof(double) * (size_t)n * CH_NS, hipMemcpyDeviceToDevice, CX.stream);
      hipMemcpyAsync(e2, E, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, CX.stream);
      hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, CX.stream, CX.counter, 8);
      hipEventRecord(a, CX.stream);
      launch_surrogate(v, (long)n, r2, e2, dt, nullptr);
      hipEventRecord(b, CX.stream);
      hipEventSynchronize(b);
      float ms = 0.f;
      hipEventElapsedTime(&ms, a, b);
      if (r > 0) tot += ms;
    }
    CX.surr_tune_ms[v] = tot / reps;
    if (CX.surr_tune_ms[v] < CX.surr_tune_ms[best]) best = v;
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  hipFree(r2);
  hipFree(e2);
  CX.surr_variant = best;
  if (ms_out)
    for (int v = 0; v < SURR_VARIANTS; v++) ms_out[v] = CX.surr_tune_ms[v];
  return hipGetLastError() == hipSuccess ? 0 : cfail("ocfd_chem_surrogate_autotune: launch failed");
}

// the tile shape in use (and force one with variant >= 0)
int32_t ocfd_chem_surrogate_tile(int32_t variant, int32_t* rows, int32_t* waves) {
  if (variant >= SURR_VARIANTS) return cfail("ocfd_chem_surrogate_tile: unknown variant");
  if (variant >= 0) CX.surr_variant = variant;
  const int v = CX.surr_variant;
  const int R[SURR_VARIANTS] = {SurrTile<0>::ROWS, SurrTile<1>::ROWS, SurrTile<2>::ROWS, SurrTile<3>::ROWS};
  const int W[SURR_VARIANTS] = {SurrTile<0>::NW, SurrTile<1>::NW, SurrTile<2>::NW, SurrTile<3>::NW};
  if (rows) *rows = R[v];
  if (waves) *waves = W[v];
  return v;
}

// T [n] (mode 0), raw features [n][12] f64 (mode 1) or normalised surrogate
// inputs [n][12] f32 (mode 2) of a batch; device pointers
int32_t ocfd_chem_cell_features(int64_t n, const double* rho, const double* E, const double* dt, int32_t mode,
                                void* out) {
  if (!CX.ready) return cfail("ocfd_chem_cell_features: call ocfd_chem_set_mechanism first");
  if (mode < 0 || mode > 2) return cfail("ocfd_chem_cell_features: mode 0, 1 or 2");
  if (mode == 2 && !CX.surr_ready) return cfail("ocfd_chem_cell_features: surrogate normalisation not set");
  if (mode > 0 && !dt) return cfail("ocfd_chem_cell_features: dt required");
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_cell_features, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, CX.stream, (long)n, rho, E, dt,
                     (int)mode, out);
  return hipGetLastError() == hipSuccess ? 0 : cfail("ocfd_chem_cell_features: launch failed");
}

int32_t ocfd_chem_surrogate_targets(int64_t n, const double* rho0, const double* rho1, float* y) {
  if (!CX.ready || !CX.surr_ready) return cfail("ocfd_chem_surrogate_targets: surrogate normalisation not set");
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_surr_targets, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, CX.stream, (long)n, rho0, rho1, y);
  return hipGetLastError() == hipSuccess ? 0 : cfail("ocfd_chem_surrogate_targets: launch failed");
}

int64_t ocfd_chem_surrogate_fallbacks(void) {
  if (!CX.ready) {
    cfail("ocfd_chem_surrogate_fallbacks: call ocfd_chem_set_mechanism first");
    return -1;
  }
  unsigned long long v = 0;
  if (hipMemcpyAsync(&v, CX.counter + 2, sizeof(v), hipMemcpyDeviceToHost, CX.stream) != hipSuccess ||
      hipStreamSynchronize(CX.stream) != hipSuccess) {
    cfail("ocfd_chem_surrogate_fallbacks: copy failed");
    return -1;
  }
  return (int64_t)v;
}

int32_t ocfd_chem_synchronize(void) {
  if (!CX.stream) return 0;
  if (hipStreamSynchronize(CX.stream) != hipSuccess) return cfail("ocfd_chem_synchronize: stream error");
  return 0;
}

void* ocfd_chem_stream(void) { return (void*)CX.stream; }

int32_t ocfd_chem_set_option(int32_t key, int32_t value) {
  switch (key) {
    case 1:
      if (value < 1) return cfail("ocfd_chem_set_option: the attempt budget must be >= 1");
      CX.coop_budget = value;
      return 0;
    case 2:
      if (value < 1 || value > 2) return cfail("ocfd_chem_set_option: fallback scheduler 1 or 2");
      CX.fallback_scheduler = value;
      return 0;
    case 3:
      if (value < 0) return cfail("ocfd_chem_set_option: the direct cooperative list size must be >= 0");
      CX.coop_direct_max = value;
      return 0;
    case 4:
      if (value < 0 || value > 1) return cfail("ocfd_chem_set_option: cell order 0 or 1");
      CX.order = value;
      return 0;
  }
  return cfail("ocfd_chem_set_option: unknown option");
}

int32_t ocfd_chem_counters(int64_t* out, int32_t count) {
  if (!CX.ready) return cfail("ocfd_chem_counters: call ocfd_chem_set_mechanism first");
  if (count < 0 || count > 8) return cfail("ocfd_chem_counters: at most 8 counters");
  unsigned long long v[8] = {};
  if (hipMemcpyAsync(v, CX.counter, sizeof(v), hipMemcpyDeviceToHost, CX.stream) != hipSuccess ||
      hipStreamSynchronize(CX.stream) != hipSuccess)
    return cfail("ocfd_chem_counters: copy failed");
  for (int i = 0; i < count; i++) out[i] = (int64_t)v[i];
  return 0;
}

int64_t ocfd_chem_unfinished(void) {
  if (!CX.ready) {
    cfail("ocfd_chem_unfinished: call ocfd_chem_set_mechanism first");
    return -1;
  }
  unsigned long long v = 0;
  if (hipMemcpyAsync(&v, CX.counter + 1, sizeof(v), hipMemcpyDeviceToHost, CX.stream) != hipSuccess ||
      hipStreamSynchronize(CX.stream) != hipSuccess) {
    cfail("ocfd_chem_unfinished: copy failed");
    return -1;
  }
  return (int64_t)v;
}

int32_t ocfd_chem_set_stream(void* stream) {
  if (!CX.ready) return cfail("ocfd_chem_set_stream: call ocfd_chem_set_mechanism first");
  CX.stream = (hipStream_t)stream;
  return 0;
}

}  // extern "C"
