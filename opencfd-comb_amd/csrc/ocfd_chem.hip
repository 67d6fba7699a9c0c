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

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/ocfd_hip.h"

#define CH_NS 9
#define CH_NR_MAX 32
#define CH_NT_MAX 4
#define CH_R0 8.314
#define CH_ATM 1.013e5

struct ChemMech {
  int ns, nr, n_tct;
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

namespace {
struct ChemCtx {
  bool ready = false;
  ChemMech h{};
  hipStream_t stream = nullptr, own = nullptr;
  unsigned long long* counter = nullptr;
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
template <bool JAC>
__device__ void chem_rates(Mech m, const double* ci, double T, double* ri, double (*A)[CH_NS]) {
  const double Rs = CH_R0 / CH_ATM * 1.0e6;
  const double logT = log(T);
  double gi[CH_NS];
#pragma unroll
  for (int i = 0; i < CH_NS; i++) gi[i] = chem_gibbs(m, i, T, logT);
#pragma unroll
  for (int i = 0; i < CH_NS; i++) {
    ri[i] = 0.0;
    if (JAC) {
#pragma unroll
      for (int j = 0; j < CH_NS; j++) A[i][j] = 0.0;
    }
  }
  for (int k = 0; k < m->nr; k++) {
    // comput_KjX (CHEM_reactions.f90:76-103)
    double dgi = 0.0;
#pragma unroll
    for (int i = 0; i < CH_NS; i++) dgi = dgi + gi[i] * (m->vr[k][i] - m->vf[k][i]);
    const double kjx = exp(-dgi / T) * ipow(Rs * T, -m->sgm[k]);
    double Ctrb = 0.0;
    if (m->tb[k] == 1) {
#pragma unroll
      for (int i = 0; i < CH_NS; i++) Ctrb = Ctrb + ci[i] * m->eff[k][i];
    }
    const double wf1a = m->arr[k][0][0] * pow(T, m->arr[k][0][1]) * exp(-m->arr[k][0][2] / T);
    double wf, wr;
    if (m->af_type[k] == 0) {
      wf = wf1a;
    } else if (m->af_type[k] == 2) {
      const double wf2 = m->arr[k][1][0] * pow(T, m->arr[k][1][1]) * exp(-m->arr[k][1][2] / T);
      wf = wf1a + wf2;
    } else {
      const double wf2 = m->arr[k][1][0] * pow(T, m->arr[k][1][1]) * exp(-m->arr[k][1][2] / T);
      const double Prs = wf1a / wf2 * Ctrb;
      const double Fc = m->fc[k];
      const double lfc = log10(Fc);
      const double ac = -0.4 - (double)0.67f * lfc;   // REAL(4) literals of the reference
      const double an = 0.75 - (double)1.27f * lfc;
      const double lp = log10(Prs);
      const double F = pow(10.0, lfc / (1.0 + q2((lp + ac) / (an - 0.14 * (lp + ac)))));
      wf = wf2 * Prs / (1.0 + Prs) * F;
    }
    wr = wf / kjx;
    if (m->tb[k] == 1 && m->af_type[k] != 1) {
      wf = wf * Ctrb;
      wr = wr * Ctrb;
    }
    // mass action (CHEM_reactions.f90:246-250).  k is uniform across the
    // wavefront, so the stoichiometry tests below are scalar branches and
    // the fully unrolled i/j loops keep ri and A in registers.
    double wf1 = wf, wr1 = wr;
#pragma unroll
    for (int i = 0; i < CH_NS; i++) {
      if (m->vf[k][i] != 0) wf1 = wf1 * ipow(ci[i], m->vf[k][i]);
      if (m->vr[k][i] != 0) wr1 = wr1 * ipow(ci[i], m->vr[k][i]);
    }
#pragma unroll
    for (int i = 0; i < CH_NS; i++) {
      const int nu = m->vr[k][i] - m->vf[k][i];
      if (nu != 0) ri[i] = ri[i] + (wf1 - wr1) * nu;
    }
    if (JAC) {
#pragma unroll
      for (int j = 0; j < CH_NS; j++) {
        if (m->vf[k][j] == 0 && m->vr[k][j] == 0) continue;
        const double Awf = m->vf[k][j] * wf1 / (ci[j] + 1.0e-20);
        const double Awr = m->vr[k][j] * wr1 / (ci[j] + 1.0e-20);
#pragma unroll
        for (int i = 0; i < CH_NS; i++) {
          const int nu = m->vr[k][i] - m->vf[k][i];
          if (nu != 0) A[i][j] = A[i][j] + (Awf - Awr) * nu;
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
__global__ __launch_bounds__(128) void k_chem_queue(long n, double* rho, double* Ecell, const double* dtcell,
                                                     ChemCtl ctl, int* nsub_out, unsigned long long* queue) {
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
        if (my < (unsigned long long)n) {
          cell = (long)my;
#pragma unroll
          for (int i = 0; i < CH_NS; i++) di[i] = rho[(long)i * n + cell];
          E = Ecell[cell];
          dt = dtcell[cell];
          t = 0.0;
          h_cap = dt;
          nsub = 0;
          nrej = 0;
        } else {
          cell = -2;
        }
      }
    }
    if (__ballot(cell >= 0) == 0) break;
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
      }
    }
  }
}

__global__ void k_chem_reset(unsigned long long* queue) {
  if (threadIdx.x < 2 && blockIdx.x == 0) queue[threadIdx.x] = 0;
}

// ===========================================================================
extern "C" {

const char* ocfd_chem_last_error(void) { return CX.err.c_str(); }

int32_t ocfd_chem_set_mechanism(int32_t ns, int32_t nr, int32_t rereac, const int32_t* af_type,
                                const int32_t* third_body, const double* fc_troe, const double* arrhenius,
                                const int32_t* vf, const int32_t* vr, const double* eff, int32_t n_tct,
                                const double* tct, double tct_max, const double* mi, const double* coef) {
  if (ns != CH_NS) return cfail("ocfd_chem_set_mechanism: this build is compiled for 9 species");
  if (nr > CH_NR_MAX || n_tct > CH_NT_MAX || n_tct < 1) return cfail("ocfd_chem_set_mechanism: too many reactions or ranges");
  if (rereac != 0) return cfail("ocfd_chem_set_mechanism: only ReReac=0 (Gibbs equilibrium constants) is supported");
  ChemMech& h = CX.h;
  h = ChemMech{};
  h.ns = ns;
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
    if (hipMalloc(&CX.counter, 2 * sizeof(unsigned long long)) != hipSuccess) return cfail("hipMalloc failed");
  }
  CX.stream = CX.own;
  // kernels still in flight read the constant bank being replaced
  if (hipDeviceSynchronize() != hipSuccess) return cfail("ocfd_chem_set_mechanism: device error");
  if (hipMemcpyToSymbol(HIP_SYMBOL(cm), &h, sizeof(ChemMech)) != hipSuccess) return cfail("hipMemcpyToSymbol failed");
  CX.ready = true;
  return 0;
}

// Chemical_time_advance over a batch of independent cells (OCFD_Time_Adv.f90:169-196
// calls chemical_timeAdv per cell); device pointers; nsub substeps of dt/nsub.
int32_t ocfd_chem_timeadv_batch(int64_t n, double* rho, double* E, const double* dt, int32_t mode, int32_t nsub) {
  if (!CX.ready) return cfail("ocfd_chem_timeadv_batch: call ocfd_chem_set_mechanism first");
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
  if (!CX.ready) return cfail("ocfd_chem_adaptive_batch: call ocfd_chem_set_mechanism first");
  if (max_sub < 1 || !(rtol_y > 0.0) || !(y_floor >= 0.0) || !(rtol_t > 0.0))
    return cfail("ocfd_chem_adaptive_batch: bad controller parameters");
  const ChemCtl ctl{rtol_y, y_floor, rtol_t, (int)max_sub};
  hipLaunchKernelGGL(k_chem_reset, dim3(1), dim3(64), 0, CX.stream, CX.counter);
  if (n <= 0) return hipGetLastError() == hipSuccess ? 0 : cfail("ocfd_chem_adaptive_batch: launch failed");
  if (scheduler == 0) {
    hipLaunchKernelGGL(k_chem_adaptive, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, CX.stream, (long)n, rho, E, dt,
                       ctl, (int*)nsub_out, CX.counter);
  } else {
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_chem_queue, 128, 0) != hipSuccess)
      return cfail("ocfd_chem_adaptive_batch: device query failed");
    if (per < 1) per = 1;
    long blocks = (long)ncu * per;
    const long need = (n + 127) / 128;
    if (blocks > need) blocks = need;
    hipLaunchKernelGGL(k_chem_queue, dim3((unsigned)blocks), dim3(128), 0, CX.stream, (long)n, rho, E, dt, ctl,
                       (int*)nsub_out, CX.counter);
  }
  if (hipGetLastError() != hipSuccess) return cfail("ocfd_chem_adaptive_batch: launch failed");
  return 0;
}

int32_t ocfd_chem_synchronize(void) {
  if (!CX.stream) return 0;
  if (hipStreamSynchronize(CX.stream) != hipSuccess) return cfail("ocfd_chem_synchronize: stream error");
  return 0;
}

void* ocfd_chem_stream(void) { return (void*)CX.stream; }

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
