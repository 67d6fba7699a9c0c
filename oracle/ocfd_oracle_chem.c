/*
 * ocfd_oracle_chem.c -- CPU restatement of OpenCFD-Comb's finite-rate
 * chemistry (TEST INFRASTRUCTURE ONLY; see ocfd_oracle.c).
 *
 * Restates code/Chemical/CHEM_reactions.f90 and CHEM_thermdyn.f90: NASA-7-form
 * thermo with Newton T(E), Arrhenius / third-body / Troe fall-off / dual
 * Arrhenius rate constants, reverse rates from Gibbs equilibrium constants,
 * the analytic rate Jacobian, Gauss elimination with partial pivoting, and
 * the explicit Euler / RK2 / RK3 and point-implicit Euler integrators of
 * chemical_timeAdv.  Reverse rates: Gibbs equilibrium constants (ReReac = 0),
 * Arrhenius fits (1) or the Park curve fit of Kc (2).  Built with -ffp-contract=off; transcendentals are glibc's.
 *
 * Pinned by tests/test_oracle_chem.py: element and mass conservation per
 * step, Gauss == numpy.linalg.solve, Arrhenius / Troe / K_c against closed
 * forms evaluated independently in numpy from the NASA coefficients, detailed
 * balance at equilibrium, the point-implicit step's exact solution of a
 * linear test system, and Newton T(E) round trips.
 *
 * Also restated here, as the definition the GPU kernels are checked against:
 * the adaptive per-cell substep controller (a build addition with no
 * reference counterpart, see chem_adaptive below).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NS_MAX 16
#define NR_MAX 64
#define NT_MAX 10
#define R0 8.314
#define ATM 1.013e5

static struct {
  int ns, nr, rereac;
  int af_type[NR_MAX], tb[NR_MAX], sgm[NR_MAX];
  double fc[NR_MAX], arr[NR_MAX][2][3], eff[NR_MAX][NS_MAX];
  double rev[NR_MAX][5];  /* ReReac = 1: Ar, br, Er; ReReac = 2: A1r..A5r */
  int rev_set;
  int vf[NR_MAX][NS_MAX], vr[NR_MAX][NS_MAX];
  int n_tct;
  double tct[NT_MAX], tct_max;
  double mi[NS_MAX], ri[NS_MAX];
  double coef[NS_MAX][NT_MAX][7];
  double ect[NS_MAX][NT_MAX];
} C;

static inline double p2(double x) { return x * x; }
static inline double p3(double x) { double x2 = x * x; return x2 * x; }
static inline double p4(double x) { double x2 = x * x; return x2 * x2; }
static inline double p5(double x) { double x2 = x * x; return (x2 * x2) * x; }
/* x**n for an integer n as LLVM/GCC expand it: repeated squaring, negative
 * powers as 1/x**|n| */
static double ipow(double x, int n) {
  if (n == 0) return 1.0;
  int m = n < 0 ? -n : n;
  double r = 1.0, b = x;
  int first = 1;
  while (m) {
    if (m & 1) { r = first ? b : r * b; first = 0; }
    m >>= 1;
    if (m) b = b * b;
  }
  return n < 0 ? 1.0 / r : r;
}

int orc_chem_set_mechanism(int ns, int nr, int rereac, const int* af_type, const int* tb, const double* fc,
                           const double* arr, const int* vf, const int* vr, const double* eff, int n_tct,
                           const double* tct, double tct_max, const double* mi, const double* coef) {
  if (ns > NS_MAX || nr > NR_MAX || n_tct > NT_MAX || rereac < 0 || rereac > 2) return 1;
  C.ns = ns; C.nr = nr; C.rereac = rereac;
  C.rev_set = 0;
  for (int j = 0; j < nr; j++) {
    C.af_type[j] = af_type[j]; C.tb[j] = tb[j]; C.fc[j] = fc[j];
    for (int a = 0; a < 2; a++) for (int b = 0; b < 3; b++) C.arr[j][a][b] = arr[(j * 2 + a) * 3 + b];
    C.sgm[j] = 0;
    for (int i = 0; i < ns; i++) {
      C.vf[j][i] = vf[j * ns + i]; C.vr[j][i] = vr[j * ns + i]; C.eff[j][i] = eff[j * ns + i];
      C.sgm[j] = C.sgm[j] + C.vr[j][i] - C.vf[j][i];   /* read_Reac, CHEM_reactions.f90:65-68 */
    }
  }
  C.n_tct = n_tct; C.tct_max = tct_max;
  for (int j = 0; j < n_tct - 1; j++) C.tct[j] = tct[j];
  for (int k = 0; k < ns; k++) {
    C.mi[k] = mi[k];
    C.ri[k] = R0 / mi[k];
    for (int j = 0; j < n_tct; j++) for (int c = 0; c < 7; c++) C.coef[k][j][c] = coef[(k * n_tct + j) * 7 + c];
    /* enthalpy continuity offsets, read_Spec CHEM_thermdyn.f90:43-47 */
    C.ect[k][0] = 0.0;
    for (int j = 0; j < n_tct - 1; j++) {
      double t = C.tct[j];
      C.ect[k][j + 1] = -((C.coef[k][j + 1][0] - C.coef[k][j][0]) * t + (C.coef[k][j + 1][1] - C.coef[k][j][1]) * p2(t) / 2.0 +
                          (C.coef[k][j + 1][2] - C.coef[k][j][2]) * p3(t) / 3.0 + (C.coef[k][j + 1][3] - C.coef[k][j][3]) * p4(t) / 4.0 +
                          (C.coef[k][j + 1][4] - C.coef[k][j][4]) * p5(t) / 5.0) + C.ect[k][j];
    }
  }
  return 0;
}

/* Reverse-rate data of ReReac = 1 or 2 (the extra line per reaction that
 * read_Reac reads, CHEM_reactions.f90:46-47): rev[j*5 .. j*5+4]. */
int orc_chem_set_reverse(const double* rev) {
  for (int j = 0; j < C.nr; j++)
    for (int a = 0; a < 5; a++) C.rev[j][a] = rev[j * 5 + a];
  C.rev_set = 1;
  return 0;
}

static int range_of(double T) {
  int jc = 0;
  for (int j = 0; j < C.n_tct - 1; j++) if (T > C.tct[j]) jc = j + 1;
  return jc;
}

/* comput_E_rgas, CHEM_thermdyn.f90:203-231 */
double orc_chem_E(const double* di, double T) {
  double A1 = 0, B1 = 0, C1 = 0, D1 = 0, E1 = 0, Ect = 0;
  int jc = range_of(T);
  for (int k = 0; k < C.ns; k++) {
    double R1 = C.ri[k];
    A1 = A1 + (C.coef[k][jc][0] - 1.0) * R1 * di[k];
    B1 = B1 + C.coef[k][jc][1] * R1 * di[k];
    C1 = C1 + C.coef[k][jc][2] * R1 * di[k];
    D1 = D1 + C.coef[k][jc][3] * R1 * di[k];
    E1 = E1 + C.coef[k][jc][4] * R1 * di[k];
    Ect = Ect + C.ect[k][jc] * R1 * di[k];
  }
  if (T <= C.tct_max) return A1 * T + B1 * p2(T) / 2.0 + C1 * p3(T) / 3.0 + D1 * p4(T) / 4.0 + E1 * p5(T) / 5.0 + Ect;
  double tm = C.tct_max;
  double Cv2 = A1 + B1 * tm + C1 * p2(tm) + D1 * p3(tm) + E1 * p4(tm);
  return A1 * tm + B1 * p2(tm) / 2.0 + C1 * p3(tm) / 3.0 + D1 * p4(tm) / 4.0 + E1 * p5(tm) / 5.0 + Ect + Cv2 * (T - tm);
}

/* comput_T_rgas, CHEM_thermdyn.f90:326-382 (Newton, 1e-3 K, <= 31 iterations) */
double orc_chem_T(const double* di, double E) {
  double A1[NT_MAX], B1[NT_MAX], C1[NT_MAX], D1[NT_MAX], E1[NT_MAX], Ect[NT_MAX];
  for (int j = 0; j < C.n_tct; j++) {
    A1[j] = B1[j] = C1[j] = D1[j] = E1[j] = Ect[j] = 0.0;
    for (int k = 0; k < C.ns; k++) {
      double R1 = C.ri[k];
      A1[j] = A1[j] + (C.coef[k][j][0] - 1.0) * R1 * di[k];
      B1[j] = B1[j] + C.coef[k][j][1] * R1 * di[k];
      C1[j] = C1[j] + C.coef[k][j][2] * R1 * di[k];
      D1[j] = D1[j] + C.coef[k][j][3] * R1 * di[k];
      E1[j] = E1[j] + C.coef[k][j][4] * R1 * di[k];
      Ect[j] = Ect[j] + C.ect[k][j] * R1 * di[k];
    }
  }
  double T1 = E / A1[0], T2 = T1;
  int ks = 0;
  while (ks <= 30) {
    int jc = range_of(T1);
    double Et1, Ex;
    if (T1 <= C.tct_max) {
      Et1 = A1[jc] * T1 + B1[jc] * p2(T1) / 2.0 + C1[jc] * p3(T1) / 3.0 + D1[jc] * p4(T1) / 4.0 + E1[jc] * p5(T1) / 5.0 + Ect[jc];
      Ex = A1[jc] + B1[jc] * T1 + C1[jc] * p2(T1) + D1[jc] * p3(T1) + E1[jc] * p4(T1);
    } else {
      double tm = C.tct_max;
      Ex = A1[jc] + B1[jc] * tm + C1[jc] * p2(tm) + D1[jc] * p3(tm) + E1[jc] * p4(tm);
      Et1 = A1[jc] * tm + B1[jc] * p2(tm) / 2.0 + C1[jc] * p3(tm) / 3.0 + D1[jc] * p4(tm) / 4.0 + E1[jc] * p5(tm) / 5.0 + Ect[jc] +
            Ex * (T1 - tm);
    }
    T2 = T1 - (Et1 - E) / Ex;
    ks = ks + 1;
    double error = fabs(T2 - T1);
    T1 = T2;
    if (error < 1.0e-3) break;
  }
  return T2;
}

/* comput_gibbs, CHEM_thermdyn.f90:296-319 (gi = G_i/R_i) */
double orc_chem_gibbs(int k, double T) {
  int jc = range_of(T);
  const double* c = C.coef[k][jc];
  if (T <= C.tct_max)
    return c[0] * (1.0 - log(T)) * T - c[1] * p2(T) / 2.0 - c[2] * p3(T) / 6.0 - c[3] * p4(T) / 12.0 - c[4] * p5(T) / 20.0 + c[5] -
           c[6] * T;
  double tm = C.tct_max;
  double Cp_max = c[0] + c[1] * tm + c[2] * p2(tm) + c[3] * p3(tm) + c[4] * p4(tm);
  double h_max = c[0] * tm + c[1] * p2(tm) / 2.0 + c[2] * p3(tm) / 3.0 + c[3] * p4(tm) / 4.0 + c[4] * p5(tm) / 5.0 + c[5];
  double S_max = c[0] * log(tm) + c[1] * tm + c[2] * p2(tm) / 2.0 + c[3] * p3(tm) / 3.0 + c[4] * p4(tm) / 4.0 + c[6];
  return h_max + Cp_max * (T - tm) - T * (S_max + Cp_max * log(T / tm));
}

/* comput_KjX, CHEM_reactions.f90:76-103 */
static void comput_kjx(double T, double* kjx) {
  const double Rs = R0 / ATM * 1.0e6;
  double gi[NS_MAX];
  for (int i = 0; i < C.ns; i++) gi[i] = orc_chem_gibbs(i, T);
  for (int j = 0; j < C.nr; j++) {
    double dgi = 0.0;
    for (int i = 0; i < C.ns; i++) dgi = dgi + gi[i] * (C.vr[j][i] - C.vf[j][i]);
    kjx[j] = exp(-dgi / T) * ipow(Rs * T, -C.sgm[j]);
  }
}

/* Arrhenius, CHEM_reactions.f90:106-111 */
static double arrhenius(double A, double b, double E, double T) { return A * pow(T, b) * exp(-E / T); }

/* ParkCurveFit, CHEM_reactions.f90:113-118: Kc = exp(A1 + A2 z + A3 z^2 +
 * A4 z^3 + A5 z^4), z = 10000/T */
static double park_curve_fit(const double* a, double T) {
  double z = 10000.0 / T;
  return exp(a[0] + a[1] * z + a[2] * p2(z) + a[3] * p3(z) + a[4] * p4(z));
}

/* Reaction_rate_coef, CHEM_reactions.f90:130-197: reverse rates from the
 * Gibbs equilibrium constant (ReReac = 0), an Arrhenius fit (1) or
 * k_f / Kc with the Park fit of Kc floored at 1e-30 (2). */
int orc_chem_rate_coef(const double* ci, double T, double* wf, double* wr, double* kjx) {
  double kj[NR_MAX];
  if (C.rereac != 0 && !C.rev_set) return 1;
  if (C.rereac == 0) comput_kjx(T, kj);
  else
    for (int j = 0; j < C.nr; j++) kj[j] = 0.0;
  for (int j = 0; j < C.nr; j++) {
    double Ctrb = 0.0;
    if (C.tb[j] == 1)
      for (int i = 0; i < C.ns; i++) Ctrb = Ctrb + ci[i] * C.eff[j][i];
    double wf1 = arrhenius(C.arr[j][0][0], C.arr[j][0][1], C.arr[j][0][2], T);
    if (C.af_type[j] == 0) {
      wf[j] = wf1;
    } else if (C.af_type[j] == 2) {
      double wf2 = arrhenius(C.arr[j][1][0], C.arr[j][1][1], C.arr[j][1][2], T);
      wf[j] = wf1 + wf2;
    } else {
      /* Troe fall-off; 0.67 and 1.27 are REAL(4) literals (CHEM_reactions.f90:170-173) */
      double wf2 = arrhenius(C.arr[j][1][0], C.arr[j][1][1], C.arr[j][1][2], T);
      double Prs = wf1 / wf2 * Ctrb;
      double Fc = C.fc[j];
      double ac = -0.4 - (double)0.67f * log10(Fc);
      double an = 0.75 - (double)1.27f * log10(Fc);
      double F = pow(10.0, log10(Fc) / (1.0 + p2((log10(Prs) + ac) / (an - 0.14 * (log10(Prs) + ac)))));
      wf[j] = wf2 * Prs / (1.0 + Prs) * F;
    }
    if (C.rereac == 0) {
      wr[j] = wf[j] / kj[j];
    } else if (C.rereac == 1) {
      wr[j] = arrhenius(C.rev[j][0], C.rev[j][1], C.rev[j][2], T);
    } else {
      kj[j] = park_curve_fit(C.rev[j], T);
      if (kj[j] < 1.0e-30) kj[j] = 1.0e-30;
      wr[j] = wf[j] / kj[j];
    }
    if (C.tb[j] == 1 && C.af_type[j] != 1) {
      wf[j] = wf[j] * Ctrb;
      wr[j] = wr[j] * Ctrb;
    }
    if (kjx) kjx[j] = kj[j];
  }
  return 0;
}

/* Reaction_rate (CHEM_reactions.f90:202-226) and, when A is given,
 * Reaction_rate_and_Matrix (:230-272): ri (mol/cm3/s), A(i,j)=d ri/d cj
 * stored A[i*ns+j] */
int orc_chem_rates(const double* ci, double T, double* ri, double* A) {
  double wf[NR_MAX], wr[NR_MAX];
  int ns = C.ns;
  if (orc_chem_rate_coef(ci, T, wf, wr, NULL)) return 1;
  for (int i = 0; i < ns; i++) ri[i] = 0.0;
  if (A) for (int i = 0; i < ns * ns; i++) A[i] = 0.0;
  for (int k = 0; k < C.nr; k++) {
    double wf1 = wf[k], wr1 = wr[k];
    for (int i = 0; i < ns; i++) {
      if (C.vf[k][i] != 0) wf1 = wf1 * ipow(ci[i], C.vf[k][i]);
      if (C.vr[k][i] != 0) wr1 = wr1 * ipow(ci[i], C.vr[k][i]);
    }
    for (int i = 0; i < ns; i++) ri[i] = ri[i] + (wf1 - wr1) * (C.vr[k][i] - C.vf[k][i]);
    if (A) {
      for (int j = 0; j < ns; j++) {
        double Awf = C.vf[k][j] * wf1 / (ci[j] + 1.0e-20);
        double Awr = C.vr[k][j] * wr1 / (ci[j] + 1.0e-20);
        for (int i = 0; i < ns; i++) A[i * ns + j] = A[i * ns + j] + (Awf - Awr) * (C.vr[k][i] - C.vf[k][i]);
      }
    }
  }
  return 0;
}

/* Gauss, CHEM_reactions.f90:523-567 (A row-major a[i*n+j], overwritten) */
int orc_gauss(int n, double* a, double* b, double* x) {
  for (int k = 0; k < n - 1; k++) {
    double fmax = fabs(a[k * n + k]);
    int kk = k;
    for (int i = k + 1; i < n; i++)
      if (fabs(a[i * n + k]) > fmax) { fmax = fabs(a[i * n + k]); kk = i; }
    for (int j = k; j < n; j++) { double t = a[k * n + j]; a[k * n + j] = a[kk * n + j]; a[kk * n + j] = t; }
    { double t = b[k]; b[k] = b[kk]; b[kk] = t; }
    for (int i = k + 1; i < n; i++) {
      double f = -a[i * n + k] / a[k * n + k];
      for (int j = 0; j < n; j++) a[i * n + j] = a[i * n + j] + f * a[k * n + j];
      b[i] = b[i] + f * b[k];
    }
  }
  x[n - 1] = b[n - 1] / a[(n - 1) * n + n - 1];
  for (int i = n - 2; i >= 0; i--) {
    x[i] = b[i];
    for (int j = i + 1; j < n; j++) x[i] = x[i] - a[i * n + j] * x[j];
    x[i] = x[i] / a[i * n + i];
  }
  return 0;
}

/* update_Et, CHEM_thermdyn.f90:184-196 */
static void update_et(const double* di, const double* dinew, double* E) {
  for (int i = 0; i < C.ns; i++) *E = *E - (dinew[i] - di[i]) * C.ri[i] * C.coef[i][0][5];
}

/* chemical_Euler_implicit, CHEM_reactions.f90:473-516 */
static int euler_implicit(double* di, double* E, double dt) {
  int ns = C.ns;
  double ci[NS_MAX], ri[NS_MAX], Ar[NS_MAX * NS_MAX], A1[NS_MAX * NS_MAX], di1[NS_MAX], dd[NS_MAX], dinew[NS_MAX];
  double T = orc_chem_T(di, *E);
  for (int i = 0; i < ns; i++) ci[i] = di[i] / C.mi[i] * 1.0e-6;
  if (orc_chem_rates(ci, T, ri, Ar)) return 1;
  for (int j = 0; j < ns; j++)
    for (int i = 0; i < ns; i++) {
      double a = (i == j) ? 1.0 : 0.0;
      A1[i * ns + j] = a - Ar[i * ns + j] * C.mi[i] / C.mi[j] * dt;
    }
  for (int i = 0; i < ns; i++) di1[i] = ri[i] * C.mi[i] * 1.0e6 * dt;
  orc_gauss(ns, A1, di1, dd);
  for (int i = 0; i < ns; i++) dinew[i] = di[i] + dd[i];
  update_et(di, dinew, E);
  for (int i = 0; i < ns; i++) di[i] = dinew[i];
  return 0;
}

/* chemical_Euler, CHEM_reactions.f90:395-425 */
static int euler_explicit(double* di, double* E, double dt) {
  int ns = C.ns;
  double ci[NS_MAX], ri[NS_MAX], dinew[NS_MAX];
  double T = orc_chem_T(di, *E);
  for (int i = 0; i < ns; i++) ci[i] = di[i] / C.mi[i] * 1.0e-6;
  if (orc_chem_rates(ci, T, ri, NULL)) return 1;
  for (int i = 0; i < ns; i++) dinew[i] = di[i] + ri[i] * C.mi[i] * 1.0e6 * dt;
  update_et(di, dinew, E);
  for (int i = 0; i < ns; i++) di[i] = dinew[i];
  return 0;
}

/* chemical_RK2, CHEM_reactions.f90:429-469 (stage 2 evaluates T from the
 * old di with the updated E, as the reference does) */
static int rk2(double* di, double* E, double dt) {
  int ns = C.ns;
  double ci[NS_MAX], ri[NS_MAX], di1[NS_MAX], dinew[NS_MAX];
  double T = orc_chem_T(di, *E);
  for (int i = 0; i < ns; i++) ci[i] = di[i] / C.mi[i] * 1.0e-6;
  if (orc_chem_rates(ci, T, ri, NULL)) return 1;
  for (int i = 0; i < ns; i++) di1[i] = di[i] + ri[i] * C.mi[i] * 1.0e6 * dt;
  update_et(di, di1, E);
  T = orc_chem_T(di, *E);
  for (int i = 0; i < ns; i++) ci[i] = di1[i] / C.mi[i] * 1.0e-6;
  if (orc_chem_rates(ci, T, ri, NULL)) return 1;
  for (int i = 0; i < ns; i++) dinew[i] = 0.5 * di[i] + 0.5 * di1[i] + 0.5 * ri[i] * C.mi[i] * 1.0e6 * dt;
  update_et(di1, dinew, E);
  for (int i = 0; i < ns; i++) di[i] = dinew[i];
  return 0;
}

/* chemical_RK3, CHEM_reactions.f90:572-602 */
static int rk3(double* di, double* E, double dt) {
  int ns = C.ns;
  const double Ra[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0}, Rb[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
  double din[NS_MAX], dinew[NS_MAX], ci[NS_MAX], ri[NS_MAX];
  for (int i = 0; i < ns; i++) din[i] = di[i];
  for (int krk = 0; krk < 3; krk++) {
    double T = orc_chem_T(di, *E);
    for (int i = 0; i < ns; i++) ci[i] = di[i] / C.mi[i] * 1.0e-6;
    if (orc_chem_rates(ci, T, ri, NULL)) return 1;
    for (int i = 0; i < ns; i++) dinew[i] = Ra[krk] * din[i] + Rb[krk] * (di[i] + ri[i] * C.mi[i] * 1.0e6 * dt);
    update_et(di, dinew, E);
    for (int i = 0; i < ns; i++) di[i] = dinew[i];
  }
  return 0;
}

/* chemical_timeAdv, CHEM_reactions.f90:374-390 (mode = CHEM_TimeAdv) */
int orc_chem_timeadv(int mode, double* di, double* E, double dt) {
  switch (mode) {
    case 1: return euler_explicit(di, E, dt);
    case 2: return rk2(di, E, dt);
    case 3: return rk3(di, E, dt);
    case -1: return euler_implicit(di, E, dt);
  }
  return 1;
}

/* Chemical_time_advance (OCFD_Time_Adv.f90:169-196) over a batch: cell c
 * advances nsub substeps of dt[c]/nsub; rho is [n][ns], E [n]. */
int orc_chem_batch(int mode, long n, double* rho, double* E, const double* dt, int nsub) {
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
  for (long c = 0; c < n; c++) {
    double h = dt[c] / nsub;
    for (int s = 0; s < nsub; s++) bad |= orc_chem_timeadv(mode, rho + c * C.ns, E + c, h);
  }
  return bad;
}

/* ------------------------------------------------------------------------
 * Adaptive per-cell substepping (build addition, no reference counterpart).
 * The chemistry step dt of one cell is covered by point-implicit Euler
 * substeps (the reference's integrator, chemical_Euler_implicit).  Each
 * attempt evaluates T, the rates w_k = ri_k M_k 1e6 (kg/m3/s) and the
 * Jacobian at the current state.  With the per-species scale
 * s_k = rtol_y (max(rho_k, 0) + y_floor rho), the thermal scale
 * s_T = rtol_t T rcv (rcv = sum_k 2.5 R_k rho_k) and the heat release
 * q = sum_k w_k R_k F_k (F_k the formation coefficient of update_Et), it
 * proposes
 *     h = min(dt - t, h_cap, min_k s_k / |w_k|, s_T / |q|).
 * The relative species bound limits h to about rtol_y / lambda while
 * radicals grow like exp(lambda t) during ignition.  Negative partial
 * densities (undershoot of a consumed species) enter the rates and the
 * Jacobian as zero: with the reference's regularised Jacobian
 * d(ri)/d(cj) ~ 1/(cj + 1e-20), the point-implicit update of a consumed
 * species has a fixed point at cj = -1e-20 mol/cm3, where that denominator
 * vanishes.  The linear solve gives
 * dd; the attempt is accepted when
 *     |dd_k| <= 2 s_k for every k  and  |sum_k dd_k R_k F_k| <= 2 s_T.
 * A rejected attempt leaves the state unchanged and halves h_cap; an
 * accepted one applies the update (dinew = di + dd, update_Et) and sets
 * h_cap = 2h.  At most max_sub substeps are accepted and at most max_sub
 * attempts rejected, so the loop ends; a cell that has not reached dt by
 * then is "unfinished" (returned through *unfinished; its state is where
 * the integration stopped).  nsub counts accepted substeps.
 * ------------------------------------------------------------------------ */
int orc_chem_adaptive(double* di, double* E, double dt, double rtol_y, double y_floor, double rtol_t, int max_sub,
                      int* nsub_out, int* unfinished) {
  int ns = C.ns, nsub = 0, nrej = 0;
  double t = 0.0, h_cap = dt;
  while (t < dt && nsub < max_sub && nrej < max_sub) {
    double ci[NS_MAX], ri[NS_MAX], Ar[NS_MAX * NS_MAX], A1[NS_MAX * NS_MAX], rhs[NS_MAX], dd[NS_MAX], dinew[NS_MAX];
    double sk[NS_MAX];
    double T = orc_chem_T(di, *E);
    double rho = 0.0;
    for (int i = 0; i < ns; i++) {
      ci[i] = (di[i] > 0.0 ? di[i] : 0.0) / C.mi[i] * 1.0e-6;
      rho = rho + di[i];
    }
    if (orc_chem_rates(ci, T, ri, Ar)) return 1;
    double h = dt - t;
    if (h_cap < h) h = h_cap;
    double q = 0.0, rcv = 0.0;
    for (int i = 0; i < ns; i++) {
      double w = ri[i] * C.mi[i] * 1.0e6;
      sk[i] = rtol_y * ((di[i] > 0.0 ? di[i] : 0.0) + y_floor * rho);
      if (fabs(w) * h > sk[i]) h = sk[i] / fabs(w);
      q = q + w * C.ri[i] * C.coef[i][0][5];
      rcv = rcv + 2.5 * C.ri[i] * di[i];
    }
    double sT = rtol_t * T * rcv;
    if (fabs(q) * h > sT) h = sT / fabs(q);
    /* chemical_Euler_implicit's linear system with step h */
    for (int j = 0; j < ns; j++)
      for (int i = 0; i < ns; i++) A1[i * ns + j] = ((i == j) ? 1.0 : 0.0) - Ar[i * ns + j] * C.mi[i] / C.mi[j] * h;
    for (int i = 0; i < ns; i++) rhs[i] = ri[i] * C.mi[i] * 1.0e6 * h;
    orc_gauss(ns, A1, rhs, dd);
    int ok = 1;
    double de = 0.0;
    for (int i = 0; i < ns; i++) {
      ok = ok && fabs(dd[i]) <= 2.0 * sk[i];
      de = de + dd[i] * C.ri[i] * C.coef[i][0][5];
    }
    ok = ok && fabs(de) <= 2.0 * sT;
    if (!ok) {
      h_cap = 0.5 * h;
      nrej++;
      continue;
    }
    for (int i = 0; i < ns; i++) dinew[i] = di[i] + dd[i];
    update_et(di, dinew, E);
    for (int i = 0; i < ns; i++) di[i] = dinew[i];
    t = t + h;
    nsub++;
    h_cap = 2.0 * h;
    if (dt - t <= 1e-12 * dt) break;
  }
  if (nsub_out) *nsub_out = nsub;
  if (unfinished) *unfinished = !(t >= dt || dt - t <= 1e-12 * dt);
  return 0;
}

int orc_chem_adaptive_batch(long n, double* rho, double* E, const double* dt, double rtol_y, double y_floor,
                            double rtol_t, int max_sub, int* nsub, int* unfinished) {
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : bad)
  for (long c = 0; c < n; c++)
    bad |= orc_chem_adaptive(rho + c * C.ns, E + c, dt[c], rtol_y, y_floor, rtol_t, max_sub, nsub + c,
                             unfinished ? unfinished + c : NULL);
  return bad;
}

/* Threads used by the batch functions (OpenMP); returns the count in use. */
int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

