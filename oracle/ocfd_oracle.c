/*
 * ocfd_oracle.c -- CPU restatement of OpenCFD-Comb's per-step flow path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against; it is linked only by tests/, by the smoke check and by the
 * CPU-baseline leg of the benchmark.  Nothing in the product calls it.
 *
 * It restates the reference Fortran (paths relative to the OpenCFD-Comb tree)
 * routine by routine, keeping the reference's loop structure (one pencil at a
 * time, hh(0:n) interface arrays, module-global state) and its floating-point
 * evaluation order (Fortran left-to-right, x**2 = x*x, no FMA contraction:
 * built with -ffp-contract=off).  The reference itself cannot be built in this
 * image (every module includes mpif.h and MPI is absent), so each function is
 * pinned by the mathematical checks in tests/test_oracle_*.py.
 *
 * Pinning status per function (what checks it):
 *   weno5/weno7 P,M ........ constant/linear exactness, mirror identity,
 *                             smooth-limit = linear optimal scheme, order
 *   omp6 P,M, ud7l ......... linear-branch closed form, polynomial exactness
 *   boundary closures ...... polynomial exactness of the one-sided formulas
 *   split SW / LLF ......... F+ + F- = physical flux, supersonic F- = 0
 *   dx0/dy0/dz0 ............ exactness on polynomials (CD6/CD8 + closures)
 *   du_inviscous ........... uniform flow -> 0, 1-D advection closed form
 *   du_viscous ............. Couette/linear-T closed forms, zero on uniform
 *   RK3 .................... du = lambda*f reproduces 1+z+z^2/2+z^3/6
 *   flow variables / bc .... perfect-gas closed forms, round trips
 *   metrics ................ uniform / stretched 1-D mesh closed forms
 *   flow_time_advance ...... conservation on a periodic box
 *   hybrid sensor type 1/2 . closed forms on analytic pressure fields
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/ocfd_hip.h"

#define NTCT_MAX 10
#define R0_GAS 8.314          /* CHEM_modules.f90:20 */
#define ATM_PA 1.013e5        /* CHEM_modules.f90:22 */

/* ------------------------------------------------------------------ */
/* module globals (MODULE_parameter.f90, MODULE_flow.f90, CHEM_modules.f90) */
static struct {
  int nx, ny, nz, lap, nxh, nyh, nzh, nspec;
  long nh, ni;                       /* halo / interior point counts */
  ocfd_cfg cfg;
  double hx, hy, hz;
  int Ka1, Kb1, Ka2, Kb2;
  int bound_index[3][2];
  double ud7l[8];
  /* flow data */
  double *f, *fn, *du, *di, *din;
  double *d, *u, *v, *w, *T, *p, *Et, *cc, *Amu, *Amk, *AmD;
  double *Axx, *Ayy, *Azz;
  double *Akx, *Aky, *Akz, *Aix, *Aiy, *Aiz, *Asx, *Asy, *Asz, *Ajac;
  double *Akx1, *Aky1, *Akz1, *Aix1, *Aiy1, *Aiz1, *Asx1, *Asy1, *Asz1;
  double *Rhybrid;
  unsigned char *maskx, *masky, *maskz;
  double *Ev1, *Ev2, *Ev3;
  double *uk, *ui, *us, *vk, *vi, *vs, *wk, *wi, *ws, *Tk, *Ti, *Ts;
  double *Qpx, *Qpy, *Qpz;
  double *dus;                       /* species RHS scratch (the reference's du(:,:,:,1)) */
  /* transport fits (TRANSPORT_modules.f90): mu_an, k_an [k][4], D_an [k][j][4] */
  double *mu_an, *k_an, *D_an;
  /* jet BC: flow_inlet2d(j,k,m), m = d,u,v,w,T,di(1..N_SPEC) */
  double *inlet2d;
  /* species */
  int n_tct;
  double tct[NTCT_MAX], tct_max;
  double *Mi, *Ri;
  double *coef;                      /* [k][j][7] */
  double *Ect;                       /* [k][j]    */
  /* boundary-layer BC */
  double *bl_inlet, *bl_upper, *bl_wall_pertb;  /* BC%flow_inlet*d, BC%flow_upper*d, BC%wall_pertb */
  long bl_nin, bl_nup;
  int *bl_nonref;
  double bl_ht;
  int num_negt;
  double sensor_dp_av;               /* last global mean of |grad p| */
  char err[256];
} G;

/* index helpers: halo (1-LAP..n+LAP) and interior (1..n) */
#define H(i, j, k) ((long)((i) + G.lap - 1) + (long)G.nxh * ((long)((j) + G.lap - 1) + (long)G.nyh * (long)((k) + G.lap - 1)))
#define I(i, j, k) ((long)((i) - 1) + (long)G.nx * ((long)((j) - 1) + (long)G.ny * (long)((k) - 1)))
#define SQ(x) ((x) * (x))
/* Constant integer powers as LLVM/GCC expand x**n (repeated squaring):
 * x**3 = (x*x)*x, x**4 = (x*x)*(x*x), x**5 = ((x*x)*(x*x))*x. */
static inline double pw3(double x) { double x2 = x * x; return x2 * x; }
static inline double pw4(double x) { double x2 = x * x; return x2 * x2; }
static inline double pw5(double x) { double x2 = x * x; return (x2 * x2) * x; }

static int set_err(const char* m) { snprintf(G.err, sizeof G.err, "%s", m); return 1; }
const char* orc_last_error(void) { return G.err; }

static double* dalloc(long n, double v0) {
  double* a = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (long i = 0; i < n; i++) a[i] = v0;
  return a;
}

/* ------------------------------------------------------------------ */
/* Thermodynamics (CHEM_thermdyn.f90) */
static double sp_c(int k, int j, int c) { return G.coef[((long)k * NTCT_MAX + j) * 7 + c]; }

/* comput_E_pgas, CHEM_thermdyn.f90:114-122 */
static double comput_E(const double* di, double T) {
  if (!G.cfg.if_real_gas) {
    double E = 0.0;
    for (int k = 0; k < G.nspec; k++) E = E + 2.5 * G.Ri[k] * di[k] * T;
    return E;
  }
  /* comput_E_rgas, CHEM_thermdyn.f90:203-231 */
  double A1 = 0, B1 = 0, C1 = 0, D1 = 0, E1 = 0, Ect = 0, E;
  int jc = 0;
  for (int j = 0; j < G.n_tct - 1; j++) if (T > G.tct[j]) jc = j + 1;
  for (int k = 0; k < G.nspec; k++) {
    double R1 = G.Ri[k];
    A1 = A1 + (sp_c(k, jc, 0) - 1.0) * R1 * di[k];
    B1 = B1 + sp_c(k, jc, 1) * R1 * di[k];
    C1 = C1 + sp_c(k, jc, 2) * R1 * di[k];
    D1 = D1 + sp_c(k, jc, 3) * R1 * di[k];
    E1 = E1 + sp_c(k, jc, 4) * R1 * di[k];
    Ect = Ect + G.Ect[k * NTCT_MAX + jc] * R1 * di[k];
  }
  if (T <= G.tct_max) {
    E = A1 * T + B1 * (T * T) / 2.0 + C1 * pw3(T) / 3.0 + D1 * pw4(T) / 4.0 +
        E1 * pw5(T) / 5.0 + Ect;
  } else {
    double tm = G.tct_max;
    double Cv2 = A1 + B1 * tm + C1 * (tm * tm) + D1 * pw3(tm) + E1 * pw4(tm);
    E = A1 * tm + B1 * (tm * tm) / 2.0 + C1 * pw3(tm) / 3.0 + D1 * pw4(tm) / 4.0 +
        E1 * pw5(tm) / 5.0 + Ect + Cv2 * (T - tm);
  }
  return E;
}

/* comput_T_pgas (CHEM_thermdyn.f90:157-168), comput_T_rgas (:326-382) */
static double comput_T(const double* di, double E) {
  if (!G.cfg.if_real_gas) {
    double rCv = 0.0;
    for (int k = 0; k < G.nspec; k++) rCv = rCv + 2.5 * G.Ri[k] * di[k];
    return E / rCv;
  }
  double A1[NTCT_MAX], B1[NTCT_MAX], C1[NTCT_MAX], D1[NTCT_MAX], E1[NTCT_MAX], Ect[NTCT_MAX];
  for (int j = 0; j < G.n_tct; j++) {
    A1[j] = B1[j] = C1[j] = D1[j] = E1[j] = Ect[j] = 0.0;
    for (int k = 0; k < G.nspec; k++) {
      double R1 = G.Ri[k];
      A1[j] = A1[j] + (sp_c(k, j, 0) - 1.0) * R1 * di[k];
      B1[j] = B1[j] + sp_c(k, j, 1) * R1 * di[k];
      C1[j] = C1[j] + sp_c(k, j, 2) * R1 * di[k];
      D1[j] = D1[j] + sp_c(k, j, 3) * R1 * di[k];
      E1[j] = E1[j] + sp_c(k, j, 4) * R1 * di[k];
      Ect[j] = Ect[j] + G.Ect[k * NTCT_MAX + j] * R1 * di[k];
    }
  }
  double T1 = E / A1[0], T2 = T1;
  int ks = 0;
  while (ks <= 30) {
    int jc = 0;
    for (int j = 0; j < G.n_tct - 1; j++) if (T1 > G.tct[j]) jc = j + 1;
    double Et1, Ex;
    if (T1 <= G.tct_max) {
      Et1 = A1[jc] * T1 + B1[jc] * (T1 * T1) / 2.0 + C1[jc] * pw3(T1) / 3.0 +
            D1[jc] * pw4(T1) / 4.0 + E1[jc] * pw5(T1) / 5.0 + Ect[jc];
      Ex = A1[jc] + B1[jc] * T1 + C1[jc] * (T1 * T1) + D1[jc] * pw3(T1) + E1[jc] * pw4(T1);
    } else {
      double tm = G.tct_max;
      Ex = A1[jc] + B1[jc] * tm + C1[jc] * (tm * tm) + D1[jc] * pw3(tm) + E1[jc] * pw4(tm);
      Et1 = A1[jc] * tm + B1[jc] * (tm * tm) / 2.0 + C1[jc] * pw3(tm) / 3.0 +
            D1[jc] * pw4(tm) / 4.0 + E1[jc] * pw5(tm) / 5.0 + Ect[jc] + Ex * (T1 - tm);
    }
    T2 = T1 - (Et1 - E) / Ex;
    ks = ks + 1;
    double error = fabs(T2 - T1);
    T1 = T2;
    if (error < 1.0e-3) break;
  }
  return T2;
}

/* comput_P, CHEM_thermdyn.f90:385-395 */
static double comput_P(const double* di, double T) {
  double p = 0.0;
  for (int k = 0; k < G.nspec; k++) p = p + G.Ri[k] * di[k] * T;
  return p;
}

/* comput_d, CHEM_thermdyn.f90:399-409 */
static double comput_d(const double* ai, double T, double p) {
  double R1 = 0.0;
  for (int k = 0; k < G.nspec; k++) R1 = R1 + G.Ri[k] * ai[k];
  return p / (T * R1);
}

/* comput_Cp_pgas (CHEM_thermdyn.f90:125-136), comput_Cp_rgas (:234-266) */
static double comput_Cp(double d, const double* di, double T) {
  if (!G.cfg.if_real_gas) {
    double Cp = 0.0;
    for (int k = 0; k < G.nspec; k++) Cp = Cp + 3.5 * G.Ri[k] * di[k];
    return Cp / d;
  }
  double A1 = 0, B1 = 0, C1 = 0, D1 = 0, E1 = 0;
  int jc = 0;
  for (int j = 0; j < G.n_tct - 1; j++) if (T > G.tct[j]) jc = j + 1;
  for (int k = 0; k < G.nspec; k++) {
    double R1 = G.Ri[k];
    A1 = A1 + sp_c(k, jc, 0) * R1 * di[k];
    B1 = B1 + sp_c(k, jc, 1) * R1 * di[k];
    C1 = C1 + sp_c(k, jc, 2) * R1 * di[k];
    D1 = D1 + sp_c(k, jc, 3) * R1 * di[k];
    E1 = E1 + sp_c(k, jc, 4) * R1 * di[k];
  }
  double t = T <= G.tct_max ? T : G.tct_max;
  return (A1 + B1 * t + C1 * (t * t) + D1 * pw3(t) + E1 * pw4(t)) / d;
}

/* ------------------------------------------------------------------ */
/* Reconstructions (code/schemes/OCFD_Schemes.f90, OCFD_Schemes_1.f90).
 * v is indexed like the Fortran dummy v(1-LAP:n+LAP): V(i) = v[i+LAP-1]. */
#define V(i) v[(i) + lap - 1]

/* OCFD_weno5P, OCFD_Schemes.f90:6-25 */
static void weno5P(const double* v, double* hh, int n, int lap, int ib, int ie) {
  (void)n;
  const double ep = 1.0e-6, C03 = 3.0 / 10.0, C13 = 3.0 / 5.0, C23 = 1.0 / 10.0;
  for (int i = ib; i <= ie; i++) {
    double S0 = 13.0 / 12.0 * SQ(V(i) - 2.0 * V(i + 1) + V(i + 2)) + 1.0 / 4.0 * SQ(3.0 * V(i) - 4.0 * V(i + 1) + V(i + 2));
    double S1 = 13.0 / 12.0 * SQ(V(i - 1) - 2.0 * V(i) + V(i + 1)) + 1.0 / 4.0 * SQ(V(i - 1) - V(i + 1));
    double S2 = 13.0 / 12.0 * SQ(V(i - 2) - 2.0 * V(i - 1) + V(i)) + 1.0 / 4.0 * SQ(V(i - 2) - 4.0 * V(i - 1) + 3.0 * V(i));
    double a0 = C03 / SQ(ep + S0), a1 = C13 / SQ(ep + S1), a2 = C23 / SQ(ep + S2);
    double am = a0 + a1 + a2;
    double q03 = 1.0 / 3.0 * V(i) + 5.0 / 6.0 * V(i + 1) - 1.0 / 6.0 * V(i + 2);
    double q13 = -(1.0 / 6.0 * V(i - 1)) + 5.0 / 6.0 * V(i) + 1.0 / 3.0 * V(i + 1);
    double q23 = 1.0 / 3.0 * V(i - 2) - 7.0 / 6.0 * V(i - 1) + 11.0 / 6.0 * V(i);
    hh[i] = (a0 * q03 + a1 * q13 + a2 * q23) / am;
  }
}

/* OCFD_weno5M, OCFD_Schemes.f90:27-50 (writes hh(i-1), i=ib+1..ie+1) */
static void weno5M(const double* v, double* hh, int n, int lap, int ib, int ie) {
  (void)n;
  const double ep = 1.0e-6, C03 = 3.0 / 10.0, C13 = 3.0 / 5.0, C23 = 1.0 / 10.0;
  for (int i = ib + 1; i <= ie + 1; i++) {
    double S0 = 13.0 / 12.0 * SQ(V(i) - 2.0 * V(i - 1) + V(i - 2)) + 1.0 / 4.0 * SQ(3.0 * V(i) - 4.0 * V(i - 1) + V(i - 2));
    double S1 = 13.0 / 12.0 * SQ(V(i + 1) - 2.0 * V(i) + V(i - 1)) + 1.0 / 4.0 * SQ(V(i + 1) - V(i - 1));
    double S2 = 13.0 / 12.0 * SQ(V(i + 2) - 2.0 * V(i + 1) + V(i)) + 1.0 / 4.0 * SQ(V(i + 2) - 4.0 * V(i + 1) + 3.0 * V(i));
    double a0 = C03 / SQ(ep + S0), a1 = C13 / SQ(ep + S1), a2 = C23 / SQ(ep + S2);
    double am = a0 + a1 + a2;
    double q03 = 1.0 / 3.0 * V(i) + 5.0 / 6.0 * V(i - 1) - 1.0 / 6.0 * V(i - 2);
    double q13 = -(1.0 / 6.0 * V(i + 1)) + 5.0 / 6.0 * V(i) + 1.0 / 3.0 * V(i - 1);
    double q23 = 1.0 / 3.0 * V(i + 2) - 7.0 / 6.0 * V(i + 1) + 11.0 / 6.0 * V(i);
    hh[i - 1] = (a0 * q03 + a1 * q13 + a2 * q23) / am;
  }
}

/* WENO7-JS constants, OCFD_Schemes.f90:62-80 */
#define W7_CONSTS                                                                              \
  const double C0 = 1.0 / 35.0, C1 = 12.0 / 35.0, C2 = 18.0 / 35.0, C3 = 4.0 / 35.0;           \
  const double a11 = -2.0 / 6.0, a12 = 9.0 / 6.0, a13 = -18.0 / 6.0, a14 = 11.0 / 6.0;         \
  const double a21 = 1.0 / 6.0, a23 = 3.0 / 6.0, a24 = 2.0 / 6.0;                              \
  const double a31 = -2.0 / 6.0, a32 = -3.0 / 6.0, a34 = -1.0 / 6.0;                           \
  const double a41 = -11.0 / 6.0, a42 = 18.0 / 6.0, a43 = -9.0 / 6.0, a44 = 2.0 / 6.0;         \
  const double b12 = 4.0, b13 = -5.0, b14 = 2.0, b22 = -2.0, b41 = 2.0, b42 = -5.0, b43 = 4.0; \
  const double c12 = 3.0, d12 = 13.0 / 12.0, d13 = 1043.0 / 960.0, d14 = 1.0 / 12.0;           \
  const double e11 = -3.0 / 12.0, e12 = 13.0 / 12.0, e13 = -23.0 / 12.0, e14 = 25.0 / 12.0;    \
  const double e21 = 1.0 / 12.0, e22 = -5.0 / 12.0, e23 = 13.0 / 12.0, e24 = 3.0 / 12.0;       \
  const double e31 = -1.0 / 12.0, e32 = 7.0 / 12.0, e33 = 7.0 / 12.0, e34 = -1.0 / 12.0;       \
  const double e41 = 3.0 / 12.0, e42 = 13.0 / 12.0, e43 = -5.0 / 12.0, e44 = 1.0 / 12.0;       \
  const double ep = 1.0e-8

/* WENO7 point form on x(-3..3) given as m3..p3 (hh_weno7P, OCFD_Schemes_1.f90:59-116) */
static double weno7_pt(double m3, double m2, double m1, double z0, double p1, double p2, double p3) {
  W7_CONSTS;
  double S10 = a11 * m3 + a12 * m2 + a13 * m1 + a14 * z0;
  double S11 = a21 * m2 - m1 + a23 * z0 + a24 * p1;
  double S12 = a31 * m1 + a32 * z0 + p1 + a34 * p2;
  double S13 = a41 * z0 + a42 * p1 + a43 * p2 + a44 * p3;
  double S20 = -m3 + b12 * m2 + b13 * m1 + b14 * z0;
  double S21 = m1 + b22 * z0 + p1;
  double S22 = z0 + b22 * p1 + p2;
  double S23 = b41 * z0 + b42 * p1 + b43 * p2 - p3;
  double S30 = -m3 + c12 * (m2 - m1) + z0;
  double S31 = -m2 + c12 * (m1 - z0) + p1;
  double S32 = -m1 + c12 * (z0 - p1) + p2;
  double S33 = -z0 + c12 * (p1 - p2) + p3;
  double S0 = S10 * S10 + d12 * S20 * S20 + d13 * S30 * S30 + d14 * S10 * S30;
  double S1 = S11 * S11 + d12 * S21 * S21 + d13 * S31 * S31 + d14 * S11 * S31;
  double S2 = S12 * S12 + d12 * S22 * S22 + d13 * S32 * S32 + d14 * S12 * S32;
  double S3 = S13 * S13 + d12 * S23 * S23 + d13 * S33 * S33 + d14 * S13 * S33;
  double a0 = C0 / SQ(ep + S0), a1 = C1 / SQ(ep + S1), a2 = C2 / SQ(ep + S2), a3 = C3 / SQ(ep + S3);
  double am = a0 + a1 + a2 + a3;
  double q0 = e11 * m3 + e12 * m2 + e13 * m1 + e14 * z0;
  double q1 = e21 * m2 + e22 * m1 + e23 * z0 + e24 * p1;
  double q2 = e31 * m1 + e32 * z0 + e33 * p1 + e34 * p2;
  double q3 = e41 * z0 + e42 * p1 + e43 * p2 + e44 * p3;
  return (a0 * q0 + a1 * q1 + a2 * q2 + a3 * q3) / am;
}

/* OCFD_weno7P, OCFD_Schemes.f90:55-124 */
static void weno7P(const double* v, double* hh, int n, int lap, int ib, int ie) {
  (void)n;
  for (int i = ib; i <= ie; i++)
    hh[i] = weno7_pt(V(i - 3), V(i - 2), V(i - 1), V(i), V(i + 1), V(i + 2), V(i + 3));
}

/* OCFD_weno7M, OCFD_Schemes.f90:128-196: the flux- formulas are those of
 * flux+ with the stencil read from the right (v(i+3) ... v(i-3)). */
static void weno7M(const double* v, double* hh, int n, int lap, int ib, int ie) {
  (void)n;
  for (int i = ib + 1; i <= ie + 1; i++)
    hh[i - 1] = weno7_pt(V(i + 3), V(i + 2), V(i + 1), V(i), V(i - 1), V(i - 2), V(i - 3));
}

/* minmod2 / minmod4, OCFD_Schemes.f90:284-299 */
static double fsign1(double x) { return signbit(x) ? -1.0 : 1.0; }
static double minmod2(double x1, double x2) {
  return 0.5 * (fsign1(x1) + fsign1(x2)) * fmin(fabs(x1), fabs(x2));
}
static double minmod4(double x1, double x2, double x3, double x4) {
  double r = 0.5 * (fsign1(x1) + fsign1(x2));
  r = r * fabs(0.5 * (fsign1(x1) + fsign1(x3)));
  r = r * fabs(0.5 * (fsign1(x1) + fsign1(x4)));
  return r * fmin(fmin(fmin(fabs(x1), fabs(x2)), fabs(x3)), fabs(x4));
}

/* OMP6 point form with v(-3..4) = s[0..7] (hh_OMP6P, OCFD_Schemes_1.f90:204-234).
 * ep = 1.e-10 is a single-precision literal in the reference
 * (OCFD_Schemes.f90:215): its value as a double is 1.00000001335143196e-10. */
static double omp6_pt(const double* s) {
  const double kappa = 4.0, ep = (double)1.0e-10f;
  const double OMP_m = 0.015, OMP_n = 0.0;
  const double OMP_a = 0.5 * (OMP_m + OMP_n), OMP_b = 0.5 * (OMP_m - OMP_n);
  const double a1 = OMP_a, a2 = 1.0 / 60.0 - OMP_b - 6.0 * OMP_a, a3 = -2.0 / 15.0 + 6.0 * OMP_b + 15.0 * OMP_a;
  const double a4 = 37.0 / 60.0 - 15.0 * OMP_b - 20.0 * OMP_a, a5 = 37.0 / 60.0 + 20.0 * OMP_b + 15.0 * OMP_a;
  const double a6 = -2.0 / 15.0 - 6.0 * OMP_a - 15.0 * OMP_b, a7 = 1.0 / 60.0 + OMP_a + 6.0 * OMP_b, a8 = -OMP_b;
  const double vm3 = s[0], vm2 = s[1], vm1 = s[2], v0 = s[3], v1 = s[4], v2 = s[5], v3 = s[6], v4 = s[7];
  double mid_nf = a1 * v4 + a2 * v3 + a3 * v2 + a4 * v1 + a5 * v0 + a6 * vm1 + a7 * vm2 + a8 * vm3;
  double mp = v0 + minmod2(v1 - v0, kappa * (v0 - vm1));
  if ((mid_nf - v0) * (mid_nf - mp) >= ep) {
    double d1 = vm2 + v0 - 2.0 * vm1;
    double d2 = vm1 + v1 - 2.0 * v0;
    double d3 = v0 + v2 - 2.0 * v1;
    double ul = v0 + kappa * (v0 - vm1);
    double md = 0.5 * (v0 + v1) - 0.5 * minmod4(4.0 * d2 - d3, 4.0 * d3 - d2, d2, d3);
    double lc = v0 + 0.5 * (v0 - vm1) + kappa * minmod4(4.0 * d1 - d2, 4.0 * d2 - d1, d2, d1) / 3.0;
    double fmin_ = fmax(fmin(fmin(v0, v1), md), fmin(fmin(v0, ul), lc));
    double fmax_ = fmin(fmax(fmax(v0, v1), md), fmax(fmax(v0, ul), lc));
    mid_nf = mid_nf + minmod2(fmax_ - mid_nf, fmin_ - mid_nf);
  }
  return mid_nf;
}

/* OCFD_OMP6P / OCFD_OMP6M, OCFD_Schemes.f90:207-281 */
static void omp6P(const double* v, double* hh, int n, int lap, int ib, int ie) {
  (void)n;
  for (int i = ib; i <= ie; i++) {
    double s[8];
    for (int o = -3; o <= 4; o++) s[o + 3] = V(i + o);
    hh[i] = omp6_pt(s);
  }
}
static void omp6M(const double* v, double* hh, int n, int lap, int ib, int ie) {
  (void)n;
  for (int i = ib; i <= ie; i++) {
    double s[8];
    for (int o = -3; o <= 4; o++) s[o + 3] = V(i + 1 - o);
    hh[i] = omp6_pt(s);
  }
}

/* OCFD_UD7L_P / _M, OCFD_Schemes.f90:302-326 */
static double ud7l_pt(double vm3, double vm2, double vm1, double v0, double v1, double v2, double v3, double v4) {
  const double* c = G.ud7l;
  return c[0] * vm3 + c[1] * vm2 + c[2] * vm1 + c[3] * v0 + c[4] * v1 + c[5] * v2 + c[6] * v3 + c[7] * v4;
}
static void ud7lP(const double* v, double* hh, int n, int lap, int ib, int ie) {
  (void)n;
  for (int i = ib; i <= ie; i++)
    hh[i] = ud7l_pt(V(i - 3), V(i - 2), V(i - 1), V(i), V(i + 1), V(i + 2), V(i + 3), V(i + 4));
}
static void ud7lM(const double* v, double* hh, int n, int lap, int ib, int ie) {
  (void)n;
  for (int i = ib; i <= ie; i++)
    hh[i] = ud7l_pt(V(i + 4), V(i + 3), V(i + 2), V(i + 1), V(i), V(i - 1), V(i - 2), V(i - 3));
}

/* hh_weno5P / hh_weno5M on a stencil centred at x(0): x[o] for o=-2..3 */
static double hh_weno5P(const double* x) { /* OCFD_Schemes_1.f90:9-30 */
  const double ep = 1.0e-6, C03 = 3.0 / 10.0, C13 = 3.0 / 5.0, C23 = 1.0 / 10.0;
  double S0 = 13.0 / 12.0 * SQ(x[0] - 2.0 * x[1] + x[2]) + 1.0 / 4.0 * SQ(3.0 * x[0] - 4.0 * x[1] + x[2]);
  double S1 = 13.0 / 12.0 * SQ(x[-1] - 2.0 * x[0] + x[1]) + 1.0 / 4.0 * SQ(x[-1] - x[1]);
  double S2 = 13.0 / 12.0 * SQ(x[-2] - 2.0 * x[-1] + x[0]) + 1.0 / 4.0 * SQ(x[-2] - 4.0 * x[-1] + 3.0 * x[0]);
  double a0 = C03 / SQ(ep + S0), a1 = C13 / SQ(ep + S1), a2 = C23 / SQ(ep + S2);
  double am = a0 + a1 + a2;
  double q03 = 1.0 / 3.0 * x[0] + 5.0 / 6.0 * x[1] - 1.0 / 6.0 * x[2];
  double q13 = -(1.0 / 6.0 * x[-1]) + 5.0 / 6.0 * x[0] + 1.0 / 3.0 * x[1];
  double q23 = 1.0 / 3.0 * x[-2] - 7.0 / 6.0 * x[-1] + 11.0 / 6.0 * x[0];
  return (a0 * q03 + a1 * q13 + a2 * q23) / am;
}
static double hh_weno5M(const double* x) { /* OCFD_Schemes_1.f90:32-54 */
  const double ep = 1.0e-6, C03 = 3.0 / 10.0, C13 = 3.0 / 5.0, C23 = 1.0 / 10.0;
  double S0 = 13.0 / 12.0 * SQ(x[1] - 2.0 * x[0] + x[-1]) + 1.0 / 4.0 * SQ(3.0 * x[1] - 4.0 * x[0] + x[-1]);
  double S1 = 13.0 / 12.0 * SQ(x[2] - 2.0 * x[1] + x[0]) + 1.0 / 4.0 * SQ(x[2] - x[0]);
  double S2 = 13.0 / 12.0 * SQ(x[3] - 2.0 * x[2] + x[1]) + 1.0 / 4.0 * SQ(x[3] - 4.0 * x[2] + 3.0 * x[1]);
  double a0 = C03 / SQ(ep + S0), a1 = C13 / SQ(ep + S1), a2 = C23 / SQ(ep + S2);
  double am = a0 + a1 + a2;
  double q03 = 1.0 / 3.0 * x[1] + 5.0 / 6.0 * x[0] - 1.0 / 6.0 * x[-1];
  double q13 = -(1.0 / 6.0 * x[2]) + 5.0 / 6.0 * x[1] + 1.0 / 3.0 * x[0];
  double q23 = 1.0 / 3.0 * x[3] - 7.0 / 6.0 * x[2] + 11.0 / 6.0 * x[1];
  return (a0 * q03 + a1 * q13 + a2 * q23) / am;
}

enum { DEL_NONE = 0, DEL_LIFT = 1, DEL_RIGHT = 2 };
/* hh_weno5P_boundary, OCFD_Schemes_1.f90:279-308 */
static double hh_weno5P_boundary(const double* x, int del) {
  const double ep = 1.0e-6, C03 = 3.0 / 10.0, C13 = 3.0 / 5.0, C23 = 1.0 / 10.0;
  double a0 = 0.0, q03 = 0.0, a2 = 0.0, q23 = 0.0;
  if (del != DEL_RIGHT) {
    double S0 = 13.0 / 12.0 * SQ(x[0] - 2.0 * x[1] + x[2]) + 1.0 / 4.0 * SQ(3.0 * x[0] - 4.0 * x[1] + x[2]);
    a0 = C03 / SQ(ep + S0);
    q03 = 1.0 / 3.0 * x[0] + 5.0 / 6.0 * x[1] - 1.0 / 6.0 * x[2];
  }
  double S1 = 13.0 / 12.0 * SQ(x[-1] - 2.0 * x[0] + x[1]) + 1.0 / 4.0 * SQ(x[-1] - x[1]);
  double a1 = C13 / SQ(ep + S1);
  double q13 = -(1.0 / 6.0 * x[-1]) + 5.0 / 6.0 * x[0] + 1.0 / 3.0 * x[1];
  if (del != DEL_LIFT) {
    double S2 = 13.0 / 12.0 * SQ(x[-2] - 2.0 * x[-1] + x[0]) + 1.0 / 4.0 * SQ(x[-2] - 4.0 * x[-1] + 3.0 * x[0]);
    a2 = C23 / SQ(ep + S2);
    q23 = 1.0 / 3.0 * x[-2] - 7.0 / 6.0 * x[-1] + 11.0 / 6.0 * x[0];
  }
  double am = a0 + a1 + a2;
  return (a0 * q03 + a1 * q13 + a2 * q23) / am;
}
/* hh_weno5M_boundary, OCFD_Schemes_1.f90:310-338 */
static double hh_weno5M_boundary(const double* x, int del) {
  const double ep = 1.0e-6, C03 = 3.0 / 10.0, C13 = 3.0 / 5.0, C23 = 1.0 / 10.0;
  double a0 = 0.0, q03 = 0.0, a2 = 0.0, q23 = 0.0;
  if (del != DEL_LIFT) {
    double S0 = 13.0 / 12.0 * SQ(x[1] - 2.0 * x[0] + x[-1]) + 1.0 / 4.0 * SQ(3.0 * x[1] - 4.0 * x[0] + x[-1]);
    a0 = C03 / SQ(ep + S0);
    q03 = 1.0 / 3.0 * x[1] + 5.0 / 6.0 * x[0] - 1.0 / 6.0 * x[-1];
  }
  double S1 = 13.0 / 12.0 * SQ(x[2] - 2.0 * x[1] + x[0]) + 1.0 / 4.0 * SQ(x[2] - x[0]);
  double a1 = C13 / SQ(ep + S1);
  double q13 = -(1.0 / 6.0 * x[2]) + 5.0 / 6.0 * x[1] + 1.0 / 3.0 * x[0];
  if (del != DEL_RIGHT) {
    double S2 = 13.0 / 12.0 * SQ(x[3] - 2.0 * x[2] + x[1]) + 1.0 / 4.0 * SQ(x[3] - 4.0 * x[2] + 3.0 * x[1]);
    a2 = C23 / SQ(ep + S2);
    q23 = 1.0 / 3.0 * x[3] - 7.0 / 6.0 * x[2] + 11.0 / 6.0 * x[1];
  }
  double am = a0 + a1 + a2;
  return (a0 * q03 + a1 * q13 + a2 * q23) / am;
}

/* OCFD_HybridP / OCFD_HybridM, OCFD_Schemes.f90:329-370 */
static void hybridP(const double* v, double* hh, int n, int lap, const unsigned char* scm, int ib, int ie) {
  (void)n;
  for (int i = ib; i <= ie; i++) {
    if (scm[i] == 1) hh[i] = ud7l_pt(V(i - 3), V(i - 2), V(i - 1), V(i), V(i + 1), V(i + 2), V(i + 3), V(i + 4));
    else if (scm[i] == 2) hh[i] = weno7_pt(V(i - 3), V(i - 2), V(i - 1), V(i), V(i + 1), V(i + 2), V(i + 3));
    else hh[i] = hh_weno5P(&V(i));
  }
}
static void hybridM(const double* v, double* hh, int n, int lap, const unsigned char* scm, int ib, int ie) {
  (void)n;
  for (int i = ib; i <= ie; i++) {
    if (scm[i] == 1) hh[i] = ud7l_pt(V(i + 4), V(i + 3), V(i + 2), V(i + 1), V(i), V(i - 1), V(i - 2), V(i - 3));
    else if (scm[i] == 2) hh[i] = weno7_pt(V(i + 4), V(i + 3), V(i + 2), V(i + 1), V(i), V(i - 1), V(i - 2));
    else hh[i] = hh_weno5M(&V(i));
  }
}

/* OCFD2d_flux1 (flux+), OCFD_inviscous.f90:209-281 */
static void flux1(const double* v, double* hh, int n, int lap, const int* bidx, const unsigned char* scm, const int* sb) {
  int Ka = G.Ka1, Kb = G.Kb1;
  int ib = (bidx[0] == 0 || sb[0] == -1) ? 0 : -Ka + 1;
  int ie = (bidx[1] == 0 || sb[1] == -1) ? n : n - Kb;
  switch (G.cfg.scheme_invis) {
    case OCFD_SCHEME_OMP6: omp6P(v, hh, n, lap, ib, ie); break;
    case OCFD_SCHEME_WENO7: weno7P(v, hh, n, lap, ib, ie); break;
    case OCFD_SCHEME_WENO5: weno5P(v, hh, n, lap, ib, ie); break;
    case OCFD_SCHEME_UD7L: ud7lP(v, hh, n, lap, ib, ie); break;
    case OCFD_SCHEME_HYBRID: hybridP(v, hh, n, lap, scm, ib, ie); break;
  }
  if (bidx[0] == 1) {
    if (sb[0] == 0) {
      for (int i = 0; i <= -Ka; i++) {
        if (i == 0 || i == 1) hh[i] = (2.0 * V(1) + 5.0 * V(2) - V(3)) / 6.0;
        else if (i == 2) hh[i] = hh_weno5P_boundary(&V(i), DEL_LIFT);
        else hh[i] = hh_weno5P(&V(i));
      }
    } else if (sb[0] == 1) {
      for (int i = 0; i <= -Ka; i++) hh[i] = hh_weno5P(&V(i));
    }
  }
  if (bidx[1] == 1) {
    if (sb[1] == 0) {
      for (int i = n - Kb + 1; i <= n; i++) {
        if (i == n) hh[i] = (2.0 * V(n - 2) - 7.0 * V(n - 1) + 11.0 * V(n)) / 6.0;
        else if (i == n - 1) hh[i] = hh_weno5P_boundary(&V(i), DEL_RIGHT);
        else hh[i] = hh_weno5P(&V(i));
      }
    } else if (sb[1] == 1) {
      for (int i = n - Kb + 1; i <= n; i++) hh[i] = hh_weno5P(&V(i));
    }
  }
}

/* OCFD2d_flux2 (flux-), OCFD_inviscous.f90:286-362 */
static void flux2(const double* v, double* hh, int n, int lap, const int* bidx, const unsigned char* scm, const int* sb) {
  int Ka = G.Ka2, Kb = G.Kb2;
  int ib = (bidx[0] == 0 || sb[0] == -1) ? 0 : -Ka + 1;
  int ie = (bidx[1] == 0 || sb[1] == -1) ? n : n - Kb;
  switch (G.cfg.scheme_invis) {
    case OCFD_SCHEME_OMP6: omp6M(v, hh, n, lap, ib, ie); break;
    case OCFD_SCHEME_WENO7: weno7M(v, hh, n, lap, ib, ie); break;
    case OCFD_SCHEME_WENO5: weno5M(v, hh, n, lap, ib, ie); break;
    case OCFD_SCHEME_UD7L: ud7lM(v, hh, n, lap, ib, ie); break;
    case OCFD_SCHEME_HYBRID: hybridM(v, hh, n, lap, scm, ib, ie); break;
  }
  if (bidx[0] == 1) {
    if (sb[0] == 0) {
      for (int i = 0; i <= -Ka; i++) {
        if (i == 0) hh[i] = (2.0 * V(3) - 7.0 * V(2) + 11.0 * V(1)) / 6.0;
        else if (i == 1) hh[i] = hh_weno5M_boundary(&V(i), DEL_LIFT);
        else hh[i] = hh_weno5M(&V(i));
      }
    } else if (sb[0] == 1) {
      for (int i = 0; i <= -Ka; i++) hh[i] = hh_weno5M(&V(i));
    }
  }
  if (bidx[1] == 1) {
    if (sb[1] == 0) {
      for (int i = n - Kb + 1; i <= n; i++) {
        if (i == n || i == n - 1) hh[i] = (2.0 * V(n) + 5.0 * V(n - 1) - V(n - 2)) / 6.0;
        else if (i == n - 2) hh[i] = hh_weno5M_boundary(&V(i), DEL_RIGHT);
        else hh[i] = hh_weno5M(&V(i));
      }
    } else if (sb[1] == 1) {
      for (int i = n - Kb + 1; i <= n; i++) hh[i] = hh_weno5M(&V(i));
    }
  }
}
#undef V

/* ------------------------------------------------------------------ */
/* Flux splitting (code/flow-solver/OCFD_flux_split.f90). Arrays indexed
 * 1-LAP..n+LAP through offset o = i+lap-1; fp/fm are [5][n+2lap]. */

/* split_Stager_Warming, OCFD_flux_split.f90:6-72 */
static void split_SW(int n, int lap, const double* Et1, const double* p1, const double* c1, const double* d1,
                     const double* u1, const double* v1, const double* w1, const double* A1, const double* A2,
                     const double* A3, double* fp, double* fm) {
  int np = n + 2 * lap;
  for (int o = 0; o < np; o++) {
    double gamma = p1[o] / Et1[o] + 1.0;
    double tmp1 = 2.0 * (gamma - 1.0);
    double tmp3 = (3.0 - gamma) / (2.0 * (gamma - 1.0));
    double ss = sqrt(A1[o] * A1[o] + A2[o] * A2[o] + A3[o] * A3[o]);
    double ss1 = 1.0 / ss;
    double ak1 = A1[o] * ss1, ak2 = A2[o] * ss1, ak3 = A3[o] * ss1;
    double vs = A1[o] * u1[o] + A2[o] * v1[o] + A3[o] * w1[o];
    double E1 = vs, E2 = vs - c1[o] * ss, E3 = vs + c1[o] * ss;
    double E1P = (E1 + fabs(E1)) * 0.5, E2P = (E2 + fabs(E2)) * 0.5, E3P = (E3 + fabs(E3)) * 0.5;
    double E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
    double tmp0 = d1[o] / (2.0 * gamma);
    double uc1 = u1[o] - c1[o] * ak1, uc2 = u1[o] + c1[o] * ak1;
    double vc1 = v1[o] - c1[o] * ak2, vc2 = v1[o] + c1[o] * ak2;
    double wc1 = w1[o] - c1[o] * ak3, wc2 = w1[o] + c1[o] * ak3;
    double vvc1 = (uc1 * uc1 + vc1 * vc1 + wc1 * wc1) * 0.5;
    double vvc2 = (uc2 * uc2 + vc2 * vc2 + wc2 * wc2) * 0.5;
    double vv = (gamma - 1.0) * (u1[o] * u1[o] + v1[o] * v1[o] + w1[o] * w1[o]);
    double W2 = tmp3 * c1[o] * c1[o];
    fp[0 * np + o] = tmp0 * (tmp1 * E1P + E2P + E3P);
    fp[1 * np + o] = tmp0 * (tmp1 * E1P * u1[o] + E2P * uc1 + E3P * uc2);
    fp[2 * np + o] = tmp0 * (tmp1 * E1P * v1[o] + E2P * vc1 + E3P * vc2);
    fp[3 * np + o] = tmp0 * (tmp1 * E1P * w1[o] + E2P * wc1 + E3P * wc2);
    fp[4 * np + o] = tmp0 * (E1P * vv + E2P * vvc1 + E3P * vvc2 + W2 * (E2P + E3P));
    fm[0 * np + o] = tmp0 * (tmp1 * E1M + E2M + E3M);
    fm[1 * np + o] = tmp0 * (tmp1 * E1M * u1[o] + E2M * uc1 + E3M * uc2);
    fm[2 * np + o] = tmp0 * (tmp1 * E1M * v1[o] + E2M * vc1 + E3M * vc2);
    fm[3 * np + o] = tmp0 * (tmp1 * E1M * w1[o] + E2M * wc1 + E3M * wc2);
    fm[4 * np + o] = tmp0 * (E1M * vv + E2M * vvc1 + E3M * vvc2 + W2 * (E2M + E3M));
  }
}

/* split_Local_LaxFriedrichs, OCFD_flux_split.f90:76-103 */
static void split_LLF(int n, int lap, const double* Et1, const double* p1, const double* c1, const double* d1,
                      const double* u1, const double* v1, const double* w1, const double* A1, const double* A2,
                      const double* A3, double* fp, double* fm) {
  int np = n + 2 * lap;
  for (int o = 0; o < np; o++) {
    double ss = sqrt(A1[o] * A1[o] + A2[o] * A2[o] + A3[o] * A3[o]);
    double un = A1[o] * u1[o] + A2[o] * v1[o] + A3[o] * w1[o];
    double lamda0 = fabs(un) + ss * c1[o];
    double E0 = Et1[o] + 0.5 * d1[o] * (u1[o] * u1[o] + v1[o] * v1[o] + w1[o] * w1[o]);
    fp[0 * np + o] = 0.5 * (d1[o] * un + lamda0 * d1[o]);
    fp[1 * np + o] = 0.5 * (d1[o] * u1[o] * un + A1[o] * p1[o] + lamda0 * d1[o] * u1[o]);
    fp[2 * np + o] = 0.5 * (d1[o] * v1[o] * un + A2[o] * p1[o] + lamda0 * d1[o] * v1[o]);
    fp[3 * np + o] = 0.5 * (d1[o] * w1[o] * un + A3[o] * p1[o] + lamda0 * d1[o] * w1[o]);
    fp[4 * np + o] = 0.5 * ((E0 + p1[o]) * un + lamda0 * E0);
    fm[0 * np + o] = 0.5 * (d1[o] * un - lamda0 * d1[o]);
    fm[1 * np + o] = 0.5 * (d1[o] * u1[o] * un + A1[o] * p1[o] - lamda0 * d1[o] * u1[o]);
    fm[2 * np + o] = 0.5 * (d1[o] * v1[o] * un + A2[o] * p1[o] - lamda0 * d1[o] * v1[o]);
    fm[3 * np + o] = 0.5 * (d1[o] * w1[o] * un + A3[o] * p1[o] - lamda0 * d1[o] * w1[o]);
    fm[4 * np + o] = 0.5 * ((E0 + p1[o]) * un - lamda0 * E0);
  }
}

/* ------------------------------------------------------------------ */
/* Periodic face fill (exchange_boundary_x/y/z with one rank, OCFD_mpi.f90:196-236) */
static void exchange_x(double* a) {
  if (!G.cfg.periodic[0]) return;
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int l = 1; l <= G.lap; l++) {
        a[H(G.nx + l, j, k)] = a[H(l, j, k)];
        a[H(1 - l, j, k)] = a[H(G.nx + 1 - l, j, k)];
      }
}
static void exchange_y(double* a) {
  if (!G.cfg.periodic[1]) return;
  for (int k = 1; k <= G.nz; k++)
    for (int l = 1; l <= G.lap; l++)
      for (int i = 1; i <= G.nx; i++) {
        a[H(i, G.ny + l, k)] = a[H(i, l, k)];
        a[H(i, 1 - l, k)] = a[H(i, G.ny + 1 - l, k)];
      }
}
static void exchange_z(double* a) {
  if (!G.cfg.periodic[2]) return;
  for (int l = 1; l <= G.lap; l++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        a[H(i, j, G.nz + l)] = a[H(i, j, l)];
        a[H(i, j, 1 - l)] = a[H(i, j, G.nz + 1 - l)];
      }
}
static void exchange_xyz(double* a) { exchange_x(a); exchange_y(a); exchange_z(a); }

/* ------------------------------------------------------------------ */
/* Central first derivatives, OCFD_dxyz_viscous.f90:3-242. f is a halo array,
 * out an interior array. */
static void dxyz0(int dir, const double* f, double* out) {
  int n = dir == 0 ? G.nx : dir == 1 ? G.ny : G.nz;
  double h = dir == 0 ? G.hx : dir == 1 ? G.hy : G.hz;
  double a1 = 1.0 / (60.0 * h), a2 = -3.0 / (20.0 * h), a3 = 3.0 / (4.0 * h);
  double c1 = 0.8 / h, c2 = -0.2 / h, c3 = 3.80952380952380952e-2 / h, c4 = -3.571428571428571428e-3 / h;
  int cd8 = G.cfg.scheme_vis == OCFD_SCHEME_CD8;
  long st = dir == 0 ? 1 : dir == 1 ? G.nxh : (long)G.nxh * G.nyh;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        const double* F = f + H(i, j, k);
#define FF(o) F[(long)(o) * st]
        double r;
        if (!cd8) r = a1 * (FF(3) - FF(-3)) + a2 * (FF(2) - FF(-2)) + a3 * (FF(1) - FF(-1));
        else r = c1 * (FF(1) - FF(-1)) + c2 * (FF(2) - FF(-2)) + c3 * (FF(3) - FF(-3)) + c4 * (FF(4) - FF(-4));
        int s = dir == 0 ? i : dir == 1 ? j : k;
        if (!G.cfg.periodic[dir]) {
          if (s == 1) r = (-3.0 * FF(0) + 4.0 * FF(1) - FF(2)) / (2.0 * h);
          else if (s == 2) r = (FF(1) - FF(-1)) / (2.0 * h);
          else if (s == 3) r = (8.0 * (FF(1) - FF(-1)) - (FF(2) - FF(-2))) / (12.0 * h);
          else if (s == 4 && cd8) r = a1 * (FF(3) - FF(-3)) + a2 * (FF(2) - FF(-2)) + a3 * (FF(1) - FF(-1));
          if (s == n) r = (FF(-2) - 4.0 * FF(-1) + 3.0 * FF(0)) / (2.0 * h);
          else if (s == n - 1) r = (FF(1) - FF(-1)) / (2.0 * h);
          else if (s == n - 2) r = (8.0 * (FF(1) - FF(-1)) - (FF(2) - FF(-2))) / (12.0 * h);
          else if (s == n - 3 && cd8) r = a1 * (FF(3) - FF(-3)) + a2 * (FF(2) - FF(-2)) + a3 * (FF(1) - FF(-1));
        }
#undef FF
        out[I(i, j, k)] = r;
      }
}
void orc_dx0(const double* f, double* out) { dxyz0(0, f, out); }
void orc_dy0(const double* f, double* out) { dxyz0(1, f, out); }
void orc_dz0(const double* f, double* out) { dxyz0(2, f, out); }

/* ------------------------------------------------------------------ */
/* comput_cc, Comput_flow_variables.f90:50-63 */
static void comput_cc(void) {
#pragma omp parallel for schedule(static)
  for (long q = 0; q < G.nh; q++) {
    double gamma = G.p[q] / G.Et[q] + 1.0;
    G.cc[q] = sqrt(gamma * G.p[q] / G.d[q]);
  }
}

/* set_hybrid_scheme, OCFD_SetHybrid.f90:2-14 */
static unsigned char set_hybrid_scheme(double R0) {
  if (R0 <= G.cfg.hybrid_para[1]) return 1;
  if (R0 <= G.cfg.hybrid_para[2]) return 2;
  return 3;
}

/* du_inviscous, OCFD_inviscous.f90:47-202 (non-characteristic form).
 * Pencils of one direction are independent (each writes its own cells and
 * interfaces), so they are shared out over OpenMP threads, each with its own
 * pencil buffers; the arithmetic per pencil is unchanged. */
static void du_inviscous(void) {
  int nmax = G.nx > G.ny ? G.nx : G.ny;
  if (G.nz > nmax) nmax = G.nz;
  int lap = G.lap, npm = nmax + 2 * lap;
  int hyb = G.cfg.scheme_invis == OCFD_SCHEME_HYBRID;
  double hx_1 = 1.0 / G.hx, hy_1 = 1.0 / G.hy, hz_1 = 1.0 / G.hz;
  comput_cc();
  for (int dir = 0; dir < 3; dir++) {
    int n = dir == 0 ? G.nx : dir == 1 ? G.ny : G.nz, np = n + 2 * lap;
    long st = dir == 0 ? 1 : dir == 1 ? G.nxh : (long)G.nxh * G.nyh;
    const double *M1 = dir == 0 ? G.Akx1 : dir == 1 ? G.Aix1 : G.Asx1;
    const double *M2 = dir == 0 ? G.Aky1 : dir == 1 ? G.Aiy1 : G.Asy1;
    const double *M3 = dir == 0 ? G.Akz1 : dir == 1 ? G.Aiz1 : G.Asz1;
    const int* sb = &G.cfg.scheme_boundary[2 * dir];
    int na = dir == 0 ? G.ny : G.nx, nb = dir == 2 ? G.ny : G.nz;
    unsigned char* mask = dir == 0 ? G.maskx : dir == 1 ? G.masky : G.maskz;
#pragma omp parallel
    {
      double *Et1 = dalloc(npm, 0), *p1 = dalloc(npm, 0), *c1 = dalloc(npm, 0), *d1 = dalloc(npm, 0);
      double *u1 = dalloc(npm, 0), *v1 = dalloc(npm, 0), *w1 = dalloc(npm, 0);
      double *A1 = dalloc(npm, 0), *A2 = dalloc(npm, 0), *A3 = dalloc(npm, 0);
      double *fp = dalloc(5L * npm, 0), *fm = dalloc(5L * npm, 0);
      double *hh1 = dalloc(nmax + 1, 0), *hh2 = dalloc(nmax + 1, 0);
      unsigned char* scm = (unsigned char*)calloc((size_t)nmax + 1, 1);
#pragma omp for schedule(static)
      for (long pen = 0; pen < (long)na * nb; pen++) {
        int a = (int)(pen % na) + 1, b = (int)(pen / na) + 1;
        int i0 = dir == 0 ? 0 : a, j0 = dir == 0 ? a : dir == 1 ? 0 : b, k0 = dir == 2 ? 0 : b;
        long base = H(i0, j0, k0); /* point at s=0 of the pencil */
        if (hyb) {
          for (int s = 0; s <= n; s++) scm[s] = set_hybrid_scheme(fmax(G.Rhybrid[base + s * st], G.Rhybrid[base + (s + 1) * st]));
        }
        for (int s = 1 - lap; s <= n + lap; s++) {
          long q = base + s * st;
          int o = s + lap - 1;
          Et1[o] = G.Et[q]; p1[o] = G.p[q]; c1[o] = G.cc[q]; d1[o] = G.d[q];
          u1[o] = G.u[q]; v1[o] = G.v[q]; w1[o] = G.w[q];
          A1[o] = M1[q]; A2[o] = M2[q]; A3[o] = M3[q];
        }
        if (G.cfg.flux_splitting == OCFD_SPLIT_SW) split_SW(n, lap, Et1, p1, c1, d1, u1, v1, w1, A1, A2, A3, fp, fm);
        else split_LLF(n, lap, Et1, p1, c1, d1, u1, v1, w1, A1, A2, A3, fp, fm);
        for (int m = 0; m < 5; m++) {
          flux1(fp + (long)m * np, hh1, n, lap, G.bound_index[dir], scm, sb);
          flux2(fm + (long)m * np, hh2, n, lap, G.bound_index[dir], scm, sb);
          for (int s = 1; s <= n; s++) {
            int i = dir == 0 ? s : a, j = dir == 1 ? s : (dir == 0 ? a : b), k = dir == 2 ? s : b;
            long c = I(i, j, k) + (long)m * G.ni;
            if (dir == 0) G.du[c] = (hh1[s] - hh1[s - 1] + hh2[s] - hh2[s - 1]) * hx_1;
            else if (dir == 1) G.du[c] = G.du[c] + (hh1[s] - hh1[s - 1] + hh2[s] - hh2[s - 1]) * hy_1;
            else G.du[c] = -(G.du[c] + (hh1[s] - hh1[s - 1] + hh2[s] - hh2[s - 1]) * hz_1) * G.Ajac[H(i, j, k)];
          }
        }
        if (hyb) {
          for (int s = 0; s <= n; s++) {
            long mi;
            if (dir == 0) mi = s + (long)(G.nx + 1) * ((a - 1) + (long)G.ny * (b - 1));
            else if (dir == 1) mi = (a - 1) + (long)G.nx * (s + (long)(G.ny + 1) * (b - 1));
            else mi = (a - 1) + (long)G.nx * ((b - 1) + (long)G.ny * s);
            mask[mi] = scm[s];
          }
        }
      }
      free(Et1); free(p1); free(c1); free(d1); free(u1); free(v1); free(w1);
      free(A1); free(A2); free(A3); free(fp); free(fm); free(hh1); free(hh2); free(scm);
    }
  }
  if (G.cfg.if_mass_force) { /* OCFD_inviscous.f90:187-199 */
    const double* mf = G.cfg.mass_force;
    for (int k = 1; k <= G.nz; k++)
      for (int j = 1; j <= G.ny; j++)
        for (int i = 1; i <= G.nx; i++) {
          long c = I(i, j, k), q = H(i, j, k);
          G.du[c + G.ni] = G.du[c + G.ni] + G.d[q] * mf[0];
          G.du[c + 2 * G.ni] = G.du[c + 2 * G.ni] + G.d[q] * mf[1];
          G.du[c + 3 * G.ni] = G.du[c + 3 * G.ni] + G.d[q] * mf[2];
          G.du[c + 4 * G.ni] = G.du[c + 4 * G.ni] + G.d[q] * (G.u[q] * mf[0] + G.v[q] * mf[1] + G.w[q] * mf[2]);
        }
  }
}

/* ------------------------------------------------------------------ */
/* Characteristic flux (IF_Scheme_Character = 1),
 * OCFD_inviscous_character.f90:3-425. */

/* S (left eigenvectors) and S1 = S^-1 (right eigenvectors) at the
 * arithmetic mean state of an interface, flux_charteric :209-257 */
static void char_matrices(double u1, double v1, double w1, double c1, double p1, double Et1, double a1, double a2,
                          double a3, double S[5][5], double S1[5][5]) {
  double ss = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
  double n1 = a1 / ss, n2 = a2 / ss, n3 = a3 / ss;
  double l1, l2, l3;
  if (fabs(n3) <= fabs(n2)) {
    ss = sqrt(n1 * n1 + n2 * n2);
    l1 = -n2 / ss; l2 = n1 / ss; l3 = 0.0;
  } else {
    ss = sqrt(n1 * n1 + n3 * n3);
    l1 = -n3 / ss; l2 = 0.0; l3 = n1 / ss;
  }
  double m1 = n2 * l3 - n3 * l2, m2 = n3 * l1 - n1 * l3, m3 = n1 * l2 - n2 * l1;
  double un = u1 * n1 + v1 * n2 + w1 * n3, ul = u1 * l1 + v1 * l2 + w1 * l3, um = u1 * m1 + v1 * m2 + w1 * m3;
  double V2 = (u1 * u1 + v1 * v1 + w1 * w1) * 0.5;
  double gamma = p1 / Et1 + 1.0;
  double KK = (gamma - 1.0) / (c1 * c1);
  double KK1 = KK * 0.5, X1 = 1.0 / (2.0 * c1), H = V2 + 1.0 / KK;
  double s[5][5] = {{1.0 - KK * V2, KK * u1, KK * v1, KK * w1, -KK},
                    {-ul, l1, l2, l3, 0.0},
                    {-um, m1, m2, m3, 0.0},
                    {KK1 * V2 + X1 * un, -X1 * n1 - KK1 * u1, -X1 * n2 - KK1 * v1, -X1 * n3 - KK1 * w1, KK1},
                    {KK1 * V2 - X1 * un, X1 * n1 - KK1 * u1, X1 * n2 - KK1 * v1, X1 * n3 - KK1 * w1, KK1}};
  double r[5][5] = {{1.0, 0.0, 0.0, 1.0, 1.0},
                    {u1, l1, m1, u1 - c1 * n1, u1 + c1 * n1},
                    {v1, l2, m2, v1 - c1 * n2, v1 + c1 * n2},
                    {w1, l3, m3, w1 - c1 * n3, w1 + c1 * n3},
                    {V2, ul, um, H - c1 * un, H + c1 * un}};
  memcpy(S, s, sizeof s);
  memcpy(S1, r, sizeof r);
}
int orc_char_matrices(const double* st, double* S, double* S1) {
  char_matrices(st[0], st[1], st[2], st[3], st[4], st[5], st[6], st[7], st[8], (double(*)[5])S, (double(*)[5])S1);
  return 0;
}

/* flux_x1 (flux+ at one interface on a window ff(mm), mm = Ka1..Kb1),
 * OCFD_inviscous_character.f90:317-371 */
static double char_flux_x1(const double* ff, int i, int n, const int* bidx, int ihyb, const int* sb) {
  int Ka = G.Ka1, Kb = G.Kb1;
  if ((bidx[0] == 0 || i > -Ka || sb[0] == -1) && (bidx[1] == 0 || i <= n - Kb || sb[1] == -1)) {
    double s8[8];
    switch (G.cfg.scheme_invis) {
      case OCFD_SCHEME_OMP6:
        for (int o = -3; o <= 4; o++) s8[o + 3] = ff[o];
        return omp6_pt(s8);
      case OCFD_SCHEME_WENO7: return weno7_pt(ff[-3], ff[-2], ff[-1], ff[0], ff[1], ff[2], ff[3]);
      case OCFD_SCHEME_WENO5: return hh_weno5P(ff);
      case OCFD_SCHEME_UD7L: return ud7l_pt(ff[-3], ff[-2], ff[-1], ff[0], ff[1], ff[2], ff[3], ff[4]);
      default: /* hybrid, OCFD_Schemes_1.f90 hh_HybridP */
        if (ihyb == 1) return ud7l_pt(ff[-3], ff[-2], ff[-1], ff[0], ff[1], ff[2], ff[3], ff[4]);
        if (ihyb == 2) return weno7_pt(ff[-3], ff[-2], ff[-1], ff[0], ff[1], ff[2], ff[3]);
        return hh_weno5P(ff);
    }
  } else if (i <= -Ka && bidx[0] == 1 && sb[0] != -1) {
    if (sb[0] == 0) {
      if (i == 1) return (2.0 * ff[0] + 5.0 * ff[1] - ff[2]) / 6.0;
      if (i == 2) return hh_weno5P_boundary(ff, DEL_LIFT);
      return hh_weno5P(ff);
    }
    return hh_weno5P(ff);
  } else if (i > n - Kb && bidx[1] == 1 && sb[1] != -1) {
    if (sb[1] == 0 && i == n - 1) return hh_weno5P_boundary(ff, DEL_RIGHT);
    return hh_weno5P(ff);
  }
  return 0.0;
}

/* flux_x2 (flux-), OCFD_inviscous_character.f90:374-425 */
static double char_flux_x2(const double* ff, int i, int n, const int* bidx, int ihyb, const int* sb) {
  int Ka = G.Ka2, Kb = G.Kb2;
  if ((bidx[0] == 0 || i > -Ka || sb[0] == -1) && (bidx[1] == 0 || i <= n - Kb || sb[1] == -1)) {
    double s8[8];
    switch (G.cfg.scheme_invis) {
      case OCFD_SCHEME_OMP6:
        for (int o = -3; o <= 4; o++) s8[o + 3] = ff[1 - o];
        return omp6_pt(s8);
      case OCFD_SCHEME_WENO7: return weno7_pt(ff[4], ff[3], ff[2], ff[1], ff[0], ff[-1], ff[-2]);
      case OCFD_SCHEME_WENO5: return hh_weno5M(ff);
      case OCFD_SCHEME_UD7L: return ud7l_pt(ff[4], ff[3], ff[2], ff[1], ff[0], ff[-1], ff[-2], ff[-3]);
      default:
        if (ihyb == 1) return ud7l_pt(ff[4], ff[3], ff[2], ff[1], ff[0], ff[-1], ff[-2], ff[-3]);
        if (ihyb == 2) return weno7_pt(ff[4], ff[3], ff[2], ff[1], ff[0], ff[-1], ff[-2]);
        return hh_weno5M(ff);
    }
  } else if (i <= -Ka && bidx[0] == 1 && sb[0] != -1) {
    if (sb[0] == 0 && i == 1) return hh_weno5M_boundary(ff, DEL_LIFT);
    return hh_weno5M(ff);
  } else if (i > n - Kb && bidx[1] == 1 && sb[1] != -1) {
    if (sb[1] == 0) {
      if (i == n - 1) return (2.0 * ff[1] + 5.0 * ff[0] - ff[-1]) / 6.0;
      if (i == n - 2) return hh_weno5M_boundary(ff, DEL_RIGHT);
      return hh_weno5M(ff);
    }
    return hh_weno5M(ff);
  }
  return 0.0;
}

/* flux_charteric, OCFD_inviscous_character.f90:165-314: hh(0:n,5) of one
 * pencil from fp/fm [5][n+2lap] and the pencil's primitives (offset o = s+lap-1) */
static void flux_charteric(int n, int lap, double* hh, const double* fp, const double* fm, const double* A1,
                           const double* A2, const double* A3, const double* EE, const double* pp, const double* cc1,
                           const double* uu, const double* vv, const double* ww, const int* bidx,
                           const unsigned char* scm, const int* sb) {
  int np = n + 2 * lap;
  int ia = (bidx[0] == 0 || sb[0] != 0) ? 0 : 1;
  int ib = (bidx[1] == 0 || sb[1] != 0) ? n : n - 1;
  static double fpc[5][9], fmc[5][9]; /* mm = -4..4 at index mm+4; keeps values between interfaces */
  double S[5][5], S1[5][5], h0[5];
#define O(s) ((s) + lap - 1)
  for (int i = ia; i <= ib; i++) {
    int sh = scm[i];
    if (sh == 1) {
      for (int mm = G.Ka1; mm <= G.Kb1; mm++)
        for (int m = 0; m < 5; m++) fpc[m][mm + 4] = fp[(long)m * np + O(i + mm)];
      for (int mm = G.Ka2; mm <= G.Kb2; mm++)
        for (int m = 0; m < 5; m++) fmc[m][mm + 4] = fm[(long)m * np + O(i + mm)];
    } else {
      int ka1 = G.Ka1, kb1 = G.Kb1, ka2 = G.Ka2, kb2 = G.Kb2;
      if (sh == 2) { ka1 = -3; kb1 = 3; ka2 = -2; kb2 = 4; }        /* OCFD_readpara.f90:404 */
      else if (sh == 3) { ka1 = -2; kb1 = 2; ka2 = -1; kb2 = 3; }   /* :405 */
      int a = O(i), b = O(i + 1);
      char_matrices((uu[a] + uu[b]) * 0.5, (vv[a] + vv[b]) * 0.5, (ww[a] + ww[b]) * 0.5, (cc1[a] + cc1[b]) * 0.5,
                    (pp[a] + pp[b]) * 0.5, (EE[a] + EE[b]) * 0.5, (A1[a] + A1[b]) * 0.5, (A2[a] + A2[b]) * 0.5,
                    (A3[a] + A3[b]) * 0.5, S, S1);
      for (int mm = ka1; mm <= kb1; mm++) {
        const double* f = fp + O(i + mm);
        for (int r = 0; r < 5; r++)
          fpc[r][mm + 4] = S[r][0] * f[0] + S[r][1] * f[np] + S[r][2] * f[2L * np] + S[r][3] * f[3L * np] +
                           S[r][4] * f[4L * np];
      }
      for (int mm = ka2; mm <= kb2; mm++) {
        const double* f = fm + O(i + mm);
        for (int r = 0; r < 5; r++)
          fmc[r][mm + 4] = S[r][0] * f[0] + S[r][1] * f[np] + S[r][2] * f[2L * np] + S[r][3] * f[3L * np] +
                           S[r][4] * f[4L * np];
      }
    }
    for (int m = 0; m < 5; m++) {
      double h1 = char_flux_x1(&fpc[m][4], i, n, bidx, sh, sb);
      double h2 = char_flux_x2(&fmc[m][4], i, n, bidx, sh, sb);
      h0[m] = h1 + h2;
    }
    for (int r = 0; r < 5; r++)
      hh[(long)r * (n + 1) + i] =
          sh == 1 ? h0[r] : S1[r][0] * h0[0] + S1[r][1] * h0[1] + S1[r][2] * h0[2] + S1[r][3] * h0[3] + S1[r][4] * h0[4];
  }
  /* non-reflecting faces, :288-312 */
  if (bidx[0] == 1 && sb[0] == 0)
    for (int m = 0; m < 5; m++) {
      const double* P = fp + (long)m * np;
      const double* M = fm + (long)m * np;
      hh[(long)m * (n + 1)] = (2.0 * P[O(1)] + 5.0 * P[O(2)] - P[O(3)] + 2.0 * M[O(3)] - 7.0 * M[O(2)] + 11.0 * M[O(1)]) / 6.0;
    }
  if (bidx[1] == 1 && sb[1] == 0)
    for (int m = 0; m < 5; m++) {
      const double* P = fp + (long)m * np;
      const double* M = fm + (long)m * np;
      hh[(long)m * (n + 1) + n] =
          (2.0 * P[O(n - 2)] - 7.0 * P[O(n - 1)] + 11.0 * P[O(n)] + 2.0 * M[O(n)] + 5.0 * M[O(n - 1)] - M[O(n - 2)]) / 6.0;
    }
#undef O
}

/* du_inviscous_Character, OCFD_inviscous_character.f90:3-162 */
static void du_inviscous_character(void) {
  int nmax = G.nx > G.ny ? G.nx : G.ny;
  if (G.nz > nmax) nmax = G.nz;
  int lap = G.lap, npm = nmax + 2 * lap;
  double *Et1 = dalloc(npm, 0), *p1 = dalloc(npm, 0), *c1 = dalloc(npm, 0), *d1 = dalloc(npm, 0);
  double *u1 = dalloc(npm, 0), *v1 = dalloc(npm, 0), *w1 = dalloc(npm, 0);
  double *A1 = dalloc(npm, 0), *A2 = dalloc(npm, 0), *A3 = dalloc(npm, 0);
  double *fp = dalloc(5L * npm, 0), *fm = dalloc(5L * npm, 0);
  double* hh = dalloc(5L * (nmax + 1), 0);
  unsigned char* scm = (unsigned char*)calloc((size_t)nmax + 1, 1);
  int hyb = G.cfg.scheme_invis == OCFD_SCHEME_HYBRID;
  double hx_1 = 1.0 / G.hx, hy_1 = 1.0 / G.hy, hz_1 = 1.0 / G.hz;
  comput_cc();
  for (int dir = 0; dir < 3; dir++) {
    int n = dir == 0 ? G.nx : dir == 1 ? G.ny : G.nz;
    long st = dir == 0 ? 1 : dir == 1 ? G.nxh : (long)G.nxh * G.nyh;
    const double *M1 = dir == 0 ? G.Akx1 : dir == 1 ? G.Aix1 : G.Asx1;
    const double *M2 = dir == 0 ? G.Aky1 : dir == 1 ? G.Aiy1 : G.Asy1;
    const double *M3 = dir == 0 ? G.Akz1 : dir == 1 ? G.Aiz1 : G.Asz1;
    const int* sb = &G.cfg.scheme_boundary[2 * dir];
    int na = dir == 0 ? G.ny : G.nx, nb = dir == 2 ? G.ny : G.nz;
    unsigned char* mask = dir == 0 ? G.maskx : dir == 1 ? G.masky : G.maskz;
    for (int b = 1; b <= nb; b++)
      for (int a = 1; a <= na; a++) {
        int i0 = dir == 0 ? 0 : a, j0 = dir == 0 ? a : dir == 1 ? 0 : b, k0 = dir == 2 ? 0 : b;
        long base = H(i0, j0, k0);
        if (hyb)
          for (int s = 0; s <= n; s++) scm[s] = set_hybrid_scheme(fmax(G.Rhybrid[base + s * st], G.Rhybrid[base + (s + 1) * st]));
        for (int s = 1 - lap; s <= n + lap; s++) {
          long q = base + s * st;
          int o = s + lap - 1;
          Et1[o] = G.Et[q]; p1[o] = G.p[q]; c1[o] = G.cc[q]; d1[o] = G.d[q];
          u1[o] = G.u[q]; v1[o] = G.v[q]; w1[o] = G.w[q];
          A1[o] = M1[q]; A2[o] = M2[q]; A3[o] = M3[q];
        }
        if (G.cfg.flux_splitting == OCFD_SPLIT_SW) split_SW(n, lap, Et1, p1, c1, d1, u1, v1, w1, A1, A2, A3, fp, fm);
        else split_LLF(n, lap, Et1, p1, c1, d1, u1, v1, w1, A1, A2, A3, fp, fm);
        flux_charteric(n, lap, hh, fp, fm, A1, A2, A3, Et1, p1, c1, u1, v1, w1, G.bound_index[dir], scm, sb);
        for (int m = 0; m < 5; m++) {
          const double* h = hh + (long)m * (n + 1);
          for (int s = 1; s <= n; s++) {
            int i = dir == 0 ? s : a, j = dir == 1 ? s : (dir == 0 ? a : b), k = dir == 2 ? s : b;
            long c = I(i, j, k) + (long)m * G.ni;
            if (dir == 0) G.du[c] = (h[s] - h[s - 1]) * hx_1;
            else if (dir == 1) G.du[c] = G.du[c] + (h[s] - h[s - 1]) * hy_1;
            else G.du[c] = -(G.du[c] + (h[s] - h[s - 1]) * hz_1) * G.Ajac[H(i, j, k)];
          }
        }
        if (hyb) {
          for (int s = 0; s <= n; s++) {
            long mi;
            if (dir == 0) mi = s + (long)(G.nx + 1) * ((a - 1) + (long)G.ny * (b - 1));
            else if (dir == 1) mi = (a - 1) + (long)G.nx * (s + (long)(G.ny + 1) * (b - 1));
            else mi = (a - 1) + (long)G.nx * ((b - 1) + (long)G.ny * s);
            mask[mi] = scm[s];
          }
        }
      }
  }
  if (G.cfg.if_mass_force) { /* :148-160 */
    const double* mf = G.cfg.mass_force;
    for (int k = 1; k <= G.nz; k++)
      for (int j = 1; j <= G.ny; j++)
        for (int i = 1; i <= G.nx; i++) {
          long c = I(i, j, k), q = H(i, j, k);
          G.du[c + G.ni] = G.du[c + G.ni] + G.d[q] * mf[0];
          G.du[c + 2 * G.ni] = G.du[c + 2 * G.ni] + G.d[q] * mf[1];
          G.du[c + 3 * G.ni] = G.du[c + 3 * G.ni] + G.d[q] * mf[2];
          G.du[c + 4 * G.ni] = G.du[c + 4 * G.ni] + G.d[q] * (G.u[q] * mf[0] + G.v[q] * mf[1] + G.w[q] * mf[2]);
        }
  }
  free(Et1); free(p1); free(c1); free(d1); free(u1); free(v1); free(w1);
  free(A1); free(A2); free(A3); free(fp); free(fm); free(hh); free(scm);
}

/* ------------------------------------------------------------------ */
/* Chemkin log fit exp(a1 + a2 lnT + a3 lnT^2 + a4 lnT^3) (TRANSPORT_modules.f90:276-277) */
static double chemkin_fit(const double* a, double lT) {
  return exp(a[0] + a[1] * lT + a[2] * (lT * lT) + a[3] * pw3(lT));
}

/* comput_Amu_Amk_AMD, Comput_flow_variables.f90:69-98, with
 * comput_vis_coeff (TRANSPORT_modules.f90:259-306: Sutherland, or Chemkin
 * fits + Wilke), comput_thermal_conductivity (:312-372: Pr, or fits with
 * the 1/2 (sum X k + 1/sum X/k) mixing) and comput_diffusion_coeff
 * (:378-424: Sc, or binary fits with (1-X_k)/sum_j X_j/D_kj atm/p). */
static void comput_Amu_Amk_AMD(void) {
  const double TA = 288.15, TC = 110.4;
  int ns = G.nspec;
#pragma omp parallel
  {
  double* di0 = dalloc(ns, 0);
  double* Xi = dalloc(ns, 0);
  double* mui = dalloc(ns, 0);
  double* ki = dalloc(ns, 0);
  double* D2 = dalloc((long)ns * ns, 0);
#pragma omp for schedule(static)
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        long q = H(i, j, k), c = I(i, j, k);
        double X0 = 0.0;
        for (int m = 0; m < ns; m++) {
          di0[m] = G.di[q + m * G.nh];
          Xi[m] = di0[m] / G.Mi[m];
          X0 = X0 + Xi[m];
        }
        for (int m = 0; m < ns; m++) Xi[m] = Xi[m] / X0;
        double T = G.T[q];
        double amu;
        if (G.cfg.flag_vis == 0) {
          double tr = T / TA;
          amu = G.cfg.amu0 * sqrt(pw3(tr)) * (TA + TC) / (TC + T);
        } else {
          for (int m = 0; m < ns; m++) mui[m] = chemkin_fit(G.mu_an + 4 * m, log(T)) * 0.1;
          double mu = 0.0;
          for (int kk = 0; kk < ns; kk++) {
            double Pk0 = 0.0;
            for (int jj = 0; jj < ns; jj++) {
              double a = 1.0 + sqrt(mui[kk] / mui[jj] * sqrt(G.Mi[jj] / G.Mi[kk]));
              double Pkj = (a * a) / sqrt(8.0 * (1.0 + G.Mi[kk] / G.Mi[jj]));
              Pk0 = Pk0 + Xi[jj] * Pkj;
            }
            mu = mu + Xi[kk] * mui[kk] / Pk0;
          }
          amu = mu;
        }
        double Cp = comput_Cp(G.d[q], di0, T);
        double amk;
        if (G.cfg.flag_cond == 0) {
          amk = amu * Cp / G.cfg.pr;
        } else {
          for (int m = 0; m < ns; m++) ki[m] = chemkin_fit(G.k_an + 4 * m, log(T)) * 1.0e-5;
          double Amk1 = 0.0, Amk2 = 0.0;
          for (int m = 0; m < ns; m++) {
            Amk1 = Amk1 + Xi[m] * ki[m];
            Amk2 = Amk2 + Xi[m] / ki[m];
          }
          amk = 0.5 * (Amk1 + 1.0 / Amk2);
        }
        G.Amu[c] = amu;
        G.Amk[c] = amk;
        if (G.cfg.flag_diff == 0) {
          double d0 = 0.0;
          for (int m = 0; m < ns; m++) d0 = d0 + Xi[m] * G.Mi[m];
          for (int m = 0; m < ns; m++)
            G.AmD[c + m * G.ni] = amu * (1.0 - Xi[m] * G.Mi[m] / d0) / ((1.0 - Xi[m]) * G.cfg.sc * G.d[q] + 1.0e-20);
        } else {
          for (int jj = 0; jj < ns; jj++)
            for (int kk = 0; kk <= jj; kk++) D2[kk * ns + jj] = chemkin_fit(G.D_an + 4 * ((long)kk * ns + jj), log(T)) * 1.0e-4;
          for (int jj = 0; jj < ns; jj++)
            for (int kk = jj + 1; kk < ns; kk++) D2[kk * ns + jj] = D2[jj * ns + kk];
          for (int kk = 0; kk < ns; kk++) {
            double ds = 0.0;
            for (int jj = 0; jj < ns; jj++)
              if (jj != kk) ds = ds + Xi[jj] / D2[kk * ns + jj];
            G.AmD[c + kk * G.ni] = (1.0 - Xi[kk]) / (ds + 1.0e-20) * ATM_PA / G.p[q];
          }
        }
      }
  free(di0); free(Xi); free(mui); free(ki); free(D2);
  }
}

/* du_viscous, OCFD_viscous.f90:27-131 */
static void du_viscous(void) {
  const double tmp2_3 = 2.0 / 3.0;
  orc_dx0(G.u, G.uk); orc_dx0(G.v, G.vk); orc_dx0(G.w, G.wk); orc_dx0(G.T, G.Tk);
  orc_dy0(G.u, G.ui); orc_dy0(G.v, G.vi); orc_dy0(G.w, G.wi); orc_dy0(G.T, G.Ti);
  orc_dz0(G.u, G.us); orc_dz0(G.v, G.vs); orc_dz0(G.w, G.ws); orc_dz0(G.T, G.Ts);
  long nh = G.nh;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        long c = I(i, j, k), q = H(i, j, k);
        double ux = G.uk[c] * G.Akx[q] + G.ui[c] * G.Aix[q] + G.us[c] * G.Asx[q];
        double vx = G.vk[c] * G.Akx[q] + G.vi[c] * G.Aix[q] + G.vs[c] * G.Asx[q];
        double wx = G.wk[c] * G.Akx[q] + G.wi[c] * G.Aix[q] + G.ws[c] * G.Asx[q];
        double Tx = G.Tk[c] * G.Akx[q] + G.Ti[c] * G.Aix[q] + G.Ts[c] * G.Asx[q];
        double uy = G.uk[c] * G.Aky[q] + G.ui[c] * G.Aiy[q] + G.us[c] * G.Asy[q];
        double vy = G.vk[c] * G.Aky[q] + G.vi[c] * G.Aiy[q] + G.vs[c] * G.Asy[q];
        double wy = G.wk[c] * G.Aky[q] + G.wi[c] * G.Aiy[q] + G.ws[c] * G.Asy[q];
        double Ty = G.Tk[c] * G.Aky[q] + G.Ti[c] * G.Aiy[q] + G.Ts[c] * G.Asy[q];
        double uz = G.uk[c] * G.Akz[q] + G.ui[c] * G.Aiz[q] + G.us[c] * G.Asz[q];
        double vz = G.vk[c] * G.Akz[q] + G.vi[c] * G.Aiz[q] + G.vs[c] * G.Asz[q];
        double wz = G.wk[c] * G.Akz[q] + G.wi[c] * G.Aiz[q] + G.ws[c] * G.Asz[q];
        double Tz = G.Tk[c] * G.Akz[q] + G.Ti[c] * G.Aiz[q] + G.Ts[c] * G.Asz[q];
        double div = ux + vy + wz;
        double Amu1 = G.Amu[c], Amuk = G.Amk[c];
        double s11 = (2.0 * ux - tmp2_3 * div) * Amu1;
        double s22 = (2.0 * vy - tmp2_3 * div) * Amu1;
        double s33 = (2.0 * wz - tmp2_3 * div) * Amu1;
        double s12 = (uy + vx) * Amu1, s13 = (uz + wx) * Amu1, s23 = (vz + wy) * Amu1;
        double uu = G.u[q], vv = G.v[q], ww = G.w[q];
        double E1 = uu * s11 + vv * s12 + ww * s13 + Amuk * Tx + G.Qpx[c];
        double E2 = uu * s12 + vv * s22 + ww * s23 + Amuk * Ty + G.Qpy[c];
        double E3 = uu * s13 + vv * s23 + ww * s33 + Amuk * Tz + G.Qpz[c];
        G.Ev1[q] = G.Akx1[q] * s11 + G.Aky1[q] * s12 + G.Akz1[q] * s13;
        G.Ev1[q + nh] = G.Akx1[q] * s12 + G.Aky1[q] * s22 + G.Akz1[q] * s23;
        G.Ev1[q + 2 * nh] = G.Akx1[q] * s13 + G.Aky1[q] * s23 + G.Akz1[q] * s33;
        G.Ev1[q + 3 * nh] = G.Akx1[q] * E1 + G.Aky1[q] * E2 + G.Akz1[q] * E3;
        G.Ev2[q] = G.Aix1[q] * s11 + G.Aiy1[q] * s12 + G.Aiz1[q] * s13;
        G.Ev2[q + nh] = G.Aix1[q] * s12 + G.Aiy1[q] * s22 + G.Aiz1[q] * s23;
        G.Ev2[q + 2 * nh] = G.Aix1[q] * s13 + G.Aiy1[q] * s23 + G.Aiz1[q] * s33;
        G.Ev2[q + 3 * nh] = G.Aix1[q] * E1 + G.Aiy1[q] * E2 + G.Aiz1[q] * E3;
        G.Ev3[q] = G.Asx1[q] * s11 + G.Asy1[q] * s12 + G.Asz1[q] * s13;
        G.Ev3[q + nh] = G.Asx1[q] * s12 + G.Asy1[q] * s22 + G.Asz1[q] * s23;
        G.Ev3[q + 2 * nh] = G.Asx1[q] * s13 + G.Asy1[q] * s23 + G.Asz1[q] * s33;
        G.Ev3[q + 3 * nh] = G.Asx1[q] * E1 + G.Asy1[q] * E2 + G.Asz1[q] * E3;
      }
  for (int m = 0; m < 4; m++) {
    exchange_x(G.Ev1 + m * nh);
    exchange_y(G.Ev2 + m * nh);
    exchange_z(G.Ev3 + m * nh);
    orc_dx0(G.Ev1 + m * nh, G.uk);
    orc_dy0(G.Ev2 + m * nh, G.ui);
    orc_dz0(G.Ev3 + m * nh, G.us);
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= G.nz; k++)
      for (int j = 1; j <= G.ny; j++)
        for (int i = 1; i <= G.nx; i++) {
          long c = I(i, j, k);
          long cm = c + (long)(m + 1) * G.ni;
          G.du[cm] = G.du[cm] + (G.uk[c] + G.ui[c] + G.us[c]) * G.Ajac[H(i, j, k)];
        }
  }
}

/* comput_hi_pgas (CHEM_thermdyn.f90:176-182), comput_hi_rgas (:271-293) */
static double comput_hi(int k, double T) {
  if (!G.cfg.if_real_gas) return 3.5 * G.Ri[k] * T;
  int jc = 0;
  for (int j = 0; j < G.n_tct - 1; j++) if (T > G.tct[j]) jc = j + 1;
  double c0 = sp_c(k, jc, 0), c1 = sp_c(k, jc, 1), c2 = sp_c(k, jc, 2), c3 = sp_c(k, jc, 3), c4 = sp_c(k, jc, 4);
  double ect = G.Ect[k * NTCT_MAX + jc];
  if (T <= G.tct_max)
    return (c0 * T + c1 * (T * T) / 2.0 + c2 * pw3(T) / 3.0 + c3 * pw4(T) / 4.0 + c4 * pw5(T) / 5.0 + ect) * G.Ri[k];
  double tm = G.tct_max;
  double Cp_max = c0 + c1 * tm + c2 * (tm * tm) + c3 * pw3(tm) + c4 * pw4(tm);
  return (c0 * tm + c1 * (tm * tm) / 2.0 + c2 * pw3(tm) / 3.0 + c3 * pw4(tm) / 4.0 + c4 * pw5(tm) / 5.0 + ect +
          Cp_max * (T - tm)) * G.Ri[k];
}

/* du_scaler_invis, OCFD_scaler_inviscous.f90:1-84: scalar Steger-Warming
 * split F+- = (vs +- |vs|)/2 fs with vs = A1 u + A2 v + A3 w, the same
 * reconstructions and closures as the flow, Dus = -(sum of differences) Ajac */
static void du_scaler_invis(const double* fs, double* Dus) {
  int nmax = G.nx > G.ny ? G.nx : G.ny;
  if (G.nz > nmax) nmax = G.nz;
  int lap = G.lap, npm = nmax + 2 * lap;
  double *fp = dalloc(npm, 0), *fm = dalloc(npm, 0), *hh1 = dalloc(nmax + 1, 0), *hh2 = dalloc(nmax + 1, 0);
  unsigned char* scm = (unsigned char*)calloc((size_t)nmax + 1, 1);
  int hyb = G.cfg.scheme_invis == OCFD_SCHEME_HYBRID;
  double hx_1 = 1.0 / G.hx, hy_1 = 1.0 / G.hy, hz_1 = 1.0 / G.hz;
  for (int dir = 0; dir < 3; dir++) {
    int n = dir == 0 ? G.nx : dir == 1 ? G.ny : G.nz;
    long st = dir == 0 ? 1 : dir == 1 ? G.nxh : (long)G.nxh * G.nyh;
    const double *M1 = dir == 0 ? G.Akx1 : dir == 1 ? G.Aix1 : G.Asx1;
    const double *M2 = dir == 0 ? G.Aky1 : dir == 1 ? G.Aiy1 : G.Asy1;
    const double *M3 = dir == 0 ? G.Akz1 : dir == 1 ? G.Aiz1 : G.Asz1;
    const int* sb = &G.cfg.scheme_boundary[2 * dir];
    int na = dir == 0 ? G.ny : G.nx, nb = dir == 2 ? G.ny : G.nz;
    for (int b = 1; b <= nb; b++)
      for (int a = 1; a <= na; a++) {
        int i0 = dir == 0 ? 0 : a, j0 = dir == 0 ? a : dir == 1 ? 0 : b, k0 = dir == 2 ? 0 : b;
        long base = H(i0, j0, k0);
        if (hyb)
          for (int s = 0; s <= n; s++) scm[s] = set_hybrid_scheme(fmax(G.Rhybrid[base + s * st], G.Rhybrid[base + (s + 1) * st]));
        for (int s = 1 - lap; s <= n + lap; s++) {
          long q = base + s * st;
          int o = s + lap - 1;
          double vs = M1[q] * G.u[q] + M2[q] * G.v[q] + M3[q] * G.w[q];
          fp[o] = 0.5 * (vs + fabs(vs)) * fs[q];
          fm[o] = 0.5 * (vs - fabs(vs)) * fs[q];
        }
        flux1(fp, hh1, n, lap, G.bound_index[dir], scm, sb);
        flux2(fm, hh2, n, lap, G.bound_index[dir], scm, sb);
        for (int s = 1; s <= n; s++) {
          int i = dir == 0 ? s : a, j = dir == 1 ? s : (dir == 0 ? a : b), k = dir == 2 ? s : b;
          long c = I(i, j, k);
          if (dir == 0) Dus[c] = (hh1[s] - hh1[s - 1] + hh2[s] - hh2[s - 1]) * hx_1;
          else if (dir == 1) Dus[c] = Dus[c] + (hh1[s] - hh1[s - 1] + hh2[s] - hh2[s - 1]) * hy_1;
          else Dus[c] = -(Dus[c] + (hh1[s] - hh1[s - 1] + hh2[s] - hh2[s - 1]) * hz_1) * G.Ajac[H(i, j, k)];
        }
      }
  }
  free(fp); free(fm); free(hh1); free(hh2); free(scm);
}

/* du_scaler_viscous, OCFD_scaler_viscous.f90:4-79: diffusive flux
 * rho D_k grad(c_k) of species k (c_k = fs/d), its enthalpy flux added to
 * Qp (zeroed at the first species), Dus += div(flux) Ajac */
static void du_scaler_viscous(int ks, const double* fs, double* Dus) {
  long nh = G.nh;
  if (ks == 0)
    for (long c = 0; c < G.ni; c++) { G.Qpx[c] = 0.0; G.Qpy[c] = 0.0; G.Qpz[c] = 0.0; }
  for (long q = 0; q < nh; q++) G.Ev1[q] = fs[q] / G.d[q];
  orc_dx0(G.Ev1, G.uk);
  orc_dy0(G.Ev1, G.ui);
  orc_dz0(G.Ev1, G.us);
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        long c = I(i, j, k), q = H(i, j, k);
        double hi = comput_hi(ks, G.T[q]);
        double cix = G.uk[c] * G.Akx[q] + G.ui[c] * G.Aix[q] + G.us[c] * G.Asx[q];
        double ciy = G.uk[c] * G.Aky[q] + G.ui[c] * G.Aiy[q] + G.us[c] * G.Asy[q];
        double ciz = G.uk[c] * G.Akz[q] + G.ui[c] * G.Aiz[q] + G.us[c] * G.Asz[q];
        double dD = G.d[q] * G.AmD[c + ks * G.ni];
        double E1 = dD * cix, E2 = dD * ciy, E3 = dD * ciz;
        G.Qpx[c] = G.Qpx[c] + E1 * hi;
        G.Qpy[c] = G.Qpy[c] + E2 * hi;
        G.Qpz[c] = G.Qpz[c] + E3 * hi;
        G.Ev1[q] = E1 * G.Akx1[q] + E2 * G.Aky1[q] + E3 * G.Akz1[q];
        G.Ev1[q + nh] = E1 * G.Aix1[q] + E2 * G.Aiy1[q] + E3 * G.Aiz1[q];
        G.Ev1[q + 2 * nh] = E1 * G.Asx1[q] + E2 * G.Asy1[q] + E3 * G.Asz1[q];
      }
  exchange_x(G.Ev1);
  exchange_y(G.Ev1 + nh);
  exchange_z(G.Ev1 + 2 * nh);
  orc_dx0(G.Ev1, G.uk);
  orc_dy0(G.Ev1 + nh, G.ui);
  orc_dz0(G.Ev1 + 2 * nh, G.us);
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        long c = I(i, j, k);
        Dus[c] = Dus[c] + (G.uk[c] + G.ui[c] + G.us[c]) * G.Ajac[H(i, j, k)];
      }
}

/* Scalar_time_adv_RK3, OCFD_Time_Adv.f90:134-166 */
static void scalar_time_adv_rk3(int krk, int ks, const double* Dus, double dt) {
  const double Ralfa[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0}, Rbeta[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        long c = I(i, j, k), q = H(i, j, k) + (long)ks * G.nh, cn = c + (long)ks * G.ni;
        if (krk == 1) G.di[q] = G.din[cn] + dt * Dus[c];
        else G.di[q] = Ralfa[krk - 1] * G.din[cn] + Rbeta[krk - 1] * (G.di[q] + dt * Dus[c]);
      }
}

/* OCFD_time_adv_RK3, OCFD_Time_Adv.f90:98-130 */
static void time_adv_rk3(int krk, double dt) {
  const double Ralfa[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0}, Rbeta[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
  long n = 5 * G.ni;
  if (krk == 1) {
#pragma omp parallel for schedule(static)
    for (long c = 0; c < n; c++) G.f[c] = G.fn[c] + dt * G.du[c];
  } else {
#pragma omp parallel for schedule(static)
    for (long c = 0; c < n; c++) G.f[c] = Ralfa[krk - 1] * G.fn[c] + Rbeta[krk - 1] * (G.f[c] + dt * G.du[c]);
  }
}

/* handle_NegativeEt, OCFD_handle_NegativeT.f90:1-20 */
static int handle_negative_et(int i, int j, int k) {
  long q = H(i, j, k);
  long sx = 1, sy = G.nxh, sz = (long)G.nxh * G.nyh;
  double Et0 = G.Et[q];
  G.Et[q] = (G.Et[q + sx] + G.Et[q - sx] + G.Et[q + sy] + G.Et[q - sy] + G.Et[q + sz] + G.Et[q - sz]) / 6.0;
  G.f[I(i, j, k) + 4 * G.ni] = G.f[I(i, j, k) + 4 * G.ni] + (G.Et[q] - Et0);
  G.num_negt = G.num_negt + 1;
  return G.num_negt > 10;
}

/* comput_T_P, Comput_flow_variables.f90:32-46 */
static void comput_T_P(void) {
#pragma omp parallel
  {
    double* di0 = dalloc(G.nspec, 0);
#pragma omp for schedule(static)
    for (int k = 1; k <= G.nz; k++)
      for (int j = 1; j <= G.ny; j++)
        for (int i = 1; i <= G.nx; i++) {
          long q = H(i, j, k);
          for (int m = 0; m < G.nspec; m++) di0[m] = G.di[q + m * G.nh];
          G.T[q] = comput_T(di0, G.Et[q]);
          G.p[q] = comput_P(di0, G.T[q]);
        }
    free(di0);
  }
}

/* comput_flow_variables + Rescaling_di, Comput_flow_variables.f90:3-28,101-121 */
static int comput_flow_variables(void) {
  int stop = 0, negative = 0;
  G.num_negt = 0;
  /* The repair of a cell with Et <= 0 reads its neighbours' Et, new for
   * those earlier in (k,j,i) order and old for later ones, so the loop runs
   * in order whenever any cell needs it; a parallel pre-scan with the same
   * arithmetic finds out. */
#pragma omp parallel for schedule(static) reduction(| : negative)
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        long c = I(i, j, k);
        double f1 = G.f[c], f2 = G.f[c + G.ni], f3 = G.f[c + 2 * G.ni], f4 = G.f[c + 3 * G.ni], f5 = G.f[c + 4 * G.ni];
        double u = f2 / f1, v = f3 / f1, w = f4 / f1;
        if (f5 - (f2 * u + f3 * v + f4 * w) * 0.5 <= 0) negative |= 1;
      }
  if (!negative) {
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= G.nz; k++)
      for (int j = 1; j <= G.ny; j++)
        for (int i = 1; i <= G.nx; i++) {
          long c = I(i, j, k), q = H(i, j, k);
          double f1 = G.f[c], f2 = G.f[c + G.ni], f3 = G.f[c + 2 * G.ni], f4 = G.f[c + 3 * G.ni], f5 = G.f[c + 4 * G.ni];
          G.d[q] = f1;
          G.u[q] = f2 / f1;
          G.v[q] = f3 / f1;
          G.w[q] = f4 / f1;
          G.Et[q] = f5 - (f2 * G.u[q] + f3 * G.v[q] + f4 * G.w[q]) * 0.5;
        }
  } else {
    for (int k = 1; k <= G.nz; k++)
      for (int j = 1; j <= G.ny; j++)
        for (int i = 1; i <= G.nx; i++) {
          long c = I(i, j, k), q = H(i, j, k);
          double f1 = G.f[c], f2 = G.f[c + G.ni], f3 = G.f[c + 2 * G.ni], f4 = G.f[c + 3 * G.ni], f5 = G.f[c + 4 * G.ni];
          G.d[q] = f1;
          G.u[q] = f2 / f1;
          G.v[q] = f3 / f1;
          G.w[q] = f4 / f1;
          G.Et[q] = f5 - (f2 * G.u[q] + f3 * G.v[q] + f4 * G.w[q]) * 0.5;
          if (G.Et[q] <= 0) stop |= handle_negative_et(i, j, k);
        }
  }
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        long q = H(i, j, k);
        double d0 = 0.0;
        for (int m = 0; m < G.nspec; m++) {
          if (G.di[q + m * G.nh] < 0.0) G.di[q + m * G.nh] = 0.0;
          d0 = d0 + G.di[q + m * G.nh];
        }
        double tmp = G.d[q] / d0;
        for (int m = 0; m < G.nspec; m++) G.di[q + m * G.nh] = G.di[q + m * G.nh] * tmp;
      }
  comput_T_P();
  return stop;
}

/* ------------------------------------------------------------------ */
/* element m (d,u,v,w,T,di(:)) of BC%flow_inlet0d/1d/2d or flow_upper0d/1d/2d
 * at face position (a, b): flow_0d(m), flow_1d(a,m), flow_2d(a,b,m) */
static double bl_value(const double* r, int kind, int a, int b, int na, int nb, int m) {
  if (kind == 0) return r[m];
  if (kind == 1) return r[(a - 1) + (long)na * m];
  return r[(a - 1) + (long)na * ((b - 1) + (long)nb * m)];
}

/* ocfd_bc_BoundaryLayer, OCFD_BC_Boundarylayers.f90:20-243 (inlet 0-D, 1-D or
 * 2-D data, :50-97; upper 0-D, 1-D or 2-D data where not non-reflecting,
 * none for BcData_upper = -1, :100-140; wall; outlet types 0/1/2). */
static void bc_boundarylayer(void) {
  int nx = G.nx, ny = G.ny, nz = G.nz, ns = G.nspec;
  long nh = G.nh;
  int inl = (int)lround(G.cfg.bc_para[0]), upp = (int)lround(G.cfg.bc_para[1]);
  int outlet = (int)lround(G.cfg.bc_para[3]);
  double Tw = G.cfg.bc_para[4], xinit = G.cfg.bc_para[5], disA = G.cfg.bc_para[7];
  if (inl >= 0 && inl <= 2) {
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++) {
        long q = H(1, j, k);
        const double* r = G.bl_inlet;
        G.d[q] = bl_value(r, inl, j, k, ny, nz, 0); G.u[q] = bl_value(r, inl, j, k, ny, nz, 1);
        G.v[q] = bl_value(r, inl, j, k, ny, nz, 2); G.w[q] = bl_value(r, inl, j, k, ny, nz, 3);
        G.T[q] = bl_value(r, inl, j, k, ny, nz, 4);
        for (int m = 0; m < ns; m++) G.di[q + m * nh] = bl_value(r, inl, j, k, ny, nz, 5 + m);
      }
  }
  for (int k = 1; k <= nz; k++)
    for (int i = 1; i <= nx; i++) {
      if (G.bl_nonref[(i - 1) + (long)nx * (k - 1)] == 0 && upp >= 0 && upp <= 2) {
        long q = H(i, ny, k);
        const double* r = G.bl_upper;
        G.d[q] = bl_value(r, upp, i, k, nx, nz, 0); G.u[q] = bl_value(r, upp, i, k, nx, nz, 1);
        G.v[q] = bl_value(r, upp, i, k, nx, nz, 2); G.w[q] = bl_value(r, upp, i, k, nx, nz, 3);
        G.T[q] = bl_value(r, upp, i, k, nx, nz, 4);
        for (int m = 0; m < ns; m++) G.di[q + m * nh] = bl_value(r, upp, i, k, nx, nz, 5 + m);
      }
    }
  double ht = G.bl_ht;
  double* di1 = dalloc(ns, 0); double* di2 = dalloc(ns, 0); double* di3 = dalloc(ns, 0);
  for (int k = 1; k <= nz; k++)
    for (int i = 1; i <= nx; i++) {
      long q1 = H(i, 1, k), q2 = H(i, 2, k), q3 = H(i, 3, k);
      if (G.Axx[q1] < xinit) {
        G.d[q1] = G.d[q2]; G.u[q1] = G.u[q2]; G.v[q1] = G.v[q2]; G.w[q1] = G.w[q2]; G.T[q1] = G.T[q2];
        for (int m = 0; m < ns; m++) G.di[q1 + m * nh] = G.di[q2 + m * nh];
      } else {
        double tmp = 1.0 / sqrt(SQ(G.Aix[q1]) + SQ(G.Aiy[q1]) + SQ(G.Aiz[q1]));
        double A0 = disA * ht * G.bl_wall_pertb[(i - 1) + (long)nx * (k - 1)];
        G.u[q1] = A0 * G.Aix[q1] * tmp;
        G.v[q1] = A0 * G.Aiy[q1] * tmp;
        G.w[q1] = A0 * G.Aiz[q1] * tmp;
        if (Tw > 0) G.T[q1] = Tw;
        else G.T[q1] = (4.0 * G.T[q2] - G.T[q3]) / 3.0;
        for (int m = 0; m < ns; m++) { di2[m] = G.di[q2 + m * nh]; di3[m] = G.di[q3 + m * nh]; }
        double p2 = comput_P(di2, G.T[q2]);
        double p3 = comput_P(di3, G.T[q3]);
        double p1 = (4.0 * p2 - p3) / 3.0;
        for (int m = 0; m < ns; m++) di1[m] = di2[m] / G.d[q2];
        double d1 = comput_d(di1, G.T[q1], p1);
        G.d[q1] = d1;
        for (int m = 0; m < ns; m++) G.di[q1 + m * nh] = di1[m] * d1;
      }
    }
  if (outlet == 1 || outlet == 2) {
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++) {
        long q = H(nx, j, k), qa = H(nx - 1, j, k), qb = H(nx - 2, j, k);
        double* arr[5] = {G.d, G.u, G.v, G.w, G.T};
        for (int a = 0; a < 5; a++) arr[a][q] = outlet == 1 ? arr[a][qa] : 2.0 * arr[a][qa] - arr[a][qb];
        for (int m = 0; m < ns; m++)
          G.di[q + m * nh] = outlet == 1 ? G.di[qa + m * nh] : 2.0 * G.di[qa + m * nh] - G.di[qb + m * nh];
      }
  }
  free(di1); free(di2); free(di3);
}

/* ocfd_bc_jet, OCFD_BC_jet.f90:2-96: 2-D inflow data on i=1 (Dirichlet,
 * or with BC_Para(1)=1 the pressure taken from i=2 where the inflow is
 * subsonic), zeroth-order outflow on i=nx */
static void bc_jet(void) {
  int ny = G.ny, nz = G.nz, ns = G.nspec;
  long nh = G.nh, pl = (long)ny * nz;
  int iflag = (int)lround(G.cfg.bc_para[0]);
  double* di1 = dalloc(ns, 0);
  double* ai1 = dalloc(ns, 0);
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      long q = H(1, j, k), r = (j - 1) + (long)ny * (k - 1);
      const double* in = G.inlet2d;
      if (iflag == 0) {
        G.d[q] = in[r]; G.u[q] = in[r + pl]; G.v[q] = in[r + 2 * pl]; G.w[q] = in[r + 3 * pl]; G.T[q] = in[r + 4 * pl];
        for (int m = 0; m < ns; m++) G.di[q + m * nh] = in[r + (5L + m) * pl];
      } else {
        G.u[q] = in[r + pl]; G.v[q] = in[r + 2 * pl]; G.w[q] = in[r + 3 * pl]; G.T[q] = in[r + 4 * pl];
        double d1 = in[r];
        double d0 = 0.0;
        for (int m = 0; m < ns; m++) { di1[m] = in[r + (5L + m) * pl]; d0 = d0 + di1[m]; }
        for (int m = 0; m < ns; m++) ai1[m] = di1[m] / d0;
        double p1 = comput_P(di1, G.T[q]);
        double Et1 = comput_E(di1, G.T[q]);
        double gamma = p1 / Et1 + 1.0;            /* comput_soundspeed, CHEM_thermdyn.f90:412-418 */
        double c1 = sqrt(gamma * p1 / d1);
        if (G.u[q] < c1) {
          G.p[q] = G.p[H(2, j, k)];
          double d1new = comput_d(ai1, G.T[q], G.p[q]);
          G.d[q] = d1new;
          for (int m = 0; m < ns; m++) G.di[q + m * nh] = d1new * ai1[m];
        } else {
          G.d[q] = d1;
          for (int m = 0; m < ns; m++) G.di[q + m * nh] = di1[m];
        }
      }
    }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      long q = H(G.nx, j, k), qa = H(G.nx - 1, j, k);
      G.d[q] = G.d[qa]; G.u[q] = G.u[qa]; G.v[q] = G.v[qa]; G.w[q] = G.w[qa];
      G.T[q] = G.T[qa]; G.p[q] = G.p[qa]; G.Et[q] = G.Et[qa];
      for (int m = 0; m < ns; m++) G.di[q + m * nh] = G.di[qa + m * nh];
    }
  free(di1); free(ai1);
}

/* face re-synthesis of one face cell, OCFD_BC.f90:33-56 (same body per face) */
static void face_update(int i, int j, int k, double* di1) {
  long q = H(i, j, k), c = I(i, j, k);
  double d0 = 0.0;
  for (int m = 0; m < G.nspec; m++) { di1[m] = G.di[q + m * G.nh]; d0 = d0 + di1[m]; }
  double dtmp = G.d[q] / d0;
  for (int m = 0; m < G.nspec; m++) { di1[m] = di1[m] * dtmp; G.di[q + m * G.nh] = di1[m]; }
  G.Et[q] = comput_E(di1, G.T[q]);
  G.p[q] = comput_P(di1, G.T[q]);
  double dd = G.d[q], uu = G.u[q], vv = G.v[q], ww = G.w[q];
  G.f[c] = dd;
  G.f[c + G.ni] = dd * uu;
  G.f[c + 2 * G.ni] = dd * vv;
  G.f[c + 3 * G.ni] = dd * ww;
  G.f[c + 4 * G.ni] = G.Et[q] + dd * (uu * uu + vv * vv + ww * ww) * 0.5;
}

static void flow_ghost_extent_2nd(void);

/* OCFD_bc, OCFD_BC.f90:2-188 (one rank: every face is a rank face) */
static int bc_all(void) {
  if (G.cfg.ibc == OCFD_BC_BOUNDARYLAYER) bc_boundarylayer();
  else if (G.cfg.ibc == OCFD_BC_JET) {
    if (!G.inlet2d) return set_err("oracle: jet inflow data not set");
    bc_jet();
  } else if (G.cfg.ibc != OCFD_BC_NONE) return set_err("oracle: unsupported boundary condition");
  double* di1 = dalloc(G.nspec, 0);
  for (int k = 1; k <= G.nz; k++) for (int j = 1; j <= G.ny; j++) face_update(1, j, k, di1);
  for (int k = 1; k <= G.nz; k++) for (int j = 1; j <= G.ny; j++) face_update(G.nx, j, k, di1);
  for (int k = 1; k <= G.nz; k++) for (int i = 1; i <= G.nx; i++) face_update(i, 1, k, di1);
  for (int k = 1; k <= G.nz; k++) for (int i = 1; i <= G.nx; i++) face_update(i, G.ny, k, di1);
  for (int j = 1; j <= G.ny; j++) for (int i = 1; i <= G.nx; i++) face_update(i, j, 1, di1);
  for (int j = 1; j <= G.ny; j++) for (int i = 1; i <= G.nx; i++) face_update(i, j, G.nz, di1);
  free(di1);
  flow_ghost_extent_2nd(); /* Flow_Ghost_boundary, OCFD_BC.f90:186 */
  return 0;
}

/* ------------------------------------------------------------------ */
/* comput_Rhybrid, OCFD_SetHybrid.f90:17-143 (patches: Hybrid_para(4)=0 only) */
static int comput_rhybrid(int ordered) {
  (void)ordered;
  int sensor = (int)lround(G.cfg.hybrid_para[0]);
  if (lround(G.cfg.hybrid_para[3]) != 0) return set_err("oracle: hybrid patches (Hybrid_para(4)>0) unsupported");
  if (sensor == 1) {
    long sx = 1, sy = G.nxh, sz = (long)G.nxh * G.nyh;
    for (int k = 1; k <= G.nz; k++)
      for (int j = 1; j <= G.ny; j++)
        for (int i = 1; i <= G.nx; i++) {
          long q = H(i, j, k);
          const double* P = G.p;
          G.Rhybrid[q] = fabs((P[q - sx] - 2.0 * P[q] + P[q + sx]) / (P[q - sx] + 2.0 * P[q] + P[q + sx])) +
                         fabs((P[q - sy] - 2.0 * P[q] + P[q + sy]) / (P[q - sy] + 2.0 * P[q] + P[q + sy])) +
                         fabs((P[q - sz] - 2.0 * P[q] + P[q + sz]) / (P[q - sz] + 2.0 * P[q] + P[q + sz]));
        }
  } else {
    double *pk = dalloc(G.ni, 0), *pi = dalloc(G.ni, 0), *ps = dalloc(G.ni, 0);
    orc_dx0(G.p, pk); orc_dy0(G.p, pi); orc_dz0(G.p, ps);
    double dp0 = 0.0;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= G.nz; k++)
      for (int j = 1; j <= G.ny; j++)
        for (int i = 1; i <= G.nx; i++) {
          long c = I(i, j, k), q = H(i, j, k);
          double px = pk[c] * G.Akx[q] + pi[c] * G.Aix[q] + ps[c] * G.Asx[q];
          double py = pk[c] * G.Aky[q] + pi[c] * G.Aiy[q] + ps[c] * G.Asy[q];
          double pz = pk[c] * G.Akz[q] + pi[c] * G.Aiz[q] + ps[c] * G.Asz[q];
          G.Rhybrid[q] = sqrt(px * px + py * py + pz * pz);
        }
    /* the global sum in the reference's sequential (k,j,i) order */
    for (int k = 1; k <= G.nz; k++)
      for (int j = 1; j <= G.ny; j++)
        for (int i = 1; i <= G.nx; i++) dp0 = dp0 + G.Rhybrid[H(i, j, k)];
    double dp_av = dp0 / (1.0 * G.nx * G.ny * G.nz);
    double dp_av1 = 1.0 / (dp_av + 1.0e-30);
    G.sensor_dp_av = dp_av;
    for (int k = 1; k <= G.nz; k++)
      for (int j = 1; j <= G.ny; j++)
        for (int i = 1; i <= G.nx; i++) G.Rhybrid[H(i, j, k)] = G.Rhybrid[H(i, j, k)] * dp_av1;
    free(pk); free(pi); free(ps);
  }
  exchange_xyz(G.Rhybrid);
  return 0;
}

/* ------------------------------------------------------------------ */
/* flow_time_advance, OCFD_Time_Adv.f90:7-92 (non-characteristic) */
int orc_flow_time_advance(double dt1) {
  int ns = G.nspec;
  memcpy(G.fn, G.f, sizeof(double) * 5 * G.ni);
  for (int m = 0; m < ns; m++)
    for (int k = 1; k <= G.nz; k++)
      for (int j = 1; j <= G.ny; j++)
        for (int i = 1; i <= G.nx; i++) G.din[I(i, j, k) + (long)m * G.ni] = G.di[H(i, j, k) + (long)m * G.nh];
  for (int krk = 1; krk <= 3; krk++) {
    comput_Amu_Amk_AMD();
    exchange_xyz(G.d); exchange_xyz(G.u); exchange_xyz(G.v); exchange_xyz(G.w);
    exchange_xyz(G.T); exchange_xyz(G.p); exchange_xyz(G.Et);
    if (ns == 1) memcpy(G.di, G.d, sizeof(double) * G.nh);
    else for (int m = 0; m < ns; m++) exchange_xyz(G.di + (long)m * G.nh);
    if (krk == 1 && G.cfg.scheme_invis == OCFD_SCHEME_HYBRID)
      if (comput_rhybrid(1)) return 1;
    if (ns != 1)
      for (int m = 0; m < ns; m++) {
        du_scaler_invis(G.di + (long)m * G.nh, G.dus);
        if (G.cfg.if_viscous) du_scaler_viscous(m, G.di + (long)m * G.nh, G.dus);
        scalar_time_adv_rk3(krk, m, G.dus, dt1);
      }
    if (G.cfg.if_scheme_character) du_inviscous_character();
    else du_inviscous();
    if (G.cfg.if_viscous) du_viscous();
    time_adv_rk3(krk, dt1);
    if (comput_flow_variables()) return set_err("oracle: more than 10 negative-Et cells");
    if (bc_all()) return 1;
  }
  return 0;
}

/* Chemical_time_advance, OCFD_Time_Adv.f90:169-196: chemical_timeAdv on every
 * interior cell (mode = CHEM_TimeAdv), f5 += Et1 - Et0, Et = Et1.  The
 * mechanism is the chemistry restatement's (orc_chem_set_mechanism). */
int orc_chem_timeadv(int mode, double* di, double* E, double dt);
int orc_chemical_time_advance(double dt2, int mode) {
  int ns = G.nspec;
  double* di1 = dalloc(ns, 0);
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        long q = H(i, j, k), c = I(i, j, k);
        for (int m = 0; m < ns; m++) di1[m] = G.di[q + (long)m * G.nh];
        double Et1 = G.Et[q], Et0 = Et1;
        if (orc_chem_timeadv(mode, di1, &Et1, dt2)) { free(di1); return set_err("oracle: chemical_timeAdv failed"); }
        for (int m = 0; m < ns; m++) G.di[q + (long)m * G.nh] = di1[m];
        G.f[c + 4 * G.ni] = G.f[c + 4 * G.ni] + (Et1 - Et0);
        G.Et[q] = Et1;
      }
  free(di1);
  return 0;
}

int orc_comput_T_P(void) { comput_T_P(); return 0; }

/* one time step of NS_solver, OCFD_NS_Solver.f90:26-37: flow dt/2, M chemistry
 * substeps of dt/M and comput_T_P, flow dt/2 */
int orc_ns_step(double dt, int nsub, int mode, int reaction) {
  double dt1 = 0.5 * dt, dt2 = dt / nsub;
  if (orc_flow_time_advance(dt1)) return 1;
  if (reaction) {
    for (int m = 0; m < nsub; m++)
      if (orc_chemical_time_advance(dt2, mode)) return 1;
    comput_T_P();
  }
  return orc_flow_time_advance(dt1);
}

/* the species right-hand sides of one stage without the update (for
 * comparisons of pieces): dus [N_SPEC][interior], and Qp after the last species */
int orc_du_species(double* dus_out) {
  for (int m = 0; m < G.nspec; m++) {
    du_scaler_invis(G.di + (long)m * G.nh, G.dus);
    if (G.cfg.if_viscous) du_scaler_viscous(m, G.di + (long)m * G.nh, G.dus);
    memcpy(dus_out + (long)m * G.ni, G.dus, sizeof(double) * G.ni);
  }
  return 0;
}

/* Qpx, Qpy, Qpz after the last species of orc_du_species: [3][interior] */
int orc_download_qp(double* out) {
  memcpy(out, G.Qpx, sizeof(double) * G.ni);
  memcpy(out + G.ni, G.Qpy, sizeof(double) * G.ni);
  memcpy(out + 2 * G.ni, G.Qpz, sizeof(double) * G.ni);
  return 0;
}

/* pieces, for unit-level comparisons */
int orc_comput_amu_amk(void) { comput_Amu_Amk_AMD(); return 0; }
int orc_du_inviscous(void) {
  if (G.cfg.if_scheme_character) du_inviscous_character();
  else du_inviscous();
  return 0;
}
int orc_du_viscous(void) { du_viscous(); return 0; }
int orc_time_adv_rk3(int krk, double dt) { time_adv_rk3(krk, dt); return 0; }
int orc_comput_flow_variables(void) { return comput_flow_variables(); }
int orc_bc(void) { return bc_all(); }
int orc_comput_rhybrid(void) { return comput_rhybrid(1); }
int orc_exchange_all(void) {
  exchange_xyz(G.d); exchange_xyz(G.u); exchange_xyz(G.v); exchange_xyz(G.w);
  exchange_xyz(G.T); exchange_xyz(G.p); exchange_xyz(G.Et);
  for (int m = 0; m < G.nspec; m++) exchange_xyz(G.di + (long)m * G.nh);
  return 0;
}
/* show_flow_msg, OCFD_Time_Adv.f90:200-239: interior means of the total
 * energy f5/J and of the kinetic energy d|u|^2/2/J (sequential sums in
 * (k,j,i) order, as one rank's loop before the MPI_REDUCE). */
int orc_show_flow_msg(double* E0) {
  double e1 = 0.0, e2 = 0.0;
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        long q = H(i, j, k), c = I(i, j, k);
        double tmp0 = 1.0 / G.Ajac[q];
        e1 = e1 + G.f[c + 4 * G.ni] * tmp0;
        e2 = e2 + G.d[q] * (G.u[q] * G.u[q] + G.v[q] * G.v[q] + G.w[q] * G.w[q]) * 0.5 * tmp0;
      }
  double nt = (double)G.nx * G.ny * G.nz;
  E0[0] = e1 / nt;
  E0[1] = e2 / nt;
  return 0;
}

/* ana_box, OCFD_ANA.f90:197-303: statistics of isotropic turbulence from the
 * current primitives and Amu (u exchanged, its derivatives by OCFD_dx0/dy0/dz0
 * with Scheme_Vis).  out[0..17] = Re_lamda, Re_lamdax, Amt, vv_rms, u_rms,
 * v_rms, w_rms, u_skewness, u_flatness, ux_skewness, ux_flatness, E_kinetic,
 * d_rms, T_aver, Amu_aver, d_aver, Ux_rms, Alamdax (the order of the
 * reference's statistics.dat line after tt). */
int orc_ana_box(double* out) {
  exchange_xyz(G.u);
  orc_dx0(G.u, G.uk);
  orc_dy0(G.u, G.ui);
  orc_dz0(G.u, G.us);
  double a[13] = {0};
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        long q = H(i, j, k), c = I(i, j, k);
        double ux = G.uk[c] * G.Akx[q] + G.ui[c] * G.Aix[q] + G.us[c] * G.Asx[q];
        double gamma = G.p[q] / G.Et[q] + 1.0;
        double c1 = sqrt(gamma * G.p[q] / G.d[q]);
        double uu = G.u[q];
        a[7] = a[7] + c1;
        a[6] = a[6] + G.d[q];
        a[8] = a[8] + G.Amu[c];
        a[10] = a[10] + G.T[q];
        a[0] = a[0] + uu * uu;
        a[1] = a[1] + G.v[q] * G.v[q];
        a[2] = a[2] + G.w[q] * G.w[q];
        a[3] = a[3] + ux * ux;
        a[4] = a[4] + (ux * ux) * ux;
        a[5] = a[5] + (ux * ux) * (ux * ux);
        a[11] = a[11] + (uu * uu) * uu;
        a[12] = a[12] + (uu * uu) * (uu * uu);
        a[9] = a[9] + (uu * uu + G.v[q] * G.v[q] + G.w[q] * G.w[q]) * G.d[q] * 0.5;
      }
  double nt = (double)G.nx * G.ny * G.nz;
  for (int m = 0; m < 13; m++) a[m] = a[m] / nt;
  double u_rms = a[0], v_rms = a[1], w_rms = a[2], Ux_rms = a[3], Ux_sk = a[4], Ux_fl = a[5], d_aver = a[6];
  double c_aver = a[7], Amu_aver = a[8], E_k = a[9], T_aver = a[10], u_sk = a[11], u_fl = a[12];
  double VV_rms = sqrt((u_rms + v_rms + w_rms) / 3.0);
  double Amt = sqrt(u_rms + v_rms + w_rms) / c_aver;
  u_rms = sqrt(u_rms); v_rms = sqrt(v_rms); w_rms = sqrt(w_rms); Ux_rms = sqrt(Ux_rms);
  double Alamda = VV_rms / Ux_rms, Alamdax = u_rms / Ux_rms;
  double Re_lamda = VV_rms * Alamda * d_aver / Amu_aver, Re_lamdax = VV_rms * Alamdax * d_aver / Amu_aver;
  Ux_sk = Ux_sk / ((Ux_rms * Ux_rms) * Ux_rms);
  Ux_fl = Ux_fl / ((Ux_rms * Ux_rms) * (Ux_rms * Ux_rms));
  u_sk = u_sk / ((u_rms * u_rms) * u_rms);
  u_fl = u_fl / ((u_rms * u_rms) * (u_rms * u_rms));
  double d_rms = 0.0;
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        double x = G.d[H(i, j, k)] - d_aver;
        d_rms = d_rms + x * x;
      }
  d_rms = sqrt(d_rms / nt);
  double r[18] = {Re_lamda, Re_lamdax, Amt, VV_rms, u_rms, v_rms, w_rms, u_sk, u_fl, Ux_sk, Ux_fl, E_k, d_rms,
                  T_aver, Amu_aver, d_aver, Ux_rms, Alamdax};
  for (int m = 0; m < 18; m++) out[m] = r[m];
  return 0;
}

/* ------------------------------------------------------------------ */
/* filtering, OCFD_filtering.f90:4-362 (Bogey & Bailly 9-point filter and
 * the shock-capturing filter), called after Istep has been advanced
 * (OCFD_NS_Solver.f90:45-48).  filter = Para%Filter(100,10) column-major:
 * Filter(k,m) at filter[(m-1)*100 + (k-1)]. One rank: the global index
 * ranges are the local ones, clipped to the box. */
static void filter_fo9p(int dir, double* f0, double s0, const int* lo, const int* hi) {
  const double d0 = 0.243527493120, d1 = -0.204788880640, d2 = 0.120007591680, d3 = -0.045211119360,
               d4 = 0.008228661760;
  int n = dir == 0 ? G.nx : dir == 1 ? G.ny : G.nz;
  int b[3] = {lo[0], lo[1], lo[2]}, e[3] = {hi[0], hi[1], hi[2]};
  if (!G.cfg.periodic[dir]) { /* filter_x3d :91-93 */
    if (b[dir] < 6) b[dir] = 6;
    if (e[dir] > n - 5) e[dir] = n - 5;
  }
  long st = dir == 0 ? 1 : dir == 1 ? G.nxh : (long)G.nxh * G.nyh;
  for (int m = 0; m < 5; m++) {
    for (int k = 1; k <= G.nz; k++)
      for (int j = 1; j <= G.ny; j++)
        for (int i = 1; i <= G.nx; i++) f0[H(i, j, k)] = G.f[I(i, j, k) + (long)m * G.ni];
    if (dir == 0) exchange_x(f0); else if (dir == 1) exchange_y(f0); else exchange_z(f0);
    for (int k = b[2]; k <= e[2]; k++)
      for (int j = b[1]; j <= e[1]; j++)
        for (int i = b[0]; i <= e[0]; i++) {
          const double* F = f0 + H(i, j, k);
#define F0(o) F[(long)(o) * st]
          G.f[I(i, j, k) + (long)m * G.ni] =
              F0(0) - s0 * (d0 * F0(0) + d1 * (F0(-1) + F0(1)) + d2 * (F0(-2) + F0(2)) + d3 * (F0(-3) + F0(3)) +
                            d4 * (F0(-4) + F0(4)));
#undef F0
        }
  }
}

static int filter_shock(int dir, double* f0, double s0, double rth, const int* lo, const int* hi) {
  const double c1 = -0.210383, c2 = 0.039617;
  int n = dir == 0 ? G.nx : dir == 1 ? G.ny : G.nz;
  int b[3] = {lo[0], lo[1], lo[2]}, e[3] = {hi[0], hi[1], hi[2]};
  if (!G.cfg.periodic[dir]) { /* filter_x3d_shock :204-206 */
    if (b[dir] < 4) b[dir] = 4;
    if (e[dir] > n - 3) e[dir] = n - 3;
  }
  if (comput_flow_variables()) return set_err("oracle: more than 10 negative-Et cells");
  if (dir == 0) exchange_x(G.p); else if (dir == 1) exchange_y(G.p); else exchange_z(G.p);
  long st = dir == 0 ? 1 : dir == 1 ? G.nxh : (long)G.nxh * G.nyh;
  double* sc0 = dalloc(n + 2, 0.0); /* sc0(0:n+1) */
  for (int m = 0; m < 5; m++) {
    for (int k = 1; k <= G.nz; k++)
      for (int j = 1; j <= G.ny; j++)
        for (int i = 1; i <= G.nx; i++) f0[H(i, j, k)] = G.f[I(i, j, k) + (long)m * G.ni];
    if (dir == 0) exchange_x(f0); else if (dir == 1) exchange_y(f0); else exchange_z(f0);
    /* the two directions across the pencil, in the reference's loop order */
    int a0 = dir == 0 ? 1 : 0, a1 = dir == 2 ? 1 : 2;
    for (int pb = b[a1]; pb <= e[a1]; pb++)
      for (int pa = b[a0]; pa <= e[a0]; pa++) {
        int ijk[3];
        ijk[a0] = pa; ijk[a1] = pb;
        for (int s = b[dir] - 1; s <= e[dir] + 1; s++) {
          ijk[dir] = s;
          const double* P = G.p + H(ijk[0], ijk[1], ijk[2]);
#define PP(o) P[(long)(o) * st]
          double dp0 = 0.25 * (-PP(1) + 2.0 * PP(0) - PP(-1));
          double dp1 = 0.25 * (-PP(2) + 2.0 * PP(1) - PP(0));
          double dp2 = 0.25 * (-PP(0) + 2.0 * PP(-1) - PP(-2));
          double ri = 0.5 * (SQ(dp0 - dp1) + SQ(dp0 - dp2)) / (PP(0) * PP(0)) + 1.0e-16;
#undef PP
          sc0[s] = 0.5 * (1.0 - rth / ri + fabs(1.0 - rth / ri));
        }
        for (int s = b[dir]; s <= e[dir]; s++) {
          ijk[dir] = s;
          double sc1 = 0.5 * (sc0[s] + sc0[s + 1]);
          double sc2 = 0.5 * (sc0[s] + sc0[s - 1]);
          const double* F = f0 + H(ijk[0], ijk[1], ijk[2]);
#define F0(o) F[(long)(o) * st]
          G.f[I(ijk[0], ijk[1], ijk[2]) + (long)m * G.ni] =
              F0(0) - s0 * (sc1 * (c1 * (F0(1) - F0(0)) + c2 * (F0(2) - F0(-1))) -
                            sc2 * (c1 * (F0(0) - F0(-1)) + c2 * (F0(1) - F0(-2))));
#undef F0
        }
      }
  }
  free(sc0);
  return 0;
}

int orc_filtering(int istep, int nfiltering, const double* filter) {
  int any = 0;
  for (int m = 0; m < nfiltering; m++)
    if (istep % (int)lround(filter[m * 100]) == 0) any = 1;
  if (!any) return 0;
  double* f0 = dalloc(G.nh, 0.0);
  int rc = 0;
  for (int m = 0; m < nfiltering && !rc; m++) {
    const double* F = filter + m * 100;
    if (istep % (int)lround(F[0]) != 0) continue;
    int on[3] = {(int)lround(F[1]), (int)lround(F[2]), (int)lround(F[3])};
    int lo[3], hi[3], n[3] = {G.nx, G.ny, G.nz};
    for (int d = 0; d < 3; d++) {
      lo[d] = (int)lround(F[4 + 2 * d]);
      hi[d] = (int)lround(F[5 + 2 * d]);
      if (lo[d] < 1) lo[d] = 1;
      if (hi[d] > n[d]) hi[d] = n[d];
    }
    int scheme = (int)lround(F[10]);
    double s0 = F[11], rth = F[12];
    for (int d = 0; d < 3 && !rc; d++) {
      if (on[d] != 1) continue;
      if (scheme == 1) filter_fo9p(d, f0, s0, lo, hi);
      else if (scheme == 2) rc = filter_shock(d, f0, s0, rth, lo, hi);
    }
  }
  free(f0);
  return rc;
}

int orc_negative_et_count(void) { return G.num_negt; }
double orc_sensor_mean(void) { return G.sensor_dp_av; }

/* ------------------------------------------------------------------ */
/* Mesh and metrics: init_mesh / comput_Jacobian3d / updata_Axyz_periodic,
 * OCFD_mesh.f90:1-185.  Coordinates are halo arrays; interior is read. */
/* ------------------------------------------------------------------ */
/* Ghost cells (Ghost_Cell(1:6) = 1 or 2), OCFD_GhostCell.f90.  A face f
 * (0 i-, 1 i+, 2 j-, 3 j+, 4 k-, 5 k+) covers the LAP ghost layers over the
 * interior tangential range.  gpos(f, l, a, b) is the halo index of ghost
 * layer l (1..LAP) at tangential (a, b); bpos the boundary point; mpos the
 * mirrored interior point 2*boundary - ghost. */
static long gidx(int dir, int s, int a, int b) {
  return dir == 0 ? H(s, a, b) : dir == 1 ? H(a, s, b) : H(a, b, s);
}
static int face_n(int dir) { return dir == 0 ? G.nx : dir == 1 ? G.ny : G.nz; }
static void face_tan(int dir, int* na, int* nb) {
  *na = dir == 0 ? G.ny : G.nx;
  *nb = dir == 2 ? G.ny : G.nz;
}

/* Jac_Ghost_Extent_1st, OCFD_GhostCell.f90:17-152 */
static void jac_ghost_1st(int f) {
  int dir = f / 2, hi = f & 1, n = face_n(dir), na, nb;
  face_tan(dir, &na, &nb);
  double* M[10] = {G.Akx, G.Aky, G.Akz, G.Aix, G.Aiy, G.Aiz, G.Asx, G.Asy, G.Asz, G.Ajac};
  double* X[3] = {G.Axx, G.Ayy, G.Azz};
  for (int b = 1; b <= nb; b++)
    for (int a = 1; a <= na; a++)
      for (int l = 1; l <= G.lap; l++) {
        int sg = hi ? n + l : 1 - l, sb0 = hi ? n : 1, sb1 = hi ? n - 1 : 2;
        long q = gidx(dir, sg, a, b), q0 = gidx(dir, sb0, a, b), q1 = gidx(dir, sb1, a, b);
        for (int m = 0; m < 10; m++) M[m][q] = M[m][q0];
        for (int c = 0; c < 3; c++)
          X[c][q] = hi ? X[c][q0] + (double)(sg - n) * (X[c][q0] - X[c][q1]) : X[c][q0] + (1.0 - sg) * (X[c][q0] - X[c][q1]);
      }
}

/* one-sided / central first derivative inside the ghost block range [r0, r1]
 * (comput_Jacobian3d_Ghost, OCFD_GhostCell.f90:506-559) */
static double ghost_deriv(const double* X, long q, long st, int p, int r0, int r1, double h) {
#define XX(o) X[q + (long)(o) * st]
  if (p == r0) return (-3.0 * XX(0) + 4.0 * XX(1) - XX(2)) / (2.0 * h);
  if (p == r1) return (XX(-2) - 4.0 * XX(-1) + 3.0 * XX(0)) / (2.0 * h);
  if (p == r0 + 1 || p == r1 - 1) return (XX(1) - XX(-1)) / (2.0 * h);
  return (8.0 * (XX(1) - XX(-1)) - (XX(2) - XX(-2))) / (12.0 * h);
#undef XX
}

/* Jac_Ghost_Extent_2nd + comput_Jacobian3d_Ghost, OCFD_GhostCell.f90:368-580 */
static void jac_ghost_2nd(int f) {
  int dir = f / 2, hi = f & 1, n = face_n(dir), na, nb;
  face_tan(dir, &na, &nb);
  double* X[3] = {G.Axx, G.Ayy, G.Azz};
  for (int b = 1; b <= nb; b++)
    for (int a = 1; a <= na; a++)
      for (int l = 1; l <= G.lap; l++) {
        int sg = hi ? n + l : 1 - l, sb0 = hi ? n : 1, sm = hi ? 2 * n - sg : 2 - sg;
        long q = gidx(dir, sg, a, b), q0 = gidx(dir, sb0, a, b), qm = gidx(dir, sm, a, b);
        for (int c = 0; c < 3; c++) X[c][q] = 2.0 * X[c][q0] - X[c][qm];
      }
  long st[3] = {1, G.nxh, (long)G.nxh * G.nyh};
  double h[3] = {G.hx, G.hy, G.hz};
  int lo[3] = {1, 1, 1}, up[3] = {G.nx, G.ny, G.nz};
  lo[dir] = hi ? n : 1 - G.lap;
  up[dir] = hi ? n + G.lap : 1;
  for (int b = 1; b <= nb; b++)
    for (int a = 1; a <= na; a++)
      for (int l = 1; l <= G.lap; l++) {
        int sg = hi ? n + l : 1 - l;
        int pos[3];
        pos[dir] = sg;
        pos[dir == 0 ? 1 : 0] = a;
        pos[dir == 2 ? 1 : 2] = b;
        long q = H(pos[0], pos[1], pos[2]);
        double dx[3][3]; /* dx[c][e]: d(coordinate c)/d(index e) */
        for (int c = 0; c < 3; c++)
          for (int e = 0; e < 3; e++) dx[c][e] = ghost_deriv(X[c], q, st[e], pos[e], lo[e], up[e], h[e]);
        double xi1 = dx[0][0], yi1 = dx[1][0], zi1 = dx[2][0];
        double xj1 = dx[0][1], yj1 = dx[1][1], zj1 = dx[2][1];
        double xk1 = dx[0][2], yk1 = dx[1][2], zk1 = dx[2][2];
        double Jac1 = 1.0 / (xi1 * yj1 * zk1 + yi1 * zj1 * xk1 + zi1 * xj1 * yk1 - zi1 * yj1 * xk1 - yi1 * xj1 * zk1 -
                             xi1 * zj1 * yk1);
        G.Ajac[q] = Jac1;
        G.Akx[q] = Jac1 * (yj1 * zk1 - zj1 * yk1);
        G.Aky[q] = Jac1 * (zj1 * xk1 - xj1 * zk1);
        G.Akz[q] = Jac1 * (xj1 * yk1 - yj1 * xk1);
        G.Aix[q] = Jac1 * (yk1 * zi1 - zk1 * yi1);
        G.Aiy[q] = Jac1 * (zk1 * xi1 - xk1 * zi1);
        G.Aiz[q] = Jac1 * (xk1 * yi1 - yk1 * xi1);
        G.Asx[q] = Jac1 * (yi1 * zj1 - zi1 * yj1);
        G.Asy[q] = Jac1 * (zi1 * xj1 - xi1 * zj1);
        G.Asz[q] = Jac1 * (xi1 * yj1 - yi1 * xj1);
      }
}

/* Flow_Ghost_Extent_2nd, OCFD_GhostCell.f90:166-364 */
static void flow_ghost_extent_2nd(void) {
  const double s1 = 0.8, s2 = 1.2;
  int ns = G.nspec;
  double* di1 = dalloc(ns, 0);
  for (int f = 0; f < 6; f++) {
    int gc = G.cfg.ghost_cell[f];
    if (gc != 1 && gc != 2) continue;
    int dir = f / 2, hi = f & 1, n = face_n(dir), na, nb;
    face_tan(dir, &na, &nb);
    for (int b = 1; b <= nb; b++)
      for (int a = 1; a <= na; a++)
        for (int l = 1; l <= G.lap; l++) {
          int sg = hi ? n + l : 1 - l, sb0 = hi ? n : 1, sm = hi ? 2 * n - sg : 2 - sg;
          long q = gidx(dir, sg, a, b), q0 = gidx(dir, sb0, a, b), qm = gidx(dir, sm, a, b);
          G.u[q] = 2.0 * G.u[q0] - G.u[qm];
          G.v[q] = 2.0 * G.v[q0] - G.v[qm];
          G.w[q] = 2.0 * G.w[q0] - G.w[qm];
          G.d[q] = 2.0 * G.d[q0] - G.d[qm];
          G.T[q] = 2.0 * G.T[q0] - G.T[qm];
          if (G.d[q] < s1 * G.d[q0]) G.d[q] = s1 * G.d[q0];
          if (G.d[q] > s2 * G.d[q0]) G.d[q] = s2 * G.d[q0];
          if (G.T[q] < s1 * G.T[q0]) G.T[q] = s1 * G.T[q0];
          if (G.T[q] > s2 * G.T[q0]) G.T[q] = s2 * G.T[q0];
          double d0 = 0.0;
          for (int m = 0; m < ns; m++) {
            di1[m] = 2.0 * G.di[q0 + (long)m * G.nh] - G.di[qm + (long)m * G.nh];
            d0 = d0 + di1[m];
          }
          double dtmp = G.d[q] / d0;
          for (int m = 0; m < ns; m++) {
            di1[m] = di1[m] * dtmp;
            G.di[q + (long)m * G.nh] = di1[m];
          }
          G.Et[q] = comput_E(di1, G.T[q]);
          G.p[q] = comput_P(di1, G.T[q]);
        }
  }
  free(di1);
}

int orc_set_mesh(const double* axx, const double* ayy, const double* azz, const double* span) {
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        long q = H(i, j, k);
        G.Axx[q] = axx[q]; G.Ayy[q] = ayy[q]; G.Azz[q] = azz[q];
      }
  exchange_xyz(G.Axx); exchange_xyz(G.Ayy); exchange_xyz(G.Azz);
  double* A[3] = {G.Axx, G.Ayy, G.Azz};
  for (int c = 0; c < 3; c++) {
    if (G.cfg.periodic[0])
      for (int k = 1; k <= G.nz; k++)
        for (int j = 1; j <= G.ny; j++)
          for (int l = 1; l <= G.lap; l++) {
            A[c][H(1 - l, j, k)] = A[c][H(1 - l, j, k)] - span[c];
            A[c][H(G.nx + l, j, k)] = A[c][H(G.nx + l, j, k)] + span[c];
          }
    if (G.cfg.periodic[1])
      for (int k = 1; k <= G.nz; k++)
        for (int l = 1; l <= G.lap; l++)
          for (int i = 1; i <= G.nx; i++) {
            A[c][H(i, 1 - l, k)] = A[c][H(i, 1 - l, k)] - span[3 + c];
            A[c][H(i, G.ny + l, k)] = A[c][H(i, G.ny + l, k)] + span[3 + c];
          }
    if (G.cfg.periodic[2])
      for (int l = 1; l <= G.lap; l++)
        for (int j = 1; j <= G.ny; j++)
          for (int i = 1; i <= G.nx; i++) {
            A[c][H(i, j, 1 - l)] = A[c][H(i, j, 1 - l)] - span[6 + c];
            A[c][H(i, j, G.nz + l)] = A[c][H(i, j, G.nz + l)] + span[6 + c];
          }
  }
  long ni = G.ni;
  double* t = dalloc(9 * ni, 0);
  double *xi = t, *xj = t + ni, *xk = t + 2 * ni, *yi = t + 3 * ni, *yj = t + 4 * ni, *yk = t + 5 * ni;
  double *zi = t + 6 * ni, *zj = t + 7 * ni, *zk = t + 8 * ni;
  orc_dx0(G.Axx, xi); orc_dx0(G.Ayy, yi); orc_dx0(G.Azz, zi);
  orc_dy0(G.Axx, xj); orc_dy0(G.Ayy, yj); orc_dy0(G.Azz, zj);
  orc_dz0(G.Axx, xk); orc_dz0(G.Ayy, yk); orc_dz0(G.Azz, zk);
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        long c = I(i, j, k), q = H(i, j, k);
        double xi1 = xi[c], xj1 = xj[c], xk1 = xk[c], yi1 = yi[c], yj1 = yj[c], yk1 = yk[c];
        double zi1 = zi[c], zj1 = zj[c], zk1 = zk[c];
        double Jac1 = 1.0 / (xi1 * yj1 * zk1 + yi1 * zj1 * xk1 + zi1 * xj1 * yk1 - zi1 * yj1 * xk1 - yi1 * xj1 * zk1 -
                             xi1 * zj1 * yk1);
        G.Ajac[q] = Jac1;
        G.Akx[q] = Jac1 * (yj1 * zk1 - zj1 * yk1);
        G.Aky[q] = Jac1 * (zj1 * xk1 - xj1 * zk1);
        G.Akz[q] = Jac1 * (xj1 * yk1 - yj1 * xk1);
        G.Aix[q] = Jac1 * (yk1 * zi1 - zk1 * yi1);
        G.Aiy[q] = Jac1 * (zk1 * xi1 - xk1 * zi1);
        G.Aiz[q] = Jac1 * (xk1 * yi1 - yk1 * xi1);
        G.Asx[q] = Jac1 * (yi1 * zj1 - zi1 * yj1);
        G.Asy[q] = Jac1 * (zi1 * xj1 - xi1 * zj1);
        G.Asz[q] = Jac1 * (xi1 * yj1 - yi1 * xj1);
      }
  free(t);
  double* M[10] = {G.Akx, G.Aky, G.Akz, G.Aix, G.Aiy, G.Aiz, G.Asx, G.Asy, G.Asz, G.Ajac};
  for (int a = 0; a < 10; a++) exchange_xyz(M[a]);
  /* Jac_Ghost_boundary, OCFD_mesh.f90:36 (first-order faces, then second-order) */
  for (int f = 0; f < 6; f++)
    if (G.cfg.ghost_cell[f] == 1) jac_ghost_1st(f);
  for (int f = 0; f < 6; f++)
    if (G.cfg.ghost_cell[f] == 2) jac_ghost_2nd(f);
  double* M1[9] = {G.Akx1, G.Aky1, G.Akz1, G.Aix1, G.Aiy1, G.Aiz1, G.Asx1, G.Asy1, G.Asz1};
  for (long q = 0; q < G.nh; q++)
    for (int a = 0; a < 9; a++) M1[a][q] = M[a][q] / G.Ajac[q];
  return 0;
}

/* ------------------------------------------------------------------ */
/* Setup: allocate_flow_data + set_parameters (OCFD_allocate_data.f90:2-68,
 * OCFD_readpara.f90:347-431). */
static void free_all(void) {
  double** arrs[] = {&G.f, &G.fn, &G.du, &G.di, &G.din, &G.d, &G.u, &G.v, &G.w, &G.T, &G.p, &G.Et, &G.cc,
                     &G.Amu, &G.Amk, &G.AmD, &G.Axx, &G.Ayy, &G.Azz, &G.Akx, &G.Aky, &G.Akz, &G.Aix, &G.Aiy,
                     &G.Aiz, &G.Asx, &G.Asy, &G.Asz, &G.Ajac, &G.Akx1, &G.Aky1, &G.Akz1, &G.Aix1, &G.Aiy1,
                     &G.Aiz1, &G.Asx1, &G.Asy1, &G.Asz1, &G.Rhybrid, &G.Ev1, &G.Ev2, &G.Ev3, &G.uk, &G.ui,
                     &G.us, &G.vk, &G.vi, &G.vs, &G.wk, &G.wi, &G.ws, &G.Tk, &G.Ti, &G.Ts, &G.Qpx, &G.Qpy,
                     &G.Qpz, &G.Mi, &G.Ri, &G.coef, &G.Ect, &G.bl_inlet, &G.bl_upper, &G.bl_wall_pertb,
                     &G.dus, &G.mu_an, &G.k_an, &G.D_an, &G.inlet2d};
  for (size_t a = 0; a < sizeof arrs / sizeof arrs[0]; a++) { free(*arrs[a]); *arrs[a] = NULL; }
  free(G.maskx); free(G.masky); free(G.maskz); free(G.bl_nonref);
  G.maskx = G.masky = G.maskz = NULL; G.bl_nonref = NULL;
}

int orc_init(const ocfd_cfg* cfg) {
  free_all();
  G.cfg = *cfg;
  if (cfg->lap != 4) return set_err("oracle: LAP must be 4");
  G.nx = cfg->nx; G.ny = cfg->ny; G.nz = cfg->nz; G.lap = cfg->lap; G.nspec = cfg->n_spec;
  G.nxh = G.nx + 2 * G.lap; G.nyh = G.ny + 2 * G.lap; G.nzh = G.nz + 2 * G.lap;
  G.nh = (long)G.nxh * G.nyh * G.nzh; G.ni = (long)G.nx * G.ny * G.nz;
  long nh = G.nh, ni = G.ni;
  G.hx = 1.0 / (G.nx - 1.0); G.hy = 1.0 / (G.ny - 1.0); G.hz = 1.0 / (G.nz - 1.0);
  for (int d = 0; d < 3; d++) {
    G.bound_index[d][0] = cfg->periodic[d] ? 0 : 1;
    G.bound_index[d][1] = cfg->periodic[d] ? 0 : 1;
  }
  switch (cfg->scheme_invis) {
    case OCFD_SCHEME_WENO5: G.Ka1 = -2; G.Kb1 = 2; G.Ka2 = -1; G.Kb2 = 3; break;
    case OCFD_SCHEME_WENO7: G.Ka1 = -3; G.Kb1 = 3; G.Ka2 = -2; G.Kb2 = 4; break;
    case OCFD_SCHEME_OMP6: case OCFD_SCHEME_UD7L: case OCFD_SCHEME_HYBRID:
      G.Ka1 = -3; G.Kb1 = 4; G.Ka2 = -3; G.Kb2 = 4; break;
    default: return set_err("oracle: unsupported inviscid scheme");
  }
  {
    /* set_scheme_para, OCFD_readpara.f90:424-431 */
    const double UD7[8] = {-3.0, 25.0, -101.0, 319.0, 214.0, -38.0, 4.0, 0.0};
    const double CD8[8] = {-3.0, 29.0, -139.0, 533.0, 533.0, -139.0, 29.0, -3.0};
    for (int a = 0; a < 8; a++) G.ud7l[a] = cfg->ud7l_diss * UD7[a] / 420.0 + (1.0 - cfg->ud7l_diss) * CD8[a] / 840.0;
  }
  G.f = dalloc(5 * ni, 0); G.fn = dalloc(5 * ni, 0); G.du = dalloc(5 * ni, 0);
  G.Amu = dalloc(ni, 0); G.Amk = dalloc(ni, 0); G.AmD = dalloc(ni * G.nspec, 0);
  G.d = dalloc(nh, 1.0); G.u = dalloc(nh, 1.0); G.v = dalloc(nh, 1.0); G.w = dalloc(nh, 1.0);
  G.T = dalloc(nh, 300.0); G.p = dalloc(nh, 1.0); G.cc = dalloc(nh, 1.0); G.Et = dalloc(nh, 1.0);
  G.di = dalloc(nh * G.nspec, 1.0); G.din = dalloc(ni * G.nspec, 1.0);
  G.Axx = dalloc(nh, 1.0); G.Ayy = dalloc(nh, 1.0); G.Azz = dalloc(nh, 1.0);
  G.Akx = dalloc(nh, 1.0); G.Aky = dalloc(nh, 1.0); G.Akz = dalloc(nh, 1.0);
  G.Aix = dalloc(nh, 1.0); G.Aiy = dalloc(nh, 1.0); G.Aiz = dalloc(nh, 1.0);
  G.Asx = dalloc(nh, 1.0); G.Asy = dalloc(nh, 1.0); G.Asz = dalloc(nh, 1.0); G.Ajac = dalloc(nh, 1.0);
  G.Akx1 = dalloc(nh, 1.0); G.Aky1 = dalloc(nh, 1.0); G.Akz1 = dalloc(nh, 1.0);
  G.Aix1 = dalloc(nh, 1.0); G.Aiy1 = dalloc(nh, 1.0); G.Aiz1 = dalloc(nh, 1.0);
  G.Asx1 = dalloc(nh, 1.0); G.Asy1 = dalloc(nh, 1.0); G.Asz1 = dalloc(nh, 1.0);
  G.Rhybrid = dalloc(nh, 0.0);
  G.maskx = (unsigned char*)calloc((size_t)(G.nx + 1) * G.ny * G.nz, 1);
  G.masky = (unsigned char*)calloc((size_t)G.nx * (G.ny + 1) * G.nz, 1);
  G.maskz = (unsigned char*)calloc((size_t)G.nx * G.ny * (G.nz + 1), 1);
  G.Ev1 = dalloc(4 * nh, 0); G.Ev2 = dalloc(4 * nh, 0); G.Ev3 = dalloc(4 * nh, 0);
  G.uk = dalloc(ni, 0); G.ui = dalloc(ni, 0); G.us = dalloc(ni, 0); G.vk = dalloc(ni, 0);
  G.vi = dalloc(ni, 0); G.vs = dalloc(ni, 0); G.wk = dalloc(ni, 0); G.wi = dalloc(ni, 0);
  G.ws = dalloc(ni, 0); G.Tk = dalloc(ni, 0); G.Ti = dalloc(ni, 0); G.Ts = dalloc(ni, 0);
  G.Qpx = dalloc(ni, 0); G.Qpy = dalloc(ni, 0); G.Qpz = dalloc(ni, 0);
  G.dus = dalloc(ni, 0);
  G.Mi = dalloc(G.nspec, 0.02897); G.Ri = dalloc(G.nspec, 0);
  for (int k = 0; k < G.nspec; k++) G.Ri[k] = R0_GAS / G.Mi[k];
  G.coef = dalloc((long)G.nspec * NTCT_MAX * 7, 0); G.Ect = dalloc((long)G.nspec * NTCT_MAX, 0);
  {
    /* init_bcdata_boundary (OCFD_BC_Boundarylayers.f90:247-430): 0-D, 1-D or
     * 2-D inlet / upper data by BC_Para(1) / BC_Para(2), variable slowest */
    int ki = (int)lround(G.cfg.bc_para[0]), ku = (int)lround(G.cfg.bc_para[1]);
    int bl = G.cfg.ibc == OCFD_BC_BOUNDARYLAYER;
    G.bl_nin = (5L + G.nspec) * (bl && ki == 2 ? (long)G.ny * G.nz : bl && ki == 1 ? (long)G.ny : 1L);
    G.bl_nup = (5L + G.nspec) * (bl && ku == 2 ? (long)G.nx * G.nz : bl && ku == 1 ? (long)G.nx : 1L);
    G.bl_inlet = dalloc(G.bl_nin, 0); G.bl_upper = dalloc(G.bl_nup, 0);
  }
  G.bl_wall_pertb = dalloc((long)G.nx * G.nz, 0);
  G.bl_nonref = (int*)calloc((size_t)G.nx * G.nz, sizeof(int));
  G.bl_ht = 1.0;
  G.err[0] = 0;
  return 0;
}

/* Chemkin transport fits (read_vis/conduct/diffusion_coefficients,
 * TRANSPORT_modules.f90:80-253): mu_an, k_an [N_SPEC][4], D_an
 * [N_SPEC][N_SPEC][4]; any may be NULL when its flag is 0 */
int orc_set_transport(const double* mu_an, const double* k_an, const double* D_an) {
  int ns = G.nspec;
  free(G.mu_an); free(G.k_an); free(G.D_an);
  G.mu_an = dalloc(4L * ns, 0); G.k_an = dalloc(4L * ns, 0); G.D_an = dalloc(4L * ns * ns, 0);
  if (mu_an) memcpy(G.mu_an, mu_an, sizeof(double) * 4 * ns);
  if (k_an) memcpy(G.k_an, k_an, sizeof(double) * 4 * ns);
  if (D_an) memcpy(G.D_an, D_an, sizeof(double) * 4 * ns * ns);
  return 0;
}

/* jet inflow data flow_inlet2d(j,k,m) (init_bcdata_boundary with 2-D data):
 * column-major (ny, nz, 5+N_SPEC), m = d, u, v, w, T, di(1..N_SPEC) */
int orc_set_bc_jet(const double* inlet2d) {
  long n = (long)G.ny * G.nz * (5 + G.nspec);
  free(G.inlet2d);
  G.inlet2d = dalloc(n, 0);
  memcpy(G.inlet2d, inlet2d, sizeof(double) * n);
  return 0;
}

/* read_Spec (CHEM_thermdyn.f90:16-51), with Ect by enthalpy continuity */
int orc_set_species(int n_tct, const double* tct, double tct_max, const double* mi, const double* coef) {
  if (n_tct < 1 || n_tct > NTCT_MAX) return set_err("oracle: bad N_Tct");
  G.n_tct = n_tct;
  for (int j = 0; j < n_tct - 1; j++) G.tct[j] = tct ? tct[j] : 0.0;
  G.tct_max = tct_max;
  for (int k = 0; k < G.nspec; k++) {
    G.Mi[k] = mi[k];
    G.Ri[k] = R0_GAS / G.Mi[k];
    if (coef)
      for (int j = 0; j < n_tct; j++)
        for (int c = 0; c < 7; c++) G.coef[((long)k * NTCT_MAX + j) * 7 + c] = coef[((long)k * n_tct + j) * 7 + c];
    G.Ect[k * NTCT_MAX] = 0.0;
    for (int j = 0; j < n_tct - 1; j++) {
      double t = G.tct[j];
      G.Ect[k * NTCT_MAX + j + 1] =
          -((sp_c(k, j + 1, 0) - sp_c(k, j, 0)) * t + (sp_c(k, j + 1, 1) - sp_c(k, j, 1)) * (t * t) / 2.0 +
            (sp_c(k, j + 1, 2) - sp_c(k, j, 2)) * pw3(t) / 3.0 +
            (sp_c(k, j + 1, 3) - sp_c(k, j, 3)) * pw4(t) / 4.0 +
            (sp_c(k, j + 1, 4) - sp_c(k, j, 4)) * pw5(t) / 5.0) +
          G.Ect[k * NTCT_MAX + j];
    }
  }
  return 0;
}

int orc_set_bc_boundarylayer(const double* inlet, const double* upper, const int* nonref, const double* pertb) {
  memcpy(G.bl_inlet, inlet, sizeof(double) * (size_t)G.bl_nin);
  if (upper) memcpy(G.bl_upper, upper, sizeof(double) * (size_t)G.bl_nup);
  memcpy(G.bl_nonref, nonref, sizeof(int) * (size_t)G.nx * G.nz);
  memcpy(G.bl_wall_pertb, pertb, sizeof(double) * (size_t)G.nx * G.nz);
  return 0;
}
int orc_set_bc_time(double ht) { G.bl_ht = ht; return 0; }

/* field access: same identifiers and layouts as the HIP library */
static double* field_ptr(int id, long* n) {
  long nh = G.nh, ni = G.ni;
  switch (id) {
    case OCFD_FIELD_F: *n = 5 * ni; return G.f;
    case OCFD_FIELD_FN: *n = 5 * ni; return G.fn;
    case OCFD_FIELD_DU: *n = 5 * ni; return G.du;
    case OCFD_FIELD_DI: *n = nh * G.nspec; return G.di;
    case OCFD_FIELD_DIN: *n = ni * G.nspec; return G.din;
    case OCFD_FIELD_D: *n = nh; return G.d;
    case OCFD_FIELD_U: *n = nh; return G.u;
    case OCFD_FIELD_V: *n = nh; return G.v;
    case OCFD_FIELD_W: *n = nh; return G.w;
    case OCFD_FIELD_T: *n = nh; return G.T;
    case OCFD_FIELD_P: *n = nh; return G.p;
    case OCFD_FIELD_ET: *n = nh; return G.Et;
    case OCFD_FIELD_AMU: *n = ni; return G.Amu;
    case OCFD_FIELD_AMK: *n = ni; return G.Amk;
    case OCFD_FIELD_RHYBRID: *n = nh; return G.Rhybrid;
    case OCFD_FIELD_AMD: *n = ni * G.nspec; return G.AmD;
  }
  double* M[19] = {G.Akx, G.Aky, G.Akz, G.Aix, G.Aiy, G.Aiz, G.Asx, G.Asy, G.Asz, G.Ajac,
                   G.Akx1, G.Aky1, G.Akz1, G.Aix1, G.Aiy1, G.Aiz1, G.Asx1, G.Asy1, G.Asz1};
  if (id >= OCFD_FIELD_AKX && id <= OCFD_FIELD_ASZ1) { *n = nh; return M[id - OCFD_FIELD_AKX]; }
  *n = 0;
  return NULL;
}
int orc_upload_field(int id, const double* host) {
  long n; double* p = field_ptr(id, &n);
  if (!p) return set_err("oracle: bad field id");
  memcpy(p, host, sizeof(double) * (size_t)n);
  return 0;
}
int orc_download_field(int id, double* host) {
  long n; double* p = field_ptr(id, &n);
  if (!p) return set_err("oracle: bad field id");
  memcpy(host, p, sizeof(double) * (size_t)n);
  return 0;
}
int orc_download_mask(int dir, unsigned char* host) {
  size_t n = dir == 0 ? (size_t)(G.nx + 1) * G.ny * G.nz : dir == 1 ? (size_t)G.nx * (G.ny + 1) * G.nz
                                                                    : (size_t)G.nx * G.ny * (G.nz + 1);
  memcpy(host, dir == 0 ? G.maskx : dir == 1 ? G.masky : G.maskz, n);
  return 0;
}

/* ------------------------------------------------------------------ */
/* 1-D entry points used by the scheme-level tests: v has n+2*lap entries
 * (v(1-lap:n+lap)), hh has n+1 entries.  bidx[2], sb[2], scm[n+1]. */
int orc_recon_flux1(int scheme, const double* v, double* hh, int n, const int* bidx, const unsigned char* scm,
                    const int* sb) {
  int saved = G.cfg.scheme_invis, ka1 = G.Ka1, kb1 = G.Kb1, ka2 = G.Ka2, kb2 = G.Kb2;
  G.cfg.scheme_invis = scheme;
  switch (scheme) {
    case OCFD_SCHEME_WENO5: G.Ka1 = -2; G.Kb1 = 2; break;
    case OCFD_SCHEME_WENO7: G.Ka1 = -3; G.Kb1 = 3; break;
    default: G.Ka1 = -3; G.Kb1 = 4;
  }
  flux1(v, hh, n, 4, bidx, scm, sb);
  G.cfg.scheme_invis = saved; G.Ka1 = ka1; G.Kb1 = kb1; G.Ka2 = ka2; G.Kb2 = kb2;
  return 0;
}
int orc_recon_flux2(int scheme, const double* v, double* hh, int n, const int* bidx, const unsigned char* scm,
                    const int* sb) {
  int saved = G.cfg.scheme_invis, ka2 = G.Ka2, kb2 = G.Kb2;
  G.cfg.scheme_invis = scheme;
  switch (scheme) {
    case OCFD_SCHEME_WENO5: G.Ka2 = -1; G.Kb2 = 3; break;
    case OCFD_SCHEME_WENO7: G.Ka2 = -2; G.Kb2 = 4; break;
    default: G.Ka2 = -3; G.Kb2 = 4;
  }
  flux2(v, hh, n, 4, bidx, scm, sb);
  G.cfg.scheme_invis = saved; G.Ka2 = ka2; G.Kb2 = kb2;
  return 0;
}
/* UD7L coefficient setter for the 1-D tests (set_scheme_para) */
int orc_set_ud7l_diss(double diss) {
  const double UD7[8] = {-3.0, 25.0, -101.0, 319.0, 214.0, -38.0, 4.0, 0.0};
  const double CD8[8] = {-3.0, 29.0, -139.0, 533.0, 533.0, -139.0, 29.0, -3.0};
  for (int a = 0; a < 8; a++) G.ud7l[a] = diss * UD7[a] / 420.0 + (1.0 - diss) * CD8[a] / 840.0;
  return 0;
}
int orc_split(int kind, int n, const double* Et1, const double* p1, const double* c1, const double* d1,
              const double* u1, const double* v1, const double* w1, const double* A1, const double* A2,
              const double* A3, double* fp, double* fm) {
  if (kind == OCFD_SPLIT_SW) split_SW(n, 4, Et1, p1, c1, d1, u1, v1, w1, A1, A2, A3, fp, fm);
  else split_LLF(n, 4, Et1, p1, c1, d1, u1, v1, w1, A1, A2, A3, fp, fm);
  return 0;
}
double orc_comput_T(const double* di, double E) { return comput_T(di, E); }
double orc_comput_E(const double* di, double T) { return comput_E(di, T); }
double orc_comput_Cp(double d, const double* di, double T) { return comput_Cp(d, di, T); }
int orc_finalize(void) { free_all(); return 0; }
