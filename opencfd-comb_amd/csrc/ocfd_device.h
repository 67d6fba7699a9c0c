// ocfd_device.h -- device-side state and arithmetic for the MI355X flow path.
//
// All arithmetic is FP64 and is compiled with -ffp-contract=off so that every
// multiply and add rounds exactly as the reference Fortran does (which is
// built for baseline x86-64, i.e. without FMA).  Expression shapes below keep
// the reference's evaluation order; the comments give the reference lines.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OCFD_LAP 4
#define OCFD_NS_MAX 9   // species of the multi-species flow path

// Reconstruction kinds.  A reconstruction reads an 8-point window w[0..7]
// holding v(s-3..s+4) for flux+ at interface s (between cells s and s+1).
// Flux- at the same interface uses the mirrored window w[q] = v(s+4-q):
// every flux- formula of the reference is the flux+ formula evaluated on the
// mirrored stencil (OCFD_Schemes.f90:27-50 vs :6-25, :128-196 vs :55-124,
// :243-281 vs :207-241, :316-326 vs :302-313), so one code path serves both
// sides and lanes of a wavefront working on + and - stay convergent.
enum ReconKind : int {
  K_UD7L = 1,          // hybrid mask value 1 (OCFD_Schemes.f90:337-339)
  K_WENO7 = 2,         // hybrid mask value 2
  K_WENO5 = 3,         // hybrid mask value 3
  K_OMP6 = 4,
  K_C3A0 = 5,          // (2v(1)+5v(2)-v(3))/6     OCFD_inviscous.f90:250
  K_C3A1 = 6,          // (2v(0)+5v(1)-v(2))/6     OCFD_inviscous.f90:250 at i=1
  K_C3B = 7,           // (2v(-2)-7v(-1)+11v(0))/6 OCFD_inviscous.f90:268
  K_W5_DROP_S2 = 8,    // hh_weno5P_boundary(DEL_LIFT)   OCFD_Schemes_1.f90:279-308
  K_W5_DROP_S0 = 9     // hh_weno5P_boundary(DEL_RIGHT)
};

struct DevState {
  int nx, ny, nz, nxh, nyh, nzh, ns;
  long nh, ni;
  long sh[3];          // halo strides
  long si[3];          // interior strides
  int n[3];
  int per[3];
  int bidx[3][2];
  int sb[6];
  int scheme, split, cd8, viscous, mass_force;
  double h_1[3];       // 1/hx, 1/hy, 1/hz
  double dh[3];        // hx, hy, hz
  double cd6c[3][3];   // a1,a2,a3 of OCFD_dx0 per direction
  double cd8c[3][4];   // c1..c4
  double ud7l[8];
  double hyb_t1, hyb_t2;
  double hyb_eta;      // relative bound on |1/mean_fast - 1/mean_ordered| of the type-2 sensor
  double hyb_inv_t1, hyb_inv_t2;
  double Ri0, Mi0;     // N_SPEC == 1 gas constant and molar mass
  double amu0, pr, sc;
  double mf[3];
  // BC (boundary layer)
  int bl_inlet, bl_upper, bl_outlet;  // BcData_inlet (0/1/2), BcData_upper (-1/0/1/2), bc_outlet
  double bl_tw, bl_xinit, bl_disA;
  const double* bl_ht;   // device scalar: ht(tt) of the current time step
  const double* bl_in;   // flow_inlet0d/1d/2d: [6], [6][ny], [6][nz][ny] (variable slowest)
  const double* bl_up;   // flow_upper0d/1d/2d: [6], [6][nx], [6][nz][nx]
  const int* bl_nonref;
  const double* bl_pertb;
  // flow arrays
  double *f, *fn, *du, *di, *din;
  double *d, *u, *v, *w, *T, *p, *Et, *Et_prev;
  double *Amu, *Amk;
  double *Axx, *Ayy, *Azz;
  double *Akx, *Aky, *Akz, *Aix, *Aiy, *Aiz, *Asx, *Asy, *Asz, *Ajac;
  double *Akx1, *Aky1, *Akz1, *Aix1, *Aiy1, *Aiz1, *Asx1, *Asy1, *Asz1;
  double* Rh;
  uint8_t* mask[3];
  double* Ev;          // 12 interior arrays: Ev1(1:4), Ev2(1:4), Ev3(1:4)
  // multi-species path (N_SPEC > 1)
  double* AmD;         // [ns][interior] diffusion coefficients
  double* dus;         // [ns][interior] species right-hand sides
  double* Evs;         // [3*ns][interior] species diffusive fluxes (xi, eta, zeta)
  double* Qp;          // [3][interior] enthalpy flux of species diffusion
  double* cs;          // [ns][interior] species mass fractions di/d
  const double* jet_in;  // jet inflow flow_inlet2d (ny, nz, 5+ns)
  int jet_iflag;       // BC_Para(1) of the jet
  double* red;         // reduction scratch
  int* negt;           // [0] = count, [1..16] = flagged cell linear indices
  // Metric tables (mt_on = 1 when they reproduce every interior metric bit
  // for bit, checked at set-up): the nine coordinate derivatives of
  // comput_Jacobian3d (OCFD_mesh.f90:74-82) for a mesh whose x depends on i
  // only, y on j only and z on k only (Iflag_Gridtype = GRID1D).  Diagonal
  // ones are 1-D; an off-diagonal one is a one-sided difference of a constant,
  // zero away from non-periodic faces, and depends on two indices.
  //   xi[nx], yj[ny], zk[nz]; xj, yi [ny][nx]; xk, zi [nz][nx]; yk, zj [nz][ny]
  int mt_on;
  const double *mt_xi, *mt_yj, *mt_zk, *mt_xj, *mt_yi, *mt_xk, *mt_zi, *mt_yk, *mt_zj;
};

__device__ __forceinline__ long hq(const DevState& S, int i, int j, int k) {
  return (long)(i + OCFD_LAP - 1) + S.nxh * ((long)(j + OCFD_LAP - 1) + (long)S.nyh * (long)(k + OCFD_LAP - 1));
}
__device__ __forceinline__ long iq(const DevState& S, int i, int j, int k) {
  return (long)(i - 1) + (long)S.nx * ((long)(j - 1) + (long)S.ny * (long)(k - 1));
}
// Periodic face wrap: the single-rank image of exchange_boundary_x/y/z
// (OCFD_mpi.f90:196-236).  Non-periodic faces read the stored ghost.
__device__ __forceinline__ int wrap(int s, int n, int per) {
  return per ? (s < 1 ? s + n : (s > n ? s - n : s)) : s;
}

__device__ __forceinline__ double sqr(double x) { return x * x; }

// ---------------------------------------------------------------------------
// Metrics from the nine coordinate derivatives, comput_Jacobian3d
// (OCFD_mesh.f90:84-100) and A*1 = A*/Ajac (:39-51), in the reference's
// expression shapes; k_metrics and the table evaluation share them, so both
// give the same bits.  Order of m[]: Akx, Aky, Akz, Aix, Aiy, Aiz, Asx, Asy,
// Asz, Ajac.
struct Metrics {
  double kx, ky, kz, ix, iy, iz, sx, sy, sz, jac;
};
__device__ __forceinline__ Metrics metrics_from_derivs(double xi1, double yi1, double zi1, double xj1, double yj1,
                                                       double zj1, double xk1, double yk1, double zk1) {
  Metrics m;
  const double Jac1 = 1.0 / (xi1 * yj1 * zk1 + yi1 * zj1 * xk1 + zi1 * xj1 * yk1 - zi1 * yj1 * xk1 - yi1 * xj1 * zk1 -
                             xi1 * zj1 * yk1);
  m.jac = Jac1;
  m.kx = Jac1 * (yj1 * zk1 - zj1 * yk1);
  m.ky = Jac1 * (zj1 * xk1 - xj1 * zk1);
  m.kz = Jac1 * (xj1 * yk1 - yj1 * xk1);
  m.ix = Jac1 * (yk1 * zi1 - zk1 * yi1);
  m.iy = Jac1 * (zk1 * xi1 - xk1 * zi1);
  m.iz = Jac1 * (xk1 * yi1 - yk1 * xi1);
  m.sx = Jac1 * (yi1 * zj1 - zi1 * yj1);
  m.sy = Jac1 * (zi1 * xj1 - xi1 * zj1);
  m.sz = Jac1 * (xi1 * yj1 - yi1 * xj1);
  return m;
}
// the table derivatives at interior cell (i, j, k)
__device__ __forceinline__ Metrics metrics_table(const DevState& S, int i, int j, int k) {
  const long a = i - 1, b = j - 1, c = k - 1;
  const long ab = a + (long)S.nx * b, ac = a + (long)S.nx * c, bc = b + (long)S.ny * c;
  return metrics_from_derivs(S.mt_xi[a], S.mt_yi[ab], S.mt_zi[ac], S.mt_xj[ab], S.mt_yj[b], S.mt_zj[bc], S.mt_xk[ac],
                             S.mt_yk[bc], S.mt_zk[c]);
}
// the interior cell's metrics: evaluated from the tables (MT, a compile-time
// choice so that a kernel carries one form only), else read from the halo
// arrays at q
template <bool MT>
__device__ __forceinline__ Metrics metrics_at(const DevState& S, long q, int i, int j, int k) {
  if constexpr (MT) return metrics_table(S, i, j, k);
  Metrics m;
  m.kx = S.Akx[q]; m.ky = S.Aky[q]; m.kz = S.Akz[q];
  m.ix = S.Aix[q]; m.iy = S.Aiy[q]; m.iz = S.Aiz[q];
  m.sx = S.Asx[q]; m.sy = S.Asy[q]; m.sz = S.Asz[q];
  m.jac = S.Ajac[q];
  return m;
}
template <bool MT>
__device__ __forceinline__ double jac_at(const DevState& S, long q, int i, int j, int k) {
  if constexpr (!MT) return S.Ajac[q];
  const long a = i - 1, b = j - 1, c = k - 1;
  const long ab = a + (long)S.nx * b, ac = a + (long)S.nx * c, bc = b + (long)S.ny * c;
  const double xi1 = S.mt_xi[a], yi1 = S.mt_yi[ab], zi1 = S.mt_zi[ac];
  const double xj1 = S.mt_xj[ab], yj1 = S.mt_yj[b], zj1 = S.mt_zj[bc];
  const double xk1 = S.mt_xk[ac], yk1 = S.mt_yk[bc], zk1 = S.mt_zk[c];
  return 1.0 / (xi1 * yj1 * zk1 + yi1 * zj1 * xk1 + zi1 * xj1 * yk1 - zi1 * yj1 * xk1 - yi1 * xj1 * zk1 -
                xi1 * zj1 * yk1);
}

// XCD-aware block order.  The hardware hands workgroup b to XCD b % 8; each
// XCD has its own L2.  Remapping so that XCD x works on one contiguous range
// of logical blocks keeps the stencil neighbours (k +- 3 planes) of a block in
// the L2 of the XCD that reads them.  Bijective on [0, nb).
__device__ __forceinline__ long xcd_block(long b, long nb) {
  const long q = nb >> 3, rem = nb & 7;
  const long x = b & 7, r = b >> 3;
  return x * q + (x < rem ? x : rem) + r;
}
// linear cell index of this thread for a 1-D launch of 256-thread blocks
__device__ __forceinline__ long cell_index() {
  return xcd_block((long)blockIdx.x, (long)gridDim.x) * blockDim.x + threadIdx.x;
}

// ---------------------------------------------------------------------------
// reconstructions on a window w[0..7] = v(-3..4)
__device__ __forceinline__ double recon_weno5(const double* w, int drop) {
  // OCFD_Schemes.f90:6-25 and hh_weno5P_boundary (OCFD_Schemes_1.f90:279-308)
  const double ep = 1.0e-6, C03 = 3.0 / 10.0, C13 = 3.0 / 5.0, C23 = 1.0 / 10.0;
  const double vm2 = w[1], vm1 = w[2], v0 = w[3], v1 = w[4], v2 = w[5];
  double a0 = 0.0, q03 = 0.0, a2 = 0.0, q23 = 0.0;
  if (drop != 0) {
    double S0 = 13.0 / 12.0 * sqr(v0 - 2.0 * v1 + v2) + 1.0 / 4.0 * sqr(3.0 * v0 - 4.0 * v1 + v2);
    a0 = C03 / sqr(ep + S0);
    q03 = 1.0 / 3.0 * v0 + 5.0 / 6.0 * v1 - 1.0 / 6.0 * v2;
  }
  double S1 = 13.0 / 12.0 * sqr(vm1 - 2.0 * v0 + v1) + 1.0 / 4.0 * sqr(vm1 - v1);
  double a1 = C13 / sqr(ep + S1);
  double q13 = -(1.0 / 6.0 * vm1) + 5.0 / 6.0 * v0 + 1.0 / 3.0 * v1;
  if (drop != 2) {
    double S2 = 13.0 / 12.0 * sqr(vm2 - 2.0 * vm1 + v0) + 1.0 / 4.0 * sqr(vm2 - 4.0 * vm1 + 3.0 * v0);
    a2 = C23 / sqr(ep + S2);
    q23 = 1.0 / 3.0 * vm2 - 7.0 / 6.0 * vm1 + 11.0 / 6.0 * v0;
  }
  double am = a0 + a1 + a2;
  return (a0 * q03 + a1 * q13 + a2 * q23) / am;
}

__device__ __forceinline__ double recon_weno7(const double* w) {
  // OCFD_Schemes.f90:55-124 (WENO7-JS, Balsara-Shu smoothness indicators)
  const double C0 = 1.0 / 35.0, C1 = 12.0 / 35.0, C2 = 18.0 / 35.0, C3 = 4.0 / 35.0;
  const double a11 = -2.0 / 6.0, a12 = 9.0 / 6.0, a13 = -18.0 / 6.0, a14 = 11.0 / 6.0;
  const double a21 = 1.0 / 6.0, a23 = 3.0 / 6.0, a24 = 2.0 / 6.0;
  const double a31 = -2.0 / 6.0, a32 = -3.0 / 6.0, a34 = -1.0 / 6.0;
  const double a41 = -11.0 / 6.0, a42 = 18.0 / 6.0, a43 = -9.0 / 6.0, a44 = 2.0 / 6.0;
  const double b12 = 4.0, b13 = -5.0, b14 = 2.0, b22 = -2.0, b41 = 2.0, b42 = -5.0, b43 = 4.0;
  const double c12 = 3.0, d12 = 13.0 / 12.0, d13 = 1043.0 / 960.0, d14 = 1.0 / 12.0;
  const double e11 = -3.0 / 12.0, e12 = 13.0 / 12.0, e13 = -23.0 / 12.0, e14 = 25.0 / 12.0;
  const double e21 = 1.0 / 12.0, e22 = -5.0 / 12.0, e23 = 13.0 / 12.0, e24 = 3.0 / 12.0;
  const double e31 = -1.0 / 12.0, e32 = 7.0 / 12.0, e33 = 7.0 / 12.0, e34 = -1.0 / 12.0;
  const double e41 = 3.0 / 12.0, e42 = 13.0 / 12.0, e43 = -5.0 / 12.0, e44 = 1.0 / 12.0;
  const double ep = 1.0e-8;
  const double m3 = w[0], m2 = w[1], m1 = w[2], z0 = w[3], p1 = w[4], p2 = w[5], p3 = w[6];
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
  double a0 = C0 / sqr(ep + S0), a1 = C1 / sqr(ep + S1), a2 = C2 / sqr(ep + S2), a3 = C3 / sqr(ep + S3);
  double am = a0 + a1 + a2 + a3;
  double q0 = e11 * m3 + e12 * m2 + e13 * m1 + e14 * z0;
  double q1 = e21 * m2 + e22 * m1 + e23 * z0 + e24 * p1;
  double q2 = e31 * m1 + e32 * z0 + e33 * p1 + e34 * p2;
  double q3 = e41 * z0 + e42 * p1 + e43 * p2 + e44 * p3;
  return (a0 * q0 + a1 * q1 + a2 * q2 + a3 * q3) / am;
}

// Fortran SIGN(1.d0,x): +1 or -1 following the sign bit of x.
__device__ __forceinline__ double sgn1(double x) { return signbit(x) ? -1.0 : 1.0; }
__device__ __forceinline__ double mm2(double x1, double x2) {  // minmod2, OCFD_Schemes.f90:284-290
  return 0.5 * (sgn1(x1) + sgn1(x2)) * fmin(fabs(x1), fabs(x2));
}
__device__ __forceinline__ double mm4(double x1, double x2, double x3, double x4) {  // minmod4, :292-299
  double r = 0.5 * (sgn1(x1) + sgn1(x2));
  r = r * fabs(0.5 * (sgn1(x1) + sgn1(x3)));
  r = r * fabs(0.5 * (sgn1(x1) + sgn1(x4)));
  return r * fmin(fmin(fmin(fabs(x1), fabs(x2)), fabs(x3)), fabs(x4));
}

__device__ __forceinline__ double recon_omp6(const double* w) {
  // OCFD_OMP6P, OCFD_Schemes.f90:207-241; ep=1.e-10 is a REAL(4) literal.
  const double kappa = 4.0, ep = (double)1.0e-10f;
  const double OMP_m = 0.015, OMP_n = 0.0;
  const double OMP_a = 0.5 * (OMP_m + OMP_n), OMP_b = 0.5 * (OMP_m - OMP_n);
  const double a1 = OMP_a, a2 = 1.0 / 60.0 - OMP_b - 6.0 * OMP_a, a3 = -2.0 / 15.0 + 6.0 * OMP_b + 15.0 * OMP_a;
  const double a4 = 37.0 / 60.0 - 15.0 * OMP_b - 20.0 * OMP_a, a5 = 37.0 / 60.0 + 20.0 * OMP_b + 15.0 * OMP_a;
  const double a6 = -2.0 / 15.0 - 6.0 * OMP_a - 15.0 * OMP_b, a7 = 1.0 / 60.0 + OMP_a + 6.0 * OMP_b, a8 = -OMP_b;
  const double vm3 = w[0], vm2 = w[1], vm1 = w[2], v0 = w[3], v1 = w[4], v2 = w[5], v3 = w[6], v4 = w[7];
  double mid = a1 * v4 + a2 * v3 + a3 * v2 + a4 * v1 + a5 * v0 + a6 * vm1 + a7 * vm2 + a8 * vm3;
  double mp = v0 + mm2(v1 - v0, kappa * (v0 - vm1));
  if ((mid - v0) * (mid - mp) >= ep) {
    double d1 = vm2 + v0 - 2.0 * vm1;
    double d2 = vm1 + v1 - 2.0 * v0;
    double d3 = v0 + v2 - 2.0 * v1;
    double ul = v0 + kappa * (v0 - vm1);
    double md = 0.5 * (v0 + v1) - 0.5 * mm4(4.0 * d2 - d3, 4.0 * d3 - d2, d2, d3);
    double lc = v0 + 0.5 * (v0 - vm1) + kappa * mm4(4.0 * d1 - d2, 4.0 * d2 - d1, d2, d1) / 3.0;
    double fmn = fmax(fmin(fmin(v0, v1), md), fmin(fmin(v0, ul), lc));
    double fmx = fmin(fmax(fmax(v0, v1), md), fmax(fmax(v0, ul), lc));
    mid = mid + mm2(fmx - mid, fmn - mid);
  }
  return mid;
}

__device__ __forceinline__ double recon(int kind, const double* w, const double* ud7l) {
  switch (kind) {
    case K_UD7L:  // OCFD_UD7L_P, OCFD_Schemes.f90:302-313
      return ud7l[0] * w[0] + ud7l[1] * w[1] + ud7l[2] * w[2] + ud7l[3] * w[3] + ud7l[4] * w[4] +
             ud7l[5] * w[5] + ud7l[6] * w[6] + ud7l[7] * w[7];
    case K_WENO7: return recon_weno7(w);
    case K_WENO5: return recon_weno5(w, 1);
    case K_OMP6: return recon_omp6(w);
    case K_C3A0: return (2.0 * w[4] + 5.0 * w[5] - w[6]) / 6.0;
    case K_C3A1: return (2.0 * w[3] + 5.0 * w[4] - w[5]) / 6.0;
    case K_C3B: return (2.0 * w[1] - 7.0 * w[2] + 11.0 * w[3]) / 6.0;
    case K_W5_DROP_S2: return recon_weno5(w, 2);
    case K_W5_DROP_S0: return recon_weno5(w, 0);
  }
  return 0.0;
}

// Reconstruction kind at interface s (0..n) for side (0 = flux+, 1 = flux-).
// Encodes OCFD2d_flux1/OCFD2d_flux2 (OCFD_inviscous.f90:209-362): inner range
// [ib,ie] from the scheme's stencil [Ka,Kb] (OCFD_readpara.f90:378-405) and
// the WENO5-type boundary closures otherwise.
__device__ __forceinline__ int recon_kind_p(int scheme, int n, int blo, int bhi, int sblo, int sbhi, int s, int side,
                                            int mask) {
  int Ka, Kb, inner;
  switch (scheme) {
    case 1: Ka = side ? -1 : -2; Kb = side ? 3 : 2; inner = K_WENO5; break;
    case 2: Ka = side ? -2 : -3; Kb = side ? 4 : 3; inner = K_WENO7; break;
    case 3: Ka = -3; Kb = 4; inner = K_OMP6; break;
    case 4: Ka = -3; Kb = 4; inner = K_UD7L; break;
    default: Ka = -3; Kb = 4; inner = mask; break;  // hybrid: 1 UD7L, 2 WENO7, 3 WENO5
  }
  int kind = inner;
  if (blo == 1 && s <= -Ka) {
    if (sblo == 0) {
      if (side == 0) kind = s == 0 ? K_C3A0 : s == 1 ? K_C3A1 : s == 2 ? K_W5_DROP_S2 : K_WENO5;
      else kind = s == 0 ? K_C3B : s == 1 ? K_W5_DROP_S0 : K_WENO5;
    } else if (sblo == 1) {
      kind = K_WENO5;
    }
  }
  if (bhi == 1 && s >= n - Kb + 1) {
    if (sbhi == 0) {
      if (side == 0) kind = s == n ? K_C3B : s == n - 1 ? K_W5_DROP_S0 : K_WENO5;
      else kind = s == n ? K_C3A0 : s == n - 1 ? K_C3A1 : s == n - 2 ? K_W5_DROP_S2 : K_WENO5;
    } else if (sbhi == 1) {
      kind = K_WENO5;
    }
  }
  return kind;
}
__device__ __forceinline__ int recon_kind(const DevState& S, int dir, int s, int side, int mask) {
  return recon_kind_p(S.scheme, S.n[dir], S.bidx[dir][0], S.bidx[dir][1], S.sb[2 * dir], S.sb[2 * dir + 1], s, side,
                      mask);
}

// ---------------------------------------------------------------------------
// Central first derivative along dir at cell s (1..n) of a field f read at
// logical positions s+o (o = -4..4) through `at(o)`:
//   OCFD_dx0/dy0/dz0, OCFD_dxyz_viscous.f90:3-242 (CD6/CD8 + one-sided closures).
// CD = 0 / 1 fixes CD6 / CD8 at compile time (fewer live registers in the
// kernels that take many derivatives); CD = -1 reads the scheme from S.
// INT = true: the caller guarantees R+1 <= s <= n-R (R = 3 for CD6, 4 for
// CD8), where no closure applies and no periodic wrap is needed; the result
// is then the central formula, in straight-line code the scheduler can
// overlap with the loads of the next derivative.
template <int CD, class Fn>
__device__ __forceinline__ double deriv1_central(const DevState& S, int dir, Fn at) {
  if (CD == 0) {
    const double* a = S.cd6c[dir];
    return a[0] * (at(3) - at(-3)) + a[1] * (at(2) - at(-2)) + a[2] * (at(1) - at(-1));
  }
  const double* c = S.cd8c[dir];
  return c[0] * (at(1) - at(-1)) + c[1] * (at(2) - at(-2)) + c[2] * (at(3) - at(-3)) + c[3] * (at(4) - at(-4));
}

// cells whose CD6 (R = 3) or CD8 (R = 4) stencil along a direction of n
// cells is the central one without a periodic wrap
__device__ __forceinline__ bool central_cell(int s, int n, int R) { return s >= R + 1 && s <= n - R; }

// Wave-uniform choice of the interior form: every active lane of the wave has
// the central stencil in the named directions.  Interior waves then run
// straight-line code (no closure branches), so the loads of all derivatives
// can be in flight together; boundary waves take the general form.
__device__ __forceinline__ bool wave_all(bool p) { return __all((int)p) != 0; }

template <int CD = -1, bool INT = false, class Fn>
__device__ __forceinline__ double deriv1(const DevState& S, int dir, int s, Fn at) {
  static_assert(!INT || CD >= 0, "the interior form needs the scheme fixed at compile time");
  if constexpr (INT) return deriv1_central<CD>(S, dir, at);
  const int n = S.n[dir];
  const double* a = S.cd6c[dir];
  const bool cd8 = CD < 0 ? (bool)S.cd8 : CD == 1;
  const bool lo = !S.per[dir] && (s <= 3 || (s == 4 && cd8));
  const bool hi = !S.per[dir] && (s >= n - 2 || (s == n - 3 && cd8));
  double r = 0.0;
  if (!lo && !hi) {
    if (!cd8) {
      r = a[0] * (at(3) - at(-3)) + a[1] * (at(2) - at(-2)) + a[2] * (at(1) - at(-1));
    } else {
      const double* c = S.cd8c[dir];
      r = c[0] * (at(1) - at(-1)) + c[1] * (at(2) - at(-2)) + c[2] * (at(3) - at(-3)) + c[3] * (at(4) - at(-4));
    }
    return r;
  }
  // one-sided closures; the i+ ones are applied after (and override) the i- ones
  const double h = S.dh[dir];
  if (lo) {
    if (s == 1) r = (-3.0 * at(0) + 4.0 * at(1) - at(2)) / (2.0 * h);
    else if (s == 2) r = (at(1) - at(-1)) / (2.0 * h);
    else if (s == 3) r = (8.0 * (at(1) - at(-1)) - (at(2) - at(-2))) / (12.0 * h);
    else r = a[0] * (at(3) - at(-3)) + a[1] * (at(2) - at(-2)) + a[2] * (at(1) - at(-1));
  }
  if (hi) {
    if (s == n) r = (at(-2) - 4.0 * at(-1) + 3.0 * at(0)) / (2.0 * h);
    else if (s == n - 1) r = (at(1) - at(-1)) / (2.0 * h);
    else if (s == n - 2) r = (8.0 * (at(1) - at(-1)) - (at(2) - at(-2))) / (12.0 * h);
    else r = a[0] * (at(3) - at(-3)) + a[1] * (at(2) - at(-2)) + a[2] * (at(1) - at(-1));
  }
  return r;
}
