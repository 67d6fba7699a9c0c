// ocfd_flow.hip -- MI355X (gfx950) kernels and C-ABI for OpenCFD-Comb's
// per-step flow path: flux-split upwind inviscid derivatives, central viscous
// terms, the TVD-RK3 update, primitive recovery and the boundary conditions.
//
// Data stay resident in HBM in the reference's Fortran layout (i fastest,
// halo arrays of width LAP=4).  The MPI halo exchange of the reference is not
// performed: periodic faces are read through wrapped indices and non-periodic
// ghosts keep the values the host uploaded.  FP64 throughout, no FMA
// contraction (-ffp-contract=off), same operation order as the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/ocfd_hip.h"
#include "ocfd_device.h"

// ocfd_chem_adaptive_batch / ocfd_chem_surrogate_batch with the longest-first
// cell order switched by the caller (ocfd_chem.hip)
int chem_adaptive_batch(long n, double* rho, double* E, const double* dt, double rtol_y, double y_floor, double rtol_t,
                        int max_sub, int* nsub_out, int scheduler, bool order);
int chem_surrogate_batch(long n, double* rho, double* E, const double* dt, bool fallback, double rtol_y, double y_floor,
                         double rtol_t, int max_sub, int* flag_out, bool order);

// ===========================================================================
// Inviscid sweep (du_inviscous, OCFD_inviscous.f90:47-202).
//
// One workgroup owns NP pencils x TS cells along DIR.  Phase 1 evaluates the
// flux splitting once per point of the tile (TS+8 points per pencil) into LDS;
// phase 2 reconstructs the 10 interface values (5 components x flux+/-) at the
// TS+1 interfaces; phase 3 differences them into du.  For DIR 0 the pencils
// are i-rows (coalesced along the sweep); for DIR 1/2 the NP pencils are
// adjacent in i, so loads coalesce across pencils.
// ===========================================================================
// Tile shapes offered to the autotuner: NP pencils per workgroup x TS cells
// per pencil, on NT threads.  NT is 4 or 8 wavefronts: a workgroup of 5
// wavefronts puts two of them on one SIMD, which then holds back every
// barrier (measured 1.5x slower on the flat plate).  NP (TS+8) <= NT makes
// the flux splitting of the tile's points one pass; for DIR 1/2, NP >= 8
// keeps the loads at 64 B or more per row.  Variant 0 is the default, the
// shape that measured fastest inside steps of the flat plate; the autotuner
// leaves it only for a shape more than 3% faster.  Every variant computes
// every interface with the same arithmetic, so results do not depend on the
// choice.
template <int DIR, int V> struct SweepTile;
template <> struct SweepTile<0, 0> { static constexpr int NP = 4, TS = 56, NT = 256; };
template <> struct SweepTile<0, 1> { static constexpr int NP = 4, TS = 64, NT = 256; };
template <> struct SweepTile<0, 2> { static constexpr int NP = 2, TS = 120, NT = 256; };
template <> struct SweepTile<0, 3> { static constexpr int NP = 8, TS = 56, NT = 512; };
template <> struct SweepTile<1, 0> { static constexpr int NP = 8, TS = 24, NT = 256; };
template <> struct SweepTile<1, 1> { static constexpr int NP = 8, TS = 32, NT = 256; };
template <> struct SweepTile<1, 2> { static constexpr int NP = 16, TS = 24, NT = 512; };
template <> struct SweepTile<1, 3> { static constexpr int NP = 8, TS = 56, NT = 512; };
template <> struct SweepTile<2, 0> { static constexpr int NP = 8, TS = 24, NT = 256; };
template <> struct SweepTile<2, 1> { static constexpr int NP = 8, TS = 32, NT = 256; };
template <> struct SweepTile<2, 2> { static constexpr int NP = 16, TS = 24, NT = 512; };
template <> struct SweepTile<2, 3> { static constexpr int NP = 8, TS = 56, NT = 512; };
#define OCFD_SWEEP_VARIANTS 4

// Interfaces per thread of the linear pass: the smallest run length L for
// which the 5 NP lines of one side (5 components x NP pencils) of NIF
// interfaces, cut into runs of L, fit in the NT / 2 threads of that side.
constexpr int sweep_run(int NP, int NIF, int NT) {
  int L = 1;
  while (5 * NP * ((NIF + L - 1) / L) > NT / 2) L++;
  return L;
}

// One run of the linear pass: UD7L at interfaces l0 .. l1-1 (at most L) of
// one line, the 8-point window sliding by one point per interface (flux+:
// w[q] = v(l-3+q); flux-: w[q] = v(l+4-q)); F / H point at the line's
// pencil in sF / sH, PS / IS are the LDS strides between neighbouring
// points / interfaces.  Unrolled, so the window slides by renaming.
template <int SIDE, int L, int PS, int IS>
__device__ __forceinline__ void ud7l_run(const double* F, double* H, int l0, int l1, const double* cf) {
  const double c0 = cf[0], c1 = cf[1], c2 = cf[2], c3 = cf[3], c4 = cf[4], c5 = cf[5], c6 = cf[6], c7 = cf[7];
  double w[8];
#pragma unroll
  for (int q = 0; q < 7; q++) {
    if (SIDE == 0) w[q + 1] = F[(l0 + q) * PS];
    else w[q] = F[(l0 + 6 - q) * PS];
  }
#pragma unroll
  for (int u = 0; u < L; u++) {
    const int l = l0 + u;
    if (l >= l1) break;
    if (SIDE == 0) {
#pragma unroll
      for (int q = 0; q < 7; q++) w[q] = w[q + 1];
      w[7] = F[(l + 7) * PS];
    } else {
#pragma unroll
      for (int q = 7; q > 0; q--) w[q] = w[q - 1];
      w[0] = F[(l + 7) * PS];
    }
    H[l * IS] = c0 * w[0] + c1 * w[1] + c2 * w[2] + c3 * w[3] + c4 * w[4] + c5 * w[5] + c6 * w[6] + c7 * w[7];
  }
}

template <int DIR>
__device__ __forceinline__ void pencil_ijk(const DevState& S, long pen, int& a0, int& a1) {
  // DIR 0: pen -> (j,k); DIR 1: pen -> (i,k); DIR 2: pen -> (i,j)
  const int m = DIR == 0 ? S.ny : S.nx;
  a0 = (int)(pen % m) + 1;
  a1 = (int)(pen / m) + 1;
}

template <int DIR>
__device__ __forceinline__ long sweep_hq(const DevState& S, int a0, int a1, int s) {
  const int t = wrap(s, S.n[DIR], S.per[DIR]);
  if (DIR == 0) return hq(S, t, a0, a1);
  if (DIR == 1) return hq(S, a0, t, a1);
  return hq(S, a0, a1, t);
}

template <int DIR>
__device__ __forceinline__ long sweep_iq(const DevState& S, int a0, int a1, int s) {
  if (DIR == 0) return iq(S, s, a0, a1);
  if (DIR == 1) return iq(S, a0, s, a1);
  return iq(S, a0, a1, s);
}

template <int DIR>
__device__ __forceinline__ long mask_index(const DevState& S, long pen, int s) {
  if (DIR == 0) return (long)s + (long)(S.nx + 1) * pen;
  if (DIR == 1) {
    const long i0 = pen % S.nx, k0 = pen / S.nx;
    return i0 + (long)S.nx * ((long)s + (long)(S.ny + 1) * k0);
  }
  return pen + (long)S.nx * S.ny * (long)s;
}

// Per-pencil base indices of a sweep tile, formed once per workgroup: the
// halo index, the interior index and the mask index of the pencil's point
// s = 1 (s = 0 for the mask); a point s of the pencil then adds (s-1) (or s)
// times the direction's stride, with no integer division in the hot loops.
template <int DIR, int NP>
struct PencilBase {
  long hq1[NP], iq1[NP], mb[NP];
};
template <int DIR, int NP>
__device__ __forceinline__ void pencil_bases(const DevState& S, long pen0, long npen, PencilBase<DIR, NP>& B) {
  const int t = threadIdx.x;
  if (t < NP && pen0 + t < npen) {
    int a0, a1;
    pencil_ijk<DIR>(S, pen0 + t, a0, a1);
    B.hq1[t] = sweep_hq<DIR>(S, a0, a1, 1);
    B.iq1[t] = sweep_iq<DIR>(S, a0, a1, 1);
    B.mb[t] = mask_index<DIR>(S, pen0 + t, 0);
  }
}
template <int DIR>
__device__ __forceinline__ long mask_stride(const DevState& S) {
  return DIR == 0 ? 1L : DIR == 1 ? (long)S.nx : (long)S.nx * S.ny;
}

// Flux splitting at one point (split_Stager_Warming, OCFD_flux_split.f90:6-72;
// split_Local_LaxFriedrichs, :76-103).  The sound speed is comput_cc's
// (Comput_flow_variables.f90:50-63), evaluated on the fly.
__device__ __forceinline__ void split_point(int kind, double Et1, double p1, double d1, double u1, double v1,
                                            double w1, double A1, double A2, double A3, double* fp, double* fm) {
  const double gamma = p1 / Et1 + 1.0;
  const double c1 = sqrt(gamma * p1 / d1);
  if (kind == OCFD_SPLIT_SW) {
    const double tmp1 = 2.0 * (gamma - 1.0);
    const double tmp3 = (3.0 - gamma) / (2.0 * (gamma - 1.0));
    const double ss = sqrt(A1 * A1 + A2 * A2 + A3 * A3);
    const double ss1 = 1.0 / ss;
    const double ak1 = A1 * ss1, ak2 = A2 * ss1, ak3 = A3 * ss1;
    const double vs = A1 * u1 + A2 * v1 + A3 * w1;
    const double E1 = vs, E2 = vs - c1 * ss, E3 = vs + c1 * ss;
    const double E1P = (E1 + fabs(E1)) * 0.5, E2P = (E2 + fabs(E2)) * 0.5, E3P = (E3 + fabs(E3)) * 0.5;
    const double E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
    const double tmp0 = d1 / (2.0 * gamma);
    const double uc1 = u1 - c1 * ak1, uc2 = u1 + c1 * ak1;
    const double vc1 = v1 - c1 * ak2, vc2 = v1 + c1 * ak2;
    const double wc1 = w1 - c1 * ak3, wc2 = w1 + c1 * ak3;
    const double vvc1 = (uc1 * uc1 + vc1 * vc1 + wc1 * wc1) * 0.5;
    const double vvc2 = (uc2 * uc2 + vc2 * vc2 + wc2 * wc2) * 0.5;
    const double vv = (gamma - 1.0) * (u1 * u1 + v1 * v1 + w1 * w1);
    const double W2 = tmp3 * c1 * c1;
    fp[0] = tmp0 * (tmp1 * E1P + E2P + E3P);
    fp[1] = tmp0 * (tmp1 * E1P * u1 + E2P * uc1 + E3P * uc2);
    fp[2] = tmp0 * (tmp1 * E1P * v1 + E2P * vc1 + E3P * vc2);
    fp[3] = tmp0 * (tmp1 * E1P * w1 + E2P * wc1 + E3P * wc2);
    fp[4] = tmp0 * (E1P * vv + E2P * vvc1 + E3P * vvc2 + W2 * (E2P + E3P));
    fm[0] = tmp0 * (tmp1 * E1M + E2M + E3M);
    fm[1] = tmp0 * (tmp1 * E1M * u1 + E2M * uc1 + E3M * uc2);
    fm[2] = tmp0 * (tmp1 * E1M * v1 + E2M * vc1 + E3M * vc2);
    fm[3] = tmp0 * (tmp1 * E1M * w1 + E2M * wc1 + E3M * wc2);
    fm[4] = tmp0 * (E1M * vv + E2M * vvc1 + E3M * vvc2 + W2 * (E2M + E3M));
  } else {
    const double ss = sqrt(A1 * A1 + A2 * A2 + A3 * A3);
    const double un = A1 * u1 + A2 * v1 + A3 * w1;
    const double lamda0 = fabs(un) + ss * c1;
    const double E0 = Et1 + 0.5 * d1 * (u1 * u1 + v1 * v1 + w1 * w1);
    fp[0] = 0.5 * (d1 * un + lamda0 * d1);
    fp[1] = 0.5 * (d1 * u1 * un + A1 * p1 + lamda0 * d1 * u1);
    fp[2] = 0.5 * (d1 * v1 * un + A2 * p1 + lamda0 * d1 * v1);
    fp[3] = 0.5 * (d1 * w1 * un + A3 * p1 + lamda0 * d1 * w1);
    fp[4] = 0.5 * ((E0 + p1) * un + lamda0 * E0);
    fm[0] = 0.5 * (d1 * un - lamda0 * d1);
    fm[1] = 0.5 * (d1 * u1 * un + A1 * p1 - lamda0 * d1 * u1);
    fm[2] = 0.5 * (d1 * v1 * un + A2 * p1 - lamda0 * d1 * v1);
    fm[3] = 0.5 * (d1 * w1 * un + A3 * p1 - lamda0 * d1 * w1);
    fm[4] = 0.5 * ((E0 + p1) * un - lamda0 * E0);
  }
}

// Kind lists: the sorted list of a tile depends only on the hybrid masks and
// the position, which change once per step, so the first sweep after the
// masks are built stores it (KC_WRITE: entries at kc + tile*NE, count at
// kct[tile]) and the later stages read it back (KC_READ) instead of
// classifying and sorting again.  A list entry is (kind << 12) | entry.
enum { KC_NONE = 0, KC_WRITE = 1, KC_READ = 2 };

template <int DIR, int V>
__global__ __launch_bounds__((SweepTile<DIR, V>::NT)) void k_sweep(DevState S, int kmode, unsigned short* kc,
                                                                   int* kct) {
  constexpr int NP = SweepTile<DIR, V>::NP, TS = SweepTile<DIR, V>::TS, NT = SweepTile<DIR, V>::NT;
  constexpr int NW = NT / 64;
  constexpr int NPT = TS + 8, NIF = TS + 1, NE = 2 * NP * NIF;
  static_assert(NT % 64 == 0, "whole wavefronts");
  static_assert(NE < 0x1000, "entry index must fit in 12 bits");
  // split fluxes (phase 1-2) and interface values (phase 2-3)
  __shared__ double sF[10 * NP * NPT];
  __shared__ double sH[10 * NP * NIF];
  __shared__ unsigned short sList[NE];
  __shared__ unsigned char sKind[NE];
  __shared__ int sCnt[NW][10];
  __shared__ int sTotal;
  __shared__ PencilBase<DIR, NP> PB;
  const int n = S.n[DIR];
  // XCD-aware tile order over the flattened 2-D grid
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

  // phase 1: split fluxes at the tile's points s = c0-4 .. c0+ncell+3
  for (int t = tid; t < NP * NPT; t += NT) {
    const int pl = DIR == 0 ? t / NPT : t % NP;
    const int sl = DIR == 0 ? t % NPT : t / NP;
    const long pen = pen0 + pl;
    const int s = c0 - 4 + sl;
    if (pen >= npen || sl > ncell + 7) continue;
    const long q = PB.hq1[pl] + (long)(wrap(s, n, S.per[DIR]) - 1) * S.sh[DIR];
    double fp[5], fm[5];
    split_point(S.split, S.Et[q], S.p[q], S.d[q], S.u[q], S.v[q], S.w[q], M1[q], M2[q], M3[q], fp, fm);
#pragma unroll
    for (int m = 0; m < 5; m++) {
      sF[m * NP * NPT + lp(pl, sl)] = fp[m];
      sF[(5 + m) * NP * NPT + lp(pl, sl)] = fm[m];
    }
  }

  // phase 2a: classify the (interface, side) entries and sort them by
  // reconstruction kind (stable: spatial order kept inside a kind), so the
  // lanes of a wavefront run one scheme; entry e = side*NP*NIF + lh(pl, l).
  const bool hyb = S.scheme == OCFD_SCHEME_HYBRID;
  const bool nearb = !S.per[DIR] && (c0 - 1 <= 4 || c0 + ncell - 1 >= n - 4);
  const bool uniform = !hyb && !nearb;
  // linear mode (hybrid or UD7L schemes): a first pass evaluates UD7L at
  // every interface with sliding windows, and the sorted list holds only the
  // entries of the other kinds, whose values overwrite it
  const bool lin = hyb || S.scheme == 4;
  const int kfirst = lin ? 2 : 1;
  const int wave = tid >> 6, lane = tid & 63;
  constexpr int PERW = (NE + NW - 1) / NW;
  const int e0 = wave * PERW, e1 = min(NE, e0 + PERW);
  if (!uniform && kmode == KC_READ) {
    const long base = lb * NE;
    const int cnt = kct[lb];
    for (int i = tid; i < cnt; i += NT) sList[i] = kc[base + i];
    if (tid == 0) sTotal = cnt;
  } else if (!uniform) {
    int cnt[10];
#pragma unroll
    for (int k = 0; k < 10; k++) cnt[k] = 0;
    for (int base = e0; base < e1; base += 64) {
      const int e = base + lane;
      int kind = 0;
      if (e < e1) {
        const int side = e / (NP * NIF);
        const int r = e % (NP * NIF);
        const int pl = DIR == 0 ? r / NIF : r % NP;
        const int l = DIR == 0 ? r % NIF : r / NP;
        const long pen = pen0 + pl;
        if (pen < npen && l <= ncell) {
          const int s_if = c0 - 1 + l;
          const int mask = hyb ? (int)S.mask[DIR][PB.mb[pl] + (long)s_if * mask_stride<DIR>(S)] : 0;
          // away from non-periodic faces the kind is the mask (hybrid: 1..3)
          kind = nearb ? recon_kind(S, DIR, s_if, side, mask) : mask;
        }
        sKind[e] = (unsigned char)kind;
      }
      // only kinds 2 and 3 occur in a hybrid tile away from the faces
#pragma unroll
      for (int k = 1; k < 10; k++)
        cnt[k] += (k < kfirst || (!nearb && k > 3)) ? 0 : __popcll(__ballot(kind == k));
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < 10; k++) sCnt[wave][k] = cnt[k];
    }
    __syncthreads();
    int off[10], ktot[10];
    int acc = 0;
#pragma unroll
    for (int k = 1; k < 10; k++) {
      int before = 0;
      ktot[k] = 0;
      if (k < kfirst) continue;
      for (int w = 0; w < NW; w++) {
        if (w < wave) before += sCnt[w][k];
        ktot[k] += sCnt[w][k];
      }
      off[k] = acc + before;
      acc += ktot[k];
    }
    if (tid == 0) sTotal = acc;
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (int base = e0; base < e1; base += 64) {
      const int e = base + lane;
      const int kind = e < e1 ? (int)sKind[e] : 0;
#pragma unroll
      for (int k = 1; k < 10; k++) {
        if (k < kfirst || ktot[k] == 0) continue;  // linear kind, or kind absent from the tile
        const unsigned long long bal = __ballot(kind == k);
        if (kind == k) sList[off[k] + __popcll(bal & lt)] = (unsigned short)((k << 12) | e);
        off[k] += __popcll(bal);
      }
    }
    if (kmode == KC_WRITE) {
      __syncthreads();
      const long base = lb * NE;
      for (int i = tid; i < acc; i += NT) kc[base + i] = sList[i];
      if (tid == 0) kct[lb] = acc;
    }
  }
  __syncthreads();
  const int total = uniform ? (lin ? 0 : NE) : sTotal;

  // phase 2a': UD7L (OCFD_UD7L_P, OCFD_Schemes.f90:302-313) at every
  // interface of the tile.  Thread = (line, run): a line is (side, component,
  // pencil), a run L consecutive interfaces (ud7l_run: one LDS read per
  // interface, the same products summed in the same order as recon()).  The
  // first NT / 2 threads take the flux+ lines and the others the flux-
  // lines, so the side is uniform over each wavefront.
  constexpr int LPS = DIR == 0 ? 1 : NP;  // LDS stride between neighbouring points of a pencil
  constexpr int LIS = DIR == 0 ? 1 : NP;  // LDS stride between neighbouring interfaces of a pencil
  if (lin) {
    constexpr int HT = NT / 2, L = sweep_run(NP, NIF, NT), G = (NIF + L - 1) / L;
    static_assert(HT % 64 == 0 && 5 * NP * G <= HT, "one side per wavefront");
    const int side = __builtin_amdgcn_readfirstlane(tid >= HT ? 1 : 0);
    const int rest = tid - side * HT;
    if (rest < 5 * NP * G) {
      const int g = rest % G, line = rest / G;
      const int pl = line % NP, sc = side * 5 + line / NP;
      const int l0 = g * L, l1 = min(l0 + L, ncell + 1);
      if (pen0 + pl < npen && l0 < l1) {
        const double* F = sF + sc * NP * NPT + lp(pl, 0);
        double* H = sH + sc * NP * NIF + lh(pl, 0);
        if (side == 0) ud7l_run<0, L, LPS, LIS>(F, H, l0, l1, S.ud7l);
        else ud7l_run<1, L, LPS, LIS>(F, H, l0, l1, S.ud7l);
      }
    }
    __syncthreads();
  }

  // phase 2b: reconstruct.  Item t is component t % 5 of sorted entry t / 5,
  // so the kind changes only where the sorted list does (a few wavefronts
  // per tile run two kinds).  The loop is kept rolled: the reconstruction
  // switch appears once in the code, which keeps the kernel inside the
  // instruction cache.
  const int inner = S.scheme == 1 ? K_WENO5 : S.scheme == 2 ? K_WENO7 : S.scheme == 3 ? K_OMP6 : K_UD7L;
#pragma unroll 1
  for (int t = tid; t < 5 * total; t += NT) {
    const int idx = t / 5, comp = t - 5 * idx;
    const int u = uniform ? idx : (int)sList[idx];
    const int e = u & 0xFFF;
    const int side = e >= NP * NIF ? 1 : 0;
    const int r = e - side * (NP * NIF);
    const int pl = DIR == 0 ? r / NIF : r % NP;
    const int l = DIR == 0 ? r % NIF : r / NP;
    if (uniform && (pen0 + pl >= npen || l > ncell)) continue;
    const int kind = uniform ? inner : u >> 12;
    // window w[qq] = v(l-3+qq) for flux+ and v(l+4-qq) for flux-
    const double* F = sF + (side * 5 + comp) * NP * NPT + lp(pl, side ? l + 7 : l);
    const int step = side ? -LPS : LPS;
    double w[8];
#pragma unroll
    for (int qq = 0; qq < 8; qq++) w[qq] = F[qq * step];
    sH[(side * 5 + comp) * NP * NIF + r] = recon(kind, w, S.ud7l);
  }
  __syncthreads();

  // phase 3: du from interface differences, one thread per cell for the
  // five components (the cell's index work done once)
  const double h1 = S.h_1[DIR];
  for (int r = tid; r < NP * TS; r += NT) {
    const int pl = DIR == 0 ? r / TS : r % NP;
    const int c = DIR == 0 ? r % TS : r / NP;
    const long pen = pen0 + pl;
    if (pen >= npen || c >= ncell) continue;
    const long ci = PB.iq1[pl] + (long)(c0 + c - 1) * S.si[DIR];
    const double* Hc = sH + lh(pl, c);
    double val[5], old[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const double hp1 = Hc[m * NP * NIF + LIS], hp0 = Hc[m * NP * NIF];
      const double hm1 = Hc[(5 + m) * NP * NIF + LIS], hm0 = Hc[(5 + m) * NP * NIF];
      val[m] = (hp1 - hp0 + hm1 - hm0) * h1;
      if (DIR != 0) old[m] = S.du[ci + (long)m * S.ni];
    }
    if (DIR == 0) {
#pragma unroll
      for (int m = 0; m < 5; m++) S.du[ci + (long)m * S.ni] = val[m];
    } else if (DIR == 1) {
#pragma unroll
      for (int m = 0; m < 5; m++) S.du[ci + (long)m * S.ni] = old[m] + val[m];
    } else {
      const double aj = S.Ajac[PB.hq1[pl] + (long)(c0 + c - 1) * S.sh[DIR]];
#pragma unroll
      for (int m = 0; m < 5; m++) S.du[ci + (long)m * S.ni] = -(old[m] + val[m]) * aj;
    }
  }
}

// OCFD2d_flux1 / OCFD2d_flux2 over a batch of pencils (the 1-D scheme entry
// points of OCFD_inviscous.f90:209-362): one thread per (pencil, interface),
// the sweep's window, kind selection and reconstructions.  v holds
// v(1-LAP:n+LAP) of each pencil, hh gets hh(0:n), scm the hybrid mask of
// each interface.
__global__ __launch_bounds__(256) void k_flux1d(DevState S, int side, int scheme, long npen, int n,
                                                const double* __restrict__ v, double* __restrict__ hh,
                                                const uint8_t* __restrict__ scm, int blo, int bhi, int sblo, int sbhi) {
  const long npts = (long)n + 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npen * npts) return;
  const long p = t / npts;
  const int l = (int)(t - p * npts);
  // a hybrid mask outside 1..3 names no scheme: such interfaces get NaN
  int mask = scm ? (int)scm[t] : 0;
  if (mask < K_UD7L || mask > K_WENO5) mask = -1;
  const int kind = recon_kind_p(scheme, n, blo, bhi, sblo, sbhi, l, side, mask);
  if (kind < 0) {
    hh[t] = __longlong_as_double(0x7ff8000000000000LL);
    return;
  }
  const double* vp = v + p * (n + 2 * OCFD_LAP);
  double w[8];
#pragma unroll
  for (int qq = 0; qq < 8; qq++) w[qq] = vp[side ? l + 7 - qq : l + qq];
  hh[t] = recon(kind, w, S.ud7l);
}

// ===========================================================================
// Characteristic flux (IF_Scheme_Character = 1): du_inviscous_Character and
// flux_charteric, OCFD_inviscous_character.f90:3-314.  At each interface the
// split fluxes of the stencil are projected on the left eigenvectors S of the
// arithmetic-mean state, reconstructed component by component (the same
// stencils and closures as the non-characteristic sweep), summed and
// projected back with S^-1.  Interfaces of the hybrid's linear scheme are not
// projected; non-periodic faces with Scheme_boundary = 0 take the
// non-reflecting one-sided formula on the physical fluxes (:288-312).
// ===========================================================================
__device__ __forceinline__ void char_matrices(double u1, double v1, double w1, double c1, double p1, double Et1,
                                              double a1, double a2, double a3, double S[5][5], double S1[5][5]) {
  double ss = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
  const double n1 = a1 / ss, n2 = a2 / ss, n3 = a3 / ss;
  double l1, l2, l3;
  if (fabs(n3) <= fabs(n2)) {
    ss = sqrt(n1 * n1 + n2 * n2);
    l1 = -n2 / ss;
    l2 = n1 / ss;
    l3 = 0.0;
  } else {
    ss = sqrt(n1 * n1 + n3 * n3);
    l1 = -n3 / ss;
    l2 = 0.0;
    l3 = n1 / ss;
  }
  const double m1 = n2 * l3 - n3 * l2, m2 = n3 * l1 - n1 * l3, m3 = n1 * l2 - n2 * l1;
  const double un = u1 * n1 + v1 * n2 + w1 * n3, ul = u1 * l1 + v1 * l2 + w1 * l3, um = u1 * m1 + v1 * m2 + w1 * m3;
  const double V2 = (u1 * u1 + v1 * v1 + w1 * w1) * 0.5;
  const double gamma = p1 / Et1 + 1.0;
  const double KK = (gamma - 1.0) / (c1 * c1);
  const double KK1 = KK * 0.5, X1 = 1.0 / (2.0 * c1), H = V2 + 1.0 / KK;
  S[0][0] = 1.0 - KK * V2; S[0][1] = KK * u1; S[0][2] = KK * v1; S[0][3] = KK * w1; S[0][4] = -KK;
  S[1][0] = -ul; S[1][1] = l1; S[1][2] = l2; S[1][3] = l3; S[1][4] = 0.0;
  S[2][0] = -um; S[2][1] = m1; S[2][2] = m2; S[2][3] = m3; S[2][4] = 0.0;
  S[3][0] = KK1 * V2 + X1 * un; S[3][1] = -X1 * n1 - KK1 * u1; S[3][2] = -X1 * n2 - KK1 * v1;
  S[3][3] = -X1 * n3 - KK1 * w1; S[3][4] = KK1;
  S[4][0] = KK1 * V2 - X1 * un; S[4][1] = X1 * n1 - KK1 * u1; S[4][2] = X1 * n2 - KK1 * v1;
  S[4][3] = X1 * n3 - KK1 * w1; S[4][4] = KK1;
  S1[0][0] = 1.0; S1[0][1] = 0.0; S1[0][2] = 0.0; S1[0][3] = 1.0; S1[0][4] = 1.0;
  S1[1][0] = u1; S1[1][1] = l1; S1[1][2] = m1; S1[1][3] = u1 - c1 * n1; S1[1][4] = u1 + c1 * n1;
  S1[2][0] = v1; S1[2][1] = l2; S1[2][2] = m2; S1[2][3] = v1 - c1 * n2; S1[2][4] = v1 + c1 * n2;
  S1[3][0] = w1; S1[3][1] = l3; S1[3][2] = m3; S1[3][3] = w1 - c1 * n3; S1[3][4] = w1 + c1 * n3;
  S1[4][0] = V2; S1[4][1] = ul; S1[4][2] = um; S1[4][3] = H - c1 * un; S1[4][4] = H + c1 * un;
}

template <int DIR>
__global__ __launch_bounds__(256) void k_sweep_char(DevState S) {
  constexpr int NP = SweepTile<DIR, 0>::NP, TS = SweepTile<DIR, 0>::TS;
  constexpr int NPT = TS + 8, NIF = TS + 1;
  __shared__ double sF[10 * NP * NPT];
  __shared__ double sH[5 * NP * NIF];
  __shared__ PencilBase<DIR, NP> PB;
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
    if (pen0 + pl >= npen || sl > ncell + 7) continue;
    const long q = PB.hq1[pl] + (long)(wrap(c0 - 4 + sl, n, S.per[DIR]) - 1) * S.sh[DIR];
    double fp[5], fm[5];
    split_point(S.split, S.Et[q], S.p[q], S.d[q], S.u[q], S.v[q], S.w[q], M1[q], M2[q], M3[q], fp, fm);
#pragma unroll
    for (int m = 0; m < 5; m++) {
      sF[m * NP * NPT + lp(pl, sl)] = fp[m];
      sF[(5 + m) * NP * NPT + lp(pl, sl)] = fm[m];
    }
  }
  __syncthreads();
  const bool hyb = S.scheme == OCFD_SCHEME_HYBRID;
  const int blo = S.bidx[DIR][0], bhi = S.bidx[DIR][1];
  const int sblo = S.sb[2 * DIR], sbhi = S.sb[2 * DIR + 1];
  for (int t = tid; t < NP * NIF; t += blockDim.x) {
    const int pl = DIR == 0 ? t / NIF : t % NP;
    const int l = DIR == 0 ? t % NIF : t / NP;
    if (pen0 + pl >= npen || l > ncell) continue;
    const int s_if = c0 - 1 + l;
    const int mask = hyb ? (int)S.mask[DIR][PB.mb[pl] + (long)s_if * mask_stride<DIR>(S)] : 0;
    double hh[5];
    const bool face_lo = blo == 1 && sblo == 0 && s_if == 0, face_hi = bhi == 1 && sbhi == 0 && s_if == n;
    if (face_lo || face_hi) {
      // tile point of position s: s - (c0 - 4)
#pragma unroll
      for (int m = 0; m < 5; m++) {
        const double* P = sF + m * NP * NPT;
        const double* M = sF + (5 + m) * NP * NPT;
        auto at = [&](const double* F, int sp) { return F[lp(pl, sp - c0 + 4)]; };
        hh[m] = face_lo ? (2.0 * at(P, 1) + 5.0 * at(P, 2) - at(P, 3) + 2.0 * at(M, 3) - 7.0 * at(M, 2) +
                           11.0 * at(M, 1)) / 6.0
                        : (2.0 * at(P, n - 2) - 7.0 * at(P, n - 1) + 11.0 * at(P, n) + 2.0 * at(M, n) +
                           5.0 * at(M, n - 1) - at(M, n - 2)) / 6.0;
      }
    } else {
      const int kP = recon_kind(S, DIR, s_if, 0, mask), kM = recon_kind(S, DIR, s_if, 1, mask);
      if (hyb && mask == 1) {
#pragma unroll
        for (int m = 0; m < 5; m++) {
          double wP[8], wM[8];
#pragma unroll
          for (int qq = 0; qq < 8; qq++) {
            wP[qq] = sF[m * NP * NPT + lp(pl, l + qq)];
            wM[qq] = sF[(5 + m) * NP * NPT + lp(pl, l + 7 - qq)];
          }
          hh[m] = recon(kP, wP, S.ud7l) + recon(kM, wM, S.ud7l);
        }
      } else {
        const long qa = PB.hq1[pl] + (long)(wrap(s_if, n, S.per[DIR]) - 1) * S.sh[DIR];
        const long qb = PB.hq1[pl] + (long)(wrap(s_if + 1, n, S.per[DIR]) - 1) * S.sh[DIR];
        const double ca = sqrt((S.p[qa] / S.Et[qa] + 1.0) * S.p[qa] / S.d[qa]);
        const double cb = sqrt((S.p[qb] / S.Et[qb] + 1.0) * S.p[qb] / S.d[qb]);
        double Sm[5][5], S1[5][5];
        char_matrices((S.u[qa] + S.u[qb]) * 0.5, (S.v[qa] + S.v[qb]) * 0.5, (S.w[qa] + S.w[qb]) * 0.5,
                      (ca + cb) * 0.5, (S.p[qa] + S.p[qb]) * 0.5, (S.Et[qa] + S.Et[qb]) * 0.5,
                      (M1[qa] + M1[qb]) * 0.5, (M2[qa] + M2[qb]) * 0.5, (M3[qa] + M3[qb]) * 0.5, Sm, S1);
        double h0[5];
#pragma unroll
        for (int r = 0; r < 5; r++) {
          double wP[8], wM[8];
#pragma unroll
          for (int qq = 0; qq < 8; qq++) {
            const int a = lp(pl, l + qq), b = lp(pl, l + 7 - qq);
            wP[qq] = Sm[r][0] * sF[a] + Sm[r][1] * sF[NP * NPT + a] + Sm[r][2] * sF[2 * NP * NPT + a] +
                     Sm[r][3] * sF[3 * NP * NPT + a] + Sm[r][4] * sF[4 * NP * NPT + a];
            wM[qq] = Sm[r][0] * sF[5 * NP * NPT + b] + Sm[r][1] * sF[6 * NP * NPT + b] +
                     Sm[r][2] * sF[7 * NP * NPT + b] + Sm[r][3] * sF[8 * NP * NPT + b] +
                     Sm[r][4] * sF[9 * NP * NPT + b];
          }
          h0[r] = recon(kP, wP, S.ud7l) + recon(kM, wM, S.ud7l);
        }
#pragma unroll
        for (int r = 0; r < 5; r++)
          hh[r] = S1[r][0] * h0[0] + S1[r][1] * h0[1] + S1[r][2] * h0[2] + S1[r][3] * h0[3] + S1[r][4] * h0[4];
      }
    }
#pragma unroll
    for (int m = 0; m < 5; m++) sH[m * NP * NIF + lh(pl, l)] = hh[m];
  }
  __syncthreads();
  const double h1 = S.h_1[DIR];
  for (int t = tid; t < 5 * NP * TS; t += blockDim.x) {
    const int m = t / (NP * TS);
    const int r = t % (NP * TS);
    const int pl = DIR == 0 ? r / TS : r % NP;
    const int c = DIR == 0 ? r % TS : r / NP;
    if (pen0 + pl >= npen || c >= ncell) continue;
    const double val = (sH[m * NP * NIF + lh(pl, c + 1)] - sH[m * NP * NIF + lh(pl, c)]) * h1;
    const long ci = PB.iq1[pl] + (long)(c0 + c - 1) * S.si[DIR] + (long)m * S.ni;
    if (DIR == 0) {
      S.du[ci] = val;
    } else if (DIR == 1) {
      S.du[ci] = S.du[ci] + val;
    } else {
      const long q = PB.hq1[pl] + (long)(c0 + c - 1) * S.sh[DIR];
      S.du[ci] = -(S.du[ci] + val) * S.Ajac[q];
    }
  }
}

#include "ocfd_species.h"
#include "ocfd_diag.h"
#include "ocfd_io.h"

// ===========================================================================
// Viscous terms, part 1 (du_viscous, OCFD_viscous.f90:27-95, with the
// transport coefficients of comput_Amu_Amk_AMD, Comput_flow_variables.f90:69-98):
// per cell, 12 central derivatives of u,v,w,T, Sutherland viscosity,
// kappa = mu*Cp/Pr, stresses and the 12 flux components Ev1/Ev2/Ev3(1:4).
// ===========================================================================
__device__ __forceinline__ double pw3(double x) { double x2 = x * x; return x2 * x; }

// 2.5-D marching launch for the per-cell kernels that differentiate in z.
// A block covers a COL_TX x COL_TY tile of (i, j) and walks a chunk of
// consecutive k planes.  The z-stencil window of every differentiated field
// stays in registers and advances one plane per step, so a plane of those
// fields is read once per chunk instead of once per stencil tap.  Block
// order: tile i fastest, then tile j, then chunk, so the tiles whose rows
// meet in the y stencil run together on one XCD.  Decoding a cell is two
// increments per plane instead of the three 64-bit divisions of the flat
// launch.
constexpr int COL_TX = 64, COL_TY = 4;

struct ColPos {
  int i, j, k0, k1;
};

__device__ __forceinline__ bool col_pos(const DevState& S, int kc, ColPos& P) {
  const int nti = (S.nx + COL_TX - 1) / COL_TX, ntj = (S.ny + COL_TY - 1) / COL_TY;
  const int L = (int)xcd_block((long)blockIdx.x, (long)gridDim.x);
  const int ti = L % nti, r = L / nti;
  const int tj = r % ntj, ch = r / ntj;
  P.i = ti * COL_TX + (int)(threadIdx.x % COL_TX) + 1;
  P.j = tj * COL_TY + (int)(threadIdx.x / COL_TX) + 1;
  P.k0 = ch * kc + 1;
  P.k1 = min(P.k0 + kc, S.nz + 1);
  return P.i <= S.nx && P.j <= S.ny;
}

inline dim3 col_grid(const DevState& S, int kc) {
  const long nti = (S.nx + COL_TX - 1) / COL_TX, ntj = (S.ny + COL_TY - 1) / COL_TY;
  return dim3((unsigned)(nti * ntj * ((S.nz + kc - 1) / kc)));
}

// Ev1..3(1:4) layout: per direction g a block of ni cells x 4 components,
// so a cell's four values of one direction are 32 contiguous bytes.
__host__ __device__ __forceinline__ long ev_index(long ni, int g, long c, int m) { return (long)g * 4 * ni + 4 * c + m; }

// MS (N_SPEC > 1): the transport coefficients come from k_transport_ms and
// the species-diffusion enthalpy flux Qp from k_species_visc_prep.
// (us, vs, ws, Ts) are the zeta derivatives of u, v, w, T at the cell; ctr
// (may be null) holds u, v, w, T at the cell when the caller has them; xt,
// when given (central xi form only), returns u, v, w, T (v = 0..3) at xi
// offset o from the cell; yt likewise along eta.
// INTX / INTY: the cell's xi / eta stencil is central (no wrap, no closure),
// so that derivative runs in straight-line code.
struct NoXTaps {
  __device__ double operator()(int, int) const { return 0.0; }
};
template <int CD, bool MS, bool INTX, bool INTY, bool MT, class XT = NoXTaps, class YT = NoXTaps>
__device__ __forceinline__ void visc_prep_cell(const DevState& S, long c, long q, int i, int j, int k, double us,
                                               double vs, double ws, double Ts, const double* ctr = nullptr,
                                               const XT* xt = nullptr, const YT* yt = nullptr) {
  double amu, amk;
  if (MS) {
    amu = S.Amu[c];
    amk = S.Amk[c];
  } else {
    // transport (N_SPEC = 1): comput_vis_coeff Sutherland (TRANSPORT_modules.f90:265-269),
    // comput_Cp_pgas (CHEM_thermdyn.f90:125-136), Amk = Amu*Cp/Pr (:326-327)
    const double TA = 288.15, TC = 110.4;
    const double T0 = ctr ? ctr[3] : S.T[q];
    const double tr = T0 / TA;
    amu = S.amu0 * sqrt(pw3(tr)) * (TA + TC) / (TC + T0);
    const double Cp = (0.0 + 3.5 * S.Ri0 * S.di[q]) / S.d[q];
    amk = amu * Cp / S.pr;
    S.Amu[c] = amu;
    S.Amk[c] = amk;
  }

  auto dcen = [&](const double* F, int dir, int s) {
    const double* Fq = F + q;
    const int st = dir == 0 ? 1 : S.nxh;
    return deriv1<CD, true>(S, dir, s, [&](int o) { return Fq[o * st]; });
  };
  auto dgen = [&](const double* F, int dir, int s) {
    const double* Fq = F + q;
    const int st = dir == 0 ? 1 : S.nxh;
    return deriv1<CD>(S, dir, s, [&](int o) {
      const int t = wrap(s + o, S.n[dir], S.per[dir]);
      return Fq[(t - s) * st];
    });
  };
  double uk, vk, wk, Tk;
  if (INTX && xt) {
    auto dx = [&](int v) { return deriv1<CD, true>(S, 0, i, [&](int o) { return (*xt)(v, o); }); };
    uk = dx(0); vk = dx(1); wk = dx(2); Tk = dx(3);
  } else if (INTX) {
    uk = dcen(S.u, 0, i); vk = dcen(S.v, 0, i); wk = dcen(S.w, 0, i); Tk = dcen(S.T, 0, i);
  } else {
    uk = dgen(S.u, 0, i); vk = dgen(S.v, 0, i); wk = dgen(S.w, 0, i); Tk = dgen(S.T, 0, i);
  }
  double ui, vi, wi, Ti;
  if (INTY && yt) {
    auto dy = [&](int v) { return deriv1<CD, true>(S, 1, j, [&](int o) { return (*yt)(v, o); }); };
    ui = dy(0); vi = dy(1); wi = dy(2); Ti = dy(3);
  } else if (INTY) {
    ui = dcen(S.u, 1, j); vi = dcen(S.v, 1, j); wi = dcen(S.w, 1, j); Ti = dcen(S.T, 1, j);
  } else {
    ui = dgen(S.u, 1, j); vi = dgen(S.v, 1, j); wi = dgen(S.w, 1, j); Ti = dgen(S.T, 1, j);
  }
  const Metrics M = metrics_at<MT>(S, q, i, j, k);
  const double Akx = M.kx, Aky = M.ky, Akz = M.kz;
  const double Aix = M.ix, Aiy = M.iy, Aiz = M.iz;
  const double Asx = M.sx, Asy = M.sy, Asz = M.sz;
  const double ux = uk * Akx + ui * Aix + us * Asx;
  const double vx = vk * Akx + vi * Aix + vs * Asx;
  const double wx = wk * Akx + wi * Aix + ws * Asx;
  const double Tx = Tk * Akx + Ti * Aix + Ts * Asx;
  const double uy = uk * Aky + ui * Aiy + us * Asy;
  const double vy = vk * Aky + vi * Aiy + vs * Asy;
  const double wy = wk * Aky + wi * Aiy + ws * Asy;
  const double Ty = Tk * Aky + Ti * Aiy + Ts * Asy;
  const double uz = uk * Akz + ui * Aiz + us * Asz;
  const double vz = vk * Akz + vi * Aiz + vs * Asz;
  const double wz = wk * Akz + wi * Aiz + ws * Asz;
  const double Tz = Tk * Akz + Ti * Aiz + Ts * Asz;
  const double div = ux + vy + wz;
  const double tmp2_3 = 2.0 / 3.0;
  const double s11 = (2.0 * ux - tmp2_3 * div) * amu;
  const double s22 = (2.0 * vy - tmp2_3 * div) * amu;
  const double s33 = (2.0 * wz - tmp2_3 * div) * amu;
  const double s12 = (uy + vx) * amu, s13 = (uz + wx) * amu, s23 = (vz + wy) * amu;
  const double uu = ctr ? ctr[0] : S.u[q], vv = ctr ? ctr[1] : S.v[q], ww = ctr ? ctr[2] : S.w[q];
  // Qpx/Qpy/Qpz are zero without species diffusion (OCFD_viscous.f90:13-15)
  const double qx = MS ? S.Qp[c] : 0.0, qy = MS ? S.Qp[c + S.ni] : 0.0, qz = MS ? S.Qp[c + 2 * S.ni] : 0.0;
  const double E1 = uu * s11 + vv * s12 + ww * s13 + amk * Tx + qx;
  const double E2 = uu * s12 + vv * s22 + ww * s23 + amk * Ty + qy;
  const double E3 = uu * s13 + vv * s23 + ww * s33 + amk * Tz + qz;
  double Akx1, Aky1, Akz1, Aix1, Aiy1, Aiz1, Asx1, Asy1, Asz1;
  if constexpr (MT) {  // Akx1 = Akx/Ajac etc. (OCFD_mesh.f90:39-51)
    Akx1 = M.kx / M.jac; Aky1 = M.ky / M.jac; Akz1 = M.kz / M.jac;
    Aix1 = M.ix / M.jac; Aiy1 = M.iy / M.jac; Aiz1 = M.iz / M.jac;
    Asx1 = M.sx / M.jac; Asy1 = M.sy / M.jac; Asz1 = M.sz / M.jac;
  } else {
    Akx1 = S.Akx1[q]; Aky1 = S.Aky1[q]; Akz1 = S.Akz1[q];
    Aix1 = S.Aix1[q]; Aiy1 = S.Aiy1[q]; Aiz1 = S.Aiz1[q];
    Asx1 = S.Asx1[q]; Asy1 = S.Asy1[q]; Asz1 = S.Asz1[q];
  }
  // Ev1..3(1:4) of the cell: 4 consecutive values per direction (ev_index),
  // two 16-byte stores each
  double2* E = (double2*)(S.Ev + ev_index(S.ni, 0, c, 0));
  const long g2 = 2 * S.ni;  // one direction's block, in double2
  E[0] = make_double2(Akx1 * s11 + Aky1 * s12 + Akz1 * s13, Akx1 * s12 + Aky1 * s22 + Akz1 * s23);
  E[1] = make_double2(Akx1 * s13 + Aky1 * s23 + Akz1 * s33, Akx1 * E1 + Aky1 * E2 + Akz1 * E3);
  E[g2] = make_double2(Aix1 * s11 + Aiy1 * s12 + Aiz1 * s13, Aix1 * s12 + Aiy1 * s22 + Aiz1 * s23);
  E[g2 + 1] = make_double2(Aix1 * s13 + Aiy1 * s23 + Aiz1 * s33, Aix1 * E1 + Aiy1 * E2 + Aiz1 * E3);
  E[2 * g2] = make_double2(Asx1 * s11 + Asy1 * s12 + Asz1 * s13, Asx1 * s12 + Asy1 * s22 + Asz1 * s23);
  E[2 * g2 + 1] = make_double2(Asx1 * s13 + Asy1 * s23 + Asz1 * s33, Asx1 * E1 + Asy1 * E2 + Asz1 * E3);
}

// one thread per cell
template <int CD, bool MS, bool MT>
__global__ __launch_bounds__(256) void k_visc_prep(DevState S) {
  constexpr int R = CD == 1 ? 4 : 3;
  const long c = cell_index();
  if (c >= S.ni) return;
  const int i = (int)(c % S.nx) + 1;
  const int j = (int)((c / S.nx) % S.ny) + 1;
  const int k = (int)(c / ((long)S.nx * S.ny)) + 1;
  const long q = hq(S, i, j, k);
  const long plh = (long)S.nxh * S.nyh;
  // central forms per direction (a wavefront at the xi faces keeps the
  // central eta and zeta forms)
  const bool ix = wave_all(central_cell(i, S.nx, R)), iy = wave_all(central_cell(j, S.ny, R));
  double zd[4];
  const double* zf[4] = {S.u, S.v, S.w, S.T};
  if (wave_all(central_cell(k, S.nz, R))) {
#pragma unroll
    for (int v = 0; v < 4; v++)
      zd[v] = deriv1<CD, true>(S, 2, k, [&](int o) { return zf[v][q + o * plh]; });
  } else {
#pragma unroll
    for (int v = 0; v < 4; v++)
      zd[v] = deriv1<CD>(S, 2, k, [&](int o) {
        const int t = wrap(k + o, S.n[2], S.per[2]);
        return zf[v][q + (long)(t - k) * plh];
      });
  }
  if (ix && iy) visc_prep_cell<CD, MS, true, true, MT>(S, c, q, i, j, k, zd[0], zd[1], zd[2], zd[3]);
  else if (iy) visc_prep_cell<CD, MS, false, true, MT>(S, c, q, i, j, k, zd[0], zd[1], zd[2], zd[3]);
  else visc_prep_cell<CD, MS, false, false, MT>(S, c, q, i, j, k, zd[0], zd[1], zd[2], zd[3]);
}

// marching over k planes with the z windows of u, v, w, T in registers.
// The xi and eta taps of interior wavefronts come from LDS: a wavefront is
// one row of COL_TX = 64 cells and posts its cells' u, v, w, T of the plane
// in its own row buffer (the first / last R lanes add the halo cells beside
// the row; no block barrier), and the block stages its COL_TY rows plus R
// halo rows above and below in a plane buffer, double-buffered (the next
// plane's own rows come from the z windows, its halo rows from memory; one
// barrier per plane).  Halo cells are clamped into the grid: only interior
// wavefronts read them, and theirs are real cells.
template <int CD, bool MS, bool MT>
__global__ __launch_bounds__(256) void k_visc_prep_col(DevState S, int kc) {
  constexpr int R = CD == 1 ? 4 : 3, W = 2 * R + 1;
  constexpr int TX = COL_TX, TY = COL_TY, YH = TY + 2 * R;
  static_assert(TX == 64, "one row per wavefront");
  // raised wave priority: the memory-bound preparation issues its loads
  // ahead of the VALU-bound sweeps that share the CUs from the other stream
  __builtin_amdgcn_s_setprio(1);
  ColPos P;
  const bool act = col_pos(S, kc, P);
  const int lane = (int)threadIdx.x % TX, wv = (int)threadIdx.x / TX;
  const int i0 = P.i - lane, j0 = P.j - wv;  // the tile's first column and row
  const int ic = min(P.i, S.nx), jc = min(P.j, S.ny);  // inactive lanes read a real cell
  const long plh = (long)S.nxh * S.nyh, pli = (long)S.nx * S.ny;
  const long qcol = hq(S, ic, jc, 0), ccol = iq(S, ic, jc, 1);
  // per-direction central forms: wavefronts at the xi faces (a fifth of them
  // on 640 cells) keep the central eta and zeta forms
  const bool xint = wave_all(!act || central_cell(P.i, S.nx, R));
  const bool yint = wave_all(!act || central_cell(P.j, S.ny, R));
  double zu[W], zv[W], zw[W], zT[W];
  __shared__ double xs[TY][4][TX + 2 * R];
  __shared__ double ys[2][4][YH][TX];
  double (*xw)[TX + 2 * R] = xs[wv];
  struct XTaps {
    const double (*xw)[TX + 2 * R];
    int lane;
    __device__ double operator()(int v, int o) const { return xw[v][lane + R + o]; }
  } xt{xw, lane};
  struct YTaps {
    const double (*yb)[YH][TX];
    int lane, wv;
    __device__ double operator()(int v, int o) const { return yb[v][wv + R + o][lane]; }
  };
  // plane p (-3 .. nz+4 before the periodic wrap) lies inside the halo box
  auto ld = [&](int slot, int p) {
    const long qp = qcol + (long)wrap(p, S.nz, S.per[2]) * plh;
    zu[slot] = S.u[qp];
    zv[slot] = S.v[qp];
    zw[slot] = S.w[qp];
    zT[slot] = S.T[qp];
  };
  // plane k into ys[b]: own rows from the window slot, halo rows from memory
  auto stage = [&](int b, int k, int slot) {
    ys[b][0][wv + R][lane] = zu[slot];
    ys[b][1][wv + R][lane] = zv[slot];
    ys[b][2][wv + R][lane] = zw[slot];
    ys[b][3][wv + R][lane] = zT[slot];
    const long qk = (long)wrap(k, S.nz, S.per[2]) * plh;
    for (int e = threadIdx.x; e < 2 * R * TX; e += TX * TY) {
      const int r = e / TX, col = e - r * TX;
      const int row = r < R ? r : TY + r;  // rows 0..R-1 and TY+R..TY+2R-1
      const int jj = min(max(j0 - R + row, 1), S.ny), ii = min(i0 + col, S.nx);
      const long qh = hq(S, ii, jj, 0) + qk;
      ys[b][0][row][col] = S.u[qh];
      ys[b][1][row][col] = S.v[qh];
      ys[b][2][row][col] = S.w[qh];
      ys[b][3][row][col] = S.T[qh];
    }
  };
#pragma unroll
  for (int o = 0; o < W - 1; o++) ld(o, P.k0 - R + o);
  stage(0, P.k0, R);
  __syncthreads();
  int b = 0;
  for (int k = P.k0; k < P.k1; k++) {
    ld(W - 1, k + R);
    if (k + 1 < P.k1) stage(b ^ 1, k + 1, R + 1);
    const long c = ccol + (long)(k - 1) * pli, q = qcol + (long)k * plh;
    const double ctr[4] = {zu[R], zv[R], zw[R], zT[R]};  // the cell's values, plane k
    if (xint) {
      // the previous plane's reads of xs are done before these writes (a
      // wavefront's LDS operations complete in order)
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int v = 0; v < 4; v++) xw[v][lane + R] = ctr[v];
      if (lane < R || lane >= TX - R) {
        // halo cell P.i -+ R, clamped into the row (only interior rows read it)
        const int ih = min(max(lane < R ? P.i - R : P.i + R, 1), S.nx);
        const long qh = qcol + (ih - ic) + (long)k * plh;
        const int slot = lane < R ? lane : lane + 2 * R;
        xw[0][slot] = S.u[qh];
        xw[1][slot] = S.v[qh];
        xw[2][slot] = S.w[qh];
        xw[3][slot] = S.T[qh];
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    if (act) {
      double zd[4];
      if (central_cell(k, S.nz, R)) {
        auto dz = [&](const double* z) { return deriv1<CD, true>(S, 2, k, [&](int o) { return z[o + R]; }); };
        zd[0] = dz(zu); zd[1] = dz(zv); zd[2] = dz(zw); zd[3] = dz(zT);
      } else {
        auto dz = [&](const double* z) { return deriv1<CD>(S, 2, k, [&](int o) { return z[o + R]; }); };
        zd[0] = dz(zu); zd[1] = dz(zv); zd[2] = dz(zw); zd[3] = dz(zT);
      }
      const YTaps yt{ys[b], lane, wv};
      if (xint && yint)
        visc_prep_cell<CD, MS, true, true, MT>(S, c, q, P.i, P.j, k, zd[0], zd[1], zd[2], zd[3], ctr, &xt, &yt);
      else if (yint)
        visc_prep_cell<CD, MS, false, true, MT>(S, c, q, P.i, P.j, k, zd[0], zd[1], zd[2], zd[3], ctr,
                                                (const XTaps*)nullptr, &yt);
      else
        visc_prep_cell<CD, MS, false, false, MT>(S, c, q, P.i, P.j, k, zd[0], zd[1], zd[2], zd[3], ctr);
    }
#pragma unroll
    for (int o = 0; o < W - 1; o++) {
      zu[o] = zu[o + 1];
      zv[o] = zv[o + 1];
      zw[o] = zw[o + 1];
      zT[o] = zT[o + 1];
    }
    __syncthreads();
    b ^= 1;
  }
}

// Transport only (when the viscous terms are off the reference still
// evaluates Amu/Amk at every stage; kept for the per-stage ABI).
__global__ __launch_bounds__(256) void k_transport(DevState S) {
  const long c = cell_index();
  if (c >= S.ni) return;
  const int i = (int)(c % S.nx) + 1;
  const int j = (int)((c / S.nx) % S.ny) + 1;
  const int k = (int)(c / ((long)S.nx * S.ny)) + 1;
  const long q = hq(S, i, j, k);
  const double TA = 288.15, TC = 110.4;
  const double T0 = S.T[q];
  const double tr = T0 / TA;
  const double amu = S.amu0 * sqrt(pw3(tr)) * (TA + TC) / (TC + T0);
  const double Cp = (0.0 + 3.5 * S.Ri0 * S.di[q]) / S.d[q];
  S.Amu[c] = amu;
  S.Amk[c] = amu * Cp / S.pr;
}

// ===========================================================================
// Per-cell update kernel, fused for the stepping path:
//   du(2:5) += (dEv1/dxi + dEv2/deta + dEv3/dzeta)*Ajac  (OCFD_viscous.f90:111-129)
//   du += mass force                                     (OCFD_inviscous.f90:187-199)
//   f = alpha*fn + beta*(f + dt*du)                      (OCFD_Time_Adv.f90:98-130)
//   d,u,v,w,Et from f; Rescaling_di; T(E); p             (Comput_flow_variables.f90:3-46,101-121)
// The flags select the pieces so the per-stage entry points share this code.
// A cell with Et <= 0 is recorded for handle_NegativeEt (k_negt_repair).
// ===========================================================================
enum { RK_MASS = 1, RK_VISC = 2, RK_WRITE_DU = 4, RK_UPDATE_F = 8, RK_PRIMS = 16, RK_DI_FROM_D = 32, RK_MS = 64 };

// The stepping path's flag sets are compile-time (F >= 0), so the kernel is
// straight-line code and its loads can be in flight together; F = -1 reads
// the flags at run time (per-stage entry points).
// zs[m] is the zeta derivative of Ev3(m+1) at the cell (read only with RK_VISC).
// INTX / INTY: the cell's xi / eta stencil is central (no wrap, no closure).
template <int CD, int F, bool INTX, bool INTY, bool MT>
__device__ __forceinline__ void rk_cell(const DevState& S, long c, long q, int i, int j, int k, int krk, double dt,
                                        int flags_rt, const double* zs) {
  const int flags = F >= 0 ? F : flags_rt;
  const long ni = S.ni;
  const double d_old = S.d[q];
  double du[5];
  if (flags & (RK_MASS | RK_VISC | RK_UPDATE_F)) {
#pragma unroll
    for (int m = 0; m < 5; m++) du[m] = S.du[c + m * ni];
  }
  // the stage's old solution, read before the flux divergence so that its
  // loads overlap the stencil loads
  double fnv[5], fov[5];
  if (flags & RK_UPDATE_F) {
#pragma unroll
    for (int m = 0; m < 5; m++) fnv[m] = S.fn[c + m * ni];
    if (krk != 1) {
#pragma unroll
      for (int m = 0; m < 5; m++) fov[m] = S.f[c + m * ni];
    }
  }
  if (flags & RK_MASS) {
    du[1] = du[1] + d_old * S.mf[0];
    du[2] = du[2] + d_old * S.mf[1];
    du[3] = du[3] + d_old * S.mf[2];
    du[4] = du[4] + d_old * (S.u[q] * S.mf[0] + S.v[q] * S.mf[1] + S.w[q] * S.mf[2]);
  }
  if (flags & RK_VISC) {
    const double ajac = jac_at<MT>(S, q, i, j, k);
    double uk[4], ui[4];
    constexpr int R = CD == 1 ? 4 : 3;
    // generic taps (closures, periodic wrap) of component m of direction g
    auto dv = [&](int g, int m, int dir, int s) {
      const double* Fc = S.Ev + ev_index(ni, g, c, m);
      const int st = dir == 0 ? 4 : 4 * S.nx;  // ev_index steps 4 per cell
      return deriv1<CD>(S, dir, s, [&](int o) {
        const int t = wrap(s + o, S.n[dir], S.per[dir]);
        return Fc[(t - s) * st];
      });
    };
    if constexpr (INTX && INTY) {
      // central taps: the four components of a neighbour in two 16-byte loads
      const double* bx = S.Ev + ev_index(ni, 0, c, 0);
      const double* by = S.Ev + ev_index(ni, 1, c, 0);
      const long sy = 4L * S.nx;
      double tx[2 * R + 1][4], ty[2 * R + 1][4];
#pragma unroll
      for (int o = -R; o <= R; o++) {
        if (o == 0) continue;
        const double2 xa = *(const double2*)(bx + 4 * o), xb = *(const double2*)(bx + 4 * o + 2);
        const double2 ya = *(const double2*)(by + o * sy), yb = *(const double2*)(by + o * sy + 2);
        tx[o + R][0] = xa.x; tx[o + R][1] = xa.y; tx[o + R][2] = xb.x; tx[o + R][3] = xb.y;
        ty[o + R][0] = ya.x; ty[o + R][1] = ya.y; ty[o + R][2] = yb.x; ty[o + R][3] = yb.y;
      }
#pragma unroll
      for (int m = 0; m < 4; m++) {
        uk[m] = deriv1<CD, true>(S, 0, i, [&](int o) { return tx[o + R][m]; });
        ui[m] = deriv1<CD, true>(S, 1, j, [&](int o) { return ty[o + R][m]; });
      }
    } else if constexpr (INTY) {
      // xi face: generic xi taps, central eta taps
#pragma unroll
      for (int m = 0; m < 4; m++) uk[m] = dv(0, m, 0, i);
      const double* by = S.Ev + ev_index(ni, 1, c, 0);
      const long sy = 4L * S.nx;
      double ty[2 * R + 1][4];
#pragma unroll
      for (int o = -R; o <= R; o++) {
        if (o == 0) continue;
        const double2 ya = *(const double2*)(by + o * sy), yb = *(const double2*)(by + o * sy + 2);
        ty[o + R][0] = ya.x; ty[o + R][1] = ya.y; ty[o + R][2] = yb.x; ty[o + R][3] = yb.y;
      }
#pragma unroll
      for (int m = 0; m < 4; m++) ui[m] = deriv1<CD, true>(S, 1, j, [&](int o) { return ty[o + R][m]; });
    } else {
      static_assert(!INTX, "a central xi form comes with a central eta form");
#pragma unroll
      for (int m = 0; m < 4; m++) {
        uk[m] = dv(0, m, 0, i);
        ui[m] = dv(1, m, 1, j);
      }
    }
#pragma unroll
    for (int m = 0; m < 4; m++) du[m + 1] = du[m + 1] + (uk[m] + ui[m] + zs[m]) * ajac;
  }
  if (flags & RK_WRITE_DU) {
#pragma unroll
    for (int m = 0; m < 5; m++) S.du[c + m * ni] = du[m];
  }
  double f[5];
  if (flags & RK_UPDATE_F) {
    if (krk == 1) {
#pragma unroll
      for (int m = 0; m < 5; m++) f[m] = fnv[m] + dt * du[m];
    } else {
      const double ra = krk == 2 ? 3.0 / 4.0 : 1.0 / 3.0;
      const double rb = krk == 2 ? 1.0 / 4.0 : 2.0 / 3.0;
#pragma unroll
      for (int m = 0; m < 5; m++) f[m] = ra * fnv[m] + rb * (fov[m] + dt * du[m]);
    }
#pragma unroll
    for (int m = 0; m < 5; m++) S.f[c + m * ni] = f[m];
  } else {
#pragma unroll
    for (int m = 0; m < 5; m++) f[m] = S.f[c + m * ni];
  }
  if (!(flags & RK_PRIMS)) return;
  const double dd = f[0];
  const double uu = f[1] / f[0], vv = f[2] / f[0], ww = f[3] / f[0];
  const double Et = f[4] - (f[1] * uu + f[2] * vv + f[3] * ww) * 0.5;
  S.Et_prev[c] = S.Et[q];
  if (Et <= 0.0) {
    const int slot = atomicAdd(&S.negt[0], 1);
    if (slot < 16) S.negt[1 + slot] = (int)c;
  }
  if (flags & RK_MS) {
    double T, p;
    ms_prims(S, q, dd, Et, T, p);
    S.d[q] = dd;
    S.u[q] = uu;
    S.v[q] = vv;
    S.w[q] = ww;
    S.Et[q] = Et;
    S.T[q] = T;
    S.p[q] = p;
    return;
  }
  // N_SPEC == 1 stepping path: di was set to d at the stage start
  // (OCFD_Time_Adv.f90:45-52), so the old d is the di being rescaled.
  double di = (flags & RK_DI_FROM_D) ? d_old : S.di[q];
  if (di < 0.0) di = 0.0;
  const double d0 = 0.0 + di;
  const double tmp = dd / d0;
  di = di * tmp;
  const double T = Et / (0.0 + 2.5 * S.Ri0 * di);   // comput_T_pgas
  const double p = 0.0 + S.Ri0 * di * T;            // comput_P
  S.d[q] = dd;
  S.u[q] = uu;
  S.v[q] = vv;
  S.w[q] = ww;
  S.Et[q] = Et;
  S.di[q] = di;
  S.T[q] = T;
  S.p[q] = p;
}

// one thread per cell
template <int CD, int F, bool MT>
__global__ __launch_bounds__(256) void k_rk(DevState S, int krk, double dt, int flags_rt) {
  constexpr int R = CD == 1 ? 4 : 3;
  const int flags = F >= 0 ? F : flags_rt;
  const long c = cell_index();
  if (c >= S.ni) return;
  const int i = (int)(c % S.nx) + 1;
  const int j = (int)((c / S.nx) % S.ny) + 1;
  const int k = (int)(c / ((long)S.nx * S.ny)) + 1;
  const long q = hq(S, i, j, k);
  const long pli = (long)S.nx * S.ny;
  double zs[4] = {0.0, 0.0, 0.0, 0.0};
  if (wave_all(central_cell(i, S.nx, R) && central_cell(j, S.ny, R) && central_cell(k, S.nz, R))) {
    if (flags & RK_VISC) {
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const double* Fv = S.Ev + ev_index(S.ni, 2, c, m);
        zs[m] = deriv1<CD, true>(S, 2, k, [&](int o) { return Fv[4 * o * pli]; });
      }
    }
    rk_cell<CD, F, true, true, MT>(S, c, q, i, j, k, krk, dt, flags_rt, zs);
  } else {
    if (flags & RK_VISC) {
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const double* Fv = S.Ev + ev_index(S.ni, 2, c, m);
        zs[m] = deriv1<CD>(S, 2, k, [&](int o) {
          const int t = wrap(k + o, S.n[2], S.per[2]);
          return Fv[4 * (long)(t - k) * pli];
        });
      }
    }
    rk_cell<CD, F, false, false, MT>(S, c, q, i, j, k, krk, dt, flags_rt, zs);
  }
}

// marching over k planes with the z windows of Ev3(1:4) in registers
template <int CD, int F, bool MT>
// (the CD6 N_SPEC = 1 stepping path, with its three interior forms, held to
// 2 waves/SIMD)
__global__ __launch_bounds__(256, (CD == 0 && F >= 0 && !(F & RK_MS)) ? 2 : 1) void k_rk_col(DevState S, int kc, int krk, double dt,
                                                                              int flags_rt) {
  constexpr int R = CD == 1 ? 4 : 3, W = 2 * R + 1;
  const int flags = F >= 0 ? F : flags_rt;
  ColPos P;
  if (!col_pos(S, kc, P)) return;
  const long plh = (long)S.nxh * S.nyh, pli = (long)S.nx * S.ny;
  const long qcol = hq(S, P.i, P.j, 0), ccol = iq(S, P.i, P.j, 1);
  const bool visc = flags & RK_VISC;
  // per-direction central forms: wavefronts at the xi faces (a fifth of them
  // on 640 cells) keep the central eta and zeta forms
  const bool xint = wave_all(central_cell(P.i, S.nx, R));
  const bool yint = wave_all(central_cell(P.j, S.ny, R));
  double z[4][W];
  // Ev is stored over the interior only: planes outside 1..nz of a
  // non-periodic z are never read by the one-sided closures and stay zero
  auto ld = [&](int slot, int p) {
    p = wrap(p, S.nz, S.per[2]);
    if (visc && p >= 1 && p <= S.nz) {
      const double* Fv = S.Ev + ev_index(S.ni, 2, ccol + (long)(p - 1) * pli, 0);
#pragma unroll
      for (int m = 0; m < 4; m++) z[m][slot] = Fv[m];
    } else {
#pragma unroll
      for (int m = 0; m < 4; m++) z[m][slot] = 0.0;
    }
  };
#pragma unroll
  for (int o = 0; o < W - 1; o++) ld(o, P.k0 - R + o);
  for (int k = P.k0; k < P.k1; k++) {
    ld(W - 1, k + R);
    const long c = ccol + (long)(k - 1) * pli, q = qcol + (long)k * plh;
    double zs[4];
    // per-direction forms on the CD6 N_SPEC = 1 stepping path; elsewhere the
    // extra code would push the kernel into register spills, so one interior
    // form
    constexpr bool XFACE = CD == 0 && F >= 0 && !(F & RK_MS);
    if constexpr (XFACE) {
      if (central_cell(k, S.nz, R)) {
#pragma unroll
        for (int m = 0; m < 4; m++) zs[m] = visc ? deriv1<CD, true>(S, 2, k, [&](int o) { return z[m][o + R]; }) : 0.0;
      } else {
#pragma unroll
        for (int m = 0; m < 4; m++) zs[m] = visc ? deriv1<CD>(S, 2, k, [&](int o) { return z[m][o + R]; }) : 0.0;
      }
      if (xint && yint) rk_cell<CD, F, true, true, MT>(S, c, q, P.i, P.j, k, krk, dt, flags_rt, zs);
      else if (yint) rk_cell<CD, F, false, true, MT>(S, c, q, P.i, P.j, k, krk, dt, flags_rt, zs);
      else rk_cell<CD, F, false, false, MT>(S, c, q, P.i, P.j, k, krk, dt, flags_rt, zs);
    } else if (xint && yint && central_cell(k, S.nz, R)) {
#pragma unroll
      for (int m = 0; m < 4; m++) zs[m] = visc ? deriv1<CD, true>(S, 2, k, [&](int o) { return z[m][o + R]; }) : 0.0;
      rk_cell<CD, F, true, true, MT>(S, c, q, P.i, P.j, k, krk, dt, flags_rt, zs);
    } else {
#pragma unroll
      for (int m = 0; m < 4; m++) zs[m] = visc ? deriv1<CD>(S, 2, k, [&](int o) { return z[m][o + R]; }) : 0.0;
      rk_cell<CD, F, false, false, MT>(S, c, q, P.i, P.j, k, krk, dt, flags_rt, zs);
    }
#pragma unroll
    for (int o = 0; o < W - 1; o++)
#pragma unroll
      for (int m = 0; m < 4; m++) z[m][o] = z[m][o + 1];
  }
}

// di = d over the whole halo box for N_SPEC == 1 (OCFD_Time_Adv.f90:45-52)
__global__ __launch_bounds__(256) void k_di_from_d(DevState S) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < S.nh) S.di[q] = S.d[q];
}

// start of a stage's negative-Et bookkeeping (Num_NegT=0, Comput_flow_variables.f90:8)
__global__ void k_negt_reset(DevState S, int all) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    S.negt[0] = 0;
    if (all) S.negt[17] = 0;
  }
}

// handle_NegativeEt (OCFD_handle_NegativeT.f90:1-20), replayed in the
// reference's (k,j,i) order by one thread: a neighbour already visited by
// the reference loop holds its new Et, a later one (and every ghost) its old.
__global__ void k_negt_repair(DevState S) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int cnt = S.negt[0];
  if (cnt > S.negt[17]) S.negt[17] = cnt;
  if (cnt == 0) return;
  int n = cnt < 16 ? cnt : 16;
  int idx[16];
  for (int a = 0; a < n; a++) idx[a] = S.negt[1 + a];
  for (int a = 1; a < n; a++) {
    int x = idx[a], b = a - 1;
    while (b >= 0 && idx[b] > x) { idx[b + 1] = idx[b]; b--; }
    idx[b + 1] = x;
  }
  for (int a = 0; a < n; a++) {
    const long c = idx[a];
    const int i = (int)(c % S.nx) + 1;
    const int j = (int)((c / S.nx) % S.ny) + 1;
    const int k = (int)(c / ((long)S.nx * S.ny)) + 1;
    const long q = hq(S, i, j, k);
    auto nb = [&](int ii, int jj, int kk) -> double {
      const bool inside = ii >= 1 && ii <= S.nx && jj >= 1 && jj <= S.ny && kk >= 1 && kk <= S.nz;
      if (inside) {
        const long cn = iq(S, ii, jj, kk);
        return cn < c ? S.Et[hq(S, ii, jj, kk)] : S.Et_prev[cn];
      }
      const int wi = wrap(ii, S.nx, S.per[0]), wj = wrap(jj, S.ny, S.per[1]), wk = wrap(kk, S.nz, S.per[2]);
      const bool img = wi >= 1 && wi <= S.nx && wj >= 1 && wj <= S.ny && wk >= 1 && wk <= S.nz;
      return img ? S.Et_prev[iq(S, wi, wj, wk)] : S.Et[hq(S, ii, jj, kk)];
    };
    const double Et0 = S.Et[q];
    const double En = (nb(i + 1, j, k) + nb(i - 1, j, k) + nb(i, j + 1, k) + nb(i, j - 1, k) + nb(i, j, k + 1) +
                       nb(i, j, k - 1)) / 6.0;
    S.Et[q] = En;
    S.f[c + 4 * S.ni] = S.f[c + 4 * S.ni] + (En - Et0);
    if (S.ns > 1) {
      double dim[OCFD_NS_MAX];
      for (int m = 0; m < S.ns; m++) dim[m] = S.di[q + (long)m * S.nh];
      const double T = gas_T(dim, En);
      S.T[q] = T;
      S.p[q] = gas_P(dim, T);
      continue;
    }
    const double di = S.di[q];
    const double T = En / (0.0 + 2.5 * S.Ri0 * di);
    S.T[q] = T;
    S.p[q] = 0.0 + S.Ri0 * di * T;
  }
}

// ===========================================================================
// Boundary conditions (OCFD_bc, OCFD_BC.f90:2-188; boundary layer case
// ocfd_bc_BoundaryLayer, OCFD_BC_Boundarylayers.f90:20-243).
// ===========================================================================
// Boundary data of one face point: variable m (d, u, v, w, T, di) of the
// 0-D, 1-D or 2-D data (kind 0/1/2) at position a (1..na) along the face's
// first direction and b (1..nb) along the second.
__device__ __forceinline__ double bl_data(const double* r, int kind, int a, int b, int na, int nb, int m) {
  if (kind == 0) return r[m];
  if (kind == 1) return r[(a - 1) + (long)na * m];
  return r[(a - 1) + (long)na * ((b - 1) + (long)nb * m)];
}

__global__ void k_bl_inlet(DevState S) {  // :50-97 (BcData_inlet = 0, 1, 2)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)S.ny * S.nz) return;
  const int j = (int)(t % S.ny) + 1, k = (int)(t / S.ny) + 1;
  const long q = hq(S, 1, j, k);
  const double* r = S.bl_in;
  const int kind = S.bl_inlet;
  S.d[q] = bl_data(r, kind, j, k, S.ny, S.nz, 0);
  S.u[q] = bl_data(r, kind, j, k, S.ny, S.nz, 1);
  S.v[q] = bl_data(r, kind, j, k, S.ny, S.nz, 2);
  S.w[q] = bl_data(r, kind, j, k, S.ny, S.nz, 3);
  S.T[q] = bl_data(r, kind, j, k, S.ny, S.nz, 4);
  S.di[q] = bl_data(r, kind, j, k, S.ny, S.nz, 5);
}

__global__ void k_bl_upper(DevState S) {  // :100-140 (BcData_upper = 0, 1, 2; Dirichlet where not non-reflecting)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)S.nx * S.nz) return;
  if (S.bl_nonref[t] != 0) return;
  const int i = (int)(t % S.nx) + 1, k = (int)(t / S.nx) + 1;
  const long q = hq(S, i, S.ny, k);
  const double* r = S.bl_up;
  const int kind = S.bl_upper;
  S.d[q] = bl_data(r, kind, i, k, S.nx, S.nz, 0);
  S.u[q] = bl_data(r, kind, i, k, S.nx, S.nz, 1);
  S.v[q] = bl_data(r, kind, i, k, S.nx, S.nz, 2);
  S.w[q] = bl_data(r, kind, i, k, S.nx, S.nz, 3);
  S.T[q] = bl_data(r, kind, i, k, S.nx, S.nz, 4);
  S.di[q] = bl_data(r, kind, i, k, S.nx, S.nz, 5);
}

__global__ void k_bl_wall(DevState S) {  // :143-201 (isothermal, non-catalytic, blow/suction)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)S.nx * S.nz) return;
  const int i = (int)(t % S.nx) + 1, k = (int)(t / S.nx) + 1;
  const long q1 = hq(S, i, 1, k), q2 = hq(S, i, 2, k), q3 = hq(S, i, 3, k);
  if (S.Axx[q1] < S.bl_xinit) {
    S.d[q1] = S.d[q2];
    S.u[q1] = S.u[q2];
    S.v[q1] = S.v[q2];
    S.w[q1] = S.w[q2];
    S.T[q1] = S.T[q2];
    S.di[q1] = S.di[q2];
    return;
  }
  const double tmp = 1.0 / sqrt(sqr(S.Aix[q1]) + sqr(S.Aiy[q1]) + sqr(S.Aiz[q1]));
  const double A0 = S.bl_disA * S.bl_ht[0] * S.bl_pertb[t];
  S.u[q1] = A0 * S.Aix[q1] * tmp;
  S.v[q1] = A0 * S.Aiy[q1] * tmp;
  S.w[q1] = A0 * S.Aiz[q1] * tmp;
  double T1;
  if (S.bl_tw > 0) T1 = S.bl_tw;
  else T1 = (4.0 * S.T[q2] - S.T[q3]) / 3.0;
  S.T[q1] = T1;
  const double di2 = S.di[q2], di3 = S.di[q3];
  const double p2 = 0.0 + S.Ri0 * di2 * S.T[q2];
  const double p3 = 0.0 + S.Ri0 * di3 * S.T[q3];
  const double p1 = (4.0 * p2 - p3) / 3.0;
  const double di1 = di2 / S.d[q2];
  const double R1 = 0.0 + S.Ri0 * di1;              // comput_d, CHEM_thermdyn.f90:399-409
  const double d1 = p1 / (T1 * R1);
  S.d[q1] = d1;
  S.di[q1] = di1 * d1;
}

__global__ void k_bl_outlet(DevState S) {  // :203-232 (types 1 and 2)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)S.ny * S.nz) return;
  const int j = (int)(t % S.ny) + 1, k = (int)(t / S.ny) + 1;
  const long q = hq(S, S.nx, j, k), qa = hq(S, S.nx - 1, j, k), qb = hq(S, S.nx - 2, j, k);
  double* arr[6] = {S.d, S.u, S.v, S.w, S.T, S.di};
  for (int a = 0; a < 6; a++) arr[a][q] = S.bl_outlet == 1 ? arr[a][qa] : 2.0 * arr[a][qa] - arr[a][qb];
}

// Face re-synthesis of conserved variables (OCFD_BC.f90:33-184), applied on
// every face in the reference's face order x-, x+, y-, y+, z-, z+; an edge or
// corner cell gets each of its faces' updates in that order.
__device__ __forceinline__ void face_update(const DevState& S, long q, long c) {
  double di = S.di[q];
  const double d0 = 0.0 + di;
  const double dtmp = S.d[q] / d0;
  di = di * dtmp;
  S.di[q] = di;
  const double T = S.T[q];
  const double Et = 0.0 + 2.5 * S.Ri0 * di * T;  // comput_E_pgas
  const double p = 0.0 + S.Ri0 * di * T;          // comput_P
  S.Et[q] = Et;
  S.p[q] = p;
  const double dd = S.d[q], uu = S.u[q], vv = S.v[q], ww = S.w[q];
  S.f[c] = dd;
  S.f[c + S.ni] = dd * uu;
  S.f[c + 2 * S.ni] = dd * vv;
  S.f[c + 3 * S.ni] = dd * ww;
  S.f[c + 4 * S.ni] = Et + dd * (uu * uu + vv * vv + ww * ww) * 0.5;
}

__global__ void k_bc_faces(DevState S) {
  const int nx = S.nx, ny = S.ny, nz = S.nz;
  const long nxi = nx > 2 ? nx - 2 : 0, nyi = ny > 2 ? ny - 2 : 0;
  const long n0 = (long)ny * nz, n1 = nx > 1 ? (long)ny * nz : 0;
  const long n2 = nxi * nz, n3 = ny > 1 ? nxi * nz : 0;
  const long n4 = nxi * nyi, n5 = nz > 1 ? nxi * nyi : 0;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (t < n0) { i = 1; j = (int)(t % ny) + 1; k = (int)(t / ny) + 1; }
  else if ((t -= n0) < n1) { i = nx; j = (int)(t % ny) + 1; k = (int)(t / ny) + 1; }
  else if ((t -= n1) < n2) { i = (int)(t % nxi) + 2; j = 1; k = (int)(t / nxi) + 1; }
  else if ((t -= n2) < n3) { i = (int)(t % nxi) + 2; j = ny; k = (int)(t / nxi) + 1; }
  else if ((t -= n3) < n4) { i = (int)(t % nxi) + 2; j = (int)(t / nxi) + 2; k = 1; }
  else if ((t -= n4) < n5) { i = (int)(t % nxi) + 2; j = (int)(t / nxi) + 2; k = nz; }
  else return;
  const long q = hq(S, i, j, k), c = iq(S, i, j, k);
  const int nf = (i == 1) + (i == nx) + (j == 1) + (j == ny) + (k == 1) + (k == nz);
  for (int a = 0; a < nf; a++) {
    if (S.ns > 1) face_update_ms(S, q, c);
    else face_update(S, q, c);
  }
}

// ===========================================================================
// Ghost cells (Ghost_Cell(1:6) = 1 or 2, OCFD_GhostCell.f90).  Face f:
// 0 i-, 1 i+, 2 j-, 3 j+, 4 k-, 5 k+; one thread per (ghost layer, tangential
// point) over the interior tangential range.  Every ghost point reads only
// interior points, so the points of a face are independent.
// ===========================================================================
struct GhostPt {
  int dir, hi, n, sg, sb, sm;  // direction, high face, extent, ghost / boundary / mirror positions
  int a, b;                    // tangential indices
};
__device__ __forceinline__ bool ghost_point(const DevState& S, int f, long t, GhostPt& g) {
  g.dir = f >> 1;
  g.hi = f & 1;
  g.n = S.n[g.dir];
  const int na = g.dir == 0 ? S.ny : S.nx, nb = g.dir == 2 ? S.ny : S.nz;
  if (t >= (long)OCFD_LAP * na * nb) return false;
  const int l = (int)(t % OCFD_LAP) + 1;
  const long r = t / OCFD_LAP;
  g.a = (int)(r % na) + 1;
  g.b = (int)(r / na) + 1;
  g.sg = g.hi ? g.n + l : 1 - l;
  g.sb = g.hi ? g.n : 1;
  g.sm = g.hi ? 2 * g.n - g.sg : 2 - g.sg;
  return true;
}
__device__ __forceinline__ long ghost_hq(const DevState& S, const GhostPt& g, int s) {
  return g.dir == 0 ? hq(S, s, g.a, g.b) : g.dir == 1 ? hq(S, g.a, s, g.b) : hq(S, g.a, g.b, s);
}

// Flow_Ghost_Extent_2nd (OCFD_GhostCell.f90:166-364): linear extrapolation of
// u, v, w, d, T and the partial densities, d and T limited to 0.8..1.2 of the
// boundary value, partial densities rescaled to d, then Et = E(di, T), p = p(di, T)
__global__ void k_flow_ghost(DevState S, int f) {
  GhostPt g;
  if (!ghost_point(S, f, (long)blockIdx.x * blockDim.x + threadIdx.x, g)) return;
  const long q = ghost_hq(S, g, g.sg), q0 = ghost_hq(S, g, g.sb), qm = ghost_hq(S, g, g.sm);
  const double s1 = 0.8, s2 = 1.2;
  S.u[q] = 2.0 * S.u[q0] - S.u[qm];
  S.v[q] = 2.0 * S.v[q0] - S.v[qm];
  S.w[q] = 2.0 * S.w[q0] - S.w[qm];
  double d = 2.0 * S.d[q0] - S.d[qm];
  double T = 2.0 * S.T[q0] - S.T[qm];
  if (d < s1 * S.d[q0]) d = s1 * S.d[q0];
  if (d > s2 * S.d[q0]) d = s2 * S.d[q0];
  if (T < s1 * S.T[q0]) T = s1 * S.T[q0];
  if (T > s2 * S.T[q0]) T = s2 * S.T[q0];
  S.d[q] = d;
  S.T[q] = T;
  if (S.ns == 1) {
    double di = 2.0 * S.di[q0] - S.di[qm];
    const double d0 = 0.0 + di;
    di = di * (d / d0);
    S.di[q] = di;
    S.Et[q] = 0.0 + 2.5 * S.Ri0 * di * T;  // comput_E_pgas
    S.p[q] = 0.0 + S.Ri0 * di * T;          // comput_P
    return;
  }
  double di1[OCFD_NS_MAX];
  double d0 = 0.0;
  for (int m = 0; m < S.ns; m++) {
    di1[m] = 2.0 * S.di[q0 + (long)m * S.nh] - S.di[qm + (long)m * S.nh];
    d0 = d0 + di1[m];
  }
  const double dtmp = d / d0;
  for (int m = 0; m < S.ns; m++) {
    di1[m] = di1[m] * dtmp;
    S.di[q + (long)m * S.nh] = di1[m];
  }
  S.Et[q] = gas_E(di1, T);
  S.p[q] = gas_P(di1, T);
}

// Jac_Ghost_Extent_1st (OCFD_GhostCell.f90:17-152): metrics copied from the
// boundary point, coordinates extrapolated linearly
__global__ void k_jac_ghost1(DevState S, int f) {
  GhostPt g;
  if (!ghost_point(S, f, (long)blockIdx.x * blockDim.x + threadIdx.x, g)) return;
  const long q = ghost_hq(S, g, g.sg), q0 = ghost_hq(S, g, g.sb), q1 = ghost_hq(S, g, g.hi ? g.n - 1 : 2);
  double* M[10] = {S.Akx, S.Aky, S.Akz, S.Aix, S.Aiy, S.Aiz, S.Asx, S.Asy, S.Asz, S.Ajac};
  for (int m = 0; m < 10; m++) M[m][q] = M[m][q0];
  double* X[3] = {S.Axx, S.Ayy, S.Azz};
  for (int c = 0; c < 3; c++)
    X[c][q] = g.hi ? X[c][q0] + (double)(g.sg - g.n) * (X[c][q0] - X[c][q1])
                   : X[c][q0] + (1.0 - g.sg) * (X[c][q0] - X[c][q1]);
}

// Jac_Ghost_Extent_2nd (OCFD_GhostCell.f90:368-470): coordinates extrapolated
// to second order ...
__global__ void k_jac_ghost2_coords(DevState S, int f) {
  GhostPt g;
  if (!ghost_point(S, f, (long)blockIdx.x * blockDim.x + threadIdx.x, g)) return;
  const long q = ghost_hq(S, g, g.sg), q0 = ghost_hq(S, g, g.sb), qm = ghost_hq(S, g, g.sm);
  S.Axx[q] = 2.0 * S.Axx[q0] - S.Axx[qm];
  S.Ayy[q] = 2.0 * S.Ayy[q0] - S.Ayy[qm];
  S.Azz[q] = 2.0 * S.Azz[q0] - S.Azz[qm];
}

// ... and the metrics of the ghost block from them (comput_Jacobian3d_Ghost,
// :473-580): per direction, one-sided 2nd order at the block's ends, central
// 2nd order next to them, central 4th order inside; the block spans the
// ghost layers plus the boundary plane along the face normal and the
// interior along the tangential directions
__global__ void k_jac_ghost2_metrics(DevState S, int f) {
  GhostPt g;
  if (!ghost_point(S, f, (long)blockIdx.x * blockDim.x + threadIdx.x, g)) return;
  int pos[3];
  pos[g.dir] = g.sg;
  pos[g.dir == 0 ? 1 : 0] = g.a;
  pos[g.dir == 2 ? 1 : 2] = g.b;
  const long q = hq(S, pos[0], pos[1], pos[2]);
  auto der = [&](const double* X, int e) {
    const int r0 = e == g.dir ? (g.hi ? g.n : 1 - OCFD_LAP) : 1;
    const int r1 = e == g.dir ? (g.hi ? g.n + OCFD_LAP : 1) : S.n[e];
    const long st = S.sh[e];
    const double h = S.dh[e];
    const int p = pos[e];
    auto x = [&](int o) { return X[q + (long)o * st]; };
    if (p == r0) return (-3.0 * x(0) + 4.0 * x(1) - x(2)) / (2.0 * h);
    if (p == r1) return (x(-2) - 4.0 * x(-1) + 3.0 * x(0)) / (2.0 * h);
    if (p == r0 + 1 || p == r1 - 1) return (x(1) - x(-1)) / (2.0 * h);
    return (8.0 * (x(1) - x(-1)) - (x(2) - x(-2))) / (12.0 * h);
  };
  const double xi1 = der(S.Axx, 0), yi1 = der(S.Ayy, 0), zi1 = der(S.Azz, 0);
  const double xj1 = der(S.Axx, 1), yj1 = der(S.Ayy, 1), zj1 = der(S.Azz, 1);
  const double xk1 = der(S.Axx, 2), yk1 = der(S.Ayy, 2), zk1 = der(S.Azz, 2);
  const double Jac1 = 1.0 / (xi1 * yj1 * zk1 + yi1 * zj1 * xk1 + zi1 * xj1 * yk1 - zi1 * yj1 * xk1 - yi1 * xj1 * zk1 -
                             xi1 * zj1 * yk1);
  S.Ajac[q] = Jac1;
  S.Akx[q] = Jac1 * (yj1 * zk1 - zj1 * yk1);
  S.Aky[q] = Jac1 * (zj1 * xk1 - xj1 * zk1);
  S.Akz[q] = Jac1 * (xj1 * yk1 - yj1 * xk1);
  S.Aix[q] = Jac1 * (yk1 * zi1 - zk1 * yi1);
  S.Aiy[q] = Jac1 * (zk1 * xi1 - xk1 * zi1);
  S.Aiz[q] = Jac1 * (xk1 * yi1 - yk1 * xi1);
  S.Asx[q] = Jac1 * (yi1 * zj1 - zi1 * yj1);
  S.Asy[q] = Jac1 * (zi1 * xj1 - xi1 * zj1);
  S.Asz[q] = Jac1 * (xi1 * yj1 - yi1 * xj1);
}

// fn = f, din = di  (OCFD_Time_Adv.f90:10-29)
__global__ __launch_bounds__(256) void k_save0(DevState S) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ni) return;
  for (int m = 0; m < 5; m++) S.fn[c + m * S.ni] = S.f[c + m * S.ni];
  const int i = (int)(c % S.nx) + 1;
  const int j = (int)((c / S.nx) % S.ny) + 1;
  const int k = (int)(c / ((long)S.nx * S.ny)) + 1;
  const long q = hq(S, i, j, k);
  for (int m = 0; m < S.ns; m++) S.din[c + (long)m * S.ni] = S.di[q + (long)m * S.nh];
}

// ===========================================================================
// Hybrid shock sensor (comput_Rhybrid, OCFD_SetHybrid.f90:17-143) and the
// per-interface scheme masks (set_hybrid_scheme, :2-14; OCFD_inviscous.f90:63-68).
// ===========================================================================
__global__ __launch_bounds__(256) void k_sensor1(DevState S) {
  const long c = cell_index();
  if (c >= S.ni) return;
  const int i = (int)(c % S.nx) + 1;
  const int j = (int)((c / S.nx) % S.ny) + 1;
  const int k = (int)(c / ((long)S.nx * S.ny)) + 1;
  const long q = hq(S, i, j, k);
  const double* P = S.p;
  const double p0 = P[q];
  const double xm = P[hq(S, wrap(i - 1, S.nx, S.per[0]), j, k)], xp = P[hq(S, wrap(i + 1, S.nx, S.per[0]), j, k)];
  const double ym = P[hq(S, i, wrap(j - 1, S.ny, S.per[1]), k)], yp = P[hq(S, i, wrap(j + 1, S.ny, S.per[1]), k)];
  const double zm = P[hq(S, i, j, wrap(k - 1, S.nz, S.per[2]))], zp = P[hq(S, i, j, wrap(k + 1, S.nz, S.per[2]))];
  S.Rh[q] = fabs((xm - 2.0 * p0 + xp) / (xm + 2.0 * p0 + xp)) + fabs((ym - 2.0 * p0 + yp) / (ym + 2.0 * p0 + yp)) +
            fabs((zm - 2.0 * p0 + zp) / (zm + 2.0 * p0 + zp));
}

template <bool MT>
__global__ __launch_bounds__(256) void k_sensor2(DevState S) {
  const long c = cell_index();
  if (c >= S.ni) return;
  const int i = (int)(c % S.nx) + 1;
  const int j = (int)((c / S.nx) % S.ny) + 1;
  const int k = (int)(c / ((long)S.nx * S.ny)) + 1;
  const long q = hq(S, i, j, k);
  auto dh = [&](int dir, int s) {
    return deriv1(S, dir, s, [&](int o) {
      const int t = wrap(s + o, S.n[dir], S.per[dir]);
      return S.p[q + (long)(t - s) * S.sh[dir]];
    });
  };
  const double pk = dh(0, i), pi = dh(1, j), ps = dh(2, k);
  const Metrics M = metrics_at<MT>(S, q, i, j, k);
  const double px = pk * M.kx + pi * M.ix + ps * M.sx;
  const double py = pk * M.ky + pi * M.iy + ps * M.sy;
  const double pz = pk * M.kz + pi * M.iz + ps * M.sz;
  S.Rh[q] = sqrt(px * px + py * py + pz * pz);
}

// Ordered mean: one lane adds in the reference's (k,j,i) order, so the
// result is bit-identical to the single-rank MPI_ALLREDUCE of :64.  The
// other 63 lanes of the (single) wave fetch the next chunk of SUMC values
// while lane 0 adds the current one from LDS, so the add chain, not the
// load latency, sets the time.
constexpr int SUMC = 1024;
__device__ void k_sum_ordered_body(DevState S) {
  __shared__ double buf[SUMC];
  if (blockIdx.x != 0) return;
  const int lane = threadIdx.x;
  const long N = S.ni;
  double r[SUMC / 64];
  auto fetch = [&](long base) {
#pragma unroll
    for (int a = 0; a < SUMC / 64; a++) {
      const long c = base + a * 64 + lane;
      if (c < N) {
        const int i = (int)(c % S.nx) + 1;
        const int j = (int)((c / S.nx) % S.ny) + 1;
        const int k = (int)(c / ((long)S.nx * S.ny)) + 1;
        r[a] = S.Rh[hq(S, i, j, k)];
      } else {
        r[a] = 0.0;
      }
    }
  };
  double dp0 = 0.0;
  fetch(0);
  for (long base = 0; base < N; base += SUMC) {
#pragma unroll
    for (int a = 0; a < SUMC / 64; a++) buf[a * 64 + lane] = r[a];
    __syncthreads();
    if (base + SUMC < N) fetch(base + SUMC);
    if (lane == 0) {
      const int m = (int)min((long)SUMC, N - base);
      int a = 0;
      for (; a + 8 <= m; a += 8) {
        double t[8];
#pragma unroll
        for (int b = 0; b < 8; b++) t[b] = buf[a + b];
#pragma unroll
        for (int b = 0; b < 8; b++) dp0 = dp0 + t[b];
      }
      for (; a < m; a++) dp0 = dp0 + buf[a];
    }
    __syncthreads();
  }
  if (lane == 0) {
    const double dp_av = dp0 / (1.0 * S.nx * S.ny * S.nz);
    S.red[0] = 1.0 / (dp_av + 1.0e-30);
    S.red[1] = dp_av;
  }
}
__global__ __launch_bounds__(64) void k_sum_ordered(DevState S) { k_sum_ordered_body(S); }

// Fast (non-ordered) mean of the type-2 sensor: per-thread strided sums, a
// wavefront tree by __shfl_xor, the four wave sums in fixed order; then one
// wave over the partials.  Deterministic for a given grid.
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
  return v;
}
__global__ __launch_bounds__(256) void k_sum_partial(DevState S, double* part) {
  __shared__ double sm[4];
  double acc = 0.0;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < S.ni; c += (long)gridDim.x * blockDim.x) {
    const int i = (int)(c % S.nx) + 1;
    const int j = (int)((c / S.nx) % S.ny) + 1;
    const int k = (int)(c / ((long)S.nx * S.ny)) + 1;
    acc += S.Rh[hq(S, i, j, k)];
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((sm[0] + sm[1]) + sm[2]) + sm[3];
}
__global__ __launch_bounds__(64) void k_sum_final(DevState S, const double* part, int nparts) {
  double acc = 0.0;
  for (int a = threadIdx.x; a < nparts; a += 64) acc += part[a];
  acc = wave_sum(acc);
  if (threadIdx.x == 0) {
    const double dp_av = acc / (1.0 * S.nx * S.ny * S.nz);
    S.red[0] = 1.0 / (dp_av + 1.0e-30);
    S.red[1] = dp_av;
    // certificate counters of k_masks<.., 2>: uncertain interfaces, and
    // the smallest relative distance of a normalised R to a threshold
    reinterpret_cast<unsigned long long*>(S.red)[4] = 0ull;
    reinterpret_cast<unsigned long long*>(S.red)[5] = 0x7ff0000000000000ull;  // +inf
  }
}

// The sequential mean again, only when the certificate of the fast one
// failed (k_masks counted an interface inside the rounding band); otherwise
// the lane exits at once.  Launched on one wave so it costs nothing when idle.
__global__ __launch_bounds__(64) void k_sum_ordered_if(DevState S) {
  if (reinterpret_cast<const unsigned long long*>(S.red)[4] == 0ull) return;
  k_sum_ordered_body(S);
}

// CFL time step (no reference counterpart; the reference takes dt from the
// namelist): lambda = sum over xi, eta, zeta of (|A1 u + A2 v + A3 w| +
// c |(A1, A2, A3)|) / h with c = sqrt(gamma p / d), gamma = p/Et + 1
// (comput_cc, Comput_flow_variables.f90:50-63); dt = CFL / max lambda.  The
// maximum is order independent, so the result does not depend on the grid.
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__global__ __launch_bounds__(256) void k_cfl_partial(DevState S, double* part) {
  __shared__ double sm[4];
  double acc = 0.0;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < S.ni; c += (long)gridDim.x * blockDim.x) {
    const int i = (int)(c % S.nx) + 1;
    const int j = (int)((c / S.nx) % S.ny) + 1;
    const int k = (int)(c / ((long)S.nx * S.ny)) + 1;
    const long q = hq(S, i, j, k);
    const double uu = S.u[q], vv = S.v[q], ww = S.w[q];
    const double cc = sqrt((S.p[q] / S.Et[q] + 1.0) * S.p[q] / S.d[q]);
    const double a[3][3] = {{S.Akx[q], S.Aky[q], S.Akz[q]}, {S.Aix[q], S.Aiy[q], S.Aiz[q]}, {S.Asx[q], S.Asy[q], S.Asz[q]}};
    double lam = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double un = a[d][0] * uu + a[d][1] * vv + a[d][2] * ww;
      const double g = sqrt(a[d][0] * a[d][0] + a[d][1] * a[d][1] + a[d][2] * a[d][2]);
      lam = lam + (fabs(un) + cc * g) * S.h_1[d];
    }
    acc = fmax(acc, lam);
  }
  acc = wave_max(acc);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
}
__global__ __launch_bounds__(64) void k_cfl_final(const double* part, int nparts, double cfl, double* out) {
  double acc = 0.0;
  for (int a = threadIdx.x; a < nparts; a += 64) acc = fmax(acc, part[a]);
  acc = wave_max(acc);
  if (threadIdx.x == 0) {
    out[0] = cfl / acc;
    out[1] = acc;
  }
}

__global__ __launch_bounds__(256) void k_sensor_norm(DevState S) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ni) return;
  const int i = (int)(c % S.nx) + 1;
  const int j = (int)((c / S.nx) % S.ny) + 1;
  const int k = (int)(c / ((long)S.nx * S.ny)) + 1;
  const long q = hq(S, i, j, k);
  S.Rh[q] = S.Rh[q] * S.red[0];
}

// Interface masks from the sensor (set_hybrid_scheme, OCFD_SetHybrid.f90:2-14):
// kind 1/2/3 as max(R_i, R_i+1) <= t1, <= t2, above.  With a normalised
// sensor (type 2) Rh still holds the raw |grad p| here and the mask uses
// fl(max(r0, r1) * inv), which equals max(fl(r0 * inv), fl(r1 * inv)) (rounding
// is monotone), i.e. the mask of the normalised field; k_sensor_norm
// normalises Rh afterwards.  SCALE: 0 type 1 (R as stored), 1 normalised,
// 2 normalised and certified: an interface whose normalised R lies within the
// relative band hyb_eta of a threshold could flip between the fast and the
// ordered mean; such interfaces are counted (red[4]) and the smallest
// relative distance to a threshold is kept (red[5]) for the report.
// One kernel builds the masks of all three directions: thread (i, j, k), 0 <=
// i <= nx, 0 <= j <= ny, 0 <= k <= nz, owns interface s = i of the x-pencil
// (j, k), s = j of the y-pencil (i, k) and s = k of the z-pencil (i, j), each
// where that pencil exists (its other two indices >= 1).  Workgroups take
// 64 x 4 (i, j) tiles of one k plane from a fixed grid, so the index
// arithmetic is 32-bit and per tile; Rh(i, j, k) is read once for the three
// interfaces.  REDO: the masks again with the ordered mean, only when the
// certified launch counted an uncertain interface (otherwise every workgroup
// reads one counter and exits).
constexpr int MASK_BLOCKS = 2048, MK_TX = 64, MK_TY = 4;
template <int SCALE>
__device__ __forceinline__ uint8_t mask_of(const DevState& S, double r0, double r1, double inv, long& nunsure,
                                           double& dist) {
  const double R0 = SCALE == 0 ? fmax(r0, r1) : fmax(r0, r1) * inv;
  if (SCALE == 2) {
    const double e = S.hyb_eta;
    const bool unsure = (R0 >= S.hyb_t1 * (1.0 - e) && R0 <= S.hyb_t1 * (1.0 + e)) ||
                        (R0 >= S.hyb_t2 * (1.0 - e) && R0 <= S.hyb_t2 * (1.0 + e));
    nunsure += unsure ? 1 : 0;
    dist = fmin(dist, fmin(fabs(R0 * S.hyb_inv_t1 - 1.0), fabs(R0 * S.hyb_inv_t2 - 1.0)));
  }
  return R0 <= S.hyb_t1 ? 1 : (R0 <= S.hyb_t2 ? 2 : 3);
}

template <int SCALE, bool REDO>
__global__ __launch_bounds__(256) void k_masks(DevState S) {
  if (REDO && reinterpret_cast<const unsigned long long*>(S.red)[4] == 0ull) return;
  const int nx = S.nx, ny = S.ny, nz = S.nz;
  const int nti = (nx + MK_TX) / MK_TX, ntj = (ny + MK_TY) / MK_TY;  // tiles over 0..nx, 0..ny
  const int ntiles = nti * ntj * (nz + 1);
  const double inv = SCALE == 0 ? 1.0 : S.red[0];
  long nunsure = 0;
  double dist = __builtin_inf();
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int ti = tile % nti, r = tile / nti;
    const int tj = r % ntj, k = r / ntj;
    const int i = ti * MK_TX + (int)(threadIdx.x % MK_TX);
    const int j = tj * MK_TY + (int)(threadIdx.x / MK_TX);
    if (i > nx || j > ny) continue;
    // the exchanged Rhybrid: periodic faces wrap; non-periodic ghosts keep 0
    const int iw = wrap(i, nx, S.per[0]), jw = wrap(j, ny, S.per[1]), kw = wrap(k, nz, S.per[2]);
    const long q = hq(S, iw, jw, kw);
    const double r0 = S.Rh[q];
    if (j >= 1 && k >= 1) {
      const double r1 = S.Rh[q + (long)(wrap(i + 1, nx, S.per[0]) - iw)];
      S.mask[0][i + (long)(nx + 1) * ((j - 1) + (long)ny * (k - 1))] = mask_of<SCALE>(S, r0, r1, inv, nunsure, dist);
    }
    if (i >= 1 && k >= 1) {
      const double r1 = S.Rh[q + (long)(wrap(j + 1, ny, S.per[1]) - jw) * S.sh[1]];
      S.mask[1][(i - 1) + (long)nx * (j + (long)(ny + 1) * (k - 1))] = mask_of<SCALE>(S, r0, r1, inv, nunsure, dist);
    }
    if (i >= 1 && j >= 1) {
      const double r1 = S.Rh[q + (long)(wrap(k + 1, nz, S.per[2]) - kw) * S.sh[2]];
      S.mask[2][(i - 1) + (long)nx * ((j - 1) + (long)ny * k)] = mask_of<SCALE>(S, r0, r1, inv, nunsure, dist);
    }
  }
  if (SCALE == 2) {
    // certificate counters, reduced in the workgroup first so that the
    // counters see one atomic per workgroup
    __shared__ unsigned long long sU[4];
    __shared__ double sD[4];
    for (int o = 32; o > 0; o >>= 1) {
      dist = fmin(dist, __shfl_xor(dist, o, 64));
      nunsure += __shfl_xor(nunsure, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      sU[threadIdx.x >> 6] = (unsigned long long)nunsure;
      sD[threadIdx.x >> 6] = dist;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const unsigned long long u = sU[0] + sU[1] + sU[2] + sU[3];
      const double d = fmin(fmin(sD[0], sD[1]), fmin(sD[2], sD[3]));
      unsigned long long* cnt = reinterpret_cast<unsigned long long*>(S.red);
      if (u) atomicAdd(cnt + 4, u);
      // non-negative doubles order as their bit patterns
      if (d < __builtin_inf()) atomicMin(cnt + 5, (unsigned long long)__double_as_longlong(d));
    }
  }
}

// ===========================================================================
// Mesh metrics (init_mesh / comput_Jacobian3d, OCFD_mesh.f90:1-107)
// ===========================================================================
// Periodic face fill of a halo array with an additive span on the ghosts
// (exchange_boundary_* + updata_Axyz_periodic, OCFD_mesh.f90:109-185).
template <int DIR>
__global__ void k_fill_faces(DevState S, double* A, int shift, double span) {
  const int n = S.n[DIR];
  const long npen = DIR == 0 ? (long)S.ny * S.nz : DIR == 1 ? (long)S.nx * S.nz : (long)S.nx * S.ny;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npen * 2 * OCFD_LAP) return;
  const long pen = t / (2 * OCFD_LAP);
  const int g = (int)(t % (2 * OCFD_LAP));
  int a0, a1;
  pencil_ijk<DIR>(S, pen, a0, a1);
  const int l = g % OCFD_LAP + 1;
  auto at = [&](int s) -> long {
    if (DIR == 0) return hq(S, s, a0, a1);
    if (DIR == 1) return hq(S, a0, s, a1);
    return hq(S, a0, a1, s);
  };
  // exchange = exact copy (keeps -0.0); coordinates then get the span shift
  if (g < OCFD_LAP) A[at(1 - l)] = shift ? A[at(n + 1 - l)] - span : A[at(n + 1 - l)];
  else A[at(n + l)] = shift ? A[at(l)] + span : A[at(l)];
}

__global__ __launch_bounds__(256) void k_metrics(DevState S) {
  const long c = cell_index();
  if (c >= S.ni) return;
  const int i = (int)(c % S.nx) + 1;
  const int j = (int)((c / S.nx) % S.ny) + 1;
  const int k = (int)(c / ((long)S.nx * S.ny)) + 1;
  const long q = hq(S, i, j, k);
  // coordinates already carry their (shifted) ghosts: read them directly
  auto dh = [&](const double* F, int dir, int s) {
    return deriv1(S, dir, s, [&](int o) { return F[q + (long)o * S.sh[dir]]; });
  };
  const double xi1 = dh(S.Axx, 0, i), yi1 = dh(S.Ayy, 0, i), zi1 = dh(S.Azz, 0, i);
  const double xj1 = dh(S.Axx, 1, j), yj1 = dh(S.Ayy, 1, j), zj1 = dh(S.Azz, 1, j);
  const double xk1 = dh(S.Axx, 2, k), yk1 = dh(S.Ayy, 2, k), zk1 = dh(S.Azz, 2, k);
  const Metrics m = metrics_from_derivs(xi1, yi1, zi1, xj1, yj1, zj1, xk1, yk1, zk1);
  S.Ajac[q] = m.jac;
  S.Akx[q] = m.kx;
  S.Aky[q] = m.ky;
  S.Akz[q] = m.kz;
  S.Aix[q] = m.ix;
  S.Aiy[q] = m.iy;
  S.Aiz[q] = m.iz;
  S.Asx[q] = m.sx;
  S.Asy[q] = m.sy;
  S.Asz[q] = m.sz;
}

// Metric tables: the coordinate derivatives of k_metrics taken once per row
// (diagonal: along the row's own direction; off-diagonal: on the face plane
// they depend on).  Thread t of a flat launch writes one table entry.
__global__ __launch_bounds__(256) void k_mettab_build(DevState S, double* xi, double* yj, double* zk, double* xj,
                                                      double* yi, double* xk, double* zi, double* yk, double* zj) {
  const long nxy = (long)S.nx * S.ny, nxz = (long)S.nx * S.nz, nyz = (long)S.ny * S.nz;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  auto d = [&](const double* F, int dir, int i, int j, int k) {
    const long q = hq(S, i, j, k);
    const int s = dir == 0 ? i : dir == 1 ? j : k;
    return deriv1(S, dir, s, [&](int o) { return F[q + (long)o * S.sh[dir]]; });
  };
  if (t < S.nx) { xi[t] = d(S.Axx, 0, (int)t + 1, 1, 1); return; }
  t -= S.nx;
  if (t < S.ny) { yj[t] = d(S.Ayy, 1, 1, (int)t + 1, 1); return; }
  t -= S.ny;
  if (t < S.nz) { zk[t] = d(S.Azz, 2, 1, 1, (int)t + 1); return; }
  t -= S.nz;
  if (t < nxy) {
    const int i = (int)(t % S.nx) + 1, j = (int)(t / S.nx) + 1;
    xj[t] = d(S.Axx, 1, i, j, 1);
    yi[t] = d(S.Ayy, 0, i, j, 1);
    return;
  }
  t -= nxy;
  if (t < nxz) {
    const int i = (int)(t % S.nx) + 1, k = (int)(t / S.nx) + 1;
    xk[t] = d(S.Axx, 2, i, 1, k);
    zi[t] = d(S.Azz, 0, i, 1, k);
    return;
  }
  t -= nxz;
  if (t < nyz) {
    const int j = (int)(t % S.ny) + 1, k = (int)(t / S.ny) + 1;
    yk[t] = d(S.Ayy, 2, 1, j, k);
    zj[t] = d(S.Azz, 1, 1, j, k);
  }
}

// Counts the interior cells where a table-evaluated metric (the ten of
// comput_Jacobian3d and the nine A*1) differs in any bit from the stored one.
__device__ __forceinline__ bool same_bits(double a, double b) { return __double_as_longlong(a) == __double_as_longlong(b); }
__global__ __launch_bounds__(256) void k_mettab_check(DevState S, unsigned long long* bad) {
  const long c = cell_index();
  bool ok = true;
  if (c < S.ni) {
    const int i = (int)(c % S.nx) + 1;
    const int j = (int)((c / S.nx) % S.ny) + 1;
    const int k = (int)(c / ((long)S.nx * S.ny)) + 1;
    const long q = hq(S, i, j, k);
    const Metrics m = metrics_table(S, i, j, k);
    ok = same_bits(m.kx, S.Akx[q]) && same_bits(m.ky, S.Aky[q]) && same_bits(m.kz, S.Akz[q]) &&
         same_bits(m.ix, S.Aix[q]) && same_bits(m.iy, S.Aiy[q]) && same_bits(m.iz, S.Aiz[q]) &&
         same_bits(m.sx, S.Asx[q]) && same_bits(m.sy, S.Asy[q]) && same_bits(m.sz, S.Asz[q]) &&
         same_bits(m.jac, S.Ajac[q]) && same_bits(m.kx / m.jac, S.Akx1[q]) && same_bits(m.ky / m.jac, S.Aky1[q]) &&
         same_bits(m.kz / m.jac, S.Akz1[q]) && same_bits(m.ix / m.jac, S.Aix1[q]) &&
         same_bits(m.iy / m.jac, S.Aiy1[q]) && same_bits(m.iz / m.jac, S.Aiz1[q]) &&
         same_bits(m.sx / m.jac, S.Asx1[q]) && same_bits(m.sy / m.jac, S.Asy1[q]) &&
         same_bits(m.sz / m.jac, S.Asz1[q]) && same_bits(jac_at<true>(S, q, i, j, k), S.Ajac[q]);
  }
  const unsigned long long nbad = __popcll(__ballot(!ok));
  if ((threadIdx.x & 63) == 0 && nbad) atomicAdd(bad, nbad);
}

__global__ __launch_bounds__(256) void k_metrics_over_jac(DevState S) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= S.nh) return;
  const double J = S.Ajac[q];
  S.Akx1[q] = S.Akx[q] / J;
  S.Aky1[q] = S.Aky[q] / J;
  S.Akz1[q] = S.Akz[q] / J;
  S.Aix1[q] = S.Aix[q] / J;
  S.Aiy1[q] = S.Aiy[q] / J;
  S.Aiz1[q] = S.Aiz[q] / J;
  S.Asx1[q] = S.Asx[q] / J;
  S.Asy1[q] = S.Asy[q] / J;
  S.Asz1[q] = S.Asz[q] / J;
}

__global__ __launch_bounds__(256) void k_copy_interior(DevState S, const double* src_halo, double* dst_halo) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ni) return;
  const int i = (int)(c % S.nx) + 1;
  const int j = (int)((c / S.nx) % S.ny) + 1;
  const int k = (int)(c / ((long)S.nx * S.ny)) + 1;
  const long q = hq(S, i, j, k);
  dst_halo[q] = src_halo[q];
}

// ===========================================================================
// filtering (OCFD_filtering.f90:4-362): the Bogey & Bailly 9-point filter and
// the shock-capturing filter along one direction over the index box [b, e]
// (already clipped to the faces the stencil may reach), reading a copy ft of
// f taken before the direction's pass, writing f.  One thread per cell of the
// box, all five conserved variables.
// ===========================================================================
struct FiltBox {
  int b[3], e[3];
  double s0, rth;
};

__global__ __launch_bounds__(256) void k_copy5(const double* __restrict__ src, double* __restrict__ dst, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[t];
}

template <int DIR>
__device__ __forceinline__ bool filt_cell(const DevState& S, const FiltBox& B, int& i, int& j, int& k) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ni = B.e[0] - B.b[0] + 1, nj = B.e[1] - B.b[1] + 1, nk = B.e[2] - B.b[2] + 1;
  if (t >= ni * nj * nk) return false;
  i = B.b[0] + (int)(t % ni);
  j = B.b[1] + (int)((t / ni) % nj);
  k = B.b[2] + (int)(t / (ni * nj));
  return true;
}

// filter_x3d / filter_y3d / filter_z3d (:78-190)
template <int DIR>
__global__ __launch_bounds__(256) void k_filter9(DevState S, FiltBox B, const double* __restrict__ ft) {
  int i, j, k;
  if (!filt_cell<DIR>(S, B, i, j, k)) return;
  const double d0 = 0.243527493120, d1 = -0.204788880640, d2 = 0.120007591680, d3 = -0.045211119360,
               d4 = 0.008228661760;
  const int s = DIR == 0 ? i : DIR == 1 ? j : k;
  const int n = S.n[DIR], per = S.per[DIR];
  const long c = iq(S, i, j, k);
  auto at = [&](const double* F, int o) { return F[c + (long)(wrap(s + o, n, per) - s) * S.si[DIR]]; };
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const double* F = ft + (long)m * S.ni;
    S.f[c + (long)m * S.ni] =
        at(F, 0) - B.s0 * (d0 * at(F, 0) + d1 * (at(F, -1) + at(F, 1)) + d2 * (at(F, -2) + at(F, 2)) +
                           d3 * (at(F, -3) + at(F, 3)) + d4 * (at(F, -4) + at(F, 4)));
  }
}

// sc0 of filter_x3d_shock (:226-232) at position s of the cell's pencil
template <int DIR>
__device__ __forceinline__ double filt_sc0(const DevState& S, long q0, int s0pos, int s, double rth) {
  const int n = S.n[DIR], per = S.per[DIR];
  auto P = [&](int o) { return S.p[q0 + (long)(wrap(s + o, n, per) - s0pos) * S.sh[DIR]]; };
  const double dp0 = 0.25 * (-P(1) + 2.0 * P(0) - P(-1));
  const double dp1 = 0.25 * (-P(2) + 2.0 * P(1) - P(0));
  const double dp2 = 0.25 * (-P(0) + 2.0 * P(-1) - P(-2));
  const double ri = 0.5 * (sqr(dp0 - dp1) + sqr(dp0 - dp2)) / (P(0) * P(0)) + 1.0e-16;
  return 0.5 * (1.0 - rth / ri + fabs(1.0 - rth / ri));
}

// filter_x3d_shock / filter_y3d_shock / filter_z3d_shock (:196-360), after
// comput_flow_variables has renewed p
template <int DIR>
__global__ __launch_bounds__(256) void k_filter_shock(DevState S, FiltBox B, const double* __restrict__ ft) {
  int i, j, k;
  if (!filt_cell<DIR>(S, B, i, j, k)) return;
  const double c1 = -0.210383, c2 = 0.039617;
  const int s = DIR == 0 ? i : DIR == 1 ? j : k;
  const int n = S.n[DIR], per = S.per[DIR];
  const long q = hq(S, i, j, k), c = iq(S, i, j, k);
  const double sm = filt_sc0<DIR>(S, q, s, s - 1, B.rth);
  const double sc = filt_sc0<DIR>(S, q, s, s, B.rth);
  const double sp = filt_sc0<DIR>(S, q, s, s + 1, B.rth);
  const double sc1 = 0.5 * (sc + sp);
  const double sc2 = 0.5 * (sc + sm);
  auto at = [&](const double* F, int o) { return F[c + (long)(wrap(s + o, n, per) - s) * S.si[DIR]]; };
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const double* F = ft + (long)m * S.ni;
    S.f[c + (long)m * S.ni] =
        at(F, 0) - B.s0 * (sc1 * (c1 * (at(F, 1) - at(F, 0)) + c2 * (at(F, 2) - at(F, -1))) -
                           sc2 * (c1 * (at(F, 0) - at(F, -1)) + c2 * (at(F, 1) - at(F, -2))));
  }
}

__global__ void k_zero_u64(unsigned long long* a) {
  if (threadIdx.x == 0) *a = 0ull;
}

__global__ void k_fill(double* a, long n, double v) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

// ===========================================================================
// Host side
// ===========================================================================
namespace {

struct KStat {
  double ms = 0.0;
  long long launches = 0;
};

struct Ctx {
  bool inited = false;
  ocfd_cfg cfg{};
  DevState S{};
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  double* part = nullptr;  // fast-sum partials
  double* diag = nullptr;  // diagnostics: block partials [ndiag][DIAG_SLOTS], then the DIAG_SLOTS results
  int ndiag = 512;
  int nparts = 0;
  int hybrid_sum_mode = 1;
  double hyb_eta0 = 0.0;      // certificate band before OCFD_OPT_HYBRID_CERT_WIDEN
  bool hyb_cert_ran = false;
  int sweep_variant[3] = {0, 0, 0};
  float tune_ms[3][8] = {};
  int col_planes = -1;    // k planes per block of the marching kernels (0: one thread per cell, -1: auto)
  unsigned short* kc[3] = {nullptr, nullptr, nullptr};  // sweep kind lists per direction
  int* kct[3] = {nullptr, nullptr, nullptr};            // their per-tile counts
  int kc_var[3] = {-1, -1, -1};  // tile variant the stored lists belong to (-1: none / masks rebuilt)
  hipStream_t side = nullptr;    // the viscous-flux preparation runs here beside the sweeps
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_qp = nullptr;
  int concurrent = 1;
  int use_graph = 0;
  int metric_tables = 0;     // OCFD_OPT_METRIC_TABLES (off: the divisions cost more than the loads)
  double* filt_tmp = nullptr;  // ocfd_filtering's copy of f
  long bl_nin = 0, bl_nup = 0;  // lengths of the boundary-layer inlet and upper data
  bool mt_valid = false;     // the tables reproduce every interior metric
  unsigned long long* mt_bad = nullptr;
  int timing = 0;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  double graph_dt = 0.0;
  int* h_negt = nullptr;  // pinned copy of the negative-Et counter
  FlowGas hg{};           // host copy of the species constant bank
  double *chem_rho = nullptr, *chem_E = nullptr, *chem_dt = nullptr;  // chemistry batch buffers
  int chem_method = OCFD_CHEM_FIXED;
  double chem_rtol_y = 0.05, chem_y_floor = 1e-6, chem_rtol_t = 0.01;
  int chem_max_sub = 4000, chem_scheduler = 1;
  std::vector<std::string> knames;
  std::map<std::string, int> kindex;
  std::vector<KStat> kstats;
  struct Pending { int id; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> evpool;
  std::string err;
};
Ctx G;

int fail(const std::string& m) {
  G.err = m;
  return 1;
}
#define HIPCHK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) return fail(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
  } while (0)

double* dalloc(long n) {
  void* p = nullptr;
  if (hipMalloc(&p, sizeof(double) * (size_t)(n > 0 ? n : 1)) != hipSuccess) return nullptr;
  G.allocs.push_back(p);
  return (double*)p;
}

hipEvent_t get_event() {
  if (!G.evpool.empty()) {
    hipEvent_t e = G.evpool.back();
    G.evpool.pop_back();
    return e;
  }
  hipEvent_t e;
  hipEventCreate(&e);
  return e;
}

void collect_timing() {
  if (G.pending.empty()) return;
  hipStreamSynchronize(G.stream);
  if (G.side) hipStreamSynchronize(G.side);
  for (auto& p : G.pending) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, p.a, p.b);
    G.kstats[p.id].ms += ms;
    G.kstats[p.id].launches += 1;
    G.evpool.push_back(p.a);
    G.evpool.push_back(p.b);
  }
  G.pending.clear();
}

int kid(const char* name) {
  auto it = G.kindex.find(name);
  if (it != G.kindex.end()) return it->second;
  int id = (int)G.knames.size();
  G.knames.push_back(name);
  G.kstats.push_back(KStat{});
  G.kindex[name] = id;
  return id;
}

template <typename K, typename... A>
void launch_s(hipStream_t st, const char* name, K kern, dim3 g, dim3 b, A... args) {
  if (G.timing) {
    hipEvent_t a = get_event(), e = get_event();
    hipEventRecord(a, st);
    hipLaunchKernelGGL(kern, g, b, 0, st, args...);
    hipEventRecord(e, st);
    G.pending.push_back({kid(name), a, e});
    if (G.pending.size() > 4096) collect_timing();
  } else {
    hipLaunchKernelGGL(kern, g, b, 0, st, args...);
  }
}
template <typename K, typename... A>
void launch(const char* name, K kern, dim3 g, dim3 b, A... args) {
  launch_s(G.stream, name, kern, g, b, args...);
}

inline dim3 grid1(long n) { return dim3((unsigned)((n + 255) / 256)); }

// ghost points of face f (LAP layers over the interior tangential range)
inline long ghost_count(int f) {
  const DevState& S = G.S;
  const int dir = f >> 1;
  return (long)OCFD_LAP * (dir == 0 ? S.ny : S.nx) * (dir == 2 ? S.ny : S.nz);
}

// Planes per block of the marching kernels.  Automatic choice: 4 for N_SPEC
// = 1; one thread per cell for N_SPEC > 1, whose per-cell bodies (species
// loops, Rescaling_di, Newton temperature) leave no registers for the z
// windows (measured on a 256x192x128 jet: 1.26 vs 1.09 ms per k_rk launch).
inline int col_planes() { return G.col_planes >= 0 ? G.col_planes : (G.S.ns > 1 ? 0 : 4); }

// k_rk with the viscous scheme and (for the stepping path's flag sets) the
// flags fixed at compile time, flat or marching
template <int CD, int F>
void launch_rk_f(int krk, double dt, int flags) {
  const DevState& S = G.S;
  const int kc = col_planes();
  if (S.mt_on) {
    if (kc > 0) launch("k_rk", k_rk_col<CD, F, true>, col_grid(S, kc), dim3(256), S, kc, krk, dt, flags);
    else launch("k_rk", k_rk<CD, F, true>, grid1(S.ni), dim3(256), S, krk, dt, flags);
  } else {
    if (kc > 0) launch("k_rk", k_rk_col<CD, F, false>, col_grid(S, kc), dim3(256), S, kc, krk, dt, flags);
    else launch("k_rk", k_rk<CD, F, false>, grid1(S.ni), dim3(256), S, krk, dt, flags);
  }
}

template <int CD>
void launch_rk_cd(int krk, double dt, int flags) {
  constexpr int STEP1 = RK_VISC | RK_UPDATE_F | RK_PRIMS | RK_DI_FROM_D;
  constexpr int STEPMS = RK_VISC | RK_UPDATE_F | RK_PRIMS | RK_MS;
  switch (flags) {
    case STEP1: launch_rk_f<CD, STEP1>(krk, dt, flags); break;
    case STEP1 & ~RK_VISC: launch_rk_f<CD, (STEP1 & ~RK_VISC)>(krk, dt, flags); break;
    case STEPMS: launch_rk_f<CD, STEPMS>(krk, dt, flags); break;
    case STEPMS & ~RK_VISC: launch_rk_f<CD, (STEPMS & ~RK_VISC)>(krk, dt, flags); break;
    default: launch_rk_f<CD, -1>(krk, dt, flags); break;
  }
}

void launch_rk(int krk, double dt, int flags) {
  if (G.S.cd8) launch_rk_cd<1>(krk, dt, flags);
  else launch_rk_cd<0>(krk, dt, flags);
}

void launch_visc_prep(bool ms, hipStream_t st = nullptr) {
  const DevState& S = G.S;
  if (!st) st = G.stream;
  const int kc = col_planes();
  const bool mt = S.mt_on;
  if (kc > 0) {
    auto kern = mt ? (ms ? (S.cd8 ? k_visc_prep_col<1, true, true> : k_visc_prep_col<0, true, true>)
                         : (S.cd8 ? k_visc_prep_col<1, false, true> : k_visc_prep_col<0, false, true>))
                   : (ms ? (S.cd8 ? k_visc_prep_col<1, true, false> : k_visc_prep_col<0, true, false>)
                         : (S.cd8 ? k_visc_prep_col<1, false, false> : k_visc_prep_col<0, false, false>));
    launch_s(st, "k_visc_prep", kern, col_grid(S, kc), dim3(256), S, kc);
  } else {
    auto kern = mt ? (ms ? (S.cd8 ? k_visc_prep<1, true, true> : k_visc_prep<0, true, true>)
                         : (S.cd8 ? k_visc_prep<1, false, true> : k_visc_prep<0, false, true>))
                   : (ms ? (S.cd8 ? k_visc_prep<1, true, false> : k_visc_prep<0, true, false>)
                         : (S.cd8 ? k_visc_prep<1, false, false> : k_visc_prep<0, false, false>));
    launch_s(st, "k_visc_prep", kern, grid1(S.ni), dim3(256), S);
  }
}

// The viscous-flux preparation reads only primitives, metrics and transport
// data and writes only Ev, Amu, Amk (and reads Qp), none of which the
// inviscid sweeps touch; k_rk needs both.  It therefore runs on a second
// stream beside the sweeps (VALU-bound sweeps, memory-bound preparation):
// fork after the work it depends on, join before k_rk.  Under graph capture
// the fork and join become parallel branches.
hipStream_t fork_side() {
  if (!G.concurrent) return G.stream;
  hipEventRecord(G.ev_fork, G.stream);
  hipStreamWaitEvent(G.side, G.ev_fork, 0);
  return G.side;
}
void join_side(hipStream_t st) {
  if (st == G.stream) return;
  hipEventRecord(G.ev_join, st);
  hipStreamWaitEvent(G.stream, G.ev_join, 0);
}

template <int DIR, int V>
void launch_sweep_v() {
  const DevState& S = G.S;
  constexpr int NP = SweepTile<DIR, V>::NP, TS = SweepTile<DIR, V>::TS;
  const long npen = DIR == 0 ? (long)S.ny * S.nz : DIR == 1 ? (long)S.nx * S.nz : (long)S.nx * S.ny;
  dim3 g((unsigned)((S.n[DIR] + TS - 1) / TS), (unsigned)((npen + NP - 1) / NP));
  static const char* names[3] = {"k_sweep_x", "k_sweep_y", "k_sweep_z"};
  // the first sweep after the masks are built (or after a tile change)
  // stores the kind lists, the following ones read them
  int mode = KC_WRITE;
  if (G.kc_var[DIR] == V) mode = KC_READ;
  G.kc_var[DIR] = V;
  launch(names[DIR], k_sweep<DIR, V>, g, dim3(SweepTile<DIR, V>::NT), S, mode, G.kc[DIR], G.kct[DIR]);
}

// entries and tiles of the largest kind-list layout over the tile variants
template <int DIR, int V>
void kc_size(long& entries, long& tiles) {
  const DevState& S = G.S;
  constexpr int NP = SweepTile<DIR, V>::NP, TS = SweepTile<DIR, V>::TS;
  const long npen = DIR == 0 ? (long)S.ny * S.nz : DIR == 1 ? (long)S.nx * S.nz : (long)S.nx * S.ny;
  const long nt = (long)((S.n[DIR] + TS - 1) / TS) * ((npen + NP - 1) / NP);
  entries = std::max(entries, nt * 2L * NP * (TS + 1));
  tiles = std::max(tiles, nt);
}
template <int DIR>
void kc_sizes(long& entries, long& tiles) {
  kc_size<DIR, 0>(entries, tiles);
  kc_size<DIR, 1>(entries, tiles);
  kc_size<DIR, 2>(entries, tiles);
  kc_size<DIR, 3>(entries, tiles);
}

template <int DIR>
void launch_sweep_variant(int v) {
  switch (v) {
    case 1: launch_sweep_v<DIR, 1>(); break;
    case 2: launch_sweep_v<DIR, 2>(); break;
    case 3: launch_sweep_v<DIR, 3>(); break;
    default: launch_sweep_v<DIR, 0>(); break;
  }
}

template <int DIR>
void launch_sweep() {
  if (G.cfg.if_scheme_character) {
    const DevState& S = G.S;
    constexpr int NP = SweepTile<DIR, 0>::NP, TS = SweepTile<DIR, 0>::TS;
    const long npen = DIR == 0 ? (long)S.ny * S.nz : DIR == 1 ? (long)S.nx * S.nz : (long)S.nx * S.ny;
    dim3 g((unsigned)((S.n[DIR] + TS - 1) / TS), (unsigned)((npen + NP - 1) / NP));
    static const char* names[3] = {"k_sweep_char_x", "k_sweep_char_y", "k_sweep_char_z"};
    launch(names[DIR], k_sweep_char<DIR>, g, dim3(256), S);
    return;
  }
  launch_sweep_variant<DIR>(G.sweep_variant[DIR]);
}

template <int SCALE, bool REDO = false>
void launch_masks() {
  const DevState& S = G.S;
  const long nt = (long)((S.nx + MK_TX) / MK_TX) * ((S.ny + MK_TY) / MK_TY) * (S.nz + 1);
  launch(REDO ? "k_masks_redo" : "k_masks", k_masks<SCALE, REDO>, dim3((unsigned)std::min<long>(nt, MASK_BLOCKS)),
         dim3(256), S);
}

void enq_rhybrid() {
  const DevState& S = G.S;
  for (int d = 0; d < 3; d++) G.kc_var[d] = -1;  // new masks: the stored kind lists are stale
  const int sensor = (int)std::lround(G.cfg.hybrid_para[0]);
  if (sensor == 1) {
    launch("k_sensor1", k_sensor1, grid1(S.ni), dim3(256), S);
    launch_masks<0>();
    return;
  }
  if (S.mt_on) launch("k_sensor2", k_sensor2<true>, grid1(S.ni), dim3(256), S);
  else launch("k_sensor2", k_sensor2<false>, grid1(S.ni), dim3(256), S);
  if (G.hybrid_sum_mode == 0) {
    launch("k_sum_ordered", k_sum_ordered, dim3(1), dim3(64), S);
    launch_masks<1>();
  } else {
    // fast mean, certified: the masks are the ordered mean's in every case
    G.hyb_cert_ran = true;
    launch("k_sum_partial", k_sum_partial, dim3(G.nparts), dim3(256), S, G.part);
    launch("k_sum_final", k_sum_final, dim3(1), dim3(64), S, (const double*)G.part, G.nparts);
    launch_masks<2>();
    launch("k_sum_ordered_if", k_sum_ordered_if, dim3(1), dim3(64), S);
    launch_masks<1, true>();
  }
  launch("k_sensor_norm", k_sensor_norm, grid1(S.ni), dim3(256), S);
}

void enq_bc() {
  const DevState& S = G.S;
  if (G.cfg.ibc == OCFD_BC_BOUNDARYLAYER) {
    launch("k_bl_inlet", k_bl_inlet, grid1((long)S.ny * S.nz), dim3(256), S);
    // BcData_upper = -1: no upper data; the face keeps its values
    if (S.bl_upper >= 0) launch("k_bl_upper", k_bl_upper, grid1((long)S.nx * S.nz), dim3(256), S);
    launch("k_bl_wall", k_bl_wall, grid1((long)S.nx * S.nz), dim3(256), S);
    if (S.bl_outlet == 1 || S.bl_outlet == 2) launch("k_bl_outlet", k_bl_outlet, grid1((long)S.ny * S.nz), dim3(256), S);
  } else if (G.cfg.ibc == OCFD_BC_JET) {
    launch("k_jet_inlet", k_jet_inlet, grid1((long)S.ny * S.nz), dim3(256), S);
    launch("k_jet_outlet", k_jet_outlet, grid1((long)S.ny * S.nz), dim3(256), S);
  }
  const long nxi = S.nx > 2 ? S.nx - 2 : 0, nyi = S.ny > 2 ? S.ny - 2 : 0;
  const long nf = 2L * S.ny * S.nz + 2L * nxi * S.nz + 2L * nxi * nyi;
  launch("k_bc_faces", k_bc_faces, grid1(nf), dim3(256), S);
  for (int f = 0; f < 6; f++)  // Flow_Ghost_boundary, OCFD_BC.f90:186
    if (G.cfg.ghost_cell[f] == 1 || G.cfg.ghost_cell[f] == 2)
      launch("k_flow_ghost", k_flow_ghost, grid1(ghost_count(f)), dim3(256), S, f);
}

void enq_stage(int krk, double dt1) {
  const DevState& S = G.S;
  // the viscous preparation does not read the hybrid masks: fork before the sensor
  const hipStream_t vs = fork_side();
  if (S.viscous) launch_visc_prep(false, vs);
  else launch_s(vs, "k_transport", k_transport, grid1(S.ni), dim3(256), S);
  // the step's saved solution (fn, din) is first read by this stage's
  // k_rk, after the join: its copy runs beside the sweeps (k_save0 is
  // launched inside stage 1 on either path)
  if (krk == 1) launch_s(vs, "k_save0", k_save0, grid1(S.ni), dim3(256), S);
  if (krk == 1 && G.cfg.scheme_invis == OCFD_SCHEME_HYBRID) enq_rhybrid();
  launch_sweep<0>();
  launch_sweep<1>();
  launch_sweep<2>();
  join_side(vs);
  const int flags = (S.mass_force ? RK_MASS : 0) | (S.viscous ? RK_VISC : 0) | RK_UPDATE_F | RK_PRIMS | RK_DI_FROM_D;
  launch("k_negt_reset", k_negt_reset, dim3(1), dim3(64), S, 0);
  launch_rk(krk, dt1, flags);
  launch("k_negt_repair", k_negt_repair, dim3(1), dim3(64), S);
  enq_bc();
}

template <int DIR>
void launch_species_sweep(hipStream_t st, int m0, int m1) {
  const DevState& S = G.S;
  constexpr int NP = SpTile<DIR>::NP, TS = SpTile<DIR>::TS;
  const long npen = DIR == 0 ? (long)S.ny * S.nz : DIR == 1 ? (long)S.nx * S.nz : (long)S.nx * S.ny;
  dim3 g((unsigned)((S.n[DIR] + TS - 1) / TS), (unsigned)((npen + NP - 1) / NP));
  static const char* names[3] = {"k_species_sweep_x", "k_species_sweep_y", "k_species_sweep_z"};
  launch_s(st, names[DIR], k_species_sweep<DIR>, g, dim3(256), S, m0, m1);
}

// species right-hand sides (krk = 0: leave them in dus) or the species RK3
// stage (OCFD_Time_Adv.f90:64-71)
// on stream st (default: the library stream); after_prep, if given, is
// recorded once Qp (species viscous preparation) is written
void enq_species(int krk, double dt1, hipStream_t st = nullptr, hipEvent_t after_prep = nullptr) {
  const DevState& S = G.S;
  if (!st) st = G.stream;
  launch_species_sweep<0>(st, 0, S.ns);
  launch_species_sweep<1>(st, 0, S.ns);
  launch_species_sweep<2>(st, 0, S.ns);
  if (S.viscous) {
    launch_s(st, "k_species_conc", k_species_conc, grid1(S.ni), dim3(256), S, 0, S.ns);
    launch_s(st, "k_species_visc_prep", S.cd8 ? k_species_visc_prep<1> : k_species_visc_prep<0>, grid1(S.ni),
             dim3(256), S, 0, S.ns);
  }
  if (after_prep) hipEventRecord(after_prep, st);
  launch_s(st, "k_species_rk", S.cd8 ? k_species_rk<1> : k_species_rk<0>, grid1(S.ni), dim3(256), S, krk, dt1,
           S.viscous, 0, S.ns);
}

// One species m (0-based) of the reference's species loop
// (OCFD_Time_Adv.f90:64-71): part 0 du_scaler_invis (Dus(m) = inviscid
// right-hand side), part 1 du_scaler_viscous (Dus(m) += div of the diffusive
// flux; Qp zeroed at m = 0, accumulated after), part 2 Scalar_time_adv_RK3
// (di(m) from din(m) and Dus(m)).  In that order over every species they
// give enq_species bit for bit.
void enq_species_one(int part, int m, int krk, double dt1) {
  const DevState& S = G.S;
  const hipStream_t st = G.stream;
  if (part == 0) {
    launch_species_sweep<0>(st, m, m + 1);
    launch_species_sweep<1>(st, m, m + 1);
    launch_species_sweep<2>(st, m, m + 1);
  } else if (part == 1) {
    launch_s(st, "k_species_conc", k_species_conc, grid1(S.ni), dim3(256), S, m, m + 1);
    launch_s(st, "k_species_visc_prep", S.cd8 ? k_species_visc_prep<1> : k_species_visc_prep<0>, grid1(S.ni),
             dim3(256), S, m, m + 1);
    launch_s(st, "k_species_rk", S.cd8 ? k_species_rk<1> : k_species_rk<0>, grid1(S.ni), dim3(256), S, 0, 0.0, 1,
             m, m + 1);
  } else {
    launch_s(st, "k_species_rk", S.cd8 ? k_species_rk<1> : k_species_rk<0>, grid1(S.ni), dim3(256), S, krk, dt1, 0,
             m, m + 1);
  }
}

// one RK3 stage for N_SPEC > 1 (OCFD_Time_Adv.f90:33-91): transport, sensor,
// species convection/diffusion and update, flow fluxes and update, primitive
// recovery with Rescaling_di, negative-Et repair, boundary conditions
void enq_stage_ms(int krk, double dt1) {
  const DevState& S = G.S;
  launch("k_transport_ms", k_transport_ms, grid1(S.ni), dim3(256), S);
  if (krk == 1 && G.cfg.scheme_invis == OCFD_SCHEME_HYBRID) enq_rhybrid();
  // the species chain (sweeps, concentrations, diffusion, update of di) on
  // the second stream beside the flow sweeps, which read no di; the flow's
  // viscous preparation waits for Qp, k_rk (Rescaling_di) for the whole chain
  const hipStream_t sp = fork_side();
  // the step's saved solution: din is first read by the species update on
  // the second stream, fn by k_rk after the join
  if (krk == 1) launch_s(sp, "k_save0", k_save0, grid1(S.ni), dim3(256), S);
  enq_species(krk, dt1, sp, sp != G.stream ? G.ev_qp : nullptr);
  launch_sweep<0>();
  launch_sweep<1>();
  launch_sweep<2>();
  if (S.viscous) {
    if (sp != G.stream) hipStreamWaitEvent(G.stream, G.ev_qp, 0);
    launch_visc_prep(true);
  }
  join_side(sp);
  const int flags = (S.mass_force ? RK_MASS : 0) | (S.viscous ? RK_VISC : 0) | RK_UPDATE_F | RK_PRIMS | RK_MS;
  launch("k_negt_reset", k_negt_reset, dim3(1), dim3(64), S, 0);
  launch_rk(krk, dt1, flags);
  launch("k_negt_repair", k_negt_repair, dim3(1), dim3(64), S);
  enq_bc();
}

void enq_step(double dt1) {
  const DevState& S = G.S;
  launch("k_negt_reset", k_negt_reset, dim3(1), dim3(64), S, 1);
  for (int krk = 1; krk <= 3; krk++) {
    if (S.ns > 1) enq_stage_ms(krk, dt1);
    else enq_stage(krk, dt1);
  }
  hipMemcpyAsync(G.h_negt, S.negt + 17, sizeof(int), hipMemcpyDeviceToHost, G.stream);
}

void drop_graph() {
  if (G.gexec) hipGraphExecDestroy(G.gexec);
  if (G.graph) hipGraphDestroy(G.graph);
  G.gexec = nullptr;
  G.graph = nullptr;
}

int check_ready() {
  if (!G.inited) return fail("ocfd: library not initialised (call ocfd_init)");
  return 0;
}

double* field_ptr(int id, long* n) {
  DevState& S = G.S;
  const long nh = S.nh, ni = S.ni;
  switch (id) {
    case OCFD_FIELD_F: *n = 5 * ni; return S.f;
    case OCFD_FIELD_FN: *n = 5 * ni; return S.fn;
    case OCFD_FIELD_DU: *n = 5 * ni; return S.du;
    case OCFD_FIELD_DI: *n = nh * S.ns; return S.di;
    case OCFD_FIELD_DIN: *n = ni * S.ns; return S.din;
    case OCFD_FIELD_D: *n = nh; return S.d;
    case OCFD_FIELD_U: *n = nh; return S.u;
    case OCFD_FIELD_V: *n = nh; return S.v;
    case OCFD_FIELD_W: *n = nh; return S.w;
    case OCFD_FIELD_T: *n = nh; return S.T;
    case OCFD_FIELD_P: *n = nh; return S.p;
    case OCFD_FIELD_ET: *n = nh; return S.Et;
    case OCFD_FIELD_AMU: *n = ni; return S.Amu;
    case OCFD_FIELD_AMK: *n = ni; return S.Amk;
    case OCFD_FIELD_RHYBRID: *n = nh; return S.Rh;
    case OCFD_FIELD_AMD: *n = ni * S.ns; return S.AmD;
    case OCFD_FIELD_DUS: *n = ni * S.ns; return S.dus;
    case OCFD_FIELD_QP: *n = 3 * ni; return S.Qp;
  }
  double* M[19] = {S.Akx, S.Aky, S.Akz, S.Aix, S.Aiy, S.Aiz, S.Asx, S.Asy, S.Asz, S.Ajac,
                   S.Akx1, S.Aky1, S.Akz1, S.Aix1, S.Aiy1, S.Aiz1, S.Asx1, S.Asy1, S.Asz1};
  if (id >= OCFD_FIELD_AKX && id <= OCFD_FIELD_ASZ1) { *n = nh; return M[id - OCFD_FIELD_AKX]; }
  *n = 0;
  return nullptr;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
// One diagnostics pass: block partials, then the means into G.diag's result
// slots, copied to out[0 .. N).  Synchronises.
template <int PASS>
static int diag_pass(double d_aver, double* out) {
  const DevState& S = G.S;
  constexpr int N = DiagN<PASS>::N;
  double* res = G.diag + (long)G.ndiag * DIAG_SLOTS;
  launch("k_diag_partial", S.cd8 ? k_diag_partial<PASS, 1> : k_diag_partial<PASS, 0>, dim3(G.ndiag), dim3(256), S,
         G.diag, d_aver);
  launch("k_diag_final", k_diag_final, dim3(N), dim3(64), S, (const double*)G.diag, G.ndiag, N, res);
  HIPCHK(hipMemcpyAsync(out, res, sizeof(double) * N, hipMemcpyDeviceToHost, G.stream));
  HIPCHK(hipStreamSynchronize(G.stream));
  return 0;
}

extern "C" {

const char* ocfd_last_error(void) { return G.err.c_str(); }

int32_t ocfd_finalize(void) {
  if (!G.inited) return 0;
  hipStreamSynchronize(G.stream);
  drop_graph();
  for (auto& p : G.pending) { hipEventDestroy(p.a); hipEventDestroy(p.b); }
  G.pending.clear();
  for (auto e : G.evpool) hipEventDestroy(e);
  G.evpool.clear();
  for (void* p : G.allocs) hipFree(p);
  G.allocs.clear();
  G.chem_rho = G.chem_E = G.chem_dt = nullptr;
  G.part = nullptr;
  G.diag = nullptr;
  G.filt_tmp = nullptr;
  G.mt_bad = nullptr;
  G.mt_valid = false;
  for (int d = 0; d < 3; d++) {
    G.kc[d] = nullptr;
    G.kct[d] = nullptr;
    G.kc_var[d] = -1;
  }
  G.hg = FlowGas{};
  G.chem_method = OCFD_CHEM_FIXED;
  if (G.h_negt) hipHostFree(G.h_negt);
  G.h_negt = nullptr;
  if (G.side) hipStreamDestroy(G.side);
  if (G.ev_fork) hipEventDestroy(G.ev_fork);
  if (G.ev_join) hipEventDestroy(G.ev_join);
  if (G.ev_qp) hipEventDestroy(G.ev_qp);
  G.side = nullptr;
  G.ev_fork = G.ev_join = G.ev_qp = nullptr;
  hipStreamDestroy(G.stream);
  G.stream = nullptr;
  G.inited = false;
  return 0;
}

int32_t ocfd_init(const ocfd_cfg* cfg) {
  ocfd_finalize();
  G.err.clear();
  // every context starts from the documented option defaults
  G.hybrid_sum_mode = 1;
  G.hyb_cert_ran = false;
  G.col_planes = -1;
  G.concurrent = 1;
  G.use_graph = 0;
  G.timing = 0;
  if (!cfg) return fail("ocfd_init: null config");
  const ocfd_cfg& c = *cfg;
  if (c.lap != OCFD_LAP) return fail("ocfd_init: LAP must be 4");
  if (c.nx < 8 || c.ny < 8 || c.nz < 8) return fail("ocfd_init: each dimension needs at least 8 points");
  if (c.n_spec < 1 || c.n_spec > OCFD_NS_MAX) return fail("ocfd_init: N_SPEC must be 1..9");
  if (c.if_scheme_character != 0 && c.if_scheme_character != 1) return fail("ocfd_init: IF_Scheme_Character must be 0 or 1");
  if (c.n_spec == 1 && c.if_real_gas != 0) return fail("ocfd_init: the N_SPEC == 1 flow path is the perfect gas (IF_RealGas=0)");
  if (c.n_spec == 1 && (c.flag_vis != 0 || c.flag_cond != 0))
    return fail("ocfd_init: N_SPEC == 1 uses Sutherland/Prandtl transport");
  if (c.if_real_gas < 0 || c.if_real_gas > 1) return fail("ocfd_init: IF_RealGas must be 0 or 1");
  if (c.flag_vis < 0 || c.flag_vis > 1 || c.flag_cond < 0 || c.flag_cond > 1 || c.flag_diff < 0 || c.flag_diff > 1)
    return fail("ocfd_init: transport flags must be 0 or 1");
  if (c.scheme_vis != OCFD_SCHEME_CD6 && c.scheme_vis != OCFD_SCHEME_CD8) return fail("ocfd_init: Scheme_Vis must be CD6 or CD8");
  if (c.flux_splitting != OCFD_SPLIT_SW && c.flux_splitting != OCFD_SPLIT_LLF) return fail("ocfd_init: bad Flux_Splitting");
  switch (c.scheme_invis) {
    case OCFD_SCHEME_WENO5: case OCFD_SCHEME_WENO7: case OCFD_SCHEME_OMP6: case OCFD_SCHEME_UD7L: case OCFD_SCHEME_HYBRID: break;
    default: return fail("ocfd_init: unsupported Scheme_Invis");
  }
  for (int a = 0; a < 6; a++) {
    if (c.scheme_boundary[a] < -1 || c.scheme_boundary[a] > 1) return fail("ocfd_init: Scheme_boundary must be -1, 0 or 1");
    if (c.ghost_cell[a] < 0 || c.ghost_cell[a] > 2) return fail("ocfd_init: Ghost_Cell must be 0, 1 or 2");
    if (c.ghost_cell[a] != 0 && c.periodic[a / 2]) return fail("ocfd_init: Ghost_Cell on a periodic face");
  }
  if (c.ibc != OCFD_BC_NONE && c.ibc != OCFD_BC_BOUNDARYLAYER && c.ibc != OCFD_BC_JET)
    return fail("ocfd_init: unsupported boundary condition");
  if (c.ibc == OCFD_BC_JET && c.n_spec < 2) return fail("ocfd_init: the jet boundary condition needs N_SPEC > 1");
  if (c.scheme_invis == OCFD_SCHEME_HYBRID && std::lround(c.hybrid_para[3]) != 0)
    return fail("ocfd_init: hybrid patches (Hybrid_para(4) > 0) are not supported");
  if (c.wall_catalysis != 0) return fail("ocfd_init: only non-catalytic walls are supported");

  G.cfg = c;
  DevState& S = G.S;
  S = DevState{};
  S.nx = c.nx; S.ny = c.ny; S.nz = c.nz; S.ns = c.n_spec;
  S.nxh = c.nx + 2 * OCFD_LAP; S.nyh = c.ny + 2 * OCFD_LAP; S.nzh = c.nz + 2 * OCFD_LAP;
  S.nh = (long)S.nxh * S.nyh * S.nzh;
  S.ni = (long)S.nx * S.ny * S.nz;
  S.sh[0] = 1; S.sh[1] = S.nxh; S.sh[2] = (long)S.nxh * S.nyh;
  S.si[0] = 1; S.si[1] = S.nx; S.si[2] = (long)S.nx * S.ny;
  S.n[0] = c.nx; S.n[1] = c.ny; S.n[2] = c.nz;
  for (int d = 0; d < 3; d++) {
    S.per[d] = c.periodic[d] ? 1 : 0;
    S.bidx[d][0] = S.bidx[d][1] = c.periodic[d] ? 0 : 1;   // set_parameters, OCFD_readpara.f90:360-369
  }
  for (int a = 0; a < 6; a++) S.sb[a] = c.scheme_boundary[a];
  S.scheme = c.scheme_invis;
  S.split = c.flux_splitting;
  S.cd8 = c.scheme_vis == OCFD_SCHEME_CD8;
  S.viscous = c.if_viscous;
  S.mass_force = c.if_mass_force;
  const double hh[3] = {1.0 / (c.nx - 1.0), 1.0 / (c.ny - 1.0), 1.0 / (c.nz - 1.0)};  // OCFD_readpara.f90:351-353
  for (int d = 0; d < 3; d++) {
    const double h = hh[d];
    S.dh[d] = h;
    S.h_1[d] = 1.0 / h;
    S.cd6c[d][0] = 1.0 / (60.0 * h);
    S.cd6c[d][1] = -3.0 / (20.0 * h);
    S.cd6c[d][2] = 3.0 / (4.0 * h);
    S.cd8c[d][0] = 0.8 / h;
    S.cd8c[d][1] = -0.2 / h;
    S.cd8c[d][2] = 3.80952380952380952e-2 / h;
    S.cd8c[d][3] = -3.571428571428571428e-3 / h;
  }
  {  // set_scheme_para, OCFD_readpara.f90:424-431
    const double UD7[8] = {-3.0, 25.0, -101.0, 319.0, 214.0, -38.0, 4.0, 0.0};
    const double CD8[8] = {-3.0, 29.0, -139.0, 533.0, 533.0, -139.0, 29.0, -3.0};
    for (int a = 0; a < 8; a++) S.ud7l[a] = c.ud7l_diss * UD7[a] / 420.0 + (1.0 - c.ud7l_diss) * CD8[a] / 840.0;
  }
  S.hyb_t1 = c.hybrid_para[1];
  S.hyb_t2 = c.hybrid_para[2];
  S.Mi0 = 0.02897;
  S.Ri0 = 8.314 / S.Mi0;
  S.amu0 = c.amu0; S.pr = c.pr; S.sc = c.sc;
  for (int a = 0; a < 3; a++) S.mf[a] = c.mass_force[a];
  S.bl_inlet = (int)std::lround(c.bc_para[0]);
  S.bl_upper = (int)std::lround(c.bc_para[1]);
  S.bl_outlet = (int)std::lround(c.bc_para[3]);
  S.bl_tw = c.bc_para[4];
  S.bl_xinit = c.bc_para[5];
  S.bl_disA = c.bc_para[7];
  S.jet_iflag = c.ibc == OCFD_BC_JET ? (int)std::lround(c.bc_para[0]) : 0;
  if (c.ibc == OCFD_BC_BOUNDARYLAYER) {
    if (S.bl_inlet < 0 || S.bl_inlet > 2)
      return fail("ocfd_init: boundary layer inlet data must be 0-D, 1-D or 2-D (BC_Para(1) = 0, 1, 2)");
    if (S.bl_upper < -1 || S.bl_upper > 2)
      return fail("ocfd_init: boundary layer upper data must be none, 0-D, 1-D or 2-D (BC_Para(2) = -1, 0, 1, 2)");
    if (c.n_spec != 1) return fail("ocfd_init: the boundary-layer BC runs with N_SPEC = 1 (perfect gas)");
  }

  if (hipStreamCreateWithFlags(&G.stream, hipStreamNonBlocking) != hipSuccess) return fail("ocfd_init: no HIP device/stream");
  if (hipStreamCreateWithFlags(&G.side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&G.ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&G.ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&G.ev_qp, hipEventDisableTiming) != hipSuccess)
    return fail("ocfd_init: stream or event creation failed");
  const long nh = S.nh, ni = S.ni, ns = S.ns, nms = S.ns > 1 ? S.ns : 0;
  struct { double** p; long n; double v; } arrs[] = {
      {&S.f, 5 * ni, 0.0}, {&S.fn, 5 * ni, 0.0}, {&S.du, 5 * ni, 0.0}, {&S.di, nh * ns, 1.0}, {&S.din, ni * ns, 1.0},
      {&S.AmD, ni * nms, 0.0}, {&S.dus, ni * nms, 0.0}, {&S.Evs, 3 * ni * nms, 0.0}, {&S.Qp, 3 * ni, 0.0}, {&S.cs, ni * nms, 0.0},
      {(double**)&S.jet_in, (long)S.ny * S.nz * (5 + ns), 1.0},
      {&S.d, nh, 1.0}, {&S.u, nh, 1.0}, {&S.v, nh, 1.0}, {&S.w, nh, 1.0}, {&S.T, nh, 300.0}, {&S.p, nh, 1.0},
      {&S.Et, nh, 1.0}, {&S.Et_prev, ni, 1.0}, {&S.Amu, ni, 0.0}, {&S.Amk, ni, 0.0},
      {&S.Axx, nh, 1.0}, {&S.Ayy, nh, 1.0}, {&S.Azz, nh, 1.0},
      {&S.Akx, nh, 1.0}, {&S.Aky, nh, 1.0}, {&S.Akz, nh, 1.0}, {&S.Aix, nh, 1.0}, {&S.Aiy, nh, 1.0},
      {&S.Aiz, nh, 1.0}, {&S.Asx, nh, 1.0}, {&S.Asy, nh, 1.0}, {&S.Asz, nh, 1.0}, {&S.Ajac, nh, 1.0},
      {&S.Akx1, nh, 1.0}, {&S.Aky1, nh, 1.0}, {&S.Akz1, nh, 1.0}, {&S.Aix1, nh, 1.0}, {&S.Aiy1, nh, 1.0},
      {&S.Aiz1, nh, 1.0}, {&S.Asx1, nh, 1.0}, {&S.Asy1, nh, 1.0}, {&S.Asz1, nh, 1.0},
      {&S.Rh, nh, 0.0}, {&S.Ev, 12 * ni, 0.0}, {&S.red, 8, 0.0}};
  G.inited = true;  // so that finalize frees partial allocations on error
  for (auto& a : arrs) {
    *a.p = dalloc(a.n);
    if (!*a.p) return fail("ocfd_init: hipMalloc failed");
    if (a.n > 0) hipLaunchKernelGGL(k_fill, grid1(a.n), dim3(256), 0, G.stream, *a.p, a.n, a.v);
  }
  for (int d = 0; d < 3; d++) {
    const long nm = (long)(S.nx + (d == 0)) * (S.ny + (d == 1)) * (S.nz + (d == 2));
    void* m = nullptr;
    if (hipMalloc(&m, (size_t)nm) != hipSuccess) return fail("ocfd_init: hipMalloc failed");
    G.allocs.push_back(m);
    hipMemsetAsync(m, 0, (size_t)nm, G.stream);
    S.mask[d] = (uint8_t*)m;
  }
  {
    void* m = nullptr;
    if (hipMalloc(&m, sizeof(int) * 32) != hipSuccess) return fail("ocfd_init: hipMalloc failed");
    G.allocs.push_back(m);
    hipMemsetAsync(m, 0, sizeof(int) * 32, G.stream);
    S.negt = (int*)m;
  }
  for (int d = 0; d < 3; d++) {
    long entries = 0, tiles = 0;
    if (d == 0) kc_sizes<0>(entries, tiles);
    else if (d == 1) kc_sizes<1>(entries, tiles);
    else kc_sizes<2>(entries, tiles);
    void* a = nullptr;
    void* b = nullptr;
    if (hipMalloc(&a, sizeof(unsigned short) * (size_t)entries) != hipSuccess) return fail("ocfd_init: hipMalloc failed");
    G.allocs.push_back(a);
    if (hipMalloc(&b, sizeof(int) * (size_t)tiles) != hipSuccess) return fail("ocfd_init: hipMalloc failed");
    G.allocs.push_back(b);
    G.kc[d] = (unsigned short*)a;
    G.kct[d] = (int*)b;
    G.kc_var[d] = -1;
  }
  G.nparts = 1024;
  G.part = dalloc(G.nparts);
  G.diag = dalloc((long)G.ndiag * DIAG_SLOTS + DIAG_SLOTS);
  if (!G.diag) return fail("ocfd_init: hipMalloc failed");
  {
    // Certificate of the fast sensor mean (k_sum_partial/k_sum_final against
    // the reference's sequential sum, k_sum_ordered).  Both sums of the same
    // ni non-negative values lie within gamma_k * (exact sum) of it, with k =
    // ni - 1 for the sequential sum and k = the depth of the tree (each
    // thread's strided run, 6 wave levels, 3 wave sums, each final lane's run
    // over the partials, 6 wave levels).  Dividing by N, adding 1e-30, the
    // reciprocal, the product R * inv and the band ends t*(1 +- eta) add one
    // rounding each per side.  eta covers all of it with a factor 1.1 of room.
    const double u = std::ldexp(1.0, -53);
    const double run = std::ceil((double)S.ni / ((double)G.nparts * 256.0));
    const double depth = run + 6.0 + 3.0 + std::ceil(G.nparts / 64.0) + 6.0;
    S.hyb_eta = G.hyb_eta0 = 1.1 * ((double)S.ni - 1.0 + depth + 16.0) * u;
    S.hyb_inv_t1 = S.hyb_t1 > 0.0 ? 1.0 / S.hyb_t1 : 0.0;
    S.hyb_inv_t2 = S.hyb_t2 > 0.0 ? 1.0 / S.hyb_t2 : 0.0;
  }
  const bool blbc = c.ibc == OCFD_BC_BOUNDARYLAYER;
  const long nin = 6L * (blbc && S.bl_inlet == 2 ? (long)S.ny * S.nz : blbc && S.bl_inlet == 1 ? (long)S.ny : 1L);
  const long nup = 6L * (blbc && S.bl_upper == 2 ? (long)S.nx * S.nz : blbc && S.bl_upper == 1 ? (long)S.nx : 1L);
  G.bl_nin = nin;
  G.bl_nup = nup;
  double* bl = dalloc(nin + nup + (long)S.nx * S.nz);
  void* nr = nullptr;
  if (!G.part || !bl || hipMalloc(&nr, sizeof(int) * (size_t)S.nx * S.nz) != hipSuccess) return fail("ocfd_init: hipMalloc failed");
  G.allocs.push_back(nr);
  hipMemsetAsync(nr, 0, sizeof(int) * (size_t)S.nx * S.nz, G.stream);
  S.bl_in = bl;
  S.bl_up = bl + nin;
  S.bl_pertb = bl + nin + nup;
  S.bl_nonref = (const int*)nr;
  {
    double* htp = dalloc(1);
    if (!htp) return fail("ocfd_init: hipMalloc failed");
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, G.stream, htp, 1L, 1.0);
    S.bl_ht = htp;
  }
  hipLaunchKernelGGL(k_fill, grid1(nin + nup + (long)S.nx * S.nz), dim3(256), 0, G.stream, bl,
                     (long)S.ny * 6 + 6 + (long)S.nx * S.nz, 0.0);
  if (hipHostMalloc((void**)&G.h_negt, sizeof(int)) != hipSuccess) return fail("ocfd_init: hipHostMalloc failed");
  *G.h_negt = 0;
  {
    // metric tables (off until a mesh has been set and checked)
    const long nxy = (long)S.nx * S.ny, nxz = (long)S.nx * S.nz, nyz = (long)S.ny * S.nz;
    double* t = dalloc((long)S.nx + S.ny + S.nz + 2 * (nxy + nxz + nyz));
    void* bad = nullptr;
    if (!t || hipMalloc(&bad, sizeof(unsigned long long)) != hipSuccess) return fail("ocfd_init: hipMalloc failed");
    G.allocs.push_back(bad);
    G.mt_bad = (unsigned long long*)bad;
    S.mt_xi = t;
    S.mt_yj = t + S.nx;
    S.mt_zk = t + S.nx + S.ny;
    t += S.nx + S.ny + S.nz;
    S.mt_xj = t;
    S.mt_yi = t + nxy;
    S.mt_xk = t + 2 * nxy;
    S.mt_zi = t + 2 * nxy + nxz;
    S.mt_yk = t + 2 * nxy + 2 * nxz;
    S.mt_zj = t + 2 * nxy + 2 * nxz + nyz;
    S.mt_on = 0;
    G.mt_valid = false;
  }
  HIPCHK(hipStreamSynchronize(G.stream));
  return 0;
}

int32_t ocfd_set_option(int32_t key, int32_t value) {
  switch (key) {
    case OCFD_OPT_HYBRID_SUM_MODE: G.hybrid_sum_mode = value; drop_graph(); return 0;
    case OCFD_OPT_USE_GRAPH: G.use_graph = value; drop_graph(); return 0;
    case OCFD_OPT_KERNEL_TIMING: collect_timing(); G.timing = value; drop_graph(); return 0;
    case OCFD_OPT_CONCURRENT: G.concurrent = value ? 1 : 0; drop_graph(); return 0;
    case OCFD_OPT_HYBRID_CERT_WIDEN:
      if (value < 0 || value > 60) return fail("ocfd_set_option: certificate widening must be 0..60");
      G.S.hyb_eta = G.hyb_eta0 * std::ldexp(1.0, value);
      drop_graph();
      return 0;
    case OCFD_OPT_COLUMN_PLANES:
      if (value < -1) return fail("ocfd_set_option: column planes must be >= -1");
      G.col_planes = value;
      drop_graph();
      return 0;
    case OCFD_OPT_METRIC_TABLES:
      G.metric_tables = value ? 1 : 0;
      G.S.mt_on = G.metric_tables && G.mt_valid;
      drop_graph();
      return 0;
  }
  return fail("ocfd_set_option: unknown key");
}

static int upload_gas() {
  HIPCHK(hipStreamSynchronize(G.stream));  // kernels in flight read the bank
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(fg), &G.hg, sizeof(FlowGas)));
  drop_graph();
  return 0;
}

int32_t ocfd_set_species(int32_t n_tct, const double* tct, double tct_max, const double* mi, const double* coef) {
  if (check_ready()) return 1;
  const int ns = G.S.ns;
  if (!mi) return fail("ocfd_set_species: null molar masses");
  for (int k = 0; k < ns; k++)
    if (!(mi[k] > 0.0)) return fail("ocfd_set_species: bad molar mass");
  if (G.cfg.if_real_gas && (n_tct < 1 || n_tct > GAS_NT_MAX || !coef || (n_tct > 1 && !tct)))
    return fail("ocfd_set_species: the real gas needs 1..4 temperature ranges and their coefficients");
  G.S.Mi0 = mi[0];
  G.S.Ri0 = 8.314 / mi[0];  // Sp%Ri = R0/Sp%Mi, CHEM_thermdyn.f90:41
  FlowGas& h = G.hg;
  h.ns = ns;
  h.real_gas = G.cfg.if_real_gas;
  h.flag_vis = G.cfg.flag_vis;
  h.flag_cond = G.cfg.flag_cond;
  h.flag_diff = G.cfg.flag_diff;
  h.n_tct = n_tct < 1 ? 1 : (n_tct > GAS_NT_MAX ? GAS_NT_MAX : n_tct);
  for (int j = 0; j < GAS_NT_MAX; j++) h.tct[j] = (tct && j < h.n_tct - 1) ? tct[j] : 0.0;
  h.tct_max = tct_max;
  auto P3 = [](double x) { double x2 = x * x; return x2 * x; };
  auto P4 = [](double x) { double x2 = x * x; return x2 * x2; };
  auto P5 = [](double x) { double x2 = x * x; return (x2 * x2) * x; };
  for (int k = 0; k < ns; k++)
    for (int j = 0; j < ns; j++) {
      h.wq[k][j] = std::sqrt(std::sqrt(mi[j] / mi[k]));       // comput_vis_coeff, TRANSPORT_modules.f90:300-301
      h.wrec[k][j] = 1.0 / std::sqrt(8.0 * (1.0 + mi[k] / mi[j]));
    }
  for (int k = 0; k < ns; k++) {
    h.Mi[k] = mi[k];
    h.Ri[k] = GAS_R0 / mi[k];
    for (int j = 0; j < GAS_NT_MAX; j++)
      for (int a = 0; a < 7; a++) h.coef[k][j][a] = (coef && j < h.n_tct) ? coef[((long)k * h.n_tct + j) * 7 + a] : 0.0;
    // energy offsets that make the internal energy continuous across the
    // range boundaries (CHEM_thermdyn.f90:42-46)
    h.ect[k][0] = 0.0;
    for (int j = 0; j + 1 < h.n_tct; j++) {
      const double t = h.tct[j];
      h.ect[k][j + 1] = -((h.coef[k][j + 1][0] - h.coef[k][j][0]) * t + (h.coef[k][j + 1][1] - h.coef[k][j][1]) * (t * t) / 2.0 +
                          (h.coef[k][j + 1][2] - h.coef[k][j][2]) * P3(t) / 3.0 +
                          (h.coef[k][j + 1][3] - h.coef[k][j][3]) * P4(t) / 4.0 +
                          (h.coef[k][j + 1][4] - h.coef[k][j][4]) * P5(t) / 5.0) +
                        h.ect[k][j];
    }
  }
  return upload_gas();
}

int32_t ocfd_set_transport(const double* mu_an, const double* k_an, const double* D_an) {
  if (check_ready()) return 1;
  const int ns = G.S.ns;
  FlowGas& h = G.hg;
  if ((G.cfg.flag_vis && !mu_an) || (G.cfg.flag_cond && !k_an) || (G.cfg.flag_diff && !D_an))
    return fail("ocfd_set_transport: a fitted coefficient set is missing");
  for (int k = 0; k < ns; k++)
    for (int a = 0; a < 4; a++) {
      h.mu_an[k][a] = mu_an ? mu_an[4 * k + a] : 0.0;
      h.k_an[k][a] = k_an ? k_an[4 * k + a] : 0.0;
      for (int j = 0; j < ns; j++) h.D_an[k][j][a] = D_an ? D_an[4 * ((long)k * ns + j) + a] : 0.0;
    }
  return upload_gas();
}

int32_t ocfd_set_bc_jet(const double* inlet2d) {
  if (check_ready()) return 1;
  if (G.cfg.ibc != OCFD_BC_JET) return fail("ocfd_set_bc_jet: the run is not configured with the jet boundary condition");
  if (!inlet2d) return fail("ocfd_set_bc_jet: null pointer");
  const DevState& S = G.S;
  HIPCHK(hipMemcpyAsync((void*)S.jet_in, inlet2d, sizeof(double) * (size_t)S.ny * S.nz * (5 + S.ns), hipMemcpyHostToDevice,
                        G.stream));
  HIPCHK(hipStreamSynchronize(G.stream));
  return 0;
}

// init_mesh (OCFD_mesh.f90:1-57) from halo-shaped host coordinate arrays:
// the coordinates' interior, periodic ghosts with the span shift
// (updata_Axyz_periodic), comput_Jacobian3d -- or, when `metrics` is given
// (GRID_AND_JACOBIAN3D), the ten interior arrays Akx..Asz, Ajac read from
// the grid file instead -- their periodic ghosts, Jac_Ghost_boundary and the
// A*1 = A*/Ajac arrays.
static int mesh_setup(const double* axx, const double* ayy, const double* azz, const double* span,
                      const double* metrics) {
  DevState& S = G.S;
  const size_t bytes = sizeof(double) * (size_t)S.nh;
  // staging buffer, freed on every return path (hipFree waits for the device)
  struct Staging {
    double* p = nullptr;
    ~Staging() { if (p) hipFree(p); }
  } stage;
  if (hipMalloc(&stage.p, bytes) != hipSuccess) return fail("ocfd_set_mesh: hipMalloc failed");
  double* tmp = stage.p;
  const double* src[3] = {axx, ayy, azz};
  double* dst[3] = {S.Axx, S.Ayy, S.Azz};
  for (int a = 0; a < 3; a++) {
    HIPCHK(hipMemcpyAsync(tmp, src[a], bytes, hipMemcpyHostToDevice, G.stream));
    hipLaunchKernelGGL(k_copy_interior, grid1(S.ni), dim3(256), 0, G.stream, S, (const double*)tmp, dst[a]);
    HIPCHK(hipStreamSynchronize(G.stream));
  }
  for (int a = 0; a < 3; a++) {
    const long npx = (long)S.ny * S.nz, npy = (long)S.nx * S.nz, npz = (long)S.nx * S.ny;
    if (S.per[0]) hipLaunchKernelGGL(k_fill_faces<0>, grid1(npx * 8), dim3(256), 0, G.stream, S, dst[a], 1, span ? span[a] : 0.0);
    if (S.per[1]) hipLaunchKernelGGL(k_fill_faces<1>, grid1(npy * 8), dim3(256), 0, G.stream, S, dst[a], 1, span ? span[3 + a] : 0.0);
    if (S.per[2]) hipLaunchKernelGGL(k_fill_faces<2>, grid1(npz * 8), dim3(256), 0, G.stream, S, dst[a], 1, span ? span[6 + a] : 0.0);
  }
  double* M[10] = {S.Akx, S.Aky, S.Akz, S.Aix, S.Aiy, S.Aiz, S.Asx, S.Asy, S.Asz, S.Ajac};
  if (!metrics) {
    hipLaunchKernelGGL(k_metrics, grid1(S.ni), dim3(256), 0, G.stream, S);
  } else {
    for (int a = 0; a < 10; a++) {
      HIPCHK(hipMemcpyAsync(tmp, metrics + (size_t)a * S.ni, sizeof(double) * (size_t)S.ni, hipMemcpyHostToDevice,
                            G.stream));
      hipLaunchKernelGGL(k_interior_to_halo, grid1(S.ni), dim3(256), 0, G.stream, S, (const double*)tmp, M[a]);
      HIPCHK(hipStreamSynchronize(G.stream));
    }
  }
  for (int a = 0; a < 10; a++) {
    const long npx = (long)S.ny * S.nz, npy = (long)S.nx * S.nz, npz = (long)S.nx * S.ny;
    if (S.per[0]) hipLaunchKernelGGL(k_fill_faces<0>, grid1(npx * 8), dim3(256), 0, G.stream, S, M[a], 0, 0.0);
    if (S.per[1]) hipLaunchKernelGGL(k_fill_faces<1>, grid1(npy * 8), dim3(256), 0, G.stream, S, M[a], 0, 0.0);
    if (S.per[2]) hipLaunchKernelGGL(k_fill_faces<2>, grid1(npz * 8), dim3(256), 0, G.stream, S, M[a], 0, 0.0);
  }
  // Jac_Ghost_boundary (OCFD_mesh.f90:36): first-order faces, then second-order
  for (int f = 0; f < 6; f++)
    if (G.cfg.ghost_cell[f] == 1)
      hipLaunchKernelGGL(k_jac_ghost1, grid1(ghost_count(f)), dim3(256), 0, G.stream, S, f);
  for (int f = 0; f < 6; f++)
    if (G.cfg.ghost_cell[f] == 2) {
      hipLaunchKernelGGL(k_jac_ghost2_coords, grid1(ghost_count(f)), dim3(256), 0, G.stream, S, f);
      hipLaunchKernelGGL(k_jac_ghost2_metrics, grid1(ghost_count(f)), dim3(256), 0, G.stream, S, f);
    }
  hipLaunchKernelGGL(k_metrics_over_jac, grid1(S.nh), dim3(256), 0, G.stream, S);
  // Metric tables: built from the coordinates as k_metrics differentiates
  // them, then compared with every stored interior metric; they replace the
  // arrays in the stepping kernels only if no bit differs (a mesh that is not
  // GRID1D, or metrics given by the caller, fail the check).
  S.mt_on = 0;
  G.mt_valid = false;
  if (!metrics) {
    const long nt = (long)S.nx + S.ny + S.nz + (long)S.nx * S.ny + (long)S.nx * S.nz + (long)S.ny * S.nz;
    hipLaunchKernelGGL(k_mettab_build, grid1(nt), dim3(256), 0, G.stream, S, (double*)S.mt_xi, (double*)S.mt_yj,
                       (double*)S.mt_zk, (double*)S.mt_xj, (double*)S.mt_yi, (double*)S.mt_xk, (double*)S.mt_zi,
                       (double*)S.mt_yk, (double*)S.mt_zj);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, G.stream, G.mt_bad);
    hipLaunchKernelGGL(k_mettab_check, grid1(S.ni), dim3(256), 0, G.stream, S, G.mt_bad);
    unsigned long long bad = 1;
    HIPCHK(hipMemcpyAsync(&bad, G.mt_bad, sizeof bad, hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    G.mt_valid = bad == 0;
    S.mt_on = G.metric_tables && G.mt_valid;
  }
  HIPCHK(hipStreamSynchronize(G.stream));
  HIPCHK(hipGetLastError());
  drop_graph();
  return 0;
}

int32_t ocfd_set_mesh(const double* axx, const double* ayy, const double* azz, const double* span) {
  if (check_ready()) return 1;
  return mesh_setup(axx, ayy, azz, span, nullptr);
}

int32_t ocfd_metric_tables_in_use(void) { return G.S.mt_on ? 1 : 0; }

// ---------------------------------------------------------------------------
// The reference's files: Fortran unformatted sequential records, each framed
// by a 4-byte length before and after the payload (records of 2 GiB or more,
// which compilers split, are refused).
// ---------------------------------------------------------------------------
struct FortranFile {
  FILE* f = nullptr;
  std::string path;
  ~FortranFile() { if (f) fclose(f); }
  int open(const char* p, const char* mode, const char* who) {
    if (!p) return fail(std::string(who) + ": null path");
    path = p;
    f = fopen(p, mode);
    if (!f) return fail(std::string(who) + ": cannot open " + path);
    return 0;
  }
  // one record; its length must be `bytes` (or at least `bytes` if `prefix`,
  // the rest of the record is then skipped)
  int read(void* buf, size_t bytes, const char* who, bool prefix = false) {
    int32_t a = 0, b = 0;
    if (fread(&a, 4, 1, f) != 1) return fail(std::string(who) + ": " + path + ": unexpected end of file");
    if (a < 0 || (prefix ? (size_t)a < bytes : (size_t)a != bytes))
      return fail(std::string(who) + ": " + path + ": record of " + std::to_string(a) + " bytes where " +
                  std::to_string(bytes) + " were expected");
    if (bytes && fread(buf, 1, bytes, f) != bytes) return fail(std::string(who) + ": " + path + ": short record");
    if ((size_t)a > bytes && fseek(f, (long)((size_t)a - bytes), SEEK_CUR) != 0)
      return fail(std::string(who) + ": " + path + ": seek failed");
    if (fread(&b, 4, 1, f) != 1 || b != a) return fail(std::string(who) + ": " + path + ": bad record marker");
    return 0;
  }
  int write(const void* buf, size_t bytes, const char* who) {
    if (bytes > 0x7fffffffu) return fail(std::string(who) + ": record too large");
    const int32_t a = (int32_t)bytes;
    if (fwrite(&a, 4, 1, f) != 1 || (bytes && fwrite(buf, 1, bytes, f) != bytes) || fwrite(&a, 4, 1, f) != 1)
      return fail(std::string(who) + ": " + path + ": write failed");
    return 0;
  }
};

// read_3d1 / write_3d1 of one field: nz records of nx*ny reals, k plane by k
// plane (OCFD_IO_mpi.f90:151-307 with one rank), between the file and a
// contiguous host array of nx*ny*nz reals
static int read_field3d(FortranFile& F, double* host, const char* who) {
  const DevState& S = G.S;
  const size_t plane = (size_t)S.nx * S.ny;
  for (int k = 0; k < S.nz; k++)
    if (F.read(host + plane * k, sizeof(double) * plane, who)) return 1;
  return 0;
}
static int write_field3d(FortranFile& F, const double* host, const char* who) {
  const DevState& S = G.S;
  const size_t plane = (size_t)S.nx * S.ny;
  for (int k = 0; k < S.nz; k++)
    if (F.write(host + plane * k, sizeof(double) * plane, who)) return 1;
  return 0;
}

// read_flow_data + init3d (OCFD_IO.f90:5-33, OCFD_init.f90:1-25)
int32_t ocfd_read_flow_data(const char* path, int32_t* istep, double* tt) {
  if (check_ready()) return 1;
  const DevState& S = G.S;
  FortranFile F;
  if (F.open(path, "rb", "ocfd_read_flow_data")) return 1;
  unsigned char head[12];
  if (F.read(head, 12, "ocfd_read_flow_data", true)) return 1;
  int32_t is;
  double t;
  std::memcpy(&is, head, 4);
  std::memcpy(&t, head + 4, 8);
  std::vector<double> host((size_t)S.ni);
  struct Staging {
    double* p = nullptr;
    ~Staging() { if (p) hipFree(p); }
  } stage;
  if (hipMalloc(&stage.p, sizeof(double) * (size_t)S.ni) != hipSuccess) return fail("ocfd_read_flow_data: hipMalloc failed");
  double* dst[5] = {S.d, S.u, S.v, S.w, S.T};
  for (int a = 0; a < 5 + S.ns; a++) {
    if (read_field3d(F, host.data(), "ocfd_read_flow_data")) return 1;
    double* to = a < 5 ? dst[a] : S.di + (long)(a - 5) * S.nh;
    HIPCHK(hipMemcpyAsync(stage.p, host.data(), sizeof(double) * (size_t)S.ni, hipMemcpyHostToDevice, G.stream));
    hipLaunchKernelGGL(k_interior_to_halo, grid1(S.ni), dim3(256), 0, G.stream, S, (const double*)stage.p, to);
    HIPCHK(hipStreamSynchronize(G.stream));
  }
  hipLaunchKernelGGL(k_init3d, grid1(S.ni), dim3(256), 0, G.stream, S);
  HIPCHK(hipStreamSynchronize(G.stream));
  HIPCHK(hipGetLastError());
  if (istep) *istep = is;
  if (tt) *tt = t;
  return 0;
}

// save_flow_data (OCFD_IO.f90:38-64): the caller names the file
// (the reference uses OCFD-comb%08d.dat with Istep)
int32_t ocfd_save_flow_data(const char* path, int32_t istep, double tt) {
  if (check_ready()) return 1;
  const DevState& S = G.S;
  FortranFile F;
  if (F.open(path, "wb", "ocfd_save_flow_data")) return 1;
  unsigned char head[12];
  std::memcpy(head, &istep, 4);
  std::memcpy(head + 4, &tt, 8);
  if (F.write(head, 12, "ocfd_save_flow_data")) return 1;
  std::vector<double> host((size_t)S.ni);
  struct Staging {
    double* p = nullptr;
    ~Staging() { if (p) hipFree(p); }
  } stage;
  if (hipMalloc(&stage.p, sizeof(double) * (size_t)S.ni) != hipSuccess) return fail("ocfd_save_flow_data: hipMalloc failed");
  const double* srcs[5] = {S.d, S.u, S.v, S.w, S.T};
  for (int a = 0; a < 5 + S.ns; a++) {
    const double* from = a < 5 ? srcs[a] : S.di + (long)(a - 5) * S.nh;
    hipLaunchKernelGGL(k_halo_to_interior, grid1(S.ni), dim3(256), 0, G.stream, S, from, stage.p);
    HIPCHK(hipMemcpyAsync(host.data(), stage.p, sizeof(double) * (size_t)S.ni, hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    if (write_field3d(F, host.data(), "ocfd_save_flow_data")) return 1;
  }
  return 0;
}

// init_mesh from OCFD-grid.dat (OCFD_mesh.f90:1-57; read_mesh1d,
// read_mesh2d_plane, read_mesh2d_AxialSymm, read_mesh3d: OCFD_IO_mpi.f90
// and OCFD_mesh.f90:187-317) for Iflag_Gridtype = 10, 20, 21, 30 or 31
int32_t ocfd_read_mesh(const char* path, int32_t gridtype, const double* span) {
  if (check_ready()) return 1;
  const DevState& S = G.S;
  const char* who = "ocfd_read_mesh";
  if (gridtype != 10 && gridtype != 20 && gridtype != 21 && gridtype != 30 && gridtype != 31)
    return fail("ocfd_read_mesh: Iflag_Gridtype must be 10, 20, 21, 30 or 31");
  FortranFile F;
  if (F.open(path, "rb", who)) return 1;
  const long nx = S.nx, ny = S.ny, nz = S.nz;
  std::vector<double> X((size_t)S.nh, 1.0), Y((size_t)S.nh, 1.0), Z((size_t)S.nh, 1.0);
  auto hidx = [&](long i, long j, long k) {  // 1-based interior index into a halo array
    return (size_t)((i + OCFD_LAP - 1) + S.nxh * ((j + OCFD_LAP - 1) + (long)S.nyh * (k + OCFD_LAP - 1)));
  };
  std::vector<double> metrics;
  if (gridtype == 10 || gridtype == 20 || gridtype == 21) {
    const bool one = gridtype == 10;
    std::vector<double> a(one ? nx : nx * ny), b(one ? ny : nx * ny), c(nz);
    if (F.read(a.data(), sizeof(double) * a.size(), who) || F.read(b.data(), sizeof(double) * b.size(), who) ||
        F.read(c.data(), sizeof(double) * c.size(), who))
      return 1;
    for (long k = 1; k <= nz; k++)
      for (long j = 1; j <= ny; j++)
        for (long i = 1; i <= nx; i++) {
          const size_t h = hidx(i, j, k);
          if (one) {
            X[h] = a[i - 1];
            Y[h] = b[j - 1];
            Z[h] = c[k - 1];
          } else {
            const size_t p = (size_t)((i - 1) + nx * (j - 1));
            X[h] = a[p];
            if (gridtype == 20) {
              Y[h] = b[p];
              Z[h] = c[k - 1];
            } else {  // y horizontal, z up; seta the angle from the y axis
              Y[h] = b[p] * std::cos(c[k - 1]);
              Z[h] = b[p] * std::sin(c[k - 1]);
            }
          }
        }
  } else {
    std::vector<double> v((size_t)S.ni);
    std::vector<double>* xyz[3] = {&X, &Y, &Z};
    for (int a = 0; a < 3; a++) {
      if (read_field3d(F, v.data(), who)) return 1;
      for (long k = 1; k <= nz; k++)
        for (long j = 1; j <= ny; j++)
          for (long i = 1; i <= nx; i++) (*xyz[a])[hidx(i, j, k)] = v[(size_t)((i - 1) + nx * ((j - 1) + ny * (k - 1)))];
    }
    if (gridtype == 31) {
      metrics.resize((size_t)10 * S.ni);
      for (int a = 0; a < 10; a++)
        if (read_field3d(F, metrics.data() + (size_t)a * S.ni, who)) return 1;
    }
  }
  return mesh_setup(X.data(), Y.data(), Z.data(), span, metrics.empty() ? nullptr : metrics.data());
}

int32_t ocfd_upload_field(int32_t field, const double* host) {
  if (check_ready()) return 1;
  long n;
  double* p = field_ptr(field, &n);
  if (!p || !host) return fail("ocfd_upload_field: bad field id or null pointer");
  HIPCHK(hipMemcpyAsync(p, host, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, G.stream));
  HIPCHK(hipStreamSynchronize(G.stream));
  return 0;
}

int32_t ocfd_download_field(int32_t field, double* host) {
  if (check_ready()) return 1;
  long n;
  double* p = field_ptr(field, &n);
  if (!p || !host) return fail("ocfd_download_field: bad field id or null pointer");
  HIPCHK(hipMemcpyAsync(host, p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, G.stream));
  HIPCHK(hipStreamSynchronize(G.stream));
  return 0;
}

int32_t ocfd_download_mask(int32_t dir, uint8_t* host) {
  if (check_ready()) return 1;
  if (dir < 0 || dir > 2 || !host) return fail("ocfd_download_mask: bad direction or null pointer");
  const DevState& S = G.S;
  const size_t nm = (size_t)(S.nx + (dir == 0)) * (S.ny + (dir == 1)) * (S.nz + (dir == 2));
  HIPCHK(hipMemcpyAsync(host, S.mask[dir], nm, hipMemcpyDeviceToHost, G.stream));
  HIPCHK(hipStreamSynchronize(G.stream));
  return 0;
}

int32_t ocfd_set_bc_boundarylayer(const double* inlet, const double* upper, const int32_t* nonreflect_upper,
                                  const double* wall_pertb) {
  if (check_ready()) return 1;
  const DevState& S = G.S;
  if (G.cfg.ibc != OCFD_BC_BOUNDARYLAYER) return fail("ocfd_set_bc_boundarylayer: the configuration has no boundary-layer BC");
  if (!inlet || (!upper && S.bl_upper >= 0) || !nonreflect_upper || !wall_pertb)
    return fail("ocfd_set_bc_boundarylayer: null pointer");
  HIPCHK(hipMemcpyAsync((void*)S.bl_in, inlet, sizeof(double) * (size_t)G.bl_nin, hipMemcpyHostToDevice, G.stream));
  if (S.bl_upper >= 0)
    HIPCHK(hipMemcpyAsync((void*)S.bl_up, upper, sizeof(double) * (size_t)G.bl_nup, hipMemcpyHostToDevice, G.stream));
  HIPCHK(hipMemcpyAsync((void*)S.bl_nonref, nonreflect_upper, sizeof(int) * (size_t)S.nx * S.nz, hipMemcpyHostToDevice, G.stream));
  HIPCHK(hipMemcpyAsync((void*)S.bl_pertb, wall_pertb, sizeof(double) * (size_t)S.nx * S.nz, hipMemcpyHostToDevice, G.stream));
  HIPCHK(hipStreamSynchronize(G.stream));
  return 0;
}

int32_t ocfd_set_bc_time(double ht) {
  // a device scalar, so a captured step graph stays valid when ht changes
  if (check_ready()) return 1;
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, G.stream, (double*)G.S.bl_ht, 1L, ht);
  HIPCHK(hipGetLastError());
  return 0;
}

static int check_negt() {
  if (*G.h_negt > 10) return fail("ocfd: number of negative-Et points > 10 (handle_NegativeEt stops the run)");
  return 0;
}

int32_t ocfd_flow_time_advance(double dt1) {
  if (check_ready()) return 1;
  if (check_negt()) return 1;
  if (G.use_graph && !G.timing) {
    if (!G.gexec || G.graph_dt != dt1) {
      drop_graph();
      HIPCHK(hipStreamBeginCapture(G.stream, hipStreamCaptureModeThreadLocal));
      enq_step(dt1);
      HIPCHK(hipStreamEndCapture(G.stream, &G.graph));
      HIPCHK(hipGraphInstantiate(&G.gexec, G.graph, nullptr, nullptr, 0));
      G.graph_dt = dt1;
    }
    HIPCHK(hipGraphLaunch(G.gexec, G.stream));
  } else {
    enq_step(dt1);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int32_t ocfd_save_stage0(void) {
  if (check_ready()) return 1;
  launch("k_negt_reset", k_negt_reset, dim3(1), dim3(64), G.S, 1);
  launch("k_save0", k_save0, grid1(G.S.ni), dim3(256), G.S);
  HIPCHK(hipGetLastError());
  return 0;
}
int32_t ocfd_comput_amu_amk(void) {
  if (check_ready()) return 1;
  if (G.S.ns > 1) launch("k_transport_ms", k_transport_ms, grid1(G.S.ni), dim3(256), G.S);
  else launch("k_transport", k_transport, grid1(G.S.ni), dim3(256), G.S);
  HIPCHK(hipGetLastError());
  return 0;
}
int32_t ocfd_stage_begin(int32_t krk) {
  if (check_ready()) return 1;
  if (G.S.ns > 1) {
    launch("k_transport_ms", k_transport_ms, grid1(G.S.ni), dim3(256), G.S);
  } else {
    launch("k_transport", k_transport, grid1(G.S.ni), dim3(256), G.S);
    launch("k_di_from_d", k_di_from_d, grid1(G.S.nh), dim3(256), G.S);
  }
  if (krk == 1 && G.cfg.scheme_invis == OCFD_SCHEME_HYBRID) enq_rhybrid();
  HIPCHK(hipGetLastError());
  return 0;
}
int32_t ocfd_comput_rhybrid(void) {
  if (check_ready()) return 1;
  enq_rhybrid();
  HIPCHK(hipGetLastError());
  return 0;
}
int32_t ocfd_du_inviscous(void) {
  if (check_ready()) return 1;
  launch_sweep<0>();
  launch_sweep<1>();
  launch_sweep<2>();
  if (G.S.mass_force) launch_rk(1, 0.0, (int)(RK_MASS | RK_WRITE_DU));
  HIPCHK(hipGetLastError());
  return 0;
}
int32_t ocfd_flux1d_batch(int32_t side, int32_t scheme, int64_t npen, int32_t n, const double* v, double* hh,
                          const uint8_t* scm, const int32_t* bound_index, const int32_t* scheme_boundary) {
  if (check_ready()) return 1;
  if (side != 0 && side != 1) return fail("ocfd_flux1d_batch: side must be 0 (flux1) or 1 (flux2)");
  if (scheme != OCFD_SCHEME_WENO5 && scheme != OCFD_SCHEME_WENO7 && scheme != OCFD_SCHEME_OMP6 &&
      scheme != OCFD_SCHEME_UD7L && scheme != OCFD_SCHEME_HYBRID)
    return fail("ocfd_flux1d_batch: unknown inviscid scheme");
  if (npen < 0 || n < 1) return fail("ocfd_flux1d_batch: npen >= 0 and n >= 1 required");
  if (npen > 0 && (!v || !hh)) return fail("ocfd_flux1d_batch: null v or hh");
  if (scheme == OCFD_SCHEME_HYBRID && npen > 0 && !scm) return fail("ocfd_flux1d_batch: the hybrid scheme needs scm");
  if (!bound_index || !scheme_boundary) return fail("ocfd_flux1d_batch: null bound_index or scheme_boundary");
  for (int a = 0; a < 2; a++) {
    if (bound_index[a] != 0 && bound_index[a] != 1) return fail("ocfd_flux1d_batch: Bound_index must be 0 or 1");
    if (scheme_boundary[a] < -1 || scheme_boundary[a] > 1) return fail("ocfd_flux1d_batch: Scheme_boundary must be -1, 0 or 1");
  }
  const long total = npen * ((long)n + 1);
  if (total > 0)
    launch("k_flux1d", k_flux1d, grid1(total), dim3(256), G.S, (int)side, (int)scheme, (long)npen, (int)n, v, hh,
           scm, (int)bound_index[0], (int)bound_index[1], (int)scheme_boundary[0], (int)scheme_boundary[1]);
  HIPCHK(hipGetLastError());
  return 0;
}

int32_t ocfd_du_viscous(void) {
  if (check_ready()) return 1;
  launch_visc_prep(G.S.ns > 1);
  launch_rk(1, 0.0, (int)(RK_VISC | RK_WRITE_DU));
  HIPCHK(hipGetLastError());
  return 0;
}
int32_t ocfd_time_adv_rk3(int32_t krk, double dt) {
  if (check_ready()) return 1;
  if (krk < 1 || krk > 3) return fail("ocfd_time_adv_rk3: KRK must be 1, 2 or 3");
  launch_rk((int)krk, dt, (int)RK_UPDATE_F);
  HIPCHK(hipGetLastError());
  return 0;
}
int32_t ocfd_comput_flow_variables(void) {
  if (check_ready()) return 1;
  launch("k_negt_reset", k_negt_reset, dim3(1), dim3(64), G.S, 0);
  launch_rk(1, 0.0, (int)(G.S.ns > 1 ? RK_PRIMS | RK_MS : RK_PRIMS));
  launch("k_negt_repair", k_negt_repair, dim3(1), dim3(64), G.S);
  hipMemcpyAsync(G.h_negt, G.S.negt + 17, sizeof(int), hipMemcpyDeviceToHost, G.stream);
  HIPCHK(hipGetLastError());
  return 0;
}
// filtering (OCFD_filtering.f90:4-76), called by NS_solver after Istep has
// been advanced (OCFD_NS_Solver.f90:45-48)
int32_t ocfd_filtering(int32_t istep, int32_t nfiltering, const double* filter) {
  if (check_ready()) return 1;
  if (nfiltering < 0 || nfiltering > 10) return fail("ocfd_filtering: NFiltering must be 0..10");
  if (nfiltering > 0 && !filter) return fail("ocfd_filtering: null Filter array");
  const DevState& S = G.S;
  bool any = false;
  for (int m = 0; m < nfiltering; m++) {
    const long every = std::lround(filter[m * 100]);
    if (every < 1) return fail("ocfd_filtering: Filter(1,m) (the step interval) must be >= 1");
    if (istep % every == 0) any = true;
  }
  if (!any) return 0;
  if (!G.filt_tmp) {
    G.filt_tmp = dalloc(5 * S.ni);
    if (!G.filt_tmp) return fail("ocfd_filtering: hipMalloc failed");
  }
  const int n[3] = {S.nx, S.ny, S.nz};
  for (int m = 0; m < nfiltering; m++) {
    const double* F = filter + m * 100;
    if (istep % std::lround(F[0]) != 0) continue;
    const int scheme = (int)std::lround(F[10]);
    FiltBox B;
    for (int d = 0; d < 3; d++) {
      B.b[d] = std::max((int)std::lround(F[4 + 2 * d]), 1);
      B.e[d] = std::min((int)std::lround(F[5 + 2 * d]), n[d]);
    }
    B.s0 = F[11];
    B.rth = F[12];
    for (int d = 0; d < 3; d++) {
      if (std::lround(F[1 + d]) != 1 || (scheme != 1 && scheme != 2)) continue;
      FiltBox Bd = B;
      if (!S.per[d]) {  // the stencil stays inside the box (:91-93, :204-206)
        Bd.b[d] = std::max(Bd.b[d], scheme == 1 ? 6 : 4);
        Bd.e[d] = std::min(Bd.e[d], n[d] - (scheme == 1 ? 5 : 3));
      }
      const long cnt = (long)std::max(0, Bd.e[0] - Bd.b[0] + 1) * std::max(0, Bd.e[1] - Bd.b[1] + 1) *
                       std::max(0, Bd.e[2] - Bd.b[2] + 1);
      if (scheme == 2 && ocfd_comput_flow_variables()) return 1;
      if (cnt <= 0) continue;
      launch("k_copy5", k_copy5, grid1(5 * S.ni), dim3(256), (const double*)S.f, G.filt_tmp, 5 * S.ni);
      const double* ft = G.filt_tmp;
      if (scheme == 1) {
        if (d == 0) launch("k_filter9", k_filter9<0>, grid1(cnt), dim3(256), S, Bd, ft);
        if (d == 1) launch("k_filter9", k_filter9<1>, grid1(cnt), dim3(256), S, Bd, ft);
        if (d == 2) launch("k_filter9", k_filter9<2>, grid1(cnt), dim3(256), S, Bd, ft);
      } else {
        if (d == 0) launch("k_filter_shock", k_filter_shock<0>, grid1(cnt), dim3(256), S, Bd, ft);
        if (d == 1) launch("k_filter_shock", k_filter_shock<1>, grid1(cnt), dim3(256), S, Bd, ft);
        if (d == 2) launch("k_filter_shock", k_filter_shock<2>, grid1(cnt), dim3(256), S, Bd, ft);
      }
    }
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int32_t ocfd_bc(void) {
  if (check_ready()) return 1;
  enq_bc();
  HIPCHK(hipGetLastError());
  return 0;
}

// ---- multi-species pieces ---------------------------------------------------
int32_t ocfd_du_species(void) {
  if (check_ready()) return 1;
  if (G.S.ns < 2) return fail("ocfd_du_species: needs N_SPEC > 1");
  enq_species(0, 0.0);
  HIPCHK(hipGetLastError());
  return 0;
}
int32_t ocfd_species_time_adv(int32_t krk, double dt) {
  if (check_ready()) return 1;
  if (G.S.ns < 2) return fail("ocfd_species_time_adv: needs N_SPEC > 1");
  if (krk < 1 || krk > 3) return fail("ocfd_species_time_adv: KRK must be 1, 2 or 3");
  enq_species(krk, dt);
  HIPCHK(hipGetLastError());
  return 0;
}
int32_t ocfd_du_scaler_invis(int32_t m) {
  if (check_ready()) return 1;
  if (G.S.ns < 2) return fail("ocfd_du_scaler_invis: needs N_SPEC > 1");
  if (m < 1 || m > G.S.ns) return fail("ocfd_du_scaler_invis: species index must be 1..N_SPEC");
  enq_species_one(0, m - 1, 0, 0.0);
  HIPCHK(hipGetLastError());
  return 0;
}
int32_t ocfd_du_scaler_viscous(int32_t m) {
  if (check_ready()) return 1;
  if (G.S.ns < 2) return fail("ocfd_du_scaler_viscous: needs N_SPEC > 1");
  if (m < 1 || m > G.S.ns) return fail("ocfd_du_scaler_viscous: species index must be 1..N_SPEC");
  if (!G.S.viscous) return fail("ocfd_du_scaler_viscous: the run is inviscid (IF_Viscous = 0)");
  enq_species_one(1, m - 1, 0, 0.0);
  HIPCHK(hipGetLastError());
  return 0;
}
int32_t ocfd_scalar_time_adv_rk3(int32_t krk, int32_t m, double dt) {
  if (check_ready()) return 1;
  if (G.S.ns < 2) return fail("ocfd_scalar_time_adv_rk3: needs N_SPEC > 1");
  if (krk < 1 || krk > 3) return fail("ocfd_scalar_time_adv_rk3: KRK must be 1, 2 or 3");
  if (m < 1 || m > G.S.ns) return fail("ocfd_scalar_time_adv_rk3: species index must be 1..N_SPEC");
  enq_species_one(2, m - 1, krk, dt);
  HIPCHK(hipGetLastError());
  return 0;
}
int32_t ocfd_comput_T_P(void) {
  if (check_ready()) return 1;
  if (G.S.ns < 2) return fail("ocfd_comput_T_P: needs N_SPEC > 1");
  launch("k_comput_T_P", k_comput_T_P, grid1(G.S.ni), dim3(256), G.S);
  HIPCHK(hipGetLastError());
  return 0;
}

// Chemical_time_advance (OCFD_Time_Adv.f90:169-196): the interior cells go
// through the chemistry batch on this library's stream.  The species count
// must be the chemistry build's (9).
int32_t ocfd_chemical_time_advance(double dt2, int32_t mode) {
  if (check_ready()) return 1;
  const DevState& S = G.S;
  if (S.ns != OCFD_NS_MAX) return fail("ocfd_chemical_time_advance: the chemistry batch is built for 9 species");
  if (!G.chem_rho) {
    G.chem_rho = dalloc(S.ni * S.ns);
    G.chem_E = dalloc(S.ni);
    G.chem_dt = dalloc(S.ni);
    if (!G.chem_rho || !G.chem_E || !G.chem_dt) return fail("ocfd_chemical_time_advance: hipMalloc failed");
  }
  launch("k_chem_pack", k_chem_pack, grid1(S.ni), dim3(256), S, G.chem_rho, G.chem_E, G.chem_dt, dt2);
  void* prev = ocfd_chem_stream();
  if (ocfd_chem_set_stream((void*)G.stream)) return fail(std::string("ocfd_chemical_time_advance: ") + ocfd_chem_last_error());
  int rc;
  if (G.chem_method == OCFD_CHEM_ADAPTIVE)
    rc = chem_adaptive_batch(S.ni, G.chem_rho, G.chem_E, G.chem_dt, G.chem_rtol_y, G.chem_y_floor, G.chem_rtol_t,
                             G.chem_max_sub, nullptr, G.chem_scheduler, false);
  else if (G.chem_method == OCFD_CHEM_SURROGATE)
    rc = chem_surrogate_batch(S.ni, G.chem_rho, G.chem_E, G.chem_dt, true, G.chem_rtol_y, G.chem_y_floor, G.chem_rtol_t,
                              G.chem_max_sub, nullptr, false);
  else
    rc = ocfd_chem_timeadv_batch(S.ni, G.chem_rho, G.chem_E, G.chem_dt, mode, 1);
  ocfd_chem_set_stream(prev);
  if (rc) return fail(std::string("ocfd_chemical_time_advance: ") + ocfd_chem_last_error());
  launch("k_chem_unpack", k_chem_unpack, grid1(S.ni), dim3(256), S, (const double*)G.chem_rho, (const double*)G.chem_E);
  HIPCHK(hipGetLastError());
  return 0;
}

int32_t ocfd_cfl_dt(double cfl, double* dt) {
  if (check_ready()) return 1;
  if (!(cfl > 0.0) || !dt) return fail("ocfd_cfl_dt: CFL must be positive and dt non-null");
  const DevState& S = G.S;
  launch("k_cfl_partial", k_cfl_partial, dim3(G.nparts), dim3(256), S, G.part);
  launch("k_cfl_final", k_cfl_final, dim3(1), dim3(64), (const double*)G.part, G.nparts, cfl, S.red + 2);
  double h[2];
  HIPCHK(hipMemcpyAsync(h, S.red + 2, sizeof(h), hipMemcpyDeviceToHost, G.stream));
  HIPCHK(hipStreamSynchronize(G.stream));
  if (!(h[1] > 0.0) || !std::isfinite(h[1])) return fail("ocfd_cfl_dt: no finite positive wave speed (NaN in the state?)");
  *dt = h[0];
  return 0;
}

int32_t ocfd_set_chem_method(int32_t method, double rtol_y, double y_floor, double rtol_t, int32_t max_sub,
                             int32_t scheduler) {
  if (method < OCFD_CHEM_FIXED || method > OCFD_CHEM_SURROGATE) return fail("ocfd_set_chem_method: unknown method");
  if (method != OCFD_CHEM_FIXED &&
      (max_sub < 1 || !(rtol_y > 0.0) || !(y_floor >= 0.0) || !(rtol_t > 0.0) || scheduler < 0 || scheduler > 2))
    return fail("ocfd_set_chem_method: bad controller parameters");
  G.chem_method = method;
  G.chem_rtol_y = rtol_y;
  G.chem_y_floor = y_floor;
  G.chem_rtol_t = rtol_t;
  G.chem_max_sub = max_sub;
  G.chem_scheduler = scheduler;
  return 0;
}

// One time step of NS_solver (OCFD_NS_Solver.f90:26-37): flow dt/2, nsub
// chemistry substeps of dt/nsub and comput_T_P, flow dt/2
int32_t ocfd_ns_step(double dt, int32_t nsub, int32_t mode, int32_t reaction) {
  if (check_ready()) return 1;
  if (nsub < 1) return fail("ocfd_ns_step: nsub must be >= 1");
  const double dt1 = 0.5 * dt, dt2 = dt / nsub;
  if (ocfd_flow_time_advance(dt1)) return 1;
  if (reaction) {
    for (int m = 0; m < nsub; m++)
      if (ocfd_chemical_time_advance(dt2, mode)) return 1;
    if (ocfd_comput_T_P()) return 1;
  }
  return ocfd_flow_time_advance(dt1);
}

// Autotuner for the inviscid sweep tiles: times every tile variant of each
// direction on the current state (the sweeps only overwrite du, which every
// stage recomputes) and keeps the fastest.  Results are bitwise independent
// of the choice.
int32_t ocfd_autotune(int32_t reps) {
  if (check_ready()) return 1;
  if (G.cfg.if_scheme_character) return 0;  // one tile shape for the characteristic sweep
  if (reps < 1) reps = 3;
  drop_graph();
  const int saved_timing = G.timing;
  G.timing = 0;
  // time the sweeps on the masks a step builds from the current state: most
  // interfaces of a hybrid scheme take the linear reconstruction and the
  // rest WENO, and masks not yet built would time neither (the masks are
  // rebuilt at the next step's first stage anyway)
  if (G.cfg.scheme_invis == OCFD_SCHEME_HYBRID) enq_rhybrid();
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  for (int dir = 0; dir < 3; dir++) {
    float best = 1e30f;
    int bestv = 0;
    for (int v = 0; v < OCFD_SWEEP_VARIANTS; v++) {
      auto run = [&]() {
        if (dir == 0) launch_sweep_variant<0>(v);
        else if (dir == 1) launch_sweep_variant<1>(v);
        else launch_sweep_variant<2>(v);
      };
      run();  // warm
      // a hybrid step rebuilds the masks once per RK3 step, so its sweeps
      // classify and store their kind lists in one stage of three and read
      // them back in the other two: time that mix; otherwise the lists are
      // stored once and every launch reads them
      const bool hyb = G.cfg.scheme_invis == OCFD_SCHEME_HYBRID;
      const int per = hyb ? 3 : 1;
      HIPCHK(hipEventRecord(a, G.stream));
      for (int r = 0; r < reps; r++) {
        if (hyb) G.kc_var[dir] = -1;
        for (int l = 0; l < per; l++) run();
      }
      HIPCHK(hipEventRecord(b, G.stream));
      HIPCHK(hipEventSynchronize(b));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, a, b));
      G.tune_ms[dir][v] = ms / (reps * per);
      // isolated launches rank near-equal shapes by noise and differ from a
      // launch inside the step (whose inputs come warm from the previous
      // kernels, 10% apart for the x sweep): keep the default shape unless
      // another is clearly faster
      if (v == 0 || ms < 0.90f * best) { best = ms; bestv = v; }
    }
    G.sweep_variant[dir] = bestv;
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  G.timing = saved_timing;
  HIPCHK(hipGetLastError());
  return 0;
}

int32_t ocfd_tile_choice(int32_t dir, int32_t* np, int32_t* ts, float* ms_per_variant) {
  if (dir < 0 || dir > 2) return fail("ocfd_tile_choice: bad direction");
  static const int tiles[3][OCFD_SWEEP_VARIANTS][2] = {
      {{SweepTile<0, 0>::NP, SweepTile<0, 0>::TS}, {SweepTile<0, 1>::NP, SweepTile<0, 1>::TS},
       {SweepTile<0, 2>::NP, SweepTile<0, 2>::TS}, {SweepTile<0, 3>::NP, SweepTile<0, 3>::TS}},
      {{SweepTile<1, 0>::NP, SweepTile<1, 0>::TS}, {SweepTile<1, 1>::NP, SweepTile<1, 1>::TS},
       {SweepTile<1, 2>::NP, SweepTile<1, 2>::TS}, {SweepTile<1, 3>::NP, SweepTile<1, 3>::TS}},
      {{SweepTile<2, 0>::NP, SweepTile<2, 0>::TS}, {SweepTile<2, 1>::NP, SweepTile<2, 1>::TS},
       {SweepTile<2, 2>::NP, SweepTile<2, 2>::TS}, {SweepTile<2, 3>::NP, SweepTile<2, 3>::TS}}};
  const int v = G.sweep_variant[dir];
  if (np) *np = tiles[dir][v][0];
  if (ts) *ts = tiles[dir][v][1];
  if (ms_per_variant)
    for (int k = 0; k < OCFD_SWEEP_VARIANTS; k++) ms_per_variant[k] = G.tune_ms[dir][k];
  return v;
}

int32_t ocfd_set_tile(int32_t dir, int32_t variant) {
  if (dir < 0 || dir > 2 || variant < 0 || variant >= OCFD_SWEEP_VARIANTS) return fail("ocfd_set_tile: bad argument");
  G.sweep_variant[dir] = variant;
  drop_graph();
  return 0;
}

int32_t ocfd_negative_et_count(void) {
  if (!G.inited) return 0;
  hipStreamSynchronize(G.stream);
  return *G.h_negt;
}

// show_flow_msg (OCFD_Time_Adv.f90:200-239)
int32_t ocfd_show_flow_msg(double* E0) {
  if (check_ready()) return 1;
  if (!E0) return fail("ocfd_show_flow_msg: null output");
  return diag_pass<0>(0.0, E0);
}

// ana_box (OCFD_ANA.f90:197-303): the device means, then the reference's
// derived statistics (:243-270) and the density rms (:272-281)
int32_t ocfd_ana_box(double* out) {
  if (check_ready()) return 1;
  if (!out) return fail("ocfd_ana_box: null output");
  double a[DIAG_SLOTS];
  if (diag_pass<1>(0.0, a)) return 1;
  double u_rms = a[0], v_rms = a[1], w_rms = a[2], Ux_rms = a[3], Ux_sk = a[4], Ux_fl = a[5], d_aver = a[6];
  const double c_aver = a[7], Amu_aver = a[8], E_k = a[9], T_aver = a[10];
  double u_sk = a[11], u_fl = a[12];
  const double VV_rms = std::sqrt((u_rms + v_rms + w_rms) / 3.0);
  const double Amt = std::sqrt(u_rms + v_rms + w_rms) / c_aver;
  u_rms = std::sqrt(u_rms);
  v_rms = std::sqrt(v_rms);
  w_rms = std::sqrt(w_rms);
  Ux_rms = std::sqrt(Ux_rms);
  const double Alamda = VV_rms / Ux_rms, Alamdax = u_rms / Ux_rms;
  const double Re_lamda = VV_rms * Alamda * d_aver / Amu_aver, Re_lamdax = VV_rms * Alamdax * d_aver / Amu_aver;
  Ux_sk = Ux_sk / ((Ux_rms * Ux_rms) * Ux_rms);
  Ux_fl = Ux_fl / ((Ux_rms * Ux_rms) * (Ux_rms * Ux_rms));
  u_sk = u_sk / ((u_rms * u_rms) * u_rms);
  u_fl = u_fl / ((u_rms * u_rms) * (u_rms * u_rms));
  double drms[DIAG_SLOTS];
  if (diag_pass<2>(d_aver, drms)) return 1;
  const double d_rms = std::sqrt(drms[0]);
  const double r[18] = {Re_lamda, Re_lamdax, Amt,   VV_rms, u_rms,  v_rms,    w_rms,    u_sk,   u_fl,
                        Ux_sk,    Ux_fl,     E_k,   d_rms,  T_aver, Amu_aver, d_aver,   Ux_rms, Alamdax};
  for (int m = 0; m < 18; m++) out[m] = r[m];
  return 0;
}

int32_t ocfd_hybrid_certificate(int64_t* uncertain, double* margin, double* eta) {
  if (check_ready()) return 1;
  if (!G.hyb_cert_ran) return fail("ocfd_hybrid_certificate: no certified sensor mean has run");
  unsigned long long v[2];
  HIPCHK(hipMemcpyAsync(v, G.S.red + 4, sizeof(v), hipMemcpyDeviceToHost, G.stream));
  HIPCHK(hipStreamSynchronize(G.stream));
  if (uncertain) *uncertain = (int64_t)v[0];
  if (margin) std::memcpy(margin, &v[1], sizeof(double));
  if (eta) *eta = G.S.hyb_eta;
  return 0;
}

int32_t ocfd_synchronize(void) {
  if (check_ready()) return 1;
  HIPCHK(hipStreamSynchronize(G.stream));
  HIPCHK(hipGetLastError());
  collect_timing();
  return check_negt();
}

void* ocfd_stream(void) { return (void*)G.stream; }

int32_t ocfd_kernel_count(void) {
  collect_timing();
  return (int32_t)G.knames.size();
}
const char* ocfd_kernel_name(int32_t i) {
  if (i < 0 || i >= (int)G.knames.size()) return "";
  return G.knames[i].c_str();
}
int32_t ocfd_kernel_stats(int32_t i, double* total_ms, int64_t* launches) {
  collect_timing();
  if (i < 0 || i >= (int)G.kstats.size()) return fail("ocfd_kernel_stats: bad index");
  if (total_ms) *total_ms = G.kstats[i].ms;
  if (launches) *launches = G.kstats[i].launches;
  return 0;
}
int32_t ocfd_kernel_stats_reset(void) {
  collect_timing();
  for (auto& k : G.kstats) k = KStat{};
  return 0;
}

}  // extern "C"
