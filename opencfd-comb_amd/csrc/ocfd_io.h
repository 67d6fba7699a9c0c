// ocfd_io.h -- device side of the reference's file formats: the restart
// (read_flow_data / save_flow_data, code/mpi-io/OCFD_IO.f90:5-64, written
// per k plane by write_3d, OCFD_IO_mpi.f90:235-307) moves between the
// interior of the resident halo arrays and a contiguous (nx, ny, nz) staging
// array; init3d (code/flow-solver/OCFD_init.f90:1-25) then rebuilds the
// conserved variables from the primitives that were read.
#pragma once

// dst(i,j,k) = src[(i-1) + nx ((j-1) + ny (k-1))] over the interior
__global__ __launch_bounds__(256) void k_interior_to_halo(DevState S, const double* src, double* dst) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ni) return;
  int i, j, k;
  cell_ijk(S, c, i, j, k);
  dst[hq(S, i, j, k)] = src[c];
}

__global__ __launch_bounds__(256) void k_halo_to_interior(DevState S, const double* src, double* dst) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ni) return;
  int i, j, k;
  cell_ijk(S, c, i, j, k);
  dst[c] = src[hq(S, i, j, k)];
}

// init3d after read_flow_data: Rescaling_di (Comput_flow_variables.f90:101-121,
// negative partial densities set to zero, then scaled so that they sum to d),
// then f = (d, d u, d v, d w, E(di, T) + d |u|^2 / 2) with comput_E
// (comput_E_pgas for N_SPEC = 1, CHEM_thermdyn.f90:114-178 / 203-231).
__global__ __launch_bounds__(256) void k_init3d(DevState S) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ni) return;
  int i, j, k;
  cell_ijk(S, c, i, j, k);
  const long q = hq(S, i, j, k);
  double di[OCFD_NS_MAX];
  double d0 = 0.0;
  for (int m = 0; m < S.ns; m++) {
    double x = S.di[q + (long)m * S.nh];
    if (x < 0.0) x = 0.0;
    di[m] = x;
    d0 = d0 + x;
  }
  const double tmp = S.d[q] / d0;
  for (int m = 0; m < S.ns; m++) {
    di[m] = di[m] * tmp;
    S.di[q + (long)m * S.nh] = di[m];
  }
  const double dd = S.d[q], uu = S.u[q], vv = S.v[q], ww = S.w[q], T = S.T[q];
  const double Et1 = S.ns > 1 ? gas_E(di, T) : 0.0 + 2.5 * S.Ri0 * di[0] * T;
  S.f[c] = dd;
  S.f[c + S.ni] = dd * uu;
  S.f[c + 2 * S.ni] = dd * vv;
  S.f[c + 3 * S.ni] = dd * ww;
  S.f[c + 4 * S.ni] = Et1 + dd * (uu * uu + vv * vv + ww * ww) * 0.5;
}
