/*
 * ocfd_hip.h -- C-ABI of the MI355X (gfx950) implementation of OpenCFD-Comb's
 * per-step hot path.
 *
 * Every entry point replaces one reference subroutine (cited as
 * file:line relative to the OpenCFD-Comb source tree).  The reference works
 * on Fortran module globals; this library keeps one global context holding
 * device-resident mirrors of those globals, so the argument lists below stay
 * as close to the reference's as a C ABI allows:
 *   - reals are double (OCFD_REAL_KIND=8, code/head/OCFD_precision.f90:7),
 *     integers int32_t (default INTEGER);
 *   - host arrays are in Fortran layout, i fastest, variable index slowest;
 *     "halo" arrays span (1-LAP:nx+LAP, 1-LAP:ny+LAP, 1-LAP:nz+LAP) and the
 *     pointer is to element (1-LAP,1-LAP,1-LAP); "interior" arrays span
 *     (1:nx,1:ny,1:nz) (code/head/OCFD_allocate_data.f90:7-22);
 *   - every call returns int32 status: 0 ok, >0 error, with the text in
 *     ocfd_last_error().  The reference prints and STOPs on errors
 *     (e.g. code/mpi-io/OCFD_readpara.f90:312-315); the Fortran shim in the
 *     integration notes maps a non-zero status to print + stop.
 *   - not thread-safe: one host thread, one HIP stream (the library's own).
 *
 * Only one GPU and one rank: the MPI halo exchange
 * (code/mpi-io/OCFD_mpi.f90:186-236) becomes on-device periodic index
 * wrapping, and MPI_ALLREDUCE in the hybrid sensor becomes a device
 * reduction.
 */
#ifndef OCFD_HIP_H
#define OCFD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scheme / splitting / BC codes: identical values to OCFD_constants
 * (code/head/MODULE_parameter.f90:14-26). */
enum {
  OCFD_SCHEME_WENO5 = 1, OCFD_SCHEME_WENO7 = 2, OCFD_SCHEME_OMP6 = 3,
  OCFD_SCHEME_UD7L = 4, OCFD_SCHEME_CD6 = 5, OCFD_SCHEME_CD8 = 6,
  OCFD_SCHEME_HYBRID = 10
};
enum { OCFD_SPLIT_SW = 1, OCFD_SPLIT_LLF = 2 };
enum { OCFD_BC_NONE = 0, OCFD_BC_BOUNDARYLAYER = 2, OCFD_BC_JET = 4 };

/* Run configuration: the subset of TYPE_Para / TYPE_Scheme / Transport
 * (code/head/MODULE_parameter.f90:49-94, code/Chemical/TRANSPORT_modules.f90:16-21)
 * that the hot path reads.  Field meanings and defaults follow the namelist
 * /control_opencfd/ (code/mpi-io/OCFD_readpara.f90:41-110). */
typedef struct ocfd_cfg {
  int32_t nx, ny, nz;          /* interior points (global == local: one GPU) */
  int32_t lap;                 /* halo width LAP; must be 4 */
  int32_t n_spec;              /* N_SPEC */
  int32_t periodic[3];         /* Iperiodic_X/Y/Z */
  int32_t scheme_invis;        /* OCFD_SCHEME_* */
  int32_t scheme_vis;          /* OCFD_SCHEME_CD6 or OCFD_SCHEME_CD8 */
  int32_t flux_splitting;      /* OCFD_SPLIT_SW or OCFD_SPLIT_LLF */
  int32_t if_viscous;          /* IF_Viscous */
  int32_t if_scheme_character; /* IF_Scheme_Character: 0 or 1 (characteristic flux) */
  int32_t scheme_boundary[6];  /* Scheme_boundary(1:6): 0, 1 or -1 */
  int32_t ghost_cell[6];       /* Ghost_Cell(1:6): 0, 1 or 2 (non-periodic faces) */
  int32_t ibc;                 /* OCFD_BC_* */
  int32_t if_mass_force;       /* IF_Mass_Force */
  int32_t if_real_gas;         /* IF_RealGas (0 perfect gas, 1 NASA fits) */
  int32_t flag_vis, flag_cond, flag_diff; /* Trans%Flag_* (0: Sutherland/Pr/Sc) */
  int32_t wall_catalysis;      /* Para%Wall_Catalysis (only 0 supported) */
  int32_t pad0;
  double  ud7l_diss;           /* UD7L_Diss */
  double  hybrid_para[16];     /* Hybrid_para(1:16) */
  double  amu0, pr, sc;        /* Trans%Amu0, Trans%Pr, Trans%Sc */
  double  mass_force[3];       /* Mass_Force(1:3) */
  double  bc_para[32];         /* BC_Para(1:32) */
} ocfd_cfg;

/* Field identifiers for ocfd_upload_field / ocfd_download_field. */
enum {
  OCFD_FIELD_F = 0,   /* f(nx,ny,nz,5)            interior x5      */
  OCFD_FIELD_FN,      /* fn(nx,ny,nz,5)           interior x5      */
  OCFD_FIELD_DU,      /* du(nx,ny,nz,5)           interior x5      */
  OCFD_FIELD_DI,      /* di(halo,N_SPEC)          halo x N_SPEC    */
  OCFD_FIELD_D, OCFD_FIELD_U, OCFD_FIELD_V, OCFD_FIELD_W,
  OCFD_FIELD_T, OCFD_FIELD_P, OCFD_FIELD_ET,            /* halo       */
  OCFD_FIELD_AMU, OCFD_FIELD_AMK,                       /* interior   */
  OCFD_FIELD_RHYBRID,                                   /* halo       */
  OCFD_FIELD_AKX, OCFD_FIELD_AKY, OCFD_FIELD_AKZ,       /* halo metrics */
  OCFD_FIELD_AIX, OCFD_FIELD_AIY, OCFD_FIELD_AIZ,
  OCFD_FIELD_ASX, OCFD_FIELD_ASY, OCFD_FIELD_ASZ, OCFD_FIELD_AJAC,
  OCFD_FIELD_AKX1, OCFD_FIELD_AKY1, OCFD_FIELD_AKZ1,
  OCFD_FIELD_AIX1, OCFD_FIELD_AIY1, OCFD_FIELD_AIZ1,
  OCFD_FIELD_ASX1, OCFD_FIELD_ASY1, OCFD_FIELD_ASZ1,
  OCFD_FIELD_DIN,     /* din(nx,ny,nz,N_SPEC)     interior x N_SPEC */
  OCFD_FIELD_AMD,     /* Amd(nx,ny,nz,N_SPEC)     interior x N_SPEC (N_SPEC > 1) */
  OCFD_FIELD_DUS,     /* species right-hand sides interior x N_SPEC (N_SPEC > 1) */
  OCFD_FIELD_QP,      /* Qpx, Qpy, Qpz            interior x 3 */
  OCFD_FIELD_COUNT
};

/* Hybrid-scheme masks (Scm_Hbx/y/z, code/flow-solver/OCFD_inviscous.f90:63-68),
 * kept as uint8 per interface: x (0:nx,ny,nz), y (nx,0:ny,nz), z (nx,ny,0:nz). */
enum { OCFD_MASK_X = 0, OCFD_MASK_Y = 1, OCFD_MASK_Z = 2 };

/* Options for ocfd_set_option. */
enum {
  /* Global mean of |grad p| for hybrid sensor type != 1
   * (MPI_ALLREDUCE at code/schemes/OCFD_SetHybrid.f90:64):
   *   0 = ordered: one sequential sum in (k,j,i) order, bit-identical to a
   *       single-rank reference run (slow: one wavefront);
   *   1 (default) = fast and certified: a parallel tree sum; every
   *       interface whose normalised sensor lies within the rounding bound
   *       of the two sums (relative eta, see ocfd_hybrid_certificate) of a
   *       threshold is counted, and when the count is not zero the
   *       sequential sum and the masks are recomputed on the device.  The
   *       masks and Rhybrid are therefore the ordered mode's, bit for bit. */
  OCFD_OPT_HYBRID_SUM_MODE = 1,
  /* 1 = capture the RK3 step in a hipGraph on first use and replay it. */
  OCFD_OPT_USE_GRAPH = 2,
  /* 1 = bracket every kernel launch with HIP events (per-kernel timing). */
  OCFD_OPT_KERNEL_TIMING = 3,
  /* k planes walked by one block of the viscous-flux and update kernels
   * (their z-stencil windows stay in registers); 0 = one thread per cell;
   * -1 (default) = 4 for N_SPEC = 1, one thread per cell for N_SPEC > 1.
   * Results are bitwise independent of the value. */
  OCFD_OPT_COLUMN_PLANES = 4,
  /* 1 (default) = the viscous-flux preparation of each stage runs on a
   * second stream beside the inviscid sweeps; 0 = one stream.  Results are
   * bitwise independent of the value. */
  OCFD_OPT_CONCURRENT = 5,
  /* Widens the certificate band of OCFD_OPT_HYBRID_SUM_MODE = 1 to
   * eta * 2^value (0 <= value <= 60; default 0).  Tests use it to force the
   * ordered recomputation; the masks stay the ordered ones for any value. */
  OCFD_OPT_HYBRID_CERT_WIDEN = 6,
  /* 1 = the stepping kernels (viscous fluxes, update, pressure sensor)
   * evaluate the metrics of comput_Jacobian3d from per-row tables of the
   * nine coordinate derivatives instead of reading the 19 metric arrays,
   * when ocfd_set_mesh found that the tables give every interior metric bit
   * for bit (GRID1D meshes); 0 (default) = always read the arrays.  The
   * tables move fewer bytes but cost ten FP64 divisions per cell in the
   * viscous preparation, which made the stepping kernels slower on MI355X
   * (flat plate 640x120x64: 558/578 us vs 487/484 us per launch).  Results
   * are bitwise independent of the value. */
  OCFD_OPT_METRIC_TABLES = 7
};

/* ---- lifetime ---------------------------------------------------------- */
/* Replaces allocate_flow_data / allocate_inviscous_data / allocate_vicous_data
 * and set_parameters (code/head/OCFD_allocate_data.f90:2-68,
 * code/flow-solver/OCFD_inviscous.f90:5-31, code/flow-solver/OCFD_viscous.f90:5-16,
 * code/mpi-io/OCFD_readpara.f90:347-431).  Device arrays get the reference's
 * initial values (d=u=v=w=p=Et=1, T=300 in the halo, metrics 1). */
int32_t ocfd_init(const ocfd_cfg* cfg);
int32_t ocfd_finalize(void);
const char* ocfd_last_error(void);
int32_t ocfd_set_option(int32_t key, int32_t value);

/* Species thermo (read_Spec, code/Chemical/CHEM_thermdyn.f90:16-51):
 * mi[n_spec] (kg/mol); coef[k][j][7] = (A,B,C,D,E,F,G) of range j of
 * species k (n_spec x n_tct x 7); tct[n_tct-1]; tct_max.  For a perfect
 * gas only mi is used (coef may be NULL).  The energy offsets Ect that keep
 * the enthalpy continuous across ranges are derived here (:42-46). */
int32_t ocfd_set_species(int32_t n_tct, const double* tct, double tct_max,
                         const double* mi, const double* coef);

/* Transport fits (read_transport / Trans%mu_an, k_an, D_an,
 * code/Chemical/TRANSPORT_modules.f90:80-253), used when Flag_Vis,
 * Flag_Cond or Flag_Diff is 1: mu_an[n_spec][4], k_an[n_spec][4] and
 * D_an[n_spec][n_spec][4] (pair (k,j) read for k <= j) of
 * exp(a1 + a2 lnT + a3 lnT^2 + a4 lnT^3).  A set whose flag is 0 may be NULL. */
int32_t ocfd_set_transport(const double* mu_an, const double* k_an, const double* D_an);

/* Jet inflow data flow_inlet2d (ocfd_bc_jet, code/boundary-condition/
 * OCFD_BC_jet.f90:2-96): column-major (ny, nz, 5+N_SPEC) with the variables
 * d, u, v, w, T, di(1..N_SPEC).  BC_Para(1) = 1 takes the pressure of i=2
 * where the inflow is subsonic. */
int32_t ocfd_set_bc_jet(const double* inlet2d);

/* Mesh: coordinates Axx, Ayy, Azz (halo arrays; only the interior is read).
 * Computes the halo coordinate fill, periodic span shift, the 9 metrics and
 * Jacobian and the Akx1 = Akx/Ajac (etc.) arrays on the device
 * (init_mesh / comput_Jacobian3d / updata_Axyz_periodic,
 * code/flow-solver/OCFD_mesh.f90:1-185).  span[9] = Periodic_ISpan(1:3),
 * Periodic_JSpan(1:3), Periodic_KSpan(1:3). */
int32_t ocfd_set_mesh(const double* axx, const double* ayy, const double* azz,
                      const double* span);
/* 1 when the stepping kernels use the metric tables (OCFD_OPT_METRIC_TABLES),
 * 0 when they read the metric arrays.  No counterpart in the reference. */
int32_t ocfd_metric_tables_in_use(void);

/* ---- state transfer ---------------------------------------------------- */
int32_t ocfd_upload_field(int32_t field, const double* host);
int32_t ocfd_download_field(int32_t field, double* host);
int32_t ocfd_download_mask(int32_t dir, uint8_t* host);

/* Boundary-layer BC data (ocfd_bc_BoundaryLayer and its init routines,
 * code/boundary-condition/OCFD_BC_Boundarylayers.f90:20-243,247-489; the
 * boundary-layer BC runs with N_SPEC = 1, so 5+N_SPEC = 6 variables
 * d,u,v,w,T,di), in the layout of BC%flow_inlet*d / flow_upper*d, column-major
 * with the variable index slowest, already rescaled so sum(di)=d
 * (init_bcdata_boundary, :247-430):
 *   inlet  by BC_Para(1) (BcData_inlet): 0 -> [6] (flow_inlet0d),
 *          1 -> (ny,6) (flow_inlet1d), 2 -> (ny,nz,6) (flow_inlet2d);
 *   upper  by BC_Para(2) (BcData_upper): -1 -> none (may be NULL; the upper
 *          face keeps its values), 0 -> [6], 1 -> (nx,6), 2 -> (nx,nz,6);
 *   nonreflect_upper[nx*nz]    (int32, 1 = non-reflecting, 0 = Dirichlet);
 *   wall_pertb[nx*nz]          (fx*gz of the Rai blow/suction).
 * The random phases of the reference are drawn by the caller. */
int32_t ocfd_set_bc_boundarylayer(const double* inlet, const double* upper,
                                  const int32_t* nonreflect_upper,
                                  const double* wall_pertb);
/* ht(tt) of get_ht_multifrequancy (OCFD_BC_Boundarylayers.f90:594-621),
 * evaluated by the caller once per time step (it depends only on tt). */
int32_t ocfd_set_bc_time(double ht);

/* ---- the hot path ------------------------------------------------------ */
/* One TVD-RK3 step of size dt1 (flow_time_advance,
 * code/flow-solver/OCFD_Time_Adv.f90:7-92).  N_SPEC == 1, or N_SPEC > 1 with
 * the species convection/diffusion and Scalar_time_adv_RK3 (:64-71,134-166). */
int32_t ocfd_flow_time_advance(double dt1);

/* One step of NS_solver (code/flow-solver/OCFD_NS_Solver.f90:26-37): flow
 * dt/2, then (reaction /= 0) nsub chemistry substeps of dt/nsub
 * (Chemical_time_advance with CHEM_TimeAdv = mode) and comput_T_P, then
 * flow dt/2.  The chemistry needs ocfd_chem_set_mechanism and N_SPEC = 9. */
int32_t ocfd_ns_step(double dt, int32_t nsub, int32_t mode, int32_t reaction);
/* Chemical_time_advance (OCFD_Time_Adv.f90:169-196) on every interior cell. */
int32_t ocfd_chemical_time_advance(double dt2, int32_t mode);
/* CFL time step (no reference counterpart: the reference reads dt from the
 * namelist, OCFD_NS_Solver.f90:20): dt = cfl / max over interior cells of
 * sum_dir (|A_dir . u| + c |A_dir|) / h_dir, with A_dir the metric rows
 * (Akx.., Aix.., Asx..), c = sqrt(gamma p/d), gamma = p/Et + 1 and
 * h_dir = 1/(n_dir - 1).  Deterministic (a maximum); synchronises. */
int32_t ocfd_cfl_dt(double cfl, double* dt);

/* Integrator of ocfd_chemical_time_advance / ocfd_ns_step (no reference
 * counterpart beyond OCFD_CHEM_FIXED): OCFD_CHEM_FIXED = chemical_timeAdv
 * with CHEM_TimeAdv = mode (default); OCFD_CHEM_ADAPTIVE = the adaptive
 * point-implicit controller of ocfd_chem_adaptive_batch with these
 * controls and scheduler (mode ignored); OCFD_CHEM_SURROGATE =
 * ocfd_chem_surrogate_batch with the adaptive integrator as fallback
 * (needs a configured surrogate network). */
enum { OCFD_CHEM_FIXED = 0, OCFD_CHEM_ADAPTIVE = 1, OCFD_CHEM_SURROGATE = 2 };
int32_t ocfd_set_chem_method(int32_t method, double rtol_y, double y_floor, double rtol_t, int32_t max_sub,
                             int32_t scheduler);
/* comput_T_P (code/flow-solver/Comput_flow_variables.f90:32-46), N_SPEC > 1. */
int32_t ocfd_comput_T_P(void);
/* Species pieces (N_SPEC > 1): the right-hand sides du_scaler_invis +
 * du_scaler_viscous of every species into OCFD_FIELD_DUS and Qp into
 * OCFD_FIELD_QP (code/flow-solver/OCFD_scaler_inviscous.f90:1-84,
 * OCFD_scaler_viscous.f90:4-79), and the same followed by
 * Scalar_time_adv_RK3 of stage krk. */
int32_t ocfd_du_species(void);
int32_t ocfd_species_time_adv(int32_t krk, double dt);
/* One species at a time, for a host that keeps the reference's species loop
 * (OCFD_Time_Adv.f90:64-71); m is the 1-based species index.  Called as
 *   for m = 1..N_SPEC: ocfd_du_scaler_invis(m); [ocfd_du_scaler_viscous(m);]
 *                      ocfd_scalar_time_adv_rk3(krk, m, dt);
 * they give ocfd_species_time_adv(krk, dt) bit for bit.
 *   ocfd_du_scaler_invis(m)        du_scaler_invis(fs,Dus), OCFD_scaler_inviscous.f90:1-84:
 *                                  Dus(m) = inviscid right-hand side of species m;
 *   ocfd_du_scaler_viscous(m)      du_scaler_viscous(K_SPEC,fs,Dus), OCFD_scaler_viscous.f90:4-79:
 *                                  Dus(m) += diffusive right-hand side; Qp zeroed at m = 1
 *                                  and accumulated for later m (needs IF_Viscous = 1);
 *   ocfd_scalar_time_adv_rk3(krk, m, dt)  Scalar_time_adv_RK3(KRK,m,dus,dt),
 *                                  OCFD_Time_Adv.f90:134-166: di(m) from din(m) and Dus(m).
 * Dus(m) is the slice m of OCFD_FIELD_DUS (the reference reuses du(:,:,:,1)). */
int32_t ocfd_du_scaler_invis(int32_t m);
int32_t ocfd_du_scaler_viscous(int32_t m);
int32_t ocfd_scalar_time_adv_rk3(int32_t krk, int32_t m, double dt);

/* Per-stage pieces, for a host that keeps the reference's own stage loop
 * (OCFD_Time_Adv.f90:33-91).  Called in this order they reproduce
 * ocfd_flow_time_advance exactly:
 *   ocfd_save_stage0();
 *   for krk = 1..3:
 *     ocfd_stage_begin(krk); [N_SPEC > 1: ocfd_species_time_adv(krk, dt1);]
 *     ocfd_du_inviscous(); ocfd_du_viscous();
 *     ocfd_time_adv_rk3(krk, dt1); ocfd_comput_flow_variables(); ocfd_bc(); */
int32_t ocfd_save_stage0(void);             /* fn=f, din=di (OCFD_Time_Adv.f90:10-29) */
/* Stage prologue, OCFD_Time_Adv.f90:35-62: comput_Amu_Amk_AMD, the halo
 * exchanges (periodic wrap, nothing to do), di=d (N_SPEC=1) and, at krk=1
 * with the hybrid scheme, comput_Rhybrid. */
int32_t ocfd_stage_begin(int32_t krk);
int32_t ocfd_comput_amu_amk(void);          /* Comput_flow_variables.f90:69-98 */
int32_t ocfd_comput_rhybrid(void);          /* OCFD_SetHybrid.f90:17-143 + masks */
int32_t ocfd_du_inviscous(void);            /* OCFD_inviscous.f90:47-202 */
int32_t ocfd_du_viscous(void);              /* OCFD_viscous.f90:27-131 */
int32_t ocfd_time_adv_rk3(int32_t krk, double dt); /* OCFD_Time_Adv.f90:98-130 */
int32_t ocfd_comput_flow_variables(void);   /* Comput_flow_variables.f90:3-46 */
int32_t ocfd_bc(void);                      /* OCFD_BC.f90:2-188 */
/* filtering (code/others/OCFD_filtering.f90:4-76), which NS_solver calls
 * after advancing Istep (code/flow-solver/OCFD_NS_Solver.f90:45-48): every
 * filter m (1..NFiltering) with mod(istep, Filter(1,m)) == 0 is applied in x,
 * y, z order where Filter(2..4,m) = 1, over the index box Filter(5..10,m)
 * (clipped to the domain), with Filter(11,m) = 1 the Bogey & Bailly 9-point
 * filter (filter_x3d, :78-116) or 2 the shock-capturing filter
 * (filter_x3d_shock, :196-246, which first renews the primitives by
 * comput_flow_variables), amplitude s0 = Filter(12,m) and threshold
 * rth = Filter(13,m).  filter = Para%Filter(100,10), column-major, as the
 * reference's module holds it.  Only f changes (and, for the shock filter,
 * the primitives), as in the reference. */
int32_t ocfd_filtering(int32_t istep, int32_t nfiltering, const double* filter);

/* OCFD2d_flux1 (side 0, flux+) and OCFD2d_flux2 (side 1, flux-) over npen
 * pencils at once (code/flow-solver/OCFD_inviscous.f90:209-281, 286-362).
 * Device pointers, run on the library stream:
 *   v   npen x (n + 2*LAP) doubles, pencil p holding f1d(1-LAP:n+LAP);
 *   hh  npen x (n + 1) doubles, receives hh(0:n);
 *   scm npen x (n + 1) bytes, Scm_Hy(0:n) (1 UD7L, 2 WENO7, 3 WENO5); read
 *       only by the hybrid scheme, may be null otherwise.  An interior
 *       interface whose hybrid mask is outside 1..3 gets NaN.
 * Host pointers: bound_index[2] (Bound_index, 0 or 1) and
 * scheme_boundary[2] (Scheme_boundary, -1, 0 or 1).  scheme is one of the
 * inviscid OCFD_SCHEME_* values; the UD7L coefficients are the context's
 * (OCFD_Config.ud7l_diss).  Results are bit-identical to the sweep's
 * reconstructions and to the reference subroutines. */
int32_t ocfd_flux1d_batch(int32_t side, int32_t scheme, int64_t npen, int32_t n, const double* v, double* hh,
                          const uint8_t* scm, const int32_t* bound_index, const int32_t* scheme_boundary);

/* Largest number of cells repaired by handle_NegativeEt in one
 * comput_flow_variables since the last ocfd_save_stage0 / step start
 * (code/others/OCFD_handle_NegativeT.f90:1-20).  More than 10 makes the next
 * synchronising call (and the next step) return an error, as the reference
 * stops.  Stepping calls are asynchronous: errors surface at
 * ocfd_synchronize, a download or the next step. */
int32_t ocfd_negative_et_count(void);

/* Certificate of the last fast sensor mean (OCFD_OPT_HYBRID_SUM_MODE = 1,
 * sensor type 2; no reference counterpart).  Synchronises.  uncertain: the
 * interfaces whose normalised max(R_i, R_i+1) fell within the relative band
 * eta of Hybrid_para(2) or (3), where the fast and the sequential mean could
 * disagree (0: the fast masks are certified equal to the ordered ones; > 0:
 * the library recomputed the ordered mean and masks).  margin: the smallest
 * |R / theta - 1| over all interfaces and both thresholds.  eta: the band in
 * use.  Any pointer may be NULL.  Fails when no certified mean has run. */
int32_t ocfd_hybrid_certificate(int64_t* uncertain, double* margin, double* eta);

/* ---- the reference's files ----------------------------------------------
 * Fortran unformatted sequential files (each record framed by a 4-byte
 * length before and after it), as the reference writes them with one rank.
 * ocfd_read_flow_data: read_flow_data + init3d (code/mpi-io/OCFD_IO.f90:5-33,
 *   code/flow-solver/OCFD_init.f90:1-25): a record (Istep int32, tt real8),
 *   then d, u, v, w, T, di(1..N_SPEC), each as nz records of nx*ny reals,
 *   into the interior of the resident arrays; Rescaling_di; f from the
 *   primitives with Et = comput_E(di, T).  As in NS_solver, call
 *   ocfd_comput_flow_variables and ocfd_bc afterwards.
 * ocfd_save_flow_data: save_flow_data (OCFD_IO.f90:38-64), the same layout
 *   (the reference names the file OCFD-comb%08d.dat after Istep).
 * ocfd_read_mesh: init_mesh (code/flow-solver/OCFD_mesh.f90:1-57) reading
 *   OCFD-grid.dat for Iflag_Gridtype 10 (GRID1D: x(nx), y(ny), z(nz)),
 *   20 (GRID2D_PLANE: x(nx,ny), y(nx,ny), z(nz)), 21 (GRID2D_AXIAL_SYMM:
 *   x(nx,ny), R(nx,ny), seta(nz); y = R cos seta, z = R sin seta),
 *   30 (GRID3D: x, y, z as nz records each) or 31 (GRID_AND_JACOBIAN3D:
 *   GRID3D plus Akx, Aky, Akz, Aix, Aiy, Aiz, Asx, Asy, Asz, Ajac, used
 *   instead of comput_Jacobian3d); span[9] = Periodic_ISpan, _JSpan, _KSpan.
 *   Then the metrics as ocfd_set_mesh. */
int32_t ocfd_read_flow_data(const char* path, int32_t* istep, double* tt);
int32_t ocfd_save_flow_data(const char* path, int32_t istep, double tt);
int32_t ocfd_read_mesh(const char* path, int32_t gridtype, const double* span);

/* ---- diagnostics (device reductions of the resident state) --------------
 * show_flow_msg (code/flow-solver/OCFD_Time_Adv.f90:200-239): E0[0] = mean
 * over the interior of f5/Ajac (total energy), E0[1] = mean of
 * d(u^2+v^2+w^2)/2/Ajac (kinetic energy).
 * ana_box (code/others/OCFD_ANA.f90:197-303), the statistics of isotropic
 * turbulence from the current primitives and Amu (call
 * ocfd_comput_amu_amk first if no stage has run since the state changed):
 * out[0..17] = Re_lamda, Re_lamdax, Amt, vv_rms, u_rms, v_rms, w_rms,
 * u_skewness, u_flatness, ux_skewness, ux_flatness, E_kinetic, d_rms,
 * T_aver, Amu_aver, d_aver, Ux_rms, Alamdax (the reference's
 * statistics.dat columns after tt).  The sums are deterministic trees, so
 * they differ from the reference's sequential sums by summation order only.
 * Both synchronise. */
int32_t ocfd_show_flow_msg(double* E0);
int32_t ocfd_ana_box(double* out);

/* ---- tile autotuner (no reference counterpart) ---------------------------
 * The inviscid sweeps come in several workgroup tile shapes (pencils per
 * workgroup x cells per pencil).  ocfd_autotune times each shape per
 * direction on the current state with HIP events (reps launches each) and
 * keeps the fastest; ocfd_set_tile forces one.  Results are bit-identical
 * for every shape.  ocfd_tile_choice returns the chosen variant index and
 * its shape, and the measured ms per launch of every variant (4 floats). */
int32_t ocfd_autotune(int32_t reps);
int32_t ocfd_set_tile(int32_t dir, int32_t variant);
int32_t ocfd_tile_choice(int32_t dir, int32_t* np, int32_t* ts, float* ms_per_variant);

/* ---- timing ------------------------------------------------------------ */
/* Synchronises the library stream. */
int32_t ocfd_synchronize(void);
/* The library's HIP stream (hipStream_t) as an opaque pointer. */
void*   ocfd_stream(void);
/* Per-kernel accumulated time (ms) and launch count since the last reset,
 * recorded when OCFD_OPT_KERNEL_TIMING is on.  Kernel names are returned
 * through ocfd_kernel_name(i) for i < ocfd_kernel_count(). */
int32_t ocfd_kernel_count(void);
const char* ocfd_kernel_name(int32_t i);
int32_t ocfd_kernel_stats(int32_t i, double* total_ms, int64_t* launches);
int32_t ocfd_kernel_stats_reset(void);

/* ==== finite-rate chemistry: batched 0-D cells ============================
 * Independent of the flow context above (its own stream and error string).
 * A "cell" is the 0-D state chemical_timeAdv integrates at constant volume:
 * partial densities rho_k (kg/m3) and the energy E (J/m3).  All arrays are
 * DEVICE pointers: rho is species-major [n_spec][n] (rho[k*n + c]), E and dt
 * are [n].  Calls are asynchronous on ocfd_chem_stream(); the state is
 * updated in place. */

/* Mechanism and thermo data in the layout read_Reac / read_Spec fill
 * (CHEM_reactions.f90:15-71, CHEM_thermdyn.f90:16-51): af_type, third_body,
 * fc_troe [nr]; arrhenius [nr][2][3] (A, b, E for the low / second set);
 * vf, vr, eff [nr][ns]; tct [n_tct-1], tct_max; mi [ns];
 * coef [ns][n_tct][7] (NASA-7 A..G per temperature range).  This build
 * takes ns = 9, nr <= 32, n_tct <= 4.  ReReac (Reaction_rate_coef,
 * CHEM_reactions.f90:130-197): 0 = reverse rates from Gibbs equilibrium
 * constants (comput_KjX, :76-103); 1 = Arrhenius reverse rates; 2 = k_f / Kc
 * with the Park curve fit of Kc floored at 1e-30 (ParkCurveFit, :113-118).
 * For 1 and 2 the batch calls fail until ocfd_chem_set_reverse_rates has
 * given the reverse data. */
int32_t ocfd_chem_set_mechanism(int32_t ns, int32_t nr, int32_t rereac, const int32_t* af_type,
                                const int32_t* third_body, const double* fc_troe, const double* arrhenius,
                                const int32_t* vf, const int32_t* vr, const double* eff, int32_t n_tct,
                                const double* tct, double tct_max, const double* mi, const double* coef);
/* Reverse-rate data of ReReac = 1 or 2: rev [nr][5], per reaction the line
 * read_Reac reads after the Arrhenius lines (CHEM_reactions.f90:46-47):
 * (A_r, b_r, E_r, -, -) for ReReac = 1, (A1r .. A5r) for ReReac = 2. */
int32_t ocfd_chem_set_reverse_rates(const double* rev);

/* Chemical_time_advance over a batch (OCFD_Time_Adv.f90:169-196 calling
 * chemical_timeAdv, CHEM_reactions.f90:374-390, per cell): nsub substeps of
 * dt[c]/nsub each; mode = CHEM_TimeAdv (1 Euler, 2 RK2, 3 RK3, -1
 * point-implicit Euler).  One cell per thread. */
int32_t ocfd_chem_timeadv_batch(int64_t n, double* rho, double* E, const double* dt, int32_t mode, int32_t nsub);

/* Adaptive substepping (no reference counterpart; the reference uses a fixed
 * Chemical_sub_step).  Point-implicit Euler substeps (chemical_Euler_implicit)
 * cover dt[c].  Each attempt evaluates the rates and Jacobian (negative
 * partial densities counted as zero) and proposes
 *   h = min(dt - t, h_cap, min_k s_k/|w_k|, s_T/|q|),
 * with s_k = rtol_y (max(rho_k,0) + y_floor rho), s_T = rtol_t T rcv,
 * rcv = sum_k 2.5 R_k rho_k, w_k the species mass rates and q = sum_k w_k R_k F_k
 * (F_k the formation coefficient of update_Et).  The increment dd of the
 * linear solve is accepted when |dd_k| <= 2 s_k and |sum_k dd_k R_k F_k| <= 2 s_T;
 * then h_cap = 2h, else the attempt is discarded and h_cap = h/2.  A cell
 * stops after max_sub accepted or max_sub rejected attempts; one that has
 * not reached dt by then counts as unfinished (ocfd_chem_unfinished).
 * nsub_out [n] (device, may be NULL) receives the accepted substeps per
 * cell.  scheduler 0: one cell per thread; 1: a persistent kernel whose
 * lanes take cells from a device work queue; 2: the work queue, where a cell
 * still running after a budget of attempts (OCFD_CHEM_OPT_COOP_BUDGET) moves
 * to a cooperative kernel that
 * integrates one cell per wavefront (lower latency for the few cells that
 * need hundreds of substeps).  All three do the same per-cell arithmetic and
 * give bit-identical results. */
int32_t ocfd_chem_adaptive_batch(int64_t n, double* rho, double* E, const double* dt, double rtol_y, double y_floor,
                                 double rtol_t, int32_t max_sub, int32_t* nsub_out, int32_t scheduler);

/* ocfd_chem_adaptive_batch with scheduler 1 (the per-lane work queue) and
 * the longest-first cell order, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its work counters, sort keys, index lists and radix-sort workspace live in
 * a static device array of about 32.5 MiB that each GPU the library runs on
 * holds until the process ends; n above 2^20 cells is refused with an
 * error.  Results are bit-identical to ocfd_chem_adaptive_batch.  Every call
 * on one GPU shares that array (the call clears its counters first), so calls
 * that overlap on one GPU race.  The chemistry stream and mechanism are those
 * of ocfd_chem_set_stream / ocfd_chem_set_mechanism (which allocates, at
 * set-up). */
int32_t ocfd_chem_adaptive_batch_noalloc(int64_t n, double* rho, double* E, const double* dt, double rtol_y,
                                         double y_floor, double rtol_t, int32_t max_sub, int32_t* nsub_out);

/* Chemistry options: OCFD_CHEM_OPT_COOP_BUDGET = attempts a cell may take on
 * one lane before scheduler 2 moves it to the cooperative kernel (default
 * 128, the fastest of 16..1024 on a 2^21-cell jet-flame batch);
 * OCFD_CHEM_OPT_FALLBACK_SCHEDULER = scheduler (1 or 2) of the
 * surrogate's fallback (default 2); OCFD_CHEM_OPT_COOP_DIRECT_MAX = a
 * fallback list of at most this many cells hands every cell that needs more
 * than one attempt to the cooperative kernel (default 16384);
 * OCFD_CHEM_OPT_ORDER = 1 (default): schedulers 1 and 2 of
 * ocfd_chem_adaptive_batch, and the surrogate's fallback, take the cells in
 * order of decreasing dt T^2 (an estimate of each cell's work; stable device
 * radix sort), so the longest integrations start first (batches of 2^31
 * cells or more run in index order; scheduler 2 refuses them); 0: index
 * order.  Either order gives bit-identical results. */
enum { OCFD_CHEM_OPT_COOP_BUDGET = 1, OCFD_CHEM_OPT_FALLBACK_SCHEDULER = 2, OCFD_CHEM_OPT_COOP_DIRECT_MAX = 3,
       OCFD_CHEM_OPT_ORDER = 4 };
int32_t ocfd_chem_set_option(int32_t key, int32_t value);
/* The chemistry work counters of the last batch call (synchronises):
 * [0] per-lane queue head, [1] unfinished cells, [2] surrogate fallbacks,
 * [3] cells handed to the cooperative kernel, [4] its queue head. */
int32_t ocfd_chem_counters(int64_t* out, int32_t count);
/* Cells left unfinished by the last ocfd_chem_adaptive_batch (synchronises
 * the chemistry stream); -1 on error. */
int64_t ocfd_chem_unfinished(void);
int32_t ocfd_chem_synchronize(void);
void*   ocfd_chem_stream(void);
/* Launch the chemistry kernels on this hipStream_t from now on (NULL is the
 * null stream).  ocfd_chem_stream() returns the stream in use; the library's
 * own stream is the one in use after ocfd_chem_set_mechanism. */
int32_t ocfd_chem_set_stream(void* stream);
const char* ocfd_chem_last_error(void);

/* ==== learned chemistry surrogate (no reference counterpart) ==============
 * A multilayer perceptron held by the library (one network per process):
 * dims[0] -> ... -> dims[n_layers], GELU (tanh form) between layers, BF16
 * MFMA GEMMs with FP32 accumulation, FP32 master weights, Adam on the
 * device.  Parameters travel in a flat float layout: per layer W [out][in]
 * row-major, then b [out].  Batch arrays are device pointers, float32,
 * row-major [B][width].  Calls run on the MLP stream (ocfd_mlp_set_stream). */
int32_t ocfd_mlp_create(int32_t n_layers, const int32_t* dims);
int32_t ocfd_mlp_destroy(void);
int64_t ocfd_mlp_param_count(void);
int32_t ocfd_mlp_set_params(const float* flat_host);
int32_t ocfd_mlp_get_params(float* flat_host);
int32_t ocfd_mlp_set_stream(void* stream);
int32_t ocfd_mlp_synchronize(void);
const char* ocfd_mlp_last_error(void);
/* y [B][dout] = net(x [B][din]) */
int32_t ocfd_mlp_forward(int64_t B, const float* x, float* y);
/* mean squared error over the B x dout outputs and its gradient (flat
 * layout, device; either output may be NULL) */
int32_t ocfd_mlp_gradients(int64_t B, const float* x, const float* y, float* grad, float* loss);
/* one bias-corrected Adam step on the batch; loss (device) may be NULL */
int32_t ocfd_mlp_train_step(int64_t B, const float* x, const float* y, float lr, float beta1, float beta2, float eps,
                            float* loss);
/* C [M][N] (ldc) = op(A) op(B) with bfloat16 device operands, op = transpose
 * when ta / tb; lda, ldb multiples of 8; split-K over `splits` chunks */
int32_t ocfd_mlp_gemm_bf16(int32_t M, int32_t N, int32_t K, int32_t ta, int32_t tb, const void* A, int64_t lda,
                           const void* B, int64_t ldb, float* C, int64_t ldc, int32_t splits);

/* Surrogate configuration: 152 doubles, in this order:
 *   in_shift[12], in_scale[12]  inputs x = (f - in_shift) * in_scale of the raw
 *                               features f = [T, ln p, cbrt(Y_0..Y_8), ln dt];
 *   out_scale[9]                dY_k = z_k * out_scale_k (z = network output);
 *   T_lo, T_hi, p_lo, p_hi, dt_lo, dt_hi   domain; cells outside fall back;
 *   y_neg_tol, elem_tol, T_out_lo, T_out_hi, dy_max  acceptance bounds
 *                               (dy_max bounds every |dY_k|; +inf: no bound);
 *   P[9][9]                     dY <- P dY (projection onto element balance);
 *   Cel[3][9]                   element moles per kg of each species (H, O, N). */
int32_t ocfd_chem_surrogate_set(const double* cfg, int32_t count);
/* Surrogate step over a batch (layout of ocfd_chem_adaptive_batch): the
 * fused MLP kernel updates every accepted cell; with fallback != 0 the cells
 * it rejects are then integrated by the plain adaptive FP64 integrator
 * (work-queue scheduler, controller arguments as ocfd_chem_adaptive_batch).
 * flag_out [n] (device, may be NULL): 0 surrogate, else the failed checks
 * (1 outside the training domain, 2 element balance, 4 negative mass
 * fraction, 8 temperature bounds, 16 an increment above dy_max) and the cell
 * went to the fallback.  With OCFD_CHEM_OPT_ORDER = 1 (the default) the
 * call reads the number of rejected cells back (it synchronises the
 * chemistry stream once), launches nothing more when there is none and
 * otherwise sorts only those cells longest first.  The
 * network must be 12 -> 256 x k -> 9. */
int32_t ocfd_chem_surrogate_batch(int64_t n, double* rho, double* E, const double* dt, int32_t fallback,
                                  double rtol_y, double y_floor, double rtol_t, int32_t max_sub, int32_t* flag_out);
/* Tile-shape autotuner of the surrogate network kernel (no reference
 * counterpart): each shape (cells per workgroup x waves: 128x4 with each
 * wave on 64 output columns of all rows, 128x4 with the waves as 2 x 2
 * blocks of 64 rows x 128 columns, 64x4, 96x4) runs reps times on scratch copies of the batch (rho, E are not
 * modified), timed with HIP events; the fastest is kept.  Results are
 * bit-identical for every shape.  ms_out (may be NULL) gets the mean ms per
 * launch of each shape (4 floats).  Synchronises. */
int32_t ocfd_chem_surrogate_autotune(int64_t n, const double* rho, const double* E, const double* dt, int32_t reps,
                                     float* ms_out);
/* Returns the shape index in use and its cells per workgroup and waves;
 * variant >= 0 forces that shape first. */
int32_t ocfd_chem_surrogate_tile(int32_t variant, int32_t* rows, int32_t* waves);
/* cells of the last ocfd_chem_surrogate_batch handed to the fallback
 * (synchronises); -1 on error */
int64_t ocfd_chem_surrogate_fallbacks(void);
/* per-cell T [n] (mode 0, f64), raw features [n][12] (mode 1, f64) or
 * normalised surrogate inputs [n][12] (mode 2, f32) of a batch */
int32_t ocfd_chem_cell_features(int64_t n, const double* rho, const double* E, const double* dt, int32_t mode,
                                void* out);
/* normalised training targets y [n][9] = ((rho1_k - rho0_k)/rho0) / out_scale_k */
int32_t ocfd_chem_surrogate_targets(int64_t n, const double* rho0, const double* rho1, float* y);

#ifdef __cplusplus
}
#endif

#endif /* OCFD_HIP_H */
