"""Algorithmic operation and byte counts of the flow step.

Counting rule: every +, -, *, / and sqrt of the reference arithmetic is one
flop; abs, min, max, sign and comparisons are free; a value computed twice
in the reference source (e.g. c*ss in split_Stager_Warming) counts twice.
Counts are per evaluation and were made line by line from the reference
formulas (which the kernels in csrc/ follow operation for operation):

  SPLIT_SW    split_Stager_Warming per point (OCFD_flux_split.f90:22-70):
              gamma 2, tmp1 2, tmp3 4, ss 6, ss1 1, ak 3, vs 5, E2/E3 4,
              E+ 6, E- 3, tmp0 2, uc..wc 12, vvc1/vvc2 12, vv 7, W2 2,
              fp 34, fm 34                                          = 139
  CC          comput_cc per point (Comput_flow_variables.f90:57-58)   = 5
  SPLIT_LLF   split_Local_LaxFriedrichs per point (:80-101)           = 61
  WENO7       OCFD_weno7P per interface (OCFD_Schemes.f90:84-121):
              S1x 26, S2x 18, S3x 16, S0..S3 40, a0..a3 12, am 3,
              q0..q3 28, hh 8                                       = 151
  WENO5       OCFD_weno5P per interface (:14-24)                     = 65
  UD7L        8 mul + 7 add                                          = 15
  OMP6        linear part + MP test on smooth data (:225-228)         = 25
  DIFF_X/Y/Z  (hh1(i)-hh1(i-1)+hh2(i)-hh2(i-1))*h_1 per variable,
              +1 add (y), +1 add +1 mul (z)                   = 4 / 5 / 6
  VISC_PREP   12 CD6 derivatives (8 each) 96, chain rule 12*5 60, div 2,
              stresses 15, E1..E3 3*9 27, Ev 12*5 60, Sutherland 9,
              Cp and kappa 5                                         = 274
  VISC_DIV    12 CD6 derivatives 96 + 4*(2 add + 1 mul + 1 add)      = 112
  RK          5 * (3 mul + 2 add) (stage 2/3 form)                   = 25
  PRIMS       d,u,v,w,Et 10, rescale 3, T 4, p 3                      = 20

The algorithmic HBM bytes of a step are the compulsory state traffic of an
ideal fused RK3 step: 64 * N_var bytes per cell with N_var = 5 for one
species (read U0 / write U1; read U1, U0 / write U2; read U2, U0 / write U3
= 8 reads/writes of 5 doubles).
"""

SPLIT_SW = 139
SPLIT_LLF = 61
CC = 5
WENO7 = 151
WENO5 = 65
UD7L = 15
OMP6 = 25
DIFF = (4, 5, 6)
VISC_PREP = 274
VISC_DIV = 112
RK = 25
PRIMS = 20

RECON = {1: UD7L, 2: WENO7, 3: WENO5}  # hybrid mask value -> flops


def recon_flops(scheme, mask_fractions=None):
    """Average flops of one reconstruction for an inviscid scheme code."""
    if scheme == 1:
        return WENO5
    if scheme == 2:
        return WENO7
    if scheme == 3:
        return OMP6
    if scheme == 4:
        return UD7L
    f = mask_fractions or {1: 0.0, 2: 1.0, 3: 0.0}
    return sum(RECON[k] * f.get(k, 0.0) for k in (1, 2, 3))


def sweep_flops_per_cell(direction, scheme, split=1, mask_fractions=None):
    """Flops per cell of one inviscid sweep: split + comput_cc at one point,
    10 reconstructions at one interface, 5 differences."""
    sp = (SPLIT_SW if split == 1 else SPLIT_LLF) + CC
    return sp + 10 * recon_flops(scheme, mask_fractions) + 5 * DIFF[direction]


def step_flops_per_cell(scheme, split=1, mask_fractions=None, viscous=True):
    """Flops per cell of one RK3 step (3 stages)."""
    stage = sum(sweep_flops_per_cell(d, scheme, split, mask_fractions) for d in range(3))
    if viscous:
        stage += VISC_PREP + VISC_DIV
    stage += RK + PRIMS
    return 3 * stage


def step_bytes_per_cell(n_var=5):
    return 64 * n_var


# Reacting 9-species jet (WENO5, Steger-Warming, CD6, Chemkin/Wilke transport,
# mixture-averaged diffusion, NASA-fit real gas), per cell per RK3 step, same
# counting rule.  Per stage: flow sweeps 3 * (SPLIT_SW + CC + 10 * WENO5 + 5 *
# DIFF) = 2457; species sweeps 9 species * 3 directions * (scalar split 11 +
# 2 * WENO5 + DIFF) = 3906; species diffusion 9 * (3 CD6 gradients 24, chain
# rule 15, rho*D*grad c 9, metric projection 15, enthalpy h_k 12 and Qp 6,
# divergence 27) = 972; transport (mole fractions 27, 9 Chemkin mu fits 63,
# Wilke 81 Phi * 8 = 648, conductivity fits and mixing 103, 45 binary D fits
# 315 and the 9 mixture sums 162) = 1318; viscous VISC_PREP + VISC_DIV = 386;
# updates RK + 9 species * 5 + primitives with Rescaling_di and the Newton
# temperature (about 10 iterations of a 9-species NASA energy) = 260.
# 3 * 9299 = 27,897, i.e. about 2.9e4 (+ about 200 exp/log per step).
JET_STEP = 2.9e4


def jet_step_bytes_per_cell(n_spec=9):
    """Compulsory state bytes of a reacting RK3 step: 5 flow + N_SPEC species
    variables, 8 reads/writes of each."""
    return step_bytes_per_cell(5 + n_spec)


# Point-implicit chemistry substep of the 9-species / 19-reaction mechanism
# (CHEM_reactions.f90:230-272, 473-567): forward rates with third-body and
# Troe falloff terms, Gibbs K_c of 9 species, rates of progress, the analytic
# Jacobian of the 9 production rates, the 9x9 implicit matrix and its Gauss
# elimination with partial pivoting, the temperature Newton iteration and
# the energy update: about 7e3 flops (+ about 70 exp/log/pow).
CHEM_SUBSTEP = 7.0e3

# Surrogate inference, 12 -> 256 x5 -> 9: 2 * (12*256 + 4*256*256 + 256*9)
# multiply-adds as flops, per cell (GELU and the checks not counted).
MLP_CELL = 2 * (12 * 256 + 4 * 256 * 256 + 256 * 9)
