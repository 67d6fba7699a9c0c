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
