"""Case set-up for the flow path: meshes, initial states and BC data.

Host-side numpy code that plays the role of the reference's case generators
(cases/isotropic/init-isotropic.f90, cases/flat-2.25/sc2-data/Grid-and-init.f90)
and of init3d (code/flow-solver/OCFD_init.f90:1-25).  It runs once, before the
hot path, and its outputs are uploaded through the C-ABI.

Synthetic inputs (no data files are read):
  * isotropic(n): decaying compressible isotropic turbulence in a periodic box
    of side 2*pi, calorically perfect air, uniform rho and T, a seeded random
    solenoidal velocity with E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 4, scaled to
    turbulent Mach 0.5 and Re_lambda = 72 (definitions of ana_box,
    code/others/OCFD_ANA.f90:197-303).
  * flat_plate(nx, ny, nz): Mach 2.25 flat-plate boundary layer, T_inf =
    169.44 K, Re = 635000/in, isothermal wall T_w = 1.9 T_inf, geometric
    wall-normal stretching, compressible laminar similarity inflow, Rai
    blow/suction strip with seeded phases, periodic span.
"""
import math

import numpy as np

from . import (LAP, BC_BOUNDARYLAYER, SCHEME_HYBRID, SCHEME_WENO7, SCHEME_CD6, SPLIT_SW, make_cfg)

R0 = 8.314          # CHEM_modules.f90:20
MI_AIR = 0.02897    # kg/mol
R_AIR = R0 / MI_AIR
GAMMA = 1.4
SUTH_TA, SUTH_TC = 288.15, 110.4   # TRANSPORT_modules.f90:265
AMU0 = 1.789e-5


def sutherland(T, amu0=AMU0):
    tr = T / SUTH_TA
    return amu0 * np.sqrt(tr * tr * tr) * (SUTH_TA + SUTH_TC) / (SUTH_TC + T)


def halo_zeros(nx, ny, nz, value=0.0):
    return np.full((nx + 2 * LAP, ny + 2 * LAP, nz + 2 * LAP), value, dtype=np.float64, order="F")


def interior(a):
    return a[LAP:-LAP, LAP:-LAP, LAP:-LAP]


def state_from_primitives(d, u, v, w, T, R=R_AIR):
    """Host image of init3d (OCFD_init.f90:1-25) for N_SPEC=1, perfect gas.

    d,u,v,w,T are interior arrays.  Returns the dict of reference arrays
    (halo arrays carry the allocate_flow_data ghost values d=u=v=w=p=Et=1,
    T=300, OCFD_allocate_data.f90:53-55)."""
    nx, ny, nz = d.shape
    st = {}
    for name, val, src in (("D", 1.0, d), ("U", 1.0, u), ("V", 1.0, v), ("W", 1.0, w), ("T", 300.0, T)):
        a = halo_zeros(nx, ny, nz, val)
        interior(a)[...] = src
        st[name] = a
    di = halo_zeros(nx, ny, nz, 1.0)
    interior(di)[...] = d
    # Rescaling_di with one species: di*(d/di)
    interior(di)[...] = interior(di) * (d / interior(di))
    st["DI"] = di[..., None].copy(order="F")
    Et = 0.0 + 2.5 * R * interior(di) * T       # comput_E_pgas
    f = np.zeros((nx, ny, nz, 5), order="F")
    f[..., 0] = d
    f[..., 1] = d * u
    f[..., 2] = d * v
    f[..., 3] = d * w
    f[..., 4] = Et + d * (u * u + v * v + w * w) * 0.5
    st["F"] = f
    st["P"] = halo_zeros(nx, ny, nz, 1.0)
    st["ET"] = halo_zeros(nx, ny, nz, 1.0)
    return st


def mesh_1d(x, y, z):
    """GRID1D coordinates as halo arrays (read_mesh1d layout)."""
    nx, ny, nz = len(x), len(y), len(z)
    axx, ayy, azz = (halo_zeros(nx, ny, nz, 1.0) for _ in range(3))
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    interior(axx)[...] = X
    interior(ayy)[...] = Y
    interior(azz)[...] = Z
    return axx, ayy, azz


# --------------------------------------------------------------------------
def isotropic(n=64, seed=1, k0=4.0, mach_t=0.5, re_lambda=72.0, T0=288.15, scheme_invis=SCHEME_WENO7,
              cfl=0.5):
    """Decaying isotropic turbulence (the 64^3 WENO7 / Steger-Warming / CD6 case)."""
    L = 2.0 * math.pi
    h = L / n
    x = np.arange(n) * h
    cfg = make_cfg(n, n, n, periodic=(1, 1, 1), scheme_invis=scheme_invis, scheme_vis=SCHEME_CD6,
                   flux_splitting=SPLIT_SW, if_viscous=1, pr=0.72, amu0=AMU0)
    rng = np.random.Generator(np.random.PCG64(seed))
    kv = np.fft.fftfreq(n, d=1.0 / n)
    KX, KY, KZ = np.meshgrid(kv, kv, kv, indexing="ij")
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    K = np.sqrt(K2)
    Ek = np.where(K > 0, K ** 4 * np.exp(-2.0 * (K / k0) ** 2), 0.0)
    # random complex field, made solenoidal by projection, amplitude per shell
    U = [rng.standard_normal(K.shape) + 1j * rng.standard_normal(K.shape) for _ in range(3)]
    K2s = np.where(K2 > 0, K2, 1.0)
    kdotu = KX * U[0] + KY * U[1] + KZ * U[2]
    U = [U[0] - KX * kdotu / K2s, U[1] - KY * kdotu / K2s, U[2] - KZ * kdotu / K2s]
    amp = np.sqrt(np.sum([np.abs(c) ** 2 for c in U], axis=0))
    scale = np.where((K > 0) & (amp > 0), np.sqrt(Ek / np.maximum(4 * np.pi * K2s, 1e-300)) / np.maximum(amp, 1e-300), 0.0)
    vel = [np.real(np.fft.ifftn(c * scale)) for c in U]
    c0 = math.sqrt(GAMMA * R_AIR * T0)
    urms2 = np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    s = mach_t * c0 / math.sqrt(urms2)
    vel = [a * s for a in vel]
    # Re_lambda = q * lambda * rho / mu, lambda = q / rms(du/dx)   (ana_box)
    q = math.sqrt(np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2) / 3.0)
    dudx = np.real(np.fft.ifftn(1j * KX * np.fft.fftn(vel[0])))
    lam = q / math.sqrt(np.mean(dudx ** 2))
    mu0 = float(sutherland(T0))
    rho0 = re_lambda * mu0 / (q * lam)
    d = np.full((n, n, n), rho0)
    T = np.full((n, n, n), T0)
    st = state_from_primitives(d, vel[0], vel[1], vel[2], T)
    mesh = mesh_1d(x, x, x)
    span = np.array([L, 0, 0, 0, L, 0, 0, 0, L], dtype=np.float64)
    umax = (np.abs(vel[0]) + c0) / h + (np.abs(vel[1]) + c0) / h + (np.abs(vel[2]) + c0) / h
    dt = cfl / float(np.max(umax))
    info = dict(name="isotropic", rho0=rho0, T0=T0, mach_t=mach_t, re_lambda=re_lambda, dt=dt, n=n)
    return cfg, mesh, span, st, info


# --------------------------------------------------------------------------
def similarity_profile(mach, T_inf, T_wall, pr=0.72, n=2001, eta_max=12.0):
    """Compressible laminar boundary-layer similarity solution (Levy-Lees
    variables, Sutherland viscosity, constant Pr):
        (C f'')' + f f'' = 0
        (C th'/Pr)' + f th' + (gamma-1) M^2 C f''^2 = 0,  C = th^-1 mu(th)/mu_inf
    with f(0)=f'(0)=0, th(0)=T_w/T_inf, f'(inf)=th(inf)=1.  Returns eta,
    f, f' and th = T/T_inf."""
    from scipy.integrate import solve_bvp

    tw = T_wall / T_inf

    def C(th):
        T = th * T_inf
        return sutherland(T) / sutherland(T_inf) / th

    def rhs(eta, y):
        f, fp, g, th, q = y  # g = C f'', q = C th'/Pr
        c = C(th)
        fpp = g / c
        thp = q * pr / c
        return np.vstack([fp, fpp, -f * fpp, thp, -f * thp - (GAMMA - 1.0) * mach ** 2 * c * fpp ** 2])

    def bc(ya, yb):
        return np.array([ya[0], ya[1], ya[3] - tw, yb[1] - 1.0, yb[3] - 1.0])

    eta = np.linspace(0.0, eta_max, 400)
    fp0 = np.tanh(eta)
    th0 = tw + (1.0 - tw) * np.tanh(eta)
    y0 = np.vstack([eta - np.tanh(eta) * 0.9, fp0, 0.5 * (1 - fp0 ** 2), th0, 0.1 * (1 - fp0 ** 2)])
    sol = solve_bvp(rhs, bc, eta, y0, tol=1e-8, max_nodes=100000)
    if not sol.success:
        raise RuntimeError("similarity solver did not converge: " + sol.message)
    e = np.linspace(0.0, eta_max, n)
    y = sol.sol(e)
    return e, y[0], y[1], y[3]


def stretched_y(ny, dy1, ly):
    """Geometric wall-normal spacing with first spacing dy1 and height ly."""
    lo, hi = 1.0 + 1e-9, 2.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        if dy1 * (r ** (ny - 1) - 1.0) / (r - 1.0) > ly:
            hi = r
        else:
            lo = r
    r = 0.5 * (lo + hi)
    return np.concatenate([[0.0], np.cumsum(dy1 * r ** np.arange(ny - 1))])


# ---------------------------------------------------------------------------
# Boundary-layer BC set-up (the init routines of OCFD_BC_Boundarylayers.f90).
# The reference draws its random phases from an unseeded random_number; here
# the caller passes a seeded numpy Generator, drawn in the reference's order.
# ---------------------------------------------------------------------------
PI_REF = 3.14159265358979  # the PI parameter of OCFD_BC_Boundarylayers.f90:493,599


def init_bc_nonref(kind, axx_top, ayy_top, aoa=0.0):
    """init_bc_nonRef (OCFD_BC_Boundarylayers.f90:406-443): the upper face's
    BC%NonReflect_upper(i,k) (1 = non-reflecting, 0 = Dirichlet data).
    kind = BC_Para(3): 1 every point non-reflecting; 2 every point Dirichlet;
    0 (auto) Dirichlet where the top surface's slope sy = dy/dx (backward
    difference, forward at i = 1) is >= tan(AoA) + 1e-6, non-reflecting
    elsewhere.  axx_top, ayy_top: (nx, nz) coordinates of the plane j = ny.
    Returns int32 (nx, nz), Fortran order."""
    axx_top = np.asarray(axx_top, dtype=np.float64)
    ayy_top = np.asarray(ayy_top, dtype=np.float64)
    nx, nz = axx_top.shape
    if kind == 1:
        return np.ones((nx, nz), dtype=np.int32, order="F")
    if kind == 2:
        return np.zeros((nx, nz), dtype=np.int32, order="F")
    if kind != 0:
        raise ValueError("BC_Para(3) (upper boundary type) must be 0, 1 or 2")
    sy0 = math.tan(aoa)
    sy = np.empty((nx, nz))
    sy[0] = (ayy_top[1] - ayy_top[0]) / (axx_top[1] - axx_top[0])
    sy[1:] = (ayy_top[1:] - ayy_top[:-1]) / (axx_top[1:] - axx_top[:-1])
    return np.asfortranarray(np.where(sy >= sy0 + 1.0e-6, 0, 1).astype(np.int32))


def _mode_weights(m):
    """Weights 1, 1/1.25, 1/1.25^2, ... normalised to sum 1 (Zl, TM)."""
    w = np.empty(m)
    for k in range(m):
        w[k] = 1.0 if k == 0 else w[k - 1] / 1.25
    return w / w.sum()


def init_bc_wall_perturbation(dis_type, xbegin, xend, mz, zl_span, axx_wall, azz_wall, rng):
    """init_bc_wall_perturbation (OCFD_BC_Boundarylayers.f90:446-562): the
    wall's BC%wall_pertb(i,k) from the coordinates of the plane j = 1.
    dis_type 0: zero.  1 (Rai's multi-wave blowing/suction): fx gz with
    fx = 4/sqrt(27) sin(seta) (1 - cos(seta)), seta = 2 pi (x - xbegin) /
    (xend - xbegin), inside [xbegin, xend] and 0 outside; gz = sum_m Zl(m)
    sin(2 pi m (z/ZL + faiz(m))) for mz > 0 (random phases faiz), 1 for
    mz = 0, sin(-2 pi mz z/ZL) for mz < 0.  2: uniform random in (-1, 1)
    inside the strip, 0 outside.  Returns (nx, nz), Fortran order."""
    X = np.asarray(axx_wall, dtype=np.float64)
    Z = np.asarray(azz_wall, dtype=np.float64)
    if dis_type == 0:
        return np.zeros(X.shape, order="F")
    inside = (X >= xbegin) & (X <= xend)
    if dis_type == 2:
        r = rng.random(X.shape[::-1]).T          # drawn in the reference's (k, i) loop order
        return np.asfortranarray(np.where(inside, 2.0 * (r - 0.5), 0.0))
    if dis_type != 1:
        raise ValueError("BC_Para(7) (wall disturbance type) must be 0, 1 or 2")
    if mz > 0:
        faiz = rng.random(mz)
        zl = _mode_weights(mz)
    seta = 2.0 * PI_REF * (X - xbegin) / (xend - xbegin)
    fx = np.where(inside, 4.0 / math.sqrt(27.0) * np.sin(seta) * (1.0 - np.cos(seta)), 0.0)
    sz = Z / zl_span
    if mz > 0:
        gz = np.zeros_like(X)
        for m in range(1, mz + 1):
            gz = gz + zl[m - 1] * np.sin(2.0 * PI_REF * m * (sz + faiz[m - 1]))
    elif mz == 0:
        gz = np.ones_like(X)
    else:
        gz = np.sin(-2.0 * PI_REF * mz * sz)
    return np.asfortranarray(fx * gz)


def get_ht_multifrequancy(mt, beta, rng):
    """get_ht_multifrequancy (OCFD_BC_Boundarylayers.f90:594-621): the
    temporal factor of the wall disturbance, ht(tt) = sum_m TM(m)
    sin(m beta tt 2 pi + 2 pi fait(m)) for beta > 0 (TM normalised weights,
    random phases fait drawn once), 1 otherwise.  Returns the function ht."""
    fait = rng.random(mt) if mt > 0 else np.zeros(0)
    tm = _mode_weights(mt) if mt > 0 else np.zeros(0)

    def ht(tt):
        if not beta > 0.0:
            return 1.0
        h = 0.0
        for m in range(1, mt + 1):
            h = h + tm[m - 1] * math.sin(m * beta * tt * 2.0 * PI_REF + 2.0 * PI_REF * fait[m - 1])
        return float(h)

    return ht


def flat_plate(nx=640, ny=120, nz=64, seed=2, x0=0.1, lx=0.2, ly=0.012, lz=0.004445, theta=(2.0, 10.0),
               cfl=0.5):
    """Mach 2.25 spatially developing flat-plate boundary layer (hybrid
    UD7L/WENO7/WENO5 scheme switched by the pressure-gradient sensor)."""
    mach, T_inf = 2.25, 169.44
    re_per_m = 635000.0 / 0.0254
    c_inf = math.sqrt(GAMMA * R_AIR * T_inf)
    U_inf = mach * c_inf
    mu_inf = float(sutherland(T_inf))
    rho_inf = re_per_m * mu_inf / U_inf
    p_inf = rho_inf * R_AIR * T_inf
    T_wall = 1.9 * T_inf
    # mesh: x uniform, y stretched (first cell y+ ~ 0.5 for Cf ~ 2e-3), z uniform periodic
    rho_w = p_inf / (R_AIR * T_wall)
    nu_w = float(sutherland(T_wall)) / rho_w
    u_tau = U_inf * math.sqrt(1.0e-3)
    dy1 = 0.5 * nu_w / u_tau
    x = x0 + np.arange(nx) * (lx / (nx - 1))
    y = stretched_y(ny, dy1, ly)
    z = np.arange(nz) * (lz / nz)
    eta, f, fp, th = similarity_profile(mach, T_inf, T_wall)
    # physical wall distance of each eta at station x: y = sqrt(2 nu x / U) * int_0^eta th
    Y_eta = np.concatenate([[0.0], np.cumsum(0.5 * (th[1:] + th[:-1]) * np.diff(eta))])

    def profile_at(xs):
        scale = math.sqrt(2.0 * mu_inf / rho_inf * xs / U_inf)
        yy = Y_eta * scale
        uu = np.interp(y, yy, fp, right=1.0) * U_inf
        tt = np.interp(y, yy, th, right=1.0) * T_inf
        vv = np.zeros_like(y)  # wall-normal velocity of the inflow profile is set to zero
        dd = p_inf / (R_AIR * tt)
        return dd, uu, vv, tt

    d = np.empty((nx, ny, nz))
    u = np.empty_like(d)
    v = np.zeros_like(d)
    w = np.zeros_like(d)
    T = np.empty_like(d)
    for i in range(nx):
        dd, uu, vv, tt = profile_at(x[i])
        d[i] = dd[:, None]
        u[i] = uu[:, None]
        T[i] = tt[:, None]
    st = state_from_primitives(d, u, v, w, T)
    mesh = mesh_1d(x, y, z)
    span = np.array([0, 0, 0, 0, 0, 0, 0, 0, lz], dtype=np.float64)
    # inlet data (ny, 5+N_SPEC): d,u,v,w,T,di (di rescaled so sum = d)
    dd, uu, vv, tt = profile_at(x[0])
    inlet = np.zeros((ny, 6), order="F")
    inlet[:, 0], inlet[:, 1], inlet[:, 2], inlet[:, 4], inlet[:, 5] = dd, uu, vv, tt, dd
    inlet[:, 5] = inlet[:, 5] * inlet[:, 0] / inlet[:, 5]
    upper = np.array([rho_inf, U_inf, 0.0, 0.0, T_inf, rho_inf])
    X = x[:, None] * np.ones((1, nz))
    Z = np.ones((nx, 1)) * z[None, :]
    # upper boundary: auto type (BC_Para(3) = 0) with a flat top ->
    # non-reflecting everywhere
    nonref = init_bc_nonref(0, X, np.full((nx, nz), y[-1]))
    # Rai blow/suction strip with 10 spanwise and 10 temporal modes, seeded phases
    rng = np.random.Generator(np.random.PCG64(seed))
    mz = mt = 10
    xb, xe = 0.1143, 0.1270
    freq = 28861.58
    pertb = init_bc_wall_perturbation(1, xb, xe, mz, lz, X, Z, rng)
    ht = get_ht_multifrequancy(mt, freq, rng)

    bc_para = [1, 0, 0, 0, T_wall, 0.0, 1, 0.01 * U_inf, xb, xe, mt, mz, lz, freq]
    hp = [2.0, theta[0], theta[1], 0.0]
    cfg = make_cfg(nx, ny, nz, periodic=(0, 0, 1), scheme_invis=SCHEME_HYBRID, scheme_vis=SCHEME_CD6,
                   flux_splitting=SPLIT_SW, if_viscous=1, ibc=BC_BOUNDARYLAYER, hybrid_para=hp, pr=0.72,
                   amu0=AMU0, bc_para=bc_para, ud7l_diss=1.0)
    dxm = min(np.min(np.diff(x)), 1.0)
    dym = np.diff(y)
    dzm = lz / nz
    c_all = np.sqrt(GAMMA * R_AIR * T)
    lam = ((np.abs(u) + c_all) / dxm + (np.abs(v) + c_all) / dym.min() + (np.abs(w) + c_all) / dzm)
    dt = cfl / float(np.max(lam))
    info = dict(name="flat_plate", U_inf=U_inf, rho_inf=rho_inf, p_inf=p_inf, T_inf=T_inf, T_wall=T_wall,
                dy1=dy1, dt=dt, ht=ht)
    bc = dict(inlet1d=inlet, upper0d=upper, nonreflect_upper=nonref, wall_pertb=pertb)
    return cfg, mesh, span, st, bc, info


# ---------------------------------------------------------------------------
# Jet-flame chemistry batches.  Streams of the jet case (cases/jet/inlet.in):
# fuel X_H2 = 0.85, X_N2 = 0.15 at 305 K; oxidiser air X_O2 = 0.21,
# X_N2 = 0.79 at 1150 K; 1 atm.
# ---------------------------------------------------------------------------
def jet_streams(mech, kind="planar"):
    """Mass fractions of the fuel and oxidiser streams: "planar" = the
    planar jet (X_H2 = 0.10, X_N2 = 0.90 against air), "round" = the
    reference's round-jet deck (X_H2 = 0.85, X_N2 = 0.15 against air)."""
    xf = np.zeros(mech.n_spec)
    xo = np.zeros(mech.n_spec)
    h2 = {"planar": 0.10, "round": 0.85}[kind]
    xf[mech.index("H2")], xf[mech.index("N2")] = h2, 1.0 - h2
    xo[mech.index("O2")], xo[mech.index("N2")] = 0.21, 0.79
    return mech.mass_fractions(xf), mech.mass_fractions(xo)


def jet_chem_states(n, seed=4, mech=None, T_range=(900.0, 2500.0), p_range_atm=(0.5, 5.0), dt_range=(1e-8, 1e-6),
                    tau_range=(1e-7, 2e-3), streams="planar"):
    """Unreacted jet-mixture cells and the data to move them along their
    0-D ignition trajectories.

    Z (mixture fraction of the fuel stream, by mass) ~ U[0, 1] between the
    planar jet's streams (``jet_streams``), T ~ U[T_range], p ~ U[p_range_atm]
    atm, dt ~ log-U[dt_range] s (the chemistry step to benchmark), tau ~
    log-U[tau_range] s (how far along its ignition trajectory each cell is
    taken before the benchmark; the caller advances the cells by tau with the
    0-D integrator, which holds density and internal energy fixed as the
    reference's chemical_timeAdv does, so the trajectories are constant
    volume).  Returns a dict of numpy arrays: rho [n_spec, n] (kg/m3), E, T,
    p, Z, dt, tau [n]."""
    from .mechanism import Mechanism
    mech = mech or Mechanism()
    rng = np.random.default_rng(seed)
    yf, yo = jet_streams(mech, streams)
    Z = rng.uniform(0.0, 1.0, n)
    Y = Z[None, :] * yf[:, None] + (1.0 - Z[None, :]) * yo[:, None]
    T = rng.uniform(T_range[0], T_range[1], n)
    p = 101300.0 * rng.uniform(p_range_atm[0], p_range_atm[1], n)
    rho = Y * mech.density(Y, T, p)[None, :]
    E = mech.energy(rho, T)
    dt = np.exp(rng.uniform(math.log(dt_range[0]), math.log(dt_range[1]), n))
    tau = np.exp(rng.uniform(math.log(tau_range[0]), math.log(tau_range[1]), n))
    return dict(rho=np.ascontiguousarray(rho), E=E, T=T, p=p, Z=Z, dt=dt, tau=tau)


# ---------------------------------------------------------------------------
# Reacting jet (N_SPEC = 9, real gas, Chemkin/Wilke transport)
# ---------------------------------------------------------------------------
def jet(nx=64, ny=48, nz=48, seed=3, lx=0.004, ly=0.003, lz=0.003, rd=0.0005, rbuf=0.0001, xh=0.001, Th=1800.0,
        epsl=0.05, rdist=0.0002, omeg=8.0, mech=None, cfl=0.3, iflag=1):
    """Round H2/N2 jet into a hot air coflow: the set-up of cases/jet
    (opencfd-comb2.in, inlet.in and the profile construction of
    jet-comb-inlet.f90) on a uniform box [0, lx] x [-ly/2, ly/2] x
    [-lz/2, lz/2].

    Fuel stream X_H2 = 0.85, X_N2 = 0.15, 900 m/s, 305 K; coflow air at
    20 m/s, 1150 K; 1 atm.  Radial profiles blend the two streams linearly
    over rbuf around r = rd; a seeded azimuthal perturbation of amplitude
    epsl (modes omeg, 2 omeg, 4 omeg) on the shear layer
    (rd +- rdist/2) feeds u, v, w.  The initial field is the inflow plane
    repeated downstream with the jet core at Th for x > xh.  OMP6 / CD6,
    Steger-Warming, real gas, Flag_Viscosity = Flag_Conductivity =
    Flag_Diffusion = 1, BC_JET with BC_Para(1) = iflag.

    Returns cfg, mesh, span, state, bc (inlet2d), info (dt, mech)."""
    from .mechanism import Mechanism
    from . import BC_JET, SCHEME_OMP6
    mech = mech or Mechanism()
    ns = mech.n_spec
    xf = np.zeros(ns)
    xo = np.zeros(ns)
    xf[mech.index("H2")], xf[mech.index("N2")] = 0.85, 0.15
    xo[mech.index("O2")], xo[mech.index("N2")] = 0.21, 0.79
    u1, T1, p1 = 900.0, 305.0, 101300.0
    u2, T2, p2 = 20.0, 1150.0, 101300.0
    x = np.linspace(0.0, lx, nx)
    y = np.linspace(-0.5 * ly, 0.5 * ly, ny)
    z = np.linspace(-0.5 * lz, 0.5 * lz, nz)
    Y2, Z2 = np.meshgrid(y, z, indexing="ij")
    rr = np.sqrt(Y2 * Y2 + Z2 * Z2)
    tmp = np.clip((rr - (rd - 0.5 * rbuf)) / rbuf, 0.0, 1.0)
    rng = np.random.Generator(np.random.PCG64(seed))
    fai = rng.random(3)
    Ri = R0 / mech.mi

    def plane(t_core):
        T0 = t_core + (T2 - t_core) * tmp
        p0 = p1 + (p2 - p1) * tmp
        u0 = u1 + (u2 - u1) * tmp
        X0 = xf[:, None, None] + (xo - xf)[:, None, None] * tmp[None]
        di = p0[None] * X0 / (Ri[:, None, None] * T0[None])
        seta = np.arccos(np.divide(Y2, rr, out=np.ones_like(rr), where=rr > 0))
        seta = np.where(Z2 < 0, 2.0 * math.pi - seta, seta)
        r1, r2 = rd - 0.5 * rdist, rd + 0.5 * rdist
        band = (rr >= r1) & (rr <= r2)
        fr = np.where(band, u0 * epsl / 1.75 * np.sin(math.pi * (rr - r1) / (r2 - r1)) *
                      (np.sin(2.0 * math.pi * (omeg * seta + fai[0])) +
                       0.5 * np.sin(2.0 * math.pi * (2.0 * omeg * seta + fai[1])) +
                       0.25 * np.sin(2.0 * math.pi * (4.0 * omeg * seta + fai[2]))), 0.0)
        safe = np.where(rr > 0, rr, 1.0)
        return di, u0 + fr, fr * Y2 / safe, fr * Z2 / safe, T0

    di_in, u_in, v_in, w_in, T_in = plane(T1)
    di_h, u_h, v_h, w_h, T_h = plane(Th)
    hot = x > xh
    di = np.where(hot[None, :, None, None], di_h[:, None], di_in[:, None])      # [ns, nx, ny, nz]
    u = np.where(hot[:, None, None], u_h[None], u_in[None])
    v = np.where(hot[:, None, None], v_h[None], v_in[None])
    w = np.where(hot[:, None, None], w_h[None], w_in[None])
    T = np.where(hot[:, None, None], T_h[None], T_in[None])
    d = di.sum(axis=0)
    st = {}
    for name, val, src in (("D", 1.0, d), ("U", 1.0, u), ("V", 1.0, v), ("W", 1.0, w), ("T", 300.0, T)):
        a = halo_zeros(nx, ny, nz, val)
        interior(a)[...] = src
        st[name] = a
    dih = np.ones((nx + 2 * LAP, ny + 2 * LAP, nz + 2 * LAP, ns), order="F")
    for m in range(ns):
        interior(dih[..., m])[...] = di[m]
    st["DI"] = dih
    E = mech.energy(di.reshape(ns, -1), T.reshape(-1)).reshape(d.shape)
    f = np.zeros((nx, ny, nz, 5), order="F")
    f[..., 0] = d
    f[..., 1] = d * u
    f[..., 2] = d * v
    f[..., 3] = d * w
    f[..., 4] = E + d * (u * u + v * v + w * w) * 0.5
    st["F"] = f
    st["P"] = halo_zeros(nx, ny, nz, 1.0)
    st["ET"] = halo_zeros(nx, ny, nz, 1.0)
    inlet = np.zeros((ny, nz, 5 + ns), order="F")
    inlet[..., 0] = di_in.sum(axis=0)
    inlet[..., 1], inlet[..., 2], inlet[..., 3], inlet[..., 4] = u_in, v_in, w_in, T_in
    for m in range(ns):
        inlet[..., 5 + m] = di_in[m]
    mesh = mesh_1d(x, y, z)
    span = np.zeros(9)
    cfg = make_cfg(nx, ny, nz, periodic=(0, 0, 0), scheme_invis=SCHEME_OMP6, scheme_vis=SCHEME_CD6,
                   flux_splitting=SPLIT_SW, if_viscous=1, ibc=BC_JET, bc_para=[iflag], n_spec=ns, if_real_gas=1,
                   flag_vis=1, flag_cond=1, flag_diff=1)
    # acoustic CFL with the frozen sound speed sqrt(gamma p / d), gamma = p/E + 1
    p = (di * (Ri[:, None, None, None] * T[None])).sum(axis=0)
    c = np.sqrt((p / E + 1.0) * p / d)
    lam = (np.abs(u) + c) / (lx / (nx - 1)) + (np.abs(v) + c) / (ly / (ny - 1)) + (np.abs(w) + c) / (lz / (nz - 1))
    dt = cfl / float(np.max(lam))
    info = dict(name="jet", dt=dt, mech=mech)
    return cfg, mesh, span, st, dict(inlet2d=inlet), info


def jet_gas(engine, mech):
    """Species thermo and transport fits of the mechanism into a solver (or
    the oracle): read_Spec and read_vis/conduct/diffusion_coefficients."""
    engine.set_species(mech.mi, n_tct=mech.n_tct, tct=mech.tct, tct_max=mech.tct_max, coef=mech.coef)
    engine.set_transport(mech.viscosity, mech.conductivity, mech.diffusion)


def planar_jet(nx=256, ny=192, nz=128, seed=5, lx=0.024, ly=0.018, lz=0.012, width=0.002, delta=0.0002,
               noise=0.05, kmax=8, ignite=True, T_ign=1800.0, x_ign=0.25, mech=None, cfl=0.5, iflag=1):
    """Reacting planar H2 jet into a hot air coflow on a uniform box
    [0, lx] x [-ly/2, ly/2] x [0, lz) (span periodic).

    Jet: X_H2 = 0.10, X_N2 = 0.90 at 300 K, 300 m/s, |y| < width/2; coflow:
    air at 1100 K, 20 m/s; 1 atm.  The streams blend with a tanh profile of
    thickness delta.  The inlet velocity carries seeded broadband noise:
    random-phase Fourier modes in (y, z) up to kmax per direction, flat
    spectrum, scaled to an RMS of `noise` x 300 m/s inside the jet, on u, v
    and w.  The initial field repeats the inflow plane downstream; with
    `ignite`, the mixing layers (blend weight 0.1 .. 0.9) for x > x_ign lx
    start at T_ign as an igniter.  WENO5, Steger-Warming, CD6, real gas,
    Chemkin/Wilke transport with mixture-averaged diffusion, BC_JET with
    BC_Para(1) = iflag, periodic span.

    Returns cfg, mesh, span, state, bc (inlet2d), info (dt, mech)."""
    from .mechanism import Mechanism
    from . import BC_JET, SCHEME_WENO5
    mech = mech or Mechanism()
    ns = mech.n_spec
    xf = np.zeros(ns)
    xo = np.zeros(ns)
    xf[mech.index("H2")], xf[mech.index("N2")] = 0.10, 0.90
    xo[mech.index("O2")], xo[mech.index("N2")] = 0.21, 0.79
    u1, T1, u2, T2, p0 = 300.0, 300.0, 20.0, 1100.0, 101300.0
    x = np.linspace(0.0, lx, nx)
    y = np.linspace(-0.5 * ly, 0.5 * ly, ny)
    z = np.arange(nz) * (lz / nz)
    s = 0.5 * (1.0 - np.tanh((np.abs(y) - 0.5 * width) / delta))          # 1 in the jet
    Y2, Z2 = np.meshgrid(y, z, indexing="ij")
    S2 = np.broadcast_to(s[:, None], Y2.shape)
    rng = np.random.Generator(np.random.PCG64(seed))
    pert = []
    for _ in range(3):
        f = np.zeros(Y2.shape)
        for ky in range(1, kmax + 1):
            for kz in range(0, kmax + 1):
                ph = rng.uniform(0.0, 2.0 * math.pi, 2)
                f += np.sin(2 * math.pi * ky * (Y2 / ly) + ph[0]) * np.cos(2 * math.pi * kz * (Z2 / lz) + ph[1])
        pert.append(f)
    inj = S2 > 0.5
    rms = math.sqrt(np.mean(np.concatenate([p[inj] for p in pert]) ** 2))
    pert = [noise * u1 * p / rms * S2 for p in pert]
    Ri = R0 / mech.mi
    T_in = T2 + (T1 - T2) * S2
    X_in = xo[:, None, None] + (xf - xo)[:, None, None] * S2[None]
    di_in = p0 * X_in / (Ri[:, None, None] * T_in[None])
    u_in = u2 + (u1 - u2) * S2 + pert[0]
    v_in, w_in = pert[1], pert[2]
    T = np.broadcast_to(T_in[None], (nx, ny, nz)).copy()
    if ignite:
        band = (S2 > 0.1) & (S2 < 0.9)
        T[(x > x_ign * lx)[:, None, None] & band[None]] = T_ign
    # constant pressure: partial densities from the local temperature
    di = p0 * np.broadcast_to(X_in[:, None], (ns, nx, ny, nz)) / (Ri[:, None, None, None] * T[None])
    u = np.broadcast_to(u_in[None], (nx, ny, nz)).copy()
    v = np.broadcast_to(v_in[None], (nx, ny, nz)).copy()
    w = np.broadcast_to(w_in[None], (nx, ny, nz)).copy()
    d = di.sum(axis=0)
    st = {}
    for name, val, src in (("D", 1.0, d), ("U", 1.0, u), ("V", 1.0, v), ("W", 1.0, w), ("T", 300.0, T)):
        a = halo_zeros(nx, ny, nz, val)
        interior(a)[...] = src
        st[name] = a
    dih = np.ones((nx + 2 * LAP, ny + 2 * LAP, nz + 2 * LAP, ns), order="F")
    for m in range(ns):
        interior(dih[..., m])[...] = di[m]
    st["DI"] = dih
    E = mech.energy(di.reshape(ns, -1), T.reshape(-1)).reshape(d.shape)
    f = np.zeros((nx, ny, nz, 5), order="F")
    f[..., 0] = d
    f[..., 1] = d * u
    f[..., 2] = d * v
    f[..., 3] = d * w
    f[..., 4] = E + d * (u * u + v * v + w * w) * 0.5
    st["F"] = f
    st["P"] = halo_zeros(nx, ny, nz, 1.0)
    st["ET"] = halo_zeros(nx, ny, nz, 1.0)
    inlet = np.zeros((ny, nz, 5 + ns), order="F")
    inlet[..., 0] = di_in.sum(axis=0)
    inlet[..., 1], inlet[..., 2], inlet[..., 3], inlet[..., 4] = u_in, v_in, w_in, T_in
    for m in range(ns):
        inlet[..., 5 + m] = di_in[m]
    mesh = mesh_1d(x, y, z)
    span = np.array([0, 0, 0, 0, 0, 0, 0, 0, lz], dtype=np.float64)
    cfg = make_cfg(nx, ny, nz, periodic=(0, 0, 1), scheme_invis=SCHEME_WENO5, scheme_vis=SCHEME_CD6,
                   flux_splitting=SPLIT_SW, if_viscous=1, ibc=BC_JET, bc_para=[iflag], n_spec=ns, if_real_gas=1,
                   flag_vis=1, flag_cond=1, flag_diff=1)
    p = (di * (Ri[:, None, None, None] * T[None])).sum(axis=0)
    c = np.sqrt((p / E + 1.0) * p / d)
    lam = (np.abs(u) + c) / (lx / (nx - 1)) + (np.abs(v) + c) / (ly / (ny - 1)) + (np.abs(w) + c) / (lz / nz)
    dt = cfl / float(np.max(lam))
    info = dict(name="planar_jet", dt=dt, mech=mech)
    return cfg, mesh, span, st, dict(inlet2d=inlet), info
