"""Reaction mechanism, species thermo and transport-fit data.

Readers for the reference's data files (``Reaction.in``, ``specie.in``,
``viscosity.in``, ``conductivity.in``, ``diffusion_coefficients.in``; formats
of read_Reac, CHEM_reactions.f90:15-71, read_Spec, CHEM_thermdyn.f90:16-51, and
the transport readers, TRANSPORT_modules.f90:80-253) and a JSON form of the
same data that travels with the package.

The readers accept the current format and the older deck format found in
the reference's cases/ directory: specie.in without the SpecFlag column and
Reaction.in without the per-reaction reverse-rate line when ReReac=0.

``data/h2_o2_n2_9sp19r.json`` is the jet case's 9-species / 19-reaction
H2-O2 mechanism with N2 inert (species order O, O2, H2, H2O, OH, H, HO2,
H2O2, N2), converted by ``python3 -m`` ... ``mechanism.convert(<cases/jet dir>)``.
"""
import json
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(HERE, "data")
JET_JSON = os.path.join(DATA, "h2_o2_n2_9sp19r.json")

# element composition (H, O, N) used by the conservation checks
ELEMENTS = {"O": (0, 1, 0), "O2": (0, 2, 0), "H2": (2, 0, 0), "H2O": (2, 1, 0), "OH": (1, 1, 0),
            "H": (1, 0, 0), "HO2": (1, 2, 0), "H2O2": (2, 2, 0), "N2": (0, 0, 2), "N": (0, 0, 1),
            "NO": (0, 1, 1), "AIR": None}


def _fnum(tok):
    return float(tok.replace("d", "e").replace("D", "e").rstrip(","))


def _data_lines(path):
    """Lines with the Fortran list-directed comment lines kept (the readers
    skip them by position)."""
    return open(path).read().splitlines()


def read_reactions(path, n_spec_expected=None):
    """read_Reac (CHEM_reactions.f90:15-71)."""
    L = _data_lines(path)
    i = 2
    n_spec, n_reac, rereac = (int(x) for x in L[i].split()[:3])
    i += 1
    names = L[i + 1].split()
    i += 4  # 4 header lines after the counts
    reacs = []
    for _ in range(n_reac):
        i += 1  # '#--- j' line
        label = L[i].strip()
        i += 1
        af_type, tr, fc = L[i].split()[:3]
        af_type, tr, fc = int(af_type), int(tr), _fnum(fc)
        i += 1
        A1, b1, E1 = (_fnum(x) for x in L[i].split()[:3])
        i += 1
        arr = [(A1, b1, E1)]
        if af_type != 0:
            arr.append(tuple(_fnum(x) for x in L[i].split()[:3]))
            i += 1
        rev = None
        # ReReac=0: the current reader skips one line; the old deck format has
        # no such line (the next line is already the integer stoichiometry)
        if rereac == 0:
            toks = L[i].split()
            if not (len(toks) == n_spec and all(re.fullmatch(r"-?\d+", t) for t in toks)):
                i += 1
        elif rereac == 1:
            rev = tuple(_fnum(x) for x in L[i].split()[:3])
            i += 1
        elif rereac == 2:
            rev = tuple(_fnum(x) for x in L[i].split()[:5])
            i += 1
        vf = [int(x) for x in L[i].split()[:n_spec]]
        i += 1
        vr = [int(x) for x in L[i].split()[:n_spec]]
        i += 1
        eff = None
        if tr == 1:
            eff = [_fnum(x) for x in L[i].split()[:n_spec]]
            i += 1
        reacs.append(dict(label=label, af_type=af_type, third_body=tr, fc_troe=fc, arrhenius=arr, reverse=rev,
                          vf=vf, vr=vr, eff=eff))
    if n_spec_expected is not None and n_spec != n_spec_expected:
        raise ValueError("N_SPEC in Reaction.in is inconsistent")
    return dict(n_spec=n_spec, n_reac=n_reac, rereac=rereac, names=names, reactions=reacs)


def read_species(path, n_spec):
    """read_Spec (CHEM_thermdyn.f90:16-51): N_Tct, Tct(1:N_Tct-1), Tct_max, then
    per species: name, Mi [SpecFlag], N_Tct lines of (A..G)."""
    toks = []
    for line in _data_lines(path):
        s = line.strip()
        if not s or s.startswith("#"):
            toks.append(None)
            continue
        toks.append(s)
    lines = [t for t in toks]
    # header: 3 lines then the N_Tct line
    body = [l for l in lines]
    k = 0
    while body[k] is None:
        k += 1
    first = body[k].split()
    n_tct = int(first[0])
    tct = [_fnum(x) for x in first[1:n_tct]]
    tct_max = _fnum(first[n_tct])
    k += 1
    species = []
    while len(species) < n_spec:
        while body[k] is None:
            k += 1
        name = body[k].strip().strip("'\"")
        k += 1
        parts = body[k].replace(",", " ").split()
        mi = _fnum(parts[0])
        flag = int(float(parts[1])) if len(parts) > 1 else 2
        k += 1
        coef = []
        for _ in range(n_tct):
            nums = [_fnum(x) for x in re.split(r"[,\s]+", body[k].strip()) if x]
            # a missing comma may glue two numbers ("-0.27551d-4    -0.31028d-8"):
            # the split on blanks already separates them
            coef.append(nums[:7])
            k += 1
        species.append(dict(name=name, mi=mi, spec_flag=flag, coef=coef))
    return dict(n_tct=n_tct, tct=tct, tct_max=tct_max, species=species)


def read_fits(path, names, ncoef=4, pairs=False):
    """Chemkin log-polynomial transport fits (TRANSPORT_modules.f90:80-253)."""
    L = [l.strip() for l in _data_lines(path)]
    L = [l for l in L if l and not l.startswith("#")]
    if pairs:
        n = int(L[1])
        out = {}
        k = 2
        for _ in range(n):
            a, b = (x.strip("'\"") for x in L[k].split()[:2])
            c = [_fnum(x) for x in L[k + 1].split()[:ncoef]]
            out[(a, b)] = c
            out[(b, a)] = c
            k += 2
        return [[out[(a, b)] for b in names] for a in names]
    n = int(L[0])
    out = {}
    k = 1
    for _ in range(n):
        nm = L[k].split()[0].strip("'\"")
        out[nm] = [_fnum(x) for x in L[k + 1].split()[:ncoef]]
        k += 2
    return [out[nm] for nm in names]


def convert(case_dir, out=JET_JSON):
    """Read a reference case directory and write the JSON form."""
    r = read_reactions(os.path.join(case_dir, "Reaction.in"))
    s = read_species(os.path.join(case_dir, "specie.in"), r["n_spec"])
    names = [sp["name"] for sp in s["species"]]
    if names != r["names"]:
        raise ValueError(f"species order differs: {names} vs {r['names']}")
    data = dict(source=os.path.basename(os.path.normpath(case_dir)), species=names, mechanism=r, thermo=s)
    for key, fname, pairs in (("viscosity", "viscosity.in", False), ("conductivity", "conductivity.in", False),
                              ("diffusion", "diffusion_coefficients.in", True)):
        p = os.path.join(case_dir, fname)
        if os.path.exists(p):
            data[key] = read_fits(p, names, pairs=pairs)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    json.dump(data, open(out, "w"), indent=1)
    return out


class Mechanism:
    """Arrays of a mechanism in the layout the C-ABI takes."""

    def __init__(self, path=JET_JSON, data=None):
        """From the JSON form at ``path``, or from the same structure given
        as ``data`` (a dict)."""
        d = data if data is not None else json.load(open(path))
        self.names = d["species"]
        m, t = d["mechanism"], d["thermo"]
        self.n_spec, self.n_reac, self.rereac = m["n_spec"], m["n_reac"], m["rereac"]
        R = m["reactions"]
        self.af_type = np.array([r["af_type"] for r in R], dtype=np.int32)
        self.third_body = np.array([r["third_body"] for r in R], dtype=np.int32)
        self.fc_troe = np.array([r["fc_troe"] for r in R])
        # Arrhenius (A, b, E) x 2 per reaction (second set zero when unused)
        arr = np.zeros((self.n_reac, 2, 3))
        for j, r in enumerate(R):
            for a, abc in enumerate(r["arrhenius"]):
                arr[j, a] = abc
        self.arrhenius = arr
        # reverse-rate data per reaction: ReReac = 1 (A_r, b_r, E_r, 0, 0),
        # ReReac = 2 the Park fit A1..A5 (read_Reac, CHEM_reactions.f90:46-47)
        rev = np.zeros((self.n_reac, 5))
        for j, r in enumerate(R):
            if r.get("reverse") is not None:
                rev[j, :len(r["reverse"])] = r["reverse"]
        self.reverse = rev
        self.vf = np.array([r["vf"] for r in R], dtype=np.int32)  # [reac][spec]
        self.vr = np.array([r["vr"] for r in R], dtype=np.int32)
        self.eff = np.array([r["eff"] if r["eff"] is not None else [0.0] * self.n_spec for r in R])
        self.sgm = (self.vr - self.vf).sum(axis=1).astype(np.int32)
        self.n_tct, self.tct, self.tct_max = t["n_tct"], np.array(t["tct"]), t["tct_max"]
        self.mi = np.array([s["mi"] for s in t["species"]])
        self.coef = np.array([s["coef"] for s in t["species"]])  # [spec][range][7]
        self.viscosity = np.array(d.get("viscosity", []))
        self.conductivity = np.array(d.get("conductivity", []))
        self.diffusion = np.array(d.get("diffusion", []))
        self.elements = np.array([ELEMENTS.get(n) or (0, 0, 0) for n in self.names], dtype=np.float64)  # [spec][H,O,N]

    def index(self, name):
        return self.names.index(name)

    # -- host thermo for setting up states (comput_E_rgas, CHEM_thermdyn.f90:203-231) --
    def ect(self):
        """Enthalpy-continuity offsets Ect[k][j] (read_Spec, CHEM_thermdyn.f90:43-47)."""
        e = np.zeros((self.n_spec, self.n_tct))
        c = self.coef
        for j in range(self.n_tct - 1):
            t = self.tct[j]
            d = c[:, j + 1] - c[:, j]
            e[:, j + 1] = -(d[:, 0] * t + d[:, 1] * (t * t) / 2.0 + d[:, 2] * ((t * t) * t) / 3.0
                            + d[:, 3] * ((t * t) * (t * t)) / 4.0 + d[:, 4] * (((t * t) * (t * t)) * t) / 5.0) + e[:, j]
        return e

    def energy(self, rho, T):
        """Internal energy per volume E (J/m3) of partial densities rho
        [n_spec, n] at temperatures T [n]."""
        rho = np.asarray(rho, dtype=np.float64)
        T = np.asarray(T, dtype=np.float64)
        jc = np.zeros(T.shape, dtype=int)
        for j in range(self.n_tct - 1):
            jc = np.where(T > self.tct[j], j + 1, jc)
        R = 8.314 / self.mi
        ect = self.ect()
        A1, B1, C1, D1, E1, Ect = (np.zeros(T.shape) for _ in range(6))
        for k in range(self.n_spec):
            w = R[k] * rho[k]
            cf = self.coef[k][jc]                               # [n, 7]
            A1 = A1 + (cf[..., 0] - 1.0) * w
            B1 = B1 + cf[..., 1] * w
            C1 = C1 + cf[..., 2] * w
            D1 = D1 + cf[..., 3] * w
            E1 = E1 + cf[..., 4] * w
            Ect = Ect + ect[k][jc] * w
        tm = self.tct_max
        Tc = np.minimum(T, tm)
        e = A1 * Tc + B1 * Tc ** 2 / 2.0 + C1 * Tc ** 3 / 3.0 + D1 * Tc ** 4 / 4.0 + E1 * Tc ** 5 / 5.0 + Ect
        cv_max = A1 + B1 * tm + C1 * tm ** 2 + D1 * tm ** 3 + E1 * tm ** 4
        return np.where(T <= tm, e, e + cv_max * (T - tm))

    def density(self, Y, T, p):
        """Mixture density from mass fractions Y [n_spec, n], T and p (Pa)."""
        return p / (T * np.sum(np.asarray(Y) * (8.314 / self.mi)[:, None], axis=0))

    def mass_fractions(self, X):
        """Mass fractions from mole fractions X [n_spec, ...]."""
        X = np.asarray(X, dtype=np.float64)
        w = X * self.mi.reshape((-1,) + (1,) * (X.ndim - 1))
        return w / np.sum(w, axis=0)
