"""Mechanism loader: a Cantera-YAML subset compiled to flat tables.

Ember hands Cantera a mechanism file (``CanteraGas::initialize``,
src/chemistry0d.cpp:455-485) and Cantera parses it.  Cantera is not part of
this build, so this module restates the subset of Cantera's YAML reader that
the hot path needs: ideal-gas phases, NASA7 two-range thermo, elementary,
three-body (with efficiencies) and falloff (Lindemann / Troe) reactions,
pressure-dependent Arrhenius (PLOG), Chebyshev and SRI falloff reactions, reversible
and irreversible, duplicates.

The output is a plain-text table file (``.emxm``) that the C-ABI library reads
in ``emx_create``.  All quantities are in SI/kmol units, matching Cantera's
internal units: A in (m^3/kmol)^(n-1)/s, activation energies as Ea/R in K,
molecular weights in kg/kmol.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np
import yaml

# Cantera's physical constants (Cantera ct_defs.h, CODATA 2018 exact values)
AVOGADRO = 6.02214076e26          # 1/kmol
BOLTZMANN = 1.380649e-23          # J/K
GAS_CONSTANT = AVOGADRO * BOLTZMANN  # J/kmol/K
ONE_ATM = 101325.0
CAL = 4.184

# Conventional atomic weights as used by Cantera's element table [kg/kmol]
ATOMIC_WEIGHTS = {"H": 1.008, "C": 12.011, "N": 14.007, "O": 15.999,
                  "Ar": 39.95, "He": 4.002602}

TYPE_ELEMENTARY, TYPE_THREE_BODY, TYPE_LINDEMANN, TYPE_TROE, TYPE_PLOG, TYPE_CHEBYSHEV, TYPE_SRI = \
    0, 1, 2, 3, 4, 5, 6

# pressure units of Cantera YAML quantities, in Pa
PRESSURE_UNITS = {"Pa": 1.0, "kPa": 1e3, "MPa": 1e6, "bar": 1e5, "atm": ONE_ATM,
                  "torr": ONE_ATM / 760.0, "Torr": ONE_ATM / 760.0, "dyn/cm^2": 0.1}

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


@dataclass
class Reaction:
    equation: str
    rtype: int
    reversible: bool
    reactants: List[int]       # species indices, repeated per unit stoichiometry
    products: List[int]
    A: float
    b: float
    EaR: float
    A0: float = 0.0
    b0: float = 0.0
    EaR0: float = 0.0
    troe: Tuple[float, float, float, float] = (0.0, 0.0, 0.0, 0.0)
    has_T2: bool = False
    default_eff: float = 1.0
    efficiencies: Dict[int, float] = field(default_factory=dict)
    duplicate: bool = False
    # PLOG: (P [Pa], A [SI kmol units], b, Ea/R [K]) per rate, in file order
    plog: List[Tuple[float, float, float, float]] = field(default_factory=list)
    # Chebyshev: (Tmin, Tmax, Pmin [Pa], Pmax [Pa]) and the coefficients of log10 k
    # (nT x nP, k in SI kmol units)
    cheb_range: Tuple[float, float, float, float] = (0.0, 0.0, 0.0, 0.0)
    cheb: Optional[np.ndarray] = None
    # SRI falloff: (a, b, c, d, e)
    sri: Tuple[float, float, float, float, float] = (0.0, 0.0, 0.0, 1.0, 0.0)
    # explicit forward orders (Cantera 'orders') per reactant position: the first position
    # of each reactant species carries that species' order, repeated positions 0; None when
    # every order equals the stoichiometric coefficient
    orders: Optional[List[float]] = None


@dataclass
class Mechanism:
    name: str
    elements: List[str]
    species: List[str]
    W: np.ndarray               # [K] kg/kmol
    composition: np.ndarray     # [K, E]
    T_ranges: np.ndarray        # [K, 3] (Tmin, Tmid, Tmax)
    nasa_low: np.ndarray        # [K, 7]
    nasa_high: np.ndarray       # [K, 7]
    reactions: List[Reaction]
    p_ref: float = ONE_ATM
    # per-species transport parameters (Cantera YAML `transport` blocks); None when the
    # file has none for some species
    transport: Optional[list] = None

    def transport_fits(self):
        """Temperature fits of the pure-species viscosities and conductivities and the
        binary diffusion coefficients (ember_amd.transport), over the temperature range
        where every species' NASA7 polynomials hold."""
        if self.transport is None:
            return None
        from . import transport as tr
        Tmin = float(self.T_ranges[:, 0].max())
        Tmax = float(self.T_ranges[:, 2].min())
        return tr.fit(self.transport, self.W, self.cp_R, Tmin, Tmax)

    @property
    def n_species(self) -> int:
        return len(self.species)

    @property
    def n_reactions(self) -> int:
        return len(self.reactions)

    def species_index(self, name: str) -> int:
        return self.species.index(name)

    # --- NASA7 thermo (Cantera NasaPoly2 semantics: low range if T <= Tmid)
    def cp_R(self, T: float) -> np.ndarray:
        a = np.where((T <= self.T_ranges[:, 1])[:, None], self.nasa_low, self.nasa_high)
        return a[:, 0] + T * (a[:, 1] + T * (a[:, 2] + T * (a[:, 3] + T * a[:, 4])))

    def h_RT(self, T: float) -> np.ndarray:
        a = np.where((T <= self.T_ranges[:, 1])[:, None], self.nasa_low, self.nasa_high)
        return (a[:, 0] + a[:, 1] * T / 2 + a[:, 2] * T**2 / 3 + a[:, 3] * T**3 / 4
                + a[:, 4] * T**4 / 5 + a[:, 5] / T)

    def s_R(self, T: float) -> np.ndarray:
        a = np.where((T <= self.T_ranges[:, 1])[:, None], self.nasa_low, self.nasa_high)
        return (a[:, 0] * np.log(T) + a[:, 1] * T + a[:, 2] * T**2 / 2 + a[:, 3] * T**3 / 3
                + a[:, 4] * T**4 / 4 + a[:, 6])

    def mixture_enthalpy_mass(self, T: float, Y: np.ndarray) -> float:
        """h [J/kg] of a mixture with mass fractions Y."""
        return float(np.sum(Y / self.W * self.h_RT(T)) * GAS_CONSTANT * T)

    def mole_to_mass(self, X: np.ndarray) -> np.ndarray:
        m = X * self.W
        return m / m.sum()

    # --- table file
    def write_table(self, path: str) -> None:
        """Write the flat ``.emxm`` table read by the C-ABI library."""
        K, I = self.n_species, self.n_reactions
        out = []
        out.append("EMXM 1")
        out.append(f"name {self.name}")
        out.append(f"counts {K} {I} {len(self.elements)}")
        out.append(f"pref {self.p_ref!r}")
        out.append("elements " + " ".join(self.elements))
        for k in range(K):
            vals = [self.W[k], *self.T_ranges[k], *self.nasa_low[k], *self.nasa_high[k],
                    *self.composition[k]]
            out.append(f"S {self.species[k]} " + " ".join(repr(float(v)) for v in vals))
        for r in self.reactions:
            if len(r.reactants) > 3 or len(r.products) > 3:
                raise ValueError(f"more than 3 reactant/product molecules: {r.equation}")
            rr = r.reactants + [-1] * (3 - len(r.reactants))
            pp = r.products + [-1] * (3 - len(r.products))
            # efficiencies in species-index order (Cantera BulkKinetics::addThirdBody
            # collects them in a std::map<size_t, double>)
            eff = sorted(r.efficiencies.items())
            fields = [r.rtype, int(r.reversible), len(r.reactants), *rr, len(r.products), *pp]
            fl = [r.A, r.b, r.EaR, r.A0, r.b0, r.EaR0, *r.troe]
            line = "R " + " ".join(str(int(v)) for v in fields) + " " + \
                " ".join(repr(float(v)) for v in fl) + f" {int(r.has_T2)} {r.default_eff!r} {len(eff)}"
            for k, e in eff:
                line += f" {k} {e!r}"
            out.append(line)
            if r.orders is not None:
                # ORD: forward reaction orders of reactant positions 0..2 (Cantera C_AnyN)
                oo = list(r.orders) + [0.0] * (3 - len(r.orders))
                out.append("ORD " + " ".join(repr(float(v)) for v in oo))
            if r.rtype == TYPE_PLOG:
                out.append(f"PLOG {len(r.plog)} " +
                           " ".join(repr(float(v)) for lvl in r.plog for v in lvl))
            elif r.rtype == TYPE_SRI:
                out.append("SRI " + " ".join(repr(float(v)) for v in r.sri))
            elif r.rtype == TYPE_CHEBYSHEV:
                nT, nP = r.cheb.shape
                out.append(f"CHEB {nT} {nP} " + " ".join(repr(float(v)) for v in r.cheb_range) + " " +
                           " ".join(repr(float(v)) for v in r.cheb.ravel()))
        fits = self.transport_fits()
        if fits is not None:
            # TRANSPORT Tmin Tmax, then per species TV k (sqrt(mu)/T^(1/4)) and TC k
            # (lambda/sqrt(T)), per pair j <= k TD j k (D_jk(P = 1 Pa)/T^(3/2)): five
            # coefficients of a polynomial in ln T each
            out.append(f"TRANSPORT {fits.Tmin!r} {fits.Tmax!r}")
            for k in range(K):
                out.append(f"TV {k} " + " ".join(repr(float(v)) for v in fits.visc[k]))
                out.append(f"TC {k} " + " ".join(repr(float(v)) for v in fits.cond[k]))
            from .transport import pair_index
            for j in range(K):
                for k in range(j, K):
                    out.append(f"TD {j} {k} " + " ".join(repr(float(v)) for v in fits.diff[pair_index(j, k, K)]))
        out.append("END")
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            f.write("\n".join(out) + "\n")
        os.replace(tmp, path)


_SPECIES_TOKEN = re.compile(r"^(\d+(?:\.\d+)?)?\s*(\S+)$")


def _parse_side(text: str, names: List[str]) -> Tuple[List[int], bool, Optional[str]]:
    """Parse one side of a Cantera equation into expanded species indices.

    Returns (indices, has_M, falloff_collider): the collider of a falloff reaction is
    "M" for "(+M)", the species name for a specific collider such as "(+AR)", and None
    otherwise.  Indices are ordered the way Cantera iterates a reaction's Composition
    map (sorted by species name), each repeated by its integer stoichiometric
    coefficient.
    """
    falloff = None
    m = re.search(r"\(\+\s*([^)\s]+)\s*\)", text)
    if m:
        falloff = m.group(1)
        if falloff != "M" and falloff not in names:
            raise ValueError(f"unknown falloff collider {falloff!r} in {text!r}")
        text = text[:m.start()] + text[m.end():]
    text = text.strip()
    # split on ' + ' (species names such as "CH2(S)" contain no spaces)
    toks = [t.strip() for t in re.split(r"\s\+\s", " " + text + " ") if t.strip()]
    counts: Dict[str, int] = {}
    has_M = False
    for t in toks:
        if t == "M":
            has_M = True
            continue
        m = re.match(r"^(\d+)\s+(\S+)$", t)
        if m:
            n, sp = int(m.group(1)), m.group(2)
        else:
            n, sp = 1, t
        if sp not in names:
            raise ValueError(f"unknown species {sp!r} in {text!r}")
        counts[sp] = counts.get(sp, 0) + n
    idx: List[int] = []
    for sp in sorted(counts):
        idx += [names.index(sp)] * counts[sp]
    return idx, has_M, falloff


def _pressure(v, units: dict) -> float:
    """A Cantera YAML pressure: a number in the file's pressure unit, or 'value unit'."""
    if isinstance(v, str):
        num, unit = v.split()
        return float(num) * PRESSURE_UNITS[unit]
    return float(v) * PRESSURE_UNITS[units.get("pressure", "Pa")]


def _unit_factors(units: dict) -> Tuple[float, float]:
    length = units.get("length", "m")
    quantity = units.get("quantity", "kmol")
    ea = units.get("activation-energy", "J/kmol")
    lf = {"cm": 0.01, "m": 1.0, "mm": 1e-3}[length]
    qf = {"mol": 1e-3, "kmol": 1.0, "molec": 1.0 / AVOGADRO}[quantity]
    conc_unit = lf**3 / qf   # one (length^3/quantity) expressed in m^3/kmol
    eaf = {"cal/mol": CAL * 1000.0, "kcal/mol": CAL * 1e6, "J/mol": 1000.0,
           "kJ/mol": 1e6, "J/kmol": 1.0, "K": GAS_CONSTANT,
           "eV": 1.602176634e-19 * AVOGADRO}[ea]
    return conc_unit, eaf


class _Yaml12Loader(yaml.SafeLoader):
    """SafeLoader with YAML 1.2 booleans: only true/false are booleans.

    Cantera reads mechanisms with a YAML 1.2 parser, where species names such
    as ``NO`` or ``ON`` are plain strings; PyYAML's default (YAML 1.1) would
    turn them into booleans.
    """


_Yaml12Loader.yaml_implicit_resolvers = {
    k: [(tag, rx) for tag, rx in v if tag != "tag:yaml.org,2002:bool"]
    for k, v in yaml.SafeLoader.yaml_implicit_resolvers.items()}
_Yaml12Loader.add_implicit_resolver(
    "tag:yaml.org,2002:bool", re.compile(r"^(?:true|True|TRUE|false|False|FALSE)$"),
    list("tTfF"))


_COMMON_KEYS = {"equation", "type", "duplicate", "note", "id"}
_TB_KEYS = {"efficiencies", "default-efficiency"}
# the fields each Cantera reaction type may carry here; anything else is refused
_KNOWN_REACTION_KEYS = {
    "elementary": _COMMON_KEYS | _TB_KEYS | {"rate-constant", "orders"},
    "three-body": _COMMON_KEYS | _TB_KEYS | {"rate-constant", "orders"},
    "falloff": _COMMON_KEYS | _TB_KEYS | {"high-P-rate-constant", "low-P-rate-constant", "Troe", "SRI",
                                          "orders"},
    "pressure-dependent-Arrhenius": _COMMON_KEYS | {"rate-constants", "orders"},
    "Chebyshev": _COMMON_KEYS | {"temperature-range", "pressure-range", "data", "orders"},
}


def _reaction_orders(rd: dict, eq: str, rev: bool, R: List[int], names: List[str]) -> Tuple[Optional[List[float]], float]:
    """Cantera's explicit forward orders ('orders'): the per-position orders of the
    reactant list R (None when every order equals the stoichiometric coefficient, as
    Cantera then keeps the plain concentration product) and the reaction's total forward
    order, which sets the units of its rate constant (Reaction::rate_units).  Refused as
    Cantera refuses them: on reversible reactions, for species that are not reactants
    (nonreactant-orders) and negative orders (negative-orders)."""
    spec = rd.get("orders")
    stoich: Dict[int, int] = {}
    for k in R:
        stoich[k] = stoich.get(k, 0) + 1
    if not spec:
        return None, float(len(R))
    if rev:
        raise ValueError(f"reaction {eq!r}: reaction orders may only be given for irreversible reactions")
    eff = {k: float(n) for k, n in stoich.items()}
    for sp, o in spec.items():
        if sp not in names or names.index(sp) not in stoich:
            raise ValueError(f"reaction {eq!r}: order for {sp!r}, which is not a reactant (nonreactant-orders)")
        o = float(o)
        if o < 0:
            raise ValueError(f"reaction {eq!r}: negative order for {sp!r} (negative-orders)")
        eff[names.index(sp)] = o
    total = float(sum(eff.values()))
    if all(eff[k] == stoich[k] for k in stoich):
        return None, total
    pos: List[float] = []
    seen = set()
    for k in R:
        pos.append(0.0 if k in seen else eff[k])
        seen.add(k)
    return pos, total


def load_yaml(path: str, phase: Optional[str] = None) -> Mechanism:
    with open(path) as f:
        doc = yaml.load(f, Loader=_Yaml12Loader)
    units = doc.get("units", {})
    conc_unit, eaf = _unit_factors(units)
    ph = doc["phases"][0] if phase is None else next(p for p in doc["phases"] if p["name"] == phase)
    if ph.get("thermo") != "ideal-gas":
        raise ValueError("only ideal-gas phases are supported")
    names = list(ph["species"])
    elements = list(ph["elements"])
    spdefs = {s["name"]: s for s in doc["species"]}
    K = len(names)
    W = np.zeros(K)
    comp = np.zeros((K, len(elements)))
    Tr = np.zeros((K, 3))
    lo = np.zeros((K, 7))
    hi = np.zeros((K, 7))
    for k, n in enumerate(names):
        s = spdefs[n]
        for el, c in s["composition"].items():
            comp[k, elements.index(el)] = c
            W[k] += c * ATOMIC_WEIGHTS[el]
        th = s["thermo"]
        if th["model"] != "NASA7":
            raise ValueError(f"species {n}: only NASA7 thermo is supported")
        tr = th["temperature-ranges"]
        if len(tr) != 3:
            raise ValueError(f"species {n}: two-range NASA7 expected")
        Tr[k] = tr
        lo[k] = th["data"][0]
        hi[k] = th["data"][1]

    def ea_R(v) -> float:
        return float(v) * eaf / GAS_CONSTANT

    rxns: List[Reaction] = []
    for rd in doc["reactions"]:
        eq = rd["equation"]
        rtype_s = rd.get("type", "elementary")
        if rtype_s not in _KNOWN_REACTION_KEYS:
            raise ValueError(f"unsupported reaction type {rtype_s!r}: {eq}")
        unknown = set(rd) - _KNOWN_REACTION_KEYS[rtype_s]
        if unknown:
            # e.g. 'orders' (explicit reaction orders) or 'negative-A': compiling the
            # reaction without them would give wrong rates, so refuse it
            raise ValueError(f"reaction {eq!r}: unsupported field(s) {sorted(unknown)}")
        rev = "<=>" in eq or ("<=" not in eq and "=>" not in eq and "=" in eq)
        if "<=>" in eq:
            lhs, rhs = eq.split("<=>")
        elif "=>" in eq:
            lhs, rhs = eq.split("=>")
            rev = False
        else:
            lhs, rhs = eq.split("=")
        R, mR, fR = _parse_side(lhs, names)
        P, mP, fP = _parse_side(rhs, names)
        rtype_s = rd.get("type", "elementary")
        # the forward order sets the rate constant's units: the explicit orders where given
        rorders, order = _reaction_orders(rd, eq, rev, R, names)
        if rtype_s == "three-body" or mR:
            rt = TYPE_THREE_BODY
            k = rd["rate-constant"]
            A = float(k["A"]) * conc_unit ** (order + 1 - 1)
            r = Reaction(eq, rt, rev, R, P, A, float(k["b"]), ea_R(k["Ea"]))
        elif rtype_s == "falloff" or fR is not None:
            hiP = rd["high-P-rate-constant"]
            loP = rd["low-P-rate-constant"]
            rt = TYPE_TROE if "Troe" in rd else TYPE_SRI if "SRI" in rd else TYPE_LINDEMANN
            r = Reaction(eq, rt, rev, R, P,
                         float(hiP["A"]) * conc_unit ** (order - 1), float(hiP["b"]), ea_R(hiP["Ea"]),
                         float(loP["A"]) * conc_unit ** order, float(loP["b"]), ea_R(loP["Ea"]))
            if "Troe" in rd:
                t = rd["Troe"]
                r.troe = (float(t["A"]), float(t["T3"]), float(t["T1"]), float(t.get("T2", 0.0)))
                r.has_T2 = "T2" in t
            if "SRI" in rd:
                t = rd["SRI"]
                r.sri = (float(t["A"]), float(t["B"]), float(t["C"]), float(t.get("D", 1.0)),
                         float(t.get("E", 0.0)))
            if fR is not None and fR != "M":
                # a specific collider: [M] is that species' concentration alone
                r.default_eff = 0.0
                r.efficiencies = {names.index(fR): 1.0}
        elif rtype_s == "elementary":
            k = rd["rate-constant"]
            r = Reaction(eq, TYPE_ELEMENTARY, rev, R, P,
                         float(k["A"]) * conc_unit ** (order - 1), float(k["b"]), ea_R(k["Ea"]))
        elif rtype_s == "pressure-dependent-Arrhenius":
            # Cantera PlogRate: Arrhenius expressions at listed pressures
            r = Reaction(eq, TYPE_PLOG, rev, R, P, 0.0, 0.0, 0.0)
            for k in rd["rate-constants"]:
                r.plog.append((_pressure(k["P"], units), float(k["A"]) * conc_unit ** (order - 1),
                               float(k["b"]), ea_R(k["Ea"])))
        elif rtype_s == "Chebyshev":
            # Cantera ChebyshevRate: log10 k = sum_tp a_tp T_t(Tr) T_p(Pr); the unit
            # conversion of k enters as log10 of its factor on a_00
            r = Reaction(eq, TYPE_CHEBYSHEV, rev, R, P, 0.0, 0.0, 0.0)
            Tmin, Tmax = (float(v) for v in rd["temperature-range"])
            Pmin, Pmax = (_pressure(v, units) for v in rd["pressure-range"])
            r.cheb_range = (Tmin, Tmax, Pmin, Pmax)
            c = np.array(rd["data"], dtype=float)
            if c.ndim != 2:
                raise ValueError(f"Chebyshev data must be a matrix: {eq}")
            c[0, 0] += np.log10(conc_unit ** (order - 1))
            r.cheb = c
        else:
            raise ValueError(f"unsupported reaction type {rtype_s!r}: {eq}")
        if r.rtype != TYPE_ELEMENTARY and not (fR is not None and fR != "M"):
            r.default_eff = float(rd.get("default-efficiency", 1.0))
            for sp, e in (rd.get("efficiencies") or {}).items():
                if sp in names:
                    r.efficiencies[names.index(sp)] = float(e)
        r.duplicate = bool(rd.get("duplicate", False))
        r.orders = rorders
        rxns.append(r)
    name = os.path.splitext(os.path.basename(path))[0]
    transport = None
    if all("transport" in spdefs[n] for n in names):
        from .transport import SpeciesTransport
        transport = [SpeciesTransport.from_yaml(spdefs[n]["transport"]) for n in names]
    return Mechanism(name, elements, names, W, comp, Tr, lo, hi, rxns, transport=transport)


def builtin_yaml(name: str) -> str:
    p = os.path.join(DATA_DIR, name if name.endswith(".yaml") else name + ".yaml")
    if not os.path.exists(p):
        raise FileNotFoundError(p)
    return p


def load(name: str) -> Mechanism:
    """Load a bundled mechanism ('gri30', 'h2o2') or a path to a YAML file."""
    if os.path.exists(name):
        return load_yaml(name)
    return load_yaml(builtin_yaml(name))


def table_path(name: str, out_dir: Optional[str] = None) -> str:
    """Compile mechanism `name` to an ``.emxm`` table (cached next to the YAML)."""
    src = name if os.path.exists(name) else builtin_yaml(name)
    base = os.path.splitext(os.path.basename(src))[0]
    out_dir = out_dir or DATA_DIR
    dst = os.path.join(out_dir, base + ".emxm")
    if not os.path.exists(dst) or os.path.getmtime(dst) < os.path.getmtime(src) \
            or os.path.getmtime(dst) < os.path.getmtime(__file__):
        load_yaml(src).write_table(dst)
    return dst
