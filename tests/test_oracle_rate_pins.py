"""Rate magnitudes and tables of the oracle pinned against the published mechanism data.

Conservation, detailed balance and equilibrium are blind to the size of the forward rate
constants.  These checks read the mechanism YAML here, convert its units with their own
table, and compare the oracle's rates of progress:

- elementary reactions: ropf = A T^b exp(-Ea / RT) prod C^nu' (Arrhenius, GRI-Mech units
  cm, mol, cal/mol converted to m, kmol, J);
- three-body reactions: ropf / (k_f prod C^nu') = [M] = sum_k eff_k C_k;
- falloff reactions: k = k_inf Pr / (1 + Pr) F with Lindemann F = 1 and the Troe F of
  Gilbert, Luther and Troe (Ber. Bunsenges. Phys. Chem. 87, 1983) written out here;
- two rate constants of GRI-Mech 3.0 worked by hand at 1500 K;
- InterpKinetics (src/chemistry0d.h:134-155, src/chemistry0d.cpp:291-331, 392-397,
  415-436): at the table nodes the interpolated k_f and 1/Kc equal the exact ones, in
  between they are the straight line through the two neighbouring nodes (which fixes the
  node index and the offset dT), and the interpolation error stays under h^2/8 max|k''|;
  a grown grid (Ember's growth rule) covers a 3600 K state;
- NASA7: the transcribed coefficient sets are continuous at Tmid, and the oracle uses the
  low range up to and including Tmid and the high range above (Cantera NasaPoly2).
"""
import re

import numpy as np
import pytest
import yaml

from ember_amd import mechanism as M
from ember_amd import workloads as W

R = 8314.46261815324                # J/kmol/K
LENGTH = {"cm": 1e-2, "m": 1.0, "mm": 1e-3}
QUANTITY = {"mol": 1e-3, "kmol": 1.0, "molec": 1.0 / 6.02214076e26}
ENERGY = {"cal/mol": 4184.0, "kcal/mol": 4.184e6, "J/mol": 1e3, "kJ/mol": 1e6, "J/kmol": 1.0}
ATOMIC = {"H": 1.008, "C": 12.011, "N": 14.007, "O": 15.999, "Ar": 39.95, "He": 4.002602}


def _raw(name):
    with open(M.builtin_yaml(name)) as f:
        return yaml.load(f, Loader=M._Yaml12Loader)


def _side(text):
    """{species: coefficient} of one side of an equation; third bodies dropped."""
    out = {}
    text = re.sub(r"\(\+\s*[A-Za-z0-9]+\s*\)", "", text)
    for tok in text.split(" + "):
        tok = tok.strip()
        if tok in ("M", ""):
            continue
        m = re.match(r"^(\d+)\s+(.+)$", tok)
        n, sp = (int(m.group(1)), m.group(2).strip()) if m else (1, tok)
        out[sp] = out.get(sp, 0) + n
    return out


def _sides(eq):
    for arrow in ("<=>", "=>", "="):
        if arrow in eq:
            lhs, rhs = eq.split(arrow)
            return _side(lhs), _side(rhs), arrow != "=>"
    raise ValueError(eq)


class Raw:
    """The mechanism as the YAML states it, in SI units by this file's own tables."""

    def __init__(self, name):
        doc = _raw(name)
        u = doc["units"]
        self.lf = LENGTH[u.get("length", "m")]
        self.qf = QUANTITY[u.get("quantity", "kmol")]
        self.ef = ENERGY[u.get("activation-energy", "J/kmol")]
        self.names = list(doc["phases"][0]["species"])
        spd = {s["name"]: s for s in doc["species"]}
        self.W = np.array([sum(c * ATOMIC[e] for e, c in spd[n]["composition"].items())
                           for n in self.names])
        self.nasa = [(spd[n]["thermo"]["temperature-ranges"], np.array(spd[n]["thermo"]["data"]))
                     for n in self.names]
        self.rxns = doc["reactions"]

    def conc_factor(self, order):
        """SI factor of a rate constant of the given concentration order, (m^3/kmol)^(order-1)."""
        return (self.lf ** 3 / self.qf) ** (order - 1)

    def arrhenius(self, k, order, T):
        return float(k["A"]) * self.conc_factor(order) * T ** float(k["b"]) * \
            np.exp(-float(k["Ea"]) * self.ef / (R * T))


def _states(name, n, seed):
    rng = np.random.default_rng(seed)
    st = W.tanh_progress_states(n, seed=seed, name=name)
    st[:, 1] = rng.uniform(700, 2600, n)
    # every species present, so that every concentration product is positive
    st[:, 2:] = st[:, 2:] + 1e-4
    st[:, 2:] /= st[:, 2:].sum(axis=1, keepdims=True)
    return st


@pytest.fixture(scope="module", params=["h2o2", "gri30"])
def case(request, oracle_mod):
    name = request.param
    raw = Raw(name)
    om = oracle_mod.Mech(M.table_path(name))
    st = _states(name, 6, 4)
    r = om.rates(st[:, 1], st[:, 2:])
    C = (r["rho"][:, None] * st[:, 2:]) / raw.W[None, :]
    return name, raw, st, r, C


def _cprod(raw, C, stoich):
    return np.prod([C[:, raw.names.index(s)] ** n for s, n in stoich.items()], axis=0)


def _M(raw, rd, C):
    eff = np.full(len(raw.names), float(rd.get("default-efficiency", 1.0)))
    for s, e in (rd.get("efficiencies") or {}).items():
        if s in raw.names:
            eff[raw.names.index(s)] = float(e)
    return C @ eff


def test_elementary_and_three_body_rates(case):
    name, raw, st, r, C = case
    T = st[:, 1]
    n_el = n_tb = 0
    for i, rd in enumerate(raw.rxns):
        reac, _, _ = _sides(rd["equation"])
        order = sum(reac.values())
        typ = rd.get("type", "elementary")
        base = _cprod(raw, C, reac)
        if typ == "elementary":
            k = raw.arrhenius(rd["rate-constant"], order, T)
            np.testing.assert_allclose(r["ropf"][:, i], k * base, rtol=1e-12, err_msg=rd["equation"])
            n_el += 1
        elif typ == "three-body":
            k = raw.arrhenius(rd["rate-constant"], order + 1, T)
            np.testing.assert_allclose(r["ropf"][:, i], k * _M(raw, rd, C) * base, rtol=1e-12,
                                       err_msg=rd["equation"])
            n_tb += 1
    assert n_el > 0 and n_tb > 0


def _troe_F(t, T, Pr):
    A, T3, T1 = float(t["A"]), float(t["T3"]), float(t["T1"])
    Fc = (1 - A) * np.exp(-T / T3) + A * np.exp(-T / T1)
    if "T2" in t:
        Fc = Fc + np.exp(-float(t["T2"]) / T)
    lFc = np.log10(np.maximum(Fc, 1e-300))
    c = -0.4 - 0.67 * lFc
    n = 0.75 - 1.27 * lFc
    lPr = np.log10(np.maximum(Pr, 1e-300))
    f1 = (lPr + c) / (n - 0.14 * (lPr + c))
    return 10.0 ** (lFc / (1.0 + f1 * f1))


def test_falloff_rates(case):
    name, raw, st, r, C = case
    T = st[:, 1]
    n = 0
    for i, rd in enumerate(raw.rxns):
        if rd.get("type") != "falloff" or "SRI" in rd:
            continue
        reac, _, _ = _sides(rd["equation"])
        order = sum(reac.values())
        kinf = raw.arrhenius(rd["high-P-rate-constant"], order, T)
        k0 = raw.arrhenius(rd["low-P-rate-constant"], order + 1, T)
        Pr = k0 * _M(raw, rd, C) / kinf
        F = _troe_F(rd["Troe"], T, Pr) if "Troe" in rd else 1.0
        k = kinf * Pr / (1 + Pr) * F
        np.testing.assert_allclose(r["ropf"][:, i], k * _cprod(raw, C, reac), rtol=1e-10,
                                   err_msg=rd["equation"])
        n += 1
    assert n > 0


def test_hand_checked_gri_rates(oracle_mod):
    """Two GRI-Mech 3.0 rate constants at 1500 K worked by hand from the published
    parameters:
      H + O2 <=> O + OH   A = 2.65e16 cm^3/mol/s, b = -0.6707, Ea = 17041 cal/mol
        k = 2.65e16 * 1500^-0.6707 * exp(-17041 / (1.98720 * 1500)) = 6.4598e11 cm^3/mol/s
          = 6.4598e8 m^3/kmol/s
      2 O + M <=> O2 + M  A = 1.2e17 cm^6/mol^2/s, b = -1, Ea = 0, efficiencies AR 0.83,
                          C2H6 3, CH4 2, CO 1.75, CO2 3.6, H2 2.4, H2O 15.4
        k = 1.2e17 / 1500 = 8.0e13 cm^6/mol^2/s = 8.0e7 m^6/kmol^2/s
    The oracle's ropf at a pure-O2 state (so [M] = C_O2 for the second) divided by the
    concentration products gives them back."""
    m = M.load("gri30")
    om = oracle_mod.Mech(M.table_path("gri30"))
    names = m.species
    Y = np.full(len(names), 1e-6)
    Y[names.index("O2")] = 1.0
    Y[names.index("O")] = 1e-3
    Y[names.index("H")] = 1e-3
    Y /= Y.sum()
    T = 1500.0
    r = om.rates(T, Y)
    C = r["rho"][0] * Y / m.W
    eqs = [rx.equation.replace(" ", "") for rx in m.reactions]
    i1 = eqs.index("H+O2<=>O+OH")
    k1 = r["ropf"][0, i1] / (C[names.index("H")] * C[names.index("O2")])
    assert abs(k1 / 6.4598e8 - 1) < 1e-4, k1
    i2 = eqs.index("2O+M<=>O2+M")
    effs = {"AR": 0.83, "C2H6": 3.0, "CH4": 2.0, "CO": 1.75, "CO2": 3.6, "H2": 2.4, "H2O": 15.4}
    Mc = sum(C[k] * effs.get(s, 1.0) for k, s in enumerate(names))
    k2 = r["ropf"][0, i2] / (C[names.index("O")] ** 2 * Mc)
    assert abs(k2 / 8.0e7 - 1) < 1e-12, k2


# ---------------------------------------------------------------------------
# InterpKinetics tables
# ---------------------------------------------------------------------------
def _kf_rkc(om, oracle_mod, T, Y, kinetics):
    """Per-reaction k_f (times third-body / falloff factors) and 1/Kc from the rates of
    progress at one composition: ropf / prod C^nu', ropr / ropf * prod C^nu' / prod C^nu''."""
    r = om.rates(T, np.broadcast_to(Y, (len(np.atleast_1d(T)), len(Y))),
                 oracle_mod.options(kinetics=kinetics))
    return r["ropf"], r["ropr"]


@pytest.fixture(scope="module")
def interp_case(oracle_mod):
    m = M.load("gri30")
    om = oracle_mod.Mech(M.table_path("gri30"))
    Y = np.full(m.n_species, 1.0 / m.n_species)
    arr = np.array([rx.rtype in (M.TYPE_ELEMENTARY, M.TYPE_THREE_BODY) for rx in m.reactions])
    rev = np.array([rx.reversible for rx in m.reactions])
    return m, om, Y, arr, rev


def test_interp_equals_exact_at_nodes(oracle_mod, interp_case):
    m, om, Y, arr, rev = interp_case
    n, lo, hi = 200, 273.0, 3500.0
    nodes = lo + np.arange(n) / (n - 1) * (hi - lo)
    T = nodes[[0, 1, 57, 120, 198, 199]]
    fi, ri = _kf_rkc(om, oracle_mod, T, Y, 1)
    fs, rs = _kf_rkc(om, oracle_mod, T, Y, 0)
    np.testing.assert_allclose(fi[:, arr], fs[:, arr], rtol=1e-11)
    np.testing.assert_allclose(ri[:, rev], rs[:, rev], rtol=1e-9)


def test_interp_is_the_chord_between_nodes(oracle_mod, interp_case):
    """At T = T_n + theta h the interpolated rate is (1 - theta) k(T_n) + theta k(T_n+1),
    with k at the nodes from the exact (standard) path.  Concentrations are held fixed by
    comparing ropf at one composition and density-independent ratios."""
    m, om, Y, arr, rev = interp_case
    n, lo, hi = 200, 273.0, 3500.0
    h = (hi - lo) / (n - 1)
    for j in (0, 33, 101, 197):
        Tn = lo + j * h
        thetas = np.array([1e-7, 0.25, 0.5, 0.999])
        T = Tn + thetas * h
        # rates of progress scale with density: compare k_f via ropf / ropf at a reference
        # state of the same T handled by the standard path is not possible, so use the
        # pure forward rate constants: elementary reactions with all reactant C at the
        # same density factor cancel in ropf_interp / ropf_std at equal T
        fi, ri = _kf_rkc(om, oracle_mod, T, Y, 1)
        fs, rs = _kf_rkc(om, oracle_mod, T, Y, 0)
        fa, ra = _kf_rkc(om, oracle_mod, np.array([Tn]), Y, 0)
        fb, rb = _kf_rkc(om, oracle_mod, np.array([Tn + h]), Y, 0)
        # k_f(T) exact = fs / prod(C(T)); the same prod(C(T)) divides fi, so the chord is
        # formed on k_f values obtained by removing each state's concentration product
        kfa = fa[0] / _conc_prod(m, om, oracle_mod, Tn, Y)
        kfb = fb[0] / _conc_prod(m, om, oracle_mod, Tn + h, Y)
        for q in range(len(thetas)):
            # the weight the table sees: the offset from the node of the rounded T
            th = (T[q] - Tn) / h
            cp = _conc_prod(m, om, oracle_mod, T[q], Y)
            kfi = fi[q] / cp
            chord = (1 - th) * kfa + th * kfb
            el = np.array([rx.rtype == M.TYPE_ELEMENTARY for rx in m.reactions])
            np.testing.assert_allclose(kfi[el], chord[el], rtol=1e-9, err_msg=f"node {j} theta {th}")
            # 1/Kc: ropr / ropf times the product ratio, independent of the concentrations
            rkci = ri[q, rev] / np.where(fi[q, rev] > 0, fi[q, rev], 1.0)
            rkca = ra[0, rev] / fa[0, rev]
            rkcb = rb[0, rev] / fb[0, rev]
            qa = _cprod_ratio(m, om, oracle_mod, Tn, Y)[rev]
            qb = _cprod_ratio(m, om, oracle_mod, Tn + h, Y)[rev]
            qi = _cprod_ratio(m, om, oracle_mod, T[q], Y)[rev]
            np.testing.assert_allclose(rkci / qi, (1 - th) * rkca / qa + th * rkcb / qb, rtol=1e-8)


def _conc(m, om, oracle_mod, T, Y):
    r = om.rates(np.array([T]), Y[None, :])
    return r["rho"][0] * Y / m.W


def _conc_prod(m, om, oracle_mod, T, Y):
    """prod_k C_k^nu'_k times [M] for three-body reactions, per reaction."""
    C = _conc(m, om, oracle_mod, T, Y)
    out = np.ones(m.n_reactions)
    for i, rx in enumerate(m.reactions):
        for k in rx.reactants:
            out[i] *= C[k]
        if rx.rtype == M.TYPE_THREE_BODY:
            eff = np.full(m.n_species, rx.default_eff)
            for k, e in rx.efficiencies.items():
                eff[k] = e
            out[i] *= C @ eff
    return out


def _cprod_ratio(m, om, oracle_mod, T, Y):
    C = _conc(m, om, oracle_mod, T, Y)
    out = np.ones(m.n_reactions)
    for i, rx in enumerate(m.reactions):
        for k in rx.products:
            out[i] *= C[k]
        for k in rx.reactants:
            out[i] /= C[k]
    return out


def test_interp_error_within_linear_bound(oracle_mod, interp_case):
    """Between nodes |k_interp - k| <= h^2/8 max|k''| on the interval (k'' by central
    differences of the exact Arrhenius expression on a fine grid)."""
    m, om, Y, arr, rev = interp_case
    n, lo, hi = 200, 273.0, 3500.0
    h = (hi - lo) / (n - 1)
    el = [i for i, rx in enumerate(m.reactions) if rx.rtype == M.TYPE_ELEMENTARY]
    for j in (2, 40, 150):
        Tn = lo + j * h
        T = Tn + 0.5 * h
        fi, _ = _kf_rkc(om, oracle_mod, np.array([T]), Y, 1)
        fs, _ = _kf_rkc(om, oracle_mod, np.array([T]), Y, 0)
        cp = _conc_prod(m, om, oracle_mod, T, Y)
        err = np.abs(fi[0] - fs[0]) / cp
        grid = np.linspace(Tn, Tn + h, 41)
        for i in el:
            rx = m.reactions[i]
            k = rx.A * grid ** rx.b * np.exp(-rx.EaR / grid)
            d2 = np.abs(np.diff(k, 2)) / (grid[1] - grid[0]) ** 2
            bound = h * h / 8 * d2.max() * 1.05 + 1e-13 * k.max()
            assert err[i] <= bound, (rx.equation, err[i], bound)


def test_interp_grown_grid_covers_hot_state(oracle_mod, interp_case):
    """A 3600 K burned state is outside Ember's default grid (status -2 from the stage); on the
    grid grown by Ember's rule (10 more nodes and 100 K more: 210 nodes over [273, 3600])
    the stage integrates and stays within the QSS tolerance of the exact kinetics."""
    m, om, Y, arr, rev = interp_case
    # burned (equilibrium) compositions heated to 3600 K: dissociation only cools them
    st = W.blend(m, "gri30", np.ones(4), np.random.default_rng(7), T_override=np.full(4, 3600.0))
    om.set_interp_range(200, 273.0, 3500.0)
    out, status, *_ = om.integrate(st, 2e-5, oracle_mod.options(kinetics=1))
    assert (status == -2).all()
    np.testing.assert_array_equal(out, st)
    om.set_interp_range(210, 273.0, 3600.0)
    try:
        out, status, *_ = om.integrate(st, 2e-5, oracle_mod.options(kinetics=1))
        ref, rstatus, *_ = om.integrate(st, 2e-5, oracle_mod.options(kinetics=0))
        assert (status == 0).all() and (rstatus == 0).all()
        assert (out[:, 1] < 3600.0).all()  # dissociation cools the gas
        np.testing.assert_allclose(out[:, 1], ref[:, 1], rtol=1e-4)
    finally:
        om.set_interp_range(200, 273.0, 3500.0)


# ---------------------------------------------------------------------------
# NASA7
# ---------------------------------------------------------------------------
def _nasa(a, T):
    cp = a[0] + a[1] * T + a[2] * T ** 2 + a[3] * T ** 3 + a[4] * T ** 4
    h = a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T
    s = a[0] * np.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + a[6]
    return cp, h, s


@pytest.mark.parametrize("name", ["h2o2", "gri30"])
def test_nasa7_continuity_at_tmid(name):
    raw = Raw(name)
    for sp, (tr, d) in zip(raw.names, raw.nasa):
        Tm = tr[1]
        cl, hl, sl = _nasa(d[0], Tm)
        ch, hh, sh = _nasa(d[1], Tm)
        assert abs(cl - ch) <= 1e-5 * abs(ch), (sp, cl, ch)
        assert abs(hl - hh) <= 1e-5 * abs(ch), (sp, hl, hh)   # h/RT against cp/R scale
        assert abs(sl - sh) <= 1e-5 * abs(sh), (sp, sl, sh)


@pytest.mark.parametrize("name", ["h2o2", "gri30"])
def test_nasa7_range_selection(oracle_mod, name):
    """The oracle's cp_k and h_k at Tmid come from the low-range coefficients and just
    above Tmid from the high-range ones (Cantera NasaPoly2: T <= Tmid is low)."""
    raw = Raw(name)
    om = oracle_mod.Mech(M.table_path(name))
    K = len(raw.names)
    checked = 0
    for Tm in sorted({tr[1] for tr, _ in raw.nasa}):
        for T, rng in ((Tm, 0), (np.nextafter(Tm, 2 * Tm), 1), (0.9 * Tm, 0), (1.5 * Tm, 1)):
            st = np.zeros((1, K + 2))
            st[0, 1] = T
            st[0, 2:] = 1.0 / K
            p = om.thermo_properties(st)
            for k, (tr, d) in enumerate(raw.nasa):
                if tr[1] != Tm:
                    continue
                checked += 1
                cp, h, s = _nasa(d[rng], T)
                assert abs(p["cpSpec"][0, k] - cp * R) <= 1e-12 * abs(cp * R), (raw.names[k], T)
                assert abs(p["hk"][0, k] - h * R * T) <= 1e-12 * max(abs(h * R * T), R * T), (raw.names[k], T)
    assert checked == 4 * K
