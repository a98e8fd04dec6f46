"""Kinetics/thermo pins for the oracle (Cantera is not importable here, so these are
identities the mathematics fixes, not reference vectors).

- element and mass conservation of net production rates;
- thermodynamic equilibrium: at the HP-equilibrium state (reached by integrating the
  kinetics) every fast reversible reaction has Delta G / RT ~ 0, computed here in numpy
  from the NASA7 tables alone -- a wrong sign, Delta n or standard-state term in 1/Kc
  moves the kinetic equilibrium off the thermodynamic one;
- c_p = dh/dT;
- falloff limits: Pr -> 0 gives k0 [M], Pr -> infinity gives k_inf;
- adiabatic flame temperature of stoichiometric CH4/air ~2226 K (GRI-Mech 3.0).
"""
import json
import os

import numpy as np
import pytest

from ember_amd import mechanism as M
from ember_amd import workloads as W


@pytest.fixture(scope="module", params=["h2o2", "gri30"])
def mech(request, oracle_mod):
    m = M.load(request.param)
    return request.param, m, oracle_mod.Mech(M.table_path(request.param))


def _states(name, m, n=8, seed=0):
    rng = np.random.default_rng(seed)
    st = W.tanh_progress_states(n, seed=seed, name=name)
    st[:, 1] = rng.uniform(800, 2500, n)
    return st


def test_conservation(mech, oracle_mod):
    name, m, om = mech
    st = _states(name, m)
    r = om.rates(st[:, 1], st[:, 2:])
    w = r["wdot"]
    scale = np.abs(r["wdotQ"] * m.W).max(axis=1)
    np.testing.assert_allclose((w * m.W).sum(axis=1) / scale, 0, atol=1e-12)
    for e in range(len(m.elements)):
        num = (w * m.composition[:, e]).sum(axis=1)
        den = (np.abs(r["wdotQ"]) * m.composition[:, e]).sum(axis=1) + 1e-300
        np.testing.assert_allclose(num / den, 0, atol=1e-12)
    # creation - destruction == net
    np.testing.assert_allclose(r["wdotQ"] - r["wdotD"], w, rtol=1e-9,
                               atol=1e-12 * np.abs(r["wdotQ"]).max())


def test_reverse_rates_detailed_balance(mech, oracle_mod):
    """ropr/ropf = exp(Delta G / RT) with Delta G = Delta G0 + RT ln Q, for every reversible
    reaction at arbitrary states (third-body and falloff factors cancel).  Delta G is
    computed here in numpy from the NASA7 tables alone."""
    name, m, om = mech
    st = _states(name, m, n=6, seed=11)
    checked = 0
    for T, Y in zip(st[:, 1], st[:, 2:]):
        r = om.rates(T, Y)
        X = Y / m.W
        X = X / X.sum()
        g_rt = m.h_RT(T) - m.s_R(T) + np.log(np.maximum(X, 1e-300))  # P = P_ref
        for i, rx in enumerate(m.reactions):
            f, b = r["ropf"][0, i], r["ropr"][0, i]
            if not rx.reversible:
                assert b == 0.0
                continue
            if f <= 1e-250 or b <= 1e-250 or X[rx.reactants].min() < 1e-200 or X[rx.products].min() < 1e-200:
                continue
            dG = g_rt[rx.products].sum() - g_rt[rx.reactants].sum()
            assert abs(np.log(b / f) - dG) < 1e-9 * max(1.0, abs(dG)), (rx.equation, np.log(b / f), dG)
            checked += 1
    assert checked > 0.5 * 6 * sum(r.reversible for r in m.reactions)


def test_equilibrium_state(mech, oracle_mod):
    """At the integrated HP-equilibrium state the fast H/O reactions have Delta G ~ 0."""
    name, m, om = mech
    Tb, Yb = W.burned(m, name)
    X = Yb / m.W
    X = X / X.sum()
    g_rt = m.h_RT(Tb) - m.s_R(Tb) + np.log(np.maximum(X, 1e-300))
    ho = {m.species_index(s) for s in ["H2", "H", "O", "O2", "OH", "H2O", "HO2", "H2O2"]}
    checked = 0
    for rx in m.reactions:
        if rx.reversible and set(rx.reactants + rx.products) <= ho:
            dG = g_rt[rx.products].sum() - g_rt[rx.reactants].sum()
            assert abs(dG) < 1e-3, (rx.equation, dG)
            checked += 1
    assert checked >= 20


def test_cp_is_dh_dT(mech, oracle_mod):
    name, m, om = mech
    st = _states(name, m, n=4, seed=3)
    for T, Y in zip(st[:, 1], st[:, 2:]):
        h = lambda TT: (om.rates(TT, Y)["hk"][0] * (Y / m.W)).sum()   # J/kg
        cp = om.rates(T, Y)["cp"][0]
        dT = 1e-3
        assert abs((h(T + dT) - h(T - dT)) / (2 * dT) - cp) < 1e-6 * cp


def test_falloff_limits(oracle_mod):
    """Lindemann falloff: Pr -> 0 gives k0 [M], Pr -> infinity gives k_inf.  Troe: the
    broadening factor F = kf (1 + Pr) / (k_inf Pr) lies in (0, 1] and equals 1 when
    Pr = 0 is approached only through Fcent; both limits checked on Lindemann."""
    m = M.load("gri30")
    om = oracle_mod.Mech(M.table_path("gri30"))
    T = 1500.0
    Y = np.full(m.n_species, 1.0 / m.n_species)
    n_lind = 0
    for P, limit in [(1e-8, "low"), (1e18, "high"), (101325.0, "mid")]:
        r = om.rates(T, Y, oracle_mod.options(pressure=P))
        rho = r["rho"][0]
        C = rho * Y / m.W
        ctot = C.sum()
        for i, rx in enumerate(m.reactions):
            if rx.rtype not in (M.TYPE_TROE, M.TYPE_LINDEMANN):
                continue
            kf = r["ropf"][0, i] / np.prod(C[rx.reactants])
            kinf = rx.A * T ** rx.b * np.exp(-rx.EaR / T)
            Mc = rx.default_eff * ctot + sum((e - rx.default_eff) * C[k] for k, e in rx.efficiencies.items())
            k0 = rx.A0 * T ** rx.b0 * np.exp(-rx.EaR0 / T)
            Pr = k0 * Mc / kinf
            F = kf * (1 + Pr) / (kinf * Pr)
            if rx.rtype == M.TYPE_LINDEMANN:
                n_lind += 1
                assert abs(F - 1) < 1e-9, rx.equation
                if limit == "low":
                    assert abs(kf / (k0 * Mc) - 1) < 1e-9, rx.equation
                if limit == "high":
                    assert abs(kf / kinf - 1) < 1e-9, rx.equation
            else:
                assert 0 < F <= 1 + 1e-12, rx.equation
    assert n_lind == 9


def test_adiabatic_flame_temperature():
    with open(os.path.join(M.DATA_DIR, "gri30_burned.json")) as f:
        d = json.load(f)
    assert abs(d["T"] - 2226.0) < 10.0
    assert abs(d["enthalpy_error_rel"]) < 1e-6


def test_element_balance_of_mechanisms():
    for name in ("h2o2", "gri30"):
        m = M.load(name)
        for r in m.reactions:
            d = m.composition[r.products].sum(0) - m.composition[r.reactants].sum(0)
            assert np.abs(d).max() == 0, r.equation


def test_bdf_matches_radau(oracle_mod):
    """The oracle BDF at rtol 1e-6 against scipy Radau at rtol 1e-12 on the same RHS."""
    from scipy.integrate import solve_ivp
    m = M.load("h2o2")
    om = oracle_mod.Mech(M.table_path("h2o2"))
    st = W.tanh_progress_states(3, seed=7, name="h2o2")
    st[:, 1] = [1100.0, 1400.0, 1800.0]
    opts = oracle_mod.options(integrator=1, atol_species=1e-12)
    out, status, *_ = om.integrate(st, 2e-5, opts)
    assert (status == 0).all()
    for j in range(3):
        sol = solve_ivp(lambda t, y: om.rhs(y, opts), (0, 2e-5), st[j], method="Radau", rtol=1e-12,
                        atol=1e-20)
        ref = sol.y[:, -1]
        assert abs(out[j, 1] - ref[1]) < 1e-5 * ref[1]
        assert np.abs(out[j, 2:] - ref[2:]).max() < 1e-5 * max(ref[2:].max(), 1e-3)
