"""Explicit forward reaction orders (Cantera "orders") in the mechanism reader and the oracle
(CPU), on the synthetic test mechanism tests/golden/h2o2_orders.yaml: H2 + O2 => 2 OH with
orders H2 0.5, O2 1.25; 2 H2 + O2 => 2 H2O with orders H2 1.0, O2 0.5; the three-body
H2O2 + M => 2 OH + M with order 1.5 for H2O2.

Pins:
- the forward rate of progress of each ordered reaction equals k_f prod_k C_k^o_k (times
  [M] from the efficiencies for the three-body one), with k_f computed here from the raw
  YAML numbers and this test's own unit table (cm, mol, cal/mol -> m, kmol, J; the rate
  constant's units follow the sum of the orders, Cantera's Reaction::rate_units);
- Cantera's C_AnyN rule: a reactant concentration <= 0 (no normalisation, as the BDF
  right-hand side sets the state) makes the ordered rate zero, while the mass-action
  reactions keep their products;
- the reader refuses orders Cantera refuses (reversible reactions, non-reactant species,
  negative orders), and orders equal to the stoichiometric coefficients leave the
  reaction a mass-action one;
- the adjoint refuses mechanisms with explicit orders.
"""
import os
import textwrap

import numpy as np
import pytest
import yaml

from ember_amd import mechanism

HERE = os.path.dirname(os.path.abspath(__file__))
YAML = os.path.join(HERE, "golden", "h2o2_orders.yaml")

CAL = 4184.0                      # J/kmol per cal/mol
R_GAS = 8314.46261815324          # J/kmol/K
CONC = 1e-3                       # m^3/kmol per cm^3/mol


@pytest.fixture(scope="module")
def mech():
    return mechanism.load(YAML)


@pytest.fixture(scope="module")
def om(oracle_mod):
    return oracle_mod.Mech(mechanism.table_path(YAML))


@pytest.fixture(scope="module")
def raw():
    with open(YAML) as f:
        return yaml.safe_load(f)


def _ordered(raw):
    return [(i, rd) for i, rd in enumerate(raw["reactions"]) if "orders" in rd]


def _kf(rd, T):
    k = rd["rate-constant"]
    order = sum(float(v) for v in rd["orders"].values())
    if rd.get("type") == "three-body":
        order += 1.0
    A = float(k["A"]) * CONC ** (order - 1.0)
    return A * T ** float(k["b"]) * np.exp(-float(k["Ea"]) * CAL / (R_GAS * T))


def _states(mech, n, seed):
    rng = np.random.default_rng(seed)
    Y = rng.uniform(0.01, 1.0, (n, mech.n_species))
    Y /= Y.sum(axis=1, keepdims=True)
    T = rng.uniform(900.0, 2500.0, n)
    return T, Y


def test_reader_orders(mech):
    ordered = [r for r in mech.reactions if r.orders is not None]
    assert len(ordered) == 3
    h2, o2, h2o2 = (mech.species_index(s) for s in ("H2", "O2", "H2O2"))
    assert ordered[0].reactants == [h2, o2] and ordered[0].orders == [0.5, 1.25]
    # repeated reactant positions: the first carries the species' order, the repeat 0
    assert ordered[1].reactants == [h2, h2, o2] and ordered[1].orders == [1.0, 0.0, 0.5]
    assert ordered[2].reactants == [h2o2] and ordered[2].orders == [1.5]
    assert all(not r.reversible for r in ordered)


def test_forward_rates_with_orders(oracle_mod, om, mech, raw):
    T, Y = _states(mech, 12, 3)
    out = om.rates(T, Y, oracle_mod.options())
    names = mech.species
    eff_h2o = {"H2O": 5.0, "AR": 0.6}
    for n in range(len(T)):
        C = out["rho"][n] * Y[n] / mech.W
        for i, rd in _ordered(raw):
            expect = _kf(rd, T[n])
            for sp, o in rd["orders"].items():
                expect *= C[names.index(sp)] ** float(o)
            if rd.get("type") == "three-body":
                M = sum(C[k] * eff_h2o.get(names[k], 1.0) for k in range(len(names)))
                expect *= M
            assert out["ropf"][n, i] == pytest.approx(expect, rel=1e-12)
            assert out["ropr"][n, i] == 0.0


def test_nonpositive_concentration_zeroes_the_ordered_rate(oracle_mod, om, mech, raw):
    T, Y = _states(mech, 2, 4)
    h2 = mech.species_index("H2")
    Y[0, h2] = 0.0
    Y[1, h2] = -1e-8  # no normalisation: the BDF right-hand side's state
    out = om.rates(T, Y, oracle_mod.options(), nonorm=True)
    for i, rd in _ordered(raw):
        if "H2" in rd["orders"]:
            assert out["ropf"][0, i] == 0.0 and out["ropf"][1, i] == 0.0
    # a mass-action reaction with H2 as reactant keeps its (zero / negative) product
    j = next(k for k, r in enumerate(mech.reactions) if r.equation == "H2 + O <=> H + OH")
    assert out["ropf"][0, j] == 0.0 and out["ropf"][1, j] < 0.0


def _mini(tmp_path, reaction):
    src = open(YAML).read()
    doc = yaml.safe_load(src)
    doc["reactions"] = [yaml.safe_load(textwrap.dedent(reaction))]
    p = tmp_path / "mini.yaml"
    p.write_text(yaml.safe_dump(doc))
    return str(p)


@pytest.mark.parametrize("reaction,msg", [
    ("""\
     equation: H2 + O2 <=> 2 OH
     rate-constant: {A: 1.0e+13, b: 0.0, Ea: 4.0e+04}
     orders: {H2: 0.5}
     """, "irreversible"),
    ("""\
     equation: H2 + O2 => 2 OH
     rate-constant: {A: 1.0e+13, b: 0.0, Ea: 4.0e+04}
     orders: {H2O: 0.5}
     """, "not a reactant"),
    ("""\
     equation: H2 + O2 => 2 OH
     rate-constant: {A: 1.0e+13, b: 0.0, Ea: 4.0e+04}
     orders: {H2: -0.5}
     """, "negative"),
])
def test_reader_refuses_what_cantera_refuses(tmp_path, reaction, msg):
    with pytest.raises(ValueError, match=msg):
        mechanism.load(_mini(tmp_path, reaction))


def test_stoichiometric_orders_stay_mass_action(tmp_path):
    m = mechanism.load(_mini(tmp_path, """\
        equation: 2 H2 + O2 => 2 H2O
        rate-constant: {A: 1.0e+13, b: 0.0, Ea: 4.0e+04}
        orders: {H2: 2, O2: 1}
        """))
    r = m.reactions[0]
    assert r.orders is None
    # rate constant units from the order sum 3: (cm^3/mol)^2/s
    assert r.A == pytest.approx(1.0e13 * CONC ** 2, rel=1e-15)


def test_adjoint_refuses_orders(oracle_mod, om, mech):
    from ember_amd import workloads as W
    st = W.flame_front_states(2, seed=5, name="h2o2")
    with pytest.raises(RuntimeError):
        om.adjoint(st, 2e-5, oracle_mod.options())
