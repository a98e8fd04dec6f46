"""The YAML mechanism reader refuses what it cannot compile.

A reaction field the compiler does not implement ('negative-A', sticking or Blowers-Masel
rates, orders for non-reactant species or negative orders) must raise instead of
compiling the reaction with mass-action orders or a different rate.  Explicit forward
orders of reactants on irreversible reactions are compiled (tests/test_oracle_orders.py).
"""
import os

import pytest
import yaml

from ember_amd import mechanism

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _with_reaction(tmp_path, extra):
    with open(mechanism.builtin_yaml("h2o2")) as f:
        doc = yaml.safe_load(f)
    doc["reactions"] = doc["reactions"][:3] + [extra]
    p = tmp_path / "mech.yaml"
    p.write_text(yaml.safe_dump(doc, sort_keys=False))
    return str(p)


def test_orders_compiled_or_refused(tmp_path):
    p = _with_reaction(tmp_path, {"equation": "H2 + O2 => 2 OH",
                                  "rate-constant": {"A": 1.0e13, "b": 0.0, "Ea": 0.0},
                                  "orders": {"H2": 0.5, "O2": 1.0}})
    assert mechanism.load_yaml(p).reactions[-1].orders == [0.5, 1.0]
    p = _with_reaction(tmp_path, {"equation": "H2 + O2 => 2 OH", "nonreactant-orders": True,
                                  "rate-constant": {"A": 1.0e13, "b": 0.0, "Ea": 0.0},
                                  "orders": {"H2": 0.5, "H2O": 1.0}})
    with pytest.raises(ValueError, match="nonreactant-orders"):
        mechanism.load_yaml(p)


def test_negative_A_refused(tmp_path):
    p = _with_reaction(tmp_path, {"equation": "H2 + O2 <=> 2 OH", "negative-A": True,
                                  "rate-constant": {"A": -1.0e13, "b": 0.0, "Ea": 0.0}})
    with pytest.raises(ValueError, match="negative-A"):
        mechanism.load_yaml(p)


def test_unknown_type_refused(tmp_path):
    p = _with_reaction(tmp_path, {"equation": "H2 + O2 <=> 2 OH", "type": "Blowers-Masel",
                                  "rate-constant": {"A": 1.0e13, "b": 0.0, "Ea0": 0.0, "w": 1e9}})
    with pytest.raises(ValueError, match="Blowers-Masel"):
        mechanism.load_yaml(p)


def test_bundled_mechanisms_still_load():
    assert len(mechanism.load("gri30").reactions) == 325
    assert len(mechanism.load("h2o2").reactions) == 29
    assert len(mechanism.load(os.path.join(ROOT, "tests", "golden", "h2o2_pdep.yaml")).reactions) == 30
