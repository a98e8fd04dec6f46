"""Thermodynamic properties of FlameSolver::updateChemicalProperties in the oracle
(orc_thermo_properties), pinned by:
- closed forms: Wmx = 1 / sum(Y'_k / W_k) with Y' = max(Y, 0) / sum max(Y, 0) (Cantera's
  setMassFractions), rho = P Wmx / (R T);
- the mixing rule cp = sum Y'_k cp_k / W_k and cp_k = d h_k / dT (central differences of
  the oracle's own h_k, away from the 1000 K polynomial midpoint);
- JANAF molar heat capacities at 1000 K (N2 32.698, O2 34.870, H2O 41.268, CO2 54.308
  J/mol/K) within 0.3%, the accuracy of the GRI-Mech NASA fits;
- the per-point values equal the ones the rates path reports (rho, cp, h_k).
"""
import numpy as np
import pytest

from ember_amd import mechanism
from ember_amd import workloads as W

R = 8314.46261815324  # J/kmol/K


@pytest.fixture(scope="module", params=["h2o2", "gri30"])
def case(request, oracle_mod):
    name = request.param
    m = mechanism.load(name)
    om = oracle_mod.Mech(mechanism.table_path(name))
    st = W.tanh_progress_states(40, seed=8, name=name)
    st[0, 2] = -1e-7        # a negative mass fraction: clipped
    st[1, 2:] *= 1.02       # unnormalised
    st[2, 1] = 400.0
    st[3, 1] = 3200.0
    return name, m, om, st


def test_closed_forms_and_mixing_rule(case):
    name, m, om, st = case
    t = om.thermo_properties(st)
    Yc = np.maximum(st[:, 2:], 0.0)
    Yc /= Yc.sum(axis=1, keepdims=True)
    Wmx = 1.0 / (Yc / m.W).sum(axis=1)
    np.testing.assert_allclose(t["Wmx"], Wmx, rtol=1e-14)
    np.testing.assert_allclose(t["rho"], 101325.0 * Wmx / (R * st[:, 1]), rtol=1e-14)
    np.testing.assert_allclose((Yc * t["cpSpec"] / m.W).sum(axis=1), t["cp"], rtol=1e-13)
    r = om.rates(st[:, 1], st[:, 2:])
    assert (t["rho"] == r["rho"]).all() and (t["cp"] == r["cp"]).all() and (t["hk"] == r["hk"]).all()


def test_cp_is_dh_dT_per_species(case):
    name, m, om, st = case
    sel = np.abs(st[:, 1] - 1000.0) > 5.0
    s = st[sel]
    dT = 1e-2
    hp = om.thermo_properties(np.column_stack([s[:, 0], s[:, 1] + dT, s[:, 2:]]))["hk"]
    hm = om.thermo_properties(np.column_stack([s[:, 0], s[:, 1] - dT, s[:, 2:]]))["hk"]
    cpk = om.thermo_properties(s)["cpSpec"]
    np.testing.assert_allclose((hp - hm) / (2 * dT), cpk, rtol=1e-7)


def test_janaf_heat_capacities(oracle_mod):
    m = mechanism.load("gri30")
    om = oracle_mod.Mech(mechanism.table_path("gri30"))
    st = np.zeros((1, m.n_species + 2))
    st[0, 1] = 1000.0
    st[0, 2:] = 1.0 / m.n_species
    cpk = om.thermo_properties(st)["cpSpec"][0] / 1000.0  # J/mol/K
    for sp, ref in (("N2", 32.698), ("O2", 34.870), ("H2O", 41.268), ("CO2", 54.308)):
        assert cpk[m.species_index(sp)] == pytest.approx(ref, rel=3e-3), sp
