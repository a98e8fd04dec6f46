"""The density derivative of the convection stage (FlameSolver::prepareIntegrators,
src/flameSolver.cpp:203-213) in the oracle, on the CPU.

Pins:
- drhodt is the time derivative of the ideal-gas density rho(T, Y) = P Wmx / (R T) along
  (dT/dt, dY/dt) at constant pressure: a central difference of the oracle's own density
  (thermo_properties) reproduces it.  The species rates conserve mass (sum_k dY_k/dt = 0),
  so the normalisation of Y in setMassFractions does not enter;
- the split rates enter only through their sum ddtConv + ddtDiff + ddtProd, and ddtCross
  adds to it with balanced splitting only.
"""
import numpy as np
import pytest

from ember_amd import mechanism
from ember_amd import workloads as W


@pytest.fixture(scope="module")
def om(oracle_mod):
    return oracle_mod.Mech(mechanism.table_path("gri30"))


def _rates(st, seed):
    rng = np.random.default_rng(seed)
    n, nv = st.shape
    ddt = np.zeros_like(st)
    ddt[:, 0] = rng.normal(0.0, 10.0, n)
    ddt[:, 1] = rng.normal(0.0, 1e5, n)
    dY = rng.normal(0.0, 1.0, (n, nv - 2)) * st[:, 2:]
    dY -= st[:, 2:] * dY.sum(axis=1, keepdims=True)  # sum_k dY_k/dt = 0 (sum_k Y_k = 1)
    ddt[:, 2:] = 50.0 * dY
    return ddt


def test_density_derivative_is_the_ideal_gas_rate(om):
    st = W.flame_front_states(24, seed=3, name="gri30")
    st[:, 2:] /= st[:, 2:].sum(axis=1, keepdims=True)
    ddt = _rates(st, 1)
    th = om.thermo_properties(st)
    z = np.zeros_like(st)
    got = om.density_derivative(st, 0.25 * ddt, 0.5 * ddt, 0.25 * ddt, th["rho"], th["Wmx"])
    eps = 1e-7
    rp = om.thermo_properties(st + eps * ddt)["rho"]
    rm = om.thermo_properties(st - eps * ddt)["rho"]
    fd = (rp - rm) / (2 * eps)
    assert (np.abs(got) > 1e-3 * th["rho"]).all()  # the rates move the density
    np.testing.assert_allclose(got, fd, rtol=1e-7)
    # a temperature rate alone: drhodt = -rho dT/dt / T
    only_T = np.zeros_like(st)
    only_T[:, 1] = ddt[:, 1]
    got_T = om.density_derivative(st, only_T, z, z, th["rho"], th["Wmx"])
    np.testing.assert_allclose(got_T, -th["rho"] * ddt[:, 1] / st[:, 1], rtol=1e-15)


def test_split_rates_enter_through_their_sum(om):
    st = W.flame_front_states(16, seed=5, name="gri30")
    ddt = _rates(st, 2)
    th = om.thermo_properties(st)
    z = np.zeros_like(st)
    ref = om.density_derivative(st, ddt, z, z, th["rho"], th["Wmx"])
    for parts in [(z, ddt, z), (z, z, ddt), (0.5 * ddt, 0.25 * ddt, 0.25 * ddt)]:
        np.testing.assert_allclose(om.density_derivative(st, *parts, th["rho"], th["Wmx"]), ref,
                                   rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    # ddtCross: added with balanced splitting, ignored with Strang splitting
    cross = _rates(st, 3)
    bal = om.density_derivative(st, ddt, z, z, th["rho"], th["Wmx"], ddt_cross=cross, balanced=True)
    both = om.density_derivative(st, ddt + cross, z, z, th["rho"], th["Wmx"])
    np.testing.assert_allclose(bal, both, rtol=1e-13, atol=1e-13 * np.abs(both).max())
    strang = om.density_derivative(st, ddt, z, z, th["rho"], th["Wmx"], ddt_cross=cross, balanced=False)
    np.testing.assert_array_equal(strang, ref)
