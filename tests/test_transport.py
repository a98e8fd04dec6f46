"""Transport properties: the kinetic-theory fits of the mechanism compiler and the oracle's
ApproxMixTransport mixture rules, pinned against published gas data and closed forms.

Published values (1 atm): viscosity and thermal conductivity of N2, O2, Ar, CO, H2 and
CO2 at 300 K (NIST Chemistry WebBook / Poling, Prausnitz and O'Connell, The Properties of
Gases and Liquids, 5th ed.), air at 300 K (mu 18.46 uPa s, lambda 0.02624 W/m/K), binary
diffusion coefficients O2-N2 (0.202 cm^2/s, 293 K) and H2-N2 (0.76 cm^2/s, 293 K) (Marrero
and Mason, J. Phys. Chem. Ref. Data 1, 3, 1972).  Closed forms: a monatomic gas's
conductivity is 15/4 R mu / W (Eucken); a pure species' mixture properties are its own; in a
binary mixture the mass-weighted mixture diffusion coefficient equals the binary one.
Cantera's own fits are not available here: the fitted polynomials and therefore the
mixture properties match Cantera's only to the accuracy of the kinetic theory they both
implement (parity with Cantera's numbers is unpinned).
"""
import numpy as np
import pytest

from ember_amd import mechanism as M
from ember_amd import transport as Tr

R = Tr.GAS_CONSTANT


@pytest.fixture(scope="module")
def gri():
    m = M.load("gri30")
    return m, {n: k for k, n in enumerate(m.species)}


PURE_300K = {  # mu [uPa s], lambda [W/m/K]
    "N2": (17.89, 0.02598), "O2": (20.65, 0.02652), "AR": (22.70, 0.01772),
    "CO": (17.80, 0.02500), "H2": (8.96, 0.1856), "CO2": (15.02, 0.01664),
}


@pytest.mark.parametrize("name", sorted(PURE_300K))
def test_pure_species_at_300K(gri, name):
    m, idx = gri
    k = idx[name]
    mu = Tr.viscosity(m.transport[k], m.W[k], 300.0)
    lam = Tr.conductivity(m.transport[k], m.W[k], 300.0, m.cp_R(300.0)[k])
    mu_ref, lam_ref = PURE_300K[name]
    assert abs(mu * 1e6 / mu_ref - 1) < 0.04, (name, mu)
    assert abs(lam / lam_ref - 1) < 0.06, (name, lam)


def test_binary_diffusion_published(gri):
    m, idx = gri
    P = 101325.0
    for a, b, T, ref in (("O2", "N2", 293.15, 0.202e-4), ("H2", "N2", 293.15, 0.76e-4)):
        ka, kb = idx[a], idx[b]
        D = Tr.binary_diffusion(m.transport[ka], m.W[ka], m.transport[kb], m.W[kb], T, P)
        assert abs(D / ref - 1) < 0.03, (a, b, D)
        # symmetric, and inversely proportional to the pressure
        D2 = Tr.binary_diffusion(m.transport[kb], m.W[kb], m.transport[ka], m.W[ka], T, 2 * P)
        assert abs(D2 * 2 / D - 1) < 1e-14


def test_monatomic_conductivity_is_eucken(gri):
    """Ar and H (cp/R = 5/2 exactly in their NASA7 data): lambda = 15/4 R mu / W."""
    m, idx = gri
    T = np.array([300.0, 1000.0, 2500.0])
    for name in ("AR", "H"):
        k = idx[name]
        cp = np.array([m.cp_R(t)[k] for t in T])
        np.testing.assert_allclose(cp, 2.5, rtol=1e-8)
        lam = Tr.conductivity(m.transport[k], m.W[k], T, np.full(3, 2.5))
        mu = Tr.viscosity(m.transport[k], m.W[k], T)
        np.testing.assert_allclose(lam, 3.75 * R * mu / m.W[k], rtol=1e-13)


def test_fits_reproduce_kinetic_theory(gri):
    m, idx = gri
    f = m.transport_fits()
    assert (f.Tmin, f.Tmax) == (300.0, 3000.0)
    T = np.linspace(300.0, 3000.0, 97)
    K = m.n_species
    for k in range(K):
        np.testing.assert_allclose(Tr.fitted_viscosity(f, k, T),
                                   Tr.viscosity(m.transport[k], m.W[k], T), rtol=5e-3)
        cp = np.array([m.cp_R(t)[k] for t in T])
        np.testing.assert_allclose(Tr.fitted_conductivity(f, k, T),
                                   Tr.conductivity(m.transport[k], m.W[k], T, cp), rtol=1.5e-2)
    for j in (0, idx["N2"], idx["H2O"]):
        for k in range(K):
            np.testing.assert_allclose(Tr.fitted_diffusion(f, j, k, K, T),
                                       Tr.binary_diffusion(m.transport[j], m.W[j], m.transport[k], m.W[k], T),
                                       rtol=5e-3)


def _state(m, idx, X, T):
    """Point-major state (U, T, Y) of mole fractions X {name: x}."""
    x = np.zeros(m.n_species)
    for n, v in X.items():
        x[idx[n]] = v
    st = np.zeros((1, m.n_species + 2))
    st[0, 1] = T
    st[0, 2:] = m.mole_to_mass(x / x.sum())
    return st


def test_mixture_rules_limits(gri, oracle_mod):
    m, idx = gri
    om = oracle_mod.Mech(M.table_path("gri30"))
    f = m.transport_fits()
    K = m.n_species
    for name in ("N2", "CH4", "H2O"):
        k = idx[name]
        for T in (400.0, 1500.0):
            mu, lam, rhoD = om.transport_properties(_state(m, idx, {name: 1.0}, T))
            np.testing.assert_allclose(mu[0], Tr.fitted_viscosity(f, k, T), rtol=1e-12)
            np.testing.assert_allclose(lam[0], Tr.fitted_conductivity(f, k, T), rtol=1e-12)
    # binary mixtures: the mass-weighted mixture diffusion coefficient is the binary one
    for a, b in (("O2", "N2"), ("H2", "H2O"), ("CH4", "AR")):
        for xa in (0.1, 0.5, 0.9):
            T, P = 1200.0, 101325.0
            st = _state(m, idx, {a: xa, b: 1 - xa}, T)
            mu, lam, rhoD = om.transport_properties(st, pressure=P)
            Wmix = 1.0 / np.sum(st[0, 2:] / m.W)
            rho = P * Wmix / (R * T)
            D12 = Tr.fitted_diffusion(f, idx[a], idx[b], K, T, P)
            np.testing.assert_allclose(rhoD[0, idx[a]] / rho, D12, rtol=1e-12)
            np.testing.assert_allclose(rhoD[0, idx[b]] / rho, D12, rtol=1e-12)


def test_air_at_300K(gri, oracle_mod):
    m, idx = gri
    om = oracle_mod.Mech(M.table_path("gri30"))
    mu, lam, _ = om.transport_properties(_state(m, idx, {"N2": 0.781, "O2": 0.2095, "AR": 0.0093}, 300.0))
    assert abs(mu[0] / 18.46e-6 - 1) < 0.03, mu
    assert abs(lam[0] / 0.02624 - 1) < 0.04, lam


def test_threshold_selects_major_species(gri, oracle_mod):
    """Species below the threshold leave the viscosity sums; a minor species' diffusion
    coefficient sums over the major species only (ApproxMixTransport)."""
    m, idx = gri
    om = oracle_mod.Mech(M.table_path("gri30"))
    st = _state(m, idx, {"N2": 0.7, "O2": 0.2, "H2O": 0.1 - 1e-7, "OH": 1e-7}, 1800.0)
    mu0, lam0, d0 = om.transport_properties(st, threshold=0.0)
    mu1, lam1, d1 = om.transport_properties(st, threshold=1e-5)
    assert lam0[0] == lam1[0]             # conductivity sums over every species
    assert mu0[0] != mu1[0] and abs(mu0[0] / mu1[0] - 1) < 1e-5
    # with OH the only minor species the major-species sums differ only by the 1e-20
    # floors of the absent species and OH itself
    np.testing.assert_allclose(d1[0, idx["OH"]], d0[0, idx["OH"]], rtol=1e-6)
