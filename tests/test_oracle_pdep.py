"""Pressure-dependent rates in the oracle (CPU): Cantera's PlogRate, ChebyshevRate and
SriRate, on the synthetic test mechanism tests/golden/h2o2_pdep.yaml (H + O2 <=> O + OH as
PLOG with levels at 0.1, 1 (two rates) and 10 atm; H + HO2 <=> 2 OH as Chebyshev;
H + O2 (+M) <=> HO2 (+M) as SRI falloff).

Pins (k_f recovered from the oracle's forward rate of progress, k_f = ropf / (C_H C_X)):
- PLOG at a listed pressure equals that level's Arrhenius sum; between levels ln k is
  linear in ln P; below the first and above the last level the end level applies;
- Chebyshev equals numpy's 2-D Chebyshev series of (Tr, Pr) as defined by Cantera;
- InterpKinetics interpolates only Arrhenius rates: both stay exact with kinetics=interp;
- SRI falloff equals k_inf Pr/(1 + Pr) F with Cantera's SRI F.
"""
import os

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb

from ember_amd import mechanism

ATM = mechanism.ONE_ATM
HERE = os.path.dirname(os.path.abspath(__file__))
YAML = os.path.join(HERE, "golden", "h2o2_pdep.yaml")


@pytest.fixture(scope="module")
def mech():
    return mechanism.load(YAML)


@pytest.fixture(scope="module")
def om(oracle_mod):
    return oracle_mod.Mech(mechanism.table_path(YAML))


def _kf(oracle_mod, om, mech, i, T, P, kinetics=0):
    """k_f of reaction i at (T, P) from the oracle's forward rate of progress."""
    r = mech.reactions[i]
    Y = np.full(mech.n_species, 1.0 / mech.n_species)
    out = om.rates(np.array([T]), Y[None, :], oracle_mod.options(pressure=P, kinetics=kinetics))
    C = out["rho"][0] * Y / mech.W
    return out["ropf"][0, i] / np.prod(C[r.reactants])


def _arr(A, b, EaR, T):
    return A * T**b * np.exp(-EaR / T)


def test_plog(oracle_mod, om, mech):
    i = next(j for j, r in enumerate(mech.reactions) if r.rtype == mechanism.TYPE_PLOG)
    lv = mech.reactions[i].plog
    levels = {}
    for P, A, b, EaR in lv:
        levels.setdefault(P, []).append((A, b, EaR))
    Ps = sorted(levels)
    for T in (800.0, 1500.0, 2400.0):
        k = {P: sum(_arr(A, b, E, T) for A, b, E in levels[P]) for P in Ps}
        for P in Ps:  # listed pressures
            assert _kf(oracle_mod, om, mech, i, T, P) == pytest.approx(k[P], rel=1e-13)
        for P in (0.3 * ATM, 3.0 * ATM):  # between levels: linear in ln P
            lo = max(q for q in Ps if q <= P)
            hi = min(q for q in Ps if q > P)
            w = (np.log(P) - np.log(lo)) / (np.log(hi) - np.log(lo))
            ref = np.exp(np.log(k[lo]) + w * (np.log(k[hi]) - np.log(k[lo])))
            assert _kf(oracle_mod, om, mech, i, T, P) == pytest.approx(ref, rel=1e-12)
        assert _kf(oracle_mod, om, mech, i, T, 0.01 * ATM) == pytest.approx(k[Ps[0]], rel=1e-13)
        assert _kf(oracle_mod, om, mech, i, T, 50.0 * ATM) == pytest.approx(k[Ps[-1]], rel=1e-13)
    # the levels differ, so the interpolation is not trivially flat
    k800 = [sum(_arr(A, b, E, 800.0) for A, b, E in levels[P]) for P in Ps]
    assert np.ptp(np.log(k800)) > 0.1


def test_chebyshev(oracle_mod, om, mech):
    i = next(j for j, r in enumerate(mech.reactions) if r.rtype == mechanism.TYPE_CHEBYSHEV)
    r = mech.reactions[i]
    Tmin, Tmax, Pmin, Pmax = r.cheb_range
    for T in (350.0, 1000.0, 2900.0):
        for P in (0.02 * ATM, ATM, 40.0 * ATM):
            Tr = (2 / T - 1 / Tmin - 1 / Tmax) / (1 / Tmax - 1 / Tmin)
            Pr = (2 * np.log10(P) - np.log10(Pmin) - np.log10(Pmax)) / (np.log10(Pmax) - np.log10(Pmin))
            ref = 10 ** npcheb.chebval2d(Tr, Pr, r.cheb)
            assert _kf(oracle_mod, om, mech, i, T, P) == pytest.approx(ref, rel=1e-12)


def test_interp_kinetics_keeps_pdep_exact(oracle_mod, om, mech):
    for rtype in (mechanism.TYPE_PLOG, mechanism.TYPE_CHEBYSHEV):
        i = next(j for j, r in enumerate(mech.reactions) if r.rtype == rtype)
        for T in (900.0, 2100.0):
            a = _kf(oracle_mod, om, mech, i, T, 2.0 * ATM, kinetics=0)
            b = _kf(oracle_mod, om, mech, i, T, 2.0 * ATM, kinetics=1)
            assert b == pytest.approx(a, rel=1e-14)


def test_adjoint_refuses_pdep(oracle_mod, om, mech):
    st = np.zeros((1, mech.n_species + 2))
    st[0, 1] = 1500.0
    st[0, 2:] = 1.0 / mech.n_species
    with pytest.raises(RuntimeError):
        om.adjoint(st, 1e-6, opts=oracle_mod.options())


def test_sri_falloff(oracle_mod, om, mech):
    """SRI falloff (Cantera SriRate): k = k_inf Pr/(1 + Pr) F with Pr = k0 [M] / k_inf and
    F = d T^e [a exp(-b/T) + exp(-T/c)]^X, X = 1/(1 + log10^2 Pr); [M] from the
    reaction's efficiencies."""
    i = next(j for j, r in enumerate(mech.reactions) if r.rtype == mechanism.TYPE_SRI)
    r = mech.reactions[i]
    a, b, c, d, e = r.sri
    Y = np.full(mech.n_species, 1.0 / mech.n_species)
    for T in (600.0, 1300.0, 2500.0):
        for P in (0.1 * ATM, ATM, 30.0 * ATM):
            out = om.rates(np.array([T]), Y[None, :], oracle_mod.options(pressure=P))
            C = out["rho"][0] * Y / mech.W
            M = sum(C[k] * r.efficiencies.get(k, r.default_eff) for k in range(mech.n_species))
            kinf = _arr(r.A, r.b, r.EaR, T)
            pr = _arr(r.A0, r.b0, r.EaR0, T) * M / kinf
            X = 1.0 / (1.0 + np.log10(pr) ** 2)
            F = d * T**e * (a * np.exp(-b / T) + np.exp(-T / c)) ** X
            ref = kinf * pr / (1.0 + pr) * F
            got = out["ropf"][0, i] / np.prod(C[r.reactants])
            assert got == pytest.approx(ref, rel=1e-12)
            assert abs(F - 1.0) > 1e-3  # the SRI factor matters here


def test_specific_collider_falloff(oracle_mod, om, mech):
    """A falloff reaction written with "(+AR)" has argon as its only collider: [M] = C_AR
    (default efficiency 0), so its rate does not change when other species are swapped
    at fixed C_AR, and is the Lindemann form with that [M]."""
    i = next(j for j, r in enumerate(mech.reactions) if "(+AR)" in r.equation)
    r = mech.reactions[i]
    ar = mech.species_index("AR")
    assert r.rtype == mechanism.TYPE_LINDEMANN and r.default_eff == 0.0 and r.efficiencies == {ar: 1.0}
    T, P = 1400.0, ATM
    Y = np.full(mech.n_species, 1.0 / mech.n_species)
    out = om.rates(np.array([T]), Y[None, :], oracle_mod.options(pressure=P))
    C = out["rho"][0] * Y / mech.W
    kinf = _arr(r.A, r.b, r.EaR, T)
    pr = _arr(r.A0, r.b0, r.EaR0, T) * C[ar] / kinf
    ref = kinf * pr / (1.0 + pr)
    assert out["ropf"][0, i] / np.prod(C[r.reactants]) == pytest.approx(ref, rel=1e-12)
