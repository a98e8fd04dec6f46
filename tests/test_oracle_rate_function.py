"""Ember's rate multiplier function in the oracle (CPU).

FlameSolver::rateMultiplierFunction is a ChebyshevFunction (src/scalarFunction.cpp:24-89)
whose coefficients the Python layer sets before every step (python/ember/_ember.pyx:533-543);
SourceSystemQSS::odefun and SourceSystemCVODE::f set every reaction's multiplier to a(t) at
their own time t (src/sourceSystem.cpp:133-135, :453-455), t = stage start + integrator time.

Pins:
- a(t) against numpy's Chebyshev series on the mapped interval, including the zero fill of
  ChebyshevFunction::setCoefficients for n = 2 and n = 3;
- a constant function gives the same stage, bit for bit, as per-reaction multipliers all
  equal to that constant (QSS and BDF);
- a function that varies over the stage, with the stage at a nonzero start time: the BDF
  stage against an independent solve (scipy Radau at rtol 1e-12) of the same ODE with the
  multiplier a(t) applied at each right-hand side's absolute time, and the QSS stage
  converging to that solve as epsmin -> 0.
"""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb

from ember_amd import mechanism
from ember_amd import workloads as W


@pytest.fixture(scope="module")
def h2(oracle_mod):
    return oracle_mod.Mech(mechanism.table_path("h2o2"))


def _cheb(data, t):
    T0, T1 = data[0], data[1]
    c = list(data[2:]) if len(data) >= 4 else ([data[2], 0.0] if len(data) == 3 else [0.0, 0.0])
    return npcheb.chebval(2 * (t - T0) / (T1 - T0) - 1, c)


@pytest.mark.parametrize("data", [
    [0.1, 0.3],
    [0.1, 0.3, 1.7],
    [0.1, 0.3, 1.2, -0.4],
    [2e-3, 2.02e-3, 1.0, 0.5, 0.25, -0.125, 0.0625],
    [-1.0, 4.0, 0.3, 0.1, -0.2, 0.05, 0.01, -0.003, 0.002, 1e-4],
])
def test_chebyshev_series(h2, data):
    h2.set_rate_function(data)
    try:
        ts = np.linspace(data[0], data[1], 41)
        got = np.array([h2.rate_function(t) for t in ts])
        np.testing.assert_allclose(got, _cheb(data, ts), rtol=1e-14, atol=1e-14)
    finally:
        h2.set_rate_function(None)
    assert h2.rate_function(0.2) == 1.0  # no function: unit multipliers


@pytest.mark.parametrize("integrator", [0, 1])
def test_constant_function_equals_uniform_multipliers(oracle_mod, h2, integrator):
    st = W.tanh_progress_states(12, seed=3, name="h2o2")
    o = oracle_mod.options(integrator=integrator, t_start=5e-3)
    m = 1.7
    h2.set_multipliers(np.full(h2.I, m))
    ref, s_ref, *_ = h2.integrate(st, 2e-5, opts=o)
    h2.set_multipliers(None)
    h2.set_rate_function([5e-3, 5.02e-3, m])
    try:
        got, s_got, *_ = h2.integrate(st, 2e-5, opts=o)
    finally:
        h2.set_rate_function(None)
    assert (s_ref == 0).all() and (s_got == 0).all()
    np.testing.assert_array_equal(got, ref)
    base, *_ = h2.integrate(st, 2e-5, opts=o)
    assert np.abs(base - ref).max() > 0  # the multiplier changed the stage


def test_varying_function_against_independent_solve(oracle_mod, h2):
    from scipy.integrate import solve_ivp
    t0, dt = 3.0e-3, 2e-5
    # a(t) rising from 0.5 to 2.5 over the stage, with curvature
    data = [t0, t0 + dt, 1.5, 1.0, 0.3]
    st = W.flame_front_states(3, seed=5, name="h2o2")
    o = oracle_mod.options(integrator=1, t_start=t0, rtol=1e-10, atol_species=1e-16,
                           atol_energy=1e-10, atol_momentum=1e-12)
    h2.set_rate_function(data)
    try:
        got, status, *_ = h2.integrate(st, dt, opts=o)
    finally:
        h2.set_rate_function(None)
    assert (status == 0).all()
    plain = oracle_mod.options(integrator=1)
    for j in range(st.shape[0]):
        def f(t, y):
            h2.set_multipliers(np.full(h2.I, _cheb(data, t0 + t)))
            return h2.rhs(y, opts=plain)
        sol = solve_ivp(f, (0.0, dt), st[j], method="Radau", rtol=1e-12, atol=1e-18)
        h2.set_multipliers(None)
        ref = sol.y[:, -1]
        assert abs(got[j, 1] - ref[1]) <= 1e-7 * ref[1]
        np.testing.assert_allclose(got[j, 2:], ref[2:], rtol=1e-6, atol=1e-12)
        # the function matters at this tolerance: a constant multiplier 1.5 misses it
        h2.set_multipliers(np.full(h2.I, 1.5))
        const = solve_ivp(lambda t, y: h2.rhs(y, opts=plain), (0.0, dt), st[j], method="Radau",
                          rtol=1e-12, atol=1e-18).y[:, -1]
        h2.set_multipliers(None)
        assert np.abs(const[2:] - ref[2:]).max() > 100 * (1e-6 * np.abs(ref[2:]).max())


def test_varying_function_qss_converges(oracle_mod, h2):
    """The QSS stage with the function converges to the independent solve as epsmin -> 0
    (QssIntegrator's error control is loose at Ember's epsmin 2e-2)."""
    from scipy.integrate import solve_ivp
    t0, dt = 3.0e-3, 2e-5
    data = [t0, t0 + dt, 1.5, 1.0, 0.3]
    st = W.flame_front_states(2, seed=5, name="h2o2")
    plain = oracle_mod.options(integrator=1)
    refs = []
    for j in range(st.shape[0]):
        def f(t, y):
            h2.set_multipliers(np.full(h2.I, _cheb(data, t0 + t)))
            return h2.rhs(y, opts=plain)
        refs.append(solve_ivp(f, (0.0, dt), st[j], method="Radau", rtol=1e-12, atol=1e-18).y[:, -1])
        h2.set_multipliers(None)
    refs = np.array(refs)
    errs = []
    for eps in (1e-3, 1e-5):
        o = oracle_mod.options(integrator=0, t_start=t0, qss_epsmin=eps, qss_dtmax=1e-7)
        h2.set_rate_function(data)
        try:
            got, status, *_ = h2.integrate(st, dt, opts=o)
        finally:
            h2.set_rate_function(None)
        assert (status == 0).all()
        errs.append(np.abs(got[:, 2:] - refs[:, 2:]).max() / np.abs(refs[:, 2:]).max())
    assert errs[1] < errs[0] / 10 and errs[1] < 3e-6, errs
