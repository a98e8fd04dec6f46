"""The oracle checked against the reference's own known-answer tests.

- test/tridiagonalIntegrator.cpp:5-10, 64-102 (6x5 BDF table from test/reference/bdf.py)
- test/test_diffusionSystem.cpp:42-200 (closed forms and invariants)
- test/old/test_qssintegrator.cpp:52-186 (CHEMEQ2 cesium problem, reference final state)
"""
import numpy as np
import pytest

# test/tridiagonalIntegrator.cpp:5-10
SOLN = np.array([
    [0.00000000000000, 0.50000000000000, 2.00000000000000, 1.00000000000000, 0.00000000000000],
    [0.09475912852595, 0.63130024184639, 1.61199522551290, 1.01579521276719, 0.22818950052384],
    [0.17271644069234, 0.69321490131769, 1.34982270336185, 1.01014852549086, 0.38863616112939],
    [0.23179316689213, 0.71176849986673, 1.17610917328091, 0.99302967646570, 0.49567553015264],
    [0.27296869366712, 0.70706202264716, 1.05947277610782, 0.97139437597267, 0.56384834470269],
    [0.29912090096054, 0.69144453968429, 0.97840324923541, 0.94893150017619, 0.60507643387595]])
KAT_A = [0, 1, 1, 1, 1]
KAT_B = [-2, -2, -2, -2, -2]
KAT_C = [1, 1, 1, 1, 0]
KAT_K = [0, 0, 0, 0.2, 0.4]
KAT_Y0 = [0, 0.5, 2.0, 1.0, 0]


def test_tridiagonal_stepwise(oracle_mod):
    for j in range(6):
        y, t, h = oracle_mod.tridiag_integrate(KAT_A, KAT_B, KAT_C, KAT_K, KAT_Y0, 0.0, 0.2, mode=0, nsteps=j)
        np.testing.assert_allclose(y, SOLN[j], atol=1e-10, rtol=0)


def test_tridiagonal_full_and_short(oracle_mod):
    y, t, h = oracle_mod.tridiag_integrate(KAT_A, KAT_B, KAT_C, KAT_K, KAT_Y0, 0.0, 0.2, t1=1.0, mode=1)
    assert abs(h - 0.2) < 1e-10 and abs(t - 1.0) < 1e-10
    np.testing.assert_allclose(y, SOLN[5], atol=1e-10, rtol=0)
    y, t, h = oracle_mod.tridiag_integrate(KAT_A, KAT_B, KAT_C, KAT_K, KAT_Y0, 0.0, 1.3, t1=0.2, mode=1)
    assert abs(t - 0.2) < 1e-10
    np.testing.assert_allclose(y, SOLN[1], atol=1e-10, rtol=0)


def _diff(O, x, B, D, y0, t0, t1, h, alpha=0, left=0, right=0):
    N = len(x)
    st = O.diffusion_integrate(np.atleast_2d(B), np.zeros(1, np.int32), D[:, None], None, x, y0[:, None],
                               t0, t1, alpha=alpha, leftBC=left, rightBC=right, h_per_var=np.array([h]))
    return st[:, 0]


N401 = 401


def test_diffusion_constant(oracle_mod):
    x = np.linspace(-2, 2, N401)
    y = _diff(oracle_mod, x, np.ones(N401), np.full(N401, 0.1), np.full(N401, 5.0), 0.0, 1.0, 0.01)
    np.testing.assert_allclose(y, 5.0, atol=1e-11, rtol=0)


def test_diffusion_gaussian(oracle_mod):
    x = np.linspace(-2, 2, N401)
    D, t0, t = 0.1, 0.2, 0.3
    y = _diff(oracle_mod, x, np.ones(N401), np.full(N401, D), np.exp(-x * x / (4 * D * t0)), t0, t, 0.001)
    np.testing.assert_allclose(y, np.sqrt(t0 / t) * np.exp(-x * x / (4 * D * t)), atol=1e-4, rtol=0)


def test_diffusion_cylindrical(oracle_mod):
    x = np.linspace(0, 2, N401)
    D, t0, t = 0.1, 0.7, 0.8
    y = _diff(oracle_mod, x, np.ones(N401), np.full(N401, D), np.exp(-x * x / (4 * D * t0)), t0, t, 0.001,
              alpha=1)
    np.testing.assert_allclose(y[1:], (t0 / t * np.exp(-x * x / (4 * D * t)))[1:], atol=2e-4, rtol=0)


@pytest.mark.parametrize("variable", [False, True])
def test_diffusion_superposition(oracle_mod, variable):
    x = np.linspace(-2, 2, N401)
    D = 0.1 + 0.02 * x if variable else np.full(N401, 0.1)
    B = 1 + 0.5 * x * x if variable else np.ones(N401)
    ya, yb = np.cos(np.pi * x), np.sin(np.pi * x / 4.0)
    ra = _diff(oracle_mod, x, B, D, ya, 0.0, 0.1, 0.001)
    rb = _diff(oracle_mod, x, B, D, yb, 0.0, 0.1, 0.001)
    r = _diff(oracle_mod, x, B, D, ya + yb, 0.0, 0.1, 0.001)
    np.testing.assert_allclose(ra + rb, r, atol=1e-12, rtol=0)


def test_diffusion_timestep(oracle_mod):
    x = np.linspace(-2, 2, N401)
    D = 0.1 + 0.02 * x
    B = 1 + 0.5 * x * x
    y0 = np.cos(np.pi * x) + np.exp(-x * x / (4 * 0.1 * 0.2))
    a = _diff(oracle_mod, x, B, D, y0, 0.0, 0.1, 0.001)
    b = _diff(oracle_mod, x, B, D, y0, 0.0, 0.1, 0.0005)
    np.testing.assert_allclose(a, b, atol=5e-6, rtol=0)


def test_diffusion_zero_gradient_conserves(oracle_mod):
    """ZeroGradient at both ends, no source: sum_{j=1..N-2} dlj r y / B is invariant."""
    rng = np.random.default_rng(0)
    N = 64
    x = np.cumsum(rng.uniform(0.5, 1.5, N))
    B = rng.uniform(0.5, 2.0, N)
    D = rng.uniform(0.05, 0.2, N)
    y0 = rng.uniform(0, 1, N)
    y = _diff(oracle_mod, x, B, D, y0, 0.0, 0.5, 0.05, left=1, right=1)
    dlj = np.zeros(N)
    dlj[1:-1] = 0.5 * (x[2:] - x[:-2])
    w = dlj / B
    assert abs(np.sum(w[1:-1] * y[1:-1]) - np.sum(w[1:-1] * y0[1:-1])) < 1e-12 * np.sum(np.abs(w * y0))


def test_diffusion_dense_bruteforce(oracle_mod):
    """One BDF1 step of h/8 equals a dense solve of (I/(h/8) - A) y1 = y0/(h/8) + k."""
    rng = np.random.default_rng(1)
    N = 12
    x = np.cumsum(rng.uniform(0.5, 1.5, N))
    B = rng.uniform(0.5, 2.0, N)
    D = rng.uniform(0.05, 0.2, N)
    y0 = rng.uniform(0, 1, N)
    for left, right in [(0, 0), (1, 1), (2, 0), (3, 1)]:
        a = np.zeros(N); b = np.zeros(N); c = np.zeros(N)
        # build A from the closed forms of DiffusionSystem::get_A
        hh = np.diff(x)
        dlj = np.zeros(N); dlj[1:-1] = 0.5 * (x[2:] - x[:-2])
        c1 = np.zeros(N); c2 = np.zeros(N)
        c1[1:-1] = 0.5 * B[1:-1] / dlj[1:-1]
        c2[1:-1] = (D[1:-1] + D[2:]) / hh[1:]
        js, je = (2 if left == 1 else 1), (N - 2 if right == 1 else N - 1)
        for j in range(js, je):
            a[j], b[j], c[j] = c1[j] * c2[j - 1], -c1[j] * (c2[j - 1] + c2[j]), c1[j] * c2[j]
        if left == 1:
            b[1], c[1] = -c1[1] * c2[1], c1[1] * c2[1]
        if left == 2:
            c0 = B[0] * (D[0] + D[1]) / (2 * hh[0] ** 2); b[0], c[0] = -c0, c0
        if left == 3:
            c0 = B[0] / hh[0]; d = 0.5 * (D[0] + D[1]); b[0], c[0] = -c0 * (d / hh[0]), d * c0 / hh[0]
        if right == 1:
            a[N - 2], b[N - 2] = c1[N - 2] * c2[N - 3], -c1[N - 2] * c2[N - 3]
        A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        h = 0.08  # one step of h -> eight BDF1 substeps of h/8
        y = y0.copy()
        for _ in range(8):
            y = np.linalg.solve(np.eye(N) / (h / 8) - A, y / (h / 8))
        got = _diff(oracle_mod, x, B, D, y0, 0.0, h, h, left=left, right=right)
        np.testing.assert_allclose(got, y, rtol=1e-12, atol=1e-14)


# test/old/test_qssintegrator.cpp:155-186: initial state and reference state at t = 1000
CS_YI = [5.200e+02, 6.200e+02, 1.000e+12, 0.0, 3.600e+14, 1.400e+15]
CS_YF = np.array([2.59139492061e+04, 7.55718460300e+04, 1.53194051722e+03, 9.99999923516e+11,
                  3.59000000051e+14, 1.40000000000e+15])


def test_qss_cesium_converges(oracle_mod):
    """CHEMEQ2 cesium problem: the error against the reference final state shrinks
    with epsmin, as the QSS error estimate promises (itermax = 5 as in the test)."""
    errs = []
    for eps in [0.1, 0.01, 0.001, 1e-4]:
        y, err, g, r = oracle_mod.qss_test_problem(0, CS_YI, 1000.0, epsmin=eps, itermax=5, dtmax=1e3)
        assert err == 0
        rel = np.abs(y - CS_YF) / np.minimum(y, CS_YF)
        errs.append(rel.max())
    assert errs[-1] < 5e-6
    assert all(b < a for a, b in zip(errs, errs[1:]))
    # electron density by charge balance (test/old/test_qssintegrator.cpp:209-210)
    ne = y[1] - y[0]
    assert abs(ne - 4.96578968239e+04) / 4.96578968239e+04 < 1e-4


def test_qss_linear_exact(oracle_mod):
    """y' = q - k y with constant q, k: QSS reproduces q/k + (y0 - q/k) exp(-k t)."""
    q = np.array([1.0, 2.0, 0.5, 3.0])
    k = np.array([1e3, 1.0, 1e-2, 50.0])
    y0 = np.array([0.0, 5.0, 1.0, 1e-3])
    y, err, g, r = oracle_mod.qss_test_problem(1, y0, 0.5, epsmin=1e-3, epsmax=10.0, itermax=1,
                                               dtmax=1e-2, qc=q, kc=k, ymin=1e-60, abstol=1e-11,
                                               stability=False)
    assert err == 0
    exact = q / k + (y0 - q / k) * np.exp(-k * 0.5)
    np.testing.assert_allclose(y, exact, rtol=1e-6)


def test_qss_dtmin_failure(oracle_mod):
    """dt <= dtmin returns -1 (src/qssintegrator.cpp:308-325)."""
    y, err, g, r = oracle_mod.qss_test_problem(0, CS_YI, 1000.0, epsmin=0.01, itermax=5, dtmax=1e3,
                                               dtmin=1e3)
    assert err == -1
