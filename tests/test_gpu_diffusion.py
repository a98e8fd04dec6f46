"""GPU diffusion stage vs the reference's known answers and the CPU oracle."""
import numpy as np
import pytest

from ember_amd import capi
from ember_amd import workloads as W

from test_oracle_kats import KAT_A, KAT_B, KAT_C, KAT_K, KAT_Y0, SOLN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(gpu):
    return capi.Context("h2o2")


def test_tridiagonal_kat(ctx):
    print("[diff] tridiagonal KAT", flush=True)
    y = ctx.tridiagonal_integrate(KAT_A, KAT_B, KAT_C, KAT_K, KAT_Y0, 0.2, 0.0, 1.0)
    np.testing.assert_allclose(y[0], SOLN[5], atol=1e-10, rtol=0)
    y = ctx.tridiagonal_integrate(KAT_A, KAT_B, KAT_C, KAT_K, KAT_Y0, 1.3, 0.0, 0.2)
    np.testing.assert_allclose(y[0], SOLN[1], atol=1e-10, rtol=0)


def _one(ctx, x, B, D, y0, t0, t1, h, alpha=0, left=0, right=0):
    st = ctx.diffusion_integrate(y0[:, None], B[None, :], D[:, None], x, t0, t1, alpha=alpha,
                                 left_bc=left, right_bc=right, h_per_var=np.array([h]))
    return st[:, 0]


def test_diffusion_closed_forms(ctx):
    print("[diff] closed forms", flush=True)
    N = 401
    x = np.linspace(-2, 2, N)
    D, t0, t = 0.1, 0.2, 0.3
    y = _one(ctx, x, np.ones(N), np.full(N, D), np.exp(-x * x / (4 * D * t0)), t0, t, 0.001)
    np.testing.assert_allclose(y, np.sqrt(t0 / t) * np.exp(-x * x / (4 * D * t)), atol=1e-4, rtol=0)
    y = _one(ctx, x, np.ones(N), np.full(N, 0.1), np.full(N, 5.0), 0.0, 1.0, 0.01)
    np.testing.assert_allclose(y, 5.0, atol=1e-11, rtol=0)
    xc = np.linspace(0, 2, N)
    t0, t = 0.7, 0.8
    y = _one(ctx, xc, np.ones(N), np.full(N, D), np.exp(-xc * xc / (4 * D * t0)), t0, t, 0.001, alpha=1)
    np.testing.assert_allclose(y[1:], (t0 / t * np.exp(-xc * xc / (4 * D * t)))[1:], atol=2e-4, rtol=0)
    Dv = 0.1 + 0.02 * x
    Bv = 1 + 0.5 * x * x
    ya, yb = np.cos(np.pi * x), np.sin(np.pi * x / 4.0)
    ra = _one(ctx, x, Bv, Dv, ya, 0.0, 0.1, 0.001)
    rb = _one(ctx, x, Bv, Dv, yb, 0.0, 0.1, 0.001)
    r = _one(ctx, x, Bv, Dv, ya + yb, 0.0, 0.1, 0.001)
    np.testing.assert_allclose(ra + rb, r, atol=1e-12, rtol=0)


@pytest.mark.parametrize("N", [37, 150, 4099, 20000])
@pytest.mark.parametrize("bc", [(0, 0), (1, 1), (2, 0), (3, 1)])
def test_diffusion_parity(ctx, oracle_mod, N, bc):
    print(f"[diff] parity N={N} bc={bc}", flush=True)
    rng = np.random.default_rng(N)
    nV = 12
    case = W.diffusion_case(n_vars=nV, n_points=N, seed=N)
    sc = rng.normal(0, 1.0, (N, nV))
    B = np.vstack([case.B[0], 1.3 * case.B[0]])
    B_row = rng.integers(0, 2, nV).astype(np.int32)
    left, right = bc
    alpha = 1 if left == 2 else 0
    kw = dict(alpha=alpha, y_inf=0.3, wall_const=2.0, wall_var=1)
    # Ember's step rule (several BDF2 steps on the denser grids); the largest grid
    # gets explicit steps so that the oracle finishes quickly
    t0, t1 = 1e-3, 1e-3 + 5e-6
    hv = None if N <= 4099 else np.full(nV, (t1 - t0) / 3)
    got = ctx.diffusion_integrate(case.y0, B, case.D, case.x, t0, t1, B_row=B_row, split_const=sc,
                                  left_bc=left, right_bc=right, h_per_var=hv, **kw)
    ref = oracle_mod.diffusion_integrate(B, B_row, case.D, sc, case.x, case.y0, t0, t1, alpha=alpha,
                                         leftBC=left, rightBC=right, yInf=0.3, wallConst=2.0, wallVar=1,
                                         multiplier=10.0, dt_global=2e-5, h_per_var=hv)
    scale = np.abs(ref).max(axis=0, keepdims=True)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12 * scale.max())


def test_diffusion_full_size(ctx, oracle_mod):
    """53 species x 65536 points, one stage: GPU vs the oracle's Thomas solves."""
    print("[diff] full size 53 x 65536", flush=True)
    case = W.diffusion_case()
    t0, t1 = 0.0, 2e-5
    h = np.full(53, 2e-5)
    got = ctx.diffusion_integrate(case.y0, case.B, case.D, case.x, t0, t1, B_row=case.B_row, h_per_var=h)
    ref = oracle_mod.diffusion_integrate(case.B, case.B_row, case.D, None, case.x, case.y0, t0, t1,
                                         h_per_var=h)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-14)
