"""Convection stage on the GPU (emx_convection_integrate[_device]) against the oracle's
ConvectionSystemSplit restatement, and the closed-form advection of a linear profile.

Flow: an opposed-flow-like field on the flame profile (U = a sqrt(rho_u / rho), inflow
rV(0) = rho_u 0.5 m/s from the left with the stagnation point inside the domain, so both
upwind directions occur), seeded density derivatives and split constants.  Both sides run
the Dormand-Prince 5(4) integrator with the same step control; sums over the workgroup run
in a different order on the GPU, so accept/reject decisions can differ in the last bits.
Bar: U, T, Y_k within 1e-6 relative (+ 1e-12 absolute for Y_k), over one step (5 us) and
over 5-14 accepted steps (0.1-0.2 ms).
"""
import numpy as np
import pytest

from ember_amd import capi, mechanism
from ember_amd import workloads as W

pytestmark = pytest.mark.gpu

R = mechanism.GAS_CONSTANT


def _case(name, n, seed, shift=0.0):
    case = W.flame_profile(n, seed=seed, name=name)
    m = mechanism.load(name)
    st = case.state.copy()
    Y = st[:, 2:]
    rho = mechanism.ONE_ATM / (R * st[:, 1] * (Y / m.W).sum(axis=1))
    rng = np.random.default_rng(seed)
    a = 150.0
    st[:, 0] = a * np.sqrt(rho[0] / rho)
    sc = np.zeros_like(st)
    sc[:, 0] = rng.normal(0, 50.0, n)
    sc[:, 1] = rng.normal(0, 1e4, n)
    sc[:, 2:] = rng.normal(0, 10.0, Y.shape) * Y
    drhodt = rng.normal(0, 1.0, n) * rho
    return case.x + shift, st, sc, drhodt, rho, a, m


@pytest.mark.parametrize("name,n,alpha,left,t1", [("h2o2", 150, 0, 0, 5e-6), ("gri30", 400, 0, 0, 5e-6),
                                                  ("h2o2", 150, 0, 0, 2e-4), ("h2o2", 150, 1, 2, 1e-4),
                                                  ("gri30", 300, 0, 2, 2e-4), ("h2o2", 2500, 0, 2, 1e-4)])
def test_convection_parity(gpu, oracle_mod, name, n, alpha, left, t1):
    print(f"[convection] {name} n={n} alpha={alpha} left_bc={left} t1={t1}", flush=True)
    x, st, sc, drhodt, rho, a, m = _case(name, n, 3, shift=0.006 if alpha else 0.0)
    Yleft = st[0, 2:].copy()
    # rV(0) = r^alpha rho v with v = 0.5 m/s at the first point
    kw = dict(alpha=alpha, beta=1, left_bc=left, right_bc=1, rVzero=rho[0] * 0.5 * x[0] ** alpha, Tleft=300.0,
              rhou=rho[0],
              strain_initial=a, strain_final=a)
    ctx = capi.Context(name)
    got, steps = ctx.convection_integrate(st, x, 0.0, t1, params=ctx.convection_params(Yleft=Yleft, **kw),
                                          split_const=sc, drhodt=drhodt)
    om = oracle_mod.Mech(mechanism.table_path(name))
    ref, rsteps = om.convection_integrate(x, st, 0.0, t1, params=oracle_mod.conv_params(**kw),
                                          split_const=sc, drhodt=drhodt, Yleft=Yleft)
    assert (steps > 0).all() and (rsteps > 0).all()
    rel = np.abs(got - ref) / (np.abs(ref) + 1e-12)
    print(f"[convection] steps gpu {steps[0]} oracle {rsteps[0]}; max rel diff U {rel[:, 0].max():.2e} "
          f"T {rel[:, 1].max():.2e} Y {rel[:, 2:].max():.2e}", flush=True)
    np.testing.assert_allclose(got[:, 0], ref[:, 0], rtol=1e-6, atol=1e-6 * np.abs(ref[:, 0]).max())
    np.testing.assert_allclose(got[:, 1], ref[:, 1], rtol=1e-6)
    dY = np.abs(got[:, 2:] - ref[:, 2:])
    assert (dY <= 1e-6 * np.abs(ref[:, 2:]) + 1e-12).all(), dY.max()
    # the stage changed the state
    assert np.abs(got[:, 1] - st[:, 1]).max() > 1e-3


def test_convection_device_entry_and_linear_advection(gpu):
    """A trace species' linear profile in uniform N2 moves at v = rVzero / rho exactly
    (first-order upwinding is exact on a linear profile; the inflow boundary is felt only
    within a few points of the left end), through the device entry point."""
    import torch
    print("[convection] linear advection, device entry", flush=True)
    m = mechanism.load("gri30")
    N = 80
    x = np.linspace(0.0, 0.008, N)
    st = np.zeros((N, m.n_species + 2))
    st[:, 1] = 300.0
    k = m.species.index("AR")
    st[:, 2 + m.species.index("N2")] = 1.0
    a0, b = 1e-9, 2e-7
    st[:, 2 + k] = a0 + b * x
    st[:, 2 + m.species.index("N2")] -= st[:, 2 + k]
    rho = mechanism.ONE_ATM * m.W[m.species.index("N2")] / (R * 300.0)
    v, t1 = 0.7, 5e-5
    ctx = capi.Context("gri30", integrator_species_abs_tol=1e-20)
    t = lambda arr: torch.from_numpy(np.ascontiguousarray(arr)).to(gpu)
    d_in = t(st)
    d_out = torch.empty_like(d_in)
    steps = torch.zeros(1 + m.n_species, dtype=torch.int32, device=gpu)
    ctx.convection_integrate_device(d_in, d_out, t(x), 0.0, t1,
                                    params=ctx.convection_params(rVzero=rho * v, rhou=rho), steps=steps)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    assert (steps.cpu().numpy() > 0).all()
    j = np.arange(8, N - 1)
    np.testing.assert_allclose(out[j, 2 + k], (a0 + b * (x - v * t1))[j], rtol=2e-8)
    np.testing.assert_allclose(out[:, 1], 300.0, rtol=1e-14)


def test_convection_refusals(gpu):
    ctx = capi.Context("h2o2")
    st = W.tanh_progress_states(3, seed=1, name="h2o2")
    with pytest.raises(capi.EmberAmdError, match="continuity"):
        ctx.convection_integrate(st, np.linspace(0, 1e-3, 3), 0.0, 1e-6,
                                 params=ctx.convection_params(continuity_bc=1))
    with pytest.raises(capi.EmberAmdError, match="4096"):
        ctx.convection_integrate(np.zeros((5000, ctx.n_vars)), np.linspace(0, 1, 5000), 0.0, 1e-6)
