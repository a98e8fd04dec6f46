"""Transport properties on the GPU (emx_transport_properties[_device]) against the oracle's
ApproxMixTransport restatement, on the same states and the same temperature fits.

Bar: mu, lambda and rho D_k within 1e-12 relative (the sums run in a different order on
the GPU: wave reductions instead of sequential loops).  States: tanh progress-variable
GRI-Mech 3.0 states (flame-front and burned points, where many species cross the
threshold), H2/O2 states, and three thresholds (0: every species major; Ember's 1e-5; 1e-3).
"""
import numpy as np
import pytest

from ember_amd import capi, mechanism
from ember_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,n", [("gri30", 700), ("h2o2", 129)])
@pytest.mark.parametrize("threshold", [0.0, 1e-5, 1e-3])
def test_transport_parity(gpu, oracle_mod, name, n, threshold):
    import torch
    print(f"[transport] {name} n={n} threshold={threshold}", flush=True)
    st = W.tanh_progress_states(n, seed=9, name=name)
    st[::7, 1] += 400.0  # temperatures off the blend line too
    ctx = capi.Context(name, transport_threshold=threshold)
    x = torch.from_numpy(st).to(gpu)
    mu = torch.empty(n, dtype=torch.float64, device=gpu)
    lam = torch.empty_like(mu)
    rhoD = torch.empty((n, ctx.n_species), dtype=torch.float64, device=gpu)
    ctx.transport_properties_device(x, mu, lam, rhoD)
    torch.cuda.synchronize()
    om = oracle_mod.Mech(mechanism.table_path(name))
    rmu, rlam, rrhoD = om.transport_properties(st, threshold=threshold)
    np.testing.assert_allclose(mu.cpu().numpy(), rmu, rtol=1e-12)
    np.testing.assert_allclose(lam.cpu().numpy(), rlam, rtol=1e-12)
    np.testing.assert_allclose(rhoD.cpu().numpy(), rrhoD, rtol=1e-12)
    # host-array entry point: the same numbers
    hmu, hlam, hrhoD = ctx.transport_properties(st)
    np.testing.assert_array_equal(hmu, mu.cpu().numpy())
    np.testing.assert_array_equal(hrhoD, rhoD.cpu().numpy())


def test_transport_refusals(gpu):
    """A mechanism without transport data is refused with a clear error."""
    import os
    print("[transport] refusals", flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    ctx = capi.Context(os.path.join(root, "tests", "golden", "h2o2_pdep.yaml"))
    st = W.tanh_progress_states(4, seed=9, name="h2o2")  # same 10 species
    with pytest.raises(capi.EmberAmdError, match="transport"):
        ctx.transport_properties(st)


def test_diffusion_coefficients_feed_the_split_step(gpu, oracle_mod):
    """emx_diffusion_coefficients_device: B = (1/rho, 1/(rho cp)) and D = (mu, lambda,
    rho D_k) per point as FlameSolver::prepareIntegrators builds them, against the oracle's
    thermo and transport properties; a device-resident split step driven by them matches
    one driven by the oracle's coefficients."""
    import torch
    print("[transport] diffusion coefficients -> split step", flush=True)
    name, n, dt = "gri30", 400, 1e-5
    case = W.flame_profile(n, seed=1, name=name)
    ctx = capi.Context(name, global_timestep=dt)
    nV = ctx.n_vars
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    x = t(case.state)
    B = torch.empty((2, n), dtype=torch.float64, device=gpu)
    D = torch.empty((n, nV), dtype=torch.float64, device=gpu)
    ctx.diffusion_coefficients_device(x, B, D)
    torch.cuda.synchronize()
    om = oracle_mod.Mech(mechanism.table_path(name))
    th = om.thermo_properties(case.state)
    mu, lam, rhoD = om.transport_properties(case.state)
    refB = np.stack([1.0 / th["rho"], 1.0 / (th["rho"] * th["cp"])])
    refD = np.concatenate([mu[:, None], lam[:, None], rhoD], axis=1)
    np.testing.assert_allclose(B.cpu().numpy(), refB, rtol=1e-12)
    np.testing.assert_allclose(D.cpu().numpy(), refD, rtol=1e-12)
    Brow = t(np.array([0, 1] + [0] * (nV - 2), np.int32))
    xs = t(case.x)
    h = np.full(nV, dt / 4)
    results = []
    for BB, DD in ((B, D), (t(refB), t(refD))):
        ss = capi.SplitState(t(case.state))
        status = torch.zeros(n, dtype=torch.int32, device=gpu)
        for k in range(3):
            ctx.split_step_device(ss, BB, DD, xs, k * dt, dt, B_row=Brow, h_per_var=h, status=status)
        torch.cuda.synchronize()
        assert (status.cpu().numpy() == 0).all()
        results.append(ss.state.cpu().numpy())
    np.testing.assert_allclose(results[0][:, 1], results[1][:, 1], rtol=1e-9)
    np.testing.assert_allclose(results[0][:, 2:], results[1][:, 2:], rtol=1e-6, atol=1e-12)
