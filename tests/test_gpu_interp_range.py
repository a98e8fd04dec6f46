"""InterpKinetics table range on the GPU.

Burned GRI-Mech 3.0 compositions at 3600 K lie outside Ember's default table grid (200
nodes over [273, 3500] K).  The device entry point reports status -2 and keeps the
states; the host-array entry point grows the grid by Ember's rule (src/chemistry0d.cpp:
279-291: 10 more nodes and 100 K more) and integrates them, matching the oracle on the
grown grid (210 nodes over [273, 3600] K) within the stage tolerance; a state at 250 K
grows the grid downwards (2 more nodes, Tmin = T - 20).  Points that did not fail keep
their results on the grid they ran on.
"""
import numpy as np
import pytest

from ember_amd import capi, mechanism
from ember_amd import workloads as W

pytestmark = pytest.mark.gpu

DT = 2e-5


def _hot_states(n=6):
    m = mechanism.load("gri30")
    return W.blend(m, "gri30", np.ones(n), np.random.default_rng(7), T_override=np.full(n, 3600.0))


def test_device_entry_reports_out_of_range(gpu):
    import torch
    print("[interp] device entry, 3600 K", flush=True)
    st = _hot_states()
    ctx = capi.Context("gri30", kinetics_model=capi.EMX_KIN_INTERP)
    x = torch.from_numpy(st).to(gpu)
    out = torch.empty_like(x)
    status = torch.zeros(len(st), dtype=torch.int32, device=gpu)
    ctx.source_integrate_device(x, out, DT, status=status)
    torch.cuda.synchronize()
    assert (status.cpu().numpy() == -2).all()
    np.testing.assert_array_equal(out.cpu().numpy(), st)
    assert ctx.interp_range() == (200, 273.0, 3500.0)


def test_host_entry_grows_tables_as_ember(gpu, oracle_mod):
    print("[interp] host entry grows the grid", flush=True)
    st = _hot_states()
    mix = np.concatenate([st, W.tanh_progress_states(4, seed=2, name="gri30")])
    ctx = capi.Context("gri30", kinetics_model=capi.EMX_KIN_INTERP)
    out, status, g, r = ctx.source_integrate(mix, DT)
    assert (status == 0).all(), status
    n, lo, hi = ctx.interp_range()
    assert (n, lo, hi) == (210, 273.0, 3600.0)
    # the hot points were integrated again on the grown grid, the others on the default one
    om = oracle_mod.Mech(mechanism.table_path("gri30"))
    ref = np.empty_like(mix)
    om.set_interp_range(n, lo, hi)
    ref[:len(st)], rstat, *_ = om.integrate(st, DT, oracle_mod.options(kinetics=1))
    assert (rstat == 0).all()
    om.set_interp_range(200, 273.0, 3500.0)
    ref[len(st):], rstat, *_ = om.integrate(mix[len(st):], DT, oracle_mod.options(kinetics=1))
    assert (rstat == 0).all()
    np.testing.assert_allclose(out[:, 1], ref[:, 1], rtol=1e-6)
    dY = np.abs(out[:, 2:] - ref[:, 2:])
    assert (dY <= 1e-6 * np.abs(ref[:, 2:]) + 1e-12).all(), dY.max()
    # the grown grid stays: the device entry now integrates the hot states too
    import torch
    x = torch.from_numpy(st).to(gpu)
    o2 = torch.empty_like(x)
    s2 = torch.zeros(len(st), dtype=torch.int32, device=gpu)
    ctx.source_integrate_device(x, o2, DT, status=s2)
    torch.cuda.synchronize()
    assert (s2.cpu().numpy() == 0).all()
    np.testing.assert_array_equal(o2.cpu().numpy(), out[:len(st)])


def test_host_entry_grows_tables_downwards(gpu, oracle_mod):
    print("[interp] host entry, 250 K", flush=True)
    m = mechanism.load("gri30")
    st = W.blend(m, "gri30", np.zeros(3), np.random.default_rng(8), T_override=np.full(3, 250.0))
    ctx = capi.Context("gri30", kinetics_model=capi.EMX_KIN_INTERP)
    out, status, *_ = ctx.source_integrate(st, DT)
    assert (status == 0).all(), status
    n, lo, hi = ctx.interp_range()
    assert (n, lo, hi) == (202, 230.0, 3500.0)
    om = oracle_mod.Mech(mechanism.table_path("gri30"))
    om.set_interp_range(n, lo, hi)
    ref, rstat, *_ = om.integrate(st, DT, oracle_mod.options(kinetics=1))
    assert (rstat == 0).all()
    np.testing.assert_allclose(out[:, 1], ref[:, 1], rtol=1e-6)
