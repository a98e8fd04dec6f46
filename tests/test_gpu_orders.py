"""Reactions with explicit forward orders (Cantera "orders") on the GPU against the oracle,
on the synthetic test mechanism tests/golden/h2o2_orders.yaml (three irreversible global
reactions with fractional orders, one of them three-body).  The kernels take them in their
ordered instantiation (selected when the mechanism has such reactions).  Bars: rates within
1e-10 relative (the device forms c^o as exp(o ln c), the oracle with std::pow); heat release
1e-9; stages within the production-stage contract, |dT| <= 1e-6 T and
|dY| <= 1e-6 |Y| + 1e-12."""
import os

import numpy as np
import pytest

from ember_amd import capi, mechanism
from ember_amd import workloads as W

pytestmark = pytest.mark.gpu

YAML = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "h2o2_orders.yaml")


def _states(n, seed):
    # the H2/air flame-front states of the base mechanism (same species)
    return W.flame_front_states(n, seed=seed, name="h2o2")


# GRI-Mech 3.0 with three made-up global reactions with explicit orders appended: the
# six-round ordered kernels (the H2/O2 mechanism above has one round)
_GRI_EXTRA = """
- equation: CH4 + 2 O2 => CO2 + 2 H2O
  rate-constant: {A: 1.1e+10, b: 0.0, Ea: 4.8e+04}
  orders: {CH4: 0.7, O2: 0.8}
- equation: CO + OH => CO2 + H
  rate-constant: {A: 2.0e+10, b: 0.0, Ea: 3.0e+04}
  orders: {CO: 1.2}
- equation: 2 H2 + O2 => 2 H2O
  rate-constant: {A: 3.0e+11, b: 0.0, Ea: 4.0e+04}
  orders: {H2: 1.0, O2: 0.5}
"""


@pytest.fixture(scope="module")
def gri_orders(tmp_path_factory):
    src = open(mechanism.builtin_yaml("gri30")).read().rstrip("\n") + "\n" + _GRI_EXTRA.lstrip("\n")
    p = tmp_path_factory.mktemp("gri_orders") / "gri30_orders.yaml"
    p.write_text(src)
    return str(p)


@pytest.mark.parametrize("integrator", [0, 1])
def test_orders_gri_stage(gpu, oracle_mod, gri_orders, integrator):
    print(f"[orders] gri30 + ordered reactions, integrator={integrator}", flush=True)
    m = mechanism.load(gri_orders)
    assert sum(r.orders is not None for r in m.reactions) == 3
    st = W.flame_front_states(32, seed=5, name="gri30")
    ctx = capi.Context(gri_orders, chemistry_integrator=integrator, integrator_species_abs_tol=1e-12)
    wQ, wD, wN, rs = ctx.rates(st[:, 1], st[:, 2:])
    om = oracle_mod.Mech(mechanism.table_path(gri_orders, out_dir=os.path.dirname(gri_orders)))
    r = om.rates(st[:, 1], st[:, 2:], oracle_mod.options())
    scale = np.abs(r["wdotQ"]).max(axis=1, keepdims=True)
    assert (rs == 0).all()
    assert (np.abs(wN - r["wdot"]) <= 1e-10 * np.abs(r["wdot"]) + 1e-13 * scale).all()
    out, status, *_ = ctx.source_integrate(st, 2e-5)
    ref, rstat, *_ = om.integrate(st, 2e-5, oracle_mod.options(integrator=integrator, atol_species=1e-12))
    assert (status == 0).all() and (rstat == 0).all()
    assert (np.abs(out[:, 1] - ref[:, 1]) <= 1e-6 * ref[:, 1]).all()
    floor = 5e-11 if integrator == 1 else 1e-12  # the BDF's one-ulp sensitivity on GRI states
    dY = np.abs(out[:, 2:] - ref[:, 2:])
    assert (dY <= 1e-6 * np.abs(ref[:, 2:]) + floor).all(), dY.max()


@pytest.mark.parametrize("kinetics", [0, 1])
@pytest.mark.parametrize("nonorm", [False, True])
def test_orders_rates(gpu, oracle_mod, kinetics, nonorm):
    print(f"[orders] rates kinetics={kinetics} nonorm={nonorm}", flush=True)
    st = _states(32, 1)
    rng = np.random.default_rng(0)
    st[:, 1] = rng.uniform(600, 2800, 32)
    Y = st[:, 2:].copy()
    if nonorm:
        # a zero and a slightly negative H2 (the BDF right-hand side's unnormalised states):
        # the ordered rates with H2 vanish
        Y[0, 0] = 0.0
        Y[1, 0] = -1e-9
    ctx = capi.Context(YAML, kinetics_model=kinetics)
    wQ, wD, wN, status = ctx.rates(st[:, 1], Y, nonorm=nonorm)
    assert (status == 0).all()
    om = oracle_mod.Mech(mechanism.table_path(YAML))
    r = om.rates(st[:, 1], Y, oracle_mod.options(kinetics=kinetics), nonorm=nonorm)
    scale = np.abs(r["wdotQ"]).max(axis=1, keepdims=True)
    for got, ref in [(wQ, r["wdotQ"]), (wD, r["wdotD"]), (wN, r["wdot"])]:
        assert (np.abs(got - ref) <= 1e-10 * np.abs(ref) + 1e-13 * scale).all()


def test_orders_heat_release(gpu, oracle_mod):
    import torch
    print("[orders] heat release", flush=True)
    ctx = capi.Context(YAML)
    st = _states(64, 3)
    d = torch.from_numpy(st).to(gpu)
    q = torch.zeros(st.shape[0], dtype=torch.float64, device=gpu)
    status = torch.zeros(st.shape[0], dtype=torch.int32, device=gpu)
    ctx.heat_release_device(d, q, status)
    torch.cuda.synchronize()
    om = oracle_mod.Mech(mechanism.table_path(YAML))
    r = om.rates(st[:, 1], st[:, 2:])
    ref = -(r["wdot"] * r["hk"]).sum(axis=1)
    assert (status.cpu().numpy() == 0).all()
    np.testing.assert_allclose(q.cpu().numpy(), ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("integrator,kinetics", [(0, 0), (0, 1), (1, 0)])
def test_orders_stage(gpu, oracle_mod, integrator, kinetics):
    print(f"[orders] stage integrator={integrator} kinetics={kinetics}", flush=True)
    st = _states(48, 5)
    ctx = capi.Context(YAML, chemistry_integrator=integrator, kinetics_model=kinetics)
    out, status, *_ = ctx.source_integrate(st, 2e-5)
    om = oracle_mod.Mech(mechanism.table_path(YAML))
    ref, rstat, *_ = om.integrate(st, 2e-5, oracle_mod.options(integrator=integrator, kinetics=kinetics))
    assert (status == 0).all() and (rstat == 0).all()
    assert (np.abs(out[:, 1] - ref[:, 1]) <= 1e-6 * ref[:, 1]).all()
    dY = np.abs(out[:, 2:] - ref[:, 2:])
    assert (dY <= 1e-6 * np.abs(ref[:, 2:]) + 1e-12).all(), dY.max()


def test_orders_refusals(gpu):
    import torch
    print("[orders] refusals", flush=True)
    ctx = capi.Context(YAML)
    d = torch.from_numpy(_states(4, 5)).to(gpu)
    with pytest.raises(Exception):
        ctx.adjoint_forward(d, torch.empty_like(d), 2e-5)
    ctx.set_kernel_variant(1)
    with pytest.raises(Exception):
        ctx.source_integrate(_states(4, 5), 2e-5)
