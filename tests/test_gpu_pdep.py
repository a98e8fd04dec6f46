"""PLOG, Chebyshev and SRI-falloff reactions on the GPU against the oracle, on the
synthetic test mechanism tests/golden/h2o2_pdep.yaml, at pressures below, between, on and
above the PLOG levels.  Bars: rates within 1e-10 relative; stages within the production-stage contract,
|dT| <= 1e-6 T and |dY| <= 1e-6 |Y| + 1e-12."""
import os

import numpy as np
import pytest

from ember_amd import capi, mechanism
from ember_amd import workloads as W

pytestmark = pytest.mark.gpu

ATM = mechanism.ONE_ATM
YAML = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "h2o2_pdep.yaml")


def _states(n, seed):
    # the H2/air flame-front states of the base mechanism (same species)
    return W.flame_front_states(n, seed=seed, name="h2o2")


@pytest.mark.parametrize("P", [0.05 * ATM, 0.3 * ATM, ATM, 3.0 * ATM, 20.0 * ATM])
@pytest.mark.parametrize("kinetics", [0, 1])
def test_pdep_rates(gpu, oracle_mod, P, kinetics):
    print(f"[pdep] rates P={P:.0f} kinetics={kinetics}", flush=True)
    st = _states(32, 1)
    rng = np.random.default_rng(0)
    st[:, 1] = rng.uniform(400, 3000, 32)
    ctx = capi.Context(YAML, pressure=P, kinetics_model=kinetics)
    wQ, wD, wN, status = ctx.rates(st[:, 1], st[:, 2:])
    assert (status == 0).all()
    om = oracle_mod.Mech(mechanism.table_path(YAML))
    r = om.rates(st[:, 1], st[:, 2:], oracle_mod.options(pressure=P, kinetics=kinetics))
    scale = np.abs(r["wdotQ"]).max(axis=1, keepdims=True)
    for got, ref in [(wQ, r["wdotQ"]), (wD, r["wdotD"]), (wN, r["wdot"])]:
        assert (np.abs(got - ref) <= 1e-10 * np.abs(ref) + 1e-13 * scale).all()


@pytest.mark.parametrize("P,integrator", [(3.0 * ATM, 0), (0.3 * ATM, 0), (3.0 * ATM, 1)])
def test_pdep_stage(gpu, oracle_mod, P, integrator):
    print(f"[pdep] stage P={P:.0f} integrator={integrator}", flush=True)
    st = _states(48, 5)
    ctx = capi.Context(YAML, pressure=P, chemistry_integrator=integrator)
    out, status, *_ = ctx.source_integrate(st, 2e-5)
    om = oracle_mod.Mech(mechanism.table_path(YAML))
    ref, rstat, *_ = om.integrate(st, 2e-5, oracle_mod.options(pressure=P, integrator=integrator))
    assert (status == 0).all() and (rstat == 0).all()
    assert (np.abs(out[:, 1] - ref[:, 1]) <= 1e-6 * ref[:, 1]).all()
    dY = np.abs(out[:, 2:] - ref[:, 2:])
    assert (dY <= 1e-6 * np.abs(ref[:, 2:]) + 1e-12).all(), dY.max()


def test_pdep_refusals(gpu):
    import torch
    print("[pdep] refusals", flush=True)
    ctx = capi.Context(YAML)
    d = torch.from_numpy(_states(4, 5)).to(gpu)
    with pytest.raises(Exception):
        ctx.adjoint_forward(d, torch.empty_like(d), 2e-5)
    ctx.set_kernel_variant(1)
    with pytest.raises(Exception):
        ctx.source_integrate(_states(4, 5), 2e-5)
