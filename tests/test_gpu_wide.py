"""The widest mechanism the lane-per-variable layout takes: 62 species, 64 state variables
(every lane of the wavefront a variable), on the GPU against the oracle.

The mechanism is GRI-Mech 3.0 with nine inert diluents appended (species D1..D9 carrying
N2's composition, thermo and transport data under new names; no reactions), and the
states are GRI flame-front states with the diluents at mass fraction 2e-3 each.  This is
where the BDF's LU in column pairs ends on a full pair (64 columns), the state row fills
the wavefront, and the transport pair table is largest.  Bars as for GRI: rates 1e-10
relative, stages within the stage contract (Y floor 5e-11 for the BDF), thermo and
transport 1e-12."""
import os

import numpy as np
import pytest

from ember_amd import capi, mechanism
from ember_amd import workloads as W

pytestmark = pytest.mark.gpu

N_EXTRA = 9


def wide_yaml_text():
    """GRI-Mech 3.0's YAML with the diluents appended (text edits: a YAML 1.1 round trip
    would read the species NO as a boolean)."""
    src = open(mechanism.builtin_yaml("gri30")).read()
    lines = src.splitlines(keepends=True)
    i0 = next(i for i, l in enumerate(lines) if l.startswith("- {name: N2,"))
    i1 = next(i for i in range(i0 + 1, len(lines)) if lines[i].startswith("- "))
    n2 = "".join(lines[i0:i1])
    names = [f"D{i}" for i in range(1, N_EXTRA + 1)]
    src = src.replace("CH2CHO, CH3CHO]", "CH2CHO, CH3CHO, " + ", ".join(names) + "]", 1)
    extra = "".join(n2.replace("{name: N2,", "{name: " + nm + ",", 1) for nm in names)
    k = src.index("\nreactions:")
    return src[:k].rstrip("\n") + "\n" + extra + src[k:]


@pytest.fixture(scope="module")
def wide(tmp_path_factory):
    p = tmp_path_factory.mktemp("wide") / "gri30_wide.yaml"
    p.write_text(wide_yaml_text())
    return str(p)


def _states(n, seed):
    st = W.flame_front_states(n, seed=seed, name="gri30")
    Y = np.concatenate([st[:, 2:], np.full((n, N_EXTRA), 2e-3)], axis=1)
    Y /= Y.sum(axis=1, keepdims=True)
    return np.concatenate([st[:, :2], Y], axis=1)


def _om(oracle_mod, path):
    return oracle_mod.Mech(mechanism.table_path(path, out_dir=os.path.dirname(path)))


def test_wide_mechanism_layout(wide):
    m = mechanism.load(wide)
    assert m.n_species == 62 and m.n_reactions == 325


def test_wide_rates(gpu, oracle_mod, wide):
    print("[wide] rates, 62 species", flush=True)
    st = _states(32, 1)
    ctx = capi.Context(wide)
    wQ, wD, wN, status = ctx.rates(st[:, 1], st[:, 2:])
    assert (status == 0).all()
    r = _om(oracle_mod, wide).rates(st[:, 1], st[:, 2:], oracle_mod.options())
    scale = np.abs(r["wdotQ"]).max(axis=1, keepdims=True)
    for got, ref in [(wQ, r["wdotQ"]), (wD, r["wdotD"]), (wN, r["wdot"])]:
        assert (np.abs(got - ref) <= 1e-10 * np.abs(ref) + 1e-13 * scale).all()


@pytest.mark.parametrize("integrator", [0, 1])
def test_wide_stage(gpu, oracle_mod, wide, integrator):
    print(f"[wide] stage, 62 species, integrator={integrator}", flush=True)
    st = _states(48, 5)
    ctx = capi.Context(wide, chemistry_integrator=integrator, integrator_species_abs_tol=1e-12)
    out, status, *_ = ctx.source_integrate(st, 2e-5)
    ref, rstat, *_ = _om(oracle_mod, wide).integrate(
        st, 2e-5, oracle_mod.options(integrator=integrator, atol_species=1e-12))
    assert (status == 0).all() and (rstat == 0).all()
    assert (np.abs(out[:, 1] - ref[:, 1]) <= 1e-6 * ref[:, 1]).all()
    floor = 5e-11 if integrator == 1 else 1e-12
    dY = np.abs(out[:, 2:] - ref[:, 2:])
    assert (dY <= 1e-6 * np.abs(ref[:, 2:]) + floor).all(), dY.max()
    # the inert diluents are not touched by the chemistry
    assert np.allclose(out[:, -N_EXTRA:], st[:, -N_EXTRA:], rtol=1e-9)


def test_wide_properties(gpu, oracle_mod, wide):
    print("[wide] thermo and transport properties, 62 species", flush=True)
    st = _states(64, 3)
    ctx = capi.Context(wide)
    om = _om(oracle_mod, wide)
    th = ctx.thermo_properties(st)
    rt = om.thermo_properties(st)
    for k in ("rho", "Wmx", "cp"):
        np.testing.assert_allclose(th[k], rt[k], rtol=1e-12)
    mu, lam, rhoD = ctx.transport_properties(st)
    ref = om.transport_properties(st)
    np.testing.assert_allclose(mu, ref[0], rtol=1e-12)
    np.testing.assert_allclose(lam, ref[1], rtol=1e-12)
    np.testing.assert_allclose(rhoD, ref[2], rtol=1e-12)


def test_wide_bdf_launch_shape(gpu, oracle_mod, wide):
    # four BDF waves of 62 species need more than a CU's 160 KB of LDS: the default shape
    # runs three per workgroup (checked above), a shape asked for explicitly is refused
    print("[wide] BDF launch shape", flush=True)
    ctx = capi.Context(wide, chemistry_integrator=1)
    ctx.set_launch(4, 0)
    with pytest.raises(capi.EmberAmdError, match="LDS"):
        ctx.source_integrate(_states(4, 5), 2e-5)
    ctx.set_launch(3, 0)
    out, status, *_ = ctx.source_integrate(_states(4, 5), 2e-5)
    assert (status == 0).all()


def test_wide_adjoint(gpu, oracle_mod, wide):
    import torch
    from ember_amd import adjoint
    print("[wide] adjoint, 62 species", flush=True)
    st = _states(16, 5)
    ctx = capi.Context(wide)
    d = torch.from_numpy(st).to("cuda:0")
    out, Q, gs, gk, status = (a.cpu().numpy() for a in adjoint.per_point_gradients(ctx, d, 2e-5))
    rout, rQ, rgs, rgk, rst = _om(oracle_mod, wide).adjoint(st, 2e-5, oracle_mod.options())
    assert (status == 0).all() and (rst == 0).all()
    np.testing.assert_allclose(out[:, 1], rout[:, 1], rtol=1e-6)
    np.testing.assert_allclose(Q, rQ, rtol=1e-6)
    for a, b in [(gs[:, 1:] * st[:, 1:], rgs[:, 1:] * st[:, 1:]), (gk, rgk)]:
        scale = np.abs(b).max(axis=1, keepdims=True)
        assert (np.abs(a - b) / scale).max() <= 1e-8
