"""Chemistry stages at the benchmark configurations' full sizes against the oracle.

- QSS and BDF on all 4096 seeded tanh-progress GRI-Mech 3.0 states (seed 2, dt = 2e-5 s,
  P = 1 atm): T and Y_k within |d| <= 1e-6 |ref| + floor, statuses equal.  Floor 1e-12
  for QSS.  For the BDF (species absolute tolerance 1e-12) the floor is 5e-11: its
  accept/reject decisions make Y_k sensitive to rounding at that level; the oracle fed the
  same states perturbed by one ulp moves Y_k by up to 1.2e-11
  (tests/test_oracle_bdf.py::test_bdf_rounding_sensitivity).
- The uneven-stiffness workload (60 % burned, 30 % flame front, 10 % pre-ignition, seed 3)
  on its first 8192 points, eager and replayed from a captured HIP graph, against the
  oracle; the graph replays equal the eager stage bit for bit.
- The autotuner's chosen launch shape gives the same results bit for bit as the default
  shape.
- Non-finite input states fail their point (negative status, state kept) and leave the
  other points alone, in both integrators.
"""
import os

import numpy as np
import pytest

from ember_amd import capi, mechanism
from ember_amd import workloads as W

pytestmark = pytest.mark.gpu

DT = 2e-5
NTHREADS = min(16, len(os.sched_getaffinity(0)))


def _assert_close(got, ref, label, atol_y=1e-12, rtol=1e-6):
    dT = np.abs(got[:, 1] - ref[:, 1])
    assert (dT <= rtol * np.abs(ref[:, 1])).all(), (label, dT.max())
    dY = np.abs(got[:, 2:] - ref[:, 2:])
    bad = dY > rtol * np.abs(ref[:, 2:]) + atol_y
    assert not bad.any(), (label, dY[bad].max(), np.argwhere(bad)[:5])


@pytest.fixture(scope="module")
def gri_states():
    return W.tanh_progress_states(4096, seed=2, name="gri30")


@pytest.fixture(scope="module")
def gri_oracle(oracle_mod):
    return oracle_mod.Mech(mechanism.table_path("gri30"))


def test_qss_full_config(gpu, oracle_mod, gri_states, gri_oracle):
    print("[full] QSS, 4096 GRI states", flush=True)
    ctx = capi.Context("gri30")
    out, status, g, r = ctx.source_integrate(gri_states, DT)
    ref, rstat, rg, rr, _ = gri_oracle.integrate(gri_states, DT, oracle_mod.options(), nthreads=NTHREADS)
    assert (status == rstat).all() and (status == 0).all()
    _assert_close(out, ref, "qss4096")
    assert np.array_equal(g, rg)  # same RHS evaluation counts: the same step sequence


def test_bdf_full_config(gpu, oracle_mod, gri_states, gri_oracle):
    print("[full] BDF, 4096 GRI states", flush=True)
    ctx = capi.Context("gri30", chemistry_integrator=capi.EMX_BDF, integrator_species_abs_tol=1e-12)
    out, status, nst, nfe = ctx.source_integrate(gri_states, DT)
    ref, rstat, rnst, rnfe, _ = gri_oracle.integrate(
        gri_states, DT, oracle_mod.options(integrator=1, atol_species=1e-12), nthreads=NTHREADS)
    assert (status == 0).all() and (rstat == 0).all()
    _assert_close(out, ref, "bdf4096", atol_y=5e-11)
    assert np.mean(nst == rnst) > 0.8


def test_uneven_graph_vs_oracle(gpu, oracle_mod, gri_oracle):
    import torch
    n = 8192
    print(f"[full] uneven workload, {n} points, eager + graph replay", flush=True)
    st, kind = W.uneven_stiffness_states(65536, seed=3)
    st = np.ascontiguousarray(st[:n])
    ctx = capi.Context("gri30")
    s = torch.cuda.Stream(gpu)
    ctx.set_stream(s.cuda_stream)
    with torch.cuda.stream(s):
        x = torch.from_numpy(st).to(gpu)
        out = torch.empty_like(x)
        status = torch.zeros(n, dtype=torch.int32, device=gpu)
        g = torch.zeros_like(status)
        stage = lambda: ctx.source_integrate_device(x, out, DT, status=status, stat1=g)
        stage()
        torch.cuda.synchronize()
        eager, eager_status, eager_g = out.clone(), status.clone(), g.clone()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            stage()
        torch.cuda.synchronize()
        for i in range(2):
            out.fill_(-1.0)
            graph.replay()
            torch.cuda.synchronize()
            assert torch.equal(out, eager) and torch.equal(status, eager_status), i
    ref, rstat, rg, rr, _ = gri_oracle.integrate(st, DT, oracle_mod.options(), nthreads=NTHREADS)
    got = eager.cpu().numpy()
    assert (eager_status.cpu().numpy() == rstat).all()
    _assert_close(got, ref, "uneven")
    counts = eager_g.cpu().numpy()
    assert np.array_equal(counts, rg)
    print(f"[full] RHS evaluations per point {counts.min()} .. {counts.max()}", flush=True)


def test_autotuned_shape_bit_identical(gpu, tmp_path, monkeypatch):
    import torch
    from ember_amd import autotune
    print("[full] autotuner", flush=True)
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tune.json"))
    st, _ = W.uneven_stiffness_states(2048, seed=3)
    ctx = capi.Context("gri30")
    x = torch.from_numpy(st).to(gpu)
    ref = torch.empty_like(x)
    ctx.source_integrate_device(x, ref, DT)
    torch.cuda.synchronize()
    res = autotune.tune(ctx, x, torch.empty_like(x), DT, reps=1)
    assert res["shape"] in [tuple(s) for s in autotune.QSS_SHAPES] and not res["cached"]
    out = torch.empty_like(x)
    ctx.source_integrate_device(x, out, DT)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    # a second tune reads the cache and applies the same shape
    again = autotune.tune(ctx, x, torch.empty_like(x), DT)
    assert again["cached"] and tuple(again["shape"]) == tuple(res["shape"])


@pytest.mark.parametrize("integrator", [capi.EMX_QSS, capi.EMX_BDF])
def test_non_finite_state_fails_point(gpu, integrator):
    """A NaN temperature or mass fraction fails that point (negative status, input state
    kept); the finite points integrate as they do alone."""
    print(f"[full] non-finite input, integrator={integrator}", flush=True)
    st = W.tanh_progress_states(8, seed=2, name="h2o2")
    bad = st.copy()
    bad[2, 1] = np.nan
    bad[5, 6] = np.nan
    ctx = capi.Context("h2o2", chemistry_integrator=integrator)
    out, status, _, _ = ctx.source_integrate(bad, DT)
    ok = np.ones(8, bool)
    ok[[2, 5]] = False
    assert (status[~ok] < 0).all(), status
    np.testing.assert_array_equal(out[~ok], bad[~ok])
    ref, rstat, _, _ = ctx.source_integrate(st, DT)
    assert (status[ok] == 0).all()
    np.testing.assert_array_equal(out[ok], ref[ok])
