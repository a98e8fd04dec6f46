"""The device-resident split step with its convection stages (emx_split_step_conv_device)
and the convection stage's density derivative (emx_density_derivative_device) on the GPU.

- the density derivative against the oracle's restatement (tests/test_oracle_density_
  derivative.py pins that one by the ideal-gas density): the same operations in the same
  order, bit for bit;
- a whole split step with convection (D1/4 C1/2 D1/4 P D1/4 C1/2 D1/4, the density
  derivative from the previous step's split rates) captured in a HIP graph after an eager
  warm-up step and replayed: every replay equals the eager step from the same start bit for
  bit, with eager work between the replays, and the convection stages moved the state
  (deltaConv nonzero, ddtConv = deltaConv / dt - splitConstConv).
The host-driven sequence against the device-resident one, bit for bit, is in
tests/test_gpu_split.py (convection case)."""
import numpy as np
import pytest

from ember_amd import capi, mechanism
from ember_amd import workloads as W

pytestmark = pytest.mark.gpu

DT = 1e-5


def test_density_derivative_device_matches_oracle(gpu, oracle_mod):
    import torch
    print("[split-conv] density derivative vs oracle", flush=True)
    st = W.flame_front_states(200, seed=4, name="gri30")
    rng = np.random.default_rng(7)
    parts = [rng.normal(0.0, 1.0, st.shape) * np.r_[10.0, 1e5, 50.0 * np.ones(st.shape[1] - 2)]
             for _ in range(4)]
    om = oracle_mod.Mech(mechanism.table_path("gri30"))
    th = om.thermo_properties(st)
    ctx = capi.Context("gri30")
    t = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(gpu)
    for balanced in (True, False):
        out = torch.zeros(st.shape[0], dtype=torch.float64, device=gpu)
        ctx.density_derivative_device(t(st), t(parts[0]), t(parts[1]), t(parts[2]), t(th["rho"]), t(th["Wmx"]),
                                      out, ddt_cross=t(parts[3]), balanced=balanced)
        torch.cuda.synchronize()
        ref = om.density_derivative(st, parts[0], parts[1], parts[2], th["rho"], th["Wmx"],
                                    ddt_cross=parts[3], balanced=balanced)
        np.testing.assert_array_equal(out.cpu().numpy(), ref)


def test_split_conv_step_graph_replays(gpu):
    import torch
    print("[split-conv] graph of a split step with convection, h2o2 150 points", flush=True)
    case = W.flame_profile(150, seed=1, name="h2o2")
    n, nV = case.state.shape
    m = mechanism.load("h2o2")
    Y0 = case.state[0, 2:]
    rho0 = 101325.0 / (mechanism.GAS_CONSTANT * case.state[0, 1] * (Y0 / m.W).sum())
    ctx = capi.Context("h2o2", global_timestep=DT)
    conv = ctx.convection_params(Yleft=Y0, rVzero=0.2 * rho0, Tleft=float(case.state[0, 1]), rhou=rho0,
                                 strain_initial=100.0, strain_final=100.0)
    s = torch.cuda.Stream(gpu)
    ctx.set_stream(s.cuda_stream)
    t = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(gpu)
    h = np.full(nV, DT / 4)  # caller-given step sizes: the step runs without host syncs
    with torch.cuda.stream(s):
        B, D, x, Brow = t(case.B), t(case.D), t(case.x), t(case.B_row.astype(np.int32))
        status = torch.zeros(n, dtype=torch.int32, device=gpu)
        rho = torch.zeros(n, dtype=torch.float64, device=gpu)
        Wmx = torch.zeros_like(rho)
        cp = torch.zeros_like(rho)
        ctx.thermo_properties_device(t(case.state), rho, cp, Wmx=Wmx)
        kw = dict(rho=rho, Wmx=Wmx, B_row=Brow, h_per_var=h, status=status)
        # eager trajectory; the second step's density derivative uses the first step's rates
        ref = capi.SplitState(t(case.state), convection=True)
        eager = []
        for k in range(3):
            ctx.split_step_conv_device(ref, conv, B, D, x, 0.0, DT, **kw)
            eager.append(ref.state.clone())
            if k == 0:
                torch.cuda.synchronize()
                assert (status.cpu().numpy() == 0).all()
                dconv = ref.delta_conv.cpu().numpy()
                assert np.abs(dconv[:, 0]).max() > 0.0  # the strain term and the inflow acted
                sc_conv = ref.work[3 * n * nV:4 * n * nV].reshape(n, nV).cpu().numpy()
                np.testing.assert_allclose(ref.ddt_conv.cpu().numpy(), dconv / DT - sc_conv, rtol=1e-15,
                                           atol=1e-300)
        torch.cuda.synchronize()
        # capture one step on a fresh split state (buffers sized by the eager run)
        ss = capi.SplitState(t(case.state), convection=True)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            ctx.split_step_conv_device(ss, conv, B, D, x, 0.0, DT, **kw)
        torch.cuda.synchronize()
        # capture records the work without running it: start from the initial state
        for f in ss.fields() + (ss.delta_conv,):
            if f is not None:
                f.zero_()
        ss.state.copy_(t(case.state))
        for k in range(3):
            graph.replay()
            torch.cuda.synchronize()
            assert torch.equal(ss.state, eager[k]), k
            assert (status.cpu().numpy() == 0).all()
            tmp = capi.SplitState(t(case.state), convection=True)
            ctx.split_step_conv_device(tmp, conv, B, D, x, 0.0, DT, **kw)
            torch.cuda.synchronize()
            assert torch.equal(tmp.state, eager[0])
    assert (eager[-1] != t(case.state)).any()


def test_split_conv_step_refusals_leave_the_state(gpu):
    import torch
    print("[split-conv] refusals", flush=True)
    case = W.flame_profile(40, seed=1, name="h2o2")
    n, nV = case.state.shape
    ctx = capi.Context("h2o2", global_timestep=DT)
    t = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(gpu)
    B, D, x, Brow = t(case.B), t(case.D), t(case.x), t(case.B_row.astype(np.int32))
    ss = capi.SplitState(t(case.state), convection=True)
    h = np.full(nV, DT / 4)
    bad = ctx.convection_params(continuity_bc=1)  # only fixedLeft is supported
    with pytest.raises(capi.EmberAmdError, match="continuity"):
        ctx.split_step_conv_device(ss, bad, B, D, x, 0.0, DT, B_row=Brow, h_per_var=h)
    rho = torch.ones(n, dtype=torch.float64, device=gpu)
    with pytest.raises(capi.EmberAmdError, match="rho and Wmx"):
        ctx.split_step_conv_device(ss, ctx.convection_params(), B, D, x, 0.0, DT, rho=rho, B_row=Brow,
                                   h_per_var=h)
    torch.cuda.synchronize()
    # refused before any stage ran: the state and the bookkeeping are as they were
    assert torch.equal(ss.state, t(case.state))
    assert not ss.delta_conv.any() and not ss.ddt_diff.any() and not ss.delta_diff.any()
