"""A whole device-resident split step captured in a HIP graph (torch.cuda.CUDAGraph).

emx_split_step_device with caller-given diffusion step sizes issues only kernels on the
context stream (split constants, the four diffusion stages with their copies and
accumulations, the cost-ordered production stage, the time derivatives): no allocation,
no memset node and no host synchronisation once the context buffers exist.  One step is
captured after an eager warm-up step and replayed; each replay advances the resident
state by one split step, and the trajectory of replays equals the eager trajectory from
the same start bit for bit, with eager work between the replays.  A graph replays the
call it captured, absolute time included (the stage lengths are formed from t, so they
round differently at another t): the eager reference steps are all taken at t = 0.
"""
import numpy as np
import pytest

from ember_amd import capi
from ember_amd import workloads as W

pytestmark = pytest.mark.gpu

DT = 1e-5


@pytest.mark.parametrize("name,n", [("h2o2", 150), ("gri30", 512)])
def test_split_step_graph_replays(gpu, name, n):
    import torch
    print(f"[split-graph] {name} {n} points", flush=True)
    case = W.flame_profile(n, seed=1, name=name)
    nV = case.state.shape[1]
    ctx = capi.Context(name, global_timestep=DT)
    s = torch.cuda.Stream(gpu)
    ctx.set_stream(s.cuda_stream)
    t = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(gpu)
    h = np.full(nV, DT / 4)  # caller-given step sizes: the step runs without host syncs
    with torch.cuda.stream(s):
        B, D, x, Brow = t(case.B), t(case.D), t(case.x), t(case.B_row.astype(np.int32))
        status = torch.zeros(n, dtype=torch.int32, device=gpu)
        # eager trajectory
        ref = capi.SplitState(t(case.state))
        eager = []
        for k in range(4):
            ctx.split_step_device(ref, B, D, x, 0.0, DT, B_row=Brow, h_per_var=h, status=status)
            eager.append(ref.state.clone())
        torch.cuda.synchronize()
        # capture one step on a fresh split state (buffers already sized by the eager run)
        ss = capi.SplitState(t(case.state))
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            ctx.split_step_device(ss, B, D, x, 0.0, DT, B_row=Brow, h_per_var=h, status=status)
        torch.cuda.synchronize()
        # capture records the work without running it: start from the initial state
        for f in ss.fields():
            if f is not None:
                f.zero_()
        ss.state.copy_(t(case.state))
        for k in range(4):
            graph.replay()
            torch.cuda.synchronize()
            assert torch.equal(ss.state, eager[k]), k
            assert (status.cpu().numpy() == 0).all()
            # eager work on the same context between replays
            tmp = capi.SplitState(t(case.state))
            ctx.split_step_device(tmp, B, D, x, 0.0, DT, B_row=Brow, h_per_var=h, status=status)
            torch.cuda.synchronize()
            assert torch.equal(tmp.state, eager[0])
    assert (eager[-1] != t(case.state)).any()
