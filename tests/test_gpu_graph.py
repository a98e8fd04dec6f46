"""The chemistry stage captured in a HIP graph (through torch.cuda.CUDAGraph) replays to the
eager result: cost-order kernel, queue reset and stage kernel, no memset nodes."""
import numpy as np
import pytest

from ember_amd import capi
from ember_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kinetics", [0, 1])
def test_graph_replays_match_eager(gpu, kinetics):
    import torch
    print(f"[graph] kinetics={kinetics}", flush=True)
    st, _ = W.uneven_stiffness_states(512, seed=3)
    ctx = capi.Context("gri30", kinetics_model=kinetics)
    s = torch.cuda.Stream(gpu)
    ctx.set_stream(s.cuda_stream)
    with torch.cuda.stream(s):
        x = torch.from_numpy(st).to(gpu)
        out = torch.empty_like(x)
        status = torch.zeros(len(st), dtype=torch.int32, device=gpu)
        g = torch.zeros_like(status)
        stage = lambda: ctx.source_integrate_device(x, out, 2e-5, status=status, stat1=g)
        stage()                      # records costs: the captured stage uses the cost order
        torch.cuda.synchronize()
        ref = out.clone()
        ref_status = status.clone()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            stage()
        torch.cuda.synchronize()
        for i in range(4):
            out.fill_(-1.0)
            status.fill_(7)
            graph.replay()
            torch.cuda.synchronize()
            assert torch.equal(out, ref), i
            assert torch.equal(status, ref_status), i
            stage()                  # eager work between replays
            torch.cuda.synchronize()
            assert torch.equal(out, ref)
