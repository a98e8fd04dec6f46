"""emx_source_integrate_device_noalloc against emx_source_integrate_device.

Both run the same QSS stage kernel; the allocation-free entry point takes its work-queue
counter, costs and order from the library's fixed per-GPU scratch instead of context
buffers.  Called through the C ABI on two stages in a row, the second over fewer points
(so the scratch holds the first call's contents), and a third repeating the second (its
queue order then comes from the second's costs): states, statuses, evaluation counts and
integrated heat release equal the allocating entry point's bit for bit.  Sizes beyond the
scratch, the BDF integrator and the plain variant are refused before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from ember_amd import capi
from ember_amd import workloads as W

pytestmark = pytest.mark.gpu

DT = 2e-5
MAX_POINTS = 262144


def _stage(fn, ctx, x, gpu):
    import torch
    n = x.shape[0]
    out = torch.empty_like(x)
    status = torch.full((n,), 7, dtype=torch.int32, device=gpu)
    g = torch.zeros(n, dtype=torch.int32, device=gpu)
    r = torch.zeros(n, dtype=torch.int32, device=gpu)
    q = torch.zeros(n, dtype=torch.float64, device=gpu)
    p = capi._tptr
    rc = fn(ctx.h, n, ctx.n_vars, p(x), None, None, 0.0, DT, p(out), p(status), p(g), p(r), p(q))
    assert rc == 0, capi.lib().emx_last_error()
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (out, status, g, r, q)]


def test_noalloc_matches_allocating_entry(gpu):
    import torch
    L = capi.lib()
    ctx = capi.Context("gri30")
    first = W.tanh_progress_states(4096, seed=2, name="gri30")
    second, _ = W.uneven_stiffness_states(1000, seed=3)
    for label, st in (("4096 tanh states", first), ("1000 uneven states", second),
                      ("1000 uneven states again", second)):
        print(f"[noalloc] {label}", flush=True)
        x = torch.from_numpy(np.ascontiguousarray(st)).to(gpu)
        ref = _stage(L.emx_source_integrate_device, ctx, x, gpu)
        got = _stage(L.emx_source_integrate_device_noalloc, ctx, x, gpu)
        assert (ref[1] == 0).all()
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)


def test_noalloc_refusals(gpu):
    L = capi.lib()
    ctx = capi.Context("gri30")
    nv = ctx.n_vars
    print("[noalloc] refusals", flush=True)
    rc = L.emx_source_integrate_device_noalloc(ctx.h, MAX_POINTS + 1, nv, None, None, None, 0.0, DT,
                                               None, None, None, None, None)
    assert rc == -1 and b"262144" in L.emx_last_error()
    bdf = capi.Context("gri30", chemistry_integrator=capi.EMX_BDF)
    rc = L.emx_source_integrate_device_noalloc(bdf.h, 4, nv, C.c_void_p(1), None, None, 0.0, DT,
                                               C.c_void_p(1), None, None, None, None)
    assert rc == -1 and b"QSS" in L.emx_last_error()
    ctx.set_kernel_variant(capi.EMX_VARIANT_PLAIN)
    rc = L.emx_source_integrate_device_noalloc(ctx.h, 4, nv, C.c_void_p(1), None, None, 0.0, DT,
                                               C.c_void_p(1), None, None, None, None)
    assert rc == -1 and b"plain" in L.emx_last_error()
