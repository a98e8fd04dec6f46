"""Launch-shape autotuner for the chemistry stage kernels.

The wavefront kernels take a launch shape (waves per workgroup, workgroups per CU).  The
best shape depends on the mechanism (LDS per workgroup, register budget), on the number of
points and on how uneven their stiffness is.  ``tune`` times the candidate shapes on the
caller's own inputs (device tensors, already resident) and applies the fastest; results are
cached per (device, mechanism, integrator, kinetics, points) in a JSON file so that later
runs skip the search.
"""
from __future__ import annotations

import json
import os
from typing import Iterable, Optional, Sequence, Tuple

from . import capi

CACHE = os.environ.get("EMBER_AMD_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "ember_amd_tune.json"))

QSS_SHAPES: Tuple[Tuple[int, int], ...] = ((2, 0), (4, 0), (4, 1), (8, 0), (12, 0))
BDF_SHAPES: Tuple[Tuple[int, int], ...] = ((1, 0), (2, 0), (4, 0))


def _key(ctx, n, device_name):
    o = ctx.options
    return f"{device_name}|{os.path.basename(ctx.table)}|int{o.chemistry_integrator}|kin{o.kinetics_model}|n{n}"


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        tmp = CACHE + ".tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, CACHE)
    except OSError:
        pass


def tune(ctx: "capi.Context", state_in, state_out, dt: float, stream=None,
         shapes: Optional[Iterable[Tuple[int, int]]] = None, reps: int = 2,
         use_cache: bool = True, **kw) -> dict:
    """Time each launch shape on (state_in -> state_out) and apply the fastest.

    state_in/state_out: contiguous CUDA float64 tensors (nPoints x nVars); kw are passed to
    Context.source_integrate_device.  Returns {"shape": (wpb, bpc), "ms": {...}, "cached": bool}.
    """
    import torch
    n = state_in.shape[0]
    dev = state_in.device
    name = torch.cuda.get_device_name(dev)
    key = _key(ctx, n, name)
    cache = _load() if use_cache else {}
    if use_cache and key in cache:
        wpb, bpc = cache[key]["shape"]
        ctx.set_launch(wpb, bpc)
        return {"shape": (wpb, bpc), "ms": cache[key]["ms"], "cached": True}
    if shapes is None:
        shapes = BDF_SHAPES if ctx.options.chemistry_integrator == capi.EMX_BDF else QSS_SHAPES
    stream = stream or torch.cuda.current_stream(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    times = {}
    for wpb, bpc in shapes:
        try:
            ctx.set_launch(wpb, bpc)
            ctx.source_integrate_device(state_in, state_out, dt, **kw)  # warm-up, cost order
            torch.cuda.synchronize()
        except capi.EmberAmdError:
            continue
        e0.record(stream)
        for _ in range(reps):
            ctx.source_integrate_device(state_in, state_out, dt, **kw)
        e1.record(stream)
        torch.cuda.synchronize()
        times[f"{wpb}x{bpc}"] = e0.elapsed_time(e1) / reps
    if not times:
        raise capi.EmberAmdError("no launch shape ran")
    best = min(times, key=times.get)
    wpb, bpc = (int(v) for v in best.split("x"))
    ctx.set_launch(wpb, bpc)
    if use_cache:
        cache[key] = {"shape": [wpb, bpc], "ms": times}
        _store(cache)
    return {"shape": (wpb, bpc), "ms": times, "cached": False}
