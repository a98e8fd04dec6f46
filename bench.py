#!/usr/bin/env python
"""Benchmark: batched chemistry production stage, GRI-Mech 3.0 (53 species, 325
reactions), N = 4096 points, dt = 2e-5 s, P = 1 atm, FP64, QSS with Ember's default
tolerances.  States: tanh progress-variable blend of CH4/air (phi = 1, 300 K) and its
adiabatic equilibrium, 1% log-normal noise on minor species (seed 2).

A step is one production stage over all points with the inputs resident in HBM.
Prints one JSON line (metric point-integrations/s).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--points", type=int, default=4096)
    p.add_argument("--integrator", choices=["qss", "bdf"], default="qss")
    p.add_argument("--kinetics", choices=["standard", "interp"], default="standard")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--no-plain", action="store_true", help="skip the plain thread-per-point kernel")
    p.add_argument("--wpb", type=int, default=0, help="waves per workgroup (0: library default)")
    p.add_argument("--bpc", type=int, default=0, help="workgroups per CU (0: as many as fit)")
    p.add_argument("--schedule", choices=["cost", "fifo"], default="cost")
    return p.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    from ember_amd import capi, flops, mechanism, workloads

    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    N = a.points
    integ = capi.EMX_QSS if a.integrator == "qss" else capi.EMX_BDF
    kin = capi.EMX_KIN_STANDARD if a.kinetics == "standard" else capi.EMX_KIN_INTERP
    opts = dict(chemistry_integrator=integ, kinetics_model=kin)
    if integ == capi.EMX_BDF:
        opts["integrator_species_abs_tol"] = 1e-12
    states = workloads.tanh_progress_states(N, seed=2, name="gri30")
    ctx = capi.Context("gri30", **opts)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    if a.wpb or a.bpc:
        ctx.set_launch(a.wpb, a.bpc)
    d_in = torch.from_numpy(states).to(dev)
    d_out = torch.empty_like(d_in)
    status = torch.zeros(N, dtype=torch.int32, device=dev)
    s1 = torch.zeros(N, dtype=torch.int32, device=dev)
    s2 = torch.zeros(N, dtype=torch.int32, device=dev)
    dt = 2e-5
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)

    def timed(steps):
        e0.record(stream)
        for _ in range(steps):
            ctx.source_integrate_device(d_in, d_out, dt, status=status, stat1=s1, stat2=s2)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    # index-order work queue (what a first stage over new points gets)
    ctx.set_schedule(capi.EMX_SCHED_FIFO)
    print(f"[bench] warmup {a.warmup} stages, N={N}, {a.integrator}/{a.kinetics}", file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        ctx.source_integrate_device(d_in, d_out, dt, status=status, stat1=s1, stat2=s2)
    torch.cuda.synchronize()
    ms_fifo = timed(a.steps)
    print(f"[bench] index-order queue: {ms_fifo:.3f} ms/stage", file=sys.stderr, flush=True)
    # cost-ordered queue: each stage takes the points in descending order of the previous
    # stage's right-hand-side evaluation counts
    ms = ms_fifo
    t0 = time.perf_counter()
    wall = 0.0
    if a.schedule == "cost":
        ctx.set_schedule(capi.EMX_SCHED_COST)
        ctx.source_integrate_device(d_in, d_out, dt, status=status, stat1=s1, stat2=s2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ms = timed(a.steps)
        print(f"[bench] cost-ordered queue: {ms:.3f} ms/stage", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    print(f"[bench] timed {a.steps} stages: {ms:.3f} ms/stage", file=sys.stderr, flush=True)
    value = N / (ms * 1e-3)
    out = d_out.cpu().numpy()
    st = status.cpu().numpy()
    g = s1.cpu().numpy().astype(np.float64)
    r = s2.cpu().numpy().astype(np.float64)

    if not a.no_plain and integ == capi.EMX_QSS:
        # plain GPU port (one thread per point) on the same states, same stream
        ctx.set_kernel_variant(capi.EMX_VARIANT_PLAIN)
        ctx.source_integrate_device(d_in, d_out, dt, status=status, stat1=s1, stat2=s2)
        torch.cuda.synchronize()
        np_steps = max(1, min(a.steps, 3))
        ms_plain = timed(np_steps)
        ctx.set_kernel_variant(capi.EMX_VARIANT_WAVE)
        print(f"[bench] plain kernel: {ms_plain:.3f} ms/stage", file=sys.stderr, flush=True)
    else:
        ms_plain = None

    mech = mechanism.load("gri30")
    fl = flops.rhs_flops(mech)
    res = {
        "metric": "GRI-scale point-integrations/s per split dt @rtol1e-6",
        "value": value, "unit": "point-integrations/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded tanh progress-variable states)",
        "config": {"workload": f"GRI-Mech 3.0 chemistry stage, {a.integrator.upper()}, {a.kinetics} kinetics",
                   "mechanism": "gri30 (53 species, 325 reactions)", "points": N, "dt_s": dt,
                   "pressure_Pa": mechanism.ONE_ATM, "seed": 2},
        "failed_points": int((st != 0).sum()),
        "schedule": a.schedule, "index_order_ms_per_step": ms_fifo,
    }
    if ms_plain:
        res["plain_gpu"] = {"value": N / (ms_plain * 1e-3), "ms_per_step": ms_plain,
                            "speedup": ms_plain / ms}

    cpu = None
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        om = O.Mech(mechanism.table_path("gri30"))
        oopts = O.options(integrator=integ, kinetics=kin,
                          **({"atol_species": 1e-12} if integ == capi.EMX_BDF else {}))
        print(f"[bench] CPU oracle on {N} points, {threads} threads", file=sys.stderr, flush=True)
        om.integrate(states[:64], dt, oopts, nthreads=threads)  # warm-up
        c0 = time.perf_counter()
        ref, rst, rg, rr, _ = om.integrate(states, dt, oopts, nthreads=threads)
        cpu_s = time.perf_counter() - c0
        cpu = {"value": N / cpu_s, "unit": "point-integrations/s", "cores": threads,
               "kind": "port", "sample": f"all {N} benchmark states, one stage, OpenMP dynamic,1 "
                                          f"(TBB grain-1 analogue), {cpu_s:.2f} s"}
        res["cpu_baseline"] = cpu
        # parity of this run
        dT = np.abs(out[:, 1] - ref[:, 1]) / np.abs(ref[:, 1])
        dY = np.abs(out[:, 2:] - ref[:, 2:]) / (np.abs(ref[:, 2:]) + 1e-6)
        res["parity"] = {"max_rel_T": float(dT.max()), "max_rel_Y_floor1e-6": float(dY.max()),
                         "status_equal": bool((st == rst).all())}
        g, r = rg.astype(np.float64), rr.astype(np.float64)

    if integ == capi.EMX_QSS:
        F = float(flops.qss_point_flops(fl, g, r).sum())
        res["roofline"] = {"bound": "mfma", "achieved": F / (ms * 1e-3) / 1e12, "peak": 78.6,
                           "unit": "TFLOP/s", "frac": F / (ms * 1e-3) / 1e12 / 78.6, "traffic": None,
                           "flops_per_launch": F, "note": "FP64 peak; exp/log count 1 op (lower bound)"}
        res["rhs_evals_per_point"] = {"mean": float(g.mean()), "max": float(g.max()), "min": float(g.min())}
    res["wall_s"] = wall
    print(json.dumps(res))


if __name__ == "__main__":
    main()
