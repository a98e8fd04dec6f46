#!/usr/bin/env python
"""Benchmark: batched chemistry production stage, GRI-Mech 3.0 (53 species, 325
reactions), N = 4096 points, dt = 2e-5 s, P = 1 atm, FP64, QSS with Ember's default
tolerances.  States: tanh progress-variable blend of CH4/air (phi = 1, 300 K) and its
adiabatic equilibrium, 1% log-normal noise on minor species (seed 2).

A step is one production stage over all points with the inputs resident in HBM.
Prints one JSON line (metric point-integrations/s).  Beside the headline the same line
carries one sub-object per other configuration, each with its own roofline, parity and
CPU rate:
  bdf        the same 4096 states with the BDF integrator (rtol 1e-6, atol 1e-12)
  uneven     N = 65536 states of very uneven stiffness (60% near-equilibrium burned, 30%
             flame front, 10% pre-ignition at 900-1200 K, seed 3): autotuned launch shape,
             hipGraph-captured stage, against the same kernel launched one wave per point
             and the thread-per-point kernel
  diffusion  53 species x 65536 points diffusion stage (seeded smooth profiles, seed 4,
             FixedValue ends, dt = 2e-5 s, h = dt: 8 BDF1 solves), HBM GB/s
  thermo     thermodynamic properties of 65536 GRI points, HBM GB/s
  transport  mixture-averaged transport properties of 65536 GRI points
  convection the convection stage (ConvectionSystemSplit) on a 400-point GRI flame profile,
             half of a 2e-5 s split step
  adjoint    gradients of the integrated heat release on 4096 flame-front states
--headline-only skips them; --workload NAME runs one of them alone.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE, each scaled by the
# counter's reading for a plain copy of known size (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0):
# profiles/qss_gri30_n4096_stage.md (stage kernel, 4096 GRI points) and
# profiles/diffusion_k53_n65536.md (one diffusion stage of 53 x 65536: the setup launch,
# 9 diff_step and 8 pcr launches, summed and divided by the stage's 8 solves).
STAGE_TRAFFIC = 1555 * 2 * 1024 + 2355 * 1024
DIFFUSION_TRAFFIC_PER_SOLVE = (416743 * 2 * 1024 + 256761 * 1024) // 8


METRIC = ("GRI-scale point-integrations/s per split dt @rtol1e-6, batched vs plain GPU; "
          "adjoint/fwd")
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector peak (AMD-published)
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E


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
    p.add_argument("--cpu-repeats", type=int, default=5, help="timed CPU runs (median reported)")
    p.add_argument("--no-plain", action="store_true", help="skip the plain thread-per-point kernel")
    p.add_argument("--wpb", type=int, default=0, help="waves per workgroup (0: library default)")
    p.add_argument("--bpc", type=int, default=0, help="workgroups per CU (0: as many as fit)")
    p.add_argument("--schedule", choices=["cost", "fifo"], default="cost")
    p.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost measurement")
    p.add_argument("--headline-only", action="store_true",
                   help="skip the bdf / uneven / diffusion / thermo sub-objects")
    p.add_argument("--workload", choices=["stage", "uneven", "diffusion", "split", "copy", "thermo",
                                          "transport", "convection"],
                   default="stage")
    p.add_argument("--mech", default="gri30", help="mechanism of the split workload")
    return p.parse_args()


def cpu_info():
    """Host CPU of this run: threads used by the CPU legs and how they were chosen.  The
    thread count is OMP_NUM_THREADS when set (the CPU share the machine gives one GPU),
    otherwise every core in this process's affinity mask."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = int(omp) if omp and omp.isdigit() and int(omp) > 0 else aff
    return {"model": model, "nproc": os.cpu_count(), "affinity_cores": aff,
            "omp_num_threads": omp, "threads_used": threads}


def timed_cpu(fn, repeats):
    """1 warm-up run, then the median of `repeats` timed runs (seconds) and the last result."""
    fn()
    ts, out = [], None
    for _ in range(max(1, repeats)):
        c0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - c0)
    ts.sort()
    return ts[len(ts) // 2], ts, out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch

    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    if a.workload != "stage":
        fn = {"uneven": run_uneven, "diffusion": run_diffusion, "copy": run_copy,
              "split": run_split, "thermo": run_thermo, "transport": run_transport,
              "convection": run_convection}[a.workload]
        print(json.dumps(fn(a, dev)))
        return
    res = run_stage(a, dev)
    if not a.headline_only and a.integrator == "qss" and a.kinetics == "standard" and a.points == 4096:
        sub = argparse.Namespace(**vars(a))
        sub.points = 4096
        sub.steps = max(1, min(a.steps, 10))
        sub.warmup = 1
        res["bdf"] = run_bdf(sub, dev)
        sub.points = 65536
        res["uneven"] = _subobject(run_uneven(sub, dev))
        # a diffusion stage takes ~0.3 ms: time 50 of them (after 3 warm-up stages)
        dsub = argparse.Namespace(**vars(sub))
        dsub.steps, dsub.warmup = 50, 3
        res["diffusion"] = _subobject(run_diffusion(dsub, dev))
        res["thermo"] = _subobject(run_thermo(sub, dev))
        res["transport"] = _subobject(run_transport(sub, dev))
        sub.points = 400
        res["convection"] = _subobject(run_convection(sub, dev))
    print(json.dumps(res))


def _subobject(r):
    """A standalone workload's result as a sub-object of the headline line."""
    drop = ("n_gpus", "higher_is_better", "scaling", "vs_baseline", "dtype")
    return {k: v for k, v in r.items() if k not in drop}


def run_stage(a, dev):
    import numpy as np
    import torch

    from ember_amd import capi, flops, mechanism, workloads
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
        # HIP events on the stream the library launches on (the context stream)
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
    # cost-ordered queue: each stage takes the points in descending order of their measured
    # integration times in the previous stage
    ms = ms_fifo
    if a.schedule == "cost":
        ctx.set_schedule(capi.EMX_SCHED_COST)
        ctx.source_integrate_device(d_in, d_out, dt, status=status, stat1=s1, stat2=s2)
        torch.cuda.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms = timed(a.steps)
    wall = time.perf_counter() - t0
    print(f"[bench] timed {a.steps} stages ({a.schedule} order): {ms:.3f} ms/stage", file=sys.stderr, flush=True)
    value = N / (ms * 1e-3)
    out = d_out.cpu().numpy()
    # the host-buffer entry point on the same states: copy in, stage, copy out (PCIe included)
    ctx.source_integrate(states, dt)
    c0 = time.perf_counter()
    n_e2e = 5
    for _ in range(n_e2e):
        ctx.source_integrate(states, dt)
    ms_e2e = (time.perf_counter() - c0) / n_e2e * 1e3
    print(f"[bench] host-buffer stage (H2D + stage + D2H): {ms_e2e:.3f} ms", file=sys.stderr, flush=True)
    st = status.cpu().numpy()
    g = s1.cpu().numpy().astype(np.float64)
    r = s2.cpu().numpy().astype(np.float64)

    if not a.no_plain and integ == capi.EMX_QSS:
        # plain GPU port (one thread per point) on the same states, same stream
        ctx.set_kernel_variant(capi.EMX_VARIANT_PLAIN)
        ctx.source_integrate_device(d_in, d_out, dt, status=status, stat1=s1, stat2=s2)
        torch.cuda.synchronize()
        ms_plain = timed(max(1, min(a.steps, 3)))
        ctx.set_kernel_variant(capi.EMX_VARIANT_WAVE_PER_POINT)
        ctx.source_integrate_device(d_in, d_out, dt, status=status, stat1=s1, stat2=s2)
        torch.cuda.synchronize()
        ms_pp = timed(max(1, min(a.steps, 5)))
        ctx.set_kernel_variant(capi.EMX_VARIANT_WAVE)
        print(f"[bench] plain kernel: {ms_plain:.3f} ms/stage, wave-per-point launch {ms_pp:.3f} ms",
              file=sys.stderr, flush=True)
    else:
        ms_plain = ms_pp = None

    adj = None
    if not a.no_adjoint and integ == capi.EMX_QSS and kin == capi.EMX_KIN_STANDARD:
        adj = adjoint_cost(ctx, N, dt, dev, stream, a.steps)

    mech = mechanism.load("gri30")
    fl = flops.rhs_flops(mech)
    res = {
        "metric": METRIC,
        "value": value, "unit": "point-integrations/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded tanh progress-variable states)",
        "config": {"workload": f"GRI-Mech 3.0 chemistry stage, {a.integrator.upper()}, {a.kinetics} kinetics",
                   "mechanism": "gri30 (53 species, 325 reactions)", "points": N, "dt_s": dt,
                   "pressure_Pa": mechanism.ONE_ATM, "seed": 2},
        "failed_points": int((st != 0).sum()),
        "schedule": a.schedule,
        "first_stage_index_order": {"ms_per_step": ms_fifo, "value": N / (ms_fifo * 1e-3),
                                    "note": "work queue in index order, as a stage over new "
                                            "points runs; the headline is the cost-ordered queue "
                                            "(the points' measured times in the previous stage)"},
        "host_buffer_ms_per_step": ms_e2e,
    }
    if ms_plain:
        res["plain_gpu"] = {"value": N / (ms_plain * 1e-3), "ms_per_step": ms_plain,
                            "speedup": ms_plain / ms,
                            "note": "one thread per point, 64 wavefronts at N = 4096"}
        res["wave_per_point_launch"] = {"ms_per_step": ms_pp, "speedup": ms_pp / ms,
                                        "note": "the same wavefront kernel, one wave per point, "
                                                "no work queue"}

    if not a.no_cpu_baseline:
        from oracle import oracle as O
        info = cpu_info()
        threads = a.cpu_threads or info["threads_used"]
        om = O.Mech(mechanism.table_path("gri30"))
        oopts = O.options(integrator=integ, kinetics=kin,
                          **({"atol_species": 1e-12} if integ == capi.EMX_BDF else {}))
        print(f"[bench] CPU oracle on {N} points, {threads} threads, {info['model']}", file=sys.stderr, flush=True)
        cpu_s, runs, (ref, rst, rg, rr, _) = timed_cpu(
            lambda: om.integrate(states, dt, oopts, nthreads=threads), a.cpu_repeats)
        cpu = {"value": N / cpu_s, "unit": "point-integrations/s", "cores": threads,
               "kind": "port",
               "sample": f"all {N} benchmark states, one stage, OpenMP dynamic,1 (TBB grain-1 "
                         f"analogue); median of {len(runs)} runs after 1 warm-up: {cpu_s:.3f} s",
               "runs_s": runs, "host": info}
        n1 = 128
        cpu1_s, runs1, _ = timed_cpu(lambda: om.integrate(states[:n1], dt, oopts, nthreads=1),
                                     a.cpu_repeats)
        cpu["one_core"] = {"value": n1 / cpu1_s, "cores": 1,
                           "sample": f"first {n1} benchmark states; median of {len(runs1)} runs: "
                                     f"{cpu1_s:.3f} s"}
        res["cpu_baseline"] = cpu
        # parity of this run
        dT = np.abs(out[:, 1] - ref[:, 1]) / np.abs(ref[:, 1])
        dY = np.abs(out[:, 2:] - ref[:, 2:]) / (np.abs(ref[:, 2:]) + 1e-6)
        # the tolerance of the stage contract: |dT| <= 1e-6 |T|, |dY| <= 1e-6 |Y| + floor
        floor = 5e-11 if integ == capi.EMX_BDF else 1e-12
        ok_T = bool((np.abs(out[:, 1] - ref[:, 1]) <= 1e-6 * np.abs(ref[:, 1])).all())
        ok_Y = bool((np.abs(out[:, 2:] - ref[:, 2:]) <= 1e-6 * np.abs(ref[:, 2:]) + floor).all())
        res["parity"] = {"max_rel_T": float(dT.max()), "max_rel_Y_floor1e-6": float(dY.max()),
                         "status_equal": bool((st == rst).all()), "y_floor": floor,
                         "within_rtol1e-6": ok_T and ok_Y}
        g, r = rg.astype(np.float64), rr.astype(np.float64)

    if integ == capi.EMX_QSS:
        F = float(flops.qss_point_flops(fl, g, r).sum())
        res["roofline"] = {"bound": "fp64-vector", "achieved": F / (ms * 1e-3) / 1e12,
                           "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": F / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                           "traffic": STAGE_TRAFFIC if (kin == capi.EMX_KIN_STANDARD and N == 4096) else None,
                           "flops_per_launch": F,
                           "note": "FP64 vector (VALU) peak; no MFMA is used; exp/log count 1 op "
                                   "(lower bound); FLOPs from the oracle's evaluation counts"}
        res["rhs_evals_per_point"] = {"mean": float(g.mean()), "max": float(g.max()), "min": float(g.min())}
    if adj:
        res["adjoint"] = adj
    res["wall_s"] = wall
    return res


def bdf_flops(fl, work, n_vars):
    """Algorithmic FP64 operations of the BDF stage from the oracle's per-point work: every
    right-hand side (Newton iterations and finite-difference Jacobian columns) at the full
    count, (2/3) n^3 per LU factorisation, 2 n^2 per pair of triangular solves."""
    import numpy as np
    w = work.astype(np.float64)
    return float((w[:, 0] * fl.full + w[:, 2] * (2.0 / 3.0) * n_vars ** 3
                  + w[:, 3] * 2.0 * n_vars ** 2).sum())


def run_bdf(a, dev):
    """The 4096 benchmark states with the BDF integrator (SourceSystemCVODE semantics,
    rtol 1e-6, species atol 1e-12)."""
    import numpy as np
    import torch

    from ember_amd import capi, flops, mechanism, workloads
    N, dt = a.points, 2e-5
    states = workloads.tanh_progress_states(N, seed=2, name="gri30")
    ctx = capi.Context("gri30", chemistry_integrator=capi.EMX_BDF, integrator_species_abs_tol=1e-12)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    d_in = torch.from_numpy(states).to(dev)
    d_out = torch.empty_like(d_in)
    status = torch.zeros(N, dtype=torch.int32, device=dev)
    call = lambda: ctx.source_integrate_device(d_in, d_out, dt, status=status)
    ctx.set_schedule(capi.EMX_SCHED_FIFO)
    call()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    call()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_first = e0.elapsed_time(e1)
    ctx.set_schedule(capi.EMX_SCHED_COST)
    call()
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(a.steps):
        call()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    print(f"[bench] BDF stage: {ms:.3f} ms (index order {ms_first:.3f} ms)", file=sys.stderr, flush=True)
    out, st = d_out.cpu().numpy(), status.cpu().numpy()
    res = {"workload": "GRI-Mech 3.0 chemistry stage, BDF (rtol 1e-6, atol 1e-12), the "
                       "4096 benchmark states",
           "value": N / (ms * 1e-3), "unit": "point-integrations/s", "ms_per_step": ms,
           "steps": a.steps, "first_stage_index_order_ms": ms_first,
           "failed_points": int((st != 0).sum())}
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        info = cpu_info()
        threads = a.cpu_threads or info["threads_used"]
        om = O.Mech(mechanism.table_path("gri30"))
        oo = O.options(integrator=1, atol_species=1e-12)
        cpu_s, runs, (ref, rst, *_) = timed_cpu(lambda: om.integrate(states, dt, oo, nthreads=threads),
                                                min(a.cpu_repeats, 3))
        work = om.bdf_work(states, dt, oo, nthreads=threads)
        F = bdf_flops(flops.rhs_flops(mechanism.load("gri30")), work, states.shape[1])
        ok_T = bool((np.abs(out[:, 1] - ref[:, 1]) <= 1e-6 * np.abs(ref[:, 1])).all())
        ok_Y = bool((np.abs(out[:, 2:] - ref[:, 2:]) <= 1e-6 * np.abs(ref[:, 2:]) + 5e-11).all())
        res["parity"] = {"max_rel_T": float((np.abs(out[:, 1] - ref[:, 1]) / ref[:, 1]).max()),
                         "max_abs_Y_excess": float((np.abs(out[:, 2:] - ref[:, 2:])
                                                    - 1e-6 * np.abs(ref[:, 2:])).max()),
                         "status_equal": bool((st == rst).all()), "y_floor": 5e-11,
                         "within_rtol1e-6": ok_T and ok_Y}
        res["cpu_baseline"] = {"value": N / cpu_s, "unit": "point-integrations/s", "cores": threads,
                               "kind": "port", "sample": f"all {N} states, median of {len(runs)} "
                                                         f"runs after 1 warm-up: {cpu_s:.3f} s"}
        res["work_per_point"] = {"rhs": float(work[:, 0].mean()), "jacobians": float(work[:, 1].mean()),
                                 "lu": float(work[:, 2].mean()), "solves": float(work[:, 3].mean())}
        res["roofline"] = {"bound": "fp64-vector", "achieved": F / (ms * 1e-3) / 1e12,
                           "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": F / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, "traffic": None,
                           "flops_per_launch": F,
                           "note": "RHS at the full count, LU 2/3 n^3, solves 2 n^2; work counts "
                                   "from the oracle on the same states"}
    return res


def _pct(x):
    import numpy as np
    q = np.percentile(x, [0, 10, 50, 90, 99, 100])
    return {"min": float(q[0]), "p10": float(q[1]), "p50": float(q[2]), "p90": float(q[3]),
            "p99": float(q[4]), "max": float(q[5]), "mean": float(np.mean(x))}


def run_uneven(a, dev):
    """Uneven-stiffness batch in one persistent launch, hipGraph-captured."""
    import numpy as np
    import torch

    from ember_amd import autotune, capi, flops, mechanism, workloads
    N = a.points if a.points != 4096 else 65536
    dt = 2e-5
    st, kind = workloads.uneven_stiffness_states(N, seed=3)
    ctx = capi.Context("gri30")
    s = torch.cuda.Stream(dev)
    ctx.set_stream(s.cuda_stream)
    with torch.cuda.stream(s):
        d_in = torch.from_numpy(st).to(dev)
        d_out = torch.empty_like(d_in)
        status = torch.zeros(N, dtype=torch.int32, device=dev)
        g = torch.zeros(N, dtype=torch.int32, device=dev)
        r = torch.zeros(N, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)

    def timed(fn, reps):
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    stage = lambda: ctx.source_integrate_device(d_in, d_out, dt, status=status, stat1=g, stat2=r)
    print(f"[bench] uneven: N={N}, autotuning launch shape", file=sys.stderr, flush=True)
    with torch.cuda.stream(s):
        # first stage over new points: index-order queue, library default shape
        ctx.set_schedule(capi.EMX_SCHED_FIFO)
        stage()
        torch.cuda.synchronize()
        ms_first = timed(stage, 1)
        ctx.set_schedule(capi.EMX_SCHED_COST)
        tune = autotune.tune(ctx, d_in, d_out, dt, stream=s, reps=2, use_cache=False)
        print(f"[bench] autotune: {tune}", file=sys.stderr, flush=True)
        stage()
        torch.cuda.synchronize()
        ms_eager = timed(stage, a.steps)
        ref_out = d_out.clone()
        ref_status = status.clone()
        # hipGraph of one stage (cost-order kernel, queue reset, stage kernel)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            stage()
        torch.cuda.synchronize()
        replays_ok = True
        for _ in range(3):  # replays against the eager result, eager work in between
            d_out.zero_()
            graph.replay()
            torch.cuda.synchronize()
            replays_ok &= bool(torch.equal(d_out, ref_out)) and bool(torch.equal(status, ref_status))
        ms_graph = timed(graph.replay, a.steps)
        print(f"[bench] first {ms_first:.2f} ms, eager {ms_eager:.2f} ms, graph {ms_graph:.2f} ms, "
              f"replays equal: {replays_ok}", file=sys.stderr, flush=True)
        counts = g.cpu().numpy().astype(np.float64)
        rej = r.cpu().numpy().astype(np.float64)
        # the same kernel launched one wave per point (no queue, no cost order)
        wpb, bpc = tune["shape"]
        ctx.set_launch(min(wpb, 8), 0)
        ctx.set_kernel_variant(capi.EMX_VARIANT_WAVE_PER_POINT)
        stage()
        torch.cuda.synchronize()
        ms_pp = timed(stage, max(1, min(a.steps, 3)))
        pp_equal = bool(torch.equal(d_out, ref_out))
        ms_plain = None
        if not a.no_plain:
            ctx.set_kernel_variant(capi.EMX_VARIANT_PLAIN)
            ms_plain = timed(stage, 1)
            print(f"[bench] thread-per-point kernel {ms_plain:.1f} ms", file=sys.stderr, flush=True)
        ctx.set_kernel_variant(capi.EMX_VARIANT_WAVE)
        ctx.set_launch(wpb, bpc)
    fl = flops.rhs_flops(mechanism.load("gri30"))
    F = float(flops.qss_point_flops(fl, counts, rej).sum())
    res = {
        "metric": "GRI-scale point-integrations/s per split dt @rtol1e-6, uneven stiffness",
        "value": N / (ms_graph * 1e-3), "unit": "point-integrations/s", "n_gpus": 1,
        "steps": a.steps, "warmup": 1, "ms_per_step": ms_graph, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded uneven-stiffness states)",
        "config": {"workload": "GRI-Mech 3.0 chemistry stage, QSS, 60% burned / 30% flame front / "
                               "10% pre-ignition (900-1200 K)", "points": N, "dt_s": dt, "seed": 3},
        "failed_points": int((ref_status.cpu().numpy() != 0).sum()),
        "launch_shape": {"waves_per_block": wpb, "blocks_per_cu": bpc, "autotune_ms": tune["ms"]},
        "ms_first_stage_index_order": ms_first, "ms_eager": ms_eager, "ms_graph": ms_graph,
        "graph_replays_equal_eager": replays_ok,
        "wave_per_point_launch": {"ms_per_step": ms_pp, "speedup": ms_pp / ms_graph,
                                  "results_equal": pp_equal},
        "rhs_evals_per_point": dict(_pct(counts), spread=float(counts.max() / counts.min())),
        "roofline": {"bound": "fp64-vector", "achieved": F / (ms_graph * 1e-3) / 1e12,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": F / (ms_graph * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                     "traffic": None, "flops_per_launch": F,
                     "note": "FLOPs from this run's evaluation counts (equal to the oracle's "
                             "when parity holds); exp/log count 1 op"},
    }
    if ms_plain:
        res["plain_gpu"] = {"ms_per_step": ms_plain, "value": N / (ms_plain * 1e-3),
                            "speedup": ms_plain / ms_graph}
    if not a.no_cpu_baseline:
        # parity of the captured stage over all N points, and the CPU rate of the same run
        from oracle import oracle as O
        info = cpu_info()
        threads = a.cpu_threads or info["threads_used"]
        om = O.Mech(mechanism.table_path("gri30"))
        print(f"[bench] uneven: CPU oracle on {N} points, {threads} threads", file=sys.stderr, flush=True)
        c0 = time.perf_counter()
        ref, rst, rg, rr, _ = om.integrate(st, dt, O.options(), nthreads=threads)
        cpu_s = time.perf_counter() - c0
        got = ref_out.cpu().numpy()
        ok_T = bool((np.abs(got[:, 1] - ref[:, 1]) <= 1e-6 * np.abs(ref[:, 1])).all())
        ok_Y = bool((np.abs(got[:, 2:] - ref[:, 2:]) <= 1e-6 * np.abs(ref[:, 2:]) + 1e-12).all())
        res["parity"] = {"max_rel_T": float((np.abs(got[:, 1] - ref[:, 1]) / ref[:, 1]).max()),
                         "status_equal": bool((ref_status.cpu().numpy() == rst).all()),
                         "evaluation_counts_equal": bool((counts == rg).all()),
                         "within_rtol1e-6_atol1e-12": ok_T and ok_Y}
        res["cpu_baseline"] = {"value": N / cpu_s, "unit": "point-integrations/s", "cores": threads,
                               "kind": "port", "sample": f"all {N} states, one run: {cpu_s:.2f} s"}
    return res


def run_diffusion(a, dev):
    """Diffusion stage of 53 species on 65536 points: 8 BDF1 tridiagonal solves per call."""
    import numpy as np
    import torch

    from ember_amd import capi, workloads
    K, N = 53, (a.points if a.points != 4096 else 65536)
    dt = 2e-5
    case = workloads.diffusion_case(K, N, seed=4)
    ctx = capi.Context("gri30")
    s = torch.cuda.current_stream(dev)
    ctx.set_stream(s.cuda_stream)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    B, D, x, y0 = t(case.B), t(case.D), t(case.x), t(case.y0)
    B_row = t(case.B_row.astype(np.int32))
    sc = torch.zeros_like(y0)  # splitConstDiff (read by the kernel as in Ember)
    y = y0.clone()
    h = np.full(K, dt)
    call = lambda: ctx.diffusion_integrate_device(y, B, D, x, 0.0, dt, B_row=B_row, split_const=sc,
                                                  h_per_var=h)
    for _ in range(a.warmup):
        y.copy_(y0)
        call()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(a.steps):
        call()
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    solves = 8
    algo = 8.0 * N * (4 * K + 5)  # bytes per solve: D, y0, splitConst, y per species + 5 grid rows
    gbs = algo * solves / (ms * 1e-3) / 1e9
    print(f"[bench] diffusion {K}x{N}: {ms:.3f} ms per stage ({solves} solves), {gbs:.0f} GB/s",
          file=sys.stderr, flush=True)
    # parity of one stage against the oracle's Thomas restatement, and the CPU rate
    y.copy_(y0)
    call()
    got = y.cpu().numpy()
    res = {
        "metric": "diffusion stage HBM GB/s (53 species x 65536 points)", "value": gbs,
        "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded smooth diffusion profiles)",
        "config": {"workload": "batched tridiagonal diffusion, 53 species x 65536 points, "
                               "FixedValue ends, dt = h = 2e-5 s (8 BDF1 solves)",
                   "points": N, "species": K, "seed": 4},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s",
                     "frac": gbs / 8000.0,
                     "traffic": DIFFUSION_TRAFFIC_PER_SOLVE if (K, N) == (53, 65536) else None,
                     "algorithmic_bytes_per_solve": algo, "solves_per_step": solves},
    }
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        threads = a.cpu_threads or cpu_info()["threads_used"]
        cpu_s, runs, ref = timed_cpu(
            lambda: O.diffusion_integrate(case.B, case.B_row, case.D, np.zeros_like(case.y0), case.x,
                                          case.y0, 0.0, dt, h_per_var=h, nthreads=threads),
            a.cpu_repeats)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        res["cpu_baseline"] = {"value": algo * solves / cpu_s / 1e9, "unit": "GB/s",
                               "cores": threads, "kind": "port",
                               "sample": f"the full stage, Thomas per species in parallel; median "
                                         f"of {len(runs)} runs after 1 warm-up: {cpu_s:.3f} s"}
        res["parity"] = {"max_abs_err_over_max": float(err)}
    return res


def run_split(a, dev):
    """Whole split steps (SplitSolver::step: diffusion 1/4 x 4 + production, balanced
    splitting) on a flame-like profile with frozen transport, device-resident
    (emx_split_step_device) against the same sequence driven from the host with host
    arrays (per-stage copies in and out, bookkeeping in numpy)."""
    import numpy as np
    import torch

    from ember_amd import capi, workloads

    name = a.mech
    N = a.points if a.points != 4096 or name == "gri30" else 150
    dt = 1e-5
    case = workloads.flame_profile(N, seed=1, name=name)
    nV = case.state.shape[1]
    ctx = capi.Context(name, global_timestep=dt)
    s = torch.cuda.current_stream(dev)
    ctx.set_stream(s.cuda_stream)
    t = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(dev)
    B, D, x, Brow = t(case.B), t(case.D), t(case.x), t(case.B_row.astype(np.int32))
    status = torch.zeros(N, dtype=torch.int32, device=dev)

    def device_run(steps):
        ss = capi.SplitState(t(case.state))
        tt = 0.0
        for _ in range(steps):
            ctx.split_step_device(ss, B, D, x, tt, dt, B_row=Brow, status=status)
            tt += dt
        return ss

    # host-driven: FlameStages' sequence with host arrays
    def host_run(steps):
        st = case.state.copy()
        z = lambda: np.zeros_like(st)
        ddt_c, ddt_d, ddt_p = z(), z(), z()
        tt = 0.0
        for _ in range(steps):
            sc_d = 0.25 * (ddt_p + ddt_c + 0.0 - 3 * ddt_d)
            sc_c = 0.25 * (ddt_p + ddt_d + 0.0 - 3 * ddt_c)
            sc_p = 0.5 * (ddt_c + ddt_d + 0.0 - ddt_p)
            dd, dp = z(), z()
            def diff(t0, t1):
                nonlocal st
                s0 = st
                st = ctx.diffusion_integrate(st, case.B, case.D, case.x, t0, t1, B_row=case.B_row,
                                             split_const=sc_d, wall_var=1)
                return st - s0
            dd += diff(tt, tt + 0.25 * dt)
            dd += diff(tt + 0.25 * dt, tt + 0.5 * dt)
            s0 = st
            st, _, _, _ = ctx.source_integrate(st, (tt + dt) - tt, split_const=sc_p, t_start=tt)
            dp += st - s0
            dd += diff(tt + 0.5 * dt, tt + 0.75 * dt)
            dd += diff(tt + 0.75 * dt, tt + dt)
            ddt_c = 0.0 / dt - sc_c
            ddt_d = dd / dt - sc_d
            ddt_p = dp / dt - sc_p
            tt += dt
        return st

    steps = max(1, a.steps)
    device_run(1)
    torch.cuda.synchronize()
    c0 = time.perf_counter()
    ss = device_run(steps)
    torch.cuda.synchronize()
    ms_dev = (time.perf_counter() - c0) / steps * 1e3
    host_run(1)
    c0 = time.perf_counter()
    ref = host_run(steps)
    ms_host = (time.perf_counter() - c0) / steps * 1e3
    got = ss.state.cpu().numpy()
    print(f"[bench] split {name} N={N}: device-resident {ms_dev:.3f} ms/step, host-driven "
          f"{ms_host:.3f} ms/step", file=sys.stderr, flush=True)
    res = {
        "metric": f"split steps/s ({name}, {N}-point flame profile)", "value": 1e3 / ms_dev,
        "unit": "split steps/s", "n_gpus": 1, "steps": steps, "warmup": 1, "ms_per_step": ms_dev,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (tanh flame profile, frozen seeded transport)",
        "config": {"workload": f"device-resident split step, {name}, {N} points, dt = 1e-5 s, "
                               "balanced splitting, QSS", "points": N, "seed": 1},
        "host_driven_ms_per_step": ms_host, "speedup_vs_host_driven": ms_host / ms_dev,
        "max_abs_diff_vs_host_driven": float(np.abs(got - ref).max()),
    }
    return res


def run_thermo(a, dev):
    """Thermodynamic properties of updateChemicalProperties (rho, Wmx, cp, cp_k, h_k) for
    GRI-Mech 3.0 on 65536 points."""
    import numpy as np
    import torch

    from ember_amd import capi, workloads
    N = a.points if a.points != 4096 else 65536
    st = workloads.tanh_progress_states(N, seed=2, name="gri30")
    ctx = capi.Context("gri30")
    s = torch.cuda.current_stream(dev)
    ctx.set_stream(s.cuda_stream)
    nV = st.shape[1]
    K = nV - 2
    f = dict(dtype=torch.float64, device=dev)
    d = torch.from_numpy(st).to(dev)
    rho, Wmx, cp = (torch.empty(N, **f) for _ in range(3))
    cpk, hk = torch.empty((N, K), **f), torch.empty((N, K), **f)
    call = lambda: ctx.thermo_properties_device(d, rho, cp, Wmx=Wmx, cpSpec=cpk, hk=hk)
    for _ in range(a.warmup):
        call()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(a.steps):
        call()
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    algo = 8.0 * N * (nV + 3 + 2 * K)  # read the state, write rho, Wmx, cp, cp_k, h_k
    gbs = algo / (ms * 1e-3) / 1e9
    print(f"[bench] thermo properties gri30 x {N}: {ms * 1e3:.1f} us per call, {gbs:.0f} GB/s",
          file=sys.stderr, flush=True)
    res = {
        "metric": "thermo properties HBM GB/s (GRI-Mech 3.0, 65536 points)", "value": gbs,
        "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded tanh progress-variable states)",
        "config": {"workload": "updateChemicalProperties thermo block: rho, Wmx, cp, cp_k, h_k "
                               "for GRI-Mech 3.0 (53 species)", "points": N, "seed": 2},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s",
                     "frac": gbs / 8000.0, "traffic": None, "algorithmic_bytes_per_call": algo},
    }
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        from ember_amd import mechanism
        threads = a.cpu_threads or cpu_info()["threads_used"]
        om = O.Mech(mechanism.table_path("gri30"))
        cpu_s, runs, ref = timed_cpu(lambda: om.thermo_properties(st, nthreads=threads), a.cpu_repeats)
        got = {"rho": rho, "Wmx": Wmx, "cp": cp, "cpSpec": cpk, "hk": hk}
        err = {k: float(np.abs(v.cpu().numpy() - ref[k]).max() / np.abs(ref[k]).max())
               for k, v in got.items()}
        res["cpu_baseline"] = {"value": algo / cpu_s / 1e9, "unit": "GB/s", "cores": threads,
                               "kind": "port",
                               "sample": f"all {N} points, OpenMP dynamic,1; median of {len(runs)} "
                                         f"runs after 1 warm-up: {cpu_s * 1e3:.1f} ms"}
        res["parity"] = {"max_abs_err_over_max": err}
    return res


def transport_flops(X, threshold):
    """FP64 operations of the ApproxMixTransport evaluation per point, from its mole
    fractions (N, K): each pair term of the mass-weighted diffusion sums 15 (the fitted
    binary coefficient 10, two quotients and two sums), each Wilke term 12, each species'
    own fits and sums 30; sqrt and division count 1."""
    import numpy as np
    K = X.shape[1]
    nm = (X >= threshold).sum(axis=1).astype(np.float64)
    pairs = nm * (K - 1) + (K - nm) * nm
    return float((pairs * 15 + nm * nm * 12 + K * 30).sum())


def run_transport(a, dev):
    """Transport block of updateChemicalProperties (ApproxMixTransport: mu, lambda,
    rho D_k) for GRI-Mech 3.0 on 65536 points."""
    import numpy as np
    import torch

    from ember_amd import capi, mechanism, workloads
    N = a.points if a.points != 4096 else 65536
    st = workloads.tanh_progress_states(N, seed=2, name="gri30")
    ctx = capi.Context("gri30")
    s = torch.cuda.current_stream(dev)
    ctx.set_stream(s.cuda_stream)
    K = st.shape[1] - 2
    f = dict(dtype=torch.float64, device=dev)
    d = torch.from_numpy(st).to(dev)
    mu, lam, rhoD = torch.empty(N, **f), torch.empty(N, **f), torch.empty((N, K), **f)
    call = lambda: ctx.transport_properties_device(d, mu, lam, rhoD)
    for _ in range(max(1, a.warmup)):
        call()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(a.steps):
        call()
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    m = mechanism.load("gri30")
    Y = np.maximum(st[:, 2:], 0.0)
    Y /= Y.sum(axis=1, keepdims=True)
    X = Y / m.W
    X = np.maximum(1e-20, X / X.sum(axis=1, keepdims=True))
    F = transport_flops(X, ctx.options.transport_threshold)
    print(f"[bench] transport properties gri30 x {N}: {ms * 1e3:.1f} us per call", file=sys.stderr, flush=True)
    res = {
        "metric": "transport properties points/s (GRI-Mech 3.0, 65536 points)", "value": N / (ms * 1e-3),
        "unit": "points/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded tanh progress-variable states)",
        "config": {"workload": "updateChemicalProperties transport block, ApproxMixTransport "
                               "(threshold 1e-5): mu, lambda, rho D_k for GRI-Mech 3.0",
                   "points": N, "seed": 2},
        "roofline": {"bound": "fp64-vector", "achieved": F / (ms * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": F / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, "traffic": None,
                     "flops_per_launch": F,
                     "note": "pair and Wilke terms counted from the states' major-species sets"},
    }
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        threads = a.cpu_threads or cpu_info()["threads_used"]
        om = O.Mech(mechanism.table_path("gri30"))
        cpu_s, runs, ref = timed_cpu(lambda: om.transport_properties(st, nthreads=threads),
                                     min(a.cpu_repeats, 3))
        got = (mu.cpu().numpy(), lam.cpu().numpy(), rhoD.cpu().numpy())
        err = {k: float((np.abs(g - r) / np.abs(r)).max()) for k, g, r in zip(("mu", "lambda", "rhoD"), got, ref)}
        res["cpu_baseline"] = {"value": N / cpu_s, "unit": "points/s", "cores": threads, "kind": "port",
                               "sample": f"all {N} points, OpenMP dynamic,1; median of {len(runs)} runs "
                                         f"after 1 warm-up: {cpu_s * 1e3:.1f} ms"}
        res["parity"] = {"max_rel_err": err}
    return res


def run_copy(a, dev):
    """Device-to-device copy of 512 MiB: the reference transfer for calibrating the
    rocprofv3 FETCH_SIZE / WRITE_SIZE counters (bytes read and written are known)."""
    import torch
    n = 512 * 1024 * 1024 // 8
    x = torch.ones(n, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    for _ in range(a.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.steps):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    gbs = 2 * n * 8 / (ms * 1e-3) / 1e9
    return ({"metric": "device copy GB/s (512 MiB read + 512 MiB written)", "value": gbs,
                      "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
                      "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
                      "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                      "config": {"workload": "torch copy_ of 512 MiB"}})


def adjoint_cost(ctx, N, dt, dev, stream, steps):
    """Cost of the discrete adjoint on N flame-front states (progress variable in
    [0.3, 0.8], seed 5): forward stage, checkpointed forward, and backward sweep giving the
    gradients of the per-point integrated heat release w.r.t. the initial state and all
    ln rate multipliers."""
    import torch

    from ember_amd import workloads
    st = workloads.flame_front_states(N, seed=5, name="gri30")
    x = torch.from_numpy(st).to(dev)
    out = torch.empty_like(x)
    Q = torch.empty(N, dtype=torch.float64, device=dev)
    status = torch.empty(N, dtype=torch.int32, device=dev)
    gs = torch.empty_like(x)
    gk = torch.empty((N, ctx.n_reactions), dtype=torch.float64, device=dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    reps = max(1, min(steps, 5))

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    stage = timed(lambda: ctx.source_integrate_device(x, out, dt))
    fwd = timed(lambda: ctx.adjoint_forward(x, out, dt, heat_release=Q, status=status))
    bwd = timed(lambda: ctx.adjoint_backward(x, dt, gs, grad_lnk=gk, status=status))
    print(f"[bench] adjoint on {N} flame-front states: stage {stage:.3f} ms, checkpointed "
          f"forward {fwd:.3f} ms, backward {bwd:.3f} ms", file=sys.stderr, flush=True)
    return {"workload": f"GRI-Mech 3.0, {N} flame-front states (c in [0.3, 0.8], seed 5), "
                        "gradients of integrated heat release w.r.t. T, Y_k and 325 ln k",
            "stage_ms": stage, "forward_ms": fwd, "backward_ms": bwd,
            "backward_over_stage": bwd / stage, "adjoint_total_over_stage": (fwd + bwd) / stage,
            "failed_points": int((status.cpu().numpy() != 0).sum())}



def convection_flops(N, K, steps_utw, steps_species):
    """Algorithmic FLOPs of one convection stage: per accepted Dormand-Prince step 6
    right-hand sides and the stage combinations, error estimate and norm per component.
    Per point and right-hand side: UTW 35 (density, continuity sum, three upwind
    differences, the three equations), species 8 (velocity interpolation, one upwind
    difference); per component and step 60 (21 FMA of the stage sums, 7 of the error
    estimate, the norm).  Rejected steps are not counted."""
    utw = steps_utw * N * (6 * 35 + 3 * 60)
    species = sum(int(s) for s in steps_species) * N * (6 * 8 + 60)
    return float(utw + species)


def run_convection(a, dev):
    """Convection stage (ConvectionSystemSplit::integrateToTime) on a GRI-Mech 3.0 flame
    profile: U = a sqrt(rho_u / rho) with a = 150 1/s, inflow rV(0) = 0.5 rho_u, seeded
    density derivatives and split constants, over half of a 2e-5 s split step."""
    import numpy as np
    import torch

    from ember_amd import capi, mechanism, workloads
    N = a.points if a.points != 4096 else 400
    case = workloads.flame_profile(N, seed=3, name="gri30")
    m = mechanism.load("gri30")
    st = case.state.copy()
    Y = st[:, 2:]
    rho = 101325.0 / (mechanism.GAS_CONSTANT * st[:, 1] * (Y / m.W).sum(axis=1))
    rng = np.random.default_rng(3)
    strain = 150.0
    st[:, 0] = strain * np.sqrt(rho[0] / rho)
    sc = np.zeros_like(st)
    sc[:, 0] = rng.normal(0, 50.0, N)
    sc[:, 1] = rng.normal(0, 1e4, N)
    sc[:, 2:] = rng.normal(0, 10.0, Y.shape) * Y
    drhodt = rng.normal(0, 1.0, N) * rho
    t1 = 1e-5
    kw = dict(alpha=0, beta=1, left_bc=2, right_bc=1, rVzero=rho[0] * 0.5, Tleft=300.0, rhou=rho[0],
              strain_initial=strain, strain_final=strain)
    ctx = capi.Context("gri30")
    s = torch.cuda.current_stream(dev)
    ctx.set_stream(s.cuda_stream)
    K = m.n_species
    t = lambda arr: torch.from_numpy(np.ascontiguousarray(arr)).to(dev)
    d_in, d_x, d_sc, d_dr = t(st), t(case.x), t(sc), t(drhodt)
    d_out = torch.empty_like(d_in)
    steps = torch.zeros(1 + K, dtype=torch.int32, device=dev)
    params = ctx.convection_params(Yleft=st[0, 2:].copy(), **kw)
    call = lambda: ctx.convection_integrate_device(d_in, d_out, d_x, 0.0, t1, params=params, split_const=d_sc,
                                                   drhodt=d_dr, steps=steps)
    for _ in range(max(1, a.warmup)):
        call()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(a.steps):
        call()
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    stp = steps.cpu().numpy()
    F = convection_flops(N, K, int(stp[0]), stp[1:])
    print(f"[bench] convection gri30 x {N}: {ms * 1e3:.1f} us per stage, steps UTW {stp[0]}, "
          f"species {stp[1:].min()}-{stp[1:].max()}", file=sys.stderr, flush=True)
    res = {
        "metric": "convection stages/s (GRI-Mech 3.0 flame profile, 400 points)", "value": 1e3 / ms,
        "unit": "stages/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded tanh flame profile, strained flow field)",
        "config": {"workload": "ConvectionSystemSplit::integrateToTime: UTW system, then 53 species systems; "
                               "continuity fixedLeft, ControlVolume inflow, strain rate 150 1/s, t = 1e-5 s",
                   "points": N, "seed": 3},
        "integrator_steps": {"utw": int(stp[0]), "species_min": int(stp[1:].min()),
                             "species_max": int(stp[1:].max())},
        "roofline": {"bound": "fp64-vector", "achieved": F / (ms * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": F / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, "traffic": None,
                     "flops_per_launch": F,
                     "note": "latency-bound by construction: one workgroup for the UTW system, one per "
                             "species, each step a chain of dependent right-hand sides"},
    }
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        threads = a.cpu_threads or cpu_info()["threads_used"]
        om = O.Mech(mechanism.table_path("gri30"))
        op = O.conv_params(**kw)
        cpu_s, runs, ref = timed_cpu(lambda: om.convection_integrate(case.x, st, 0.0, t1, params=op, split_const=sc,
                                                                     drhodt=drhodt, Yleft=st[0, 2:], nthreads=threads),
                                     a.cpu_repeats)
        got = d_out.cpu().numpy()
        rel = np.abs(got - ref[0]) / (np.abs(ref[0]) + 1e-12)
        res["cpu_baseline"] = {"value": 1.0 / cpu_s, "unit": "stages/s", "cores": threads, "kind": "port",
                               "sample": f"the same stage, species systems over OpenMP threads; median of "
                                         f"{len(runs)} runs after 1 warm-up: {cpu_s * 1e3:.2f} ms"}
        res["parity"] = {"max_rel_U_T": float(rel[:, :2].max()), "max_rel_Y_floor1e-12": float(rel[:, 2:].max()),
                         "within_rtol1e-6": bool((np.abs(got - ref[0]) <= 1e-6 * np.abs(ref[0]) + 1e-12).all())}
    return res


if __name__ == "__main__":
    main()
