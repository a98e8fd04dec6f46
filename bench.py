#!/usr/bin/env python
"""Benchmark: batched chemistry production stage, GRI-Mech 3.0 (53 species, 325
reactions), N = 4096 points, dt = 2e-5 s, P = 1 atm, FP64, QSS with Ember's default
tolerances.  States: tanh progress-variable blend of CH4/air (phi = 1, 300 K) and its
adiabatic equilibrium, 1% log-normal noise on minor species (seed 2).

A step is one production stage over all points with the inputs resident in HBM.
Prints one JSON line (metric point-integrations/s).

Other workloads (one JSON line each, not the headline):
  --workload uneven     N = 65536 states of very uneven stiffness (60% near-equilibrium
                        burned, 30% flame front, 10% pre-ignition at 900-1200 K, seed 3):
                        autotuned launch shape, hipGraph-captured stage, against the same
                        kernel launched one wave per point and the thread-per-point kernel.
  --workload diffusion  53 species x 65536 points diffusion stage (seeded smooth profiles,
                        seed 4, FixedValue ends, dt = 2e-5 s, h = dt: 8 BDF1 solves), HBM GB/s.
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
# profiles/diffusion_k53_n65536.md (diff_step + pcr kernels, one solve of 53 x 65536).
STAGE_TRAFFIC = 1555 * 2 * 1024 + 2355 * 1024
DIFFUSION_TRAFFIC_PER_SOLVE = (48650 + 5220) * 2 * 1024 + (42200 + 1696) * 1024


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
    p.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost measurement")
    p.add_argument("--workload", choices=["stage", "uneven", "diffusion", "split", "copy", "thermo"],
                   default="stage")
    p.add_argument("--mech", default="gri30", help="mechanism of the split workload")
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
    if a.workload == "uneven":
        return run_uneven(a, dev)
    if a.workload == "diffusion":
        return run_diffusion(a, dev)
    if a.workload == "copy":
        return run_copy(a, dev)
    if a.workload == "split":
        return run_split(a, dev)
    if a.workload == "thermo":
        return run_thermo(a, dev)
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
        np_steps = max(1, min(a.steps, 3))
        ms_plain = timed(np_steps)
        ctx.set_kernel_variant(capi.EMX_VARIANT_WAVE)
        print(f"[bench] plain kernel: {ms_plain:.3f} ms/stage", file=sys.stderr, flush=True)
    else:
        ms_plain = None

    adj = None
    if not a.no_adjoint and integ == capi.EMX_QSS and kin == capi.EMX_KIN_STANDARD:
        adj = adjoint_cost(ctx, N, dt, dev, stream, a.steps)

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
        "host_buffer_ms_per_step": ms_e2e,
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
        # one core, on the first 128 states
        c0 = time.perf_counter()
        om.integrate(states[:128], dt, oopts, nthreads=1)
        cpu1_s = time.perf_counter() - c0
        cpu["one_core"] = {"value": 128 / cpu1_s, "sample": f"first 128 benchmark states, {cpu1_s:.2f} s"}
        res["cpu_baseline"] = cpu
        # parity of this run
        dT = np.abs(out[:, 1] - ref[:, 1]) / np.abs(ref[:, 1])
        dY = np.abs(out[:, 2:] - ref[:, 2:]) / (np.abs(ref[:, 2:]) + 1e-6)
        # the tolerance of the stage contract: |dT| <= 1e-6 |T|, |dY| <= 1e-6 |Y| + 1e-12
        ok_T = bool((np.abs(out[:, 1] - ref[:, 1]) <= 1e-6 * np.abs(ref[:, 1])).all())
        ok_Y = bool((np.abs(out[:, 2:] - ref[:, 2:]) <= 1e-6 * np.abs(ref[:, 2:]) + 1e-12).all())
        res["parity"] = {"max_rel_T": float(dT.max()), "max_rel_Y_floor1e-6": float(dY.max()),
                         "status_equal": bool((st == rst).all()),
                         "within_rtol1e-6_atol1e-12": ok_T and ok_Y}
        g, r = rg.astype(np.float64), rr.astype(np.float64)

    if integ == capi.EMX_QSS:
        F = float(flops.qss_point_flops(fl, g, r).sum())
        res["roofline"] = {"bound": "mfma", "achieved": F / (ms * 1e-3) / 1e12, "peak": 78.6,
                           "unit": "TFLOP/s", "frac": F / (ms * 1e-3) / 1e12 / 78.6,
                           "traffic": STAGE_TRAFFIC if (integ == capi.EMX_QSS and kin == capi.EMX_KIN_STANDARD
                                                        and N == 4096) else None,
                           "flops_per_launch": F, "note": "FP64 peak; exp/log count 1 op (lower bound)"}
        res["rhs_evals_per_point"] = {"mean": float(g.mean()), "max": float(g.max()), "min": float(g.min())}
    if adj:
        res["adjoint"] = adj
    res["wall_s"] = wall
    print(json.dumps(res))


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
        "roofline": {"bound": "mfma", "achieved": F / (ms_graph * 1e-3) / 1e12, "peak": 78.6,
                     "unit": "TFLOP/s", "frac": F / (ms_graph * 1e-3) / 1e12 / 78.6,
                     "traffic": None, "flops_per_launch": F,
                     "note": "FLOPs from this run's evaluation counts; exp/log count 1 op"},
    }
    if ms_plain:
        res["plain_gpu"] = {"ms_per_step": ms_plain, "value": N / (ms_plain * 1e-3),
                            "speedup": ms_plain / ms_graph}
    print(json.dumps(res))


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
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        c0 = time.perf_counter()
        ref = O.diffusion_integrate(case.B, case.B_row, case.D, np.zeros_like(case.y0), case.x,
                                    case.y0, 0.0, dt, h_per_var=h, nthreads=threads)
        cpu_s = time.perf_counter() - c0
        err = np.abs(got - ref).max() / np.abs(ref).max()
        res["cpu_baseline"] = {"value": algo * solves / cpu_s / 1e9, "unit": "GB/s",
                               "cores": threads, "kind": "port",
                               "sample": f"the full stage, Thomas per species in parallel, {cpu_s:.2f} s"}
        res["parity"] = {"max_abs_err_over_max": float(err)}
    print(json.dumps(res))


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
    print(json.dumps(res))


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
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        om = O.Mech(mechanism.table_path("gri30"))
        om.thermo_properties(st[:64], nthreads=threads)
        c0 = time.perf_counter()
        ref = om.thermo_properties(st, nthreads=threads)
        cpu_s = time.perf_counter() - c0
        got = {"rho": rho, "Wmx": Wmx, "cp": cp, "cpSpec": cpk, "hk": hk}
        err = {k: float(np.abs(v.cpu().numpy() - ref[k]).max() / np.abs(ref[k]).max())
               for k, v in got.items()}
        res["cpu_baseline"] = {"value": algo / cpu_s / 1e9, "unit": "GB/s", "cores": threads,
                               "kind": "port",
                               "sample": f"all {N} points, OpenMP dynamic,1, {cpu_s * 1e3:.1f} ms"}
        res["parity"] = {"max_abs_err_over_max": err}
    print(json.dumps(res))


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
    print(json.dumps({"metric": "device copy GB/s (512 MiB read + 512 MiB written)", "value": gbs,
                      "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
                      "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
                      "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                      "config": {"workload": "torch copy_ of 512 MiB"}}))


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


if __name__ == "__main__":
    main()
