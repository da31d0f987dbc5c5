#!/usr/bin/env python3
"""Benchmark of the NuFI rho path on one MI355X.

Headline workload: 3D3V linear Landau damping, the plain FP64 method:
x in [0,4pi)^3, f0 = (1 + 0.01 sum_d cos(0.5 x_d)) exp(-|v|^2/2)/(2pi)^(3/2),
v in [-6,6]^3, 16^3 order-4 tensor B-splines, 32^3 velocity nodes per point,
dt = 1/8, T = 20 (161 time steps, 1.73e12 characteristic steps per run).

One bench step is one device time step of that run (the trace of all 1.34e8
nodes through the stored field history, the rho partial sums, the Poisson
solve and spline fit, the slice stored) at n = 80, the step whose cost is the
run's mean (sum_n n / 161 = 80), recomputed from the run's own history (the
isolated-step pattern of bin/test_nufi_cpu_3d_isolated.cpp:169-228): W untimed
steps, then K timed; `value` is the char-steps per second of the K steps.
Beside it, per form (FP64 and the FP32-compensated trace): time-to-T of the
whole run, the FP32C errors against the FP64 run (maximum over the time steps
of the normwise relative error of the mean-free rho, of E at the nodes and of
the electric energy), the trace kernel's roofline from HIP events on the
context stream, the CPU baseline (the oracle on isolated steps at a sample of
spatial points, extrapolated over the run) and parity against the oracle at
full size.  The other workloads (1D two-stream, 1D Landau, the 512-run batched
sweep, the adjoint) are reported under `forms` with their own CPU baselines.

Prints ONE JSON line (N = 1).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "char-steps/s + time-to-T at fixed E-energy error; FP32-comp vs FP64, 1 GPU"
FLOPS_1D = 30.0      # algorithmic FLOPs per characteristic step, order 4, 1D
FLOPS_3D = 475.0     # 3D
PEAK_FP64 = 78.6     # TFLOP/s, MI355X FP64 vector
PEAK_FP32 = 157.3    # TFLOP/s, MI355X FP32 vector
CUS, CLOCK_GHZ = 256, 2.4
SHADER_ENGINES = 32  # 8 XCDs x 4; SQ_BUSY_CYCLES counts per shader engine
LDS_PEAK_TBS = 128 * CUS * CLOCK_GHZ * 1e9 / 1e12   # 128 B per clock per CU
PROFILE_KERNEL = "trace"                              # summarize_profiles: the kernel whose traffic is kept


def progress(msg):
    """One progress line on stderr per finished leg (the JSON line stays alone on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cpu_threads():
    """Threads for the CPU baseline: the physical cores lscpu reports, capped by
    this process's CPU affinity and OMP_NUM_THREADS; returns (threads, physical
    cores, CPUs in the affinity mask)."""
    phys = None
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
        phys = len({ln for ln in out.splitlines() if ln and not ln.startswith("#")}) or None
    except Exception:
        pass
    avail = len(os.sched_getaffinity(0))
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(v for v in (phys, avail, env) if v)
    return threads, phys, avail


def mean_free(a):
    return a - a.mean()


def rel(a, b):
    import numpy as np
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


# ---- 3D3V Landau (headline) ---------------------------------------------------

def landau3d_form(conf, precision, steps, warmup, autotune, flops_per_cs, kernel):
    """One context: (optionally autotuned) whole run to T (time-to-T, the
    production path: a replayed hipGraph of the steps), then the bench steps at
    n = Nt/2, each trace launch of which is timed by HIP events on the context
    stream (the trace kernel's duration is affine in n, so its mean over the
    run equals its duration at n = Nt/2).  Returns (figures, energies, history,
    scheduler)."""
    import numericalflowiteration_amd as nufi
    n_mid = conf.Nt // 2
    s = nufi.GpuScheduler(conf, precision=precision)
    tuned = None
    if autotune:
        s.run(0, n_mid + 1)
        tuned = s.autotune(n_mid, repeats=2)
    s.synchronize()
    sync()
    t0 = time.perf_counter()
    energies = s.run()
    s.synchronize()
    t_run = time.perf_counter() - t0
    history = s.history()
    for _ in range(warmup):
        s.step(n_mid, wait=False)
    s.set_timing(True)            # the trace launches of the timed steps
    s.synchronize()
    sync()                        # device-wide synchronisation on both sides of the timed steps
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(n_mid, wait=False)
    s.synchronize()
    sync()
    t_steps = time.perf_counter() - t0
    roof = trace_roofline(s, flops_per_cs, kernel, PEAK_FP64) if steps else None
    s.set_timing(False)
    cs_run = conf.char_steps(0, conf.Nt + 1)
    cs_step = conf.n_nodes() * n_mid
    r = {"time_to_T_s": t_run, "cs_per_s_time_to_T": cs_run / t_run, "char_steps_per_run": cs_run,
         "bench_step_n": n_mid, "char_steps_per_step": cs_step, "ms_per_step": t_steps / max(steps, 1) * 1e3,
         "cs_per_s_steps": cs_step * steps / t_steps if steps else None,
         "launch_shape": list(s.launch_shape()), "trace_roofline": roof,
         "autotune_probe": None if tuned is None else {"n": n_mid, "best_ms": tuned[1]}}
    return r, energies, history, s


def trace_roofline(s, flops_per_cs, kernel, peak):
    """The trace kernel over every launch timed on the context (HIP events on its
    stream): algorithmic FLOPs per launch over the mean launch duration."""
    launches, kernel_ms, csteps = s.trace_timing()
    flops = flops_per_cs * csteps / launches
    avg_s = kernel_ms / launches / 1e3
    return {"kernel": kernel, "launches": launches, "avg_launch_ms": avg_s * 1e3,
            "flops_per_launch": flops, "achieved": flops / avg_s / 1e12, "peak": peak,
            "frac": flops / avg_s / 1e12 / peak, "cs_per_s_kernel": csteps / (kernel_ms / 1e3)}


def cpu_landau3d(conf, hist, s_gpu, threads, budget_points=48):
    """The oracle (the reference's dim3 eval_rho, rho.hpp:428-462, OpenMP over the
    sampled points) on isolated steps n = 1, 40, 80, 120, 160 of the GPU FP64
    history at `budget_points` spatial points each; char-steps/s per step, the
    rate interpolated over n for the CPU time-to-T.  The same points give the
    full-size parity of the GPU's rho (reference summation order, and the
    exact v-sum at 16 points of steps 80 and 160)."""
    import numpy as np
    from tests import oracle_lib as O
    st = conf.stride_t()
    h = np.zeros(hist.size + 8)
    h[: hist.size] = hist
    rng = np.random.default_rng(7)
    pts = np.concatenate([[0, conf.n_space() - 1],
                          rng.choice(np.arange(1, conf.n_space() - 1), budget_points - 2, replace=False)])
    nvel = conf.n_vel()
    steps = [1, conf.Nt // 4, conf.Nt // 2, 3 * conf.Nt // 4, conf.Nt]
    rates, par_ref, par_exact = {}, {}, {}
    total_cs, total_t = 0, 0.0
    for n in steps:
        rho = np.zeros(conf.n_space())
        s_gpu.compute_rho(n)
        s_gpu.download_rho(rho)
        t = time.perf_counter()
        ref = O.rho3d_points(conf, n, h, pts, nthreads=threads)
        dt = time.perf_counter() - t
        rates[n] = pts.size * nvel * n / dt
        total_cs += pts.size * nvel * n
        total_t += dt
        scale = np.linalg.norm(mean_free(rho)[pts])
        par_ref[n] = float(np.linalg.norm((rho[pts] + 1.0) - ref) / scale)
        if n in (conf.Nt // 2, conf.Nt):
            ex = O.rho3d_points(conf, n, h, pts[:16], nthreads=threads, exact_sum=True)
            par_exact[n] = float(np.linalg.norm((rho[pts[:16]] + 1.0) - ex) / np.linalg.norm(mean_free(rho)[pts[:16]]))
    ns = np.array(steps, dtype=float)
    rs = np.array([rates[n] for n in steps])
    t_T = sum(conf.n_nodes() * n / float(np.interp(n, ns, rs)) for n in range(1, conf.Nt + 1))
    cpu = {"value": conf.char_steps(0, conf.Nt + 1) / t_T, "unit": "char-steps/s", "cores": threads, "kind": "port",
           "time_to_T_s_extrapolated": t_T,
           "sample": f"the oracle's rho (reference summation order) of isolated time steps n={steps} of the 3D "
                     f"Landau run at {pts.size} of the {conf.n_space()} spatial points (all {nvel} velocities each), "
                     f"from the GPU FP64 history; {total_cs:.3e} char-steps in {total_t:.1f} s; rate per step "
                     f"interpolated over n = 1..{conf.Nt} for the time-to-T (Poisson and fit excluded: O(N) per step)",
           "rate_by_step": {str(n): rates[n] for n in steps}}
    parity = {"rho_rel_err_vs_oracle_reference_sum": par_ref, "rho_rel_err_vs_oracle_exact_sum_16pts": par_exact,
              "points": int(pts.size), "note": "GPU FP64 rho(n) on the run's history vs the oracle at the sampled "
                                               "points, scaled by the mean-free GPU rho there"}
    return cpu, parity


# ---- 1D two-stream ---------------------------------------------------------------

def timed_runs(conf, precision, runs, warmup, tune_n=None, graph=True):
    """`runs` whole runs to T (a replayed hipGraph of the steps unless graph is
    False) after `warmup` untimed ones (autotuned at tune_n); then one more run
    with every trace launch timed by HIP events for the kernel figures."""
    import numericalflowiteration_amd as nufi
    with nufi.GpuScheduler(conf, precision=precision) as s:
        s.set_run_graph(graph)
        for _ in range(warmup):
            s.run()
        if tune_n is not None:
            if warmup == 0:
                s.run(0, tune_n)
            s.autotune(tune_n, repeats=2)
            s.run()                                  # the graph of the tuned shape
        s.synchronize()
        sync()
        t0 = time.perf_counter()
        energies = None
        for _ in range(runs):
            energies = s.run()
        s.synchronize()
        sync()
        wall = time.perf_counter() - t0
        s.set_timing(True)
        s.run()
        roof = trace_roofline(s, FLOPS_1D, "trace1d", PEAK_FP64)
        s.set_timing(False)
        history = s.history()
        shape = s.launch_shape()
    cs_run = conf.char_steps(0, conf.Nt + 1)
    return {"runs": runs, "ms_per_run": wall / runs * 1e3, "char_steps_per_run": cs_run, "graph": graph,
            "cs_per_s_time_to_T": cs_run * runs / wall, "cs_per_s_kernel": roof["cs_per_s_kernel"],
            "trace_avg_ms": roof["avg_launch_ms"], "achieved_tflops": roof["achieved"],
            "launch_shape": list(shape)}, energies, history


def twostream(args, threads, with_cpu):
    import numpy as np
    import numericalflowiteration_amd as nufi
    from numericalflowiteration_amd import diagnostics as D
    conf = nufi.twostream_1d()
    tune = None if args.no_autotune else conf.Nt
    r64, e64, h64 = timed_runs(conf, nufi.FP64, args.runs_1d, 1, tune)
    r32, e32, h32 = timed_runs(conf, nufi.FP32C, args.runs_1d, 1, tune)
    r32.update(D.run_errors(conf, h32, e32, h64, e64))
    r32["speedup_vs_fp64"] = r32["cs_per_s_time_to_T"] / r64["cs_per_s_time_to_T"]
    out = {"workload": "1D two-stream, x in [0,10pi), (1+0.001cos(0.2x)) v^2 M(v), v in [-8,8], Nx=256 order 4, "
                       "Nv=1024, dt=1/32, T=50 (1601 steps)", "fp64": r64, "fp32c": r32,
           "host_round_trip": host_round_trip(conf, e64)}
    progress(f"two-stream FP64 {r64['ms_per_run']:.1f} ms, FP32C {r32['ms_per_run']:.1f} ms per run")
    if with_cpu:
        from tests import oracle_lib as O
        st = conf.stride_t()
        hist = np.zeros(h64.size + 8)
        hist[: h64.size] = h64
        par, total_cs, total_t, done = {}, 0, 0.0, []
        with nufi.GpuScheduler(conf) as s:
            s.upload_history(h64[: (conf.Nt + 1) * st], 0, conf.Nt + 1)
            for n in range(conf.Nt, 0, -50):
                rho = np.zeros(conf.Nx)
                s.compute_rho(n)
                s.download_rho(rho)
                t = time.perf_counter()
                ref = O.rho(conf, n, hist, nthreads=threads)
                dt = time.perf_counter() - t
                total_cs += conf.n_nodes() * n
                total_t += dt
                done.append(n)
                par[n] = rel(mean_free(rho + 1.0), mean_free(ref))
                if total_t > 12.0:
                    break
            exact = {}
            for n in (conf.Nt, conf.Nt // 2, 1):
                rho = np.zeros(conf.Nx)
                s.compute_rho(n)
                s.download_rho(rho)
                exact[n] = rel(mean_free(rho), mean_free(O.rho(conf, n, hist, nthreads=threads, exact_sum=True)))
        rate = total_cs / total_t
        out["cpu_baseline"] = {"value": rate, "unit": "char-steps/s", "cores": threads, "kind": "port",
                               "time_to_T_s_extrapolated": conf.char_steps(0, conf.Nt + 1) / rate,
                               "sample": f"the oracle's rho of isolated steps n={done} of the GPU FP64 history "
                                         f"(all 262144 nodes each); {total_cs:.3e} char-steps in {total_t:.1f} s"}
        out["parity_vs_oracle"] = {"rho_rel_err_reference_sum": {str(k): v for k, v in par.items()},
                                   "rho_rel_err_exact_sum": {str(k): v for k, v in exact.items()},
                                   "note": "mean-free rho, GPU FP64 vs the oracle on the same history"}
        progress("two-stream CPU baseline done")
    return out


def host_round_trip(conf, e_dev):
    """The reference driver's step loop through the C ABI (bin/test_nufi_gpu_1d.cpp:
    241-253): compute_rho -> download_rho (PCIe) -> Poisson and fit on host
    buffers (each a copy in and out) -> upload_phi (PCIe), driven from Python;
    the whole run to T, its energies against the device-resident run."""
    import numpy as np
    import numericalflowiteration_amd as nufi
    st = conf.stride_t()
    coeffs = np.zeros((conf.Nt + 1) * st)
    e = np.zeros(conf.Nt + 1)
    with nufi.GpuScheduler(conf) as s:
        rho = np.zeros(conf.n_space())
        t0 = time.perf_counter()
        for n in range(conf.Nt + 1):
            rho[:] = 0.0
            s.compute_rho(n)
            s.download_rho(rho)
            e[n] = s.poisson_solve(rho)
            coeffs[n * st:(n + 1) * st] = s.interpolate(rho)
            s.upload_phi(n, coeffs)
        wall = time.perf_counter() - t0
    cs = conf.char_steps(0, conf.Nt + 1)
    return {"ms_per_run": wall * 1e3, "cs_per_s_time_to_T": cs / wall,
            "energy_max_rel_err_vs_device_loop": float(np.max(np.abs(e - e_dev) / np.abs(e_dev))),
            "note": "host buffers cross PCIe every step (rho down, slice up) and each call synchronises; "
                    "never the headline value"}


# ---- 1D Landau ----------------------------------------------------------------------

def landau1d(args, threads, with_cpu):
    import numpy as np
    import numericalflowiteration_amd as nufi
    from tests import oracle_lib as O
    conf = nufi.landau_1d()
    tune = None if args.no_autotune else conf.Nt
    r64, e64, _ = timed_runs(conf, nufi.FP64, max(args.runs_1d, 5), 1, tune)
    rd, _, _ = timed_runs(conf, nufi.FP64, max(args.runs_1d, 5), 1, tune, graph=False)
    out = {"workload": "1D Landau, x in [0,4pi), (1+0.01cos(0.5x)) M(v), v in [-10,10], Nx=64 order 4, Nv=256, "
                       "dt=1/16, T=30 (481 steps)", "fp64": r64, "fp64_launched_per_step": rd}
    progress(f"1D Landau FP64 {r64['ms_per_run']:.2f} ms per run (graph), {rd['ms_per_run']:.2f} ms launched per step")
    if with_cpu:
        h = O.new_history(conf, conf.Nt + 1)
        t = time.perf_counter()
        e = O.run(conf, 0, conf.Nt + 1, h, nthreads=threads)
        dt = time.perf_counter() - t
        cs = conf.char_steps(0, conf.Nt + 1)
        out["cpu_baseline"] = {"value": cs / dt, "unit": "char-steps/s", "cores": threads, "kind": "port",
                               "time_to_T_s": dt, "sample": "the whole run to T on the CPU (the oracle: rho, "
                                                            "DHT Poisson, circulant fit per step)"}
        out["parity_vs_oracle"] = {"energy_max_rel_err": float(np.max(np.abs(e64 - e) / np.abs(e))),
                                   "note": "energies of the whole GPU FP64 run vs the oracle's run"}
        progress("1D Landau CPU baseline done")
    return out


# ---- batched sweep -------------------------------------------------------------------

def batch_sweep(args, threads, with_cpu):
    """The 512-run sweep (alpha~U[0.001,0.5], k~U[0.3,0.6], L=2pi/k, Nx in {32..256},
    Nv=4Nx, v in [-10,10], Maxwellian or two-stream, dt=1/16, T=20) as one batch:
    shared launches per step, hipGraph replay, autotuned shape; FP32C and FP64."""
    import numpy as np
    import numericalflowiteration_amd as nufi
    from numericalflowiteration_amd import diagnostics as D
    confs = [nufi.sweep_config(r) for r in nufi.sweep_runs()]
    res = {"workload": "512 independent 1D runs (seeded sweep), dt=1/16, T=20, one batch",
           "runs": len(confs), "nx_histogram": {str(n): sum(c.Nx == n for c in confs) for n in (32, 64, 128, 256)}}
    energies, hists = {}, {}
    for name, prec in (("fp32c", nufi.FP32C), ("fp64", nufi.FP64)):
        with nufi.BatchScheduler(confs, precision=prec) as b:
            b.run(graph=True)                          # warm-up, captures the graph
            b.synchronize()
            shape, _ = b.autotune(n_probe=confs[0].Nt, repeats=1)
            b.run(graph=True)                          # recapture for the tuned shape
            b.synchronize()
            t0 = time.perf_counter()
            b.run(graph=True)
            b.synchronize()
            wall = time.perf_counter() - t0
            cs = b.char_steps(0, max(c.Nt for c in confs) + 1)
            energies[name] = [b.energies(r) for r in range(len(confs))]
            hists[name] = [b.history(r) for r in range(len(confs))]
        res[name] = {"wall_s": wall, "char_steps": cs, "cs_per_s_time_to_T": cs / wall, "launch_shape": list(shape)}
    errs = [D.run_errors(c, h32, e32, h64, e64) for c, h32, e32, h64, e64 in
            zip(confs, hists["fp32c"], energies["fp32c"], hists["fp64"], energies["fp64"])]
    for key in ("rho_max_rel_err", "E_max_rel_err", "energy_max_rel_err", "energy_normwise_rel_err"):
        res["fp32c"][key + "_max_over_runs"] = max(e[key] for e in errs)
    res["fp32c"]["speedup_vs_fp64"] = res["fp32c"]["cs_per_s_time_to_T"] / res["fp64"]["cs_per_s_time_to_T"]
    progress("batched sweep done")
    if with_cpu:
        from tests import oracle_lib as O
        picks = [next(i for i, c in enumerate(confs) if c.Nx == nx) for nx in (32, 64, 128, 256)]
        total_cs, total_t = 0, 0.0
        for i in picks:
            c = confs[i]
            h = np.zeros(hists["fp64"][i].size + 8)
            h[: hists["fp64"][i].size] = hists["fp64"][i]
            t = time.perf_counter()
            O.rho(c, c.Nt, h, nthreads=threads)
            total_t += time.perf_counter() - t
            total_cs += c.n_nodes() * c.Nt
        rate = total_cs / total_t
        cs_all = res["fp64"]["char_steps"]
        res["cpu_baseline"] = {"value": rate, "unit": "char-steps/s", "cores": threads, "kind": "port",
                               "time_to_T_s_extrapolated": cs_all / rate,
                               "sample": f"the oracle's rho of the last step (n=320) of the first run with each Nx "
                                         f"(runs {picks}), serially, from the GPU FP64 histories; "
                                         f"{total_cs:.3e} char-steps in {total_t:.1f} s; scaled by the batch's "
                                         f"char-steps"}
        progress("batch CPU baseline done")
    return res


# ---- adjoint -------------------------------------------------------------------------

def adjoint_case(threads, with_cpu):
    """Gradient of J = sum_n dt*energy_n w.r.t. the 8 multimode amplitudes (Nx=128,
    Nv=512, x in [0,4pi), v in [-10,10], dt=1/16, T=20) by the device adjoint of
    each form (FP64, FP32C), checked against central differences (h=1e-6) of the
    device FP64 forward model."""
    import numpy as np
    import numericalflowiteration_amd as nufi
    from numericalflowiteration_amd import adjoint as A
    a0 = nufi.adjoint_modes()
    conf = nufi.multimode_1d(a0)
    out = {"workload": "adjoint: 1D multimode (8 amplitudes), Nx=128, Nv=512, dt=1/16, T=20 (321 steps)",
           "forward_char_steps": conf.char_steps(0, conf.Nt + 1)}
    h, fd = 1e-6, []
    for m in range(len(a0)):
        ap, am = list(a0), list(a0)
        ap[m] += h
        am[m] -= h
        fd.append((A.energy_and_grad(conf, ap)[0] - A.energy_and_grad(conf, am)[0]) / (2 * h))
    fd = np.array(fd)
    out["fd_fp64_device_h1e-6"] = fd.tolist()
    for name, prec in (("fp64", nufi.FP64), ("fp32c", nufi.FP32C)):
        A.energy_and_grad(conf, a0, precision=prec)                      # warm-up
        t0 = time.perf_counter()
        J, g = A.energy_and_grad(conf, a0, precision=prec)
        wall = time.perf_counter() - t0
        out[name] = {"J": J, "grad": g.tolist(),
                     "max_rel_err_vs_fd_fp64": float(np.max(np.abs(g - fd)) / np.max(np.abs(fd))),
                     "forward_plus_adjoint_s": wall}
    g64, g32 = np.array(out["fp64"]["grad"]), np.array(out["fp32c"]["grad"])
    out["fp32c"]["grad_max_rel_err_vs_fp64_adjoint"] = float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64)))
    out["fp32c"]["J_rel_err_vs_fp64"] = abs(out["fp32c"]["J"] - out["fp64"]["J"]) / abs(out["fp64"]["J"])
    progress("adjoint done")
    if with_cpu:
        from tests import oracle_lib as O
        h = O.new_history(conf, conf.Nt + 1)
        t = time.perf_counter()
        e = O.run(conf, 0, conf.Nt + 1, h, nthreads=threads)
        dt = time.perf_counter() - t
        Jc = conf.dt * float(np.sum(e))
        out["cpu_baseline"] = {"value": conf.char_steps(0, conf.Nt + 1) / dt, "unit": "char-steps/s",
                               "cores": threads, "kind": "port", "time_to_T_s": dt,
                               "sample": "the whole forward run on the CPU (the oracle)"}
        out["parity_vs_oracle"] = {"J_rel_err_fp64": abs(out["fp64"]["J"] - Jc) / abs(Jc),
                                   "J_rel_err_fp32c": abs(out["fp32c"]["J"] - Jc) / abs(Jc)}
        # the gradient on the full grid against the oracle's forward-mode tangent
        # (dual numbers), over the first 81 steps: the tangent to T costs minutes
        short = nufi.multimode_1d(a0, T=5.0)
        _, dJ = O.tangent_run(short, nthreads=threads)
        dJ = np.asarray(dJ[:len(a0)])
        for name, prec in (("fp64", nufi.FP64), ("fp32c", nufi.FP32C)):
            _, g = A.energy_and_grad(short, a0, precision=prec)
            out["parity_vs_oracle"][f"grad_rel_err_vs_oracle_tangent_{name}"] = \
                float(np.max(np.abs(g - dJ)) / np.max(np.abs(dJ)))
        out["parity_vs_oracle"]["tangent_sample"] = "Nx=128, Nv=512, dt=1/16, T=5 (81 steps)"
        progress("adjoint CPU baseline done")
    return out


PEAK_PROBE_CMD = "python3 bench.py --peak-probe"


def peak_probe():
    """Measured MI355X peaks from libnufi_probes.so: FP64 and non-packed FP32 FMA
    throughput with the core clock read inside the kernel, and LDS read
    bandwidth for 8- and 16-byte reads with coherent or scattered lanes.
    Writes profiles/peaks_mi355x.md and profiles/peaks_mi355x.json."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "numericalflowiteration_amd", "libnufi_probes.so"))
    d = ctypes.c_double
    out = {}
    for name, fp64 in (("fp64_fma", 1), ("fp32_fma", 0)):
        t, g = d(), d()
        rc = lib.nufi_probe_fma(fp64, 100000, ctypes.byref(t), ctypes.byref(g))
        if rc:
            raise RuntimeError(f"nufi_probe_fma failed: HIP error {rc}")
        out[name] = {"tflops": t.value, "core_ghz": g.value, "of_published": t.value / (PEAK_FP64 if fp64 else PEAK_FP32),
                     "insts_per_simd_clock": t.value * 1e12 / (2 * 64 * 4 * CUS * g.value * 1e9) if g.value else None}
    r, g = d(), d()
    rc = lib.nufi_probe_int(100000, ctypes.byref(r), ctypes.byref(g))
    if rc:
        raise RuntimeError(f"nufi_probe_int failed: HIP error {rc}")
    out["int32_xor_add"] = {"insts_per_simd_clock": r.value, "core_ghz": g.value}
    for width in (8, 16):
        for scattered in (0, 1):
            t, g = d(), d()
            rc = lib.nufi_probe_lds(width, scattered, 20000, ctypes.byref(t), ctypes.byref(g))
            if rc:
                raise RuntimeError(f"nufi_probe_lds failed: HIP error {rc}")
            key = f"lds_read_b{width * 8}_{'scattered' if scattered else 'coherent'}"
            out[key] = {"tbs": t.value, "core_ghz": g.value,
                        "bytes_per_clk_per_cu": t.value * 1e12 / (CUS * g.value * 1e9) if g.value else None}
    lines = ["# Measured MI355X peaks (microbenchmarks on the box)", "",
             "Command (from the repository root on the MI355X box):", "", f"    {PEAK_PROBE_CMD}", "",
             "FMA (and int32): 8 independent chains per lane, 1024 threads per CU, one launch; the core clock is the "
             "shader clock counter (s_memtime) against the 100 MHz real-time counter (s_memrealtime) over each "
             "workgroup. LDS: 256-thread workgroups, 4 per CU, 8 reads per iteration from a 32 KiB array, lanes "
             "on consecutive entries (coherent) or hashed over the array (scattered).", "",
             "| probe | rate | core clock GHz | note |", "|---|---|---|---|"]
    for k in ("fp64_fma", "fp32_fma"):
        v = out[k]
        lines.append(f"| {k} | {v['tflops']:.1f} TFLOP/s | {v['core_ghz']:.2f} | {v['of_published']:.2f} of the "
                     f"published {PEAK_FP64 if k == 'fp64_fma' else PEAK_FP32} TFLOP/s; "
                     f"{v['insts_per_simd_clock']:.3f} wave instructions per SIMD per clock |")
    v = out["int32_xor_add"]
    lines.append(f"| int32_xor_add | {v['insts_per_simd_clock']:.3f} wave instructions per SIMD per clock | "
                 f"{v['core_ghz']:.2f} | 32-bit integer chains, same shape |")
    for k, v in out.items():
        if k.startswith("lds"):
            lines.append(f"| {k} | {v['tbs']:.1f} TB/s | {v['core_ghz']:.2f} | "
                         f"{v['bytes_per_clk_per_cu']:.0f} B per clock per CU |")
    with open(os.path.join(ROOT, "profiles", "peaks_mi355x.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(ROOT, "profiles", "peaks_mi355x.json"), "w") as f:
        json.dump(out, f, indent=1)
    return out


def read_measured_peak(key):
    """A measured peak from the committed probe summary (profiles/peaks_mi355x.json), if any."""
    try:
        with open(os.path.join(ROOT, "profiles", "peaks_mi355x.json")) as f:
            return json.load(f)[key]["tflops"]
    except Exception:
        return None


def read_traffic(name):
    """HBM bytes per trace launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch")
    except Exception:
        return None


def read_executed_flops(name, kernel, char_steps_per_launch):
    """FLOPs the kernel executes per char-step (64 x SQ_INSTS_VALU_FLOPS_FP64 over its mean duration, times
    that duration, over the char-steps of a launch) from the committed rocprofv3 --pmc summary, if any."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            counters = json.load(f)["counters"]
        c = next(v for k, v in counters.items() if kernel in k)
        return c["fp64_tflops"] * 1e12 * c["mean_us"] * 1e-6 / char_steps_per_launch
    except Exception:
        return None


def summarize_profiles(outdir, tag, commands):
    """Turn rocprofv3 result databases (kernel trace + FETCH_SIZE / WRITE_SIZE
    passes + the copy-kernel calibration passes) into the committed summaries
    profiles/<tag>.md and profiles/<tag>_pmc.json."""
    import glob
    import sqlite3

    def db(name):
        hits = sorted(glob.glob(os.path.join(outdir, name, "**", "*_results.db"), recursive=True))
        return sqlite3.connect(hits[0]) if hits else None

    lines = [f"# rocprofv3 summary: {tag}", "", "Commands (from the repository root on the MI355X box):", ""]
    lines += [f"    {c}" for c in commands] + [""]
    kt = db("prof_kt")
    if kt is not None:
        lines += ["## Kernel trace (--kernel-trace --stats)", "",
                  "| kernel | calls | total us | average us | % |", "|---|---|---|---|---|"]
        for name, calls, tot, avg, pct in kt.execute("select name, total_calls, total_duration, average, "
                                                      "percentage from top_kernels limit 12"):
            lines.append(f"| `{name[:90]}` | {calls} | {tot:.1f} | {avg:.3f} | {pct:.2f} |")
        lines.append("")

    def per_kernel(con):
        return {r[0]: (r[1], r[2]) for r in con.execute(
            "select kernel_name, count(*), avg(value) from counters_collection group by kernel_name")}

    def per_counter(con):
        out = {}
        for k, name, v in con.execute("select kernel_name, counter_name, avg(s) from (select kernel_name, "
                                      "counter_name, dispatch_id, sum(value) as s from counters_collection "
                                      "group by dispatch_id, counter_name) group by kernel_name, counter_name"):
            out.setdefault(k, {})[name] = v
        return out

    fetch, write = db("prof_fetch"), db("prof_write")
    cfetch, cwrite = db("prof_cfetch"), db("prof_cwrite")
    result = {}
    if fetch is not None and write is not None and cfetch is not None and cwrite is not None:
        cf, cw = per_kernel(cfetch), per_kernel(cwrite)
        copy = [k for k in cf if "copyBuffer" in k][0]
        copy_kib = float(COPY_PROBE_MIB * 1024)
        f_scale = copy_kib / cf[copy][1]
        w_scale = copy_kib / cw[copy][1]
        pf, pw = per_kernel(fetch), per_kernel(write)
        lines += ["## HBM traffic per dispatch (--pmc FETCH_SIZE, then --pmc WRITE_SIZE; KiB)", "",
                  f"Calibration: a device-to-device copy of {COPY_PROBE_MIB} MiB reads FETCH_SIZE "
                  f"{cf[copy][1]:.1f} KiB and writes WRITE_SIZE {cw[copy][1]:.1f} KiB, so the scale factors are "
                  f"fetch x{f_scale:.4f}, write x{w_scale:.4f}.", "",
                  "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | scaled HBM bytes |",
                  "|---|---|---|---|---|"]
        for k in pf:
            if k in pw and "nufi" in k:
                b = pf[k][1] * f_scale * 1024 + pw[k][1] * w_scale * 1024
                lines.append(f"| `{k[:80]}` | {pf[k][0]} | {pf[k][1]:.1f} | {pw[k][1]:.1f} | {b:.4g} |")
                if PROFILE_KERNEL in k:
                    result = {"kernel": k, "dispatches": pf[k][0], "fetch_kib_raw": pf[k][1],
                              "write_kib_raw": pw[k][1], "fetch_scale": f_scale, "write_scale": w_scale,
                              "hbm_bytes_per_launch": b}
        lines.append("")
    valu, l2 = db("prof_valu"), db("prof_l2")
    if kt is not None and ( valu is not None or l2 is not None ):
        dur = {r[0]: r[1] for r in kt.execute("select name, average from top_kernels")}   # us
        pv = per_counter(valu) if valu is not None else {}
        pl = per_counter(l2) if l2 is not None else {}
        lines += ["## VALU FLOPs, LDS bandwidth and L2 hit rate (--pmc, one pass each; per dispatch)", "",
                  f"Peaks (MI355X, {CUS} CUs at {CLOCK_GHZ} GHz): FP64 vector {PEAK_FP64} TFLOP/s, FP32 vector "
                  f"{PEAK_FP32} TFLOP/s, LDS {LDS_PEAK_TBS:.1f} TB/s (128 B per clock per CU, the CDNA3 figure; "
                  "profiles/peaks_mi355x.md has the FMA and LDS rates measured on the box). The counters are "
                  "means per dispatch; rates divide them by the kernel's mean duration from the kernel trace. "
                  "SQ_INSTS_VALU_FLOPS_FP64/FP32 add the FLOPs of one lane per wave instruction, so the executed "
                  "rate is 64 times the counter (a wave has 64 lanes); it counts what the kernel executes, "
                  "which for the 1D polynomial step is fewer FLOPs than the algorithmic 30 per char-step. "
                  "LDS bytes are 64 x SQ_INSTS_LDS_{LOAD,STORE}_BANDWIDTH; the L2 hit rate is "
                  "TCC_HIT / (TCC_HIT + TCC_MISS). The shader clock is SQ_BUSY_CYCLES per shader engine "
                  f"({SHADER_ENGINES}) over the mean duration (the box's clock while the kernel ran; the peaks "
                  f"assume {CLOCK_GHZ} GHz); VALU issue is SQ_INSTS_VALU (wave instructions) per SIMD "
                  f"({4 * CUS} SIMDs) and shader clock. The counter passes run the profiled workload alone, the kernel "
                  "trace the whole command: for kernels other workloads launch too (field_kernel) the rate "
                  "divides by a mean over all of them and is indicative only.", "",
                  "| kernel | mean us | shader clock GHz | FP64 TFLOP/s executed | FP32 TFLOP/s executed | VALU instructions per SIMD per clock | LDS load+store TB/s (frac of peak) | LDS bank-conflict / VALU-busy cycles | L2 hit rate |",
                  "|---|---|---|---|---|---|---|---|---|"]
        for k in sorted(set(pv) | set(pl)):
            if "nufi" not in k or k not in dur:
                continue
            us = dur[k]
            c = pv.get(k, {})
            f64 = 64.0 * c.get("SQ_INSTS_VALU_FLOPS_FP64", 0.0) / ( us * 1e-6 ) / 1e12
            f32 = 64.0 * c.get("SQ_INSTS_VALU_FLOPS_FP32", 0.0) / ( us * 1e-6 ) / 1e12
            lds_b = 64.0 * ( c.get("SQ_INSTS_LDS_LOAD_BANDWIDTH", 0.0) + c.get("SQ_INSTS_LDS_STORE_BANDWIDTH", 0.0) )
            lds = lds_b / ( us * 1e-6 ) / 1e12
            ghz = c.get("SQ_BUSY_CYCLES", 0.0) / SHADER_ENGINES / ( us * 1e-6 ) / 1e9
            ipc = c.get("SQ_INSTS_VALU", 0.0) / ( 4 * CUS * us * 1e-6 * ( ghz if ghz > 0 else CLOCK_GHZ ) * 1e9 )
            busy = c.get("SQ_ACTIVE_INST_VALU", 0.0)
            conf = c.get("SQ_LDS_BANK_CONFLICT", 0.0) / busy if busy else float("nan")
            h = pl.get(k, {})
            hit, miss = h.get("TCC_HIT_sum", 0.0), h.get("TCC_MISS_sum", 0.0)
            hr = hit / ( hit + miss ) if hit + miss else float("nan")
            lines.append(f"| `{k[:70]}` | {us:.1f} | {ghz:.2f} | {f64:.2f} | {f32:.2f} | {ipc:.3f} | "
                         f"{lds:.2f} ({lds / LDS_PEAK_TBS:.3f}) | {conf:.3f} | {hr:.4f} |")
            result.setdefault("counters", {})[k] = {"mean_us": us, "shader_clock_ghz": ghz,
                                                    "fp64_tflops": f64, "fp32_tflops": f32,
                                                    "valu_insts_per_simd_clock": ipc,
                                                    "lds_tbs": lds, "lds_frac": lds / LDS_PEAK_TBS,
                                                    "l2_hit_rate": hr}
        lines.append("")
    with open(os.path.join(ROOT, "profiles", f"{tag}.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    if result:
        with open(os.path.join(ROOT, "profiles", f"{tag}_pmc.json"), "w") as f:
            json.dump(result, f, indent=1)
    return result


COPY_PROBE_MIB = 256


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed bench steps (3D Landau time step n=80)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed bench steps before them")
    ap.add_argument("--runs-1d", type=int, default=5, help="timed whole runs of the 1D workloads")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines and oracle parity")
    ap.add_argument("--no-extra", action="store_true", help="skip the 1D, batch and adjoint workloads")
    ap.add_argument("--only", choices=["fp64", "fp32c", "both"], default="both", help="3D Landau forms to run")
    ap.add_argument("--case", choices=["all", "landau3d", "twostream"], default="all",
                    help="landau3d / twostream: that workload alone (profiling one kernel at a time)")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the 3D Landau launch shapes at n=80 first (the defaults are the tuned shapes)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes of the 1D workloads")
    ap.add_argument("--dump-energies", default="", help="directory for the per-step energy arrays (.npy)")
    ap.add_argument("--peak-probe", action="store_true",
                    help="only measure the FMA and LDS peaks (writes profiles/peaks_mi355x.*), then exit")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy N MiB device-to-device (counter calibration), then exit")
    ap.add_argument("--summarize-profiles", nargs=2, metavar=("OUTDIR", "TAG"),
                    help="write profiles/TAG.md (+ _pmc.json) from rocprofv3 databases under OUTDIR")
    ap.add_argument("--profile-commands", default="", help="';;'-separated commands recorded in the summary")
    ap.add_argument("--profile-kernel", default="trace", help="summary: the kernel whose HBM traffic is kept")
    args = ap.parse_args()

    if args.summarize_profiles:
        global PROFILE_KERNEL
        PROFILE_KERNEL = args.profile_kernel
        r = summarize_profiles(args.summarize_profiles[0], args.summarize_profiles[1],
                               [c for c in args.profile_commands.split(";;") if c])
        print(json.dumps(r))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the rho path runs on a single device; "
              "the multi-GPU split is a design note, not built) — run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.peak_probe:
        print(json.dumps(peak_probe()))
        return

    if args.copy_probe:
        import torch
        a = torch.empty(args.copy_probe * (1 << 20) // 8, dtype=torch.float64, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 8}))
        return

    # the CPU baseline's OpenMP runtime reads these when the oracle library loads
    threads, phys, avail = cpu_threads()
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    with_cpu = not args.no_cpu

    import numpy as np
    import numericalflowiteration_amd as nufi
    from numericalflowiteration_amd import diagnostics as D

    if args.case == "twostream":
        print(json.dumps({"case": "twostream", "forms": twostream(args, threads, with_cpu)}))
        return

    c5 = nufi.landau_3d()
    l3 = {"workload": "3D3V Landau, x in [0,4pi)^3, (1+0.01 sum cos(0.5x_d)) M3(v), v in [-6,6]^3, "
                      "16^3 x 32^3 nodes, dt=1/8, T=20 (161 steps)"}
    cpu = parity = None
    e64 = h64 = None
    if args.only in ("fp64", "both"):
        r64, e64, h64, s64 = landau3d_form(c5, nufi.FP64, args.steps, args.warmup, args.autotune, FLOPS_3D,
                                           "trace3d_xm_fast_kernel")
        l3["fp64"] = r64
        progress(f"3D Landau FP64: {r64['time_to_T_s']:.2f} s to T, {r64['ms_per_step']:.1f} ms per step n=80")
        if with_cpu and args.case == "all":
            cpu, parity = cpu_landau3d(c5, h64, s64, threads)
            progress("3D Landau CPU baseline and parity done")
        s64.close()
    if args.only in ("fp32c", "both"):
        r32, e32, h32, s32 = landau3d_form(c5, nufi.FP32C, args.steps, args.warmup, args.autotune, FLOPS_3D,
                                           "trace3d_fp32c_kernel")
        s32.close()
        if e64 is not None:
            r32.update(D.run_errors(c5, h32, e32, h64, e64))
            r32["speedup_vs_fp64_time_to_T"] = l3["fp64"]["time_to_T_s"] / r32["time_to_T_s"]
            r32["speedup_vs_fp64_steps"] = l3["fp64"]["ms_per_step"] / r32["ms_per_step"]
        l3["fp32c"] = r32
        progress(f"3D Landau FP32C: {r32['time_to_T_s']:.2f} s to T, {r32['ms_per_step']:.1f} ms per step n=80")
        if args.dump_energies:
            np.save(os.path.join(args.dump_energies, "landau3d_energy_fp32c.npy"), e32)
    if args.dump_energies and e64 is not None:
        np.save(os.path.join(args.dump_energies, "landau3d_energy_fp64.npy"), e64)

    forms = {}
    if not args.no_extra and args.case == "all":
        forms["twostream_1d"] = twostream(args, threads, with_cpu)
        forms["landau_1d"] = landau1d(args, threads, with_cpu)
        forms["batch_sweep"] = batch_sweep(args, threads, with_cpu)
        forms["adjoint"] = adjoint_case(threads, with_cpu)

    prim = l3.get("fp64") or l3.get("fp32c")
    prim_prec = "fp64" if "fp64" in l3 else "fp32c"
    roofline = dict(prim["trace_roofline"])
    roofline.update({
        "bound": "valu", "unit": "TFLOP/s", "flops_per_char_step": FLOPS_3D,
        "traffic": read_traffic("trace3d_landau3d_pmc.json"),
        "traffic_note": "HBM bytes per trace launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary "
                        "(profiles/trace3d_landau3d.md)",
        "achieved_note": "algorithmic FLOPs (475 per char-step) per launch over the mean launch duration, HIP "
                         "events on the context stream around every trace launch of the run to T and of the "
                         "bench steps",
        "peak_measured": read_measured_peak("fp64_fma"),
        "peak_measured_note": "FMA-chain probe on the box (profiles/peaks_mi355x.md); peak stays the published figure",
    })
    if prim_prec == "fp64":
        ex = read_executed_flops("trace3d_landau3d_pmc.json", "trace3d_xm_fast_kernel", prim["char_steps_per_step"])
        if ex is not None:
            roofline["executed_flops_per_char_step"] = ex
            roofline["achieved_executed"] = ex * prim["char_steps_per_step"] / (roofline["avg_launch_ms"] * 1e-3) / 1e12
            roofline["executed_note"] = ("FP64 FLOPs the kernel executes per char-step (64 x SQ_INSTS_VALU_FLOPS_FP64 "
                                         "from the committed --pmc summary) times this run's char-step rate: "
                                         "the closed-form 6B/2D bases and folded kick constants execute fewer "
                                         "FLOPs than the algorithmic 475; frac stays algorithmic")
    if cpu is not None:
        cpu["cores_physical_lscpu"] = phys
        cpu["cpus_in_affinity_mask"] = avail
        cpu["omp_proc_bind"] = os.environ.get("OMP_PROC_BIND")
    line = {
        "metric": METRIC,
        "value": prim["cs_per_s_steps"],
        "unit": "char-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": prim["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prim_prec == "fp64" else "f32 (compensated) + f64",
        "data": "synthetic: analytic initial condition, no dataset",
        "config": {"workload": "3D3V linear Landau damping, x in [0,4pi)^3, v in [-6,6]^3, 16^3 order-4 tensor "
                               "B-splines x 32^3 velocity nodes, dt=1/8, T=20 (161 time steps), FP64",
                   "step": "one device time step (trace of all nodes + rho + Poisson + fit + store) at n=80, the "
                           "run's mean-cost step, recomputed from the run's own history",
                   "char_steps_per_step": prim["char_steps_per_step"],
                   "time_to_T_s": prim["time_to_T_s"], "cs_per_s_time_to_T": prim["cs_per_s_time_to_T"],
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
        "landau_3d": l3,
        "forms": forms,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
