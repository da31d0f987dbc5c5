#!/usr/bin/env python3
"""Benchmark of the NuFI rho path on one MI355X.

Workload (one "step" of this benchmark = one whole run to time T, i.e. every
time step 0..Nt of the device-resident loop: trace kernel + field kernel per
time step): 1D two-stream instability, x in [0,10pi), f0 = (1+0.001cos(0.2x))
v^2 exp(-v^2/2)/sqrt(2pi), v in [-8,8], Nx=256 order-4 splines, Nv=1024,
dt=1/32, T=50 (1600 steps).  It runs in both forms: the plain FP64 method and
the FP32-compensated trace; the reported `value` is the FP64 form's
characteristic steps per second end to end (time-to-T), the FP32C form and
its electric-energy error against the FP64 run are reported beside it.
The 3D3V Landau case (16^3 x 32^3, dt=1/8, T=20) runs once per form as well
(--no-3d skips it), and so do the 512-run batched sweep and the adjoint case
(--no-extra skips them); their figures are reported in `forms`.

Prints ONE JSON line (rank 0, N = 1).
"""
from __future__ import annotations

import argparse
import json
import os
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
LDS_PEAK_TBS = 128 * CUS * CLOCK_GHZ * 1e9 / 1e12   # 128 B per clock per CU


def progress(msg):
    """One progress line on stderr per finished leg (the JSON line stays alone on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def sync_all(sched):
    sched.synchronize()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def timed_runs(conf, precision, steps, warmup, flops_per_cs, tune_n=None, shape=None, keep_history=False):
    import numericalflowiteration_amd as nufi
    with nufi.GpuScheduler(conf, precision=precision) as s:
        if shape is not None:
            s.set_launch_shape(*shape)
        for _ in range(warmup):
            s.run()
        tuned = None
        if tune_n is not None:          # untimed: pick the launch shape on this history
            if warmup == 0:             # no run yet: compute the slices the probe step traces through
                s.run(0, tune_n)
            tuned = s.autotune(tune_n, repeats=2)
        sync_all(s)
        s.set_timing(True)
        t0 = time.perf_counter()
        energies = None
        for _ in range(steps):
            energies = s.run()
        sync_all(s)
        wall = time.perf_counter() - t0
        launches, kernel_ms, csteps = s.trace_timing()
        shape = s.launch_shape()
        history = s.history() if keep_history else None
    cs_run = conf.char_steps(0, conf.Nt + 1)
    achieved = flops_per_cs * csteps / launches / (kernel_ms / 1e3 / launches) / 1e12   # TFLOP/s
    return dict(
        wall_s=wall, ms_per_run=wall / steps * 1e3, char_steps_per_run=cs_run,
        cs_per_s_e2e=cs_run * steps / wall,
        cs_per_s_kernel=csteps / (kernel_ms / 1e3),
        trace_launches=launches, trace_avg_ms=kernel_ms / launches,
        trace_flops_per_launch=flops_per_cs * csteps / launches,
        achieved_tflops=achieved, launch_shape=list(shape),
        autotune_probe=None if tuned is None else {"n": tune_n, "best_ms": tuned[1]},
    ), energies, history


def landau3d_forms(args):
    """The 3D3V Landau case, one run to T per form (FP64, then FP32C)."""
    import numpy as np
    import numericalflowiteration_amd as nufi
    c5 = nufi.landau_3d()
    t3 = None if args.no_autotune else c5.Nt // 2     # a mid-run step: short probes pick chunks that lose later
    r64, f64, h64 = timed_runs(c5, nufi.FP64, 1, 0, FLOPS_3D, t3, keep_history=True)
    progress(f"3D3V Landau FP64 {r64['ms_per_run'] / 1e3:.1f} s per run")
    r32, f32, h32 = timed_runs(c5, nufi.FP32C, 1, 0, FLOPS_3D, t3, keep_history=True)
    progress(f"3D3V Landau FP32C {r32['ms_per_run'] / 1e3:.1f} s per run")
    r32.update(energy_errors(f32, f64))
    r32.update(field_errors(h32, h64, c5.stride_t()))
    r32["speedup_vs_fp64"] = r32["cs_per_s_e2e"] / r64["cs_per_s_e2e"]
    if args.dump_energies:
        np.save(os.path.join(args.dump_energies, "landau3d_energy_fp64.npy"), f64)
        np.save(os.path.join(args.dump_energies, "landau3d_energy_fp32c.npy"), f32)
    return {"fp64": r64, "fp32c": r32,
            "workload": "3D3V Landau, x in [0,4pi)^3, (1+0.01 sum cos(0.5x_d)) M3(v), "
                        "v in [-6,6]^3, 16^3 x 32^3 nodes, dt=1/8, T=20 (161 steps)"}


def field_errors(h, ref, stride):
    """Normwise relative error of the E-field spline coefficients (the history
    slices) of a run against the FP64 run, maximum over the time steps."""
    import numpy as np
    a, b = h.reshape(-1, stride), ref.reshape(-1, stride)
    errs = np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)
    return {"field_coeff_max_rel_err_vs_fp64": float(errs.max()), "field_coeff_err_at_step": int(errs.argmax())}


def energy_errors(e, ref):
    """Electric-energy error of a run against the FP64 run: the pointwise maximum
    of |e-ref|/ref, and the time-normwise ||e-ref||_2/||ref||_2.  The pointwise
    figure is ill-conditioned where the energy of an oscillating mode passes near
    zero (Landau damping); both are reported."""
    import numpy as np
    d = np.abs(e - ref)
    return {"energy_max_rel_err_vs_fp64": float(np.max(d / np.abs(ref))),
            "energy_normwise_rel_err_vs_fp64": float(np.linalg.norm(e - ref) / np.linalg.norm(ref)),
            "energy_err_at_max_pointwise": int(np.argmax(d / np.abs(ref)))}


def batch_sweep():
    """The 512-run sweep (alpha~U[0.001,0.5], k~U[0.3,0.6], L=2pi/k, Nx in {32..256},
    Nv=4Nx, v in [-10,10], Maxwellian or two-stream, dt=1/16, T=20) as one batch:
    shared launches per step, hipGraph replay, autotuned shape; FP32C and FP64."""
    import numpy as np
    import numericalflowiteration_amd as nufi
    confs = [nufi.sweep_config(r) for r in nufi.sweep_runs()]
    res = {"workload": "512 independent 1D runs (seeded sweep), dt=1/16, T=20, one batch",
           "runs": len(confs), "nx_histogram": {str(n): sum(c.Nx == n for c in confs) for n in (32, 64, 128, 256)}}
    energies = {}
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
        res[name] = {"wall_s": wall, "char_steps": cs, "cs_per_s_e2e": cs / wall, "launch_shape": list(shape)}
    errs = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(energies["fp32c"], energies["fp64"])]
    res["fp32c"]["energy_normwise_rel_err_vs_fp64_max_over_runs"] = max(errs)
    res["fp32c"]["speedup_vs_fp64"] = res["fp32c"]["cs_per_s_e2e"] / res["fp64"]["cs_per_s_e2e"]
    return res


def adjoint_case():
    """Gradient of J = sum_n dt*energy_n w.r.t. the 8 multimode amplitudes (Nx=128,
    Nv=512, x in [0,4pi), v in [-10,10], dt=1/16, T=20) by the device adjoint,
    checked against central differences (h=1e-6) of the device forward model."""
    import numpy as np
    import numericalflowiteration_amd as nufi
    from numericalflowiteration_amd import adjoint as A
    a0 = nufi.adjoint_modes()
    conf = nufi.multimode_1d(a0)
    A.energy_and_grad(conf, a0)                           # warm-up
    t0 = time.perf_counter()
    J, g = A.energy_and_grad(conf, a0)
    wall = time.perf_counter() - t0
    h, fd = 1e-6, []
    for m in range(len(a0)):
        ap, am = list(a0), list(a0)
        ap[m] += h
        am[m] -= h
        Jp = A.energy_and_grad(conf, ap)[0]
        Jm = A.energy_and_grad(conf, am)[0]
        fd.append((Jp - Jm) / (2 * h))
    fd = np.array(fd)
    return {"workload": "adjoint: 1D multimode (8 amplitudes), Nx=128, Nv=512, dt=1/16, T=20 (321 steps)",
            "J": J, "grad": g.tolist(), "fd_device_h1e-6": fd.tolist(),
            "max_rel_err_vs_fd": float(np.max(np.abs(g - fd)) / np.max(np.abs(fd))),
            "forward_plus_adjoint_s": wall,
            "forward_char_steps": conf.char_steps(0, conf.Nt + 1)}


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


def cpu_baseline(conf, hist_gpu, seconds_budget=15.0):
    """The oracle (CPU restatement of the reference path, OpenMP over x as
    bin/test_nufi_cpu_1d.cpp:65-70) timed on isolated steps n of this workload,
    each from the GPU FP64 run's own history (bin/test_nufi_cpu_3d_isolated.cpp
    pattern).  Returns char-steps/s."""
    import numpy as np
    from tests import oracle_lib as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    st = conf.stride_t()
    hist = np.zeros(hist_gpu.size + 8)
    hist[: hist_gpu.size] = hist_gpu
    total_cs, total_t, steps_done = 0, 0.0, []
    for n in range(conf.Nt, 0, -50):
        t = time.perf_counter()
        O.rho(conf, n, hist, nthreads=threads)
        dt = time.perf_counter() - t
        total_cs += conf.n_nodes() * n
        total_t += dt
        steps_done.append(n)
        if total_t > seconds_budget:
            break
    assert hist_gpu.size >= (conf.Nt + 1) * st
    return dict(value=total_cs / total_t, unit="char-steps/s", cores=threads, kind="port",
                sample=f"rho of isolated time steps n={steps_done} of the 1D two-stream run (Nx=256, Nv=1024), "
                       f"each from the GPU FP64 history; {total_cs:.3e} char-steps in {total_t:.1f} s; "
                       f"OpenMP over x ({threads} threads)")


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
                if "trace" in k:
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
                  "TCC_HIT / (TCC_HIT + TCC_MISS). VALU issue is SQ_INSTS_VALU (wave instructions) per SIMD "
                  f"({4 * CUS} SIMDs) and clock.", "",
                  "| kernel | mean us | FP64 TFLOP/s executed | FP32 TFLOP/s executed | VALU instructions per SIMD per clock | LDS load+store TB/s (frac of peak) | LDS bank-conflict / VALU-busy cycles | L2 hit rate |",
                  "|---|---|---|---|---|---|---|---|"]
        for k in sorted(set(pv) | set(pl)):
            if "nufi" not in k or k not in dur:
                continue
            us = dur[k]
            c = pv.get(k, {})
            f64 = 64.0 * c.get("SQ_INSTS_VALU_FLOPS_FP64", 0.0) / ( us * 1e-6 ) / 1e12
            f32 = 64.0 * c.get("SQ_INSTS_VALU_FLOPS_FP32", 0.0) / ( us * 1e-6 ) / 1e12
            lds_b = 64.0 * ( c.get("SQ_INSTS_LDS_LOAD_BANDWIDTH", 0.0) + c.get("SQ_INSTS_LDS_STORE_BANDWIDTH", 0.0) )
            lds = lds_b / ( us * 1e-6 ) / 1e12
            ipc = c.get("SQ_INSTS_VALU", 0.0) / ( 4 * CUS * us * 1e-6 * CLOCK_GHZ * 1e9 )
            busy = c.get("SQ_ACTIVE_INST_VALU", 0.0)
            conf = c.get("SQ_LDS_BANK_CONFLICT", 0.0) / busy if busy else float("nan")
            h = pl.get(k, {})
            hit, miss = h.get("TCC_HIT_sum", 0.0), h.get("TCC_MISS_sum", 0.0)
            hr = hit / ( hit + miss ) if hit + miss else float("nan")
            lines.append(f"| `{k[:70]}` | {us:.1f} | {f64:.2f} | {f32:.2f} | {ipc:.3f} | "
                         f"{lds:.2f} ({lds / LDS_PEAK_TBS:.3f}) | {conf:.3f} | {hr:.4f} |")
            result.setdefault("counters", {})[k] = {"mean_us": us, "fp64_tflops": f64, "fp32_tflops": f32,
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
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D3V Landau case")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched sweep and the adjoint case")
    ap.add_argument("--shape", default="", help="B,N,C: fixed launch shape of the primary FP64 run (no autotune)")
    ap.add_argument("--only", choices=["fp64", "fp32c", "both"], default="both")
    ap.add_argument("--case", choices=["all", "twostream", "landau3d"], default="all",
                    help="twostream: only the 1D two-stream runs; landau3d: only the 3D3V Landau runs "
                         "(for profiling one kernel at a time; the JSON line then carries that case)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--dump-energies", default="", help="directory for the per-step energy arrays (.npy)")
    ap.add_argument("--peak-probe", action="store_true",
                    help="only measure the FMA and LDS peaks (writes profiles/peaks_mi355x.*), then exit")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy N MiB device-to-device (counter calibration), then exit")
    ap.add_argument("--summarize-profiles", nargs=2, metavar=("OUTDIR", "TAG"),
                    help="write profiles/TAG.md (+ _pmc.json) from rocprofv3 databases under OUTDIR")
    ap.add_argument("--profile-commands", default="", help="';;'-separated commands recorded in the summary")
    args = ap.parse_args()

    if args.summarize_profiles:
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

    import numpy as np
    import numericalflowiteration_amd as nufi

    conf = nufi.twostream_1d()
    out = {}
    e64 = e32 = None
    hist = None
    tune = None if ( args.no_autotune or args.shape ) else conf.Nt
    shape = tuple( int( v ) for v in args.shape.split( "," ) ) if args.shape else None
    if args.case == "landau3d":
        print(json.dumps({"case": "landau3d", "forms": landau3d_forms(args)}))
        return

    if args.only in ("fp64", "both"):
        out["fp64"], e64, _ = timed_runs(conf, nufi.FP64, args.steps, args.warmup, FLOPS_1D, tune, shape)
        progress(f"two-stream FP64 {out['fp64']['ms_per_run']:.1f} ms per run")
        with nufi.GpuScheduler(conf) as s:
            s.run()
            hist = s.history()
    if args.only in ("fp32c", "both"):
        out["fp32c"], e32, h32 = timed_runs(conf, nufi.FP32C, args.steps, args.warmup, FLOPS_1D, tune,
                                            keep_history=True)
        progress(f"two-stream FP32C {out['fp32c']['ms_per_run']:.1f} ms per run")
        if e64 is not None:
            out["fp32c"].update(energy_errors(e32, e64))
            out["fp32c"].update(field_errors(h32, hist, conf.stride_t()))
            out["fp32c"]["speedup_vs_fp64"] = out["fp32c"]["cs_per_s_e2e"] / out["fp64"]["cs_per_s_e2e"]

    if not args.no_3d and args.only == "both" and args.case == "all":
        out["landau_3d"] = landau3d_forms(args)
        progress("3D3V Landau done")
    if args.dump_energies and e64 is not None:
        np.save(os.path.join(args.dump_energies, "twostream_energy_fp64.npy"), e64)
        if e32 is not None:
            np.save(os.path.join(args.dump_energies, "twostream_energy_fp32c.npy"), e32)

    if not args.no_extra and args.only == "both" and args.case == "all":
        out["batch_sweep"] = batch_sweep()
        progress("batched sweep done")
        out["adjoint"] = adjoint_case()
        progress("adjoint done")

    prim = out.get("fp64") or out.get("fp32c")
    prim_prec = "fp64" if "fp64" in out else "fp32c"
    peak = PEAK_FP64 if prim_prec == "fp64" else PEAK_FP32
    roofline = {
        "bound": "valu", "achieved": prim["achieved_tflops"], "peak": peak, "unit": "TFLOP/s",
        "frac": prim["achieved_tflops"] / peak,
        "traffic": read_traffic(f"trace1d_twostream_{prim_prec}_pmc.json"),
        "traffic_note": "HBM bytes per trace launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary",
        "kernel": "trace1d_poly_kernel" if prim_prec == "fp64" else "trace1d_fp32c_kernel",
        "flops_per_char_step": FLOPS_1D,
        "peak_measured": read_measured_peak("fp64_fma" if prim_prec == "fp64" else "fp32_fma"),
        "peak_measured_note": "FMA-chain probe on the box (profiles/peaks_mi355x.md); peak stays the published figure",
    }
    cpu = None
    if not args.no_cpu and hist is not None:
        cpu = cpu_baseline(conf, hist)
        progress("CPU baseline done")

    line = {
        "metric": METRIC,
        "value": prim["cs_per_s_e2e"],
        "unit": "char-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": prim["ms_per_run"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prim_prec == "fp64" else "f32 (compensated) + f64",
        "data": "synthetic: analytic initial condition, no dataset",
        "config": {"workload": "1D two-stream instability, Nx=256 order-4 B-splines, Nv=1024, "
                               "v in [-8,8], x in [0,10pi), dt=1/32, T=50 (1601 time steps incl. n=0)",
                   "step": "one whole run to T (time-to-T); value = char-steps of the run / time-to-T",
                   "char_steps_per_run": prim["char_steps_per_run"],
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "forms": out,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
