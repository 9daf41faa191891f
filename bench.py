"""Benchmark of the hydro2d time step on one MI355X.

Workload: Rayleigh-Taylor instability on the 1024x2048 grid (Atwood 0.5,
heavy fluid on top, cosine interface of amplitude 0.01 plus seeded 1e-4
noise).  A step is hydro::step (hydro2d.hpp:1531-1621): the SIMPLE loop
(momentum rows, red-black symmetric GS velocity sweep, Rhie-Chow, pressure
solve to 1e-8 relative residual, correction), 10 TVD advection sub-steps and
the property update, all on the GPU with fields resident in HBM.

Two paths run on the same state:
  fp64      plain FP64 Jacobi-PCG pressure solve;
  headline  FP64 flexible CG preconditioned by the FP32 multigrid V-cycle
            plus the learned correction, which is trained on the simulation's
            own pressure operator after the warm-up and kept only if it beats
            the V-cycle alone on a held-out solve (fcg_learned, mixed).
`value` is the headline steps/s.  The pressure solve to 1e-8 is also timed
on the last SIMPLE pressure system of the fp64 path for every solver.
`roofline` is the kernel with the largest share of the headline step's time
(launches per step x average duration, each duration measured with HIP
events around back-to-back launches on the kernel's own stream), with its
algorithmic bytes from hydro_amd/manifest.py and, where measured, its HBM
traffic from rocprof counters.  `batch_c5` steps 64 independent cases at
once.
`cpu_baseline` times the reference module built from its sources
(oracle/_ref/hydro_ref_fast, OpenMP) on one step of the same grid.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE on the RT 1024x2048 state, each scaled by what the counter reads
# for a 256 MiB device-to-device copy (FETCH_SIZE x2.0, WRITE_SIZE x1.0).
MEASURED_TRAFFIC = {"KUp<double, true>": 82.79e6, "KDown<double>": 58.70e6,
                    "KAdvStep<2, false>": 120.91e6, "KMomentum": 579.68e6,
                    "KAdvStep<2, false> mixed": 104.12e6, "KMomentum mixed": 370.78e6}


def run_path(text, steps, warmup, label, train=False):
    from hydro_amd.hydro import Hydro
    from hydro_amd import compare, workloads
    h = Hydro(text)
    workloads.set_rt_state(h)
    h.volumes0 = compare.phase_volumes(h)
    for i in range(warmup):
        st = h.step()
        log(f"{label} warmup step {i + 1}: SIMPLE {st.simple_iterations}, "
            f"pressure iters {st.pressure_solve_iters_total}")
    train_stats = None
    if train:
        ts = h.train_pressure_net()
        train_stats = {"pairs": ts.pairs, "steps": ts.steps,
                       "loss_first": round(ts.loss_first, 4), "loss_last": round(ts.loss_last, 4),
                       "accepted": bool(ts.accepted),
                       "heldout_iters": ts.val_iters, "heldout_ms": round(ts.val_ms, 3),
                       "heldout_iters_vcycle_alone": ts.val_iters_baseline,
                       "heldout_ms_vcycle_alone": round(ts.val_ms_baseline, 3),
                       "train_ms": round(ts.ms, 1)}
        log(f"{label} learned correction: {train_stats}")
    h.timing()  # waits for the warm-up's last kernels
    t0 = time.perf_counter()
    iters = []
    for i in range(steps):
        st = h.step()
        iters.append((st.simple_iterations, st.pressure_solve_iters_total))
    # hg_step returns with the step's advection and property kernels still
    # queued: the clock stops once they have finished
    h.timing()
    dt = time.perf_counter() - t0
    log(f"{label}: {steps} steps in {dt:.3f} s, (SIMPLE, pressure iters) per step {iters}")
    if train:
        # step kernels on the final state, HIP events around back-to-back launches
        train_stats = dict(train_stats or {})
        h.kernel_ms = {k: h.time_kernel(k, reps=20) for k in ("KAdvStep", "KMomentum")}
    return h, dt, iters, train_stats


def cpu_info():
    """CPU model and socket count of this host (for the baselines)."""
    model, sockets = None, set()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model is None:
                model = line.split(":", 1)[1].strip()
            if line.startswith("physical id"):
                sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return {"cpu_model": model, "sockets": max(1, len(sockets)),
            "threads_used": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}


def cpu_pressure_solves(nx, ny, kx, ky, rhs, sor_sweeps=600, reps=3):
    """The pressure solve to 1e-8 on the CPU, same system, zero start, same
    stopping test (oracle/hydro_oracle.cpp or_poisson_baseline, -O2 -fopenmp):
    OpenMP Jacobi-PCG to the end (median of `reps`), and the reference's
    serial lexicographic SOR (omega 1.9) as a bounded sample of `sor_sweeps`
    sweeps, with the sweeps to 1e-8 extrapolated from its residual decay
    between the sample's midpoint and its end."""
    from oracle import pyoracle
    nfx = (nx + 1) * ny
    out = {}
    secs, its, rel = [], None, None
    for _ in range(reps):
        _, sec, its, rel = pyoracle.poisson_baseline(nx, ny, kx[:nfx], kx[nfx:] if ky is None
                                                     else ky, rhs, "pcg", tol=1e-8)
        secs.append(sec)
        if sec > 10.0:  # keep the CPU leg bounded: one run of a long solve
            break
    out["omp_jacobi_pcg"] = {"ms": round(1e3 * float(np.median(secs)), 1), "iterations": its,
                             "rel_residual": rel, "runs": len(secs)}
    half = sor_sweeps // 2
    _, sec_h, _, rel_h = pyoracle.poisson_baseline(nx, ny, kx[:nfx], kx[nfx:] if ky is None
                                                   else ky, rhs, "sor", tol=1e-8,
                                                   max_iters=half, check_every=half)
    _, sec, it, rel = pyoracle.poisson_baseline(nx, ny, kx[:nfx], kx[nfx:] if ky is None else ky,
                                                rhs, "sor", tol=1e-8, max_iters=sor_sweeps,
                                                check_every=10)
    ms_sweep = 1e3 * sec / max(it, 1)
    est = None
    if rel > 1e-8 and 0 < rel < rel_h:
        per_sweep = (np.log(rel) - np.log(rel_h)) / (it - half)
        est = int(it + (np.log(1e-8) - np.log(rel)) / per_sweep)
    out["serial_sor_sample"] = {
        "sweeps": it, "ms": round(1e3 * sec, 1), "ms_per_sweep": round(ms_sweep, 3),
        "rel_residual": rel, "rel_residual_at_half": rel_h,
        "sweeps_to_1e-8_extrapolated": est,
        "ms_to_1e-8_extrapolated": round(est * ms_sweep, 0) if est else None}
    return out


def cpu_baseline(seconds_cap=150):
    """One step of the reference module (its default solvers, OpenMP) on the
    same 1024x2048 grid with the RT densities; flat interface at y = 1
    (the reference has no way to read the perturbed state)."""
    from oracle import pyoracle
    exe = pyoracle.REF_BIN_FAST
    if not os.path.exists(exe):
        return None
    from hydro_amd import workloads
    general = os.path.join(ROOT, "oracle", "_ref", "general.hydroconf")
    text = open(general).read() + workloads.rt(solver_p="gauss_seidel", solver_v="lu") + """
set vect A1 (0, 1)
set vect B1 (1, 2)
set bool no_output 1
"""
    text = text.replace("set string precision_gpu fp64", "")
    path = "/tmp/hydro_bench_ref.conf"
    open(path, "w").write(text)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    try:
        r = subprocess.run([exe, path, "1", "-", "0"], capture_output=True, text=True,
                           timeout=seconds_cap, env=env)
    except subprocess.TimeoutExpired:
        return {"value": None, "unit": "steps/s", "cores": cores, "kind": "reference",
                "sample": f"1 step of 1024x2048 RT-like case did not finish in {seconds_cap} s"}
    m = re.search(r"STEP s=1 iters=(\d+) t=([0-9.]+)", r.stdout)
    if not m:
        return None
    t = float(m.group(2))
    sweeps = [int(x) for x in re.findall(r"iter = (\d+), diff", r.stdout)]
    info = cpu_info()
    return {"value": round(1.0 / t, 5), "unit": "steps/s", "cores": cores, "kind": "reference",
            "cpu_model": info["cpu_model"], "sockets": info["sockets"],
            "sample": f"1 step of the reference module (its default gauss_seidel + lu, OpenMP, "
                      f"-O3) on the 1024x2048 two-phase RT densities with a flat interface: "
                      f"{m.group(1)} SIMPLE iterations, SOR sweeps {sweeps} (the reference's "
                      f"SOR stops at max|update| <= 1e-3, not at a 1e-8 residual), {t:.2f} s"}


def net_throughput(nx, ny):
    """BF16 MFMA network alone at the RT grid: forward and training step."""
    from hydro_amd.hydro import Net
    n = Net(nx, ny)
    t = n.time(reps=10)
    n.close()
    fwd = t["forward_flops"] / (t["forward_ms"] * 1e-3) / 1e12
    trn = t["train_flops"] / (t["train_ms"] * 1e-3) / 1e12
    return {"forward_ms": round(t["forward_ms"], 4), "train_step_ms": round(t["train_ms"], 4),
            "forward_tflops": round(fwd, 2), "train_tflops": round(trn, 2),
            "bf16_peak_tflops": BF16_PEAK_TFLOPS,
            "forward_frac": round(fwd / BF16_PEAK_TFLOPS, 4)}


def batch_c5(steps=3):
    """64 independent cavity/dam cases (sides 64..512, seeded, parameters
    jittered), FCG + V-cycle pressure, stepped concurrently from 8 host
    threads (one HIP stream per case) and, for comparison, one after the
    other."""
    import re as _re
    from hydro_amd import workloads
    from hydro_amd.hydro import Hydro, batch_run
    texts = workloads.c5_cases(count=64)
    cells = sum(int(_re.search(r"Nx (\d+)", t).group(1)) * int(_re.search(r"Ny (\d+)", t).group(1))
                for t in texts)
    hs = [Hydro(t) for t in texts]
    batch_run(hs, 1, threads=8)  # warm-up
    _, ms_conc = batch_run(hs, steps, threads=8)
    _, ms_seq = batch_run(hs, steps, threads=1)
    for h in hs:
        h.close()
    return {"cases": 64, "cells": cells, "steps": steps,
            "case_steps_per_s": round(64 * steps / (ms_conc * 1e-3), 1),
            "cell_steps_per_s": round(cells * steps / (ms_conc * 1e-3), 0),
            "ms_concurrent_8_threads": round(ms_conc, 2), "ms_one_after_another": round(ms_seq, 2)}


def step_3d(n=128, steps=5, warmup=2, cpu_n=64):
    """The 3D step (hg3_*, the reference's hydro3d): a lid-driven cavity
    n^3, FP64, the reference's jacobi as the velocity solver (bitwise the
    reference's); the pressure by FP64 Jacobi-PCG (cg) and by FP64 flexible
    CG with the FP32 multigrid V-cycle (fcg_mg), both to a 1e-8 residual,
    and the two runs' fields compared after the same steps; beside them the
    3D oracle (one thread) on one step of the same case at cpu_n^3."""
    from hydro_amd import capi, workloads
    from hydro_amd.hydro import Hydro3
    from oracle import pyoracle
    fids = (("u", capi.HG_FIELD_VELOCITY_X), ("v", capi.HG_FIELD_VELOCITY_Y),
            ("w", capi.HG_FIELD_VELOCITY_Z), ("p", capi.HG_FIELD_PRESSURE))

    def run(solver_p):
        h = Hydro3(workloads.cavity3d(n, n, n, solver_p=solver_p, solver_v="jacobi"))
        for _ in range(warmup):
            h.step()
        t0 = time.perf_counter()
        st = [h.step() for _ in range(steps)]   # hg3_step returns after its work
        dt = (time.perf_counter() - t0) / steps
        fields = {k: h.get(f) for k, f in fids}
        h.close()
        return {"steps_per_s": round(1. / dt, 3), "ms_per_step": round(dt * 1e3, 2),
                "cell_steps_per_s": round(n ** 3 / dt, 0),
                "simple_iters_per_step": float(np.mean([s.simple_iterations for s in st])),
                "pressure_iters_per_step":
                    float(np.mean([s.pressure_solve_iters_total for s in st]))}, fields

    out = {"workload": f"cavity {n}^3, FP64, jacobi velocity, pressure to 1e-8; "
                       f"{warmup} warm-up and {steps} timed steps"}
    out["fcg_mg"], fa = run("fcg_mg")
    out["cg"], fb = run("cg")
    err = {}
    for k, _ in fids:
        a, b = fa[k], fb[k]
        if k == "p":  # the all-wall operator is singular: p up to a constant
            a, b = a - a.mean(), b - b.mean()
        err[k] = float(f"{np.linalg.norm(a - b) / np.linalg.norm(b):.3e}")
    out["fcg_mg_vs_cg_rel_l2"] = err
    out["speedup_fcg_mg_vs_cg"] = round(out["cg"]["ms_per_step"] / out["fcg_mg"]["ms_per_step"], 2)
    small = workloads.cavity3d(cpu_n, cpu_n, cpu_n, solver_p="cg", solver_v="jacobi")
    o = pyoracle.Oracle3(capi.parse_config3(small))
    o.step()
    t1 = time.perf_counter()
    o.step()
    cpu_s = time.perf_counter() - t1
    o.close()
    out["cpu_oracle_one_thread"] = {"case": f"cavity {cpu_n}^3 (cg), one step", "s": round(cpu_s, 3),
                                    "cell_steps_per_s": round(cpu_n ** 3 / cpu_s, 0)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from hydro_amd import capi, manifest
    from hydro_amd.hydro import Poisson
    from hydro_amd import workloads

    nx, ny = 1024, 2048
    text64 = workloads.rt(nx, ny, solver_p="cg", solver_v="lu_rb", precision="fp64")
    textmx = workloads.rt(nx, ny, solver_p="fcg_learned", solver_v="lu_rb", precision="mixed")

    from hydro_amd import compare
    h64, t64, it64, _ = run_path(text64, args.steps, args.warmup, "fp64")
    # the SIMPLE pressure system of the last step, for the solve comparison
    kf = h64.get(capi.HG_FIELD_PRESSURE_FACE_COEF)
    rhs = h64.get(capi.HG_FIELD_PRESSURE_RHS)
    snap64, vol64 = compare.snapshot(h64), (h64.volumes0, compare.phase_volumes(h64))
    h64.close()
    hmx, tmx, itmx, train = run_path(textmx, args.steps, args.warmup, "headline", train=True)
    step_kernel_ms = hmx.kernel_ms
    substeps = 10
    # the headline path's fields against the FP64 path's after the same
    # warmup + steps from the same state
    field_err = compare.field_errors(compare.snapshot(hmx), snap64)
    field_err = {k: float(f"{v:.3e}") for k, v in field_err.items()}
    field_err["volume_drift_headline"] = [float(f"{d:.2e}") for d in compare.volume_drift(
        hmx.volumes0, compare.phase_volumes(hmx))]
    field_err["volume_drift_fp64"] = [float(f"{d:.2e}") for d in compare.volume_drift(*vol64)]
    field_err["after_steps"] = args.warmup + args.steps
    log(f"field error of the headline path vs the fp64 path: {field_err}")
    hmx.close()

    nfx = (nx + 1) * ny
    ps = Poisson(nx, ny, kf[:nfx], kf[nfx:])
    solves = {}
    # fcg_mg precision 0: the same V-cycle on FP64 levels (FCG is FP64 in both)
    for label, solver, prec in (("fp64_jacobi_pcg", "cg", 0), ("mixed_jacobi_pcg", "cg", 1),
                                ("mixed_fcg_vcycle", "fcg_mg", 1), ("fcg_fp64_vcycle", "fcg_mg", 0)):
        ms, its, res = [], None, None
        for rep in range(4):
            _, st = ps.solve(rhs, solver=solver, tol=1e-8, precision=prec)
            if rep > 0:
                ms.append(st.ms)
            its, res = st.iterations, st.rel_residual
        solves[label] = {"ms": round(float(np.median(ms)), 3), "iterations": int(its),
                         "rel_residual": float(res)}
        log(f"pressure solve {label}: {solves[label]}")

    cpu_solves = None
    if not args.no_cpu:
        cpu_solves = cpu_pressure_solves(nx, ny, kf, None, rhs)
        cpu_solves.update(cpu_info())
        log(f"CPU pressure solves (RT system): {cpu_solves}")

    # C4: the 512x512 variable-coefficient Poisson solve (density ratio 1000)
    kx4, ky4, rhs4 = workloads.poisson_problem(512, seed=0)
    ps4 = Poisson(512, 512, kx4, ky4)
    c4 = {}
    for label, solver, prec in (("fp64_jacobi_pcg", "cg", 0), ("mixed_fcg_vcycle", "fcg_mg", 1)):
        ms = []
        for rep_ in range(4):
            _, st = ps4.solve(rhs4, solver=solver, tol=1e-8, precision=prec)
            if rep_ > 0:
                ms.append(st.ms)
        c4[label] = {"ms": round(float(np.median(ms)), 3), "iterations": int(st.iterations),
                     "rel_residual": float(st.rel_residual)}
    ps4.close()
    c4["speedup_mixed_fcg_vs_fp64"] = round(c4["fp64_jacobi_pcg"]["ms"] /
                                            c4["mixed_fcg_vcycle"]["ms"], 2)
    if not args.no_cpu:
        from oracle import pyoracle
        secs = []
        for _ in range(3):
            _, sec, its, rel = pyoracle.poisson_baseline(512, 512, kx4, ky4, rhs4, "pcg", tol=1e-8)
            secs.append(sec)
        c4["cpu_omp_jacobi_pcg"] = {"ms": round(1e3 * float(np.median(secs)), 1),
                                    "iterations": its, "rel_residual": rel}
    log(f"C4 Poisson 512x512: {c4}")

    # kernel timings on the RT operator (the V-cycle's finest level and its
    # one-workgroup coarse cycle), HIP events around back-to-back launches
    kernels = {}
    for name in ("KUp<double, true>", "KDown<double>"):
        ms = ps.time_kernel(name, reps=50)
        gbs = manifest.kernel_bytes(name, nx, ny) / (ms * 1e-3) / 1e9
        kernels[name] = {"avg_us": round(ms * 1e3, 2), "gbs": round(gbs, 1),
                         "frac": round(gbs / HBM_PEAK_GBS, 4)}
    kernels["KCoarseCycle"] = {"avg_us": round(ps.time_kernel("KCoarseCycle", reps=50) * 1e3, 2),
                               "bound": "latency (one workgroup, levels <= 2048 cells)"}
    ps.close()
    cg_ms_iter = solves["fp64_jacobi_pcg"]["ms"] / max(1, solves["fp64_jacobi_pcg"]["iterations"])
    cg_gbs = manifest.cg_iteration_bytes(nx, ny) / (cg_ms_iter * 1e-3) / 1e9
    kernels["fp64 Jacobi-PCG iteration (KPcgSpmv + KPcgUpdate)"] = {
        "avg_us": round(cg_ms_iter * 1e3, 2), "gbs": round(cg_gbs, 1),
        "frac": round(cg_gbs / HBM_PEAK_GBS, 4)}
    log(f"kernels: {kernels}")

    for name, man in (("KAdvStep", "KAdvStep<2, false> mixed"), ("KMomentum", "KMomentum mixed")):
        ms = step_kernel_ms[name]
        gbs = manifest.kernel_bytes(man, nx, ny) / (ms * 1e-3) / 1e9
        kernels[man] = {"avg_us": round(ms * 1e3, 2), "gbs": round(gbs, 1),
                        "frac": round(gbs / HBM_PEAK_GBS, 4)}
    # the dominant kernel of the headline step: launches per step x duration
    simple_per_step = float(np.mean([s_ for s_, _ in itmx]))
    fcg_per_step = float(np.mean([p for _, p in itmx]))
    per_step = {"KAdvStep<2, false> mixed": substeps, "KMomentum mixed": simple_per_step,
                "KUp<double, true>": fcg_per_step, "KDown<double>": fcg_per_step}
    dom = max(per_step, key=lambda k: per_step[k] * kernels[k]["avg_us"])
    achieved = kernels[dom]["gbs"]
    traffic = MEASURED_TRAFFIC.get(dom)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(traffic) if traffic else None,
                "kernel": f"{dom}: {per_step[dom]:.1f} launches per step, "
                          f"{manifest.kernel_bytes(dom, nx, ny) / 1e6:.1f} MB algorithmic per launch"}
    log(f"dominant kernel {dom}: {roofline}")

    # the whole headline step against the HBM roofline: the algorithmic bytes
    # of all its kernels (launch counts of the timed steps) over its time
    step_bytes = manifest.headline_step_bytes(nx, ny, simple_per_step, fcg_per_step)
    step_gbs = step_bytes / (tmx / args.steps) / 1e9
    step_roofline = {"algorithmic_bytes": round(step_bytes), "achieved": round(step_gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4)}
    log(f"step roofline: {step_roofline}")

    batch = batch_c5()
    log(f"C5: {batch}")
    s3d = step_3d()
    log(f"3D step: {s3d}")
    log(f"batch: {batch}")

    net = net_throughput(nx, ny)
    log(f"net: {net}")
    cpu = None if args.no_cpu else cpu_baseline()
    steps_mx = args.steps / tmx
    steps_64 = args.steps / t64
    out = {
        "metric": "steps/s & pressure-solve ms to 1e-8 resid: learned-PC+mixed vs plain FP64",
        "value": round(steps_mx, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * tmx / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32 (mixed)",
        "data": "synthetic",
        "config": {"workload": "Rayleigh-Taylor 1024x2048, Atwood 0.5, cos amplitude 0.01 + "
                               "1e-4 seeded noise; lu_rb velocity, pressure to 1e-8",
                   "grid": [nx, ny],
                   "path": ("fcg_learned, mixed precision: FP64 flexible CG preconditioned by "
                            "the FP32 multigrid V-cycle " +
                            ("plus the accepted learned correction" if train and
                             train.get("accepted") else
                             "alone (the learned correction was trained and rejected by its "
                             "held-out acceptance solve)") +
                            "; FP32 face fields, gradients, momentum rows, velocity sweep and "
                            "advection limiter; FP64 state, residuals and corrections"),
                   "timed": f"steps {args.warmup + 1}..{args.warmup + args.steps} of RT "
                            f"started from rest at dt 2e-4 (t <= "
                            f"{2e-4 * (args.warmup + args.steps):.4f}), where one SIMPLE "
                            f"iteration converges per step; clock stopped after the last "
                            f"step's kernels finished"},
        "steps_per_s_fp64": round(steps_64, 4),
        "ms_per_step_fp64": round(1e3 * t64 / args.steps, 3),
        "pressure_iters_per_step": {"headline": [p for _, p in itmx], "fp64": [p for _, p in it64]},
        "learned_correction": train,
        "pressure_solve_ms": dict(solves, speedup_mixed_fcg_vs_fp64=round(
            solves["fp64_jacobi_pcg"]["ms"] / solves["mixed_fcg_vcycle"]["ms"], 2),
            speedup_fp32_vcycle_vs_fp64_vcycle=round(
                solves["fcg_fp64_vcycle"]["ms"] / solves["mixed_fcg_vcycle"]["ms"], 3),
            cpu=cpu_solves),
        "poisson_c4_512": c4,
        "field_error_vs_fp64": field_err,
        "step_roofline": step_roofline,
        "kernels": kernels,
        "batch_c5": batch,
        "step_3d": s3d,
        "net": net,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
