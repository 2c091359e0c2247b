"""Benchmark of Navier2D::update on one MI355X (one JSON line on stdout).

Workload (default): the large periodic case, Navier2D::new_periodic(1024, 1025, Ra=1e9,
Pr=1, dt=2e-4, aspect=2), temperature = conductive profile + seeded N(0, 1e-3^2) noise,
velocity zero.  A step is one Integrate::update (navier.rs:737-765); only update() is
timed, as benches/benchmark_navier.rs:34 does.  Inputs are resident on the GPU when the
timed region starts.

value      : steps/s of the GPU step over the timed region (K steps, the captured step
             replayed as a HIP graph); mixed precision by default (--precision fp64: all
             FP64), with the FP64 step timed beside it
roofline   : the kernel class with the largest summed time over the same K steps run
             with HIP events recorded on the handle's stream around every launch;
             achieved = algorithmic work per launch / average launch duration
cpu_baseline: the CPU oracle (FP64 restatement of the reference algorithm) on the host,
             a bounded sample of the same workload
statistics_and_snapshot: after the timed steps, the device statistics update
             (src/navier/statistics.rs:130-153) per call and a snapshot write / read
             (navier.rs:961-1008), with the oracle's statistics update beside them
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Navier2D steps/sec on 1 MI355X: mixed vs FP64 GPU at stated field/Nu error"
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

CONFIGS = {
    # name: (periodic, nx, ny, ra, pr, dt, aspect, noise kind, amplitude, velocity)
    "c5": (True, 1024, 1025, 1e9, 1.0, 2e-4, 2.0, "gaussian", 1e-3, None),
    "c3": (True, 256, 257, 1e7, 1.0, 2e-3, 2.0, "gaussian", 1e-3, None),
    "c1": (False, 64, 64, 1e5, 1.0, 2e-2, 1.0, "uniform", 0.05, (0.2, 1.0, 1.0)),
}
WORKLOAD_NAMES = {
    "c5": "Navier2D periodic RBC 1024x1025 Ra1e9 Pr1 aspect2 dt2e-4 (FP64)",
    "c3": "Navier2D periodic RBC 256x257 Ra1e7 Pr1 aspect2 dt2e-3 (FP64)",
    "c1": "Navier2D confined RBC 64x64 Ra1e5 Pr1 aspect1 dt2e-2 adiabatic (FP64)",
}
# timing classes of rbc_timing_read; all but 3 (the remaining timed kernels) are one kernel each
KERNEL_CLASSES = {0: "gemm_mfma_kernel", 1: "fdma_rows_lds_kernel", 2: "irfft_cols_kernel", 3: "other kernels",
                  4: "cheb_y_bwd_kernel", 5: "rfft_cols_kernel", 6: "cheb_y_fwd_kernel", 7: "cheb_y_diff_kernel",
                  8: "tdma_rows_lds_kernel", 9: "band_y_kernel", 10: "pressure_rows_kernel",
                  11: "cheb_y_conv_kernel", 12: "gemm_mfma_kernel (FP32)", 13: "cheb_y_conv3_kernel"}
MIXED_CLASS = 3


def step_roofline(stats, steps, ms_per_step):
    """Whole-step roofline: the sum over kernel classes of the time each class's algorithmic
    work needs at its peak (GEMM flops at the FP64 MFMA peak, bytes at the HBM peak), against
    the measured step time."""
    t_roof = 0.0
    for k, (n, ms, w) in stats.items():
        if k == 0:
            t_roof += w / (PEAK_FP64_TFLOPS * 1e12)
        elif k == 12:
            t_roof += w / (PEAK_FP32_TFLOPS * 1e12)
        else:
            t_roof += w / (PEAK_HBM_GBS * 1e9)
    t_roof_us = 1e6 * t_roof / steps
    t_kernels_us = 1e3 * sum(v[1] for v in stats.values()) / steps
    return {"t_roof_us": t_roof_us, "t_step_us": 1e3 * ms_per_step, "frac": t_roof_us / (1e3 * ms_per_step),
            "t_kernels_us": t_kernels_us,
            "note": "sum over kernel classes of algorithmic work / peak (FP64 or FP32 MFMA for GEMM flops, HBM for "
                    "bytes) per step, over the graph-replayed step time; t_kernels_us is the summed "
                    "event-timed kernel time per step"}


def make(cfg, lib_mod):
    from rustpde_amd import Navier2D, seeded
    periodic, nx, ny, ra, pr, dt, aspect, kind, amp, vel = CONFIGS[cfg]
    if periodic:
        n = lib_mod.Navier2D.new_periodic(nx, ny, ra, pr, dt, aspect)
    else:
        n = lib_mod.Navier2D.new(nx, ny, ra, pr, dt, aspect, True)
    pert = seeded.interior_perturbation(nx, ny, amp, 5, kind=kind, periodic=periodic)
    if vel is not None:
        n.set_velocity(*vel)
    if isinstance(n, Navier2D):
        n.temp.v = pert
        n.temp.forward()
    else:
        n.set_phys_forward("temp", pert)
    return n


def nu_run(cfg, t_end, lib_mod):
    """Time-averaged Nu of the mixed and the FP64 step over t_end time units from the same
    seeded fields (Nu sampled once per time unit, as integrate() with save_intervall = 1);
    prints one JSON line and a progress line per 10 time units.  A run whose Nu turns NaN
    (the explicit convective term past its stability limit) stops there and reports the
    time; the averages then cover the time units before it."""
    import numpy as np
    dt = CONFIGS[cfg][5]
    per = int(round(1.0 / dt))
    out = {"workload": WORKLOAD_NAMES[cfg], "time_units": t_end, "steps_per_unit": per}
    series = {}
    for prec in ("fp64", "mixed"):
        nav = make(cfg, lib_mod)
        if prec == "mixed":
            nav.set_option(nav.OPT_MIXED, 1)
        nu = []
        t0 = time.perf_counter()
        for u in range(int(t_end)):
            nav.update(per)
            v = nav.eval_nu()
            if not np.isfinite(v):
                out["%s_diverged_at" % prec] = u + 1
                break
            nu.append(v)
            if (u + 1) % 10 == 0:
                print("nu-run %s t=%d Nu=%.6f (%.0f s)" % (prec, u + 1, nu[-1], time.perf_counter() - t0),
                      file=sys.stderr, flush=True)
        series[prec] = np.array(nu)
        out["%s_nu_mean" % prec] = float(series[prec].mean()) if len(nu) else float("nan")
        out["%s_nu_final" % prec] = float(series[prec][-1]) if len(nu) else float("nan")
        del nav
    common = min(len(series["fp64"]), len(series["mixed"]))
    if common < len(series["fp64"]) or common < len(series["mixed"]):
        for prec in ("fp64", "mixed"):
            out["%s_nu_mean" % prec] = float(series[prec][:common].mean()) if common else float("nan")
    out["time_units_compared"] = common
    out["nu_mean_rel_diff"] = abs(out["mixed_nu_mean"] - out["fp64_nu_mean"]) / abs(out["fp64_nu_mean"])
    for t in (10, 20, 50, 100):  # the early (deterministic) part of the run
        if t <= common:
            a, b = series["fp64"][:t], series["mixed"][:t]
            out["nu_mean_rel_diff_first_%d" % t] = float(abs(b.mean() - a.mean()) / abs(a.mean()))
    print(json.dumps(out), flush=True)


def _cpu_info():
    """CPU model, logical CPUs and the oracle's compiler and flags (oracle/Makefile)."""
    import platform
    import re
    import subprocess
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    flags = "unknown"
    try:
        with open(os.path.join(ROOT, "oracle", "Makefile")) as f:
            m = re.search(r"^CXXFLAGS \?= (.*)$", f.read(), re.M)
            flags = m.group(1) if m else flags
    except OSError:
        pass
    try:
        cxx = subprocess.run(["g++", "-dumpfullversion"], capture_output=True, text=True, timeout=10).stdout.strip()
    except (OSError, subprocess.SubprocessError):
        cxx = "unknown"
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "compiler": "g++ " + cxx, "flags": flags}


def _cpu_rate(cfg, threads, budget_s):
    """Oracle steps/s at `threads` threads: one warm-up step, then samples of k steps (k such
    that a sample takes >= 50 ms) until the budget is spent (at least 3 samples); the median
    sample rate."""
    from oracle import oracle as O
    O.set_threads(threads)

    class _M:
        Navier2D = O.Navier2D

    o = make(cfg, _M)
    t0 = time.perf_counter()
    o.update(1)
    one = time.perf_counter() - t0
    k = max(1, int(0.05 / max(one, 1e-9)) + 1)
    rates, t_start = [], time.perf_counter()
    while len(rates) < 3 or (time.perf_counter() - t_start < budget_s and len(rates) < 10):
        t0 = time.perf_counter()
        o.update(k)
        rates.append(k / (time.perf_counter() - t0))
    rates.sort()
    return rates[len(rates) // 2], len(rates), k, o


def cpu_baseline(cfg, budget_s):
    """The oracle (the FP64 restatement of the reference algorithm) on the host: at all the
    host's CPUs available to this process (OMP_NUM_THREADS when set, the box's CPU share) and
    at one thread, OPENBLAS_NUM_THREADS=1 exported (README.md:13-16 of the reference; the
    restatement links no BLAS)."""
    from oracle import oracle as O
    if not os.path.exists(O._LIB_PATH):
        O.build()
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    cpus = os.cpu_count() or 1
    env_threads = os.environ.get("OMP_NUM_THREADS", "")
    threads = max(1, min(cpus, int(env_threads))) if env_threads.isdigit() else cpus
    rate, nsamp, k, o = _cpu_rate(cfg, threads, budget_s)
    # one statistics update of the oracle at the same state (statistics.rs:130-153)
    o.stats_enable(1.0, 10.0)
    t1 = time.perf_counter()
    o.stats_update()
    stats_ms = 1e3 * (time.perf_counter() - t1)
    rate1, nsamp1, k1, _ = _cpu_rate(cfg, 1, budget_s)
    info = _cpu_info()
    info.update({"OPENBLAS_NUM_THREADS": "1", "OMP_NUM_THREADS": env_threads or None})
    return {"value": rate, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": "median of %d samples of %d steps after 1 warm-up step (the oracle, oracle/rbc_oracle.cpp: the "
                      "reference algorithm restated in C++, OpenMP over lanes; not the Rust binary)" % (nsamp, k),
            "single_thread": {"value": rate1, "unit": "steps/s", "cores": 1,
                              "sample": "median of %d samples of %d steps" % (nsamp1, k1)},
            "host": info}, stats_ms


def solid_step(cfg, lib_mod, steps=20):
    """The step with a solid obstacle (navier.rs:551-608; roughness elements of
    solid_masks.rs:61-101 on both plates) on the same workload: the FP64 step (batched for a
    periodic handle, field by field for a confined one) with the penalisation in its
    convective terms."""
    from rustpde_amd.checkpoint import coords
    from rustpde_amd.solid_masks import solid_roughness_sinusoid
    nav = make(cfg, lib_mod)
    x = coords(nav.nx, nav.periodic, nav.scale[0])
    nav.solid = solid_roughness_sinusoid(x, coords(nav.ny, False, 1.0), 0.1, 3.0)
    nav.update(3)
    nav.synchronize()
    t0 = time.perf_counter()
    nav.update(steps)
    nav.synchronize()
    el = time.perf_counter() - t0
    ok = not nav.exit()
    nav.close()
    return {"ms_per_step": 1e3 * el / steps, "steps": steps, "finite": ok,
            "mask": "solid_roughness_sinusoid(height 0.1, wavenumber 3)",
            "note": "FP64 step with the penalisation (batched when periodic; the FP32 convective pipeline has no "
                    "penalised form)"}


def stats_and_snapshot(nav, updates=20):
    """After the timed steps: the device statistics update (statistics.rs:130-153, as
    callback calls it) per call, and a snapshot write and read (navier.rs:961-1008) through
    host memory to a temporary file."""
    import tempfile

    from rustpde_amd import checkpoint
    st = nav.init_statistics(1.0, 10.0)
    st.update()
    nav.synchronize()
    t0 = time.perf_counter()
    for _ in range(updates):
        st.update()
    nav.synchronize()
    upd_ms = 1e3 * (time.perf_counter() - t0) / updates
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "snap" + checkpoint._SUFFIX)
        t0 = time.perf_counter()
        checkpoint.write_navier(nav, f)
        w_ms = 1e3 * (time.perf_counter() - t0)
        size = os.path.getsize(f)
        t0 = time.perf_counter()
        checkpoint.read_navier(nav, f)
        nav.synchronize()
        r_ms = 1e3 * (time.perf_counter() - t0)
    return {"stats_update_ms": upd_ms, "stats_updates_timed": updates, "snapshot_write_ms": w_ms,
            "snapshot_read_ms": r_ms, "snapshot_bytes": size, "snapshot_format": checkpoint._SUFFIX,
            "note": "statistics update: wall time per call with the queue drained at the end (FP64 on the "
                    "device); snapshot: device -> host -> file (HDF5 through libhdf5, or .npz where it is "
                    "absent) in a temporary directory and back"}


def bench_hholtz_batch(args):
    """Helmholtz batch: Hholtz on cheb_dirichlet(512)^2, c = 1e-3, 256 right-hand sides."""
    import ctypes

    import numpy as np

    from rustpde_amd import HHOLTZ, Solver, _lib, seeded
    from rustpde_amd.funspace import _Dev
    n, nb = 512, 256
    L = _lib.load()
    g = Solver(HHOLTZ, (1, n), (1, n), (1e-3, 1e-3))
    k = np.arange(n, dtype=float)
    w = (1.0 + k[:, None] + k[None, :]) ** -2
    f = np.stack([seeded.gaussian(1 + b, (n, n), 1.0) * w for b in range(nb)])
    din, dout = _Dev(f.nbytes), _Dev(8 * nb * (n - 2) ** 2)
    _lib.check(L.rbc_memcpy_h2d(din.p, f.ctypes.data, f.nbytes))
    for _ in range(max(1, args.warmup // 5)):
        g.solve_device(din.p, dout.p, nb)
    reps = max(1, args.steps // 50)
    t0 = time.perf_counter()
    for _ in range(reps):
        g.solve_device(din.p, dout.p, nb)  # synchronous
    el = (time.perf_counter() - t0) / reps
    # per right-hand side: fwd P_x [n-2, n] x [n, n] and bwd [n-2, n-2] x [n-2, n-2] GEMMs
    flops = nb * (2.0 * (n - 2) * n * n + 2.0 * (n - 2) * (n - 2) * (n - 2))
    out = {"metric": "Hholtz 512^2 right-hand sides solved per second (FP64)", "value": nb / el, "unit": "rhs/s",
           "n_gpus": 1, "steps": reps, "warmup": max(1, args.warmup // 5), "ms_per_step": 1e3 * el,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "Hholtz cheb_dirichlet 512x512 c=1e-3, 256 rhs per step (FP64)", "batch": nb},
           "roofline": {"bound": "mfma", "achieved": flops / el / 1e12, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                        "frac": flops / el / 1e12 / PEAK_FP64_TFLOPS, "traffic": None,
                        "kernel": "whole solve (eigen-GEMMs + rows + banded), GEMM flops only"}}
    # the mixed form: FP32 eigen-transforms with one FP64-residual refinement (two chained
    # solves), and the residual of both forms' solutions (first 8 right-hand sides)
    out["fp64_residual"] = g.residual(f[:8], _download(L, dout, (8, n - 2, n - 2)))
    g.set_mixed(2)
    g.solve_device(din.p, dout.p, nb)
    t0 = time.perf_counter()
    for _ in range(reps):
        g.solve_device(din.p, dout.p, nb)
    el32 = (time.perf_counter() - t0) / reps
    out["mixed_refined"] = {"rhs_per_s": nb / el32, "ms_per_batch": 1e3 * el32, "speedup_vs_fp64": el / el32,
                            "residual": g.residual(f[:8], _download(L, dout, (8, n - 2, n - 2))),
                            "note": "fwd / bwd eigen-transforms as FP32 matrix-core GEMMs, FP64 row solves, one "
                                    "refinement step against the FP64 residual of the tensor equation (2 solves)"}
    g.set_mixed(0)
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        threads = max(1, min(16, os.cpu_count() or 1))
        O.set_threads(threads)
        o = O.Solver("hholtz", O.CD, n, O.CD, n, [1e-3, 1e-3])
        o.solve(f[0])
        cnt = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < args.cpu_budget and cnt < nb:
            o.solve(f[cnt])
            cnt += 1
        out["cpu_baseline"] = {"value": cnt / (time.perf_counter() - t0), "unit": "rhs/s", "cores": threads,
                               "kind": "port", "sample": "%d right-hand sides of the same batch" % cnt}
    print(json.dumps(out), flush=True)


def _download(L, dev, shape):
    import numpy as np

    from rustpde_amd import _lib
    out = np.zeros(shape)
    _lib.check(L.rbc_memcpy_d2h(out.ctypes.data, dev.p, out.nbytes))
    return out


def c4_descs():
    """96 confined instances: N in {64, 96, 128, 192}, Ra log-uniform in [1e4, 1e7], Pr = 1,
    aspect 1, adiabatic sidewalls, dt = 0.01 * 64 / N (seeded draws)."""
    import numpy as np
    rng = np.random.default_rng(2024)
    sizes = rng.choice([64, 96, 128, 192], size=96)
    ras = 10.0 ** rng.uniform(4.0, 7.0, size=96)
    return [(int(nn), int(nn), float(ra), 1.0, 0.01 * 64.0 / nn, 1.0, True) for nn, ra in zip(sizes, ras)]


def make_ensemble(descs, mixed):
    from rustpde_amd import Ensemble, Navier2D, seeded
    ens = Ensemble(descs)
    for i, d in enumerate(descs):
        nav = ens[i]
        nav.set_velocity(0.2, 1.0, 1.0)
        nav.temp.v = seeded.interior_perturbation(d[0], d[1], 1e-2, 1000 + i, kind="uniform", periodic=False)
        nav.temp.forward()
    if mixed:
        ens.set_option(Navier2D.OPT_MIXED, 1)
    return ens


def timed_ensemble(ens, args):
    ens.update(args.warmup)
    t0 = time.perf_counter()
    ens.update(args.steps)
    return time.perf_counter() - t0


def ensemble_mixed_errors(descs):
    """Largest relative max-norm difference per field over the instances, mixed against FP64
    ensembles after 10 steps from the same seeded fields."""
    import numpy as np
    a = make_ensemble(descs, False)
    b = make_ensemble(descs, True)
    a.update(10)
    b.update(10)
    err = {"temp": 0.0, "ux": 0.0, "uy": 0.0, "pres": 0.0, "nu": 0.0}
    for i in range(len(descs)):
        for name in ("temp", "ux", "uy"):
            x, y = getattr(a[i], name).vhat, getattr(b[i], name).vhat
            err[name] = max(err[name], float(np.abs(x - y).max() / np.abs(x).max()))
        x, y = a[i].pres[0].vhat, b[i].pres[0].vhat
        err["pres"] = max(err["pres"], float(np.abs(x - y).max() / np.abs(x).max()))
        nua, nub = a[i].eval_nu(), b[i].eval_nu()
        err["nu"] = max(err["nu"], abs(nua - nub) / abs(nua))
    a.close()
    b.close()
    return err


def bench_ensemble(args):
    """Ensemble of 96 confined instances advanced together (rbc_create_ensemble); with
    --precision mixed the convective terms' dense transforms run as FP32 matrix-core GEMMs
    (the FP64 ensemble is timed beside it)."""
    descs = c4_descs()
    mixed = args.precision == "mixed"
    extra = {}
    if mixed:
        ens64 = make_ensemble(descs, False)
        el64 = timed_ensemble(ens64, args)
        ens64.close()
        extra["fp64_instance_steps_per_s"] = len(descs) * args.steps / el64
    ens = make_ensemble(descs, mixed)
    el = timed_ensemble(ens, args)
    if mixed:
        extra["mixed_speedup_vs_fp64"] = el64 / el
    n = len(descs)
    points = sum(d[0] * d[1] for d in descs)
    # per-launch timing of every group over the same number of steps (eager launches)
    ens.timing(True)
    t1 = time.perf_counter()
    ens.update(args.steps)
    el_timed = time.perf_counter() - t1
    stats = {k: ens.timing_read(k) for k in KERNEL_CLASSES}
    ens.timing(False)
    groups = ens.groups
    ens.close()
    if mixed:
        extra["mixed_error_vs_fp64_after_10_steps"] = ensemble_mixed_errors(descs)
    sroof = step_roofline(stats, args.steps, 1e3 * el / args.steps)
    sroof["note"] += "; summed over the groups, which run concurrently on three streams"
    dom = max((k for k in stats if k != MIXED_CLASS), key=lambda k: stats[k][1])
    ln, lms, lw = stats[dom]
    if dom in (0, 12):
        peak = PEAK_FP64_TFLOPS if dom == 0 else PEAK_FP32_TFLOPS
        achieved = lw / (lms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None}
    else:
        achieved = lw / (lms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": None}
    roof.update({"kernel": KERNEL_CLASSES[dom], "launches": ln, "avg_launch_us": 1e3 * lms / max(ln, 1),
                 "work_per_launch": lw / max(ln, 1), "share_of_kernel_time": lms / max(1e-30, sum(v[1] for v in stats.values())),
                 "note": "summed over the groups, which run concurrently on three streams"})
    prec = "mixed FP32/FP64" if mixed else "FP64"
    out = {"metric": "ensemble instance-steps/s (96 confined RBC instances, %s)" % prec, "value": n * args.steps / el,
           "unit": "instance-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64 mixed" if mixed else "f64", "data": "synthetic",
           "config": {"workload": "ensemble of 96 confined RBC instances, N in {64,96,128,192}, Ra log-uniform "
                                  "[1e4,1e7], dt = 0.01*64/N (%s)" % prec, "instances": n, "groups": groups,
                      "precision": "mixed" if mixed else "fp64",
                      "sizes": {str(k): sum(1 for d in descs if d[0] == k) for k in (64, 96, 128, 192)}},
           "grid_point_updates_per_s": points * args.steps / el, "roofline": roof, "step_roofline": sroof,
           "ms_per_step_with_kernel_events": 1e3 * el_timed / args.steps,
           "kernel_classes": {KERNEL_CLASSES[k]: {"launches": v[0], "ms": v[1]} for k, v in stats.items()}}
    out.update(extra)
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        threads = max(1, min(16, os.cpu_count() or 1))
        O.set_threads(threads)
        per = {}
        for nn in (64, 96, 128, 192):
            o = O.Navier2D.new(nn, nn, 1e5, 1.0, 0.01 * 64.0 / nn, 1.0, True)
            o.update(1)
            k = 3
            t0 = time.perf_counter()
            o.update(k)
            per[nn] = (time.perf_counter() - t0) / k
        cpu_step = sum(per[d[0]] for d in descs)  # seconds for one step of all 96 instances
        out["cpu_baseline"] = {"value": n / cpu_step, "unit": "instance-steps/s", "cores": threads, "kind": "port",
                               "sample": "3 oracle steps per size (64, 96, 128, 192), weighted by the ensemble's "
                                         "size counts"}
    print(json.dumps(out), flush=True)
    ens.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c5", choices=sorted(CONFIGS) + ["c2", "c4"])
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"],
                    help="mixed: convective terms in FP32, the rest FP64 (periodic configs and the c4 ensemble); "
                         "the FP64 step is "
                         "timed beside it and the field / Nu differences after 10 steps are reported")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU-baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the mixed-error, statistics / snapshot and solid-obstacle legs (profiler runs of "
                         "the step alone)")
    ap.add_argument("--nu-time", type=float, default=0.0, metavar="T",
                    help="instead of timing: the mixed and the FP64 step over T time units, Nu averaged "
                         "(sampled once per time unit)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB device to device 3 times: the known-size copy that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters of a profiler run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.config == "c2":
        bench_hholtz_batch(args)
        return
    if args.config == "c4":
        bench_ensemble(args)
        return
    if args.nu_time > 0:
        import rustpde_amd
        nu_run(args.config, args.nu_time, rustpde_amd)
        return

    import ctypes

    import rustpde_amd
    from rustpde_amd import _lib
    L = _lib.load()
    periodic, nx, ny = CONFIGS[args.config][:3]
    mixed = args.precision == "mixed" and periodic  # the mixed step is the periodic batched step's

    def timed_steps(nav):
        nav.update(args.warmup)
        nav.synchronize()
        t0 = time.perf_counter()  # K production steps (the captured step replayed as a HIP graph)
        nav.update(args.steps)
        nav.synchronize()
        el = time.perf_counter() - t0
        if nav.exit():
            print("bench.py: divergence is NaN after the timed steps", file=sys.stderr)
            sys.exit(3)
        return el

    nav = make(args.config, rustpde_amd)
    if mixed:
        nav.set_option(nav.OPT_MIXED, 1)
    elapsed = timed_steps(nav)
    extra = {"per_step_calls": per_step_calls(nav, args.steps)}
    if mixed:
        nav64 = make(args.config, rustpde_amd)
        el64 = timed_steps(nav64)
        del nav64
        extra["fp64_steps_per_s"] = args.steps / el64
        extra["mixed_speedup_vs_fp64"] = el64 / elapsed
        if not args.no_extras:
            extra["mixed_error_vs_fp64"] = mixed_errors(args.config, rustpde_amd)

    # the same K steps launched one kernel at a time with HIP events recorded on the
    # handle's stream around every launch: per-kernel-class durations for the roofline
    _lib.check(L.rbc_timing_enable(nav._h, 1))
    t0 = time.perf_counter()
    nav.update(args.steps)
    nav.synchronize()
    elapsed_events = time.perf_counter() - t0
    stats = {}
    for k in KERNEL_CLASSES:
        n, ms, w = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double()
        _lib.check(L.rbc_timing_read(nav._h, k, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(w)))
        stats[k] = (n.value, ms.value, w.value)
    _lib.check(L.rbc_timing_enable(nav._h, 0))

    ms_per_step = 1e3 * elapsed / args.steps
    dom = max((k for k in stats if k != MIXED_CLASS), key=lambda k: stats[k][1])
    n, ms, w = stats[dom]
    avg_ms = ms / max(n, 1)
    if dom in (0, 12):
        peak = PEAK_FP64_TFLOPS if dom == 0 else PEAK_FP32_TFLOPS
        achieved = (w / n) / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None}
    else:
        achieved = (w / n) / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": None}
    roof.update({"kernel": KERNEL_CLASSES[dom], "launches": n, "avg_launch_us": 1e3 * avg_ms,
                 "work_per_launch": w / n, "share_of_step": ms / (1e3 * elapsed_events)})

    value = args.steps / elapsed
    out = {
        "metric": METRIC, "value": value, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64 mixed" if mixed else "f64", "data": "synthetic",
        "config": {"workload": WORKLOAD_NAMES[args.config].replace("(FP64)", "(mixed FP32/FP64)" if mixed else "(FP64)"),
                   "nx": nx, "ny": ny, "precision": "mixed" if mixed else "fp64", "parallelism": "single-gpu"},
        "grid_point_updates_per_s": value * nx * ny,
        "ms_per_step_with_kernel_events": 1e3 * elapsed_events / args.steps,
        "kernel_classes": {KERNEL_CLASSES[k]: {"launches": v[0], "ms": v[1]} for k, v in stats.items()},
        "roofline": roof,
        "step_roofline": step_roofline(stats, args.steps, ms_per_step),
    }
    out.update(extra)
    traffic = committed_traffic(args.config, roof["kernel"])
    if traffic is not None:
        roof["traffic"] = traffic["bytes_per_launch"]
        roof["traffic_source"] = traffic["source"]
    if not args.no_extras:
        out["statistics_and_snapshot"] = stats_and_snapshot(nav)
        out["solid_step"] = solid_step(args.config, rustpde_amd)
    if not args.no_cpu_baseline:
        out["cpu_baseline"], stats_cpu_ms = cpu_baseline(args.config, args.cpu_budget)
        if not args.no_extras:
            out["statistics_and_snapshot"]["stats_update_cpu_ms"] = stats_cpu_ms
    if args.copy_probe > 0:
        copy_probe(L, args.copy_probe)
    print(json.dumps(out), flush=True)


def per_step_calls(nav, steps):
    """Steps/s of `steps` calls of update() with one step each, as integrate() and the
    reference's criterion bench (benches/benchmark_navier.rs:34) make them: from the second
    call on each call replays the captured step."""
    nav.update(1)
    nav.update(1)
    nav.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        nav.update(1)
    nav.synchronize()
    el = time.perf_counter() - t0
    return {"steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps, "calls": steps}


def mixed_errors(cfg, lib_mod, develop=1000):
    """Relative max-norm differences of the mixed step from the FP64 step (spectral
    coefficients, per field) and of Nu, from a moving state: the seeded fields with
    set_velocity(0.05, 2, 1) (a convecting flow, so the FP32 convective terms are live),
    advanced `develop` FP64 steps; then 10 and 100 steps of each precision from that state."""
    import numpy as np
    src = make(cfg, lib_mod)
    if CONFIGS[cfg][9] is None:
        src.set_velocity(0.05, 2.0, 1.0)
    src.update(develop)
    state = {k: getattr(src, k).vhat for k in ("temp", "ux", "uy")}
    state["pres"] = src.pres[0].vhat
    t_state = src.get_time()
    del src

    def start(mixed):
        nav = make(cfg, lib_mod)
        for k in ("temp", "ux", "uy"):
            getattr(nav, k).vhat = state[k]
        nav.pres[0].vhat = state["pres"]
        nav.time = t_state
        if mixed:
            nav.set_option(nav.OPT_MIXED, 1)
        return nav

    a, b = start(False), start(True)
    out = {"start": "seeded fields + set_velocity(0.05, 2, 1) after %d FP64 steps (t = %.3f)" % (develop, t_state)}
    done = 0
    for n in (10, 100):
        a.update(n - done)
        b.update(n - done)
        done = n
        err = {}
        for name in ("temp", "ux", "uy"):
            x, y = getattr(a, name).vhat, getattr(b, name).vhat
            err[name] = float(np.abs(x - y).max() / np.abs(x).max())
        x, y = a.pres[0].vhat, b.pres[0].vhat
        err["pres"] = float(np.abs(x - y).max() / np.abs(x).max())
        nua, nub = a.eval_nu(), b.eval_nu()
        err["nu"] = abs(nua - nub) / abs(nua)
        out["after_%d_steps" % n] = err
    return out


def committed_traffic(cfg, kernel):
    """HBM bytes per launch of `kernel` on workload `cfg` from the committed counter summary
    (profiles/hbm_traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, calibrated on a copy)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    entry = table.get(cfg, {}).get(kernel)
    if not entry:
        return None
    return {"bytes_per_launch": entry["bytes_per_launch"], "source": "profiles/hbm_traffic.json"}


def copy_probe(L, mib):
    """mib MiB copied device to device three times (hipMemcpy: one copy kernel per call)."""
    import ctypes

    from rustpde_amd import _lib
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.check(L.rbc_dev_malloc(ctypes.byref(a), n))
    _lib.check(L.rbc_dev_malloc(ctypes.byref(b), n))
    for _ in range(3):
        _lib.check(L.rbc_memcpy_d2d(b, a, n))
    _lib.check(L.rbc_dev_free(a))
    _lib.check(L.rbc_dev_free(b))


if __name__ == "__main__":
    main()
