#!/usr/bin/env python3
"""Throughput of the fused RHS + 2N-RK3 step on one MI355X.

Workload (default): isothermal MHD, 512^3 periodic box, 7 FP64 variables
(uu, lnrho, aa), nu = eta = 1e-3, cs = 1, fixed dt = 1e-3, synthetic seeded
Gaussian-noise state (seed 7).  A step is one full RK3 step (3 fused kernel
launches).  The timed region runs --steps steps with the state resident in
HBM, bracketed by device synchronisation.  Reported:
  value       Mcell-updates/s of the FP64 form (interior cells x steps / s)
  roofline    the fused kernel's achieved algorithmic HBM bandwidth
              (80 B per variable per cell-update, averaged over the 3
              launches of a step) against 8 TB/s, timed with HIP events
  mixed       the same for the mixed FP32-on-offset form
  cpu_baseline the CPU oracle (C port of the reference pencil loop, OpenMP)
              on a bounded sample of the same physics on this host
Prints exactly one JSON line on stdout.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pencil-code_amd"))

import numpy as np  # noqa: E402

import pcmx  # noqa: E402
import pcmx_ics  # noqa: E402

HBM_PEAK_GBS = 8000.0
BYTES_PER_VAR = 80  # 10 FP64 state accesses per variable per RK3 step


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_form(n, nvar_mag, precision, steps, warmup, dt, kernel, f):
    par = pcmx.make_params(n, lmagnetic=nvar_mag, nu=1e-3, eta=1e-3, precision=precision, kernel=kernel)
    pcmx.initialize_gpu(f, par, nt=1)
    try:
        t = 0.0
        if warmup:
            t, _ = pcmx.time_step(warmup, t, dt, ldiag=True)
        pcmx.set_timing(True)
        t0 = time.perf_counter()
        t, _ = pcmx.time_step(steps, t, dt, ldiag=True)   # synchronises at the end
        sec = time.perf_counter() - t0
        kms, nk = pcmx.kernel_timing()
        red = pcmx.get_gpu_reduced_vars()
    finally:
        pcmx.finalize_gpu()
    ncell = float(n) ** 3
    value = ncell * steps / sec / 1e6
    nvar = par.mvar
    bytes_per_launch = BYTES_PER_VAR * nvar * ncell / 3.0
    achieved = bytes_per_launch / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    return dict(value=value, sec=sec, ms_per_step=sec * 1e3 / steps, kernel_ms=kms, nkernels=nk,
                achieved=achieved, urms=float(red[0]), brms=float(red[2]))


def copy_calibration(mib, reps):
    """Device-to-device copies of a known size (the runtime's blit kernel),
    run beside the RHS kernel under rocprofv3 --pmc so that FETCH_SIZE and
    WRITE_SIZE can be scaled to bytes: each copy reads and writes mib MiB."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed in copy calibration")
    try:
        hip.hipMemset(a, 1, ctypes.c_size_t(n))
        for _ in range(reps):
            if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) != 0:  # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed in copy calibration")
        hip.hipDeviceSynchronize()
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    log(f"bench: copy calibration {reps} x {mib} MiB device-to-device")


def cpu_baseline(threads):
    """Oracle (C port of the reference pencil loop) on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    n, steps = 256, 30
    p = O.make_params(n, lmagnetic=True, nu=1e-3, eta=1e-3, dt=1e-3)
    f0 = pcmx_ics.config_state("noise", n, 7, lmagnetic=True)
    O.advance(f0, p, 1, nthreads=threads)          # warm caches / threads
    t0 = time.perf_counter()
    O.advance(f0, p, steps, nthreads=threads)
    sec = time.perf_counter() - t0
    return dict(value=n ** 3 * steps / sec / 1e6, unit="Mcell-updates/s", cores=threads, kind="port",
                sample=f"isothermal MHD {n}^3, {steps} RK3 steps, FP64, oracle pencil loop on {threads} "
                       f"OpenMP threads ({sec:.1f} s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--kernel", type=int, default=None)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="also run 4 device-to-device copies of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "decomposition is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    n = args.n
    log(f"bench: building {n}^3 isothermal MHD state (seed 7)")
    interior = pcmx_ics.config_state("noise", n, 7, lmagnetic=True)
    f = pcmx.farray_from_interior(interior)
    del interior

    log("bench: FP64 form")
    r64 = run_form(n, True, pcmx.PREC_FP64, args.steps, args.warmup, args.dt, args.kernel, f)
    log(f"bench: FP64 {r64}")
    rmx = None
    if not args.no_mixed:
        log("bench: mixed form")
        rmx = run_form(n, True, pcmx.PREC_MIXED, args.steps, args.warmup, args.dt, args.kernel, f)
        log(f"bench: mixed {rmx}")
    del f
    if args.copy_calib:
        copy_calibration(args.copy_calib, 4)
    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"bench: CPU baseline on {threads} threads")
        cpu = cpu_baseline(threads)
        log(f"bench: cpu {cpu}")

    out = {
        "metric": "Mcell-updates/s per RK3 step, 512^3 MHD, FP64 vs mixed; % of HBM peak",
        "value": round(r64["value"], 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r64["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"isothermal MHD {n}^3 periodic, 7 FP64 vars, nu=eta=1e-3, cs=1, "
                               f"fixed dt={args.dt}, seeded Gaussian noise IC (seed 7)",
                   "grid": n, "nvar": 7, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(r64["achieved"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(r64["achieved"] / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel_ms": round(r64["kernel_ms"], 4),
                     "bytes_per_launch": BYTES_PER_VAR * 7 * n ** 3 / 3.0},
        "cpu_baseline": cpu,
    }
    if rmx:
        out["mixed"] = {"value": round(rmx["value"], 3), "ms_per_step": round(rmx["ms_per_step"], 4),
                        "speedup_vs_fp64": round(rmx["value"] / r64["value"], 4),
                        "roofline_frac": round(rmx["achieved"] / HBM_PEAK_GBS, 4),
                        "kernel_ms": round(rmx["kernel_ms"], 4), "dtype": "f32 stencils, f64 state"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
