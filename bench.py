#!/usr/bin/env python3
"""Throughput of the fused RHS + 2N-RK3 step on one MI355X.

Default workload: isothermal MHD, 512^3 periodic box of side 2pi, 7 FP64
variables (uu, lnrho, aa), nu = eta = 1e-3, cs = 1, mu0 = 1, fixed
dt = 1e-3, synthetic state: uu and aa band-limited random fields on the
shells 1 <= |k| <= 4 (rms 0.05 and 0.02), lnrho = 0, seed 7.  A step is one
full RK3 step (3 fused kernel launches).  The timed region runs --steps steps
with the state resident in HBM, bracketed by device synchronisation.
Reported on one JSON line:
  value        Mcell-updates/s of the FP64 form (interior cells x steps / s)
  roofline     the fused kernel's achieved algorithmic HBM bandwidth (80 B
               per variable per cell-update, averaged over the 3 launches of
               a step) against 8 TB/s, timed with HIP events on the library
               stream; `traffic` is the measured L2<->memory bytes per launch
               from the committed rocprofv3 counter profile of this workload
  mixed        the same for the mixed FP32-on-offset form
  cpu_baseline the CPU oracle (C port of the reference pencil loop, OpenMP)
               on a bounded sample of the same physics on this host

--config K instead runs one of the reference configurations and prints its
throughput and the per-variable errors (FP64 against the CPU oracle where the
oracle finishes in under a minute, the mixed form against the FP64 form).
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
import pcmx_forcing  # noqa: E402
import pcmx_ics  # noqa: E402

HBM_PEAK_GBS = 8000.0
BYTES_PER_VAR = 80  # 10 FP64 state accesses per variable per RK3 step
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "rhs_zsplit_traffic_mhd512.json")
METRIC = "Mcell-updates/s per RK3 step, 512^3 MHD, FP64 vs mixed; % of HBM peak"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


def rel_linf(a, b):
    out = []
    for v in range(a.shape[0]):
        den = np.abs(b[v]).max()
        out.append(float(np.abs(a[v] - b[v]).max() / den) if den > 0 else float(np.abs(a[v]).max()))
    return out


def run_form(f, par, steps, warmup, dt, keep_state=False, tune=True, graphs=False):
    """warmup + timed steps of one engine (tile shape and z-chunk length tuned
    first when the split kernel runs); returns timings and, optionally, the
    final interior state.  With graphs, the timed steps run without kernel
    timing, so that fixed steps replay captured HIP graphs, and the kernel
    time comes from a few more steps run with timing (the state is kept only
    from the timed steps)."""
    pcmx.initialize_gpu(f, par, nt=1)
    plan = None
    try:
        if tune:
            try:
                plan = pcmx.tune()
            except pcmx.PcmxError:
                plan = None  # not the split kernel (grid not tileable)
        t = 0.0
        if warmup:
            t, dt = pcmx.time_step(warmup, t, dt, ldiag=True)
        pcmx.set_timing(not graphs)
        t0 = time.perf_counter()
        t, dt = pcmx.time_step(steps, t, dt, ldiag=True)   # synchronises at the end
        sec = time.perf_counter() - t0
        red = pcmx.get_gpu_reduced_vars()
        state = None
        if keep_state:
            out = np.zeros_like(f)
            pcmx.copy_farray_from_gpu(out)
            state = pcmx.interior_of(out, par.mvar).copy()
        if graphs:
            pcmx.set_timing(True)
            pcmx.time_step(min(steps, 4), t, dt, ldiag=True)
        kms, nk = pcmx.kernel_timing()
    finally:
        pcmx.finalize_gpu()
    ncell = float(par.nx) * par.ny * par.nz
    bytes_per_launch = BYTES_PER_VAR * par.mvar * ncell / 3.0
    achieved = bytes_per_launch / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    return dict(value=ncell * steps / sec / 1e6, sec=sec, ms_per_step=sec * 1e3 / steps, kernel_ms=kms,
                nkernels=nk, achieved=achieved, urms=float(red[0]), brms=float(red[2]), t=t, dt=dt,
                state=state, plan=(f"{plan[1]}x{256 // plan[1]} tiles, {plan[2]}-plane z chunks" if plan
                                   else (None if tune else "default plan (not tuned)")))


def measured_traffic(n):
    """L2<->memory bytes per launch of the default workload, from the
    committed rocprofv3 counter profile (None if absent or another grid)."""
    try:
        with open(TRAFFIC_FILE) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None
    return d.get("traffic_bytes_per_launch") if d.get("grid") == n else None


def cpu_baseline(threads):
    """Oracle (C port of the reference pencil loop) on a bounded sample."""
    O = oracle()
    n, steps = 256, 30
    p = O.make_params(n, lmagnetic=True, nu=1e-3, eta=1e-3, dt=1e-3)
    f0 = pcmx_ics.config_state("shells", n, 7, lmagnetic=True)
    O.advance(f0, p, 1, nthreads=threads)          # warm caches / threads
    t0 = time.perf_counter()
    O.advance(f0, p, steps, nthreads=threads)
    sec = time.perf_counter() - t0
    return dict(value=n ** 3 * steps / sec / 1e6, unit="Mcell-updates/s", cores=threads, kind="port",
                sample=f"isothermal MHD {n}^3, {steps} RK3 steps, FP64, oracle pencil loop on {threads} "
                       f"OpenMP threads ({sec:.1f} s)")


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


# ---------------------------------------------------------------- configs
# The reference configurations (grid, variables, transport coefficients,
# initial state, step control, number of steps).
CONFIGS = {
    1: dict(desc="isothermal hydro 32^3, 4 vars, nu=5e-3, fixed dt=1e-3, velocity noise 1e-3, 100 steps",
            n=32, mag=False, ent=False, nu=5e-3, eta=0.0, chi=0.0, gamma=1.0, dt=1e-3, ldt=False,
            ic="noise", seed=1, steps=100, oracle_steps=100),
    2: dict(desc="isothermal MHD 128^3, 7 vars, nu=eta=5e-3, fixed dt=2e-3, band-limited uu/aa, 100 steps",
            n=128, mag=True, ent=False, nu=5e-3, eta=5e-3, chi=0.0, gamma=1.0, dt=2e-3, ldt=False,
            ic="shells", seed=2, steps=100, oracle_steps=100),
    3: dict(desc="MHD + entropy 256^3, 8 vars, gamma=5/3, nu=eta=chi=2e-3, Courant dt (cdt=0.4), "
                 "band-limited uu/aa + ss noise 1e-3, 100 steps",
            n=256, mag=True, ent=True, nu=2e-3, eta=2e-3, chi=2e-3, gamma=5 / 3, dt=0.0, ldt=True, cdt=0.4,
            ic="shells", seed=3, steps=100, oracle_steps=10),
    4: dict(desc="isothermal MHD 512^3, 7 vars, nu=eta=1e-3, fixed dt=1e-3, band-limited uu/aa (seed 7), "
                 "50 steps", n=512, mag=True, ent=False, nu=1e-3, eta=1e-3, chi=0.0, gamma=1.0, dt=1e-3,
            ldt=False, ic="shells", seed=7, steps=50, oracle_steps=0),
}


def run_config(k, threads):
    c = CONFIGS[k]
    n = c["n"]
    log(f"bench: config {k}: {c['desc']}")
    f0 = pcmx_ics.config_state(c["ic"], n, c["seed"], lmagnetic=c["mag"], lentropy=c["ent"])
    f = pcmx.farray_from_interior(f0)
    kw = dict(lmagnetic=c["mag"], lentropy=c["ent"], nu=c["nu"], eta=c["eta"], chi=c["chi"], gamma=c["gamma"],
              ldt=c["ldt"], cdt=c.get("cdt", 0.9))
    out = {"config": k, "workload": c["desc"], "unit": "Mcell-updates/s"}
    res = {}
    for name, prec in (("fp64", pcmx.PREC_FP64), ("mixed", pcmx.PREC_MIXED)):
        par = pcmx.make_params(n, precision=prec, **kw)
        # timed after two warmup steps (plan tuned, graphs captured); the
        # states compared below come from untimed runs from t = 0
        r = run_form(f, par, c["steps"], 2, c["dt"], graphs=True)
        r["state"] = run_form(f, par, c["steps"], 0, c["dt"], keep_state=True, tune=False)["state"]
        res[name] = r
        out[name] = {"value": round(r["value"], 2), "kernel_ms": round(r["kernel_ms"], 4),
                     "roofline_frac": round(r["achieved"] / HBM_PEAK_GBS, 4), "plan": r["plan"]}
        log(f"bench: config {k} {name}: {out[name]}")
    out["mixed_vs_fp64_rel_linf"] = rel_linf(res["mixed"]["state"], res["fp64"]["state"])
    if c["oracle_steps"]:
        O = oracle()
        m = c["oracle_steps"]
        op = O.make_params(n, lmagnetic=c["mag"], lentropy=c["ent"], nu=c["nu"], eta=c["eta"], chi=c["chi"],
                           gamma=c["gamma"], cdt=c.get("cdt", 0.9), dt=0.0 if c["ldt"] else c["dt"])
        t0 = time.perf_counter()
        o, to, _, _ = O.advance(f0, op, m, nthreads=threads)
        osec = time.perf_counter() - t0
        g = {}
        for name, prec in (("fp64", pcmx.PREC_FP64), ("mixed", pcmx.PREC_MIXED)):
            par = pcmx.make_params(n, precision=prec, **kw)
            g[name] = run_form(f, par, m, 0, c["dt"], keep_state=True)
        out["oracle_steps"] = m
        out["fp64_vs_oracle_rel_linf"] = rel_linf(g["fp64"]["state"], o)
        out["mixed_vs_oracle_rel_linf"] = rel_linf(g["mixed"]["state"], o)
        out["t_oracle"] = to
        out["oracle_sec"] = round(osec, 2)
        out["oracle_value"] = round(float(n) ** 3 * m / osec / 1e6, 2)
    return out


def run_config4_large(n):
    """Configuration 4 at a grid too large for host-side oracle or state
    comparisons (1024^3: 60 GB of state, 180 GB resident on the device): the
    band-limited state is written straight into one f-array, the FP64 and
    mixed forms are timed from it, and they are compared through the
    size-independent diagnostics (urms, brms after the run)."""
    c = CONFIGS[4]
    g = pcmx.NGHOST
    desc = c["desc"].replace("512^3", f"{n}^3")
    log(f"bench: config 4 at {n}^3: {desc}")
    f = np.zeros((7, n + 2 * g, n + 2 * g, n + 2 * g))
    t0 = time.perf_counter()
    pcmx_ics.bandlimited_into(f[0:3, g:-g, g:-g, g:-g], 1, 4, 0.05, c["seed"], progress=log)
    pcmx_ics.bandlimited_into(f[4:7, g:-g, g:-g, g:-g], 1, 4, 0.02, c["seed"] + 2, progress=log)
    log(f"bench: initial state built in {time.perf_counter() - t0:.1f} s")
    kw = dict(lmagnetic=True, nu=c["nu"], eta=c["eta"])
    out = {"config": 4, "grid": n, "workload": desc, "unit": "Mcell-updates/s"}
    res = {}
    for name, prec in (("fp64", pcmx.PREC_FP64), ("mixed", pcmx.PREC_MIXED)):
        r = run_form(f, pcmx.make_params(n, precision=prec, **kw), c["steps"], 0, c["dt"])
        res[name] = r
        out[name] = {"value": round(r["value"], 2), "kernel_ms": round(r["kernel_ms"], 4),
                     "roofline_frac": round(r["achieved"] / HBM_PEAK_GBS, 4), "t": r["t"], "plan": r["plan"],
                     "urms": r["urms"], "brms": r["brms"]}
        log(f"bench: config 4 {n}^3 {name}: {out[name]}")
    out["mixed_vs_fp64_rel_diag"] = {k: abs(res["mixed"][k] - res["fp64"][k]) / abs(res["fp64"][k])
                                     for k in ("urms", "brms")}
    return out


def run_io(n):
    """Snapshot output at n^3 (isothermal MHD, 7 variables): var.dat written
    from the device (pcmx_write_snapshot, one variable at a time through
    pinned chunks) against the route of the reference's GPU module, the
    whole f-array copied to the host (copy_farray_from_GPU) and written by
    the host (the same record layout, pcmx_write_record).  Both files are
    compared byte for byte and deleted."""
    import tempfile

    par = pcmx.make_params(n, lmagnetic=True, nu=1e-3, eta=1e-3)
    f = pcmx.farray_from_interior(pcmx_ics.config_state("shells", n, 7, lmagnetic=True))
    d = tempfile.mkdtemp(prefix="pcmx_io_", dir=os.environ.get("TMPDIR", "/tmp"))
    a, b = os.path.join(d, "var_device.dat"), os.path.join(d, "var_host.dat")
    pcmx.initialize_gpu(f, par, nt=1)
    try:
        t, _ = pcmx.time_step(1, 0.0, 1e-3)
        t0 = time.perf_counter()
        pcmx.write_snapshot(a, t)
        sec_dev = time.perf_counter() - t0
        t0 = time.perf_counter()
        host = np.empty_like(f)
        pcmx.copy_farray_from_gpu(host)
        sec_copy = time.perf_counter() - t0
        g = pcmx.NGHOST
        m = [np.arange(1, k + 2 * g + 1) - 3.5 for k in (n, n, n)]
        grid = [np.array([t])] + [-math.pi + mm * par.dx for mm in m] + [np.array([par.dx, par.dy, par.dz])]
        pcmx.write_record(b, [host[:par.mvar]])
        pcmx.write_record(b, grid, append=True)
        sec_host = time.perf_counter() - t0
    finally:
        pcmx.finalize_gpu()
    size = os.path.getsize(a)
    same = size == os.path.getsize(b)
    if same:
        with open(a, "rb") as fa, open(b, "rb") as fb:
            while same:
                x, y = fa.read(1 << 26), fb.read(1 << 26)
                same = x == y
                if not x:
                    break
    for p in (a, b):
        os.remove(p)
    os.rmdir(d)
    return {"io": {"grid": n, "bytes": size, "identical": same,
                   "device_snapshot_s": round(sec_dev, 3), "device_snapshot_GBps": round(size / sec_dev / 1e9, 2),
                   "copy_then_write_s": round(sec_host, 3), "of_which_copy_s": round(sec_copy, 3),
                   "host_f_array_bytes": int(f.nbytes)}}


# the reference's published CPU timings at 512^3 for this physics
# (doc/timings/notes_2021.tex:40,42: microseconds per point per step on Dardel)
PUBLISHED_512 = {"512 procs": 1.0 / 1.571e-3, "2048 procs": 1.0 / 5.508e-4}


def run_published(n, steps, warmup, tune=True):
    """The physics of the reference's published timings
    (doc/timings/N4096_32x32x32/run.in): isothermal MHD with upwinded
    advection (lupw_uu = lupw_lnrho = T), nu = eta = 2e-5, Courant dt with
    cdt = 0.9, helical forcing of the vector potential (iforce='helical',
    force = 5e-4, relhel = 1, lmagnetic_forcing = T) with a new wave vector,
    phase and polarisation every step, as forcing_hel draws them.  The
    wave vectors are those of the shell 7 <= |k| < 8 of a 2 pi box (the
    published 4096^3 run forces at |k| ~ 60, i.e. the same scale relative to
    the grid at 512^3); the draws come from numpy, not the reference's
    generator.  Initial state as the default line (band-limited uu/aa, seed
    7).  Timed: steps full RK3 steps with the host-side forcing update
    between steps."""
    kav, kvecs = pcmx_forcing.shell_vectors(7.0, 8.0)
    interior = pcmx_ics.config_state("shells", n, 7, lmagnetic=True)
    f = pcmx.farray_from_interior(interior)
    del interior
    desc = (f"isothermal MHD {n}^3, 7 FP64 vars, nu=eta=2e-5, Courant dt (cdt=0.9), lupw_uu=lupw_lnrho=T, "
            f"helical forcing of A (force=5e-4, relhel=1, {len(kvecs)} wave vectors 7<=|k|<8, new draw each step)")
    log(f"bench: published-timing physics: {desc}")
    out = {"workload": desc, "unit": "Mcell-updates/s", "grid": n,
           "published_cpu_512": {k: round(v, 1) for k, v in PUBLISHED_512.items()}}
    for name, prec in (("fp64", pcmx.PREC_FP64), ("mixed", pcmx.PREC_MIXED)):
        par = pcmx.make_params(n, lmagnetic=True, nu=2e-5, eta=2e-5, ldt=True, cdt=0.9, precision=prec,
                               lupw_uu=True, lupw_lnrho=True)
        rng = np.random.default_rng(11)
        pcmx.initialize_gpu(f, par, nt=1)
        try:
            plan = pcmx.tune() if tune else None

            def step(t, dt):
                ik, phase, phi = pcmx_forcing.draw(rng, len(kvecs))
                fp = pcmx_forcing.helical_pars(kvecs[ik], phase, phi, 5e-4, 0.0, kav, scale_sqrt_dt=True)
                pcmx.set_forcing(target=par.iax, **fp)
                return pcmx.time_step(1, t, dt, ldiag=True)

            t, dt = 0.0, 0.0
            for _ in range(warmup):
                t, dt = step(t, dt)
            pcmx.set_timing(True)
            kms = []
            t0 = time.perf_counter()
            for _ in range(steps):
                t, dt = step(t, dt)
                kms.append(pcmx.kernel_timing()[0])
            sec = time.perf_counter() - t0
            red = pcmx.get_gpu_reduced_vars()
        finally:
            pcmx.finalize_gpu()
        value = float(n) ** 3 * steps / sec / 1e6
        km = sum(kms) / len(kms)
        achieved = BYTES_PER_VAR * 7 * float(n) ** 3 / 3.0 / (km * 1e-3) / 1e9
        out[name] = {"value": round(value, 2), "ms_per_step": round(sec * 1e3 / steps, 4),
                     "kernel_ms": round(km, 4), "roofline_frac": round(achieved / HBM_PEAK_GBS, 4),
                     "dt_last": dt, "t": t, "urms": float(red[0]), "brms": float(red[2]),
                     "plan": (f"{plan[1]}x{256 // plan[1]} tiles, {plan[2]}-plane z chunks" if plan
                              else "default plan (not tuned)"),
                     "vs_published_cpu_512": {k: round(value / v, 2) for k, v in PUBLISHED_512.items()}}
        log(f"bench: published physics {name}: {out[name]}")
    return out


CONVSLAB_BC = ["s", "s", "a", "a2", "c1:cT"]


def run_convslab(n, steps, warmup, tune=True):
    """The physics of samples/conv-slab at nx = ny = n, nz = 2n: entropy
    hydro in a 1 x 1 x 2 box (xyz0 = -0.5, -0.5, -0.68) with gravz = -1,
    rotation Omega = 0.1, nu = 4e-3, K-profile conduction (hcond0 = 8e-3,
    mpoly0 = 1, mpoly1 = 3, mpoly2 = 0, widthss = 0.05, z1 = 0, z2 = 1),
    top cooling (cool = 15, wcool = 0.2, cs2cool = 0.333333), bcz = s, s, a,
    a2, c1:cT, lupw_lnrho = lupw_ss = T, Courant steps with cdt = 0.4,
    cdtv = 0.3, gamma = 1.6666666, cs0 = 0.57735 (samples/conv-slab/start.in,
    run.in), from the sample's own initial state (piecewise polytrope and
    'up-down' u_z, pcmx_ics.conv_slab_state).  Runs on the split kernel (its
    entropy variant with lnrho + ss upwinding and the layer terms), tuned
    first."""
    dims = (n, n, 2 * n)
    interior = pcmx_ics.conv_slab_state(n, 2 * n)
    f = pcmx.farray_from_interior(interior)
    del interior
    lay = pcmx.layer_defaults(1.6666666, 1.0, -1.0, "c1:cT", hcond0=8e-3, mpoly0=1.0, mpoly1=3.0, mpoly2=0.0,
                              widthss=0.05, z1=0.0, z2=1.0, cool=15.0, wcool=0.2, cs2cool=0.333333)
    desc = (f"samples/conv-slab physics {n}x{n}x{2 * n}, 5 FP64 vars: gravz=-1, Omega=0.1, nu=4e-3, K-profile "
            "(hcond0=8e-3, mpoly 1/3/0), cooling (cool=15, wcool=0.2), bcz s,s,a,a2,c1:cT, "
            "lupw_lnrho=lupw_ss=T, Courant dt (cdt=0.4, cdtv=0.3); the sample's piecewise-polytrope initial state")
    log(f"bench: conv-slab physics: {desc}")
    out = {"workload": desc, "unit": "Mcell-updates/s", "grid": list(dims)}
    for name, prec in (("fp64", pcmx.PREC_FP64), ("mixed", pcmx.PREC_MIXED)):
        par = pcmx.make_params(dims, lentropy=True, nu=4e-3, gamma=1.6666666, cs0=0.57735, ldt=True, cdt=0.4,
                               cdtv=0.3, precision=prec, L=(1.0, 1.0, 2.0), gravz=-1.0, bcz=CONVSLAB_BC,
                               omega=0.1, layers=lay, z0=-0.68, lupw_lnrho=True, lupw_ss=True)
        r = run_form(f, par, steps, warmup, 0.0, tune=tune)
        out[name] = {"value": round(r["value"], 2), "ms_per_step": round(r["ms_per_step"], 4),
                     "kernel_ms": round(r["kernel_ms"], 4), "roofline_frac": round(r["achieved"] / HBM_PEAK_GBS, 4),
                     "dt_last": r["dt"], "t": r["t"], "urms": r["urms"], "plan": r["plan"]}
        log(f"bench: conv-slab {name}: {out[name]}")
    # the oracle (the CPU restatement of the reference's pencil loop) on the
    # same physics and state recipe, a bounded 64 x 64 x 128 sample, 2 steps
    try:
        O = oracle()
        cdims = (64, 64, 128)
        clay = O.layer_defaults(1.6666666, 1.0, -1.0, "c1:cT", hcond0=8e-3, mpoly0=1.0, mpoly1=3.0, mpoly2=0.0,
                                widthss=0.05, z1=0.0, z2=1.0, cool=15.0, wcool=0.2, cs2cool=0.333333)
        op = O.make_params(cdims, lentropy=True, nu=4e-3, gamma=1.6666666, cs0=0.57735, cdt=0.4, cdtv=0.3, dt=0.0,
                           L=(1.0, 1.0, 2.0), gravz=-1.0, bcz=CONVSLAB_BC, omega=0.1, z0=-0.68, lupw_lnrho=True,
                           lupw_ss=True, layers=clay)
        cstate = pcmx_ics.conv_slab_state(64, 128)
        threads = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        O.advance(cstate, op, 2, nthreads=threads)
        csec = time.perf_counter() - t0
        out["cpu_oracle"] = {"value": round(64 * 64 * 128 * 2 / csec / 1e6, 2), "cores": threads,
                             "sample": "64x64x128, 2 Courant steps, same physics and state recipe"}
        log(f"bench: conv-slab oracle on the host: {out['cpu_oracle']}")
    except (ImportError, OSError) as exc:
        out["cpu_oracle"] = None
        log(f"bench: conv-slab oracle not available: {exc}")
    return out


ZBOUND_BC = ["s", "s", "a", "a2", "a", "a", "s"]


def run_zbound(n, steps, warmup, tune=True):
    """Isothermal MHD with non-periodic z boundaries and constant gravity:
    bcz = s, s, a, a2, a, a, s for (uu, lnrho, aa) (bc_sym_z), gravz = -0.5,
    box 2pi x 2pi x 2pi with boundary points on the first and last z plane,
    nu = eta = 1e-3, fixed dt = 1e-3; band-limited uu/aa (seed 7) tapered by
    sin^2(z/2) so that they meet the boundary conditions, and the hydrostatic
    lnrho = -z/2.  The same state run periodic in z without gravity is timed
    beside it."""
    interior = pcmx_ics.config_state("shells", n, 7, lmagnetic=True)
    z = np.linspace(0.0, 2 * math.pi, n)[:, None, None]
    env = np.sin(z / 2) ** 2
    interior[0:3] *= env
    interior[4:7] *= env
    interior[3] = -0.5 * z
    f = pcmx.farray_from_interior(interior)
    del interior, env
    desc = (f"isothermal MHD {n}^3, 7 FP64 vars, z boundaries {','.join(ZBOUND_BC)}, gravz=-0.5, nu=eta=1e-3, "
            "fixed dt=1e-3, band-limited uu/aa (seed 7) tapered by sin^2(z/2), lnrho=-z/2")
    log(f"bench: z boundaries + gravity: {desc}")
    out = {"workload": desc, "unit": "Mcell-updates/s", "grid": n}
    for name, prec, zb in (("fp64", pcmx.PREC_FP64, True), ("mixed", pcmx.PREC_MIXED, True),
                           ("fp64_periodic_same_state", pcmx.PREC_FP64, False)):
        par = pcmx.make_params(n, lmagnetic=True, nu=1e-3, eta=1e-3, precision=prec,
                               gravz=-0.5 if zb else 0.0, bcz=ZBOUND_BC if zb else None)
        r = run_form(f, par, steps, warmup, 1e-3, tune=tune)
        out[name] = {"value": round(r["value"], 2), "ms_per_step": round(r["ms_per_step"], 4),
                     "kernel_ms": round(r["kernel_ms"], 4), "roofline_frac": round(r["achieved"] / HBM_PEAK_GBS, 4),
                     "urms": r["urms"], "plan": r["plan"]}
        log(f"bench: z boundaries {name}: {out[name]}")
    return out


ENSEMBLE_DESC = ("ensemble of 64 isothermal MHD boxes, sizes uniform in {32,48,64,96}^3 (side 2pi), "
                 "eta log-uniform in [1e-3,1e-2], Pm=nu/eta log-uniform in [0.5,10], band-limited uu/aa with "
                 "distinct seeds, fixed dt per box, 200 steps; mixed form = FP64 state with FP32 RHS")


def ensemble_setup(nbox=64, seed=5):
    rng = np.random.default_rng(seed)
    sizes = rng.choice([32, 48, 64, 96], size=nbox)
    boxes, states = [], []
    for i, n in enumerate(sizes):
        eta = float(10 ** rng.uniform(-3, -2))
        nu = float(eta * 10 ** rng.uniform(math.log10(0.5), 1.0))
        dx = 2 * math.pi / n
        dt = min(2e-3, 0.1 * dx * dx / max(nu, eta))   # well inside the diffusive limit
        boxes.append(pcmx.make_box(int(n), nu, eta, dt))
        states.append(pcmx_ics.config_state("shells", int(n), 1000 + i, lmagnetic=True))
    return boxes, states


def run_ensemble(boxes, states, precision, steps, tune=True):
    par = pcmx.make_params(32, lmagnetic=True, precision=precision)
    pcmx.ens_create(par, boxes)
    try:
        for i, f in enumerate(states):
            pcmx.ens_load(i, f)
        if tune:
            pcmx.ens_tune(2)
        pcmx.ens_step(2, ldiag=True)                       # warm-up (graphs built)
        for i, f in enumerate(states):
            pcmx.ens_load(i, f)
        pcmx.ens_set_timing(True)
        t0 = time.perf_counter()
        pcmx.ens_step(steps, ldiag=True)
        sec = time.perf_counter() - t0
        info = pcmx.ens_timing()
        out = [pcmx.ens_copy(i, f.shape) for i, f in enumerate(states)]
    finally:
        pcmx.ens_destroy()
    return sec, info, out


def run_config5(threads):
    steps = 200
    log(f"bench: config 5: {ENSEMBLE_DESC}")
    boxes, states = ensemble_setup()
    ncell = sum(b.nx * b.ny * b.nz for b in boxes)
    out = {"config": 5, "workload": ENSEMBLE_DESC, "unit": "Mcell-updates/s", "cells": ncell}
    res = {}
    for name, prec in (("mixed", pcmx.PREC_MIXED), ("fp64", pcmx.PREC_FP64)):
        sec, info, st = run_ensemble(boxes, states, prec, steps)
        res[name] = st
        bytes_per_launch = BYTES_PER_VAR * 7 * ncell / 3.0
        out[name] = {"value": round(ncell * steps / sec / 1e6, 2), "ms_per_step": round(sec * 1e3 / steps, 4),
                     "substep_ms": round(info[0], 4),
                     "roofline_frac": round(bytes_per_launch / (info[0] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "plan": {"tile": f"{info[2]}x{256 // info[2]}", "lz_max": info[3], "blocks_per_cu": info[4],
                              "work_items": info[5]}}
        log(f"bench: config 5 {name}: {out[name]}")
    err = np.zeros(7)
    for a, b in zip(res["mixed"], res["fp64"]):
        for v in range(7):
            den = np.abs(b[v]).max()
            err[v] = max(err[v], np.abs(a[v] - b[v]).max() / den if den > 0 else 0.0)
    out["mixed_vs_fp64_rel_linf"] = err.tolist()
    # the same boxes one after another through the single-box engine
    m = 16
    t = 0.0
    for b, f in zip(boxes[:m], states[:m]):
        par = pcmx.make_params((b.nx, b.ny, b.nz), lmagnetic=True, nu=b.nu, eta=b.eta)
        r = run_form(pcmx.farray_from_interior(f), par, 20, 2, b.dt)
        t += r["sec"]
    cells_m = sum(b.nx * b.ny * b.nz for b in boxes[:m])
    out["one_box_at_a_time"] = {"value": round(cells_m * 20 / t / 1e6, 2),
                                "sample": f"first {m} boxes, 20 steps each, FP64 single-box engine"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512,
                    help="grid of the default bench line; with --config 4 and N > 512, configuration 4 at N^3")
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--kernel", type=int, default=None)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--io", type=int, default=0, metavar="N",
                    help="time var.dat output at N^3: from the device against copy-back plus host write")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default plan (32x8 tiles, chunk length from the grid) instead of tuning; "
                         "counter profiles use it so that every pass runs the same plan")
    ap.add_argument("--config", type=int, default=0, choices=[0, 1, 2, 3, 4, 5],
                    help="run reference configuration K instead of the default bench line")
    ap.add_argument("--published", type=int, default=0, metavar="N",
                    help="time the physics of the reference's published CPU timings (upwinding, helical "
                         "forcing of A, Courant dt) at N^3")
    ap.add_argument("--convslab", type=int, default=0, metavar="N",
                    help="time the samples/conv-slab physics at N x N x 2N instead")
    ap.add_argument("--zbound", type=int, default=0, metavar="N",
                    help="time isothermal MHD at N^3 with non-periodic z boundaries and gravity")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="also run 4 device-to-device copies of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "decomposition is a design, not a measured path", file=sys.stderr)
        sys.exit(2)
    threads = min(16, os.cpu_count() or 1)
    if args.io:
        print(json.dumps(run_io(args.io)), flush=True)
        return
    if args.convslab:
        print(json.dumps(run_convslab(args.convslab, args.steps, args.warmup, tune=not args.no_tune)), flush=True)
        return
    if args.zbound:
        print(json.dumps(run_zbound(args.zbound, args.steps, args.warmup, tune=not args.no_tune)), flush=True)
        return
    if args.published:
        print(json.dumps(run_published(args.published, args.steps, args.warmup, tune=not args.no_tune)),
              flush=True)
        return
    if args.config:
        if args.config == 5:
            res = run_config5(threads)
        elif args.config == 4 and args.n > CONFIGS[4]["n"]:
            res = run_config4_large(args.n)
        else:
            res = run_config(args.config, threads)
        print(json.dumps(res), flush=True)
        return

    n = args.n
    log(f"bench: building {n}^3 isothermal MHD state (band-limited uu/aa, seed 7)")
    interior = pcmx_ics.config_state("shells", n, 7, lmagnetic=True)
    f = pcmx.farray_from_interior(interior)
    del interior

    kw = dict(lmagnetic=True, nu=1e-3, eta=1e-3, kernel=args.kernel)
    log("bench: FP64 form")
    r64 = run_form(f, pcmx.make_params(n, precision=pcmx.PREC_FP64, **kw), args.steps, args.warmup, args.dt,
                   tune=not args.no_tune)
    log(f"bench: FP64 { {k: v for k, v in r64.items() if k != 'state'} }")
    rmx = None
    if not args.no_mixed:
        log("bench: mixed form")
        rmx = run_form(f, pcmx.make_params(n, precision=pcmx.PREC_MIXED, **kw), args.steps, args.warmup,
                       args.dt, tune=not args.no_tune)
        log(f"bench: mixed { {k: v for k, v in rmx.items() if k != 'state'} }")
    del f
    if args.copy_calib:
        copy_calibration(args.copy_calib, 4)
    cpu = None
    if not args.no_cpu:
        log(f"bench: CPU baseline on {threads} threads")
        cpu = cpu_baseline(threads)
        log(f"bench: cpu {cpu}")

    out = {
        "metric": METRIC,
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
                               f"fixed dt={args.dt}, band-limited uu/aa (1<=|k|<=4, rms 0.05/0.02), "
                               "lnrho=0, seed 7",
                   "grid": n, "nvar": 7, "parallelism": "single GPU"},
        "plan": r64["plan"],
        "roofline": {"bound": "hbm", "achieved": round(r64["achieved"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(r64["achieved"] / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(n),
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
