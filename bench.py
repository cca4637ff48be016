#!/usr/bin/env python3
"""Benchmark: batched mixed-precision Parareal on one MI355X, 384^3 grid with 64 time slices.

Workload (the largest configuration; one GPU holds it): 3-D advection-diffusion on the
periodic unit cube, 384^3 grid, initial field = 8 Gaussians (SplitMix64 seed 0), c uniform
in [-1, 1]^3 (seed 1), nu = 1e-3, T = 2, 64 time slices batched on one GPU.  F: FP64 RK4,
dt = 2e-5 (100000 steps rounded up to 100032, 1563 per slice).  G: RK4 with BF16 storage
and FP32 arithmetic, coarse step 2/1024 split into 2 sub-steps (the whole step is unstable
on this grid), difference-form correction in FP64.  Stop once the relative L2 increment of
the slice states is <= 1e-6 (at most 8 iterations).

One "step" is one batched fused RK4 fine step over all 64 slices (64 x 384^3 lattice
updates in FP64, input and output resident in HBM): the loop that makes up ~97 % of every
Parareal iteration.  `value` = lattice updates per second of the timed steps; the roofline
is that kernel's, timed with HIP events on its launch stream, with the fraction of the
measured copy bandwidth beside the fraction of HBM peak.

Timed once beside the steps (not part of `value`):
  * the whole Parareal solve stopped by its 1e-6 increment test, with the fine sweep and
    the coarse chain of every iteration timed by HIP events;
  * the serial FP64 fine solve on the same GPU (also the error reference), and the
    Parareal speedup over it;
  * plain all-FP64 Parareal in the reference's execution structure (reference-order sweeps,
    slice after slice, FP64 G, host launches), extrapolated from measured step times;
  * the CPU baseline: the oracle on the host cores (child process, threads bound to
    physical cores), a bounded sample of the serial fine sweep on the same grid, and the
    32^3 / 4-slice CPU configuration's Parareal in full.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--no-solve] [--no-cpu]
                    [--workload cfg4|cfg3|cfg2|cfg5|ref] [--roofline-only]
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

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "Parareal wall-time vs serial-FP64-fine at ≤1e-6 rel L2; stencil GLUP/s + %HBM peak"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------------------
# pieces shared by the workloads

def conf_of(w, **kw):
    from stencilparareal_amd.config import RuntimeConfiguration
    d = dict(gridSize=w.N, endTime=w.T, timeStepsFine=w.stepsF, timeStepsCoarse=w.stepsG, nu0=w.nu,
             cx=w.c[0], cy=w.c[1], cz=w.c[2], kmax=w.kmax, async_=False, timeSlices=w.slices)
    d.update(kw)
    return RuntimeConfiguration(**d)


class Batch2:
    """Two device arrays of S slices (ping-pong) and the batched fused fine step over them."""

    def __init__(self, conv, init, S, dtype="f64"):
        import torch
        self.conv, self.S = conv, S
        self.lay = init.layout if dtype == "f64" else conv.field("layout_only", dtype=dtype).layout
        self.n = init.layout.ni * init.layout.nj * init.layout.nk
        tdt = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}[dtype]
        self.a = torch.empty(S * self.n, dtype=tdt, device="cuda")
        self.b = torch.empty_like(self.a)
        src = init.interior().reshape(-1).to(tdt)
        for s in range(S):
            self.a[s * self.n:(s + 1) * self.n].copy_(src)
        conv.ctx.sync_stream()

    def step(self, src, dst, dt):
        import ctypes as C
        from stencilparareal_amd import _lib as L
        L.call("sp_do_rk4_batched", self.conv.h, self.S, C.byref(self.lay), C.c_void_p(src.data_ptr()),
               C.byref(self.lay), C.c_void_p(dst.data_ptr()), self.n, 0.0, dt, 1)

    def run(self, k, dt, events=None):
        """k steps, each one launch of the fused kernel, alternating a -> b -> a."""
        import torch
        if events:
            events[0].record()
        for i in range(k):
            src, dst = (self.a, self.b) if i % 2 == 0 else (self.b, self.a)
            self.step(src, dst, dt)
        if events:
            events[1].record()
        return torch

    def copy_gbs(self, reps=10):
        """Bandwidth of a plain device copy of the same bytes (read S*n elements, write as many)."""
        import torch
        self.b.copy_(self.a)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            self.b.copy_(self.a)
        e1.record()
        e1.synchronize()
        return 2.0 * self.a.element_size() * self.S * self.n * reps / (e0.elapsed_time(e1) / 1e3) / 1e9

    def free(self):
        import torch
        del self.a, self.b
        torch.cuda.empty_cache()


def timed_steps(bt, a, dt):
    """The timed region: W untimed steps, then K steps between device synchronisations;
    wall clock for ms_per_step, HIP events on the launch stream for the kernel average."""
    import torch
    bt.run(a.warmup, dt)
    torch.cuda.synchronize()
    ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
    t0 = time.perf_counter()
    bt.run(a.steps, dt, ev)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall / a.steps * 1e3, ev[0].elapsed_time(ev[1]) / a.steps


def committed_traffic(name):
    """HBM bytes per launch of the dominant kernel from the committed PMC summary
    profiles/<name> (FETCH_SIZE and WRITE_SIZE runs scaled by a copy kernel); None if absent."""
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def roofline(k_ms, alg_bytes, copy_gbs, traffic, kernel):
    achieved = alg_bytes / (k_ms / 1e3) / 1e9
    actual = traffic / (k_ms / 1e3) / 1e9 if traffic else None
    # what bounds the kernel: HBM once the counted traffic streams near the copy rate,
    # otherwise the on-chip work (LDS and FP64 VALU issue of the fused stages)
    limiter = None
    if actual:
        limiter = "hbm" if actual >= 0.85 * copy_gbs else "on-chip (LDS + FP64 VALU issue, latency)"
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "frac_of_copy": round(achieved / copy_gbs, 4), "copy_gbs": round(copy_gbs, 1),
            "traffic_rate_gbs": round(actual, 1) if actual else None, "limiter": limiter,
            "kernel": kernel, "avg_launch_ms": round(k_ms, 4), "bytes_per_launch": alg_bytes}


def cpu_baseline(N, nu, c, dt, gauss_cfg, budget=12.0):
    """The oracle's serial fine RK4 (the reference's OpenMP structure) on the host cores, a
    bounded sample on the benchmark grid, plus the CPU configuration's Parareal in full; run
    as a child process with OMP_NUM_THREADS = physical cores, OMP_PROC_BIND=close."""
    from oracle.cpu_bench import child_env
    env, phys, logical = child_env()
    cmd = [sys.executable, "-m", "oracle.cpu_bench", "--grid", str(N), "--nu", repr(nu), "--c",
           *[repr(x) for x in c], "--dt", repr(dt), "--gauss-cfg", str(gauss_cfg), "--budget", str(budget)]
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        log("cpu baseline failed: " + r.stderr[-500:])
        return None
    d = json.loads(r.stdout.strip().splitlines()[-1])
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    sf = d["serial_fine"]
    return {"value": sf["glups"], "unit": "GLUP/s", "cores": sf["threads"], "kind": "port",
            "sample": f"serial fine RK4 (oracle, reference OpenMP structure), {N}^3, {sf['steps']} steps, "
                      f"{sf['seconds']} s, OMP_PROC_BIND=close on {phys} physical cores",
            "one_thread": d["one_thread"], "parareal_cfg1": d.get("parareal_cfg1"),
            "physical_cores": phys, "logical_cpus": logical, "cpu_model": model}


def plain_step_ms(w, dt, steps=10, gdt=None):
    """One slice's reference-structure RK4 step (4 RHS + 3 Euler + 1 RK sweep, host-launched),
    timed with HIP events."""
    import torch
    import stencilparareal_amd as sp
    conv = sp.Convection(w.N, w.N, w.N, 1.0 / w.N, w.nu, 0.0, *w.c, path="staged")
    a, b = conv.field("pa"), conv.field("pb")
    g = w.gauss
    sp.fill_gaussians(a, g.xc, g.yc, g.zc, g.sigma, g.amp)
    conv.DoRK4(a, b, 0.0, dt, 2)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    conv.DoRK4(a, b, 0.0, dt, steps)
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / steps


# ---------------------------------------------------------------------------------------
# the default workload: 384^3, 64 slices, BF16-storage G, stop at 1e-6

def cfg4(a):
    import torch
    import stencilparareal_amd as sp
    from stencilparareal_amd.config import workload
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    w = workload(4)
    N, S = w.N, w.slices
    conf = conf_of(w)
    tsF, dtF = conf.timeStepsFinePerTimeSlice(), conf.dtFine()
    conv = sp.Convection(N, N, N, conf.dx(), conf.nu0, 0.0, conf.cx, conf.cy, conf.cz)
    init = conv.field("qinitial")
    g = w.gauss
    sp.fill_gaussians(init, g.xc, g.yc, g.zc, g.sigma, g.amp)
    log("autotune the fused tile for 64 slices")
    conv.autotune(S)
    tile = conv.get_tile(S, True)
    bt = Batch2(conv, init, S)
    log(f"warmup {a.warmup}, timed {a.steps} batched steps")
    ms, k_ms = timed_steps(bt, a, dtF)
    copy = bt.copy_gbs()
    bt.free()
    lup = S * N ** 3
    kernel = (f"k_rk4_fused_{tile['variant']} {tile['tx']}x{tile['ty']} tiles, {tile['nthreads']} threads "
              f"(FP64, {S} slices of {N}^3, kchunk {tile['kchunk']})")
    rl = roofline(k_ms, 16.0 * lup, copy,
                  committed_traffic(f"pmc_rk4_fused_{tile['variant']}_64x384cubed_fp64.txt"), kernel)
    out = {
        "metric": METRIC, "value": round(lup / (ms / 1e3) / 1e9, 3), "unit": "GLUP/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (8 seeded Gaussians, SplitMix64 seed 0; velocity seed 1)",
        "config": {"workload": "384^3 grid, 64 time slices batched on one GPU, F FP64 RK4 dt=2e-5 (1563 steps "
                               "per slice), G RK4 BF16 storage / FP32 arithmetic dt=2/1024 in 2 sub-steps, "
                               "difference-form correction, stop at rel. increment 1e-6 (K<=8), T=2, nu=1e-3",
                   "grid": N, "slices": S, "c": [round(x, 6) for x in w.c],
                   "step": "one batched fused RK4 fine step over all 64 slices (64 x 384^3 lattice updates)",
                   "parallelism": "slices-as-batch"},
        "roofline": rl,
    }
    if not a.no_solve:
        out.update(cfg4_solve(a, w, conf, conv, init))
    out["cpu_baseline"] = None if a.no_cpu else cpu_baseline(N, w.nu, w.c, dtF, 4)
    print(json.dumps(out), flush=True)


def parareal_solve(conv, init, conf, w, sol, serial, gdt):
    """One tolerance-stopped Parareal solve with G in storage type `gdt` ("bf16:init64": BF16 G
    with the initial coarse run in FP64; "f64:spectral": the FP64 RK4 coarse map applied as its
    Fourier multiplier), driven iteration by iteration (DoParallel,
    Parareal.h:110-171: begin = the initial coarse chain, iterations until the increment
    test, the last slice's state), each iteration logged."""
    import torch
    import stencilparareal_amd as sp
    dt, _, mode = gdt.partition(":")
    log(f"Parareal solve to {w.tol:g}, G {gdt} ({w.substeps} sub-steps)")
    pr = sp.Parareal(conv, init, sol, conf, coarse="rk4_spectral" if mode == "spectral" else "rk4",
                     coarse_dtype=dt, substeps=w.substeps, tol=w.tol, use_graphs=True, autotune=True,
                     initial_fp64=mode == "init64")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pr.begin()
    t_init = time.perf_counter() - t0
    iters = []
    for k in range(w.kmax):
        inc = pr.iterate()
        f_s, c_s, _ = pr.iteration_times()
        iters.append({"increment": float(f"{inc:.3e}"), "fine_s": round(f_s, 4), "chain_s": round(c_s, 4)})
        log(f"  iteration {k + 1}: increment {inc:.3e}, fine sweep {f_s:.2f} s, coarse chain {c_s:.3f} s")
        if inc <= w.tol:
            break
    pr.solution(sol)
    torch.cuda.synchronize()
    t_par = time.perf_counter() - t0
    del pr
    K = len(iters)
    N, S = w.N, w.slices
    fine_sum = sum(x["fine_s"] for x in iters)
    chain_sum = sum(x["chain_s"] for x in iters)
    return {"g_dtype": gdt, "wall_s": round(t_par, 3), "iterations": K, "tol": w.tol,
            "init_coarse_s": round(t_init, 4), "per_iteration": iters,
            "iteration_over_fine_sweep": round((fine_sum + chain_sum) / fine_sum, 4),
            "rel_l2_err": float(f"{sp.rel_l2(sol, serial):.3e}"),
            "rel_linf_err": float(f"{sp.error_reference(sol, serial):.3e}"),
            "fine_glups_in_solve": round(K * S * conf.timeStepsFinePerTimeSlice() * N ** 3 / fine_sum / 1e9, 3)}


def cfg4_solve(a, w, conf, conv, init):
    """The once-timed runs beside the steps: serial fine, the tolerance-stopped Parareal solve
    as configured (BF16-storage G), the same with FP64 G (plain all-FP64 Parareal, batched:
    its iteration count is the plain comparator's) and with FP32 G, and the plain comparator
    in the reference's execution structure."""
    import torch
    N, S = w.N, w.slices
    tsF, dtF = conf.timeStepsFinePerTimeSlice(), conf.dtFine()
    sol, serial = conv.field("q"), conv.field("qreference")
    log("autotune the one-slice FP64 tile (DoSerial); serial FP64 fine solve")
    conv.autotune(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    conv.DoRK4(init, serial, 0.0, dtF, S * tsF)  # Parareal.h:250-253 (DoSerial)
    torch.cuda.synchronize()
    t_serial = time.perf_counter() - t0
    runs = {}
    for gdt in ([w.coarse_dtype] + [g for g in a.also_g.split(",") if g and g != w.coarse_dtype]):
        runs[gdt] = parareal_solve(conv, init, conf, w, sol, serial, gdt)
    main = runs[w.coarse_dtype]
    # plain all-FP64 Parareal in the reference's execution structure, from measured step
    # times: K iterations of S slices x tsF reference-order fine steps, S-1 coarse segments
    # of tsG FP64 reference-order RK4 sub-steps, plus the initial coarse chain (S segments);
    # K = the iterations FP64 G takes on this problem (measured above when run)
    f64_runs = [g for g in runs if g.split(":")[0] == "f64"]
    K_plain = runs[f64_runs[0]]["iterations"] if f64_runs else main["iterations"]
    log("plain comparator: reference-order step times")
    tsG = conf.timeStepsCoarsePerTimeSlice() * w.substeps
    dtG = conf.dtCoarse() / w.substeps
    f_ms = plain_step_ms(w, dtF)
    g_ms = plain_step_ms(w, dtG)
    plain_s = (K_plain * (S * tsF * f_ms + (S - 1) * tsG * g_ms) + S * tsG * g_ms) / 1e3
    t_par = main["wall_s"]
    return {
        "parareal": main,
        "parareal_other_g": [runs[g] for g in runs if g != w.coarse_dtype],
        "serial_fine": {"wall_s": round(t_serial, 3), "steps": S * tsF,
                        "glups": round(S * tsF * N ** 3 / t_serial / 1e9, 3)},
        "speedup_vs_serial_fine": round(t_serial / t_par, 4),
        "plain_fp64_parareal_reference_structure": {
            "wall_s_extrapolated": round(plain_s, 1), "iterations": K_plain,
            "iterations_from": f"the {f64_runs[0]} solve above" if f64_runs else "the configured solve",
            "fine_step_ms": round(f_ms, 4), "coarse_substep_ms": round(g_ms, 4),
            "how": "reference-order sweeps (4 RHS + 3 Euler + 1 RK kernels per step), one slice after another, "
                   "FP64 G; measured step times x the solve's step counts"},
        "speedup_vs_plain_fp64_parareal": round(plain_s / t_par, 2),
        "speedup_vs_plain_fp64_parareal_by_g": {g: round(plain_s / runs[g]["wall_s"], 2) for g in runs},
    }


# ---------------------------------------------------------------------------------------
# the 256^3 / 32-slice FP32-G workload (configuration 3)

def per_iterate_solve(conv, init, conf, w, sol, serial, gdt, tol=1e-6):
    """Parareal with G in storage type `gdt`, kmax iterations, the error of every iterate
    (the last slice's state, as DoParallel returns it) against the serial fine solution;
    the elapsed solve time counts begin and the iterations only."""
    import torch
    import stencilparareal_amd as sp
    log(f"Parareal, G {gdt}, K = {w.kmax}: error of every iterate")
    pr = sp.Parareal(conv, init, sol, conf, coarse="rk4", coarse_dtype=gdt, use_graphs=True, autotune=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pr.begin()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    pr.solution(sol)
    rows = [{"k": 0, "rel_l2_err": float(f"{sp.rel_l2(sol, serial):.3e}"),
             "rel_linf_err": float(f"{sp.error_reference(sol, serial):.3e}"), "elapsed_s": round(t, 3)}]
    first = None
    for k in range(1, w.kmax + 1):
        t0 = time.perf_counter()
        inc = pr.iterate()
        torch.cuda.synchronize()
        t += time.perf_counter() - t0
        f_s, c_s, _ = pr.iteration_times()
        pr.solution(sol)
        e = sp.rel_l2(sol, serial)
        rows.append({"k": k, "increment": float(f"{inc:.3e}"), "rel_l2_err": float(f"{e:.3e}"),
                     "rel_linf_err": float(f"{sp.error_reference(sol, serial):.3e}"),
                     "fine_s": round(f_s, 4), "chain_s": round(c_s, 4), "elapsed_s": round(t, 3)})
        log(f"  iterate {k}: increment {inc:.3e}, rel. L2 error {e:.3e}, fine sweep {f_s:.2f} s, chain {c_s:.3f} s")
        if first is None and e <= tol:
            first = {"iterate": k, "wall_s": round(t, 3)}
    del pr
    return {"g_dtype": gdt, "iterations": w.kmax, "wall_s": round(t, 3), "per_iterate": rows,
            f"first_within_{tol:g}": first}


def cfg3(a):
    """Configuration 3: 256^3 grid, 32 slices batched on one GPU, 8 seeded Gaussians, c
    uniform in [-1, 1]^3 (seed 1), nu = 2e-3, T = 1; F FP64 RK4 dt = 2.5e-5 (1250 steps per
    slice), G RK4 in FP32 with dt = 1e-3 (difference-form correction), K = 5.  value = the
    batched fine step's lattice updates per second; beside it, once: the serial FP64 fine
    solve, the Parareal solve with G in FP32 (as configured) and in FP64 with the error of
    every iterate against the serial solution, and plain all-FP64 Parareal in the reference's
    structure, extrapolated from measured step times at the iterate count FP64 G needs for
    1e-6."""
    import torch
    import stencilparareal_amd as sp
    from stencilparareal_amd.config import workload
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    w = workload(3)
    N, S = w.N, w.slices
    conf = conf_of(w)
    tsF, dtF = conf.timeStepsFinePerTimeSlice(), conf.dtFine()
    conv = sp.Convection(N, N, N, conf.dx(), conf.nu0, 0.0, conf.cx, conf.cy, conf.cz)
    init = conv.field("qinitial")
    g = w.gauss
    sp.fill_gaussians(init, g.xc, g.yc, g.zc, g.sigma, g.amp)
    log(f"autotune the fused tile for {S} slices")
    conv.autotune(S)
    tile = conv.get_tile(S, True)
    bt = Batch2(conv, init, S)
    log(f"warmup {a.warmup}, timed {a.steps} batched steps")
    ms, k_ms = timed_steps(bt, a, dtF)
    copy = bt.copy_gbs()
    bt.free()
    lup = S * N ** 3
    kernel = (f"k_rk4_fused_{tile['variant']} {tile['tx']}x{tile['ty']} tiles, {tile['nthreads']} threads "
              f"(FP64, {S} slices of {N}^3, kchunk {tile['kchunk']})")
    rl = roofline(k_ms, 16.0 * lup, copy,
                  committed_traffic(f"pmc_rk4_fused_{tile['variant']}_32x256cubed_fp64.txt"), kernel)
    out = {
        "metric": METRIC, "value": round(lup / (ms / 1e3) / 1e9, 3), "unit": "GLUP/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (8 seeded Gaussians, SplitMix64 seed 0; velocity seed 1)",
        "config": {"workload": "256^3 grid, 32 time slices batched on one GPU, F FP64 RK4 dt=2.5e-5 (1250 steps "
                               "per slice), G RK4 FP32 dt=1e-3, difference-form correction, K=5, T=1, nu=2e-3",
                   "grid": N, "slices": S, "c": [round(x, 6) for x in w.c],
                   "step": f"one batched fused RK4 fine step over all {S} slices ({S} x {N}^3 lattice updates)",
                   "parallelism": "slices-as-batch"},
        "roofline": rl,
    }
    if not a.no_solve:
        sol, serial = conv.field("q"), conv.field("qreference")
        log("autotune the one-slice FP64 tile (DoSerial); serial FP64 fine solve")
        conv.autotune(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        conv.DoRK4(init, serial, 0.0, dtF, S * tsF)  # Parareal.h:250-253 (DoSerial)
        torch.cuda.synchronize()
        t_serial = time.perf_counter() - t0
        runs = [per_iterate_solve(conv, init, conf, w, sol, serial, gdt) for gdt in (w.coarse_dtype, "f64")]
        f64 = runs[1]["first_within_1e-06"]
        K_plain = f64["iterate"] if f64 else w.kmax
        log("plain comparator: reference-order step times")
        tsG, dtG = conf.timeStepsCoarsePerTimeSlice(), conf.dtCoarse()
        f_ms, g_ms = plain_step_ms(w, dtF), plain_step_ms(w, dtG)
        plain_s = (K_plain * (S * tsF * f_ms + (S - 1) * tsG * g_ms) + S * tsG * g_ms) / 1e3
        out.update({
            "parareal": runs[0], "parareal_other_g": runs[1:],
            "serial_fine": {"wall_s": round(t_serial, 3), "steps": S * tsF,
                            "glups": round(S * tsF * N ** 3 / t_serial / 1e9, 3)},
            "plain_fp64_parareal_reference_structure": {
                "wall_s_extrapolated": round(plain_s, 1), "iterations": K_plain,
                "iterations_from": "the FP64-G solve above (first iterate within 1e-6)",
                "fine_step_ms": round(f_ms, 4), "coarse_step_ms": round(g_ms, 4)},
            "to_1e-6_vs_serial_and_plain": {
                r["g_dtype"]: ({"wall_s": r["first_within_1e-06"]["wall_s"],
                                "speedup_vs_serial_fine": round(t_serial / r["first_within_1e-06"]["wall_s"], 3),
                                "speedup_vs_plain_fp64_parareal": round(plain_s / r["first_within_1e-06"]["wall_s"], 2)}
                               if r["first_within_1e-06"] else None) for r in runs},
        })
    out["cpu_baseline"] = None if a.no_cpu else cpu_baseline(N, w.nu, w.c, dtF, 3)
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------------------
# the 128^3 / 16-slice FP64 workload (configuration 2)

def cfg2(a):
    import torch
    import stencilparareal_amd as sp
    from stencilparareal_amd.config import workload
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    w = workload(2)
    N, S = w.N, w.slices
    conf = conf_of(w)
    tsF = conf.timeStepsFinePerTimeSlice()
    conv = sp.Convection(N, N, N, conf.dx(), conf.nu0, 0.0, conf.cx, conf.cy, conf.cz)
    init, sol, serial = conv.field("qinitial"), conv.field("q"), conv.field("qreference")
    g = w.gauss
    sp.fill_gaussians(init, g.xc, g.yc, g.zc, g.sigma, g.amp)
    pr = sp.Parareal(conv, init, sol, conf, coarse="rk4", use_graphs=True, autotune=True)
    tile = conv.get_tile(S, True)
    bt = Batch2(conv, init, S)
    ms, k_ms = timed_steps(bt, a, conf.dtFine())
    copy = bt.copy_gbs()
    bt.free()
    lup = S * N ** 3
    kernel = (f"k_rk4_fused_{tile['variant']} {tile['tx']}x{tile['ty']} tiles, {tile['nthreads']} threads "
              f"(FP64, {S} slices of {N}^3, kchunk {tile['kchunk']})")
    rl = roofline(k_ms, 16.0 * lup, copy,
                  committed_traffic(f"pmc_rk4_fused_{tile['variant']}_16x128cubed_fp64.txt"), kernel)
    pr.q = serial
    pr.DoSerial()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pr.DoSerial()
    torch.cuda.synchronize()
    t_serial = time.perf_counter() - t0
    pr.q = sol
    pr.DoParallel()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep = pr.DoParallel()
    torch.cuda.synchronize()
    t_par = time.perf_counter() - t0
    K = rep.iterations
    err = sp.rel_l2(sol, serial)
    del pr
    # the same solve with the coarse RK4 map applied as its Fourier multiplier per segment
    sol2 = conv.field("q_spectral")
    pr2 = sp.Parareal(conv, init, sol2, conf, coarse="rk4_spectral", use_graphs=True, autotune=True)
    pr2.DoParallel()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep2 = pr2.DoParallel()
    torch.cuda.synchronize()
    t_spec = time.perf_counter() - t0
    K2 = rep2.iterations
    spectral = {"wall_s": round(t_spec, 4), "iterations": K2,
                "per_iteration": [{"fine_s": round(rep2.t_fine[k], 4), "chain_s": round(rep2.t_coarse[k], 4)}
                                  for k in range(K2)],
                "iteration_over_fine_sweep": round(sum(rep2.t_fine[k] + rep2.t_coarse[k] for k in range(K2)) /
                                                   sum(rep2.t_fine[k] for k in range(K2)), 4),
                "rel_l2_err": float(f"{sp.rel_l2(sol2, serial):.3e}"),
                "rel_diff_vs_stencil_g": float(f"{sp.rel_l2(sol2, sol):.3e}")}
    del pr2
    res = {"metric": METRIC, "value": round(lup / (ms / 1e3) / 1e9, 3), "unit": "GLUP/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (8 seeded Gaussians, SplitMix64 seed 0)",
           "config": {"workload": "128^3 grid, 16 time slices batched, FP64 fine RK4 dt=5e-5 + FP64 coarse RK4 "
                                  "dt=1/1008, K=5, T=1", "grid": N, "slices": S,
                      "step": "one batched fused RK4 fine step over all 16 slices", "parallelism": "slices-as-batch"},
           "roofline": rl,
           "parareal": {"wall_s": round(t_par, 4), "iterations": K,
                        "per_iteration": [{"fine_s": round(rep.t_fine[k], 4), "chain_s": round(rep.t_coarse[k], 4)}
                                          for k in range(K)],
                        "iteration_over_fine_sweep": round(sum(rep.t_fine[k] + rep.t_coarse[k] for k in range(K)) /
                                                           sum(rep.t_fine[k] for k in range(K)), 4),
                        "rel_l2_err": float(f"{err:.3e}")},
           "parareal_spectral_g": spectral,
           "serial_fine_wall_s": round(t_serial, 4), "cpu_baseline": None}
    print(json.dumps(res), flush=True)


# ---------------------------------------------------------------------------------------
# the uneven batch (configuration 5)

def cfg5(a):
    """24 independent problems (64^3 .. 160^3, 8 .. 24 slices, nu and c drawn from seed 0) in
    one Parareal batch, F in FP64, G in FP32 (difference-form correction), tiles autotuned,
    K = 5.  value = lattice updates of the solves (fine + coarse) per second; per-problem
    error against its serial FP64 fine solution on the same GPU."""
    import torch
    import stencilparareal_amd as sp
    from stencilparareal_amd.config import uneven_batch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    ups = uneven_batch(24, seed=0)
    probs = [sp.Problem(u.n, u.slices, u.nu, u.c, u.end_time, u.steps_fine, u.steps_coarse) for u in ups]
    K = 5
    b = sp.Batch(probs, kmax=K, coarse_dtype="f32", use_graphs=True, autotune=True)
    u0, q, ref = b.fields("u"), b.fields("q"), b.fields("serial")
    for f, u in zip(u0, ups):
        g = u.gauss
        sp.fill_gaussians(f, g.xc, g.yc, g.zc, g.sigma, g.amp)
    for _ in range(a.warmup):
        b.run(u0, q)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        rep = b.run(u0, q)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / a.steps * 1e3
    per = []
    t_serial = 0.0
    for i, u in enumerate(ups):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        b.serial(i, u0[i], ref[i])
        torch.cuda.synchronize()
        t_serial += time.perf_counter() - t1
        per.append({"n": u.n, "slices": u.slices, "nu": float(f"{u.nu:.4e}"),
                    "rel_l2_err": float(f"{sp.rel_l2(q[i], ref[i]):.3e}")})
    bt = sp.Batch(probs, kmax=K, coarse_dtype="f32", tol=1e-6, use_graphs=True, autotune=True)
    qt = bt.fields("qt")
    bt.run(u0, qt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    rep_t = bt.run(u0, qt)
    torch.cuda.synchronize()
    t_tol = time.perf_counter() - t1
    worst = max(sp.rel_l2(qt[i], ref[i]) for i in range(len(ups)))
    # the fine sweep of the timed runs' last solve: every problem's fine steps (grouped 2x2
    # launches), timed apart from the chains by HIP events; 16 algorithmic bytes per update
    fine = [b.iteration_times(k)[0] for k in range(rep.iterations)]
    fine_s = sum(fine) / len(fine)
    sweep_bytes = 16.0 * sum(u.n ** 3 * u.slices * b.info(i).steps_fine_per_slice for i, u in enumerate(ups))
    state = sum(u.n ** 3 * u.slices for u in ups)
    x = torch.empty(state, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        y.copy_(x)
    e1.record()
    e1.synchronize()
    copy = 2.0 * 8 * state * 5 / (e0.elapsed_time(e1) / 1e3) / 1e9
    del x, y
    achieved = sweep_bytes / fine_s / 1e9
    rl = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
          "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
          "frac_of_copy": round(achieved / copy, 4), "copy_gbs": round(copy, 1),
          "kernel": "the fine sweep of one iteration: grouped k_rk4_fused_quad launches over the 24 problems",
          "avg_sweep_ms": round(fine_s * 1e3, 3), "bytes_per_sweep": sweep_bytes,
          "chain_ms_per_iteration": round(1e3 * sum(b.iteration_times(k)[1] for k in range(rep.iterations))
                                          / rep.iterations, 3)}
    out = {"metric": METRIC, "value": round(rep.lattice_updates / (ms / 1e3) / 1e9, 3), "unit": "GLUP/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (4 seeded Gaussians per problem, SplitMix64 seed 0)",
           "config": {"workload": "uneven batch: 24 independent problems, n in {64,96,128,160}, slices in "
                                  "{8,16,24}, nu log-uniform [1e-3,1e-2], c uniform [-1,1]^3, coarse RK4 dt = half its "
                                  "stability limit, fine dt = coarse/16, 2 coarse steps per slice, F FP64, G FP32, K=5",
                      "problems": len(ups), "kmax": K, "parallelism": "problems-and-slices-as-batch",
                      "step": "one batch solve"},
           "iterations": rep.iterations,
           "max_increment_per_iter": [float(f"{x:.3e}") for x in rep.increment[:rep.iterations]],
           "serial_fine_wall_s_all_problems": round(t_serial, 4),
           "batch_to_tol": {"tol": 1e-6, "wall_s": round(t_tol, 4), "iterations": rep_t.iterations,
                            "max_rel_l2_err": float(f"{worst:.3e}"),
                            "speedup_vs_serial_fine": round(t_serial / t_tol, 3)},
           "per_problem": per, "roofline": rl, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------------------
# the reference's own speedup job on one GPU

def refjob(a):
    """The reference's GPU speedup job (jobs/paperJobs/04-SpeedupGPU.job:42-66) on one GPU:
    128^3, nu0 = 0.1, c = (1,1,1), T = 0.1, sin-product initial field, F = 32768 RK4 steps,
    G = 2048 forward-Euler steps with first-order upwind advection, K = 3, blocking schedule,
    4 ... 128 slices batched.  Each slice count runs with the stencil Euler G and with the
    same G applied per segment as its Fourier multiplier, beside the reference's published
    wall time (one GPU per slice) and K = 3 error (tests/golden/published.json, from the
    job's run_<N>_0.log).  value = fine lattice updates per second of the 128-slice solve
    with the Fourier-multiplier G."""
    import torch
    import stencilparareal_amd as sp
    from stencilparareal_amd.config import RuntimeConfiguration
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    with open(os.path.join(ROOT, "tests", "golden", "published.json")) as f:
        pub = json.load(f)
    st, runs = pub["setup"], pub["runs"]["04-SpeedupGPU"]
    N, K = st["gridSize"], 3

    def conf_for(S):
        return RuntimeConfiguration(gridSize=N, endTime=st["endTime"], timeStepsFine=st["timeStepsFine"],
                                    timeStepsCoarse=st["timeStepsCoarse"], nu0=st["nu0"], nufreq=0.0, cx=st["cx"],
                                    cy=st["cy"], cz=st["cz"], kmax=K, async_=False, timeSlices=S)

    c4 = conf_for(4)
    conv = sp.Convection(N, N, N, c4.dx(), c4.nu0, 0.0, c4.cx, c4.cy, c4.cz, upwind=True)
    init, ref, q = conv.field("qinitial"), conv.field("qreference"), conv.field("q")
    sp.fill_exact(init, c4.nu0, 0.0, c4.cx, c4.cy, c4.cz, 0.0)

    def timed(fn):
        fn()  # first call: autotuning memo, graph capture
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    pr = sp.Parareal(conv, init, ref, c4, coarse="euler", use_graphs=True, autotune=True)
    log("serial fine solve (32768 steps)")
    t_serial, _ = timed(pr.DoSerial)
    del pr
    rows = []
    for S in (4, 8, 16, 32, 64, 128):
        pubS = runs[str(S)]
        row = {"slices": S, "published_wall_s": pubS["Parallel run time"], "published_error": pubS["Error at end"]}
        for g in ("euler", "euler_spectral"):
            log(f"{S} slices, coarse {g}")
            pr = sp.Parareal(conv, init, q, conf_for(S), coarse=g, use_graphs=True, autotune=True)
            t, rep = timed(pr.DoParallel)
            err = sp.error_reference(q, ref)
            row[g] = {"wall_s": round(t, 4), "error": float(f"{err:.6e}"),
                      "fine_s": round(sum(rep.t_fine[:K]), 4), "chain_s": round(sum(rep.t_coarse[:K]), 4),
                      "vs_published_wall": round(pubS["Parallel run time"] / t, 2)}
            del pr
        rows.append(row)
    last = rows[-1]["euler_spectral"]["wall_s"]
    fine_lup = K * st["timeStepsFine"] * N ** 3
    out = {"metric": METRIC, "value": round(fine_lup / last / 1e9, 3), "unit": "GLUP/s", "n_gpus": 1,
           "steps": 1, "warmup": 1, "ms_per_step": round(last * 1e3, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "the reference job's own initial field (sin product, fillQ at t = 0)",
           "config": {"workload": "reference speedup job: 128^3, nu0=0.1, c=(1,1,1), T=0.1, F 32768 RK4 steps, "
                                  "G 2048 Euler upwind steps, K=3, 4..128 slices batched on one GPU",
                      "grid": N, "kmax": K, "parallelism": "slices-as-batch",
                      "step": "one K=3 Parareal solve over 128 slices (Fourier-multiplier Euler G)"},
           "serial_fine": {"wall_s": round(t_serial, 4), "published_wall_s": runs["4"]["Serial run time"]},
           "by_slices": rows, "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def roofline_only(a):
    """Only the dominant kernel and a copy kernel of the same bytes (the PMC scale), for
    profilers: a.reps batched steps of the workload's slices."""
    if a.workload not in ("cfg4", "cfg3", "cfg2"):
        sys.exit("--roofline-only runs the batched step of cfg4, cfg3 or cfg2")
    import torch
    import stencilparareal_amd as sp
    from stencilparareal_amd.config import workload
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    w = workload({"cfg4": 4, "cfg3": 3, "cfg2": 2}[a.workload])
    N, S = w.N, a.slices or w.slices
    conf = conf_of(w)
    conv = sp.Convection(N, N, N, conf.dx(), conf.nu0, 0.0, conf.cx, conf.cy, conf.cz)
    if a.variant == "auto":
        conv.autotune(S, a.dtype, a.dtype)
    else:
        # a fixed kernel: under counter collection the autotuner's timings are not the
        # kernels' own, so profiles name the kernel they measure
        conv.set_tile(32, 32, 512, 0, variant=a.variant)
    init = conv.field("qinitial")
    g = w.gauss
    sp.fill_gaussians(init, g.xc, g.yc, g.zc, g.sigma, g.amp)
    bt = Batch2(conv, init, S, a.dtype)
    bt.run(2, conf.dtFine())
    torch.cuda.synchronize()
    ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
    bt.run(a.reps, conf.dtFine(), ev)
    ev[1].synchronize()
    k_ms = ev[0].elapsed_time(ev[1]) / a.reps
    copy = bt.copy_gbs(reps=3)
    alg = 2.0 * {"f64": 8, "f32": 4, "bf16": 2}[a.dtype] * S * N ** 3
    print(json.dumps({"kernel_avg_ms": round(k_ms, 4), "bytes_per_launch": alg, "dtype": a.dtype,
                      "tile": conv.get_tile(S, a.dtype == "f64"),
                      "copy_bytes_read": alg / 2, "copy_bytes_written": alg / 2, "copy_gbs": round(copy, 1),
                      "reps": a.reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-solve", action="store_true", help="skip the once-timed Parareal, serial and comparator runs")
    ap.add_argument("--also-g", default="f64:spectral,bf16:init64",
                    help="cfg4: G variants solved beside the configured BF16 one (comma list of f64, f32, bf16; "
                         "':init64' for the initial coarse run in FP64, 'f64:spectral' for the FP64 RK4 coarse "
                         "map as a Fourier multiplier; '' for none)")
    ap.add_argument("--workload", choices=["cfg4", "cfg3", "cfg2", "cfg5", "ref"], default="cfg4",
                    help="cfg4: 384^3 x 64 slices, BF16 G (default); cfg3: 256^3 x 32 slices, FP32 G; "
                         "cfg2: 128^3 x 16 slices, FP64; cfg5: the 24-problem uneven batch; ref: the "
                         "reference's own speedup job (128^3, Euler G, K = 3, 4..128 slices)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="run only the batched RK4 step and a copy kernel of the same size (for profilers)")
    ap.add_argument("--reps", type=int, default=5, help="launches of the kernel with --roofline-only")
    ap.add_argument("--variant", choices=["auto", "quad", "pair"], default="auto",
                    help="fused kernel for --roofline-only: autotuned (default) or a fixed 32x32x512 variant")
    ap.add_argument("--dtype", choices=["f64", "f32", "bf16"], default="f64",
                    help="storage type of the step with --roofline-only (f32 / bf16: FP32 arithmetic, "
                         "the coarse propagator's kernels)")
    ap.add_argument("--slices", type=int, default=0,
                    help="slices per launch with --roofline-only (default: the workload's; 1: the "
                         "single-slice step of the coarse chain and the serial solve)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}): the time slices are the batch "
              "index of a single device; multi-GPU slice sharding is not part of this build.",
              file=sys.stderr)
        sys.exit(2)
    if a.roofline_only:
        return roofline_only(a)
    return {"cfg4": cfg4, "cfg3": cfg3, "cfg2": cfg2, "cfg5": cfg5, "ref": refjob}[a.workload](a)


if __name__ == "__main__":
    main()
