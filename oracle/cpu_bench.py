"""CPU baseline of the benchmark: the oracle on the host cores (TEST INFRASTRUCTURE).

Run by bench.py's cpu_baseline leg as a child process, so that the OpenMP settings
(OMP_NUM_THREADS = physical cores of this process's CPU set, OMP_PROC_BIND=close,
OMP_PLACES=cores) take effect before the OpenMP runtime starts:

    python -m oracle.cpu_bench --grid 384 --nu 1e-3 --c 0.1 0.2 0.3 --dt 2e-5 --budget 12

Prints one JSON object:
  serial_fine      the oracle's serial fine RK4 (the reference's OpenMP structure: JIK
                   storage, 8x8 ij blocks, one sweep per stage, serial halo loops) on the
                   benchmark grid, a bounded number of steps, GLUP/s;
  one_thread       the same sweep on one thread, a shorter sample;
  parareal_cfg1    plain all-FP64 Parareal on the 32^3 / 4-slice CPU configuration in full
                   (T = 0.5, fine dt 1e-4, coarse RK4 dt 1e-3, K = 3), the S ranks emulated
                   one after another as in orc_parareal_parallel, against the serial fine
                   solve of the same configuration.
"""
from __future__ import annotations

import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def physical_cores() -> tuple[int, int]:
    """(physical cores, logical CPUs) of this process's CPU set, from sysfs topology."""
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology"
        try:
            pkg = open(f"{base}/physical_package_id").read().strip()
            core = open(f"{base}/core_id").read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", str(c)))
    return max(1, len(cores)), len(cpus)


def child_env() -> tuple[dict, int, int]:
    phys, logical = physical_cores()
    env = dict(os.environ)
    env.update({"OMP_NUM_THREADS": str(phys), "OMP_PROC_BIND": "close", "OMP_PLACES": "cores"})
    return env, phys, logical


def run(a) -> dict:
    sys.path.insert(0, os.path.dirname(HERE))
    import oracle as O
    from stencilparareal_amd.config import workload

    N = a.grid
    g = workload(a.gauss_cfg).gauss
    q0 = O.fill_gaussians(N, g.xc, g.yc, g.zc, g.sigma, g.amp)
    conv = O.OConvection(N, N, N, 1.0 / N, a.nu, 0.0, *a.c)
    threads = O.max_threads()
    t0 = O.wtime()
    conv.rk4(q0, 0.0, a.dt, 2)
    probe = max(O.wtime() - t0, 1e-3) / 2
    n = max(2, min(100000, int(a.budget / probe)))
    t0 = O.wtime()
    conv.rk4(q0, 0.0, a.dt, n)
    t = O.wtime() - t0
    O.set_threads(1)
    try:
        n1 = max(2, int(a.budget / 4 / (t / n * threads)))
        t1 = O.wtime()
        conv.rk4(q0, 0.0, a.dt, n1)
        t1 = O.wtime() - t1
    finally:
        O.set_threads(threads)
    out = {"serial_fine": {"grid": N, "steps": n, "seconds": round(t, 3), "glups": round(N ** 3 * n / t / 1e9, 5),
                           "threads": threads},
           "one_thread": {"steps": n1, "seconds": round(t1, 3), "glups": round(N ** 3 * n1 / t1 / 1e9, 5)}}
    if not a.no_parareal:
        w = workload(1)
        from stencilparareal_amd.config import rounded_steps
        S, K = w.slices, w.kmax
        sf, sc = rounded_steps(w.stepsF, S), rounded_steps(w.stepsG, S)
        dtF, dtG = w.T / sf, w.T / sc
        gg = w.gauss
        u0 = O.fill_gaussians(w.N, gg.xc, gg.yc, gg.zc, gg.sigma, gg.amp)
        c1 = O.OConvection(w.N, w.N, w.N, 1.0 / w.N, w.nu, 0.0, *w.c)
        t0 = O.wtime()
        ser = c1.serial(u0, dtF, sf)
        ts = O.wtime() - t0
        t0 = O.wtime()
        par, _ = c1.parareal(u0, S, K, w.T, dtF, sf // S, dtG, sc // S, "rk4", keep_iters=False)
        tp = O.wtime() - t0
        out["parareal_cfg1"] = {
            "workload": "32^3, 4 slices, T=0.5, fine RK4 dt=1e-4, coarse RK4 dt=1e-3, K=3, all FP64",
            "serial_fine_s": round(ts, 4), "parareal_rank_order_s": round(tp, 4),
            "rel_l2_err": float(f"{O.rel_l2(par, ser):.3e}"),
            "note": "the 4 ranks run one after another in one process: the reference overlaps them on 4 nodes"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=384)
    ap.add_argument("--nu", type=float, default=1e-3)
    ap.add_argument("--c", type=float, nargs=3, default=(1.0, 0.5, 0.25))
    ap.add_argument("--dt", type=float, default=2e-5)
    ap.add_argument("--gauss-cfg", type=int, default=4, help="workload whose Gaussians seed the field")
    ap.add_argument("--budget", type=float, default=12.0, help="seconds of the multi-threaded sample")
    ap.add_argument("--no-parareal", action="store_true")
    a = ap.parse_args()
    print(json.dumps(run(a)), flush=True)


if __name__ == "__main__":
    main()
