"""CPU tests of the host logic and of the C-ABI library's surface (no compute calls).

* RuntimeConfiguration reproduces the banner values the reference printed in its own
  runs (jobs/paperJobs/*/run_*_0.log), including the step rounding of
  RuntimeConfiguration.cpp:133-151.
* libsp_hip.so loads and exports every function include/stencilparareal.h declares,
  and the header compiles as C and as C++.
* The product package does not reference the oracle.
"""
import json
import os
import re
import subprocess
import sys

import pytest

from stencilparareal_amd import _lib as L
from stencilparareal_amd.config import (RuntimeConfiguration, max_stable_rk4_dt, random_gaussians,
                                        rounded_steps, workload)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "stencilparareal.h")
PUB = json.load(open(os.path.join(ROOT, "tests", "golden", "published.json")))


@pytest.mark.parametrize("run", ["04-SpeedupGPU", "05-SpeedupCPU"])
@pytest.mark.parametrize("slices", ["4", "32", "128"])
def test_runtime_configuration_matches_reference_banner(run, slices):
    log = PUB["runs"][run][slices]
    conf = RuntimeConfiguration(gridSize=128, endTime=0.1, timeStepsCoarse=2048, timeStepsFine=32768,
                                nu0=0.1, kmax=3, async_=False, timeSlices=int(slices))
    assert f"{conf.dx():g}" == log["spatial discretization step"]
    assert f"{conf.cflFine():g}" == log["CFL fine"]
    assert f"{conf.cflCoarse():g}" == log["CFL coarse"]
    assert f"{conf.dtFine():g}" == log["timestep size fine"]
    assert f"{conf.dtCoarse():g}" == log["timestep size coarse"]
    assert f"{conf.timeSliceSize():g}" == log["time slice size"]
    assert str(conf.timeStepsFinePerTimeSlice()) == log["timesteps per slice fine propagator"]
    assert str(conf.timeStepsCoarsePerTimeSlice()) == log["timesteps per slice coarse propagator"]


def test_step_rounding_rule():
    conf = RuntimeConfiguration(timeStepsFine=100, timeStepsCoarse=10, timeSlices=16)
    assert conf.timeStepsFine == 112 and conf.timeStepsCoarse == 16
    assert len(conf.warnings) == 2
    assert rounded_steps(1000, 16) == 1008 and rounded_steps(1000, 8) == 1000
    conf = RuntimeConfiguration.from_argv(["--gridSize", "64", "--timeSlices", "8", "--noasync",
                                           "--mode", "Timing", "--timeStepsFine", "9"])
    assert conf.gridSize == 64 and not conf.async_ and conf.mode == "timing" and conf.timeStepsFine == 16
    conf = RuntimeConfiguration.from_argv(["--mode", "bogus"])
    assert conf.mode == "compare" and "not recognized" in conf.warnings[0]


def test_workloads_and_stability():
    w2 = workload(2)
    assert w2.N == 128 and w2.slices == 16 and rounded_steps(w2.stepsG, w2.slices) == 1008
    g = random_gaussians(8, 0)
    assert ((g.sigma >= 0.04) & (g.sigma <= 0.1)).all() and ((g.amp >= 0.5) & (g.amp <= 1.0)).all()
    assert g.xc.tolist() == random_gaussians(8, 0).xc.tolist()  # seeded, reproducible
    # the coarse RK4 step of the 384^3 workload is unstable; two sub-steps are stable
    w4 = workload(4)
    dt_g = w4.T / rounded_steps(w4.stepsG, w4.slices)
    lim = max_stable_rk4_dt(w4.N, w4.nu, (1, 1, 1))
    assert dt_g > lim and dt_g / w4.substeps <= lim
    # the fine steps of the benchmark workloads are stable
    for i in (1, 2, 3, 4):
        w = workload(i)
        assert w.T / rounded_steps(w.stepsF, w.slices) <= max_stable_rk4_dt(w.N, w.nu, (1, 1, 1))


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for line in text.splitlines():
        m = re.match(r"^(?:[A-Za-z_]\w*[\s*]+)+(sp_\w+)\s*\(", line)
        if m:
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_exactly_the_bound_symbols():
    assert header_functions() == sorted(L.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(L.LIB_PATH):
        from stencilparareal_amd import build
        build.build()
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (sp_\w+)$", out, flags=re.M))
    missing = [f for f in header_functions() if f not in exported]
    assert not missing, missing
    lib = L.load()  # dlopen works without a GPU; no compute is called
    assert lib.sp_version().decode().startswith("stencilparareal-gfx950")


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(ImportError):
        L.load(str(tmp_path / "libsp_hip.so"))


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_header_compiles(lang, tmp_path):
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text('#include "stencilparareal.h"\nint main(void){sp_layout l; (void)l; return SP_OK;}\n')
    cc = "gcc" if lang == "c" else "g++"
    subprocess.run([cc, "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-fsyntax-only", str(src)],
                   check=True)


def test_product_package_does_not_use_the_oracle():
    pkg = os.path.join(ROOT, "stencilparareal_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(import|from)\s+oracle\b", text, flags=re.M), f
                assert "sp_oracle" not in text and "libsp_oracle" not in text, f


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_bench_refuses_more_than_one_gpu():
    # one GPU per box: the slices are the batch index of a single device
    r = _bench("--gpus", "2")
    assert r.returncode == 2 and "one GPU only" in r.stderr


def test_bench_roofline_only_refuses_workloads_it_does_not_run():
    r = _bench("--roofline-only", "--workload", "cfg5")
    assert r.returncode != 0 and "--roofline-only runs the batched step of cfg4, cfg3 or cfg2" in r.stderr


def test_bench_lists_its_workloads():
    r = _bench("--help")
    assert r.returncode == 0
    for w in ("cfg4", "cfg3", "cfg2", "cfg5", "ref"):
        assert w in r.stdout
