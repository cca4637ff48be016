"""The C++ `parareal` driver (stencilparareal_amd/host/parareal_main.cpp), which mirrors
the reference's src/Parareal/main.cpp: option names, run modes and the stdout lines the
reference's job scripts parse ("Serial run time:", "Parallel run time:", "Speedup:",
"Maximal speedup:", "Error at end:", "Error of parareal:", "Times:"), and the C++
`convection` accuracy driver (host/convection_main.cpp, mirroring src/Convection/main.cpp).

CPU tests: the executables are built, resolve libsp_hip.so and answer --help without
touching a device.  GPU tests: every mode runs, the errors printed match the oracle's on
the same configuration, and the MAT files written read back (scipy) with those errors.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from stencilparareal_amd.config import RuntimeConfiguration

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "stencilparareal_amd", "bin", "parareal")
CONVECTION = os.path.join(ROOT, "stencilparareal_amd", "bin", "convection")

ARGS = dict(gridSize=24, timeSlices=4, kmax=2, endTime=0.02, timeStepsFine=256, timeStepsCoarse=32, nu0=0.1,
            nufreq=0.0, cx=1.0, cy=1.0, cz=1.0)


def argv(mode, **over):
    a = dict(ARGS, **over)
    out = [DRIVER, "--mode", mode]
    for k, v in a.items():
        out += [f"--{k}", str(v)]
    return out


def run(cmd, timeout=120, cwd=None):
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=timeout, cwd=cwd)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stdout
    return r.stdout


def field(out, name):
    m = re.search(rf"^{re.escape(name)}:\s*([-+0-9.eE]+)", out, re.M)
    assert m, f"'{name}:' missing from the output"
    return float(m.group(1))


def oracle_error(**over):
    a = dict(ARGS, **over)
    conf = RuntimeConfiguration(gridSize=a["gridSize"], endTime=a["endTime"], timeStepsFine=a["timeStepsFine"],
                                timeStepsCoarse=a["timeStepsCoarse"], nu0=a["nu0"], nufreq=a["nufreq"],
                                cx=a["cx"], cy=a["cy"], cz=a["cz"], kmax=a["kmax"], async_=False,
                                timeSlices=a["timeSlices"])
    N = conf.gridSize
    conv = O.OConvection(N, N, N, conf.dx(), conf.nu0, conf.nufreq, conf.cx, conf.cy, conf.cz)
    u0 = O.fill_exact(N, conf.nu0, conf.nufreq, conf.cx, conf.cy, conf.cz, 0.0)
    par, _ = conv.parareal(u0, conf.timeSlices, conf.kmax, conf.endTime, conf.dtFine(),
                           conf.timeStepsFinePerTimeSlice(), conf.dtCoarse(), conf.timeStepsCoarsePerTimeSlice(),
                           "euler", keep_iters=False)
    ser = conv.serial(u0, conf.dtFine(), conf.timeStepsFinePerTimeSlice() * conf.timeSlices)
    return O.error_reference(par, ser)


@pytest.mark.parametrize("exe", [DRIVER, CONVECTION], ids=["parareal", "convection"])
def test_driver_built_and_linked(exe):
    assert os.access(exe, os.X_OK), f"{exe} is not built"
    ldd = subprocess.run(["ldd", exe], stdout=subprocess.PIPE, text=True).stdout
    line = [ln for ln in ldd.splitlines() if "libsp_hip.so" in ln]
    assert line and "not found" not in line[0], ldd
    assert os.path.join(ROOT, "stencilparareal_amd") in os.path.realpath(line[0].split("=>")[1].split()[0])


def test_convection_help_lists_reference_options():
    out = run([CONVECTION, "--help"])
    for opt in ("--nu0", "--nufreq", "--cx", "--cy", "--cz", "--gridSize", "--timeSteps", "--endTime", "--cfl",
                "--order", "--mat"):
        assert opt in out


def test_driver_help_lists_reference_options():
    out = run([DRIVER, "--help"])
    for opt in ("--nu0", "--nufreq", "--cx", "--cy", "--cz", "--gridSize", "--endTime", "--timeStepsFine",
                "--timeStepsCoarse", "--kmax", "--mat", "--noasync", "--mode"):
        assert opt in out


@pytest.mark.gpu
def test_driver_compare_mode_matches_oracle(tmp_path):
    out = run(argv("compare"), cwd=tmp_path)
    for name in ("Serial run time", "Parallel run time", "Speedup", "Maximal speedup"):
        assert field(out, name) > 0
    want = oracle_error()
    got = field(out, "Error at end")
    # printed with 6 significant digits; the fused fine propagator agrees with the oracle
    # to 1e-12 of the field, far below the Parareal error itself (~1e-5 here)
    assert abs(got - want) <= 2e-5 * want, (got, want)
    # result.mat (main.cpp:251-253): q and qreference, whose error is the one printed
    sio = pytest.importorskip("scipy.io")
    m = sio.loadmat(str(tmp_path / "result.mat"))
    n = ARGS["gridSize"]
    assert m["q"].shape == m["qreference"].shape == (n, n, n)
    e = np.abs(m["q"] - m["qreference"]).max() / np.abs(m["q"]).max()
    assert abs(e - got) <= 1e-5 * got


def oracle_convection_error(N, nu0, nufreq, c, T, steps, order):
    conv = O.OConvection(N, N, N, 1.0 / N, nu0, nufreq, *c)
    q0 = O.fill_exact(N, nu0, nufreq, *c, 0.0)
    q = conv.rk4(q0, 0.0, T / steps, steps) if order == 4 else conv.euler(q0, 0.0, T / steps, steps)
    ex = O.fill_exact(N, nu0, nufreq, *c, T)
    return np.abs(q - ex).max() / np.abs(ex).max(), q, ex


@pytest.mark.gpu
@pytest.mark.parametrize("order", [4, 1])
@pytest.mark.filterwarnings("ignore:Duplicate variable name")  # two arrays named q, as the reference writes
def test_convection_driver_matches_oracle(tmp_path, order):
    N, T, steps, nu0, nufreq, c = 20, 0.01, 40, 0.1, 30.0, (1.0, -0.5, 0.25)
    out = run([CONVECTION, "--gridSize", str(N), "--timeSteps", str(steps), "--endTime", str(T), "--nu0", str(nu0),
               "--nufreq", str(nufreq), "--cx", str(c[0]), "--cy", str(c[1]), "--cz", str(c[2]), "--order",
               str(order), "--mat"], cwd=tmp_path)
    want, q, ex = oracle_convection_error(N, nu0, nufreq, c, T, steps, order)
    got = float(re.search(r"Total relative error:\s*([-+0-9.eE]+)", out).group(1))
    assert abs(got - want) <= 1e-5 * want, (got, want)
    assert "solver: " + ("Runge-Kutta 4" if order == 4 else "explicit Euler") in out
    sio = pytest.importorskip("scipy.io")
    init = sio.loadmat(str(tmp_path / "init.mat"))["qIn"]
    assert np.allclose(init, O.fill_exact(N, nu0, nufreq, *c, 0.0).transpose(2, 1, 0), rtol=0, atol=1e-15)
    res = sio.loadmat(str(tmp_path / "result_0_0_0.mat"))
    # q appears twice (qIn and qOut share the name "q" as in the reference); the reader
    # keeps the last, the result
    qz = q.transpose(2, 1, 0)
    assert np.abs(res["q"] - qz).max() <= 1e-12 * np.abs(qz).max()
    assert np.array_equal(res["error"], res["q"] - res["exact"])
    assert np.allclose(res["exact"], ex.transpose(2, 1, 0), rtol=0, atol=1e-15)
    cell = sio.loadmat(str(tmp_path / "convection_0_0_0.mat"))["q"]
    assert np.array_equal(cell[0, 0], init)


@pytest.mark.gpu
def test_driver_parallel_serial_timing_modes():
    out = run(argv("parallel", nufreq=20.0))
    got = field(out, "Error of parareal")
    want = oracle_error(nufreq=20.0)
    assert abs(got - want) <= 2e-4 * want, (got, want)  # printed as %.4e
    assert "Parallel run time:" in out
    out = run(argv("serial"))
    assert field(out, "Serial run time") > 0
    out = run(argv("timing"))
    times = [float(x) for x in out.split("Times:")[1].split()]
    assert len(times) == 5 * ARGS["kmax"] + 2
    assert all(b >= a for a, b in zip(times, times[1:])), times


@pytest.mark.gpu
def test_driver_euler_spectral_coarse_matches_oracle():
    # --coarse eulerspectral: the reference's Euler coarse propagator applied per segment as
    # its Fourier multiplier; the Parareal error is the stencil Euler one's (the oracle's)
    out = run(argv("parallel", nufreq=20.0) + ["--coarse", "eulerspectral"])
    got = field(out, "Error of parareal")
    want = oracle_error(nufreq=20.0)
    assert abs(got - want) <= 2e-4 * want, (got, want)  # printed as %.4e
