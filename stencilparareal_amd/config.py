"""Run configuration and the benchmark workloads.

`RuntimeConfiguration` mirrors src/Parareal/RuntimeConfiguration.{h,cpp}: the same
options, defaults, derived quantities and the rounding of step totals up to a
multiple of the number of time slices.  The reference takes the slice count from the
MPI world size; here it is an option (`--timeSlices`), since the slices are the batch
index of one GPU.
"""
from __future__ import annotations

import argparse
import math
from dataclasses import dataclass, field

import numpy as np

MODES = ("compare", "serial", "parallel", "timing")


@dataclass
class RuntimeConfiguration:
    # defaults: RuntimeConfiguration.cpp:7-12
    nu0: float = 1.0
    nufreq: float = 0.0
    cx: float = 1.0
    cy: float = 1.0
    cz: float = 1.0
    gridSize: int = 32
    endTime: float = 0.05
    timeStepsFine: int = 128
    timeStepsCoarse: int = 32
    kmax: int = 5
    mat: bool = False
    async_: bool = True
    mode: str = "compare"
    timeSlices: int = 1
    warnings: list = field(default_factory=list)

    def __post_init__(self):
        self.round_steps()

    def round_steps(self):
        """RuntimeConfiguration.cpp:133-151: totals rounded up to multiples of the slices."""
        S = self.timeSlices
        if self.timeStepsFine % S != 0:
            self.timeStepsFine = (self.timeStepsFine // S + 1) * S
            self.warnings.append("Warning: timeStepsFine is not a multiple of timeSlices\n"
                                 f" -- Changing timeStepsFine to {self.timeStepsFine}\n")
        if self.timeStepsCoarse % S != 0:
            self.timeStepsCoarse = (self.timeStepsCoarse // S + 1) * S
            self.warnings.append("Warning: timeStepsCoarse is not a multiple of timeSlices\n"
                                 f" -- Changing timeStepsCoarse to {self.timeStepsCoarse}\n\n")

    # RuntimeConfiguration.h:54-111
    def dx(self) -> float:
        return 1. / self.gridSize

    def dtFine(self) -> float:
        return self.endTime / self.timeStepsFine

    def dtCoarse(self) -> float:
        return self.endTime / self.timeStepsCoarse

    def timeStepsFinePerTimeSlice(self) -> int:
        return self.timeStepsFine // self.timeSlices

    def timeStepsCoarsePerTimeSlice(self) -> int:
        return self.timeStepsCoarse // self.timeSlices

    def timeSliceSize(self) -> float:
        return self.endTime / self.timeSlices

    def _cfl(self, dt: float) -> float:
        c = max(max(self.cx, self.cy), self.cz)
        dt_dx = dt / self.dx()
        dt_dx2 = dt_dx / self.dx()
        return max(3. / 2. * self.nu0 * dt_dx2, c * dt_dx)

    def cflFine(self) -> float:
        return self._cfl(self.dtFine())

    def cflCoarse(self) -> float:
        return self._cfl(self.dtCoarse())

    @classmethod
    def from_argv(cls, argv=None) -> "RuntimeConfiguration":
        """Parses the reference's options (RuntimeConfiguration.cpp:21-131) plus --timeSlices."""
        ap = argparse.ArgumentParser(description="Parareal on one GPU", allow_abbrev=False)
        ap.add_argument("--nu0", type=float)
        ap.add_argument("--nufreq", type=float)
        ap.add_argument("--cx", type=float)
        ap.add_argument("--cy", type=float)
        ap.add_argument("--cz", type=float)
        ap.add_argument("--gridSize", type=int)
        ap.add_argument("--endTime", type=float)
        ap.add_argument("--timeStepsFine", type=int)
        ap.add_argument("--timeStepsCoarse", type=int)
        ap.add_argument("--kmax", type=int)
        ap.add_argument("--mat", action="store_true")
        ap.add_argument("--noasync", action="store_true")
        ap.add_argument("--mode", type=str)
        ap.add_argument("--timeSlices", type=int, default=1)
        a, rest = ap.parse_known_args(argv)
        kw = {k: v for k, v in vars(a).items() if v is not None and k not in ("mat", "noasync", "mode")}
        conf = cls(**kw, mat=a.mat, async_=not a.noasync)
        if a.mode is not None:
            m = a.mode.lower()
            if m in MODES:
                conf.mode = m
            else:
                conf.warnings.insert(0, f"Mode {a.mode} not recognized.\nDefaulting to Compare\n")
                conf.mode = "compare"
        conf.rest = rest
        return conf


# ---------------------------------------------------------------------------------
# Benchmark workloads (the five configurations the benchmark is quoted on).

def splitmix64(seed: int):
    """SplitMix64 stream of uint64 values (Steele, Lea & Flood 2014)."""
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def uniform_stream(seed: int):
    """Doubles in [0, 1) from the top 53 bits of SplitMix64."""
    for v in splitmix64(seed):
        yield (v >> 11) * (1.0 / (1 << 53))


@dataclass
class Gaussians:
    xc: np.ndarray
    yc: np.ndarray
    zc: np.ndarray
    sigma: np.ndarray
    amp: np.ndarray


def random_gaussians(n: int, seed: int, sigma=(0.04, 0.1), amp=(0.5, 1.0)) -> Gaussians:
    u = uniform_stream(seed)
    rows = [[next(u) for _ in range(5)] for _ in range(n)]
    r = np.array(rows)
    return Gaussians(r[:, 0], r[:, 1], r[:, 2], sigma[0] + (sigma[1] - sigma[0]) * r[:, 3],
                     amp[0] + (amp[1] - amp[0]) * r[:, 4])


def random_velocity(seed: int):
    u = uniform_stream(seed)
    return tuple(-1. + 2. * next(u) for _ in range(3))


@dataclass
class Workload:
    name: str
    N: int
    slices: int
    T: float
    c: tuple
    nu: float
    stepsF: int        # requested total fine steps
    stepsG: int        # requested total coarse steps
    kmax: int
    gauss: Gaussians
    coarse: str = "rk4"
    coarse_dtype: str = "f64"
    substeps: int = 1
    tol: float = 0.0


def workload(idx: int, seed: int = 0) -> Workload:
    """The benchmark configurations 1-4 with concrete random draws (seed stated).

    1: 32^3, 4 slices, T=0.5, one Gaussian (centre 0.5, sigma 0.08), all FP64 (CPU case).
    2: 128^3, 16 slices, T=1, 8 random Gaussians, c=(1,0.5,0.25), nu=5e-3, FP64, K=5.
    3: 256^3, 32 slices, T=1, 8 random Gaussians, c ~ U[-1,1]^3, nu=2e-3, G in FP32.
    4: 384^3, 64 slices, T=2, nu=1e-3, G in BF16 storage, K<=8, stop at 1e-6 increment;
       the coarse RK4 step 2e-3 is unstable at this grid, so it runs 2 sub-steps.
    """
    if idx == 1:
        g = Gaussians(np.array([.5]), np.array([.5]), np.array([.5]), np.array([.08]), np.array([1.]))
        return Workload("cfg1-32^3-S4", 32, 4, 0.5, (1.0, 0.5, 0.25), 1e-2, 5000, 500, 3, g)
    if idx == 2:
        return Workload("cfg2-128^3-S16", 128, 16, 1.0, (1.0, 0.5, 0.25), 5e-3, 20000, 1000, 5,
                        random_gaussians(8, seed))
    if idx == 3:
        return Workload("cfg3-256^3-S32", 256, 32, 1.0, random_velocity(seed + 1), 2e-3, 40000, 1000, 5,
                        random_gaussians(8, seed), coarse_dtype="f32")
    if idx == 4:
        return Workload("cfg4-384^3-S64", 384, 64, 2.0, random_velocity(seed + 1), 1e-3, 100000, 1000, 8,
                        random_gaussians(8, seed), coarse_dtype="bf16", substeps=2, tol=1e-6)
    raise ValueError(idx)


@dataclass
class UnevenProblem:
    n: int
    slices: int
    nu: float
    c: tuple
    end_time: float
    steps_fine: int
    steps_coarse: int
    gauss: Gaussians


def uneven_batch(count: int = 24, seed: int = 0, sizes=(64, 96, 128, 160), slices=(8, 16, 24),
                 coarse_per_slice: int = 2, ratio: int = 16) -> list[UnevenProblem]:
    """The uneven batch: `count` independent problems, grid n from `sizes`, nu log-uniform
    in [1e-3, 1e-2], c uniform in [-1, 1]^3, slices from `slices`, all drawn from
    SplitMix64(seed).  Time steps (the configuration leaves them open): the coarse RK4 step
    is half the largest stable step for that grid, nu and c (a scan of the operator's
    symbol), the fine step is `ratio` times smaller, each slice spans `coarse_per_slice`
    coarse steps, so T_p = S_p * coarse_per_slice * dt_coarse.  Initial field: 4 Gaussians
    (centres uniform in [0,1]^3, sigma in [0.04, 0.1], amplitude in [0.5, 1])."""
    u = uniform_stream(seed)
    out = []
    for i in range(count):
        n = sizes[min(int(next(u) * len(sizes)), len(sizes) - 1)]
        nu = 10.0 ** (-3.0 + next(u))
        c = tuple(-1.0 + 2.0 * next(u) for _ in range(3))
        S = slices[min(int(next(u) * len(slices)), len(slices) - 1)]
        dtG = 0.5 * max_stable_rk4_dt(n, nu, c)
        T = S * coarse_per_slice * dtG
        out.append(UnevenProblem(n, S, nu, c, T, S * coarse_per_slice * ratio, S * coarse_per_slice,
                                 random_gaussians(4, seed * 1000 + 17 + i)))
    return out


def rounded_steps(total: int, slices: int) -> int:
    return total if total % slices == 0 else (total // slices + 1) * slices


def max_stable_rk4_dt(N: int, nu: float, c) -> float:
    """Largest dt with |R(dt*lambda)| <= 1 for every Fourier mode of the 4th-order operator,
    by a scan over the symbol (used to pick coarse sub-steps)."""
    th = np.linspace(0, math.pi, 33)
    tx, ty, tz = np.meshgrid(th, th, th, indexing="ij")
    dx = 1.0 / N

    def lam(t, cd):
        return (-1j * cd * (8 * np.sin(t) - np.sin(2 * t)) / (6 * dx)
                + nu * (-2 * np.cos(2 * t) + 32 * np.cos(t) - 30) / (12 * dx * dx))

    L_ = (lam(tx, abs(c[0])) + lam(ty, abs(c[1])) + lam(tz, abs(c[2]))).ravel()
    # the RK4 stability region lies within |z| < 2.9, so dt < 2.9 / max|lambda|
    lo, hi = 0.0, 2.9 / np.abs(L_).max()
    for _ in range(48):
        mid = 0.5 * (lo + hi)
        z = mid * L_
        R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
        if np.abs(R).max() <= 1 + 1e-12:
            lo = mid
        else:
            hi = mid
    return lo
