"""`Parareal`: the time-parallel driver, mirroring src/Parareal/Parareal.h on one GPU.

The reference runs one time slice per MPI rank (Parareal.h:110-171).  Here the S
slices are the batch index of one device: the fine propagation of all slices is one
kernel launch per step, the coarse correction chain runs slice after slice on the
device, and the update is fused with the increment norm.  The interface keeps the
reference's names: DoParallel, DoTimedParallel, DoSerial.
"""
from __future__ import annotations

import ctypes as C

from . import _lib as L
from .config import RuntimeConfiguration
from .convection import Convection
from .field import ConvectionField


class Parareal:
    """Parareal(propagator, initial, solution, conf) (Parareal.h:48-108).

    coarse: "euler" (the reference's DoEuler coarse propagator), "rk4", "rk4_spectral" or
    "euler_spectral" (the same RK4 / Euler coarse map applied per segment as its exact Fourier
    multiplier: FFT, product, inverse FFT; FP64).
    coarse_dtype: "f64", "f32" or "bf16" (storage; f32 arithmetic) for the RK4 coarse.
    tol: stop once the relative L2 increment of the slice states is <= tol (0 = run
    exactly kmax iterations, as the reference does, Parareal.h:116).
    autotune: time the fused kernel's tile shapes for F and G on these sizes at creation.
    correction: "standard" (the reference's G(U_n^{k+1}) + F(U_n^k) - G(U_n^k)), "delta"
    (F(U_n^k) + G(U_n^{k+1} - U_n^k), the same for this linear G, with G's rounding
    proportional to the increment) or "auto" (delta for f32 / bf16 coarse storage).
    initial_fp64: with f32 / bf16 coarse storage and the delta correction, run the initial
    coarse prediction U_n^0 = G^n(u0) in FP64 (the iterations' coarse steps keep coarse_dtype).
    """

    def __init__(self, propagator: Convection, initial: ConvectionField, solution: ConvectionField,
                 conf: RuntimeConfiguration, *, coarse: str = "euler", coarse_dtype: str = "f64",
                 substeps: int = 1, tol: float = 0.0, use_graphs: bool = False, autotune: bool = False,
                 correction: str = "auto", initial_fp64: bool = False):
        self.prop = propagator
        self.qinitial = initial
        self.q = solution
        self.conf = conf
        cfg = L.PararealCfg()
        cfg.slices = conf.timeSlices
        cfg.kmax = conf.kmax
        cfg.end_time = conf.endTime
        cfg.steps_fine = conf.timeStepsFine
        cfg.steps_coarse = conf.timeStepsCoarse
        cfg.coarse_kind = {"euler": L.SP_COARSE_EULER, "rk4": L.SP_COARSE_RK4,
                           "rk4_spectral": L.SP_COARSE_RK4_SPECTRAL,
                           "euler_spectral": L.SP_COARSE_EULER_SPECTRAL}[coarse]
        cfg.coarse_dtype = L.DTYPE_CODES[coarse_dtype]
        cfg.coarse_substeps = substeps
        cfg.tol = tol
        cfg.async_schedule = 1 if conf.async_ else 0
        cfg.use_graphs = 1 if use_graphs else 0
        cfg.autotune = 1 if autotune else 0
        if correction == "auto":
            correction = "standard" if coarse_dtype == "f64" else "delta"
        cfg.delta_correction = {"standard": 0, "delta": 1}[correction]
        cfg.initial_fp64 = 1 if initial_fp64 else 0
        self.cfg = cfg
        h = C.c_void_p()
        L.call("sp_parareal_create", propagator.h, C.byref(cfg), C.byref(h))
        self.h = h
        self.report = None

    # -- the reference interface ----------------------------------------------------
    def DoParallel(self) -> L.PararealReport:
        """Parareal.h:110-171; the last slice's result lands in `solution`."""
        self.prop.ctx.sync_stream()
        rep = L.PararealReport()
        L.call("sp_parareal_run", self.h, self.qinitial.lp(), self.qinitial.vp(), self.q.lp(),
               self.q.vp(), C.byref(rep))
        self.report = rep
        return rep

    def DoTimedParallel(self, times: list) -> L.PararealReport:
        """Parareal.h:173-248: 5*kmax+2 cumulative stamps (init, then per iteration after
        fine, receive, coarse, update, send).  Fine and coarse-chain times come from HIP
        events on the stream; the chain splits into coarse propagation and update where the
        corrections were timed (iterations without graphs).  Receive and send move no data
        on one device and take no time."""
        rep = self.DoParallel()
        times.clear()
        t = rep.t_init
        times.append(t)
        for k in range(self.conf.kmax):
            if k < rep.iterations:
                t += rep.t_fine[k]
                times.extend([t, t])
                t += rep.t_coarse[k] - rep.t_update[k]
                times.append(t)
                t += rep.t_update[k]
                times.extend([t, t])
            else:
                times.extend([t] * 5)
        times.append(rep.t_total)
        return rep

    def DoSerial(self):
        """Parareal.h:250-253: DoRK4(qinitial -> q, 0, dtFine, tsFine*S)."""
        self.prop.ctx.sync_stream()
        L.call("sp_parareal_serial", self.h, self.qinitial.lp(), self.qinitial.vp(), self.q.lp(),
               self.q.vp())

    # -- iterate-level access (error per iterate) -------------------------------------
    def begin(self):
        self.prop.ctx.sync_stream()
        L.call("sp_parareal_begin", self.h, self.qinitial.lp(), self.qinitial.vp())

    def iterate(self) -> float:
        inc = C.c_double()
        L.call("sp_parareal_iterate", self.h, C.byref(inc))
        return inc.value

    def iteration_times(self) -> tuple[float, float, float]:
        """(fine sweep, coarse chain with corrections, corrections alone) of the last
        iterate() in seconds, from HIP events; the last is 0 for graph iterations."""
        f, c, u = C.c_double(), C.c_double(), C.c_double()
        L.call("sp_parareal_iteration_times", self.h, C.byref(f), C.byref(c), C.byref(u))
        return f.value, c.value, u.value

    def solution(self, out: ConvectionField | None = None) -> ConvectionField:
        out = out or self.q
        L.call("sp_parareal_solution", self.h, out.lp(), out.vp())
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None):
                L.load().sp_parareal_destroy(self.h)
        except Exception:
            pass
