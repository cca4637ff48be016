"""`Batch`: independent Parareal problems of uneven sizes run together on one GPU.

No counterpart in the reference, which runs one problem per MPI job.  Each launch of the
library covers every problem (grouped launches: one per fine step, one per coarse
sub-step at each chain position, one per correction position).  F is FP64, G is RK4 in
f64 / f32 / bf16 storage, the correction is the difference form.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import _lib as L
from .field import Context, ConvectionField


@dataclass
class Problem:
    n: int
    slices: int
    nu: float
    c: tuple
    end_time: float
    steps_fine: int
    steps_coarse: int


class Batch:
    def __init__(self, problems: list[Problem], *, kmax: int, coarse_dtype: str = "f32", substeps: int = 1,
                 tol: float = 0.0, use_graphs: bool = True, autotune: bool = False, ctx: Context | None = None,
                 device: int = 0):
        self.ctx = ctx or Context.default(device)
        self.problems = list(problems)
        arr = (L.Problem * len(problems))()
        for a, p in zip(arr, problems):
            a.n, a.slices, a.nu = p.n, p.slices, p.nu
            a.cx, a.cy, a.cz = p.c
            a.end_time, a.steps_fine, a.steps_coarse = p.end_time, p.steps_fine, p.steps_coarse
        cfg = L.BatchCfg()
        cfg.kmax, cfg.coarse_dtype, cfg.coarse_substeps = kmax, L.DTYPE_CODES[coarse_dtype], substeps
        cfg.tol, cfg.use_graphs, cfg.autotune = tol, 1 if use_graphs else 0, 1 if autotune else 0
        self.ctx.sync_stream()
        h = C.c_void_p()
        L.call("sp_batch_create", self.ctx.h, len(problems), arr, C.byref(cfg), C.byref(h))
        self.h = h

    def fields(self, name="q") -> list[ConvectionField]:
        return [ConvectionField(f"{name}{i}", p.n, p.n, p.n, device=self.ctx.device) for i, p in enumerate(self.problems)]

    def run(self, u0: list[ConvectionField], q: list[ConvectionField]) -> L.BatchReport:
        n = len(self.problems)
        assert len(u0) == n and len(q) == n
        for f, p in zip(list(u0) + list(q), self.problems + self.problems):
            assert f.layout.ni == p.n and f.layout.halo == 0 and f.dtype == L.SP_F64, "compact FP64 fields"
        a, b = (C.c_void_p * n)(*[f.ptr for f in u0]), (C.c_void_p * n)(*[f.ptr for f in q])
        self.ctx.sync_stream()
        rep = L.BatchReport()
        L.call("sp_batch_run", self.h, a, b, C.byref(rep))
        self.report = rep
        return rep

    def increments(self, p: int) -> list[float]:
        out = (C.c_double * L.SP_MAX_ITERS)()
        L.call("sp_batch_increments", self.h, p, out)
        return list(out[:self.report.iterations])

    def iteration_times(self, k: int) -> tuple[float, float]:
        """(fine sweep, coarse chains with the corrections) of iteration k of the last run,
        in seconds (HIP events)."""
        f, c = C.c_double(), C.c_double()
        L.call("sp_batch_iteration_times", self.h, k, C.byref(f), C.byref(c))
        return f.value, c.value

    def info(self, p: int) -> L.BatchInfo:
        i = L.BatchInfo()
        L.call("sp_batch_problem_info", self.h, p, C.byref(i))
        return i

    def serial(self, p: int, u0: ConvectionField, q: ConvectionField):
        self.ctx.sync_stream()
        L.call("sp_batch_serial", self.h, p, u0.vp(), q.vp())

    def __del__(self):
        try:
            if getattr(self, "h", None):
                L.load().sp_batch_destroy(self.h)
        except Exception:
            pass
