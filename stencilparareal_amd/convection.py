"""`Convection`: the advection-diffusion propagator, mirroring the reference class
(src/Convection/Convection.h:22-108) on top of libsp_hip.so."""
from __future__ import annotations

import ctypes as C

from . import _lib as L
from .field import Context, ConvectionField


class Convection:
    """Convection(isize, jsize, ksize, dx, nu0, nufreq, cx, cy, cz) (Convection.h:25).

    upwind: Advection2 branch of DoEuler (True = the reference build's default,
    CMakeLists.txt:101-105).  path: "fused" (one kernel per RK4 step) or "staged"
    (the reference's 4 RHS + 3 Euler + 1 RK sweeps, bit-for-bit its operation order);
    None keeps the library's choice (fused, or staged when ksize < 8).
    """

    def __init__(self, isize, jsize, ksize, dx, nu0, nufreq, cx, cy, cz, *, upwind=True,
                 path=None, ctx: Context | None = None, device: int = 0):
        self.ctx = ctx or Context.default(device)
        self.sizes = (isize, jsize, ksize)
        self.dx, self.nu0, self.nufreq = dx, nu0, nufreq
        self.cx, self.cy, self.cz = cx, cy, cz
        h = C.c_void_p()
        L.call("sp_prop_create", self.ctx.h, isize, jsize, ksize, dx, nu0, nufreq, cx, cy, cz, C.byref(h))
        self.h = h
        self.set_upwind(upwind)
        if path is not None:
            self.set_path(path)

    def set_upwind(self, upwind: bool):
        L.call("sp_prop_set_upwind", self.h, 1 if upwind else 0)

    def set_path(self, path: str):
        L.call("sp_prop_set_path", self.h, {"fused": L.SP_PATH_FUSED, "staged": L.SP_PATH_STAGED}[path])

    def set_tile(self, tx=0, ty=0, nthreads=0, kchunk=0, variant=None):
        """Fused-kernel tile.  variant: "scalar", "pair", "quad" (FP64 32x32x512 only), "split"
        (FP64, two launches of a two-stage kernel) or None (built-in choice when tx == 0, else
        scalar)."""
        if tx == 0 and variant is None:
            v = L.SP_KERNEL_AUTO
        else:
            v = {"scalar": L.SP_KERNEL_SCALAR, "pair": L.SP_KERNEL_PAIR, "quad": L.SP_KERNEL_QUAD,
                 "split": L.SP_KERNEL_SPLIT, None: L.SP_KERNEL_SCALAR}[variant]
        L.call("sp_prop_set_tile", self.h, v, tx, ty, nthreads, kchunk)

    def get_tile(self, nslices: int = 1, f64: bool = True) -> dict:
        """The fused-kernel tile a call with `nslices` slices launches."""
        v = [C.c_int() for _ in range(5)]
        L.call("sp_prop_get_tile", self.h, nslices, 1 if f64 else 0, *[C.byref(x) for x in v])
        var, tx, ty, nt, kc = (x.value for x in v)
        names = {L.SP_KERNEL_SCALAR: "scalar", L.SP_KERNEL_PAIR: "pair", L.SP_KERNEL_QUAD: "quad",
                 L.SP_KERNEL_SPLIT: "split"}
        return {"variant": names.get(var, var), "tx": tx,
                "ty": ty, "nthreads": nt, "kchunk": kc}

    def autotune(self, nslices: int = 1, in_dtype: str = "f64", out_dtype: str = "f64") -> float:
        """Times the fused kernel's tile shapes and k chunks for `nslices` slices and keeps
        the fastest (sp_prop_autotune); returns its step time in microseconds."""
        self.ctx.sync_stream()
        us = C.c_double()
        L.call("sp_prop_autotune", self.h, nslices, L.DTYPE_CODES[in_dtype], L.DTYPE_CODES[out_dtype], C.byref(us))
        return us.value

    def computenu(self, t: float) -> float:
        import math
        return self.nu0 * (1. + .5 * math.sin(self.nufreq * t))

    def field(self, name="q", kind="compact", dtype="f64") -> ConvectionField:
        ni, nj, nk = self.sizes
        return ConvectionField(name, ni, nj, nk, kind=kind, dtype=dtype,
                               device=self.ctx.device)

    # -- the reference interface ---------------------------------------------------
    def DoRK4(self, inputField: ConvectionField, outputField: ConvectionField, tStart: float,
              dt: float, timesteps: int):
        """Convection.h:44-45 / Convection.cpp:292-303."""
        self.ctx.sync_stream()
        fn = "sp_do_rk4" if inputField.dtype == outputField.dtype == L.SP_F64 else "sp_do_rk4_ex"
        L.call(fn, self.h, inputField.lp(), inputField.vp(), outputField.lp(), outputField.vp(),
               tStart, dt, timesteps)

    def DoEuler(self, inputField: ConvectionField, outputField: ConvectionField, tStart: float,
                dt: float, timesteps: int):
        """Convection.h:63-64 / Convection.cpp:305-316."""
        self.ctx.sync_stream()
        L.call("sp_do_euler", self.h, inputField.lp(), inputField.vp(), outputField.lp(),
               outputField.vp(), tStart, dt, timesteps)

    def DoRK4Timestep(self, inputField, outputField, t, dt):
        """Convection.h:69-70 / Convection.cpp:318-370."""
        self.ctx.sync_stream()
        L.call("sp_do_rk4_timestep", self.h, inputField.lp(), inputField.vp(), outputField.lp(),
               outputField.vp(), t, dt)

    def DoEulerTimestep(self, inputField, outputField, t, dt):
        """Convection.h:74-75 / Convection.cpp:372-392."""
        self.ctx.sync_stream()
        L.call("sp_do_euler_timestep", self.h, inputField.lp(), inputField.vp(), outputField.lp(),
               outputField.vp(), t, dt)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                L.load().sp_prop_destroy(self.h)
        except Exception:
            pass
