"""ctypes front end of the CPU oracle (``oracle/libsp_oracle.so``).

TEST INFRASTRUCTURE ONLY.  The oracle restates the reference numerics on the
CPU so that the HIP library can be checked against it; only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` import
this module.  The product package ``stencilparareal_amd`` never does.

Arrays exchanged with the oracle are interior-only float64 numpy arrays of
shape ``(nk, nj, ni)`` (i fastest), the same convention as the HIP library's
compact fields.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsp_oracle.so")

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(LIB_PATH):
        cmd = ["make", "-C", _HERE]
        if force:
            cmd.insert(1, "-B")
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_field_new": (_vp, [C.c_int] * 3),
            "orc_field_free": (None, [_vp]),
            "orc_field_set": (None, [_vp, _dp]),
            "orc_field_get": (None, [_vp, _dp]),
            "orc_field_get_storage": (None, [_vp, _dp]),
            "orc_field_set_storage": (None, [_vp, _dp]),
            "orc_field_copy": (None, [_vp, _vp]),
            "orc_apply_periodic": (None, [_dp] + [C.c_int] * 12),
            "orc_field_periodic": (None, [_vp]),
            "orc_conv_new": (_vp, [C.c_int] * 3 + [C.c_double] * 6 + [C.c_int]),
            "orc_conv_free": (None, [_vp]),
            "orc_do_rk4": (None, [_vp, _vp, _vp, C.c_double, C.c_double, C.c_int]),
            "orc_do_euler": (None, [_vp, _vp, _vp, C.c_double, C.c_double, C.c_int]),
            "orc_do_rk4_timestep": (None, [_vp, _vp, _vp, C.c_double, C.c_double]),
            "orc_do_euler_timestep": (None, [_vp, _vp, _vp, C.c_double, C.c_double]),
            "orc_rhs4": (None, [_vp, _vp, _vp, C.c_double]),
            "orc_rhs2": (None, [_vp, _vp, _vp, C.c_double]),
            "orc_computenu": (C.c_double, [_vp, C.c_double]),
            "orc_parareal_parallel": (None, [_vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                             C.c_double, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int]),
            "orc_parareal_serial": (None, [_vp, C.c_double, C.c_int, _dp, _dp]),
            "orc_error_reference": (C.c_double, [_dp, _dp, C.c_long]),
            "orc_rel_l2": (C.c_double, [_dp, _dp, C.c_long]),
            "orc_exactq": (C.c_double, [C.c_double] * 9),
            "orc_fill_exact": (None, [_dp] + [C.c_int] * 3 + [C.c_double] * 6),
            "orc_fill_gaussians": (None, [_dp] + [C.c_int] * 4 + [_dp] * 5),
            "orc_wtime": (C.c_double, []),
            "orc_max_threads": (C.c_int, []),
            "orc_set_threads": (None, [C.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class OField:
    """Reference-layout field: 2-cell halo, JIK storage (k unit stride)."""

    def __init__(self, ni: int, nj: int, nk: int, data: np.ndarray | None = None):
        self.shape = (nk, nj, ni)
        self.h = lib().orc_field_new(ni, nj, nk)
        if data is not None:
            self.set(data)

    def set(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape
        lib().orc_field_set(self.h, _p(a))

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        lib().orc_field_get(self.h, _p(out))
        return out

    def storage(self) -> np.ndarray:
        nk, nj, ni = self.shape
        # JIK order: j slowest, then i, then k
        out = np.empty((nj + 4, ni + 4, nk + 4), dtype=np.float64)
        lib().orc_field_get_storage(self.h, _p(out))
        return out

    def set_storage(self, s: np.ndarray):
        s = np.ascontiguousarray(s, dtype=np.float64)
        lib().orc_field_set_storage(self.h, _p(s))

    def periodic(self):
        lib().orc_field_periodic(self.h)

    def __del__(self):
        try:
            lib().orc_field_free(self.h)
        except Exception:
            pass


class OConvection:
    """Restatement of ``Convection`` (src/Convection/Convection.h:22-108)."""

    def __init__(self, ni, nj, nk, dx, nu0, nufreq, cx, cy, cz, upwind=True):
        self.shape = (nk, nj, ni)
        self.h = lib().orc_conv_new(ni, nj, nk, dx, nu0, nufreq, cx, cy, cz, 1 if upwind else 0)

    def field(self, data=None) -> OField:
        nk, nj, ni = self.shape
        return OField(ni, nj, nk, data)

    def do_rk4(self, fin: OField, fout: OField, t0: float, dt: float, steps: int):
        lib().orc_do_rk4(self.h, fin.h, fout.h, t0, dt, steps)

    def do_euler(self, fin: OField, fout: OField, t0: float, dt: float, steps: int):
        lib().orc_do_euler(self.h, fin.h, fout.h, t0, dt, steps)

    def rk4(self, q0: np.ndarray, t0: float, dt: float, steps: int) -> np.ndarray:
        a, b = self.field(q0), self.field(q0)
        self.do_rk4(a, b, t0, dt, steps)
        return b.get()

    def euler(self, q0: np.ndarray, t0: float, dt: float, steps: int) -> np.ndarray:
        a, b = self.field(q0), self.field(q0)
        self.do_euler(a, b, t0, dt, steps)
        return b.get()

    def rhs4(self, q: np.ndarray, nu: float) -> np.ndarray:
        a, k = self.field(q), self.field()
        lib().orc_rhs4(self.h, a.h, k.h, nu)
        return k.get()

    def rhs2(self, q: np.ndarray, nu: float) -> np.ndarray:
        a, k = self.field(q), self.field()
        lib().orc_rhs2(self.h, a.h, k.h, nu)
        return k.get()

    def computenu(self, t: float) -> float:
        return lib().orc_computenu(self.h, t)

    def parareal(self, u0, S, kmax, T, dtF, tsF, dtG, tsG, g_kind="euler", keep_iters=True,
                 async_=False):
        """Emulated ``Parareal::DoParallel`` (blocking schedule, or the async schedule's
        fine-step times with async_=True).  Returns (final, iters) with iters[k] the last
        rank's q after k iterations."""
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        out = np.empty(self.shape, dtype=np.float64)
        iters = np.empty((kmax + 1,) + self.shape, dtype=np.float64) if keep_iters else None
        g = {"euler": 0, "rk4": 1}[g_kind]
        lib().orc_parareal_parallel(self.h, S, kmax, T, dtF, tsF, dtG, tsG, g, _p(u0),
                                    _p(iters) if iters is not None else None, _p(out), 1 if async_ else 0)
        return out, iters

    def serial(self, u0, dtF, steps_total):
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        out = np.empty(self.shape, dtype=np.float64)
        lib().orc_parareal_serial(self.h, dtF, steps_total, _p(u0), _p(out))
        return out

    def __del__(self):
        try:
            lib().orc_conv_free(self.h)
        except Exception:
            pass


def apply_periodic(storage: np.ndarray, ni, nj, nk, halos, strides):
    """Restated ``ApplyImpl`` on an arbitrary strided float64 buffer (in place)."""
    assert storage.dtype == np.float64 and storage.flags["C_CONTIGUOUS"]
    lib().orc_apply_periodic(_p(storage), ni, nj, nk, *halos, *strides)


def error_reference(f: np.ndarray, r: np.ndarray) -> float:
    f = np.ascontiguousarray(f, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    return lib().orc_error_reference(_p(f), _p(r), f.size)


def rel_l2(f: np.ndarray, r: np.ndarray) -> float:
    f = np.ascontiguousarray(f, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    return lib().orc_rel_l2(_p(f), _p(r), f.size)


def exactq(nu0, nufreq, cx, cy, cz, x, y, z, t) -> float:
    return lib().orc_exactq(nu0, nufreq, cx, cy, cz, x, y, z, t)


def fill_exact(n, nu0, nufreq, cx, cy, cz, t) -> np.ndarray:
    ni, nj, nk = (n, n, n) if np.isscalar(n) else n
    out = np.empty((nk, nj, ni), dtype=np.float64)
    lib().orc_fill_exact(_p(out), ni, nj, nk, nu0, nufreq, cx, cy, cz, t)
    return out


def fill_gaussians(n, xc, yc, zc, sigma, amp) -> np.ndarray:
    ni, nj, nk = (n, n, n) if np.isscalar(n) else n
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (xc, yc, zc, sigma, amp)]
    out = np.empty((nk, nj, ni), dtype=np.float64)
    lib().orc_fill_gaussians(_p(out), ni, nj, nk, len(arrs[0]), *[_p(a) for a in arrs])
    return out


def wtime() -> float:
    return lib().orc_wtime()


def max_threads() -> int:
    return lib().orc_max_threads()


def set_threads(n: int) -> None:
    """OpenMP threads for later oracle calls (the CPU baseline's one-thread sample)."""
    lib().orc_set_threads(int(n))
