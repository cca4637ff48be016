"""Device fields and the library context.

`ConvectionField` mirrors the reference's field type (src/Convection/Convection.h:13-20):
a 3-D scalar on an ni x nj x nk interior, with either the compact layout (no halo,
i fastest) or the STELLA CUDA storage of the reference GPU build (2-cell halo, i
padded to 32 doubles), described to the C ABI by an `sp_layout`.  Device memory is a
torch tensor (plumbing only); every computation goes through libsp_hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L

_TORCH_DT = {L.SP_F64: torch.float64, L.SP_F32: torch.float32, L.SP_BF16: torch.bfloat16}


class Context:
    """An sp_ctx bound to one device; work runs on torch's current stream."""

    _default: dict[int, "Context"] = {}

    def __init__(self, device: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("stencilparareal_amd needs a HIP device (torch.cuda.is_available() is False)")
        self.device = device
        h = C.c_void_p()
        L.call("sp_ctx_create", device, C.byref(h))
        self.h = h
        self.sync_stream()

    @classmethod
    def default(cls, device: int = 0) -> "Context":
        if device not in cls._default:
            cls._default[device] = Context(device)
        return cls._default[device]

    def sync_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        L.call("sp_ctx_set_stream", self.h, C.c_void_p(s))

    def synchronize(self):
        L.call("sp_ctx_sync", self.h)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                L.load().sp_ctx_destroy(self.h)
        except Exception:
            pass


def compact_layout(ni: int, nj: int, nk: int, dtype: int = L.SP_F64) -> L.Layout:
    lay = L.Layout()
    L.call("sp_layout_compact", ni, nj, nk, dtype, C.byref(lay))
    return lay


def stella_layout(ni: int, nj: int, nk: int) -> L.Layout:
    lay = L.Layout()
    L.call("sp_layout_stella_cuda", ni, nj, nk, C.byref(lay))
    return lay


class ConvectionField:
    """A field on the device.  `kind` is "compact" or "stella" (reference GPU storage)."""

    def __init__(self, name: str, isize: int, jsize: int, ksize: int, kind: str = "compact",
                 dtype: str = "f64", device: int = 0):
        self.name = name
        code = L.DTYPE_CODES[dtype]
        if kind == "compact":
            self.layout = compact_layout(isize, jsize, ksize, code)
        elif kind == "stella":
            if code != L.SP_F64:
                raise ValueError("the STELLA CUDA layout holds doubles")
            self.layout = stella_layout(isize, jsize, ksize)
        else:
            raise ValueError(f"unknown layout kind {kind!r}")
        self.kind = kind
        self.dtype = code
        n = int(L.load().sp_layout_storage_elems(C.byref(self.layout)))
        self.storage = torch.zeros(n, dtype=_TORCH_DT[code], device=f"cuda:{device}")

    # -- reference-style accessors ------------------------------------------------
    def calculationDomain(self):
        return (self.layout.ni, self.layout.nj, self.layout.nk)

    @property
    def shape(self):
        return (self.layout.nk, self.layout.nj, self.layout.ni)

    @property
    def ptr(self) -> int:
        return self.storage.data_ptr()

    def interior(self) -> torch.Tensor:
        """A (nk, nj, ni) strided view of the interior."""
        lay = self.layout
        return torch.as_strided(self.storage, (lay.nk, lay.nj, lay.ni), (lay.sk, lay.sj, 1), lay.origin)

    def set(self, a):
        t = torch.as_tensor(a).to(device=self.storage.device)
        self.interior().copy_(t.to(self.storage.dtype).reshape(self.shape))

    def get(self) -> np.ndarray:
        return self.interior().to(torch.float64).cpu().numpy()

    def SwapWith(self, other: "ConvectionField"):
        for attr in ("layout", "storage", "kind", "dtype"):
            a, b = getattr(self, attr), getattr(other, attr)
            setattr(self, attr, b)
            setattr(other, attr, a)

    def lp(self):
        return C.byref(self.layout)

    def vp(self):
        return C.c_void_p(self.ptr)


def periodic_fill(field: ConvectionField, ctx: Context | None = None):
    """ApplyImpl (src/Communication/Periodicity.h:103) on the field's halo."""
    ctx = ctx or Context.default(field.storage.device.index or 0)
    ctx.sync_stream()
    L.call("sp_periodic_fill", ctx.h, field.lp(), field.vp())


def norms(a: ConvectionField, b: ConvectionField, ctx: Context | None = None):
    """(max|a-b|, max|a|, sum (a-b)^2, sum b^2) over the interior, on the device."""
    ctx = ctx or Context.default(a.storage.device.index or 0)
    ctx.sync_stream()
    out = (C.c_double * 4)()
    L.call("sp_norms", ctx.h, a.lp(), a.vp(), b.lp(), b.vp(), out)
    return tuple(out)


def error_reference(field: ConvectionField, reference: ConvectionField, ctx: Context | None = None) -> float:
    """computeErrorReference (src/Parareal/main.cpp:98-139): max|f-r| / max|f|."""
    e, m, _, _ = norms(field, reference, ctx)
    return e / m


def rel_l2(field: ConvectionField, reference: ConvectionField, ctx: Context | None = None) -> float:
    _, _, d2, r2 = norms(field, reference, ctx)
    return float(np.sqrt(d2) / np.sqrt(r2))


def fill_exact(field: ConvectionField, nu0, nufreq, cx, cy, cz, t=0.0, ctx: Context | None = None):
    """fillQ (src/Convection/ConvectionSolution.h:23-52) on the device."""
    ctx = ctx or Context.default(field.storage.device.index or 0)
    ctx.sync_stream()
    L.call("sp_fill_exact", ctx.h, field.lp(), field.vp(), nu0, nufreq, cx, cy, cz, t)


def fill_gaussians(field: ConvectionField, xc, yc, zc, sigma, amp, ctx: Context | None = None):
    ctx = ctx or Context.default(field.storage.device.index or 0)
    ctx.sync_stream()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (xc, yc, zc, sigma, amp)]
    ptrs = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
    L.call("sp_fill_gaussians", ctx.h, field.lp(), field.vp(), len(arrs[0]), *ptrs)
