"""stencilparareal_amd -- the StencilParareal hot path on AMD MI355X (gfx950).

The RK4 advection-diffusion stencil propagator (reference: src/Convection) used as
fine and coarse propagator, and the Parareal correction sweep (src/Parareal), as
hand-written HIP kernels behind a C ABI (include/stencilparareal.h, libsp_hip.so).
This package is the Python host mirror of the reference interfaces:

    Convection(isize, jsize, ksize, dx, nu0, nufreq, cx, cy, cz).DoRK4(in, out, t0, dt, n)
    Parareal(convection, initial, solution, conf).DoParallel() / DoSerial()
    Batch([Problem(...), ...], kmax=...).run(u0_fields, q_fields)   # uneven problem batches

There is no CPU fallback: importing the compute classes requires the built library.
"""
from . import _lib  # noqa: F401
from .config import RuntimeConfiguration, workload  # noqa: F401


def __getattr__(name):
    # GPU-facing classes import torch lazily, so the configuration logic stays usable
    # (and testable) on machines without a GPU.
    if name in ("Convection",):
        from .convection import Convection
        return Convection
    if name in ("Parareal",):
        from .parareal import Parareal
        return Parareal
    if name in ("Batch", "Problem"):
        from . import batch
        return getattr(batch, name)
    if name in ("ConvectionField", "Context", "periodic_fill", "norms", "error_reference", "rel_l2",
                "fill_exact", "fill_gaussians"):
        from . import field
        return getattr(field, name)
    raise AttributeError(name)
