"""sp_do_rk4_noalloc, the allocation-free DoRK4, against sp_do_rk4 through the C ABI.

Two calls in a row on one device, the second on a smaller field (the device's scratch
array keeps the first call's contents), each bit for bit equal to sp_do_rk4 on the same
input: out-of-place with an odd and an even step count, and in place.  Fields above the
scratch's 4194304 points and the reference-order path are refused with an error.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def sp():
    import stencilparareal_amd as spm
    from stencilparareal_amd import field
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    field.Context.default(0)
    return spm


def rk4(L, fn, conv, a, b, steps, dt=2e-4, t0=0.1):
    conv.ctx.sync_stream()
    L.call(fn, conv.h, a.lp(), a.vp(), b.lp(), b.vp(), t0, dt, steps)
    torch.cuda.synchronize()


@pytest.mark.parametrize("n_first,n_second", [(40, 24)])
def test_noalloc_matches_rk4_twice_second_smaller(sp, n_first, n_second):
    from stencilparareal_amd import _lib as L
    rng = np.random.default_rng(7)
    for n in (n_first, n_second):
        # nufreq != 0: each step's nu depends on its start time
        conv = sp.Convection(n, n, n, 1.0 / n, 2e-3, 30.0, 0.7, -0.4, 0.2)
        u = rng.standard_normal((n, n, n))
        for steps in (1, 3, 4):
            ref_in, ref_out = conv.field("ri"), conv.field("ro")
            got_in, got_out = conv.field("gi"), conv.field("go")
            for f in (ref_in, got_in):
                f.set(u)
            rk4(L, "sp_do_rk4", conv, ref_in, ref_out, steps)
            rk4(L, "sp_do_rk4_noalloc", conv, got_in, got_out, steps)
            assert np.array_equal(got_out.get(), ref_out.get()), (n, steps, "out of place")
            assert np.array_equal(got_in.get(), u), (n, steps, "input untouched")
            # in place (in == out)
            ref_in.set(u)
            got_in.set(u)
            rk4(L, "sp_do_rk4", conv, ref_in, ref_in, steps)
            rk4(L, "sp_do_rk4_noalloc", conv, got_in, got_in, steps)
            assert np.array_equal(got_in.get(), ref_in.get()), (n, steps, "in place")
        # steps == 0 leaves out untouched
        z = conv.field("z")
        z.set(np.full(z.shape, 3.0))
        rk4(L, "sp_do_rk4_noalloc", conv, got_in, z, 0)
        assert np.all(z.get() == 3.0)


def test_noalloc_reference_layout_output(sp):
    # the reference's padded STELLA layout on the output side, as sp_do_rk4 takes it
    from stencilparareal_amd import _lib as L
    n = 32
    conv = sp.Convection(n, n, n, 1.0 / n, 5e-3, 0.0, 1.0, 0.5, 0.25)
    u = np.random.default_rng(3).standard_normal((n, n, n))
    a = conv.field("a")
    a.set(u)
    r, g = conv.field("r", kind="stella"), conv.field("g", kind="stella")
    rk4(L, "sp_do_rk4", conv, a, r, 5)
    rk4(L, "sp_do_rk4_noalloc", conv, a, g, 5)
    assert np.array_equal(g.get(), r.get())


def test_noalloc_refuses_what_it_cannot_hold(sp):
    from stencilparareal_amd import _lib as L
    big = sp.Convection(162, 162, 162, 1.0 / 162, 1e-3, 0.0, 1.0, 0.5, 0.25)  # 4251528 > 4194304 points
    a, b = big.field("a"), big.field("b")
    with pytest.raises(L.SPError) as ei:
        rk4(L, "sp_do_rk4_noalloc", big, a, b, 2)
    assert ei.value.code == L.SP_ERR_UNSUPPORTED and "exceeds the scratch" in str(ei.value)
    del a, b, big
    staged = sp.Convection(16, 16, 16, 1.0 / 16, 1e-3, 0.0, 1.0, 0.5, 0.25, path="staged")
    a, b = staged.field("a"), staged.field("b")
    with pytest.raises(L.SPError) as ei:
        rk4(L, "sp_do_rk4_noalloc", staged, a, b, 2)
    assert ei.value.code == L.SP_ERR_UNSUPPORTED


def test_noalloc_with_split_tile_matches_rk4(sp):
    # a split tile set on the propagator: sp_do_rk4 runs two launches per step through its
    # work field, sp_do_rk4_noalloc the built-in one-launch tile; the results agree bit for bit
    from stencilparareal_amd import _lib as L
    n = 32
    conv = sp.Convection(n, n, n, 1.0 / n, 5e-3, 30.0, 1.0, 0.5, 0.25)
    conv.set_tile(32, 16, 256, 8, variant="split")
    u = np.random.default_rng(5).standard_normal((n, n, n))
    for steps in (1, 4):
        a, r, g = conv.field("a"), conv.field("r"), conv.field("g")
        a.set(u)
        rk4(L, "sp_do_rk4", conv, a, r, steps)
        rk4(L, "sp_do_rk4_noalloc", conv, a, g, steps)
        assert np.array_equal(g.get(), r.get()), steps
