"""Batches of independent Parareal problems (sp_batch_*, stencilparareal_amd.Batch).

* A batch of small problems with FP64 G against the oracle's Parareal (the reference's
  algorithm), problem by problem: the grouped launches and the difference form of the
  correction give each problem the reference's iterate to 1e-12.
* The same with FP32 G against single-problem GPU Parareal runs (difference form).
* Problems of the uneven-batch configuration's sizes (64^3 .. 160^3, 8 .. 24 slices,
  BF16-storage G): every problem converges to its own serial FP64 fine solution.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle as O  # noqa: E402
from stencilparareal_amd.config import RuntimeConfiguration, uneven_batch  # noqa: E402


@pytest.fixture(scope="module")
def sp():
    import stencilparareal_amd as spm
    from stencilparareal_amd import field
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    field.Context.default(0)
    return spm


def relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


SMALL = [  # n, slices, nu, c, T, steps_fine, steps_coarse
    (8, 2, 0.03, (1.0, -0.5, 0.25), 0.01, 40, 4),
    (12, 3, 0.01, (-0.7, 0.2, 0.9), 0.02, 66, 9),
    (16, 4, 0.02, (0.3, 0.8, -0.6), 0.01, 64, 8),
    (10, 1, 0.05, (0.5, 0.5, 0.5), 0.005, 20, 2),
]


def small_problems(sp):
    return [sp.Problem(n, S, nu, c, T, sf, sc) for n, S, nu, c, T, sf, sc in SMALL]


def gauss(n, i):
    return O.fill_gaussians(n, [0.3 + 0.1 * i, 0.7], [0.4, 0.2], [0.6, 0.5 + 0.05 * i], [0.1, 0.07], [1.0, 0.6])


@pytest.mark.parametrize("graphs", [False, True])
def test_batch_fp64_coarse_matches_oracle_per_problem(sp, graphs):
    K = 3
    probs = small_problems(sp)
    b = sp.Batch(probs, kmax=K, coarse_dtype="f64", use_graphs=graphs)
    u0, q = b.fields("u"), b.fields("q")
    for i, (f, p) in enumerate(zip(u0, probs)):
        f.set(gauss(p.n, i))
    rep = b.run(u0, q)
    assert rep.iterations == K
    assert all(f > 0 and c > 0 for f, c in (b.iteration_times(k) for k in range(K)))
    for i, p in enumerate(probs):
        conf = RuntimeConfiguration(gridSize=p.n, endTime=p.end_time, timeStepsFine=p.steps_fine,
                                    timeStepsCoarse=p.steps_coarse, nu0=p.nu, nufreq=0.0, cx=p.c[0], cy=p.c[1],
                                    cz=p.c[2], kmax=K, async_=False, timeSlices=p.slices)
        oc = O.OConvection(p.n, p.n, p.n, 1.0 / p.n, p.nu, 0.0, *p.c)
        out, _ = oc.parareal(gauss(p.n, i), p.slices, K, p.end_time, conf.dtFine(),
                             conf.timeStepsFinePerTimeSlice(), conf.dtCoarse(), conf.timeStepsCoarsePerTimeSlice(),
                             "rk4", keep_iters=False)
        e = relerr(q[i].get(), out)
        assert e <= 1e-12, (i, e)
        info = b.info(i)
        assert info.steps_fine_per_slice == conf.timeStepsFinePerTimeSlice()
        assert len(b.increments(i)) == K


def test_batch_fp32_coarse_matches_single_problem_runs(sp):
    K = 4
    probs = small_problems(sp)
    b = sp.Batch(probs, kmax=K, coarse_dtype="f32", use_graphs=True)
    u0, q = b.fields("u"), b.fields("q")
    for i, (f, p) in enumerate(zip(u0, probs)):
        f.set(gauss(p.n, i))
    b.run(u0, q)
    for i, p in enumerate(probs):
        conf = RuntimeConfiguration(gridSize=p.n, endTime=p.end_time, timeStepsFine=p.steps_fine,
                                    timeStepsCoarse=p.steps_coarse, nu0=p.nu, nufreq=0.0, cx=p.c[0], cy=p.c[1],
                                    cz=p.c[2], kmax=K, async_=False, timeSlices=p.slices)
        conv = sp.Convection(p.n, p.n, p.n, 1.0 / p.n, p.nu, 0.0, *p.c)
        sol = conv.field()
        pr = sp.Parareal(conv, u0[i], sol, conf, coarse="rk4", coarse_dtype="f32", correction="delta")
        pr.DoParallel()
        e = relerr(q[i].get(), sol.get())
        assert e <= 1e-12, (i, e)
        del pr


def test_batch_uneven_sizes_converge_to_serial_fine(sp):
    # six problems drawn as in the uneven-batch configuration (its sizes, slices, nu, c)
    ups = uneven_batch(count=6, seed=3)
    probs = [sp.Problem(u.n, u.slices, u.nu, u.c, u.end_time, u.steps_fine, u.steps_coarse) for u in ups]
    K = 6
    b = sp.Batch(probs, kmax=K, coarse_dtype="bf16", tol=1e-10, use_graphs=True)
    u0, q = b.fields("u"), b.fields("q")
    for f, u in zip(u0, ups):
        g = u.gauss
        sp.fill_gaussians(f, g.xc, g.yc, g.zc, g.sigma, g.amp)
    rep = b.run(u0, q)
    print("iterations", rep.iterations, "max increments", ["%.2e" % x for x in rep.increment[:rep.iterations]],
          "GLUP", rep.lattice_updates / 1e9, "s", rep.t_total)
    ref = b.fields("serial")
    for i, u in enumerate(ups):
        b.serial(i, u0[i], ref[i])
        e = sp.rel_l2(q[i], ref[i])
        inc = b.increments(i)
        print(f"problem {i}: n={u.n} S={u.slices} nu={u.nu:.2e} rel L2 {e:.2e} increments",
              ["%.1e" % x for x in inc])
        assert e <= 1e-8, (i, e)
        assert all(y < x or y <= 1e-13 for x, y in zip(inc, inc[1:]))
    assert rep.iterations <= K


def test_batch_mixed_kernels_match_oracle(sp):
    # sizes that put different kernels in one batch: 2x2 blocks (96^3 x 32 slices: 288
    # blocks of 32x32) and pair 32x16 (fewer blocks); n is even in a batch
    K = 2
    spec = [(96, 32, 0.01, (0.9, -0.4, 0.3), 0.002, 64, 32),
            (48, 12, 0.02, (-0.5, 0.7, 0.2), 0.003, 24, 6),
            (34, 4, 0.015, (0.2, 0.3, -0.8), 0.004, 16, 4),
            (40, 20, 0.01, (0.6, 0.1, 0.5), 0.002, 40, 20)]
    probs = [sp.Problem(n, S, nu, c, T, sf, sc) for n, S, nu, c, T, sf, sc in spec]
    b = sp.Batch(probs, kmax=K, coarse_dtype="f64", use_graphs=False)
    u0, q = b.fields("u"), b.fields("q")
    for i, (f, p) in enumerate(zip(u0, probs)):
        f.set(gauss(p.n, i))
    b.run(u0, q)
    variants = set()
    for i, p in enumerate(probs):
        variants.add(b.info(i).fine_variant)
        conf = RuntimeConfiguration(gridSize=p.n, endTime=p.end_time, timeStepsFine=p.steps_fine,
                                    timeStepsCoarse=p.steps_coarse, nu0=p.nu, nufreq=0.0, cx=p.c[0], cy=p.c[1],
                                    cz=p.c[2], kmax=K, async_=False, timeSlices=p.slices)
        oc = O.OConvection(p.n, p.n, p.n, 1.0 / p.n, p.nu, 0.0, *p.c)
        out, _ = oc.parareal(gauss(p.n, i), p.slices, K, p.end_time, conf.dtFine(),
                             conf.timeStepsFinePerTimeSlice(), conf.dtCoarse(), conf.timeStepsCoarsePerTimeSlice(),
                             "rk4", keep_iters=False)
        e = relerr(q[i].get(), out)
        assert e <= 1e-12, (i, e)
    assert variants == {1, 2}, variants  # grouped pair and 2x2 kernels both ran



def _uneven_fields(sp, b, ups):
    u0, q = b.fields("u"), b.fields("q")
    for f, u in zip(u0, ups):
        g = u.gauss
        sp.fill_gaussians(f, g.xc, g.yc, g.zc, g.sigma, g.amp)
    return u0, q


def test_cfg5_uneven_batch_as_stated(sp):
    # the uneven-batch configuration exactly as the benchmark runs it: 24 problems drawn
    # with seed 0 (n in {64, 96, 128, 160}, S in {8, 16, 24}, nu log-uniform in [1e-3, 1e-2],
    # c uniform in [-1, 1]^3), F in FP64, G in FP32 (difference-form correction), tiles
    # autotuned, one hipGraph per iteration, K = 5
    ups = uneven_batch(24, seed=0)
    assert len(ups) == 24 and {u.n for u in ups} <= {64, 96, 128, 160} and {u.slices for u in ups} <= {8, 16, 24}
    probs = [sp.Problem(u.n, u.slices, u.nu, u.c, u.end_time, u.steps_fine, u.steps_coarse) for u in ups]
    K = 5
    b = sp.Batch(probs, kmax=K, coarse_dtype="f32", use_graphs=True, autotune=True)
    u0, q = _uneven_fields(sp, b, ups)
    rep = b.run(u0, q)
    assert rep.iterations == K
    # the fine sweep and the chains of every iteration, timed apart (two graphs)
    times = [b.iteration_times(k) for k in range(K)]
    print("fine sweep / chains per iteration (s)", ["%.4f / %.4f" % t for t in times])
    assert all(f > 0 and c > 0 for f, c in times)
    assert sum(f + c for f, c in times) <= rep.t_total
    with pytest.raises(Exception):
        b.iteration_times(K)
    inc = list(rep.increment[:K])
    print("largest increment per iteration", ["%.2e" % x for x in inc])
    # the largest increment over the problems falls every iteration until the FP64 floor
    assert all(y < x or y <= 1e-13 for x, y in zip(inc, inc[1:])), inc
    ref = b.fields("serial")
    worst = 0.0
    for i, u in enumerate(ups):
        b.serial(i, u0[i], ref[i])
        e = sp.rel_l2(q[i], ref[i])
        worst = max(worst, e)
        assert e <= 1e-10, (i, u.n, u.slices, e)
    print(f"24 problems: worst rel L2 against the serial FP64 fine solutions {worst:.2e}")


def test_cfg5_uneven_batch_fp64_coarse_vs_oracle(sp):
    # the same 24 problems (grids, slice counts, nu, c, initial fields) with FP64 G on a
    # shortened horizon: 2 fine steps of the configuration's fine step per slice, one
    # coarse step per slice; every problem's iterate K against the oracle's Parareal
    ups = uneven_batch(24, seed=0)
    K = 2
    spec = []
    for u in ups:
        dtF = u.end_time / u.steps_fine
        spec.append((u.n, u.slices, u.nu, u.c, 2 * u.slices * dtF, 2 * u.slices, u.slices))
    probs = [sp.Problem(n, S, nu, c, T, sf, sc) for n, S, nu, c, T, sf, sc in spec]
    b = sp.Batch(probs, kmax=K, coarse_dtype="f64", use_graphs=True)
    u0, q = _uneven_fields(sp, b, ups)
    b.run(u0, q)
    for i, ((n, S, nu, c, T, sf, sc), u) in enumerate(zip(spec, ups)):
        g = u.gauss
        init = O.fill_gaussians(n, g.xc, g.yc, g.zc, g.sigma, g.amp)
        oc = O.OConvection(n, n, n, 1.0 / n, nu, 0.0, *c)
        out, _ = oc.parareal(init, S, K, T, T / sf, sf // S, T / sc, sc // S, "rk4", keep_iters=False)
        e = relerr(q[i].get(), out)
        assert e <= 1e-12, (i, n, S, e)
