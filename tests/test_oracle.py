"""Pins the CPU oracle (oracle/sp_oracle.c) before it is trusted as the checker.

Each reference function the oracle restates is checked against something the
mathematics or the reference's own data fixes, chosen so that a plausible slip (a
dropped term, a wrong sign or offset, a transposed axis) fails:
  * RHS4 / RHS2 (central and upwind): exact discrete Fourier symbols, per axis with
    distinct velocities so an axis swap or sign error shows;
  * RK4 / Euler steppers: the amplification polynomial R(dt*lambda)^n;
  * Periodicity: brute force with the testPeriodicity.cu pattern (value = k + 100j + 10000i);
  * spatial order: convergence to exactQ at fourth order;
  * Parareal: k = S iterations reproduce the serial fine solution, G = F converges in
    one iteration, and the first iterate equals the formula evaluated by hand;
  * norms: the definition of computeErrorReference (normalised by the candidate).
"""
import math

import numpy as np
import pytest

import oracle as O


def mode(N, m):
    th = [2 * np.pi * mm / N for mm in m]
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    return th, np.exp(1j * (th[0] * i + th[1] * j + th[2] * k))


def lam4(th, c, nu, dx):
    s = 0
    for d in range(3):
        t = th[d]
        s += -1j * c[d] * (8 * np.sin(t) - np.sin(2 * t)) / (6 * dx)
        s += nu * (-2 * np.cos(2 * t) + 32 * np.cos(t) - 30) / (12 * dx * dx)
    return s


def lam2(th, c, nu, dx, upwind):
    s = 0
    for d in range(3):
        t = th[d]
        if upwind:
            s += (-c[d] / dx * (1 - np.exp(-1j * t))) if c[d] > 0 else (-c[d] / dx * (np.exp(1j * t) - 1))
        else:
            s += -1j * c[d] * np.sin(t) / dx
        s += nu * (2 * np.cos(t) - 2) / dx ** 2
    return s


C3 = (1.0, -0.5, 0.25)


@pytest.mark.parametrize("m", [(1, 2, 3), (3, 0, 5), (7, 7, 1)])
def test_rhs4_fourier_symbol(oracle_lib, m):
    N, nu = 16, 0.013
    dx = 1. / N
    conv = O.OConvection(N, N, N, dx, nu, 0.0, *C3)
    th, v = mode(N, m)
    lam = lam4(th, C3, nu, dx)
    for part, want in ((v.real, (lam * v).real), (v.imag, (lam * v).imag)):
        got = conv.rhs4(np.ascontiguousarray(part), nu)
        assert np.abs(got - want).max() <= 1e-12 * abs(lam)


@pytest.mark.parametrize("upwind", [True, False])
def test_rhs2_fourier_symbol(oracle_lib, upwind):
    N, nu = 12, 0.02
    dx = 1. / N
    conv = O.OConvection(N, N, N, dx, nu, 0.0, *C3, upwind=upwind)
    th, v = mode(N, (1, 4, 2))
    lam = lam2(th, C3, nu, dx, upwind)
    got = conv.rhs2(np.ascontiguousarray(v.real), nu)
    assert np.abs(got - (lam * v).real).max() <= 1e-12 * abs(lam)


def test_rk4_and_euler_amplification(oracle_lib):
    N, nu, dt, n = 16, 0.01, 1e-3, 9
    dx = 1. / N
    conv = O.OConvection(N, N, N, dx, nu, 0.0, *C3)
    th, v = mode(N, (2, 1, 3))
    z = dt * lam4(th, C3, nu, dx)
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    got = conv.rk4(np.ascontiguousarray(v.real), 0.0, dt, n)
    assert np.abs(got - (R ** n * v).real).max() < 1e-13
    # a plausible slip: RK4 truncated after the cubic term would differ by ~|z|^4/24 per step
    R3 = 1 + z + z ** 2 / 2 + z ** 3 / 6
    assert np.abs((R ** n - R3 ** n)).max() > 1e-9
    ze = dt * lam2(th, C3, nu, dx, True)
    got = conv.euler(np.ascontiguousarray(v.real), 0.0, dt, n)
    assert np.abs(got - ((1 + ze) ** n * v).real).max() < 1e-13


def test_nu_evaluated_once_per_step_at_step_start(oracle_lib):
    # DoRK4Timestep computes nu(t) once (Convection.cpp:321); with nufreq != 0 the result of
    # n steps equals the product of per-step amplifications with nu(t0 + i*dt).
    N, nu0, freq, dt, n = 8, 0.05, 40.0, 2e-3, 5
    dx = 1. / N
    conv = O.OConvection(N, N, N, dx, nu0, freq, *C3)
    th, v = mode(N, (1, 2, 1))
    R = 1.0 + 0j
    for i in range(n):
        nu = nu0 * (1 + .5 * math.sin(freq * (0.1 + i * dt)))
        assert conv.computenu(0.1 + i * dt) == nu
        z = dt * lam4(th, C3, nu, dx)
        R *= 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    got = conv.rk4(np.ascontiguousarray(v.real), 0.1, dt, n)
    assert np.abs(got - (R * v).real).max() < 1e-13


def test_steps_zero_leaves_output_untouched_and_aliasing(oracle_lib):
    N = 8
    conv = O.OConvection(N, N, N, 1. / N, 0.01, 0.0, *C3)
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal((N, N, N)), rng.standard_normal((N, N, N))
    fa, fb = conv.field(a), conv.field(b)
    conv.do_rk4(fa, fb, 0.0, 1e-3, 0)
    assert np.array_equal(fb.get(), b)
    conv.do_euler(fa, fb, 0.0, 1e-3, 0)
    assert np.array_equal(fb.get(), b)
    # in == out gives the same bits as in != out
    ref = conv.rk4(a, 0.0, 1e-3, 3)
    fa = conv.field(a)
    conv.do_rk4(fa, fa, 0.0, 1e-3, 3)
    assert np.array_equal(fa.get(), ref)


def test_periodic_fill_bruteforce(oracle_lib):
    # testPeriodicity.cu:11-33 pattern, 4^3 interior with 2-cell halos, checked against
    # modulo indexing (edges and corners included).
    n, h = 4, 2
    f = O.OField(n, n, n)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    f.set((k + 100 * j + 10000 * i).astype(np.float64))
    f.periodic()
    s = f.storage()  # [j][i][k] with halo
    for jj in range(-h, n + h):
        for ii in range(-h, n + h):
            for kk in range(-h, n + h):
                want = (kk % n) + 100 * (jj % n) + 10000 * (ii % n)
                assert s[jj + h, ii + h, kk + h] == want


def test_periodic_fill_generic_strides_stella_layout(oracle_lib):
    # The same restated ApplyImpl on an i-fastest, i-padded buffer (the STELLA CUDA layout).
    ni, nj, nk, h = 5, 3, 4, 2
    pi = 32
    sj, sk = pi, pi * (nj + 2 * h)
    buf = np.full(sk * (nk + 2 * h), -1.0)
    for kk in range(nk):
        for jj in range(nj):
            for ii in range(ni):
                buf[(ii + h) + (jj + h) * sj + (kk + h) * sk] = ii + 10 * jj + 100 * kk
    O.apply_periodic(buf, ni, nj, nk, (h, h, h, h, h, h), (1, sj, sk))
    for kk in range(-h, nk + h):
        for jj in range(-h, nj + h):
            for ii in range(-h, ni + h):
                assert buf[(ii + h) + (jj + h) * sj + (kk + h) * sk] == (ii % ni) + 10 * (jj % nj) + 100 * (kk % nk)
    # padding beyond the halo is untouched
    assert buf[ni + 2 * h] == -1.0


def test_fourth_order_convergence_to_exact_solution(oracle_lib):
    nu0, c, T = 0.1, (1.0, 1.0, 1.0), 0.004
    errs = []
    for N in (8, 16):
        dt = 2e-5
        steps = int(round(T / dt))
        conv = O.OConvection(N, N, N, 1. / N, nu0, 0.0, *c)
        q0 = O.fill_exact(N, nu0, 0.0, *c, 0.0)
        q = conv.rk4(q0, 0.0, dt, steps)
        ex = O.fill_exact(N, nu0, 0.0, *c, steps * dt)
        errs.append(np.abs(q - ex).max() / np.abs(ex).max())
    ratio = errs[0] / errs[1]
    assert 12.0 < ratio < 20.0, (errs, ratio)


def test_fillq_matches_exactq(oracle_lib):
    N = 6
    q = O.fill_exact(N, 0.1, 3.0, 1.0, -0.5, 0.25, 0.3)
    for (k, j, i) in [(0, 0, 0), (5, 2, 1), (3, 4, 5)]:
        x, y, z = (i + .5) / N, (j + .5) / N, (k + .5) / N
        pi = 3.14159265358979
        a = 0.1 * 0.3 - (0.1 / 6.0) * (math.cos(0.9) - 1.)
        want = (math.sin(2 * pi * (x - 0.3)) * math.sin(2 * pi * (y + 0.15)) * math.sin(2 * pi * (z - 0.075))
                * math.exp(-12 * pi * pi * a))
        assert abs(q[k, j, i] - want) < 1e-15


def test_gaussian_fill_periodic_minimum_image(oracle_lib):
    N = 8
    q = O.fill_gaussians(N, [0.05], [0.5], [0.5], [0.1], [2.0])
    # the point at x = 15/16 is 0.1125 away from 0.05 through the periodic boundary
    want = 2.0 * math.exp(-(0.1125 ** 2 + (1 / 16) ** 2 * 2) / (2 * 0.01))
    assert abs(q[4, 4, 7] - want) < 1e-15


def test_error_reference_normalises_by_candidate(oracle_lib):
    f = np.array([1.0, -4.0, 2.0])
    r = np.array([1.5, -4.0, 0.0])
    assert O.error_reference(f, r) == 2.0 / 4.0
    assert abs(O.rel_l2(f, r) - math.sqrt(0.25 + 4.0) / math.sqrt(2.25 + 16.0)) < 1e-16


def _small_parareal(g_kind):
    N, S = 8, 4
    conv = O.OConvection(N, N, N, 1. / N, 0.1, 0.0, 1.0, 1.0, 1.0)
    u0 = O.fill_exact(N, 0.1, 0.0, 1.0, 1.0, 1.0, 0.0)
    T, tsF, tsG = 0.01, 8, 2
    return conv, u0, S, T, tsF, tsG


@pytest.mark.parametrize("g_kind", ["euler", "rk4"])
def test_parareal_exact_after_S_iterations(oracle_lib, g_kind):
    conv, u0, S, T, tsF, tsG = _small_parareal(g_kind)
    ser = conv.serial(u0, T / (S * tsF), S * tsF)
    _, it = conv.parareal(u0, S, S, T, T / (S * tsF), tsF, T / (S * tsG), tsG, g_kind)
    errs = [O.error_reference(it[k], ser) for k in range(S + 1)]
    assert errs[S] < 1e-14
    if g_kind == "euler":
        assert errs[0] > 1e-4           # the first-order coarse prediction is far off
        assert all(errs[k + 1] < errs[k] for k in range(S - 1))
    else:
        assert errs[0] > 1e-11          # RK4 at 4x the fine step: close, not exact
        assert errs[1] < errs[0]


def test_parareal_coarse_equal_fine_converges_in_one_iteration(oracle_lib):
    conv, u0, S, T, tsF, _ = _small_parareal("rk4")
    dtF = T / (S * tsF)
    ser = conv.serial(u0, dtF, S * tsF)
    _, it = conv.parareal(u0, S, 2, T, dtF, tsF, dtF, tsF, "rk4")
    assert O.error_reference(it[0], ser) < 1e-14   # G == F: the prediction is already exact
    assert O.error_reference(it[1], ser) < 1e-14


def test_parareal_first_iterate_by_hand(oracle_lib):
    # Recompute iteration 1 of DoParallel from the propagators alone.
    conv, u0, S, T, tsF, tsG = _small_parareal("euler")
    dtF, dtG, sl = T / (S * tsF), T / (S * tsG), T / S
    U = [u0]
    for r in range(1, S):
        U.append(conv.euler(u0, 0.0, dtG, r * tsG))
    Gold = [conv.euler(U[r], sl * r, dtG, tsG) for r in range(S)]
    F = [conv.rk4(U[r], sl * r, dtF, tsF) for r in range(S)]
    q = None
    for r in range(S):
        if r > 0:
            U[r] = q
        Gnew = conv.euler(U[r], sl * r, dtG, tsG)
        q = Gnew + F[r] - Gold[r]
    _, it = conv.parareal(u0, S, 1, T, dtF, tsF, dtG, tsG, "euler")
    assert np.array_equal(it[1], q)
    assert np.array_equal(it[0], Gold[S - 1])


def test_parareal_async_schedule_times(oracle_lib):
    # The async schedule (Parareal.h:119-124) changes only the fine-step times of ranks
    # r > 0: step 0 at tStart, steps i >= 1 at (tStart + dtF) + (i-1)*dtF instead of
    # tStart + i*dtF.  With constant nu the times never enter: bit-for-bit the blocking
    # result.  With nu(t) the two differ by rounding of t only, far below what a shifted or
    # skipped step would change.
    N, S, T, tsF, tsG = 8, 4, 0.02, 8, 2
    dtF, dtG = T / (S * tsF), T / (S * tsG)
    u0 = O.fill_exact(N, 0.1, 0.0, 1.0, 1.0, 1.0, 0.0)
    c0 = O.OConvection(N, N, N, 1. / N, 0.05, 0.0, 1.0, -0.5, 0.25)
    _, blk = c0.parareal(u0, S, 2, T, dtF, tsF, dtG, tsG, "rk4")
    _, asy = c0.parareal(u0, S, 2, T, dtF, tsF, dtG, tsG, "rk4", async_=True)
    assert np.array_equal(blk, asy)
    c1 = O.OConvection(N, N, N, 1. / N, 0.05, 80.0, 1.0, -0.5, 0.25)
    _, blk = c1.parareal(u0, S, 2, T, dtF, tsF, dtG, tsG, "rk4")
    _, asy = c1.parareal(u0, S, 2, T, dtF, tsF, dtG, tsG, "rk4", async_=True)
    assert O.error_reference(asy[2], blk[2]) < 1e-13
    # a plausible mistake: the second DoRK4 started one step late (tStart + 2 dtF)
    sl = T / S
    U1 = c1.rk4(u0, 0.0, dtG, tsG)  # a start state for rank 1
    good = c1.rk4(c1.rk4(U1, sl, dtF, 1), sl + dtF, dtF, tsF - 1)
    late = c1.rk4(c1.rk4(U1, sl, dtF, 1), sl + 2 * dtF, dtF, tsF - 1)
    assert O.error_reference(late, good) > 1e-9
    assert O.error_reference(good, c1.rk4(U1, sl, dtF, tsF)) < 1e-13
