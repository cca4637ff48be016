/*
 * sp_oracle.c -- CPU restatement of the StencilParareal hot path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the HIP library, never the
 * thing measured or shipped.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.
 *
 * Built with -ffp-contract=off so every expression rounds in the order written
 * here, which is the order of the reference stencil functions.  Each function
 * names the reference file:line it restates (paths relative to the reference
 * repository root).
 *
 * Storage mirrors the reference OpenMP backend (src/Convection/Convection.h:13-20,
 * STELLA/src/shared_infrastructure/SharedInfrastructure.h:76): 2-cell halo on
 * every side, JIK order (k unit stride, then i, then j).  The sweeps run over
 * 8x8 ij blocks with k innermost, like StencilOpenMP::applyImpl
 * (STELLA/src/stencil_framework/StencilOpenMP.h:186-272) and
 * BlockLoopFunctorOpenMP (BlockLoopFunctorOpenMP.h:50-67).
 */
#include "sp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define HALO 2 /* convectionBoundaryLines, src/Convection/Convection.h:13 */
#define BLK 8  /* BlockSize<8,8>, src/Convection/Convection.cpp:15 */

struct orc_field {
    int ni, nj, nk;
    long si, sj, sk; /* element strides: JIK order */
    double* base;    /* storage corner (-2,-2,-2) */
};

static inline double* AT(const orc_field* f, int i, int j, int k)
{
    return f->base + (long)(i + HALO) * f->si + (long)(j + HALO) * f->sj + (long)(k + HALO) * f->sk;
}

static long storage_elems(const orc_field* f)
{
    return (long)(f->ni + 2 * HALO) * (f->nj + 2 * HALO) * (f->nk + 2 * HALO);
}

orc_field* orc_field_new(int ni, int nj, int nk)
{
    orc_field* f = (orc_field*)calloc(1, sizeof(orc_field));
    f->ni = ni;
    f->nj = nj;
    f->nk = nk;
    f->sk = 1;
    f->si = nk + 2 * HALO;
    f->sj = (long)(ni + 2 * HALO) * (nk + 2 * HALO);
    f->base = (double*)calloc((size_t)storage_elems(f), sizeof(double));
    return f;
}

void orc_field_free(orc_field* f)
{
    if (!f) return;
    free(f->base);
    free(f);
}

void orc_field_set(orc_field* f, const double* a)
{
    for (int k = 0; k < f->nk; ++k)
        for (int j = 0; j < f->nj; ++j)
            for (int i = 0; i < f->ni; ++i)
                *AT(f, i, j, k) = a[((long)k * f->nj + j) * f->ni + i];
}

void orc_field_get(const orc_field* f, double* a)
{
    for (int k = 0; k < f->nk; ++k)
        for (int j = 0; j < f->nj; ++j)
            for (int i = 0; i < f->ni; ++i)
                a[((long)k * f->nj + j) * f->ni + i] = *AT(f, i, j, k);
}

void orc_field_get_storage(const orc_field* f, double* s)
{
    memcpy(s, f->base, (size_t)storage_elems(f) * sizeof(double));
}

void orc_field_set_storage(orc_field* f, const double* s)
{
    memcpy(f->base, s, (size_t)storage_elems(f) * sizeof(double));
}

void orc_field_copy(orc_field* dst, const orc_field* src)
{
    memcpy(dst->base, src->base, (size_t)storage_elems(src) * sizeof(double));
}

/* ---- Periodicity: src/Communication/Periodicity.cpp:3-135 ------------------ */

static inline double* acc(double* p, int i, int si, int j, int sj, int k, int sk)
{
    return p + (long)i * si + (long)j * sj + (long)k * sk; /* Periodicity.cpp:3-8 */
}

void orc_apply_periodic(double* p, int ni, int nj, int nk,
                        int him, int hip, int hjm, int hjp, int hkm, int hkp,
                        int si, int sj, int sk)
{
    /* ApplyPeriodicI, Periodicity.cpp:10-38: interior j,k */
    for (int k = hkm; k < hkm + nk; ++k)
        for (int j = hjm; j < hjm + nj; ++j) {
            for (int i = 0; i < him; ++i)
                *acc(p, i, si, j, sj, k, sk) = *acc(p, i + ni, si, j, sj, k, sk);
            for (int i = him; i < him + hip; ++i)
                *acc(p, i + ni, si, j, sj, k, sk) = *acc(p, i, si, j, sj, k, sk);
        }
    /* ApplyPeriodicJ, Periodicity.cpp:40-68: all i, interior k */
    for (int i = 0; i < him + ni + hip; ++i)
        for (int k = hkm; k < hkm + nk; ++k) {
            for (int j = 0; j < hjm; ++j)
                *acc(p, i, si, j, sj, k, sk) = *acc(p, i, si, j + nj, sj, k, sk);
            for (int j = hjm; j < hjm + hjp; ++j)
                *acc(p, i, si, j + nj, sj, k, sk) = *acc(p, i, si, j, sj, k, sk);
        }
    /* ApplyPeriodicK, Periodicity.cpp:70-98: all i, all j */
    for (int i = 0; i < him + ni + hip; ++i)
        for (int j = 0; j < hjm + nj + hjp; ++j) {
            for (int k = 0; k < hkm; ++k)
                *acc(p, i, si, j, sj, k, sk) = *acc(p, i, si, j, sj, k + nk, sk);
            for (int k = hkm; k < hkm + hkp; ++k)
                *acc(p, i, si, j, sj, k + nk, sk) = *acc(p, i, si, j, sj, k, sk);
        }
}

void orc_field_periodic(orc_field* f)
{
    orc_apply_periodic(f->base, f->ni, f->nj, f->nk, HALO, HALO, HALO, HALO, HALO, HALO,
                       (int)f->si, (int)f->sj, (int)f->sk);
}

/* ---- propagator ----------------------------------------------------------- */

struct orc_conv {
    int ni, nj, nk;
    double nu0, nufreq, cx, cy, cz, dx, dx2; /* Convection.h:89 */
    int upwind;
    orc_field *qInternal, *k1, *k2, *k3, *k4; /* Convection.h:94-95 */
};

orc_conv* orc_conv_new(int ni, int nj, int nk, double dx, double nu0, double nufreq,
                       double cx, double cy, double cz, int upwind)
{
    /* Convection::Convection, Convection.cpp:131-143; dx2_(dx*dx) */
    orc_conv* c = (orc_conv*)calloc(1, sizeof(orc_conv));
    c->ni = ni;
    c->nj = nj;
    c->nk = nk;
    c->nu0 = nu0;
    c->nufreq = nufreq;
    c->cx = cx;
    c->cy = cy;
    c->cz = cz;
    c->dx = dx;
    c->dx2 = dx * dx;
    c->upwind = upwind;
    c->qInternal = orc_field_new(ni, nj, nk);
    c->k1 = orc_field_new(ni, nj, nk);
    c->k2 = orc_field_new(ni, nj, nk);
    c->k3 = orc_field_new(ni, nj, nk);
    c->k4 = orc_field_new(ni, nj, nk);
    return c;
}

void orc_conv_free(orc_conv* c)
{
    if (!c) return;
    orc_field_free(c->qInternal);
    orc_field_free(c->k1);
    orc_field_free(c->k2);
    orc_field_free(c->k3);
    orc_field_free(c->k4);
    free(c);
}

/* Convection::computenu, Convection.h:81-84 */
double orc_computenu(const orc_conv* c, double t)
{
    return c->nu0 * (1. + .5 * sin(c->nufreq * t));
}

/* Interior sweep over 8x8 ij blocks, k innermost (StencilOpenMP.h:192-226). */
#define SWEEP_BEGIN(ni, nj, nk)                                                   \
    {                                                                             \
        const int nbi_ = ((ni) + BLK - 1) / BLK, nbj_ = ((nj) + BLK - 1) / BLK;   \
        _Pragma("omp parallel for schedule(static)")                              \
        for (int b_ = 0; b_ < nbi_ * nbj_; ++b_) {                                \
            const int bi_ = b_ % nbi_, bj_ = b_ / nbi_;                           \
            for (int j = bj_ * BLK; j < bj_ * BLK + BLK && j < (nj); ++j)         \
                for (int i = bi_ * BLK; i < bi_ * BLK + BLK && i < (ni); ++i)     \
                    for (int k = 0; k < (nk); ++k) {
#define SWEEP_END \
    }             \
    }             \
    }

/* RHS4Stage (Convection.cpp:57-82) = Advection4 (ConvectionFunctions.h:66-104)
 * + nu * Laplace4 (ConvectionFunctions.h:135-166). */
static void rhs4_sweep(const orc_conv* c, const orc_field* q, orc_field* kf, double nu)
{
    const double cx = c->cx, cy = c->cy, cz = c->cz, dx = c->dx, dx2 = c->dx2;
    SWEEP_BEGIN(c->ni, c->nj, c->nk)
        const double q0 = *AT(q, i, j, k);
        const double qim2 = *AT(q, i - 2, j, k), qim1 = *AT(q, i - 1, j, k);
        const double qip1 = *AT(q, i + 1, j, k), qip2 = *AT(q, i + 2, j, k);
        const double qjm2 = *AT(q, i, j - 2, k), qjm1 = *AT(q, i, j - 1, k);
        const double qjp1 = *AT(q, i, j + 1, k), qjp2 = *AT(q, i, j + 2, k);
        const double qkm2 = *AT(q, i, j, k - 2), qkm1 = *AT(q, i, j, k - 1);
        const double qkp1 = *AT(q, i, j, k + 1), qkp2 = *AT(q, i, j, k + 2);
        /* Advection4 */
        const double dq_di = -1. * qim2 + 8. * qim1 - 8. * qip1 + 1. * qip2;
        const double dq_dj = -1. * qjm2 + 8. * qjm1 - 8. * qjp1 + 1. * qjp2;
        const double dq_dk = -1. * qkm2 + 8. * qkm1 - 8. * qkp1 + 1. * qkp2;
        const double tensadv = (cx * dq_di + cy * dq_dj + cz * dq_dk) / (12. * dx);
        /* Laplace4 */
        const double d2q_di2 = -1. * qim2 + 16. * qim1 - 30. * q0 + 16. * qip1 - 1. * qip2;
        const double d2q_dj2 = -1. * qjm2 + 16. * qjm1 - 30. * q0 + 16. * qjp1 - 1. * qjp2;
        const double d2q_dk2 = -1. * qkm2 + 16. * qkm1 - 30. * q0 + 16. * qkp1 - 1. * qkp2;
        const double lapl = (d2q_di2 + d2q_dj2 + d2q_dk2) / (12. * dx2);
        const double tenslapl = nu * lapl;
        *AT(kf, i, j, k) = tensadv + tenslapl;
    SWEEP_END
}

/* RHS2Stage (Convection.cpp:30-55) = Advection2 (ConvectionFunctions.h:4-64,
 * upwind branch :18-42 / central :44-60) + nu * Laplace2 (:106-132). */
static void rhs2_sweep(const orc_conv* c, const orc_field* q, orc_field* kf, double nu)
{
    const double cx = c->cx, cy = c->cy, cz = c->cz, dx = c->dx, dx2 = c->dx2;
    const int upwind = c->upwind;
    SWEEP_BEGIN(c->ni, c->nj, c->nk)
        const double q0 = *AT(q, i, j, k);
        const double qim1 = *AT(q, i - 1, j, k), qip1 = *AT(q, i + 1, j, k);
        const double qjm1 = *AT(q, i, j - 1, k), qjp1 = *AT(q, i, j + 1, k);
        const double qkm1 = *AT(q, i, j, k - 1), qkp1 = *AT(q, i, j, k + 1);
        double tensadv;
        if (upwind) {
            const double dq_di = cx > 0. ? (q0 - qim1) : (qip1 - q0);
            const double dq_dj = cy > 0. ? (q0 - qjm1) : (qjp1 - q0);
            const double dq_dk = cz > 0. ? (q0 - qkm1) : (qkp1 - q0);
            tensadv = -(cx * dq_di + cy * dq_dj + cz * dq_dk) / dx;
        } else {
            const double dq_di = +1. * qim1 - 1. * qip1;
            const double dq_dj = +1. * qjm1 - 1. * qjp1;
            const double dq_dk = +1. * qkm1 - 1. * qkp1;
            tensadv = (cx * dq_di + cy * dq_dj + cz * dq_dk) / (2. * dx);
        }
        const double d2q_di2 = +1. * qim1 - 2. * q0 + 1. * qip1;
        const double d2q_dj2 = +1. * qjm1 - 2. * q0 + 1. * qjp1;
        const double d2q_dk2 = +1. * qkm1 - 2. * q0 + 1. * qkp1;
        const double lapl = (d2q_di2 + d2q_dj2 + d2q_dk2) / (dx2);
        const double tenslapl = nu * lapl;
        *AT(kf, i, j, k) = tensadv + tenslapl;
    SWEEP_END
}

/* EulerStage, Convection.cpp:84-101 */
static void euler_sweep(const orc_conv* c, const orc_field* qin, const orc_field* kf,
                        orc_field* qout, double dt_)
{
    SWEEP_BEGIN(c->ni, c->nj, c->nk)
        *AT(qout, i, j, k) = *AT(qin, i, j, k) + dt_ * *AT(kf, i, j, k);
    SWEEP_END
}

/* RKStage, Convection.cpp:103-127 */
static void rk_sweep(const orc_conv* c, const orc_field* qin, orc_field* qout, double dt)
{
    const orc_field *k1 = c->k1, *k2 = c->k2, *k3 = c->k3, *k4 = c->k4;
    SWEEP_BEGIN(c->ni, c->nj, c->nk)
        *AT(qout, i, j, k) = *AT(qin, i, j, k) + dt / 6. *
            (*AT(k1, i, j, k) + 2. * *AT(k2, i, j, k) + 2. * *AT(k3, i, j, k) + *AT(k4, i, j, k));
    SWEEP_END
}

/* Convection::DoRK4Timestep, Convection.cpp:318-370 */
void orc_do_rk4_timestep(orc_conv* c, orc_field* in, orc_field* out, double t, double dt)
{
    const double dthalf = dt * .5;
    const double nu = orc_computenu(c, t);
    /* k1 = RHS(in) */
    orc_field_periodic(in);
    rhs4_sweep(c, in, c->k1, nu);
    /* qInternal = in + dt/2 k1; k2 = RHS(qInternal) */
    euler_sweep(c, in, c->k1, c->qInternal, dthalf);
    orc_field_periodic(c->qInternal);
    rhs4_sweep(c, c->qInternal, c->k2, nu);
    /* qInternal = in + dt/2 k2; k3 = RHS(qInternal) */
    euler_sweep(c, in, c->k2, c->qInternal, dthalf);
    orc_field_periodic(c->qInternal);
    rhs4_sweep(c, c->qInternal, c->k3, nu);
    /* qInternal = in + dt k3; k4 = RHS(qInternal) */
    euler_sweep(c, in, c->k3, c->qInternal, dt);
    orc_field_periodic(c->qInternal);
    rhs4_sweep(c, c->qInternal, c->k4, nu);
    /* out = in + dt/6 (k1 + 2k2 + 2k3 + k4) */
    rk_sweep(c, in, out, dt);
}

/* Convection::DoEulerTimestep, Convection.cpp:372-392 */
void orc_do_euler_timestep(orc_conv* c, orc_field* in, orc_field* out, double t, double dt)
{
    const double nu = orc_computenu(c, t);
    orc_field_periodic(in);
    rhs2_sweep(c, in, c->k1, nu);
    euler_sweep(c, in, c->k1, out, dt);
}

/* Convection::DoRK4, Convection.cpp:292-303: steps == 0 leaves out untouched */
void orc_do_rk4(orc_conv* c, orc_field* in, orc_field* out, double tStart, double dt, int steps)
{
    if (steps > 0) orc_do_rk4_timestep(c, in, out, tStart, dt);
    int i = 0;
    for (--steps; steps > 0; --steps) orc_do_rk4_timestep(c, out, out, tStart + (++i) * dt, dt);
}

/* Convection::DoEuler, Convection.cpp:305-316 */
void orc_do_euler(orc_conv* c, orc_field* in, orc_field* out, double tStart, double dt, int steps)
{
    if (steps > 0) orc_do_euler_timestep(c, in, out, tStart, dt);
    int i = 0;
    for (--steps; steps > 0; --steps) orc_do_euler_timestep(c, out, out, tStart + (++i) * dt, dt);
}

void orc_rhs4(orc_conv* c, orc_field* q, orc_field* k, double nu)
{
    orc_field_periodic(q);
    rhs4_sweep(c, q, k, nu);
}

void orc_rhs2(orc_conv* c, orc_field* q, orc_field* k, double nu)
{
    orc_field_periodic(q);
    rhs2_sweep(c, q, k, nu);
}

/* ---- Parareal --------------------------------------------------------------- */

static void coarse(orc_conv* c, int g_kind, orc_field* in, orc_field* out, double t0, double dt, int steps)
{
    if (g_kind == 1)
        orc_do_rk4(c, in, out, t0, dt, steps);
    else
        orc_do_euler(c, in, out, t0, dt, steps);
}

static void coarse_step(orc_conv* c, int g_kind, orc_field* in, orc_field* out, double t, double dt)
{
    if (g_kind == 1)
        orc_do_rk4_timestep(c, in, out, t, dt);
    else
        orc_do_euler_timestep(c, in, out, t, dt);
}

/* UpdateStage, Parareal.h:19-35: q = qcoarsenew + qfine - qcoarseold (interior) */
static void update_sweep(const orc_conv* c, orc_field* q, const orc_field* fine,
                         const orc_field* cnew, const orc_field* cold)
{
    SWEEP_BEGIN(c->ni, c->nj, c->nk)
        *AT(q, i, j, k) = *AT(cnew, i, j, k) + *AT(fine, i, j, k) - *AT(cold, i, j, k);
    SWEEP_END
}

/* Parareal::DoParallel (Parareal.h:110-171), emulated on S ranks in rank order.  Rank r
 * starts at tStart_r = (T/S)*r (src/Parareal/main.cpp:157).  The blocking schedule
 * (--noasync) runs the fine propagation as one DoRK4 (Parareal.h:127); the async one, on
 * ranks r > 0, as one step from tStart and tsF-1 steps from tStartFine = tStart + dtF
 * (Parareal.h:106,121-123), the receive posted in between. */
void orc_parareal_parallel(orc_conv* c, int S, int kmax, double T,
                           double dtF, int tsF, double dtG, int tsG, int g_kind,
                           const double* u0, double* q_iter, double* q_out, int async_sched)
{
    const int ni = c->ni, nj = c->nj, nk = c->nk;
    const long n = (long)ni * nj * nk;
    orc_field** U = (orc_field**)malloc(sizeof(orc_field*) * S);
    orc_field** F = (orc_field**)malloc(sizeof(orc_field*) * S);
    orc_field** Gold = (orc_field**)malloc(sizeof(orc_field*) * S);
    orc_field* Gnew = orc_field_new(ni, nj, nk);
    orc_field* q = orc_field_new(ni, nj, nk);
    orc_field* chain = orc_field_new(ni, nj, nk);
    const double slice = T / S; /* RuntimeConfiguration::timeSliceSize */

    /* Initialisation, Parareal.h:113: U_r = G^{r*tsG}(u0) run from t = 0.  One
     * continuous coarse run visits exactly the states and times of every
     * rank's own run (step j at t = 0 + j*dtG), so rank r's value is the
     * snapshot after r*tsG steps. */
    orc_field_set(chain, u0);
    for (int r = 0; r < S; ++r) {
        U[r] = orc_field_new(ni, nj, nk);
        F[r] = orc_field_new(ni, nj, nk);
        Gold[r] = orc_field_new(ni, nj, nk);
        if (r > 0)
            for (int s = 0; s < tsG; ++s) {
                const long j = (long)(r - 1) * tsG + s;
                coarse_step(c, g_kind, chain, chain, 0. + (double)j * dtG, dtG);
            }
        orc_field_copy(U[r], chain);
    }
    /* Parareal.h:114: qcoarseold = G(U_r) from tStart_r */
    for (int r = 0; r < S; ++r) coarse(c, g_kind, U[r], Gold[r], slice * r, dtG, tsG);
    if (q_iter) orc_field_get(Gold[S - 1], q_iter);

    for (int k = 0; k < kmax; ++k) {
        /* Fine propagation on every rank */
        for (int r = 0; r < S; ++r) {
            if (async_sched && r > 0) {
                orc_do_rk4(c, U[r], F[r], slice * r, dtF, 1);
                orc_do_rk4(c, F[r], F[r], slice * r + dtF, dtF, tsF - 1);
            } else {
                orc_do_rk4(c, U[r], F[r], slice * r, dtF, tsF);
            }
        }
        /* Receive, coarse propagation, update, send, swap (Parareal.h:131-164) */
        for (int r = 0; r < S; ++r) {
            if (r > 0) orc_field_copy(U[r], q);
            coarse(c, g_kind, U[r], Gnew, slice * r, dtG, tsG);
            update_sweep(c, q, F[r], Gnew, Gold[r]);
            orc_field_copy(Gold[r], Gnew);
        }
        if (q_iter) orc_field_get(q, q_iter + (long)(k + 1) * n);
    }
    if (q_out) orc_field_get(kmax > 0 ? q : Gold[S - 1], q_out);

    for (int r = 0; r < S; ++r) {
        orc_field_free(U[r]);
        orc_field_free(F[r]);
        orc_field_free(Gold[r]);
    }
    free(U);
    free(F);
    free(Gold);
    orc_field_free(Gnew);
    orc_field_free(q);
    orc_field_free(chain);
}

/* Parareal::DoSerial, Parareal.h:250-253 */
void orc_parareal_serial(orc_conv* c, double dtF, int steps_total, const double* u0, double* q_out)
{
    orc_field* in = orc_field_new(c->ni, c->nj, c->nk);
    orc_field* out = orc_field_new(c->ni, c->nj, c->nk);
    orc_field_set(in, u0);
    orc_field_set(out, u0);
    orc_do_rk4(c, in, out, 0., dtF, steps_total);
    orc_field_get(out, q_out);
    orc_field_free(in);
    orc_field_free(out);
}

/* ---- norms and initial conditions -------------------------------------------- */

/* computeErrorReference, src/Parareal/main.cpp:98-139 */
double orc_error_reference(const double* f, const double* r, long n)
{
    double error = 0., maxfield = 0.;
    for (long p = 0; p < n; ++p) {
        const double e = fabs(f[p] - r[p]);
        error = error > e ? error : e;
        const double a = fabs(f[p]);
        maxfield = a > maxfield ? a : maxfield;
    }
    return error / maxfield;
}

double orc_rel_l2(const double* f, const double* r, long n)
{
    double num = 0., den = 0.;
    for (long p = 0; p < n; ++p) {
        const double d = f[p] - r[p];
        num += d * d;
        den += r[p] * r[p];
    }
    return sqrt(num) / sqrt(den);
}

/* exactQ, src/Convection/ConvectionSolution.h:8-20 (its truncated pi) */
double orc_exactq(double nu0, double nufreq, double cx, double cy, double cz,
                  double x, double y, double z, double t)
{
    const double pi = 3.14159265358979;
    const double a = nufreq > 0 ? nu0 * t - (nu0 / (2. * nufreq)) * (cos(nufreq * t) - 1.) : nu0 * t;
    return sin(2. * pi * (x - cx * t)) * sin(2. * pi * (y - cy * t)) * sin(2. * pi * (z - cz * t)) *
           exp(-12. * pi * pi * a);
}

/* fillQ, ConvectionSolution.h:23-52, on [0,1]^3 */
void orc_fill_exact(double* a, int ni, int nj, int nk, double nu0, double nufreq,
                    double cx, double cy, double cz, double t)
{
    const double dxhalf = (1. - 0.) / ni / 2.;
    const double dyhalf = (1. - 0.) / nj / 2.;
    const double dzhalf = (1. - 0.) / nk / 2.;
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const double x = 0. + (2 * i + 1) * dxhalf;
                const double y = 0. + (2 * j + 1) * dyhalf;
                const double z = 0. + (2 * k + 1) * dzhalf;
                a[((long)k * nj + j) * ni + i] = orc_exactq(nu0, nufreq, cx, cy, cz, x, y, z, t);
            }
}

static inline double min_image(double u)
{
    return u - floor(u + 0.5);
}

void orc_fill_gaussians(double* a, int ni, int nj, int nk, int ng,
                        const double* xc, const double* yc, const double* zc,
                        const double* sigma, const double* amp)
{
    const double dxhalf = 1. / ni / 2., dyhalf = 1. / nj / 2., dzhalf = 1. / nk / 2.;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const double x = (2 * i + 1) * dxhalf, y = (2 * j + 1) * dyhalf, z = (2 * k + 1) * dzhalf;
                double s = 0.;
                for (int g = 0; g < ng; ++g) {
                    const double ddx = min_image(x - xc[g]);
                    const double ddy = min_image(y - yc[g]);
                    const double ddz = min_image(z - zc[g]);
                    const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    s += amp[g] * exp(-d2 / (2. * sigma[g] * sigma[g]));
                }
                a[((long)k * nj + j) * ni + i] = s;
            }
}

double orc_wtime(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.;
#endif
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
