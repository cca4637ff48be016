// sp_staged.hip -- reference-structured FP64 sweeps for gfx950.
//
// One kernel per reference stencil stage, each evaluating its expression in the
// operation order of the reference stencil functions with floating-point
// contraction disabled (this file is compiled with -ffp-contract=off and every
// kernel repeats `#pragma clang fp contract(off)`), so that on identical inputs
// the results are bit-for-bit those of the reference sweeps:
//   RHS4Stage   src/Convection/Convection.cpp:57-82  (Advection4 ConvectionFunctions.h:66-104,
//                                                      Laplace4 :135-166)
//   RHS2Stage   src/Convection/Convection.cpp:30-55  (Advection2 :4-64, Laplace2 :106-132)
//   EulerStage  src/Convection/Convection.cpp:84-101
//   RKStage     src/Convection/Convection.cpp:103-127
// Periodicity is index arithmetic: a neighbour at i-2 of i = 0 reads i = ni-2,
// the value the reference's halo fill (Periodicity.cpp:100-135) copies there
// for ni >= 2.
//
// Launch shape: 64 x 4 threads per block, one lattice point per thread, grid
// (ceil(ni/64), ceil(nj/4), nk * nslices).  Each wave reads 64 consecutive i
// (512 B), so every row access is one coalesced segment.

#include "sp_internal.hpp"

namespace sp {
namespace {

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

struct Pt {
    int i, j, k, slice;
    bool ok;
};

__device__ __forceinline__ Pt point(int ni, int nj, int nk)
{
    Pt p;
    p.i = blockIdx.x * 64 + threadIdx.x;
    p.j = blockIdx.y * 4 + threadIdx.y;
    p.k = blockIdx.z % nk;
    p.slice = blockIdx.z / nk;
    p.ok = p.i < ni && p.j < nj;
    return p;
}

__device__ __forceinline__ const double* cptr(const View& v, const Pt& p)
{
    return (const double*)v.p + p.slice * v.slice;
}
__device__ __forceinline__ double* mptr(const View& v, const Pt& p) { return (double*)v.p + p.slice * v.slice; }

__global__ void __launch_bounds__(256) k_rhs4(StagedParams P, View q, View kf, double nu)
{
#pragma clang fp contract(off)
    const Pt p = point(P.ni, P.nj, P.nk);
    if (!p.ok) return;
    const double* Q = cptr(q, p);
    const long long sj = q.sj, sk = q.sk;
    const int i = p.i, j = p.j, k = p.k;
    const int im2 = wrap(i - 2, P.ni), im1 = wrap(i - 1, P.ni), ip1 = wrap(i + 1, P.ni), ip2 = wrap(i + 2, P.ni);
    const int jm2 = wrap(j - 2, P.nj), jm1 = wrap(j - 1, P.nj), jp1 = wrap(j + 1, P.nj), jp2 = wrap(j + 2, P.nj);
    const int km2 = wrap(k - 2, P.nk), km1 = wrap(k - 1, P.nk), kp1 = wrap(k + 1, P.nk), kp2 = wrap(k + 2, P.nk);
    const long long r = j * sj + k * sk;
    const double q0 = Q[r + i];
    const double qim2 = Q[r + im2], qim1 = Q[r + im1], qip1 = Q[r + ip1], qip2 = Q[r + ip2];
    const long long ck = k * sk + i;
    const double qjm2 = Q[ck + jm2 * sj], qjm1 = Q[ck + jm1 * sj], qjp1 = Q[ck + jp1 * sj], qjp2 = Q[ck + jp2 * sj];
    const long long cj = j * sj + i;
    const double qkm2 = Q[cj + km2 * sk], qkm1 = Q[cj + km1 * sk], qkp1 = Q[cj + kp1 * sk], qkp2 = Q[cj + kp2 * sk];
    // Advection4
    const double dq_di = -1. * qim2 + 8. * qim1 - 8. * qip1 + 1. * qip2;
    const double dq_dj = -1. * qjm2 + 8. * qjm1 - 8. * qjp1 + 1. * qjp2;
    const double dq_dk = -1. * qkm2 + 8. * qkm1 - 8. * qkp1 + 1. * qkp2;
    const double tensadv = (P.cx * dq_di + P.cy * dq_dj + P.cz * dq_dk) / (12. * P.dx);
    // Laplace4
    const double d2q_di2 = -1. * qim2 + 16. * qim1 - 30. * q0 + 16. * qip1 - 1. * qip2;
    const double d2q_dj2 = -1. * qjm2 + 16. * qjm1 - 30. * q0 + 16. * qjp1 - 1. * qjp2;
    const double d2q_dk2 = -1. * qkm2 + 16. * qkm1 - 30. * q0 + 16. * qkp1 - 1. * qkp2;
    const double lapl = (d2q_di2 + d2q_dj2 + d2q_dk2) / (12. * P.dx2);
    const double tenslapl = nu * lapl;
    mptr(kf, p)[i + j * kf.sj + k * kf.sk] = tensadv + tenslapl;
}

__device__ __forceinline__ double rhs2_point(const StagedParams& P, const double* Q, long long sj, long long sk,
                                             int i, int j, int k, double nu, int upwind)
{
#pragma clang fp contract(off)
    const int im1 = wrap(i - 1, P.ni), ip1 = wrap(i + 1, P.ni);
    const int jm1 = wrap(j - 1, P.nj), jp1 = wrap(j + 1, P.nj);
    const int km1 = wrap(k - 1, P.nk), kp1 = wrap(k + 1, P.nk);
    const long long r = j * sj + k * sk;
    const double q0 = Q[r + i];
    const double qim1 = Q[r + im1], qip1 = Q[r + ip1];
    const double qjm1 = Q[i + jm1 * sj + k * sk], qjp1 = Q[i + jp1 * sj + k * sk];
    const double qkm1 = Q[i + j * sj + km1 * sk], qkp1 = Q[i + j * sj + kp1 * sk];
    double tensadv;
    if (upwind) {
        const double dq_di = P.cx > 0. ? (q0 - qim1) : (qip1 - q0);
        const double dq_dj = P.cy > 0. ? (q0 - qjm1) : (qjp1 - q0);
        const double dq_dk = P.cz > 0. ? (q0 - qkm1) : (qkp1 - q0);
        tensadv = -(P.cx * dq_di + P.cy * dq_dj + P.cz * dq_dk) / P.dx;
    } else {
        const double dq_di = +1. * qim1 - 1. * qip1;
        const double dq_dj = +1. * qjm1 - 1. * qjp1;
        const double dq_dk = +1. * qkm1 - 1. * qkp1;
        tensadv = (P.cx * dq_di + P.cy * dq_dj + P.cz * dq_dk) / (2. * P.dx);
    }
    const double d2q_di2 = +1. * qim1 - 2. * q0 + 1. * qip1;
    const double d2q_dj2 = +1. * qjm1 - 2. * q0 + 1. * qjp1;
    const double d2q_dk2 = +1. * qkm1 - 2. * q0 + 1. * qkp1;
    const double lapl = (d2q_di2 + d2q_dj2 + d2q_dk2) / (P.dx2);
    const double tenslapl = nu * lapl;
    return tensadv + tenslapl;
}

__global__ void __launch_bounds__(256) k_rhs2(StagedParams P, View q, View kf, double nu, int upwind)
{
#pragma clang fp contract(off)
    const Pt p = point(P.ni, P.nj, P.nk);
    if (!p.ok) return;
    const double r = rhs2_point(P, cptr(q, p), q.sj, q.sk, p.i, p.j, p.k, nu, upwind);
    mptr(kf, p)[p.i + p.j * kf.sj + p.k * kf.sk] = r;
}

// EulerStage: out = q + dt*k.  DoEulerTimestep (Convection.cpp:372-392) as one
// pass: k is the RHS2 value of this point, then the Euler update.
__global__ void __launch_bounds__(256) k_euler_fused(StagedParams P, View qin, View qout, double nu, double dt,
                                                     int upwind)
{
#pragma clang fp contract(off)
    const Pt p = point(P.ni, P.nj, P.nk);
    if (!p.ok) return;
    const double* Q = cptr(qin, p);
    const double k = rhs2_point(P, Q, qin.sj, qin.sk, p.i, p.j, p.k, nu, upwind);
    const double q0 = Q[p.i + p.j * qin.sj + p.k * qin.sk];
    mptr(qout, p)[p.i + p.j * qout.sj + p.k * qout.sk] = q0 + dt * k;
}

__global__ void __launch_bounds__(256) k_euler_stage(StagedParams P, View qin, View kf, View qout, double dt)
{
#pragma clang fp contract(off)
    const Pt p = point(P.ni, P.nj, P.nk);
    if (!p.ok) return;
    const double q0 = cptr(qin, p)[p.i + p.j * qin.sj + p.k * qin.sk];
    const double k = cptr(kf, p)[p.i + p.j * kf.sj + p.k * kf.sk];
    const double dt_ = dt;
    mptr(qout, p)[p.i + p.j * qout.sj + p.k * qout.sk] = q0 + dt_ * k;
}

__global__ void __launch_bounds__(256) k_rk_stage(StagedParams P, View qin, View k1, View k2, View k3, View k4,
                                                  View qout, double dt)
{
#pragma clang fp contract(off)
    const Pt p = point(P.ni, P.nj, P.nk);
    if (!p.ok) return;
    const long long o = p.i + p.j * qin.sj + p.k * qin.sk;
    const long long ok = p.i + p.j * k1.sj + p.k * k1.sk;
    const double q0 = cptr(qin, p)[o];
    const double a1 = cptr(k1, p)[ok], a2 = cptr(k2, p)[ok], a3 = cptr(k3, p)[ok], a4 = cptr(k4, p)[ok];
    mptr(qout, p)[p.i + p.j * qout.sj + p.k * qout.sk] = q0 + dt / 6. * (a1 + 2. * a2 + 2. * a3 + a4);
}

dim3 grid_of(const StagedParams& P, int nslices)
{
    return dim3((P.ni + 63) / 64, (P.nj + 3) / 4, P.nk * nslices);
}

}  // namespace

int launch_rhs4(hipStream_t s, const StagedParams& P, View q, View k, double nu)
{
    hipLaunchKernelGGL(k_rhs4, grid_of(P, 1), dim3(64, 4), 0, s, P, q, k, nu);
    SP_HIP(hipGetLastError());
    return SP_OK;
}

int launch_rhs2(hipStream_t s, const StagedParams& P, View q, View k, double nu, int upwind)
{
    hipLaunchKernelGGL(k_rhs2, grid_of(P, 1), dim3(64, 4), 0, s, P, q, k, nu, upwind);
    SP_HIP(hipGetLastError());
    return SP_OK;
}

int launch_euler_stage(hipStream_t s, const StagedParams& P, View qin, View k, View qout, double dt)
{
    hipLaunchKernelGGL(k_euler_stage, grid_of(P, 1), dim3(64, 4), 0, s, P, qin, k, qout, dt);
    SP_HIP(hipGetLastError());
    return SP_OK;
}

int launch_rk_stage(hipStream_t s, const StagedParams& P, View qin, View k1, View k2, View k3, View k4, View qout,
                    double dt)
{
    hipLaunchKernelGGL(k_rk_stage, grid_of(P, 1), dim3(64, 4), 0, s, P, qin, k1, k2, k3, k4, qout, dt);
    SP_HIP(hipGetLastError());
    return SP_OK;
}

int launch_euler_fused(hipStream_t s, const StagedParams& P, View qin, View qout, double nu, double dt, int upwind,
                       int nslices)
{
    hipLaunchKernelGGL(k_euler_fused, grid_of(P, nslices), dim3(64, 4), 0, s, P, qin, qout, nu, dt, upwind);
    SP_HIP(hipGetLastError());
    return SP_OK;
}

}  // namespace sp
