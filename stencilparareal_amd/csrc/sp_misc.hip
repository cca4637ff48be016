// sp_misc.hip -- halo fill, Parareal update/correction, norms, conversions and
// initial conditions for gfx950.  All are bandwidth-bound element sweeps; the
// reductions use a fixed grid and a fixed in-block tree, so their results do not
// depend on scheduling.

#include "sp_internal.hpp"

namespace sp {
namespace {

// ---- periodic halo fill: ApplyImpl, src/Communication/Periodicity.cpp:100-135 ------------
// phase 0: I faces over interior j,k; phase 1: J faces over all i (halo included),
// interior k; phase 2: K faces over all i, j.  Each phase reads only cells the previous
// phases have completed (launch order), so edges and corners match the reference.
__global__ void __launch_bounds__(256) k_periodic(double* o, int ni, int nj, int nk, int h, long long sj,
                                                  long long sk, int phase)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k, si, sjj, sk2;
    if (phase == 0) {
        const long long n = 2LL * h * nj * nk;
        if (idx >= n) return;
        const int a = (int)(idx % (2 * h));
        j = (int)((idx / (2 * h)) % nj);
        k = (int)(idx / (2LL * h * nj));
        if (a < h) { i = a - h; si = i + ni; }      // F(i) = F(i + ni), storage i in [0,h)
        else { i = ni + (a - h); si = a - h; }      // F(i + ni) = F(i), storage i in [h,2h)
        sjj = j;
        sk2 = k;
    } else if (phase == 1) {
        const long long wi = ni + 2 * h;
        const long long n = wi * 2 * h * nk;
        if (idx >= n) return;
        i = (int)(idx % wi) - h;
        const int a = (int)((idx / wi) % (2 * h));
        k = (int)(idx / (wi * 2 * h));
        if (a < h) { j = a - h; sjj = j + nj; }
        else { j = nj + (a - h); sjj = a - h; }
        si = i;
        sk2 = k;
    } else {
        const long long wi = ni + 2 * h, wj = nj + 2 * h;
        const long long n = wi * wj * 2 * h;
        if (idx >= n) return;
        i = (int)(idx % wi) - h;
        j = (int)((idx / wi) % wj) - h;
        const int a = (int)(idx / (wi * wj));
        if (a < h) { k = a - h; sk2 = k + nk; }
        else { k = nk + (a - h); sk2 = a - h; }
        si = i;
        sjj = j;
    }
    o[i + j * sj + k * sk] = o[si + sjj * sj + sk2 * sk];
}

// ---- UpdateStage, src/Parareal/Parareal.h:19-35 ----------------------------------------------
__global__ void __launch_bounds__(256) k_update(int ni, int nj, int nk, View q, View fine, View cnew, View cold)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= ni || j >= nj) return;
    const double a = ((const double*)cnew.p)[i + j * cnew.sj + k * cnew.sk];
    const double b = ((const double*)fine.p)[i + j * fine.sj + k * fine.sk];
    const double c = ((const double*)cold.p)[i + j * cold.sj + k * cold.sk];
    ((double*)q.p)[i + j * q.sj + k * q.sk] = a + b - c;
}

// ---- block reduction helpers ---------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], const bool (&is_max)[NV], double* out,
                                             int block = -1)
{
    __shared__ double sh[NV][256];
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[q][t] = v[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int q = 0; q < NV; ++q)
                sh[q][t] = is_max[q] ? fmax(sh[q][t], sh[q][t + s]) : sh[q][t] + sh[q][t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) out[(block < 0 ? (int)blockIdx.x : block) * NV + q] = sh[q][0];
    }
}

// Parareal correction for one slice (Parareal.h:143-164):
//   q = qcoarsenew + qfine - qcoarseold   (f64, reference order),
//   qcoarseold <- qcoarsenew              (the swap of :164, done as a copy so that
//                                          buffer addresses stay fixed for graph replay),
//   dst <- q, and per-block partial sums of (q - dst_old)^2 and q^2.
template <typename SG>
__global__ void __launch_bounds__(256) k_correction(int ni, int nj, int nk, View cnew, View cold, View fine, View dst,
                                                    double* partials)
{
#pragma clang fp contract(off)
    const long long n = (long long)ni * nj * nk;
    double d2 = 0., q2 = 0.;
    for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long long)gridDim.x * 256) {
        const int i = (int)(p % ni), j = (int)((p / ni) % nj), k = (int)(p / ((long long)ni * nj));
        const SG* cn = (const SG*)cnew.p + i + j * cnew.sj + k * cnew.sk;
        SG* co = (SG*)cold.p + i + j * cold.sj + k * cold.sk;
        const double a = (double)Store<SG>::load(cn);
        const double c = (double)Store<SG>::load(co);
        const double b = ((const double*)fine.p)[i + j * fine.sj + k * fine.sk];
        const double q = a + b - c;
        double* d = (double*)dst.p + i + j * dst.sj + k * dst.sk;
        const double diff = q - *d;
        *d = q;
        *co = *cn;
        d2 += diff * diff;
        q2 += q * q;
    }
    double v[2] = {d2, q2};
    const bool mx[2] = {false, false};
    block_reduce<2>(v, mx, partials);
}

// Correction in difference form, for a linear coarse propagator G (this one is: the
// operator is linear in q and nu(t) does not depend on q), where
//   G(U_n^{k+1}) - G(U_n^k) = G(U_n^{k+1} - U_n^k) = G(d_n):
//   q = fine + G(d_n)  (gd = G(d_n) in the coarse storage type; a null gd.p stands for
//   G(0) = 0, slice 0, whose start state never changes), d_{n+1} = q - dst_old into dnext
//   (skipped when dnext.p is null), dst <- q, and partial sums of d_{n+1}^2 and q^2.
// Rounding of a reduced-precision G then scales with |d|, which vanishes as Parareal
// converges, instead of with |U|, which does not.
template <typename SG>
__device__ __forceinline__ void correction_delta_body(int ni, int nj, int nk, const View& gd, const View& fine,
                                                      const View& dst, const View& dnext, double* partials, int b,
                                                      int nb)
{
#pragma clang fp contract(off)
    const long long n = (long long)ni * nj * nk;
    double d2 = 0., q2 = 0.;
    for (long long p = (long long)b * 256 + threadIdx.x; p < n; p += (long long)nb * 256) {
        const int i = (int)(p % ni), j = (int)((p / ni) % nj), k = (int)(p / ((long long)ni * nj));
        const double g = gd.p ? (double)Store<SG>::load((const SG*)gd.p + i + j * gd.sj + k * gd.sk) : 0.;
        const double bf = ((const double*)fine.p)[i + j * fine.sj + k * fine.sk];
        const double q = bf + g;
        double* d = (double*)dst.p + i + j * dst.sj + k * dst.sk;
        const double diff = q - *d;
        *d = q;
        if (dnext.p) ((double*)dnext.p)[i + j * dnext.sj + k * dnext.sk] = diff;
        d2 += diff * diff;
        q2 += q * q;
    }
    double v[2] = {d2, q2};
    const bool mx[2] = {false, false};
    block_reduce<2>(v, mx, partials, b);
}

template <typename SG>
__global__ void __launch_bounds__(256) k_correction_delta(int ni, int nj, int nk, View gd, View fine, View dst,
                                                          View dnext, double* partials)
{
    correction_delta_body<SG>(ni, nj, nk, gd, fine, dst, dnext, partials, blockIdx.x, gridDim.x);
}

// The same for several slices of independent problems in one launch: record e owns blocks
// [e*nb, (e+1)*nb).  Records are read through the constant address space (scalar loads).
template <typename SG>
__global__ void __launch_bounds__(256) k_correction_delta_grouped(const CorrRecord* recs, int nb)
{
    typedef __attribute__((address_space(4))) const CorrRecord CRec;
    const int e = blockIdx.x / nb;
    const CorrRecord& r = *(const CorrRecord*)((CRec*)recs + e);
    correction_delta_body<SG>(r.ni, r.nj, r.nk, r.gd, r.fine, r.dst, r.dnext, r.partials, blockIdx.x % nb, nb);
}

__global__ void __launch_bounds__(256) k_norms(int ni, int nj, int nk, View a, View b, double* partials)
{
    const long long n = (long long)ni * nj * nk;
    double emax = 0., amax = 0., d2 = 0., b2 = 0.;
    for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long long)gridDim.x * 256) {
        const int i = (int)(p % ni), j = (int)((p / ni) % nj), k = (int)(p / ((long long)ni * nj));
        const double x = ((const double*)a.p)[i + j * a.sj + k * a.sk];
        const double y = ((const double*)b.p)[i + j * b.sj + k * b.sk];
        const double d = x - y;
        emax = fmax(emax, fabs(d));
        amax = fmax(amax, fabs(x));
        d2 += d * d;
        b2 += y * y;
    }
    double v[4] = {emax, amax, d2, b2};
    const bool mx[4] = {true, true, false, false};
    block_reduce<4>(v, mx, partials);
}

template <typename SI, typename SO>
__global__ void __launch_bounds__(256) k_convert(int ni, int nj, int nk, View in, View out)
{
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int k = blockIdx.z % nk, s = blockIdx.z / nk;
    if (i >= ni || j >= nj) return;
    const SI* src = (const SI*)in.p + s * in.slice + i + j * in.sj + k * in.sk;
    SO* dst = (SO*)out.p + s * out.slice + i + j * out.sj + k * out.sk;
    Store<SO>::store(dst, Store<SI>::load(src));
}

// exactQ / fillQ, src/Convection/ConvectionSolution.h:8-52 (its truncated pi).
__global__ void __launch_bounds__(256) k_fill_exact(int ni, int nj, int nk, View o, double nu0, double nufreq,
                                                    double cx, double cy, double cz, double t)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= ni || j >= nj) return;
    const double pi = 3.14159265358979;
    const double dxh = (1. - 0.) / ni / 2., dyh = (1. - 0.) / nj / 2., dzh = (1. - 0.) / nk / 2.;
    const double x = 0. + (2 * i + 1) * dxh, y = 0. + (2 * j + 1) * dyh, z = 0. + (2 * k + 1) * dzh;
    const double a = nufreq > 0 ? nu0 * t - (nu0 / (2. * nufreq)) * (cos(nufreq * t) - 1.) : nu0 * t;
    ((double*)o.p)[i + j * o.sj + k * o.sk] =
        sin(2. * pi * (x - cx * t)) * sin(2. * pi * (y - cy * t)) * sin(2. * pi * (z - cz * t)) *
        exp(-12. * pi * pi * a);
}

__global__ void __launch_bounds__(256) k_fill_gauss(int ni, int nj, int nk, View o, int ng, const double* g)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= ni || j >= nj) return;
    const double x = (2 * i + 1) * (1. / ni / 2.), y = (2 * j + 1) * (1. / nj / 2.), z = (2 * k + 1) * (1. / nk / 2.);
    double s = 0.;
    for (int q = 0; q < ng; ++q) {
        double ddx = x - g[q], ddy = y - g[ng + q], ddz = z - g[2 * ng + q];
        ddx = ddx - floor(ddx + 0.5);
        ddy = ddy - floor(ddy + 0.5);
        ddz = ddz - floor(ddz + 0.5);
        const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
        const double sg = g[3 * ng + q];
        s += g[4 * ng + q] * exp(-d2 / (2. * sg * sg));
    }
    ((double*)o.p)[i + j * o.sj + k * o.sk] = s;
}

}  // namespace

int launch_periodic_fill(hipStream_t s, const sp_layout& l, void* base)
{
    double* o = (double*)base + l.origin;
    const int h = l.halo;
    const long long n0 = 2LL * h * l.nj * l.nk;
    const long long n1 = (long long)(l.ni + 2 * h) * 2 * h * l.nk;
    const long long n2 = (long long)(l.ni + 2 * h) * (l.nj + 2 * h) * 2 * h;
    const long long ns[3] = {n0, n1, n2};
    for (int ph = 0; ph < 3; ++ph) {
        const unsigned blocks = (unsigned)((ns[ph] + 255) / 256);
        hipLaunchKernelGGL(k_periodic, dim3(blocks), dim3(256), 0, s, o, l.ni, l.nj, l.nk, h, l.sj, l.sk, ph);
        SP_HIP(hipGetLastError());
    }
    return SP_OK;
}

int launch_update(hipStream_t s, int ni, int nj, int nk, View q, View fine, View cnew, View cold)
{
    hipLaunchKernelGGL(k_update, dim3((ni + 63) / 64, (nj + 3) / 4, nk), dim3(64, 4), 0, s, ni, nj, nk, q, fine,
                       cnew, cold);
    SP_HIP(hipGetLastError());
    return SP_OK;
}

int launch_convert(hipStream_t s, int ni, int nj, int nk, int nslices, int idt, View in, int odt, View out)
{
    const dim3 g((ni + 63) / 64, (nj + 3) / 4, nk * nslices), b(64, 4);
#define SP_CONV(A, B, TA, TB)                                                           \
    if (idt == A && odt == B) {                                                         \
        hipLaunchKernelGGL((k_convert<TA, TB>), g, b, 0, s, ni, nj, nk, in, out);       \
        SP_HIP(hipGetLastError());                                                      \
        return SP_OK;                                                                   \
    }
    SP_CONV(SP_F64, SP_F64, double, double)
    SP_CONV(SP_F64, SP_F32, double, float)
    SP_CONV(SP_F64, SP_BF16, double, bf16)
    SP_CONV(SP_F32, SP_F64, float, double)
    SP_CONV(SP_F32, SP_F32, float, float)
    SP_CONV(SP_BF16, SP_F64, bf16, double)
    SP_CONV(SP_BF16, SP_BF16, bf16, bf16)
#undef SP_CONV
    return fail(SP_ERR_UNSUPPORTED, "convert: unsupported dtype pair");
}

int launch_correction(hipStream_t s, int ni, int nj, int nk, int g_dtype, View cnew, View cold, View fine, View dst,
                      double* partials, int nblocks)
{
    if (g_dtype == SP_F64)
        hipLaunchKernelGGL(k_correction<double>, dim3(nblocks), dim3(256), 0, s, ni, nj, nk, cnew, cold, fine, dst,
                           partials);
    else if (g_dtype == SP_F32)
        hipLaunchKernelGGL(k_correction<float>, dim3(nblocks), dim3(256), 0, s, ni, nj, nk, cnew, cold, fine, dst,
                           partials);
    else
        hipLaunchKernelGGL(k_correction<bf16>, dim3(nblocks), dim3(256), 0, s, ni, nj, nk, cnew, cold, fine, dst,
                           partials);
    SP_HIP(hipGetLastError());
    return SP_OK;
}

int launch_correction_delta(hipStream_t s, int ni, int nj, int nk, int g_dtype, View gd, View fine, View dst,
                            View dnext, double* partials, int nblocks)
{
    if (g_dtype == SP_F64)
        hipLaunchKernelGGL(k_correction_delta<double>, dim3(nblocks), dim3(256), 0, s, ni, nj, nk, gd, fine, dst,
                           dnext, partials);
    else if (g_dtype == SP_F32)
        hipLaunchKernelGGL(k_correction_delta<float>, dim3(nblocks), dim3(256), 0, s, ni, nj, nk, gd, fine, dst,
                           dnext, partials);
    else
        hipLaunchKernelGGL(k_correction_delta<bf16>, dim3(nblocks), dim3(256), 0, s, ni, nj, nk, gd, fine, dst,
                           dnext, partials);
    SP_HIP(hipGetLastError());
    return SP_OK;
}

int launch_correction_delta_grouped(hipStream_t s, int g_dtype, const CorrRecord* recs, int n, int nb)
{
    if (n <= 0) return SP_OK;
    const dim3 grid((unsigned)(n * nb));
    if (g_dtype == SP_F64)
        hipLaunchKernelGGL(k_correction_delta_grouped<double>, grid, dim3(256), 0, s, recs, nb);
    else if (g_dtype == SP_F32)
        hipLaunchKernelGGL(k_correction_delta_grouped<float>, grid, dim3(256), 0, s, recs, nb);
    else
        hipLaunchKernelGGL(k_correction_delta_grouped<bf16>, grid, dim3(256), 0, s, recs, nb);
    SP_HIP(hipGetLastError());
    return SP_OK;
}

int launch_norms(hipStream_t s, int ni, int nj, int nk, View a, View b, double* partials, int nblocks)
{
    hipLaunchKernelGGL(k_norms, dim3(nblocks), dim3(256), 0, s, ni, nj, nk, a, b, partials);
    SP_HIP(hipGetLastError());
    return SP_OK;
}

int launch_fill_exact(hipStream_t s, const sp_layout& l, void* base, double nu0, double nufreq, double cx, double cy,
                      double cz, double t)
{
    const View o = view_of(l, base);
    hipLaunchKernelGGL(k_fill_exact, dim3((l.ni + 63) / 64, (l.nj + 3) / 4, l.nk), dim3(64, 4), 0, s, l.ni, l.nj,
                       l.nk, o, nu0, nufreq, cx, cy, cz, t);
    SP_HIP(hipGetLastError());
    return SP_OK;
}

int launch_fill_gaussians(hipStream_t s, const sp_layout& l, void* base, int ng, const double* gdev)
{
    const View o = view_of(l, base);
    hipLaunchKernelGGL(k_fill_gauss, dim3((l.ni + 63) / 64, (l.nj + 3) / 4, l.nk), dim3(64, 4), 0, s, l.ni, l.nj,
                       l.nk, o, ng, gdev);
    SP_HIP(hipGetLastError());
    return SP_OK;
}

}  // namespace sp
