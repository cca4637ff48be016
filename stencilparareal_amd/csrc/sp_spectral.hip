// sp_spectral.hip -- the coarse propagators applied as their exact Fourier multipliers.
//
// On the periodic grid the stencil operators are diagonal in the discrete Fourier basis:
// L e^{i theta.x} = lambda(theta) e^{i theta.x}.  For the RK4 propagator (Advection4 +
// nu Laplace4, ConvectionFunctions.h:66-104 and :135-166)
//   lambda = sum_d [ -i c_d (8 sin th_d - sin 2 th_d) / (6 dx)
//                    + nu (-2 cos 2 th_d + 32 cos th_d - 30) / (12 dx^2) ],
// and for the Euler propagator (Advection2 + nu Laplace2, ConvectionFunctions.h:4-64 and
// :106-132)
//   lambda = sum_d [ -|c_d| (1 - cos th_d) / dx - i c_d sin th_d / dx      (upwind)
//                    | -i c_d sin th_d / dx                                 (central)
//                    + nu (2 cos th_d - 2) / dx^2 ].
// (Upwind: c > 0 takes q0 - q(-1), whose symbol is (1 - cos th) + i sin th; otherwise
// q(+1) - q0, (cos th - 1) + i sin th; both times -c / dx.)  One fused RK4 step multiplies
// each mode by R(z) = 1 + z (1 + z/2 (1 + z/3 (1 + z/4))), z = h lambda, the polynomial of
// the nested stages; one Euler step (DoEulerTimestep, Convection.cpp:372-392) by 1 + z.  A
// coarse segment of n sub-steps multiplies by the product of its n factors (nu(t) may
// change from step to step, never in space).  So a whole segment G is one real-to-complex
// FFT (hipFFT), one pointwise product, one inverse FFT: the same linear map as n stencil
// steps, to rounding, at a cost that does not grow with n.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <string>

#include "sp_internal.hpp"

namespace sp {

struct SpectralPlan {
    int ni = 0, nj = 0, nk = 0;
    hipfftHandle fwd = 0, inv = 0;
    double2* X = nullptr;  // nk x nj x (ni/2 + 1) complex coefficients
    double2* M = nullptr;  // the multiplier of the last constant-nu segment (scale included)
    SpectralStep m_key{};  // ... and the segment it belongs to
    bool m_valid = false;
};

namespace {

const char* fft_err(hipfftResult r)
{
    switch (r) {
    case HIPFFT_SUCCESS: return "success";
    case HIPFFT_INVALID_PLAN: return "invalid plan";
    case HIPFFT_ALLOC_FAILED: return "allocation failed";
    case HIPFFT_INVALID_VALUE: return "invalid value";
    case HIPFFT_EXEC_FAILED: return "execution failed";
    case HIPFFT_SETUP_FAILED: return "setup failed";
    case HIPFFT_INVALID_SIZE: return "invalid size";
    default: return "error";
    }
}

#define SP_FFT(call)                                                                                  \
    do {                                                                                              \
        const hipfftResult r_ = (call);                                                               \
        if (r_ != HIPFFT_SUCCESS) return fail(SP_ERR_HIP, std::string("hipFFT: ") + fft_err(r_) + " (" #call ")"); \
    } while (0)

struct Cplx {
    double re, im;
};
__device__ __forceinline__ Cplx cmul(Cplx a, Cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// R(z) in the nested form of the fused stages
__device__ __forceinline__ Cplx rk4_factor(double zr, double zi)
{
    Cplx p{1. + zr * 0.25, zi * 0.25};                      // 1 + z/4
    Cplx t = cmul({zr / 3., zi / 3.}, p);                   // z/3 (1 + z/4)
    p = {1. + t.re, t.im};
    t = cmul({zr * 0.5, zi * 0.5}, p);
    p = {1. + t.re, t.im};
    t = cmul({zr, zi}, p);
    return {1. + t.re, t.im};
}

// The sub-step's factor at z = h lambda
__device__ __forceinline__ Cplx step_factor(int kind, double zr, double zi)
{
    return kind == kSpectralRK4 ? rk4_factor(zr, zi) : Cplx{1. + zr, zi};
}

// One thread per coefficient: the multiplier prod_i F(h lambda(theta; nu_i)) times the
// inverse transform's 1/N, applied to X (M == nullptr) or stored in M (X == nullptr).
// lambda = a + nu b (a: advection, b: Laplacian).  nu_tab == nullptr: constant nu, the
// product by binary powering.
__global__ void __launch_bounds__(256) k_spectral_multiplier(double2* __restrict__ X, double2* __restrict__ M, int ni,
                                                             int nj, int nk, SpectralStep st, double scale)
{
    const int nic = ni / 2 + 1;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)nic * nj * nk;
    if (idx >= total) return;
    const int i = (int)(idx % nic);
    const int j = (int)((idx / nic) % nj);
    const int k = (int)(idx / ((long long)nic * nj));
    const double two_pi = 6.283185307179586476925286766559;
    const double th[3] = {two_pi * i / ni, two_pi * j / nj, two_pi * k / nk};
    const double c[3] = {st.cx, st.cy, st.cz};
    double a_re = 0., a_im = 0., b = 0.;
    for (int d = 0; d < 3; ++d) {
        double s1, c1;
        sincos(th[d], &s1, &c1);
        if (st.kind == kSpectralRK4) {
            double s2, c2;
            sincos(2. * th[d], &s2, &c2);
            a_im -= c[d] * (8. * s1 - s2) / (6. * st.dx);
            b += (-2. * c2 + 32. * c1 - 30.) / (12. * st.dx * st.dx);
        } else {
            if (st.kind == kSpectralEulerUpwind) a_re -= fabs(c[d]) * (1. - c1) / st.dx;
            a_im -= c[d] * s1 / st.dx;
            b += (2. * c1 - 2.) / (st.dx * st.dx);
        }
    }
    const double ha_re = st.h * a_re, ha_im = st.h * a_im;
    Cplx m{1., 0.};
    if (st.nu_tab) {
        for (int s = 0; s < st.steps; ++s) {
            const double nu = st.nu_tab[(long long)s * st.nu_stride];
            m = cmul(m, step_factor(st.kind, st.h * nu * b + ha_re, ha_im));
        }
    } else {
        Cplx r = step_factor(st.kind, st.h * st.nu * b + ha_re, ha_im);
        for (int n = st.steps; n > 0; n >>= 1) {
            if (n & 1) m = cmul(m, r);
            r = cmul(r, r);
        }
    }
    m = {scale * m.re, scale * m.im};
    if (M) {
        M[idx] = make_double2(m.re, m.im);
        return;
    }
    const double2 x = X[idx];
    X[idx] = make_double2(x.x * m.re - x.y * m.im, x.x * m.im + x.y * m.re);
}

// X *= M, coefficient by coefficient (a constant-nu segment whose multiplier is kept)
__global__ void __launch_bounds__(256) k_spectral_apply(double2* __restrict__ X, const double2* __restrict__ M,
                                                        long long total)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const double2 x = X[idx], m = M[idx];
    X[idx] = make_double2(x.x * m.x - x.y * m.y, x.x * m.y + x.y * m.x);
}

bool same_segment(const SpectralStep& a, const SpectralStep& b)
{
    return a.kind == b.kind && a.dx == b.dx && a.cx == b.cx && a.cy == b.cy && a.cz == b.cz && a.h == b.h &&
           a.steps == b.steps && a.nu == b.nu;
}

}  // namespace

int spectral_create(SpectralPlan** out, int ni, int nj, int nk)
{
    SpectralPlan* sp = new SpectralPlan;
    sp->ni = ni;
    sp->nj = nj;
    sp->nk = nk;
    int r = SP_OK;
    const size_t coefs = (size_t)nk * nj * (ni / 2 + 1);
    if (hipMalloc(&sp->X, coefs * sizeof(double2)) != hipSuccess || hipMalloc(&sp->M, coefs * sizeof(double2)) != hipSuccess)
        r = fail(SP_ERR_NOMEM, "spectral G: coefficient buffers");
    if (r == SP_OK && hipfftPlan3d(&sp->fwd, nk, nj, ni, HIPFFT_D2Z) != HIPFFT_SUCCESS)
        r = fail(SP_ERR_HIP, "spectral G: hipfftPlan3d (D2Z)");
    if (r == SP_OK && hipfftPlan3d(&sp->inv, nk, nj, ni, HIPFFT_Z2D) != HIPFFT_SUCCESS)
        r = fail(SP_ERR_HIP, "spectral G: hipfftPlan3d (Z2D)");
    if (r != SP_OK) {
        spectral_destroy(sp);
        return r;
    }
    *out = sp;
    return SP_OK;
}

void spectral_destroy(SpectralPlan* sp)
{
    if (!sp) return;
    if (sp->fwd) (void)hipfftDestroy(sp->fwd);
    if (sp->inv) (void)hipfftDestroy(sp->inv);
    if (sp->X) (void)hipFree(sp->X);
    if (sp->M) (void)hipFree(sp->M);
    delete sp;
}

int spectral_apply(SpectralPlan* sp, hipStream_t s, const double* in, double* out, const SpectralStep& st)
{
    SP_CHECK(sp && in && out, SP_ERR_INVALID, "spectral G: null argument");
    SP_CHECK(st.steps >= 0, SP_ERR_INVALID, "spectral G: steps < 0");
    SP_CHECK(st.kind == kSpectralRK4 || st.kind == kSpectralEulerUpwind || st.kind == kSpectralEulerCentral,
             SP_ERR_INVALID, "spectral G: kind");
    SP_FFT(hipfftSetStream(sp->fwd, s));
    SP_FFT(hipfftSetStream(sp->inv, s));
    SP_FFT(hipfftExecD2Z(sp->fwd, const_cast<double*>(in), (hipfftDoubleComplex*)sp->X));
    const long long total = (long long)sp->nk * sp->nj * (sp->ni / 2 + 1);
    const double scale = 1.0 / ((double)sp->ni * sp->nj * sp->nk);
    const dim3 grid((unsigned)((total + 255) / 256));
    if (st.nu_tab) {
        hipLaunchKernelGGL(k_spectral_multiplier, grid, dim3(256), 0, s, sp->X, (double2*)nullptr, sp->ni, sp->nj, sp->nk, st,
                           scale);
    } else {
        // constant nu: every segment of the same length has the same multiplier
        if (!sp->m_valid || !same_segment(sp->m_key, st)) {
            hipLaunchKernelGGL(k_spectral_multiplier, grid, dim3(256), 0, s, (double2*)nullptr, sp->M, sp->ni, sp->nj, sp->nk,
                               st, scale);
            SP_HIP(hipGetLastError());
            sp->m_key = st;
            sp->m_valid = true;
        }
        hipLaunchKernelGGL(k_spectral_apply, grid, dim3(256), 0, s, sp->X, (const double2*)sp->M, total);
    }
    SP_HIP(hipGetLastError());
    SP_FFT(hipfftExecZ2D(sp->inv, (hipfftDoubleComplex*)sp->X, out));
    return SP_OK;
}

}  // namespace sp
