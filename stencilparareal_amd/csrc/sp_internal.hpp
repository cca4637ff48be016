// sp_internal.hpp -- shared definitions of the gfx950 StencilParareal library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/stencilparareal.h"

namespace sp {

// ---- errors -------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define SP_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return ::sp::fail(SP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define SP_CHECK(cond, code, msg)                    \
    do {                                             \
        if (!(cond)) return ::sp::fail((code), (msg)); \
    } while (0)

#define SP_TRY(call)                 \
    do {                             \
        int r_ = (call);             \
        if (r_ != SP_OK) return r_;  \
    } while (0)

// ---- storage types ---------------------------------------------------------------
struct bf16 {
    uint16_t v;
};

template <typename S>
struct Store;

template <>
struct Store<double> {
    using T = double;
    __device__ __forceinline__ static double load(const double* p) { return *p; }
    template <typename C>
    __device__ __forceinline__ static void store(double* p, C x) { *p = (double)x; }
};

template <>
struct Store<float> {
    __device__ __forceinline__ static float load(const float* p) { return *p; }
    template <typename C>
    __device__ __forceinline__ static void store(float* p, C x) { *p = (float)x; }
};

// bf16 storage: widening is exact; narrowing rounds to nearest even (deterministic).
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <>
struct Store<bf16> {
    __device__ __forceinline__ static float load(const bf16* p) { return bf16_to_f32(p->v); }
    template <typename C>
    __device__ __forceinline__ static void store(bf16* p, C x) { p->v = f32_to_bf16((float)x); }
};

inline int dtype_size(int dt) { return dt == SP_F64 ? 8 : dt == SP_F32 ? 4 : 2; }

// ---- a field view passed to kernels: pointer at the interior origin ------------------
struct View {
    void* p;
    long long sj, sk;
    long long slice;  // element stride between batched slices (0 for one field)
};

inline View view_of(const sp_layout& l, const void* base)
{
    View v;
    v.p = (char*)base + (size_t)l.origin * dtype_size(l.dtype);
    v.sj = l.sj;
    v.sk = l.sk;
    v.slice = 0;
    return v;
}

int check_layout(const sp_layout* l, const char* what);

// ---- stencil coefficients ------------------------------------------------------------
// Fourth-order operator L = Advection4 + nu*Laplace4 (ConvectionFunctions.h:66-104,
// 135-166) written as one 13-point star: per axis offsets -2,-1,+1,+2 and a centre.
struct Coef13 {
    double x[4], y[4], z[4], c;  // offsets order: -2, -1, +1, +2
};
__host__ __device__ inline Coef13 make_coef13(double nu, double cx, double cy, double cz, double dx)
{
    // Advection4: c*(-q[-2] + 8q[-1] - 8q[+1] + q[+2]) / (12 dx)
    // Laplace4:   (-q[-2] + 16q[-1] - 30q[0] + 16q[+1] - q[+2]) / (12 dx^2), summed over 3 axes
    const double a = 1. / (12. * dx), l = nu / (12. * dx * dx);
    const double adv[4] = {-1., 8., -8., 1.};
    const double lap[4] = {-1., 16., 16., -1.};
    Coef13 c;
    for (int o = 0; o < 4; ++o) {
        c.x[o] = cx * adv[o] * a + lap[o] * l;
        c.y[o] = cy * adv[o] * a + lap[o] * l;
        c.z[o] = cz * adv[o] * a + lap[o] * l;
    }
    c.c = -90. * l;
    return c;
}

// ---- kernel launchers (sp_staged.hip, sp_fused.hip, sp_misc.hip) -----------------------
struct StagedParams {
    int ni, nj, nk;
    double cx, cy, cz, dx, dx2;
};

// Reference-structured sweeps (bit-for-bit the reference operation order).
int launch_rhs4(hipStream_t s, const StagedParams& P, View q, View k, double nu);
int launch_rhs2(hipStream_t s, const StagedParams& P, View q, View k, double nu, int upwind);
int launch_euler_stage(hipStream_t s, const StagedParams& P, View qin, View k, View qout, double dt);
int launch_rk_stage(hipStream_t s, const StagedParams& P, View qin, View k1, View k2, View k3, View k4,
                    View qout, double dt);
// Fused forward Euler with RHS2: qout = qin + dt*(Adv2 + nu*Lap2) in reference order.
int launch_euler_fused(hipStream_t s, const StagedParams& P, View qin, View qout, double nu, double dt,
                       int upwind, int nslices);

// Fused RK4 step over `nslices` batched slices.  Types by dtype codes.
struct FusedTile {
    int tx, ty, nt, kchunk;
    int variant;  // SP_KERNEL_SCALAR, SP_KERNEL_PAIR, SP_KERNEL_QUAD or SP_KERNEL_SPLIT
};
struct FusedStep {
    int ni, nj, nk;
    int nslices;
    double dt;
    double cx, cy, cz, dx;
    const void* coef_dev;  // per-slice coefficient records for this step (device) or nullptr
    double nu;             // used when coef_dev == nullptr
    void* work = nullptr;  // SP_KERNEL_SPLIT: FP64 field of nslices compact slices (v2)
    size_t work_bytes = 0;
};
// A coarse propagator as a Fourier multiplier (sp_spectral.hip): one segment of `steps`
// sub-steps of length h on a compact FP64 field, nu per sub-step from a device table
// (nu_tab[s * nu_stride]) or constant (nu_tab == nullptr).  kind: the sub-step's map.
enum SpectralKind : int {
    kSpectralRK4 = 0,           // DoRK4Timestep: RK4 of Advection4 + nu Laplace4
    kSpectralEulerUpwind = 1,   // DoEulerTimestep, first-order upwind Advection2 + nu Laplace2
    kSpectralEulerCentral = 2,  // DoEulerTimestep, central Advection2 + nu Laplace2
};
struct SpectralStep {
    double dx, cx, cy, cz;
    double h;
    int steps;
    double nu;
    const double* nu_tab;
    long long nu_stride;
    int kind = kSpectralRK4;
};
struct SpectralPlan;
int spectral_create(SpectralPlan** out, int ni, int nj, int nk);
void spectral_destroy(SpectralPlan* sp);
int spectral_apply(SpectralPlan* sp, hipStream_t s, const double* in, double* out, const SpectralStep& st);

// Bytes of the work field a SP_KERNEL_SPLIT step over `nslices` slices needs.
inline size_t split_work_bytes(int ni, int nj, int nk, int nslices)
{
    return (size_t)ni * nj * nk * nslices * 8;
}
// Coefficient record of one fused RK4 step (stencil weights and stage factors for dt).
size_t coefs_size();
void fill_coefs(void* dst, double nu, double cx, double cy, double cz, double dx, double h);
int launch_rk4_fused(hipStream_t s, const FusedStep& st, const FusedTile& tile, int in_dtype, View in,
                     int out_dtype, View out);
bool fused_tile_supported(const FusedTile& t, int compute_f64);
FusedTile default_tile(int ni, int nj, int nk, int nslices, int compute_f64);
// Grouped launches: one launch over several batches with their own sizes, views, steps and
// coefficients.  fused_record_fill writes the kernel's argument record for one batch (of
// fused_record_size() bytes) and its block count; the caller uploads the records and the
// prefix sums of the block counts (first block of each record) to device memory.
size_t fused_record_size();
int fused_record_fill(void* rec, const FusedStep& st, const FusedTile& tile, int in_dtype, View in, int out_dtype,
                      View out, long long* nblocks);
bool fused_record_pair_ok(const FusedStep& st, View in, View out, int in_dtype, int out_dtype);
int launch_rk4_grouped(hipStream_t s, const FusedTile& tile, int in_dtype, int out_dtype, const void* records,
                       const int* first, int n, long long nblocks);
// Tile shapes and k chunks the autotuner times for this size and batch.
void fused_candidates(int ni, int nj, int nk, int nslices, int compute_f64, std::vector<FusedTile>& out);

// Misc kernels.
int launch_periodic_fill(hipStream_t s, const sp_layout& l, void* base);
int launch_update(hipStream_t s, int ni, int nj, int nk, View q, View fine, View cnew, View cold);
int launch_convert(hipStream_t s, int ni, int nj, int nk, int nslices, int in_dtype, View in, int out_dtype,
                   View out);
// Parareal correction: q = cnew + fine - cold (f64), cold <- cnew, increment partials.
int launch_correction(hipStream_t s, int ni, int nj, int nk, int g_dtype, View cnew, View cold, View fine,
                      View dst, double* partials, int nblocks);
int launch_correction_delta(hipStream_t s, int ni, int nj, int nk, int g_dtype, View gd, View fine, View dst,
                            View dnext, double* partials, int nblocks);
// One slice's difference-form correction as a record of a grouped launch (device memory).
struct CorrRecord {
    View gd, fine, dst, dnext;
    int ni, nj, nk;
    double* partials;  // nb blocks x 2 partial sums
};
int launch_correction_delta_grouped(hipStream_t s, int g_dtype, const CorrRecord* recs, int n, int nb);
int launch_norms(hipStream_t s, int ni, int nj, int nk, View a, View b, double* partials, int nblocks);
int launch_fill_exact(hipStream_t s, const sp_layout& l, void* base, double nu0, double nufreq, double cx,
                      double cy, double cz, double t);
int launch_fill_gaussians(hipStream_t s, const sp_layout& l, void* base, int ng, const double* gdev);

constexpr int kReduceBlocks = 512;
// Smallest k extent the fused RK4 kernel streams (its plane wrap covers [-3nk, 4nk)).
constexpr int kFusedMinK = 8;

}  // namespace sp
