// sp_fused.hip -- one RK4 step of the advection-diffusion operator in one kernel.
//
// What it computes (reference: Convection::DoRK4Timestep, src/Convection/Convection.cpp:318-370,
// stages :57-127).  The right-hand side L q = Advection4(q) + nu*Laplace4(q) is linear and nu is
// fixed for the whole step (Convection.cpp:321), so the four RK4 stages
//   k1 = Lq, k2 = L(q + h/2 k1), k3 = L(q + h/2 k2), k4 = L(q + h k3),
//   q' = q + h/6 (k1 + 2k2 + 2k3 + k4)
// equal the nested form (same polynomial P(hL) = 1 + hL + (hL)^2/2 + (hL)^3/6 + (hL)^4/24):
//   v1 = q + (h/4) L q,  v2 = q + (h/3) L v1,  v3 = q + (h/2) L v2,  q' = q + h L v3.
// Results agree with the reference sweeps to rounding (a few ulp per step); the
// reference-order kernels in sp_staged.hip give the bit-exact path.
//
// How (gfx950, wave64):
//  * A block owns a TX x TY tile of output columns and a chunk of CZ planes in k.  It
//    streams planes k = z0-8 .. z0+CZ+7 of the input once (8 = 4 stages x radius 2).
//  * Stage s reads its input plane from LDS (x-y neighbours) and keeps the k-direction
//    in registers as four running partial sums per column: when plane p arrives it
//    completes point p-2 and starts p+2.  Stage s therefore lags 2 planes behind stage
//    s-1; its output plane goes to the next stage's LDS buffer.
//  * The x-y region a stage computes shrinks by 2 per side per stage (6, 4, 2, 0 halo),
//    so only the ring needed by later stages is recomputed.  Each stage maps its own
//    region flat onto the block's threads, so inactive lanes are only the tail.
//  * The state q needed by each stage's update (v_s = q + ...) is re-read from global
//    memory (an L2 hit: the block loaded it a few planes earlier), keeping the per-column
//    register state at 4 partial sums per stage.
//  * HBM traffic per step: one read and one write of the state (2*sizeof(storage) bytes
//    per lattice update) plus the x-y halo overlap of neighbouring tiles.
//  * The plane for iteration t+1 is loaded into registers while iteration t computes.

#include "sp_internal.hpp"

namespace sp {

namespace {

struct FusedArgs {
    View in, out;
    int ni, nj, nk;
    int kchunk;
    int tiles_x;
    double h;
    double cx, cy, cz, dx;
    const double* nu_dev;
    double nu;
};

__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// The 13 star coefficients of L, unscaled: every stage shares them (uniform values that
// live in scalar registers); the stage factor h/4, h/3, h/2, h enters in the update.
template <typename T>
struct StarW {
    T xm2, xm1, xp1, xp2, ym2, ym1, yp1, yp2, zm2, zm1, zp1, zp2, c;
};

template <typename T>
__device__ __forceinline__ StarW<T> star_w(const Coef13& k)
{
    StarW<T> w;
    w.xm2 = (T)k.x[0];
    w.xm1 = (T)k.x[1];
    w.xp1 = (T)k.x[2];
    w.xp2 = (T)k.x[3];
    w.ym2 = (T)k.y[0];
    w.ym1 = (T)k.y[1];
    w.yp1 = (T)k.y[2];
    w.yp2 = (T)k.y[3];
    w.zm2 = (T)k.z[0];
    w.zm1 = (T)k.z[1];
    w.zp1 = (T)k.z[2];
    w.zp2 = (T)k.z[3];
    w.c = (T)k.c;
    return w;
}

__device__ __forceinline__ int pmod(int a, int n)
{
    int r = a % n;
    return r < 0 ? r + n : r;
}

// Stage geometry: stage s (1..4) computes a (TX+2h) x (TY+2h) region, h = 8-2s, from an
// input plane of (TX+2h+4) x (TY+2h+4) held in LDS.
template <int TX, int TY, int S>
struct Geo {
    static constexpr int HO = 8 - 2 * S;
    static constexpr int WO = TX + 2 * HO, HOo = TY + 2 * HO, AO = WO * HOo;
    static constexpr int WI = WO + 4, HI = HOo + 4, AI = WI * HI;
};

// One stage at one plane: consume input plane p (in LDS), emit plane p-2.
template <typename T, int TX, int TY, int S, int NT, int M>
__device__ __forceinline__ void stage_step(const T* __restrict__ P, const StarW<T>& w, T (&acc)[M][4],
                                           const int (&cidx)[M], T (&done)[M])
{
    using G = Geo<TX, TY, S>;
    constexpr int WI = G::WI;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int f = threadIdx.x + m * NT;
        if (M * NT > G::AO && f >= G::AO) break;
        const int ci = cidx[m];
        const T u = P[ci];
        T a1 = acc[m][2];
        a1 = fmaT(w.c, u, a1);
        a1 = fmaT(w.xm2, P[ci - 2], a1);
        a1 = fmaT(w.xm1, P[ci - 1], a1);
        a1 = fmaT(w.xp1, P[ci + 1], a1);
        a1 = fmaT(w.xp2, P[ci + 2], a1);
        a1 = fmaT(w.ym2, P[ci - 2 * WI], a1);
        a1 = fmaT(w.ym1, P[ci - WI], a1);
        a1 = fmaT(w.yp1, P[ci + WI], a1);
        a1 = fmaT(w.yp2, P[ci + 2 * WI], a1);
        done[m] = fmaT(w.zp2, u, acc[m][0]);      // point p-2: its +2 neighbour is p
        acc[m][0] = fmaT(w.zp1, u, acc[m][1]);    // point p-1
        acc[m][1] = a1;                          // point p
        acc[m][2] = fmaT(w.zm1, u, acc[m][3]);    // point p+1
        acc[m][3] = w.zm2 * u;                   // point p+2
    }
}

template <typename SI, typename SO, typename T, int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) k_rk4_fused(FusedArgs a)
{
    using G1 = Geo<TX, TY, 1>;
    using G2 = Geo<TX, TY, 2>;
    using G3 = Geo<TX, TY, 3>;
    using G4 = Geo<TX, TY, 4>;
    constexpr int A0 = G1::AI;  // q plane (stage-1 input) incl. 8-cell halo
    constexpr int M0 = (A0 + NT - 1) / NT;
    constexpr int M1 = (G1::AO + NT - 1) / NT;
    constexpr int M2 = (G2::AO + NT - 1) / NT;
    constexpr int M3 = (G3::AO + NT - 1) / NT;
    constexpr int M4 = (G4::AO + NT - 1) / NT;

    __shared__ T P0[G1::AI];
    __shared__ T P1[G2::AI];
    __shared__ T P2[G3::AI];
    __shared__ T P3[G4::AI];

    const int tid = threadIdx.x;
    const int ni = a.ni, nj = a.nj, nk = a.nk;
    const int tile = blockIdx.x;
    const int x0 = (tile % a.tiles_x) * TX, y0 = (tile / a.tiles_x) * TY;
    const int z0 = blockIdx.y * a.kchunk;
    const int zend = min(z0 + a.kchunk, nk);
    const int slice = blockIdx.z;
    const SI* __restrict__ qin = (const SI*)a.in.p + slice * a.in.slice;
    SO* __restrict__ qout = (SO*)a.out.p + slice * a.out.slice;
    const long long isj = a.in.sj, isk = a.in.sk, osj = a.out.sj, osk = a.out.sk;

    const double nu = a.nu_dev ? a.nu_dev[slice] : a.nu;
    const Coef13 K = make_coef13(nu, a.cx, a.cy, a.cz, a.dx);
    const StarW<T> w = star_w<T>(K);
    const T f1 = (T)(0.25 * a.h), f2 = (T)(a.h / 3.), f3 = (T)(0.5 * a.h), f4 = (T)a.h;

    // Per-column offsets, fixed for the whole k stream.
    int off0[M0];
#pragma unroll
    for (int m = 0; m < M0; ++m) {
        const int f = min(tid + m * NT, A0 - 1);
        const int ax = f % G1::WI, by = f / G1::WI;
        off0[m] = pmod(x0 - 8 + ax, ni) + pmod(y0 - 8 + by, nj) * (int)isj;
    }
    int c1[M1], c2[M2], c3[M3], c4[M4];
    int q1[M1], q2[M2], q3[M3], q4[M4], o4[M4];
    bool w4ok[M4];
#pragma unroll
    for (int m = 0; m < M1; ++m) {
        const int f = min(tid + m * NT, G1::AO - 1);
        const int ax = f % G1::WO, by = f / G1::WO;
        c1[m] = (by + 2) * G1::WI + ax + 2;
        q1[m] = pmod(x0 - G1::HO + ax, ni) + pmod(y0 - G1::HO + by, nj) * (int)isj;
    }
#pragma unroll
    for (int m = 0; m < M2; ++m) {
        const int f = min(tid + m * NT, G2::AO - 1);
        const int ax = f % G2::WO, by = f / G2::WO;
        c2[m] = (by + 2) * G2::WI + ax + 2;
        q2[m] = pmod(x0 - G2::HO + ax, ni) + pmod(y0 - G2::HO + by, nj) * (int)isj;
    }
#pragma unroll
    for (int m = 0; m < M3; ++m) {
        const int f = min(tid + m * NT, G3::AO - 1);
        const int ax = f % G3::WO, by = f / G3::WO;
        c3[m] = (by + 2) * G3::WI + ax + 2;
        q3[m] = pmod(x0 - G3::HO + ax, ni) + pmod(y0 - G3::HO + by, nj) * (int)isj;
    }
#pragma unroll
    for (int m = 0; m < M4; ++m) {
        const int f = min(tid + m * NT, G4::AO - 1);
        const int ax = f % G4::WO, by = f / G4::WO;
        c4[m] = (by + 2) * G4::WI + ax + 2;
        const int gx = x0 + ax, gy = y0 + by;
        q4[m] = pmod(gx, ni) + pmod(gy, nj) * (int)isj;
        o4[m] = gx + gy * (int)osj;
        w4ok[m] = (tid + m * NT < G4::AO) && gx < ni && gy < nj;
    }

    T acc1[M1][4], acc2[M2][4], acc3[M3][4], acc4[M4][4];
    T d1[M1], d2[M2], d3[M3], d4[M4];
#pragma unroll
    for (int m = 0; m < M1; ++m) acc1[m][0] = acc1[m][1] = acc1[m][2] = acc1[m][3] = T(0);
#pragma unroll
    for (int m = 0; m < M2; ++m) acc2[m][0] = acc2[m][1] = acc2[m][2] = acc2[m][3] = T(0);
#pragma unroll
    for (int m = 0; m < M3; ++m) acc3[m][0] = acc3[m][1] = acc3[m][2] = acc3[m][3] = T(0);
#pragma unroll
    for (int m = 0; m < M4; ++m) acc4[m][0] = acc4[m][1] = acc4[m][2] = acc4[m][3] = T(0);

    const int tbeg = z0 - 8, tend = zend + 8;
    T pre[M0];
    {
        const long long zo = pmod(tbeg, nk) * isk;
#pragma unroll
        for (int m = 0; m < M0; ++m) pre[m] = (T)Store<SI>::load(qin + zo + off0[m]);
    }

    for (int t = tbeg; t < tend; ++t) {
        // Plane t of q into LDS; start loading plane t+1.
#pragma unroll
        for (int m = 0; m < M0; ++m) {
            const int f = tid + m * NT;
            if (M0 * NT == A0 || f < A0) P0[f] = pre[m];
        }
        if (t + 1 < tend) {
            const long long zo = pmod(t + 1, nk) * isk;
#pragma unroll
            for (int m = 0; m < M0; ++m) pre[m] = (T)Store<SI>::load(qin + zo + off0[m]);
        }
        __syncthreads();

        // Stage 1: input q plane t, emits v1 plane t-2 into P1.
        {
            stage_step<T, TX, TY, 1, NT, M1>(P0, w, acc1, c1, d1);
            const long long zo = pmod(t - 2, nk) * isk;
#pragma unroll
            for (int m = 0; m < M1; ++m) {
                const int f = tid + m * NT;
                if (M1 * NT == G1::AO || f < G1::AO)
                    P1[f] = fmaT(f1, d1[m], (T)Store<SI>::load(qin + zo + q1[m]));
            }
        }
        __syncthreads();

        // Stage 2: input v1 plane t-2, emits v2 plane t-4.
        if (t >= tbeg + 4) {
            stage_step<T, TX, TY, 2, NT, M2>(P1, w, acc2, c2, d2);
            const long long zo = pmod(t - 4, nk) * isk;
#pragma unroll
            for (int m = 0; m < M2; ++m) {
                const int f = tid + m * NT;
                if (M2 * NT == G2::AO || f < G2::AO)
                    P2[f] = fmaT(f2, d2[m], (T)Store<SI>::load(qin + zo + q2[m]));
            }
        }
        __syncthreads();

        // Stage 3: input v2 plane t-4, emits v3 plane t-6.
        if (t >= tbeg + 8) {
            stage_step<T, TX, TY, 3, NT, M3>(P2, w, acc3, c3, d3);
            const long long zo = pmod(t - 6, nk) * isk;
#pragma unroll
            for (int m = 0; m < M3; ++m) {
                const int f = tid + m * NT;
                if (M3 * NT == G3::AO || f < G3::AO)
                    P3[f] = fmaT(f3, d3[m], (T)Store<SI>::load(qin + zo + q3[m]));
            }
        }
        __syncthreads();

        // Stage 4: input v3 plane t-6, emits q' plane t-8.
        if (t >= tbeg + 12) {
            stage_step<T, TX, TY, 4, NT, M4>(P3, w, acc4, c4, d4);
            const int zp = t - 8;
            if (zp >= z0) {
                const long long zi = pmod(zp, nk) * isk;
                const long long zo = (long long)zp * osk;
#pragma unroll
                for (int m = 0; m < M4; ++m)
                    if (w4ok[m]) Store<SO>::store(qout + zo + o4[m], fmaT(f4, d4[m], (T)Store<SI>::load(qin + zi + q4[m])));
            }
        }
    }
}

template <typename SI, typename SO, typename T>
int launch_typed(hipStream_t s, const FusedStep& st, const FusedTile& tile, View in, View out)
{
    FusedArgs a;
    a.in = in;
    a.out = out;
    a.ni = st.ni;
    a.nj = st.nj;
    a.nk = st.nk;
    a.kchunk = tile.kchunk > 0 ? min(tile.kchunk, st.nk) : st.nk;
    a.tiles_x = (st.ni + tile.tx - 1) / tile.tx;
    const int tiles_y = (st.nj + tile.ty - 1) / tile.ty;
    a.h = st.dt;
    a.cx = st.cx;
    a.cy = st.cy;
    a.cz = st.cz;
    a.dx = st.dx;
    a.nu_dev = st.nu_dev;
    a.nu = st.nu;
    const dim3 grid(a.tiles_x * tiles_y, (st.nk + a.kchunk - 1) / a.kchunk, st.nslices);
#define SP_FUSED_CASE(TX_, TY_, NT_)                                                             \
    if (tile.tx == TX_ && tile.ty == TY_ && tile.nt == NT_) {                                    \
        hipLaunchKernelGGL((k_rk4_fused<SI, SO, T, TX_, TY_, NT_>), grid, dim3(NT_), 0, s, a);   \
        SP_HIP(hipGetLastError());                                                               \
        return SP_OK;                                                                            \
    }
    SP_FUSED_CASE(32, 32, 1024)
    SP_FUSED_CASE(32, 16, 512)
    SP_FUSED_CASE(16, 16, 256)
    SP_FUSED_CASE(24, 24, 512)
#undef SP_FUSED_CASE
    return fail(SP_ERR_UNSUPPORTED, "fused RK4: unsupported tile shape");
}

}  // namespace

bool fused_tile_supported(const FusedTile& t, int /*compute_f64*/)
{
    return (t.tx == 32 && t.ty == 32 && t.nt == 1024) || (t.tx == 32 && t.ty == 16 && t.nt == 512) ||
           (t.tx == 16 && t.ty == 16 && t.nt == 256) || (t.tx == 24 && t.ty == 24 && t.nt == 512);
}

FusedTile default_tile(int ni, int nj, int nk, int nslices, int /*compute_f64*/)
{
    FusedTile t{32, 32, 1024, 0};
    // Aim for at least ~4 blocks per CU (256 CUs) by chunking k.
    const long long tiles = (long long)((ni + 31) / 32) * ((nj + 31) / 32) * nslices;
    const long long want = 1024;
    int chunks = (int)((want + tiles - 1) / tiles);
    if (chunks < 1) chunks = 1;
    int kc = (nk + chunks - 1) / chunks;
    if (kc < 16) kc = 16;
    t.kchunk = kc >= nk ? 0 : kc;
    return t;
}

int launch_rk4_fused(hipStream_t s, const FusedStep& st, const FusedTile& tile, int in_dtype, View in,
                     int out_dtype, View out)
{
    if (in_dtype == SP_F64 && out_dtype == SP_F64) return launch_typed<double, double, double>(s, st, tile, in, out);
    if (in_dtype == SP_F64 && out_dtype == SP_F32) return launch_typed<double, float, float>(s, st, tile, in, out);
    if (in_dtype == SP_F32 && out_dtype == SP_F32) return launch_typed<float, float, float>(s, st, tile, in, out);
    if (in_dtype == SP_F64 && out_dtype == SP_BF16) return launch_typed<double, bf16, float>(s, st, tile, in, out);
    if (in_dtype == SP_BF16 && out_dtype == SP_BF16) return launch_typed<bf16, bf16, float>(s, st, tile, in, out);
    if (in_dtype == SP_F32 && out_dtype == SP_F64) return launch_typed<float, double, float>(s, st, tile, in, out);
    if (in_dtype == SP_BF16 && out_dtype == SP_F64) return launch_typed<bf16, double, float>(s, st, tile, in, out);
    if (in_dtype == SP_BF16 && out_dtype == SP_F32) return launch_typed<bf16, float, float>(s, st, tile, in, out);
    if (in_dtype == SP_F32 && out_dtype == SP_BF16) return launch_typed<float, bf16, float>(s, st, tile, in, out);
    return fail(SP_ERR_UNSUPPORTED, "fused RK4: unsupported dtype pair");
}

}  // namespace sp
