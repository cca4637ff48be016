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
//    memory one plane ahead of its use (an L2 hit: the block loaded it a few planes
//    earlier), keeping the per-column register state at 4 partial sums per stage.
//  * The four running sums of a column are a circular buffer indexed by plane mod 4 and
//    the k loop is unrolled by 4, so advancing a plane moves no registers.
//  * The 13 stencil coefficients and the stage factors are computed on the host and
//    arrive as kernel arguments (scalar registers); global accesses are buffer
//    instructions with the plane offset in a scalar register.
//  * HBM traffic per step: one read and one write of the state (2*sizeof(storage) bytes
//    per lattice update) plus the x-y halo overlap of neighbouring tiles.
//  * The plane for iteration t+1 is loaded into registers while iteration t computes.
// That is the scalar kernel (any sizes and storage types).  The FP64 x-pair and 2x2
// variants further down keep the scheme and change the lane layout (16-byte accesses,
// shared neighbour reads), the way q reaches each stage and the prefetch depth; their
// own comments give the details.

#include <algorithm>
#include <cstring>

#include "sp_internal.hpp"

namespace sp {

namespace {

// Host-computed coefficients of one step: the 13 star weights of L (offsets -2,-1,+1,+2
// per axis, then the centre) and the stage factors h/4, h/3, h/2, h, in both precisions.
struct Coefs {
    double wd[13];
    double fd[4];
    float wf[13];
    float ff[4];
};

struct FusedArgs {
    View in, out;
    unsigned in_bytes, out_bytes;  // bytes one slice's storage spans from its origin
    int ni, nj, nk;
    int kchunk, chunks;
    int tiles_x, tiles;
    Coefs c;                 // used when cdev == nullptr
    const Coefs* cdev;       // per-slice coefficients (time-dependent nu), or nullptr
};

// A grouped launch: n argument records in device memory and their first block indices.
struct GroupArgs {
    const FusedArgs* __restrict__ args;
    const int* __restrict__ first;
    int n;
};

__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T>
struct StarW {
    T xm2, xm1, xp1, xp2, ym2, ym1, yp1, yp2, zm2, zm1, zp1, zp2, c;
    T f1, f2, f3, f4;
};

template <typename T>
__device__ __forceinline__ StarW<T> load_w(const Coefs& k);
template <>
__device__ __forceinline__ StarW<double> load_w<double>(const Coefs& k)
{
    return StarW<double>{k.wd[0], k.wd[1], k.wd[2], k.wd[3], k.wd[4], k.wd[5], k.wd[6], k.wd[7], k.wd[8],
                         k.wd[9], k.wd[10], k.wd[11], k.wd[12], k.fd[0], k.fd[1], k.fd[2], k.fd[3]};
}
template <>
__device__ __forceinline__ StarW<float> load_w<float>(const Coefs& k)
{
    return StarW<float>{k.wf[0], k.wf[1], k.wf[2], k.wf[3], k.wf[4], k.wf[5], k.wf[6], k.wf[7], k.wf[8],
                        k.wf[9], k.wf[10], k.wf[11], k.wf[12], k.ff[0], k.ff[1], k.ff[2], k.ff[3]};
}

__device__ __forceinline__ int pmod(int a, int n)
{
    int r = a % n;
    return r < 0 ? r + n : r;
}

// Plane index p in [-3n, 4n) wrapped into [0, n) without branches (scalar selects).  The
// stream touches planes z0-18 .. zend+16, inside that range for n >= 8; the host sends
// smaller k extents to the reference-order path.
__device__ __forceinline__ int wrapz(int p, int n)
{
    p += p < 0 ? n : 0;
    p += p < 0 ? n : 0;
    p += p < 0 ? n : 0;
    p -= p >= n ? n : 0;
    p -= p >= n ? n : 0;
    p -= p >= n ? n : 0;
    return p;
}

// Stage geometry: stage s (1..4) computes a (TX+2h) x (TY+2h) region, h = 8-2s, from an
// input plane of (TX+2h+4) x (TY+2h+4) held in LDS.
template <int TX, int TY, int S>
struct Geo {
    static constexpr int HO = 8 - 2 * S;
    static constexpr int WO = TX + 2 * HO, HOo = TY + 2 * HO, AO = WO * HOo;
    static constexpr int WI = WO + 4, HI = HOo + 4, AI = WI * HI;
};

// One stage at one plane: consume input plane p (in LDS), emit the completed sum of
// plane p-2 in done[].  The running sums of points p-2 .. p+1 sit in slots (x mod 4)
// of acc; R is the slot of point p.  Every slot m is processed (slots past the region's
// end repeat its last column), so the stream has no per-lane branches.
template <typename T, int TX, int TY, int S, int NT, int M, int R>
__device__ __forceinline__ void stage_step(const T* __restrict__ P, const StarW<T>& w, T (&acc)[M][4], T (&done)[M])
{
    using G = Geo<TX, TY, S>;
    constexpr int WI = G::WI;
    constexpr int sm2 = (R + 2) & 3, sm1 = (R + 3) & 3, s0 = R & 3, sp1 = (R + 1) & 3;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        // LDS centre of this slot's column (recomputed: constant divisors, saves registers)
        const int f = min((int)threadIdx.x + m * NT, G::AO - 1);
        const T* c = P + (f / G::WO + 2) * WI + f % G::WO + 2;
        const T u = c[0];
        const T xm2 = c[-2], xm1 = c[-1], xp1 = c[1], xp2 = c[2];
        const T ym2 = c[-2 * WI], ym1 = c[-WI], yp1 = c[WI], yp2 = c[2 * WI];
        done[m] = fmaT(w.zp2, u, acc[m][sm2]);            // point p-2: its +2 neighbour is p
        acc[m][sm2] = w.zm2 * u;                          // point p+2 starts in the same slot
        acc[m][sm1] = fmaT(w.zp1, u, acc[m][sm1]);        // point p-1
        // point p: centre and x-y neighbours, as two independent chains (x, y)
        T ax = fmaT(w.c, u, acc[m][s0]);
        T ay = w.ym2 * ym2;
        ax = fmaT(w.xm2, xm2, ax);
        ay = fmaT(w.ym1, ym1, ay);
        ax = fmaT(w.xm1, xm1, ax);
        ay = fmaT(w.yp1, yp1, ay);
        ax = fmaT(w.xp1, xp1, ax);
        ay = fmaT(w.yp2, yp2, ay);
        ax = fmaT(w.xp2, xp2, ax);
        acc[m][s0] = ax + ay;
        acc[m][sp1] = fmaT(w.zm1, u, acc[m][sp1]);        // point p+1
        // one slot's neighbour values in flight at a time keeps the register footprint
        // small enough for 16 waves per CU; the other waves hide the LDS latency
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Buffer-resource access: one SGPR descriptor per field, 32-bit per-lane byte offsets
// and the plane offset in an SGPR, so the k stream needs no 64-bit address arithmetic.
// Out-of-range offsets read 0 / drop the store instead of faulting.
#define SP_RSRC_DWORD3 0x00020000  // gfx9 family raw buffer (as composable_kernel uses)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, SP_RSRC_DWORD3);
}

template <typename S>
struct Buf;
template <>
struct Buf<double> {
    using T = double;
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    __device__ __forceinline__ static double ld(__amdgpu_buffer_rsrc_t r, int vo, int so)
    {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
    }
    __device__ __forceinline__ static void st(__amdgpu_buffer_rsrc_t r, int vo, int so, double x)
    {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, vo, so, 0);
    }
};
template <>
struct Buf<float> {
    __device__ __forceinline__ static float ld(__amdgpu_buffer_rsrc_t r, int vo, int so)
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
    }
    __device__ __forceinline__ static void st(__amdgpu_buffer_rsrc_t r, int vo, int so, float x)
    {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, vo, so, 0);
    }
};
template <>
struct Buf<bf16> {
    __device__ __forceinline__ static float ld(__amdgpu_buffer_rsrc_t r, int vo, int so)
    {
        return bf16_to_f32((uint16_t)__builtin_amdgcn_raw_buffer_load_b16(r, vo, so, 0));
    }
    __device__ __forceinline__ static void st(__amdgpu_buffer_rsrc_t r, int vo, int so, float x)
    {
        __builtin_amdgcn_raw_buffer_store_b16((unsigned short)f32_to_bf16(x), r, vo, so, 0);
    }
};

// Per-column constants of one stage: LDS centre index and the byte offset of the column
// in the input field (for the q value of the update).  Slots past the region's end repeat
// its last column.
template <int M>
struct Cols {
    int q[M];
};

template <typename SI, int TX, int TY, int S, int NT, int M>
__device__ __forceinline__ void stage_cols(Cols<M>& cl, int x0, int y0, int ni, int nj, int sj)
{
    using G = Geo<TX, TY, S>;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int f = min((int)threadIdx.x + m * NT, G::AO - 1);
        const int ax = f % G::WO, by = f / G::WO;
        cl.q[m] = (pmod(x0 - G::HO + ax, ni) + pmod(y0 - G::HO + by, nj) * sj) * (int)sizeof(SI);
    }
}

// Loads the q values a stage needs for its update at the plane whose byte offset is zo.
template <typename SI, typename T, int M>
__device__ __forceinline__ void stage_qload(T (&qv)[M], __amdgpu_buffer_rsrc_t r, int zo, const Cols<M>& cl)
{
#pragma unroll
    for (int m = 0; m < M; ++m) qv[m] = (T)Buf<SI>::ld(r, cl.q[m], zo);
}

// Blocks are handed to the 8 XCDs round-robin (block b runs on XCD b % 8).  Renumber so
// that each XCD receives a contiguous range of tiles: neighbouring tiles, which read each
// other's halo columns, then share one L2.
__device__ __forceinline__ int xcd_block(int b, int n)
{
    if (n % 8 != 0) return b;
    return (b % 8) * (n / 8) + b / 8;
}

template <typename SI, typename SO, typename T, int TX, int TY, int NT>
__device__ __forceinline__ void rk4_scalar_body(const FusedArgs& a, const int bid)
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

    // Every plane buffer is doubled: in iteration t stage s reads what stage s-1 wrote in
    // iteration t-1, so all four stages run between one barrier per plane.
    __shared__ T P0[2][G1::AI];
    __shared__ T P1[2][G2::AI];
    __shared__ T P2[2][G3::AI];
    __shared__ T P3[2][G4::AI];

    const int tid = threadIdx.x;
    const int ni = a.ni, nj = a.nj, nk = a.nk;
    // block -> (tile, k chunk, slice); tiles fastest
    const int tile = bid % a.tiles;
    const int chunk = (bid / a.tiles) % a.chunks;
    const int slice = bid / (a.tiles * a.chunks);
    const int x0 = (tile % a.tiles_x) * TX, y0 = (tile / a.tiles_x) * TY;
    const int z0 = chunk * a.kchunk;
    const int zend = min(z0 + a.kchunk, nk);
    const __amdgpu_buffer_rsrc_t rin = make_rsrc((const SI*)a.in.p + slice * a.in.slice, a.in_bytes);
    const __amdgpu_buffer_rsrc_t rout = make_rsrc((SO*)a.out.p + slice * a.out.slice, a.out_bytes);
    const int iskb = (int)a.in.sk * (int)sizeof(SI), oskb = (int)a.out.sk * (int)sizeof(SO);
    const int isj = (int)a.in.sj, osj = (int)a.out.sj;
    const StarW<T> w = a.cdev ? load_w<T>(a.cdev[slice]) : load_w<T>(a.c);

    // Per-column offsets, fixed for the whole k stream.
    int off0[M0];
#pragma unroll
    for (int m = 0; m < M0; ++m) {
        const int f = min(tid + m * NT, A0 - 1);
        const int ax = f % G1::WI, by = f / G1::WI;
        off0[m] = (pmod(x0 - 8 + ax, ni) + pmod(y0 - 8 + by, nj) * isj) * (int)sizeof(SI);
    }
    Cols<M1> c1;
    Cols<M2> c2;
    Cols<M3> c3;
    Cols<M4> c4;
    stage_cols<SI, TX, TY, 1, NT, M1>(c1, x0, y0, ni, nj, isj);
    stage_cols<SI, TX, TY, 2, NT, M2>(c2, x0, y0, ni, nj, isj);
    stage_cols<SI, TX, TY, 3, NT, M3>(c3, x0, y0, ni, nj, isj);
    stage_cols<SI, TX, TY, 4, NT, M4>(c4, x0, y0, ni, nj, isj);
    int o4[M4];
#pragma unroll
    for (int m = 0; m < M4; ++m) {
        const int fc = min(tid + m * NT, G4::AO - 1);
        const int gx = x0 + fc % G4::WO, gy = y0 + fc / G4::WO;
        // columns of a partial tile past the domain edge are written to their periodic image
        // (where another block writes the same bits)
        o4[m] = (pmod(gx, ni) + pmod(gy, nj) * osj) * (int)sizeof(SO);
    }

    T acc1[M1][4], acc2[M2][4], acc3[M3][4], acc4[M4][4];
    T d1[M1], d2[M2], d3[M3], d4[M4];
    T qv1[M1], qv2[M2], qv3[M3], qv4[M4];  // q for the update, loaded one plane ahead
#pragma unroll
    for (int m = 0; m < M1; ++m) acc1[m][0] = acc1[m][1] = acc1[m][2] = acc1[m][3] = T(0);
#pragma unroll
    for (int m = 0; m < M2; ++m) acc2[m][0] = acc2[m][1] = acc2[m][2] = acc2[m][3] = T(0);
#pragma unroll
    for (int m = 0; m < M3; ++m) acc3[m][0] = acc3[m][1] = acc3[m][2] = acc3[m][3] = T(0);
#pragma unroll
    for (int m = 0; m < M4; ++m) acc4[m][0] = acc4[m][1] = acc4[m][2] = acc4[m][3] = T(0);

    // Iteration t: stage 1 takes q plane t and emits v1 plane t-2; stage 2 takes v1 plane
    // t-3 (written in iteration t-1) and emits v2 plane t-5; stage 3: v2 t-6 -> v3 t-8;
    // stage 4: v3 t-9 -> q' t-11.  Stage s runs while its input planes are valid.
    const int tq0 = z0 - 8, tq1 = zend + 8;  // q planes to stream
    const int tend = zend + 11;
    {
        // q plane tq0 into P0[tq0 & 1]
        const int zo = wrapz(tq0, nk) * iskb;
#pragma unroll
        for (int m = 0; m < M0; ++m) P0[tq0 & 1][min(tid + m * NT, A0 - 1)] = (T)Buf<SI>::ld(rin, off0[m], zo);
        stage_qload<SI, T, M1>(qv1, rin, wrapz(tq0 - 2, nk) * iskb, c1);
        stage_qload<SI, T, M2>(qv2, rin, wrapz(z0 - 8, nk) * iskb, c2);
        stage_qload<SI, T, M3>(qv3, rin, wrapz(z0 - 6, nk) * iskb, c3);
        stage_qload<SI, T, M4>(qv4, rin, wrapz(z0 - 4, nk) * iskb, c4);
    }
    T pre[M0];
    if (tq0 + 1 < tq1) {
        const int zo = wrapz(tq0 + 1, nk) * iskb;
#pragma unroll
        for (int m = 0; m < M0; ++m) pre[m] = (T)Buf<SI>::ld(rin, off0[m], zo);
    }
    __syncthreads();

    // One plane of the stream; R = (t - tq0) mod 4 is a compile-time constant so the slots of
    // the running sums are static register names.  Slot of point x = (x - tq0) mod 4.
    auto plane = [&](auto Rc, int t) {
        constexpr int R = decltype(Rc)::value;
        const int cur = t & 1, prv = cur ^ 1;
        __builtin_amdgcn_sched_barrier(0);  // keep planes apart: bounded live ranges
        // next q plane from registers into the other P0 buffer; start loading the one after
        if (t + 1 < tq1) {
#pragma unroll
            for (int m = 0; m < M0; ++m) P0[prv][min(tid + m * NT, A0 - 1)] = pre[m];
            if (t + 2 < tq1) {
                const int zo = wrapz(t + 2, nk) * iskb;
#pragma unroll
                for (int m = 0; m < M0; ++m) pre[m] = (T)Buf<SI>::ld(rin, off0[m], zo);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        // Stage 1: q plane t -> v1 plane t-2.
        if (t < tq1) {
            stage_step<T, TX, TY, 1, NT, M1, R & 3>(P0[cur], w, acc1, d1);
#pragma unroll
            for (int m = 0; m < M1; ++m) P1[cur][min(tid + m * NT, G1::AO - 1)] = fmaT(w.f1, d1[m], qv1[m]);
            stage_qload<SI, T, M1>(qv1, rin, wrapz(t - 1, nk) * iskb, c1);
            __builtin_amdgcn_sched_barrier(0);
        }
        // Stage 2: v1 plane t-3 -> v2 plane t-5.
        if (t >= tq0 + 5 && t < tq1 + 1) {
            stage_step<T, TX, TY, 2, NT, M2, (R + 1) & 3>(P1[prv], w, acc2, d2);
#pragma unroll
            for (int m = 0; m < M2; ++m) P2[cur][min(tid + m * NT, G2::AO - 1)] = fmaT(w.f2, d2[m], qv2[m]);
            stage_qload<SI, T, M2>(qv2, rin, wrapz(t - 4, nk) * iskb, c2);
            __builtin_amdgcn_sched_barrier(0);
        }
        // Stage 3: v2 plane t-6 -> v3 plane t-8.
        if (t >= tq0 + 10 && t < tq1 + 2) {
            stage_step<T, TX, TY, 3, NT, M3, (R + 2) & 3>(P2[prv], w, acc3, d3);
#pragma unroll
            for (int m = 0; m < M3; ++m) P3[cur][min(tid + m * NT, G3::AO - 1)] = fmaT(w.f3, d3[m], qv3[m]);
            stage_qload<SI, T, M3>(qv3, rin, wrapz(t - 7, nk) * iskb, c3);
            __builtin_amdgcn_sched_barrier(0);
        }
        // Stage 4: v3 plane t-9 -> q' plane t-11.
        if (t >= tq0 + 15) {
            stage_step<T, TX, TY, 4, NT, M4, (R + 3) & 3>(P3[prv], w, acc4, d4);
            const int zp = t - 11;
            if (zp >= z0 && zp < zend) {
                const int zo = zp * oskb;
#pragma unroll
                for (int m = 0; m < M4; ++m) Buf<SO>::st(rout, o4[m], zo, fmaT(w.f4, d4[m], qv4[m]));
            }
            stage_qload<SI, T, M4>(qv4, rin, wrapz(t - 10, nk) * iskb, c4);
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
    };

    // Stage s runs while its input planes are valid.  The loop is unrolled by 4 so that
    // the phase R of every plane is a compile-time constant.
    for (int t = tq0; t < tend; t += 4) {
        plane(std::integral_constant<int, 0>{}, t);
        if (t + 1 < tend) plane(std::integral_constant<int, 1>{}, t + 1);
        if (t + 2 < tend) plane(std::integral_constant<int, 2>{}, t + 2);
        if (t + 3 < tend) plane(std::integral_constant<int, 3>{}, t + 3);
    }
}


// ---- FP64 x-pair variant ----------------------------------------------------------------
// Each lane owns two x-adjacent columns, so every LDS and global access moves 16 bytes
// (ds_read_b128 / buffer_load_dwordx4): per pair and stage, 7 LDS reads give the two centres,
// the four x and eight y neighbours.  Needs even ni and even, 16-byte aligned strides.
typedef double d2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ d2 ld2(__amdgpu_buffer_rsrc_t r, int vo, int so)
{
    return __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, 0));
}
__device__ __forceinline__ void st2(__amdgpu_buffer_rsrc_t r, int vo, int so, d2 x)
{
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, x), r, vo, so, 0);
}

// Non-temporal variants (cache policy nt): the output plane of a step is not read again in
// that step, and a block's lag-8 reload of q is that plane's last use, so neither needs to
// stay in the XCD's L2 beside the planes the stream will still reuse (the 2x2 kernel's
// counted reads at 64 x 384^3: 3.26x -> 2.73x the state).
__device__ __forceinline__ d2 ld2_last(__amdgpu_buffer_rsrc_t r, int vo, int so)
{
    return __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, 2));
}
__device__ __forceinline__ void st2_stream(__amdgpu_buffer_rsrc_t r, int vo, int so, d2 x)
{
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, x), r, vo, so, 2);
}
// 16 bytes per lane from global memory straight into LDS: lane l of the wave writes
// lds + 16 l (lds: the wave's base, the same for every lane).
__device__ __forceinline__ void ld2_to_lds(__amdgpu_buffer_rsrc_t r, double* lds, int vo, int so)
{
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)lds, 16, vo, so, 0, 0);
}

// Two x-adjacent values of compute type T (d2 for FP64), and their global access for
// storage type S: 2 * sizeof(S) bytes per lane (FP64 16, FP32 8, BF16 4), converted to T.
template <typename T>
struct V2 {
    typedef T type __attribute__((ext_vector_type(2)));
};
template <typename S, typename T>
struct PairIO;
template <>
struct PairIO<double, double> {
    __device__ __forceinline__ static d2 ld(__amdgpu_buffer_rsrc_t r, int vo, int so) { return ld2(r, vo, so); }
    __device__ __forceinline__ static void st(__amdgpu_buffer_rsrc_t r, int vo, int so, d2 x) { st2(r, vo, so, x); }
};
template <>
struct PairIO<float, float> {
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    __device__ __forceinline__ static f2 ld(__amdgpu_buffer_rsrc_t r, int vo, int so)
    {
        return __builtin_bit_cast(f2, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
    }
    __device__ __forceinline__ static void st(__amdgpu_buffer_rsrc_t r, int vo, int so, f2 x)
    {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, vo, so, 0);
    }
};
template <>
struct PairIO<double, float> {  // FP64 input of a reduced-precision step: rounded to FP32 on load
    typedef float f2 __attribute__((ext_vector_type(2)));
    __device__ __forceinline__ static f2 ld(__amdgpu_buffer_rsrc_t r, int vo, int so)
    {
        const d2 v = ld2(r, vo, so);
        return f2{(float)v.x, (float)v.y};
    }
};
template <>
struct PairIO<bf16, float> {
    typedef float f2 __attribute__((ext_vector_type(2)));
    __device__ __forceinline__ static f2 ld(__amdgpu_buffer_rsrc_t r, int vo, int so)
    {
        const unsigned v = __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0);
        return f2{bf16_to_f32((uint16_t)(v & 0xffffu)), bf16_to_f32((uint16_t)(v >> 16))};
    }
    __device__ __forceinline__ static void st(__amdgpu_buffer_rsrc_t r, int vo, int so, f2 x)
    {
        const unsigned v = (unsigned)f32_to_bf16(x.x) | ((unsigned)f32_to_bf16(x.y) << 16);
        __builtin_amdgcn_raw_buffer_store_b32(v, r, vo, so, 0);
    }
};

// Lane-slot layout of stage S's output pairs (WP pairs x H rows).  Slot g maps to pair
// (ax / 2, by), row-major with VW slots per row.  With VW equal to the input plane's row
// stride in pairs (WI / 2), consecutive slots read consecutive 16-byte LDS chunks across
// row ends too, so no 16-lane group of a ds_read_b128 touches a bank twice; the slots past
// a row's last pair repeat it (same address: a broadcast, and the same value written).
// That layout is used whenever it needs no more passes of NT threads than the dense one
// (VW = WP), whose groups straddling a row end hit two banks twice, unless rows are whole
// 16-lane groups already.
template <int TX, int TY, int S, int NT>
struct PairMap {
    using G = Geo<TX, TY, S>;
    static constexpr int WP = G::WO / 2, H = G::HOo;
    static constexpr int DENSE = WP * H, PADDED = (G::WI / 2) * H;
    static constexpr bool PAD = WP % 16 != 0 && (PADDED + NT - 1) / NT == (DENSE + NT - 1) / NT;
    static constexpr int VW = PAD ? G::WI / 2 : WP;
    static constexpr int SLOTS = VW * H;
    static constexpr int M = (SLOTS + NT - 1) / NT;  // passes
    // pair of slot g: column ax (doubles) and row by of the output region
    __device__ static void at(int g, int& ax, int& by)
    {
        g = min(g, SLOTS - 1);
        by = g / VW;
        ax = 2 * min(g % VW, WP - 1);
    }
    // its centre in the input plane (row stride WIN), and its place in the output plane
    template <int WIN = G::WI>
    __device__ static int centre(int g)
    {
        int ax, by;
        at(g, ax, by);
        return (by + 2) * WIN + ax + 2;
    }
    __device__ static int out(int g)
    {
        int ax, by;
        at(g, ax, by);
        return by * G::WO + ax;
    }
    // slot g is a pair of its own: not past the last slot, not a row's padding slot
    __device__ static bool live(int g) { return g < SLOTS && g % VW < WP; }
};

template <int TX, int TY, int S, int NT, int M, int R, int WI = Geo<TX, TY, S>::WI, typename T>
__device__ __forceinline__ void pair_step(const T* __restrict__ P, const StarW<T>& w, T (&acc)[M][2][4],
                                          T (&done)[M][2], typename V2<T>::type (&cen)[M])
{
    using v2 = typename V2<T>::type;
    constexpr int sm2 = (R + 2) & 3, sm1 = (R + 3) & 3, s0 = R & 3, sp1 = (R + 1) & 3;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const T* c = P + PairMap<TX, TY, S, NT>::template centre<WI>((int)threadIdx.x + m * NT);
        // x group first (centres and x neighbours), then the y group: at most 4 pairs of
        // loaded values are live at a time
        const v2 C = *(const v2*)c, Lx = *(const v2*)(c - 2), Rx = *(const v2*)(c + 2);
        cen[m] = C;
        T ax[2];
        {
            const T u[2] = {C.x, C.y};
            // x neighbours of the two columns: (Lx.x, Lx.y, C.y, Rx.x) and (Lx.y, C.x, Rx.x, Rx.y)
            const T xm2[2] = {Lx.x, Lx.y}, xm1[2] = {Lx.y, C.x}, xp1[2] = {C.y, Rx.x}, xp2[2] = {Rx.x, Rx.y};
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                done[m][e] = fmaT(w.zp2, u[e], acc[m][e][sm2]);
                acc[m][e][sm2] = w.zm2 * u[e];
                acc[m][e][sm1] = fmaT(w.zp1, u[e], acc[m][e][sm1]);
                acc[m][e][sp1] = fmaT(w.zm1, u[e], acc[m][e][sp1]);
                T a = fmaT(w.c, u[e], acc[m][e][s0]);
                a = fmaT(w.xm2, xm2[e], a);
                a = fmaT(w.xm1, xm1[e], a);
                a = fmaT(w.xp1, xp1[e], a);
                ax[e] = fmaT(w.xp2, xp2[e], a);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        {
            const v2 Ym2 = *(const v2*)(c - 2 * WI), Ym1 = *(const v2*)(c - WI);
            const v2 Yp1 = *(const v2*)(c + WI), Yp2 = *(const v2*)(c + 2 * WI);
            const T ym2[2] = {Ym2.x, Ym2.y}, ym1[2] = {Ym1.x, Ym1.y}, yp1[2] = {Yp1.x, Yp1.y},
                    yp2[2] = {Yp2.x, Yp2.y};
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                T a = w.ym2 * ym2[e];
                a = fmaT(w.ym1, ym1[e], a);
                a = fmaT(w.yp1, yp1[e], a);
                a = fmaT(w.yp2, yp2[e], a);
                acc[m][e][s0] = ax[e] + a;
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Byte offsets (element size ES) of stage S's output pairs in a field (periodic images).
template <int TX, int TY, int S, int NT, int M, int ES = 8>
__device__ __forceinline__ void pair_offsets(int (&q)[M], int x0, int y0, int ni, int nj, int sj)
{
    using G = Geo<TX, TY, S>;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        int ax, by;
        PairMap<TX, TY, S, NT>::at((int)threadIdx.x + m * NT, ax, by);
        q[m] = (pmod(x0 - G::HO + ax, ni) + pmod(y0 - G::HO + by, nj) * sj) * ES;
    }
}

template <typename SI, int M, typename T>
__device__ __forceinline__ void pair_qload(T (&qv)[M], __amdgpu_buffer_rsrc_t r, int zo, const int (&q)[M])
{
#pragma unroll
    for (int m = 0; m < M; ++m) qv[m] = PairIO<SI, decltype(qv[m].x + 0)>::ld(r, q[m], zo);
}

// Index of pair g of stage S's output region inside a plane of region stage S+d
// (d = 0: the same region; d = 1: the next, smaller one, or `dump` when the pair lies in
// the halo that the next stage does not compute).
template <int TX, int TY, int S, int NT, int M>
__device__ __forceinline__ void pair_qidx(int (&same)[M], int (&next)[M])
{
    using G = Geo<TX, TY, S>;
    constexpr int WN = G::WO - 4, AN = WN * (G::HOo - 4);
#pragma unroll
    for (int m = 0; m < M; ++m) {
        int ax, by;
        PairMap<TX, TY, S, NT>::at((int)threadIdx.x + m * NT, ax, by);
        same[m] = by * G::WO + ax;
        const bool in = ax >= 2 && ax < G::WO - 2 && by >= 2 && by < G::HOo - 2;
        next[m] = in ? (by - 2) * WN + ax - 2 : AN;  // AN: a dump pair past the plane
    }
}

// RK4 step, two x-adjacent columns per lane: FP64 (SI = SO = T = double), or FP32 arithmetic
// over FP32 / BF16 storage (T = float; SI = double for a step from an FP64 input).  The
// arithmetic is the scalar kernel's for the same types, so both give the same bits.  The
// stage updates v_s = q + f_s L v_{s-1}
// need q at each stage's output plane, which lags the q plane being loaded by 2, 5, 8 and
// 11 planes.  Reloading it from global memory at those lags keeps a 12-plane window per
// block live in L2, more than an XCD's L2 holds for its resident blocks.  Instead:
//   stage 1 keeps the centre values of the q planes it reads from LDS (2-deep register
//     ring), stage 2 reloads q at lag 5 (a window the L2 holds), and writes it into an
//     LDS plane Q3 beside its output; stage 3 reads q from Q3 with its input plane, keeps
//     it 2 planes in registers and forwards it through Q4 to stage 4 the same way.
template <typename SI, typename SO, typename T, int TX, int TY, int NT>
__device__ __forceinline__ void rk4_pair_body(const FusedArgs& a, const int bid)
{
    using v2 = typename V2<T>::type;
    constexpr int ESI = sizeof(SI), ESO = sizeof(SO);
    using G1 = Geo<TX, TY, 1>;
    using G2 = Geo<TX, TY, 2>;
    using G3 = Geo<TX, TY, 3>;
    using G4 = Geo<TX, TY, 4>;
    constexpr int A0P = G1::AI / 2;  // pairs of the q plane incl. its 8-cell halo
    constexpr int M0 = (A0P + NT - 1) / NT;
    using PM1 = PairMap<TX, TY, 1, NT>;
    using PM2 = PairMap<TX, TY, 2, NT>;
    using PM3 = PairMap<TX, TY, 3, NT>;
    using PM4 = PairMap<TX, TY, 4, NT>;
    constexpr int M1 = PM1::M, M2 = PM2::M, M3 = PM3::M, M4 = PM4::M;

    __shared__ __attribute__((aligned(16))) T P0[2][G1::AI];
    __shared__ __attribute__((aligned(16))) T P1[2][G2::AI];
    __shared__ __attribute__((aligned(16))) T P2[2][G3::AI];
    __shared__ __attribute__((aligned(16))) T P3[2][G4::AI];
    __shared__ __attribute__((aligned(16))) T Q3[2][G3::AO + 2];  // q on stage 3's region
    __shared__ __attribute__((aligned(16))) T Q4[2][G4::AO + 2];  // q on stage 4's region

    const int tid = threadIdx.x;
    const int ni = a.ni, nj = a.nj, nk = a.nk;
    const int tile = bid % a.tiles;
    const int chunk = (bid / a.tiles) % a.chunks;
    const int slice = bid / (a.tiles * a.chunks);
    const int x0 = (tile % a.tiles_x) * TX, y0 = (tile / a.tiles_x) * TY;
    const int z0 = chunk * a.kchunk;
    const int zend = min(z0 + a.kchunk, nk);
    const __amdgpu_buffer_rsrc_t rin = make_rsrc((const SI*)a.in.p + slice * a.in.slice, a.in_bytes);
    const __amdgpu_buffer_rsrc_t rout = make_rsrc((SO*)a.out.p + slice * a.out.slice, a.out_bytes);
    const int iskb = (int)a.in.sk * ESI, oskb = (int)a.out.sk * ESO;
    const int isj = (int)a.in.sj, osj = (int)a.out.sj;
    const StarW<T> w = a.cdev ? load_w<T>(a.cdev[slice]) : load_w<T>(a.c);

    int off0[M0];
#pragma unroll
    for (int m = 0; m < M0; ++m) {
        const int g = min(tid + m * NT, A0P - 1);
        const int ax = 2 * (g % (G1::WI / 2)), by = g / (G1::WI / 2);
        off0[m] = (pmod(x0 - 8 + ax, ni) + pmod(y0 - 8 + by, nj) * isj) * ESI;
    }
    int q2[M2], o4[M4];
    pair_offsets<TX, TY, 2, NT, M2, ESI>(q2, x0, y0, ni, nj, isj);
#pragma unroll
    for (int m = 0; m < M4; ++m) {
        int ax, by;
        PM4::at(tid + m * NT, ax, by);
        o4[m] = (pmod(x0 + ax, ni) + pmod(y0 + by, nj) * osj) * ESO;  // periodic image for partial tiles
    }
    int s2[M2], w3[M2], r3x[M3], w4[M3], r4x[M4], n4[M4];
    // slots past a stage's last pair and the padding slots of padded rows repeat a pair:
    // they write nothing (several lanes writing one LDS address serialise).  The 32x32
    // tile has no registers left for the masks and writes the repeats.
    constexpr bool MASK = TX * TY <= 512;
    bool lv1[M1], lv2[M2], lv3[M3], lv4[M4];
    if constexpr (MASK) {
#pragma unroll
        for (int m = 0; m < M1; ++m) lv1[m] = PM1::live(tid + m * NT);
#pragma unroll
        for (int m = 0; m < M2; ++m) lv2[m] = PM2::live(tid + m * NT);
#pragma unroll
        for (int m = 0; m < M3; ++m) lv3[m] = PM3::live(tid + m * NT);
#pragma unroll
        for (int m = 0; m < M4; ++m) lv4[m] = PM4::live(tid + m * NT);
    }
    const auto wr0 = [&](int m) { return !MASK || m < M0 - 1 || tid + m * NT < A0P; };  // loader tail
    pair_qidx<TX, TY, 2, NT, M2>(s2, w3);
    pair_qidx<TX, TY, 3, NT, M3>(r3x, w4);
    pair_qidx<TX, TY, 4, NT, M4>(r4x, n4);

    T acc1[M1][2][4], acc2[M2][2][4], acc3[M3][2][4], acc4[M4][2][4];
    T d1[M1][2], d2v[M2][2], d3[M3][2], d4[M4][2];
    v2 qv2[M2];
    v2 c1[M1], c2[M2], c3[M3], c4[M4];                   // centres read this plane
    // q one and two planes back, in two-slot rings indexed by the plane's parity (R & 1 in
    // the unrolled loop): slot B holds the value stored two planes ago
    v2 r1[2][M1], r3[2][M3], r4[2][M4];
#pragma unroll
    for (int m = 0; m < M1; ++m) {
        r1[0][m] = r1[1][m] = v2{T(0), T(0)};
        for (int e = 0; e < 2; ++e) acc1[m][e][0] = acc1[m][e][1] = acc1[m][e][2] = acc1[m][e][3] = T(0);
    }
#pragma unroll
    for (int m = 0; m < M2; ++m)
        for (int e = 0; e < 2; ++e) acc2[m][e][0] = acc2[m][e][1] = acc2[m][e][2] = acc2[m][e][3] = T(0);
#pragma unroll
    for (int m = 0; m < M3; ++m) {
        r3[0][m] = r3[1][m] = v2{T(0), T(0)};
        for (int e = 0; e < 2; ++e) acc3[m][e][0] = acc3[m][e][1] = acc3[m][e][2] = acc3[m][e][3] = T(0);
    }
#pragma unroll
    for (int m = 0; m < M4; ++m) {
        r4[0][m] = r4[1][m] = v2{T(0), T(0)};
        for (int e = 0; e < 2; ++e) acc4[m][e][0] = acc4[m][e][1] = acc4[m][e][2] = acc4[m][e][3] = T(0);
    }

    // Same skewed schedule as k_rk4_fused: iteration t, stage 1: q t -> v1 t-2; stage 2:
    // v1 t-3 -> v2 t-5; stage 3: v2 t-6 -> v3 t-8; stage 4: v3 t-9 -> q' t-11.  Outputs
    // of a stage's first iterations are incomplete and never consumed by a later stage's
    // complete outputs (each stage starts when its first complete input plane exists).
    const int tq0 = z0 - 8, tq1 = zend + 8;
    const int tend = zend + 11;
    {
        const int zo = wrapz(tq0, nk) * iskb;
#pragma unroll
        for (int m = 0; m < M0; ++m)
            if (wr0(m)) *(v2*)&P0[tq0 & 1][2 * min(tid + m * NT, A0P - 1)] = PairIO<SI, T>::ld(rin, off0[m], zo);
        pair_qload<SI>(qv2, rin, wrapz(z0 - 8, nk) * iskb, q2);
    }
    v2 pre[M0];
    if (tq0 + 1 < tq1) {
        const int zo = wrapz(tq0 + 1, nk) * iskb;
#pragma unroll
        for (int m = 0; m < M0; ++m) pre[m] = PairIO<SI, T>::ld(rin, off0[m], zo);
    }
    __syncthreads();

    auto plane = [&](auto Rc, int t) {
        constexpr int R = decltype(Rc)::value;
        constexpr int B = R & 1;  // ring slot of this plane's parity
        const int cur = t & 1, prv = cur ^ 1;
        __builtin_amdgcn_sched_barrier(0);  // keep planes apart: bounded live ranges
        if (t + 1 < tq1) {
#pragma unroll
            for (int m = 0; m < M0; ++m)
                if (wr0(m)) *(v2*)&P0[prv][2 * min(tid + m * NT, A0P - 1)] = pre[m];
            if (t + 2 < tq1) {
                const int zo = wrapz(t + 2, nk) * iskb;
#pragma unroll
                for (int m = 0; m < M0; ++m) pre[m] = PairIO<SI, T>::ld(rin, off0[m], zo);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        if (t < tq1) {
            // centres of plane t are q(t); r1[B] = q(t-2) is the update's q
            pair_step<TX, TY, 1, NT, M1, R & 3>(P0[cur], w, acc1, d1, c1);
#pragma unroll
            for (int m = 0; m < M1; ++m) {
                if (!MASK || lv1[m])
                    *(v2*)&P1[cur][PM1::out(tid + m * NT)] =
                        v2{fmaT(w.f1, d1[m][0], r1[B][m].x), fmaT(w.f1, d1[m][1], r1[B][m].y)};
                r1[B][m] = c1[m];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (t >= tq0 + 5 && t < tq1 + 1) {
            pair_step<TX, TY, 2, NT, M2, (R + 1) & 3>(P1[prv], w, acc2, d2v, c2);
#pragma unroll
            for (int m = 0; m < M2; ++m) {
                if (!MASK || lv2[m]) {
                    *(v2*)&P2[cur][PM2::out(tid + m * NT)] =
                        v2{fmaT(w.f2, d2v[m][0], qv2[m].x), fmaT(w.f2, d2v[m][1], qv2[m].y)};
                    *(v2*)&Q3[cur][w3[m]] = qv2[m];  // q(t-5) for stage 3
                }
            }
            pair_qload<SI>(qv2, rin, wrapz(t - 4, nk) * iskb, q2);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (t >= tq0 + 10 && t < tq1 + 2) {
            pair_step<TX, TY, 3, NT, M3, (R + 2) & 3>(P2[prv], w, acc3, d3, c3);
#pragma unroll
            for (int m = 0; m < M3; ++m) {
                const v2 qn = *(const v2*)&Q3[prv][r3x[m]];  // q(t-6), written by stage 2 last plane
                if (!MASK || lv3[m]) {
                    *(v2*)&P3[cur][PM3::out(tid + m * NT)] =
                        v2{fmaT(w.f3, d3[m][0], r3[B][m].x), fmaT(w.f3, d3[m][1], r3[B][m].y)};
                    *(v2*)&Q4[cur][w4[m]] = r3[B][m];  // q(t-8) for stage 4
                }
                r3[B][m] = qn;
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (t >= tq0 + 15) {
            pair_step<TX, TY, 4, NT, M4, (R + 3) & 3>(P3[prv], w, acc4, d4, c4);
            const int zp = t - 11;
            if (zp >= z0 && zp < zend) {
                const int zo = zp * oskb;
#pragma unroll
                for (int m = 0; m < M4; ++m)
                    if (!MASK || lv4[m])
                        PairIO<SO, T>::st(rout, o4[m], zo,
                                          v2{fmaT(w.f4, d4[m][0], r4[B][m].x), fmaT(w.f4, d4[m][1], r4[B][m].y)});
            }
#pragma unroll
            for (int m = 0; m < M4; ++m) r4[B][m] = *(const v2*)&Q4[prv][r4x[m]];  // q(t-9)
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
    };

    // Stage s runs while its input planes are valid.  The loop is unrolled by 4 so that
    // the phase R of every plane is a compile-time constant.
    for (int t = tq0; t < tend; t += 4) {
        plane(std::integral_constant<int, 0>{}, t);
        if (t + 1 < tend) plane(std::integral_constant<int, 1>{}, t + 1);
        if (t + 2 < tend) plane(std::integral_constant<int, 2>{}, t + 2);
        if (t + 3 < tend) plane(std::integral_constant<int, 3>{}, t + 3);
    }
}

// Waits for every outstanding vector-memory access of this wave (s_waitcnt vmcnt(0), the
// other counters left alone).  The compiler's own waits treat loads and stores as completing
// out of order, so with a store pending every use of a loaded register waits for all loads
// issued so far, the prefetches included; one explicit wait where all the data a stretch of
// planes consumes is known to have been requested lets the loads issued after it stay in flight.
__device__ __forceinline__ void vm_wait_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// ---- FP64 RK4 step in two launches of a two-stage kernel ----------------------------------
// The four-stage stream fills its pipeline with 19 planes per block.  A slice small enough to
// stay in the last-level cache (one 128^3 slice: 16.8 MB) is chunked in k to give every CU
// a block, and then a block streams 35 planes for 16 outputs, its stage-1 region 2.4x its
// tile.  Split in two launches, each block streams kchunk + 9 planes over the regions of
// stages 3 and 4 (tile + 4, tile):
//   launch A: u = q,  v_a = q + f1 L u,  out = q + f2 L v_a   (= v2, into a work field)
//   launch B: u = v2, v_a = q + f3 L u,  out = q + f4 L v_a   (= q')
// The work field adds one write and two reads of a field per step, cheap while the fields
// stay cached.  Each stage is pair_step with the same operands in the same order as in the
// four-stage kernel, so q' equals that kernel's bit for bit.
// Schedule, iteration t: stage a reads u plane t and emits v_a(t-2) into LDS; stage b reads
// v_a(t-3) and emits out(t-5).  q for the updates: in launch A stage a takes it from the
// centres it read two planes before; otherwise it is loaded from the q field.
struct HalfArgs {
    FusedArgs a;       // a.in: u; a.out: the output
    View q;            // the update base (launch A: the same field as a.in)
    unsigned q_bytes;  // bytes one slice of q spans from its origin
};

template <int TX, int TY, int NT, bool FIRST>
__device__ __forceinline__ void rk4_half_pair_body(const HalfArgs& h, const int bid)
{
    const FusedArgs& a = h.a;
    using G3 = Geo<TX, TY, 3>;
    constexpr int A0P = G3::AI / 2;  // pairs of the u plane incl. its 4-cell halo
    constexpr int M0 = (A0P + NT - 1) / NT;
    using PMa = PairMap<TX, TY, 3, NT>;
    using PMb = PairMap<TX, TY, 4, NT>;
    constexpr int Ma = PMa::M, Mb = PMb::M;

    __shared__ __attribute__((aligned(16))) double P0[2][G3::AI];
    __shared__ __attribute__((aligned(16))) double Pa[2][Geo<TX, TY, 4>::AI];

    const int tid = threadIdx.x;
    const int ni = a.ni, nj = a.nj, nk = a.nk;
    const int tile = bid % a.tiles;
    const int chunk = (bid / a.tiles) % a.chunks;
    const int slice = bid / (a.tiles * a.chunks);
    const int x0 = (tile % a.tiles_x) * TX, y0 = (tile / a.tiles_x) * TY;
    const int z0 = chunk * a.kchunk;
    const int zend = min(z0 + a.kchunk, nk);
    const __amdgpu_buffer_rsrc_t rin = make_rsrc((const double*)a.in.p + slice * a.in.slice, a.in_bytes);
    const __amdgpu_buffer_rsrc_t rq = make_rsrc((const double*)h.q.p + slice * h.q.slice, h.q_bytes);
    const __amdgpu_buffer_rsrc_t rout = make_rsrc((double*)a.out.p + slice * a.out.slice, a.out_bytes);
    const int iskb = (int)a.in.sk * 8, qskb = (int)h.q.sk * 8, oskb = (int)a.out.sk * 8;
    const int isj = (int)a.in.sj, qsj = (int)h.q.sj, osj = (int)a.out.sj;
    const StarW<double> w = a.cdev ? load_w<double>(a.cdev[slice]) : load_w<double>(a.c);
    const double fa = FIRST ? w.f1 : w.f3, fb = FIRST ? w.f2 : w.f4;

    int off0[M0];
#pragma unroll
    for (int m = 0; m < M0; ++m) {
        const int g = min(tid + m * NT, A0P - 1);
        const int ax = 2 * (g % (G3::WI / 2)), by = g / (G3::WI / 2);
        off0[m] = (pmod(x0 - 4 + ax, ni) + pmod(y0 - 4 + by, nj) * isj) * 8;
    }
    int qa[Ma], qb[Mb], ob[Mb];
    if constexpr (!FIRST) pair_offsets<TX, TY, 3, NT, Ma>(qa, x0, y0, ni, nj, qsj);
    pair_offsets<TX, TY, 4, NT, Mb>(qb, x0, y0, ni, nj, qsj);
#pragma unroll
    for (int m = 0; m < Mb; ++m) {
        int ax, by;
        PMb::at(tid + m * NT, ax, by);
        ob[m] = (pmod(x0 + ax, ni) + pmod(y0 + by, nj) * osj) * 8;  // periodic image for partial tiles
    }
    constexpr bool MASK = TX * TY <= 512;
    bool lva[Ma], lvb[Mb];
    if constexpr (MASK) {
#pragma unroll
        for (int m = 0; m < Ma; ++m) lva[m] = PMa::live(tid + m * NT);
#pragma unroll
        for (int m = 0; m < Mb; ++m) lvb[m] = PMb::live(tid + m * NT);
    }
    const auto wr0 = [&](int m) { return !MASK || m < M0 - 1 || tid + m * NT < A0P; };

    double acca[Ma][2][4], accb[Mb][2][4];
    double da[Ma][2], db[Mb][2];
    d2 ca[Ma], cb[Mb];
    // Global loads are issued four planes before use, in rings of four slots indexed by
    // (plane - tq0) & 3, and waited for once every four planes (vm_wait_all at phase 0):
    // everything a plane consumes was loaded before that wait, so no other wait stalls.
    d2 ra[2][Ma];     // launch A: q one and two planes back (ring by plane parity)
    d2 pre[4][M0];    // u planes t+1 .. t+4
    d2 qva[4][Ma];    // launch B: q planes t-2 .. t+1 for stage a
    d2 qvb[4][Mb];    // q planes t-5 .. t-2 for stage b
#pragma unroll
    for (int m = 0; m < Ma; ++m) {
        ra[0][m] = ra[1][m] = d2{0., 0.};
        for (int e = 0; e < 2; ++e) acca[m][e][0] = acca[m][e][1] = acca[m][e][2] = acca[m][e][3] = 0.;
    }
#pragma unroll
    for (int m = 0; m < Mb; ++m)
        for (int e = 0; e < 2; ++e) accb[m][e][0] = accb[m][e][1] = accb[m][e][2] = accb[m][e][3] = 0.;

    const int tq0 = z0 - 4, tq1 = zend + 4;  // u planes read: tq0 .. tq1 - 1
    const int tend = zend + 5;
    {
        const int zo = wrapz(tq0, nk) * iskb;
#pragma unroll
        for (int m = 0; m < M0; ++m)
            if (wr0(m)) *(d2*)&P0[tq0 & 1][2 * min(tid + m * NT, A0P - 1)] = ld2(rin, off0[m], zo);
#pragma unroll
        for (int j = 1; j <= 4; ++j) {
            const int zj = wrapz(tq0 + j, nk) * iskb;  // planes past tq1 - 1 are read and not used
#pragma unroll
            for (int m = 0; m < M0; ++m) pre[j & 3][m] = ld2(rin, off0[m], zj);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if constexpr (!FIRST) {
                const int za = wrapz(tq0 - 2 + j, nk) * qskb;
#pragma unroll
                for (int m = 0; m < Ma; ++m) qva[(j + 2) & 3][m] = ld2(rq, qa[m], za);
            }
            const int zb = wrapz(tq0 + j, nk) * qskb;  // first used at tq0 + 5
#pragma unroll
            for (int m = 0; m < Mb; ++m) qvb[j][m] = ld2(rq, qb[m], zb);
        }
    }
    __syncthreads();

    auto plane = [&](auto Rc, int t) {
        constexpr int R = decltype(Rc)::value;
        constexpr int B = R & 1;
        constexpr int S1 = (R + 1) & 3, SA = (R + 2) & 3, SB = (R + 3) & 3;  // slots of planes t+1, t-2, t-5
        const int cur = t & 1, prv = cur ^ 1;
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (R == 0) vm_wait_all();
        if (t + 1 < tq1) {
#pragma unroll
            for (int m = 0; m < M0; ++m)
                if (wr0(m)) *(d2*)&P0[prv][2 * min(tid + m * NT, A0P - 1)] = pre[S1][m];
            if (t + 5 < tq1) {
                const int zo = wrapz(t + 5, nk) * iskb;
#pragma unroll
                for (int m = 0; m < M0; ++m) pre[S1][m] = ld2(rin, off0[m], zo);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        if (t < tq1) {
            pair_step<TX, TY, 3, NT, Ma, R & 3>(P0[cur], w, acca, da, ca);
#pragma unroll
            for (int m = 0; m < Ma; ++m) {
                const d2 q = FIRST ? ra[B][m] : qva[SA][m];
                if (!MASK || lva[m])
                    *(d2*)&Pa[cur][PMa::out(tid + m * NT)] = d2{fmaT(fa, da[m][0], q.x), fmaT(fa, da[m][1], q.y)};
                if constexpr (FIRST) ra[B][m] = ca[m];
            }
            if constexpr (!FIRST) {
                const int zo = wrapz(t + 2, nk) * qskb;
#pragma unroll
                for (int m = 0; m < Ma; ++m) qva[SA][m] = ld2(rq, qa[m], zo);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (t >= tq0 + 5) {
            pair_step<TX, TY, 4, NT, Mb, (R + 1) & 3>(Pa[prv], w, accb, db, cb);
            const int zp = t - 5;
            if (zp >= z0 && zp < zend) {
                const int zo = zp * oskb;
#pragma unroll
                for (int m = 0; m < Mb; ++m)
                    if (!MASK || lvb[m])
                        st2(rout, ob[m], zo,
                            d2{fmaT(fb, db[m][0], qvb[SB][m].x), fmaT(fb, db[m][1], qvb[SB][m].y)});
            }
            const int zo = wrapz(t - 1, nk) * qskb;
#pragma unroll
            for (int m = 0; m < Mb; ++m) qvb[SB][m] = ld2(rq, qb[m], zo);
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
    };

    for (int t = tq0; t < tend; t += 4) {
        plane(std::integral_constant<int, 0>{}, t);
        if (t + 1 < tend) plane(std::integral_constant<int, 1>{}, t + 1);
        if (t + 2 < tend) plane(std::integral_constant<int, 2>{}, t + 2);
        if (t + 3 < tend) plane(std::integral_constant<int, 3>{}, t + 3);
    }
}

// ---- FP64 2x2 variant (32x32 tiles, 512 threads) -----------------------------------------
// Stages 1-3 give each lane a 2x2 block of columns (an x pair in two consecutive rows).  The
// centres of one row are y neighbours of the other, so 10 ds_read_b128 per block and stage
// give what the pair layout reads with 14; the points per lane (4 + 4 + 4 + 2) and with
// them the register budget are the pair kernel's.  Stage 4 keeps the pair layout.
// Bank conflicts: a lane row of VW blocks steps W 16-byte chunks per block row (two plane
// rows of W/2 chunks), so a plane stride W = VW (mod 16) puts consecutive lanes on
// consecutive chunks across row ends, for reads (input plane) and writes (output plane)
// alike.  VW = 22 (stage 1's 22 blocks per row), plane strides 54 and 38 doubles.  In
// stages 2 and 3 the lanes past a row's last block are spare: they read the input plane's
// padding columns (so their groups keep the same chunk pattern) and write nothing.
// q: stage 1 keeps it in registers, stages 2 and 3 load it from global memory at lags 5
// and 8 (one plane ahead), stage 3 forwards it to stage 4 through the LDS plane Q4.
// The q plane for P0 is loaded two planes ahead: at one plane (~1.9 us at 16 x 128^3) the
// loads' HBM latency still showed, 4-9 % of the step.
struct QuadCfg {
    static constexpr int TX = 32, TY = 32, NT = 512, VW = 22;
    static constexpr int W0 = 54, W1 = 54, W2 = 54, W3 = 38;  // strides of P0..P3
    static constexpr int WQ = 38, HQ = 34;                     // Q4: stage 4's region + spare rows
    static constexpr int VW0 = 27;                             // lanes per row of P0's loader
};

template <int S>
struct QuadMap {
    using G = Geo<QuadCfg::TX, QuadCfg::TY, S>;
    static constexpr int U = G::WO / 2;  // blocks per row = block rows
    static constexpr int VW = QuadCfg::VW, SLOTS = VW * U;
    static constexpr int WIN = S == 1 ? QuadCfg::W0 : S == 2 ? QuadCfg::W1 : QuadCfg::W2;
    static constexpr int WOUT = S == 1 ? QuadCfg::W1 : S == 2 ? QuadCfg::W2 : QuadCfg::W3;
    static_assert(G::WO == G::HOo && U <= VW && SLOTS <= QuadCfg::NT, "one pass of square blocks");
    static_assert((WIN - VW) % 16 == 0 && (WOUT - VW) % 16 == 0, "conflict-free plane strides");
    static_assert(WIN >= G::WI && WOUT >= G::WO, "strides cover the regions");
    static_assert(WIN >= 2 * VW + 4, "spare lanes read inside the padded rows");
    // block of slot g: column 2*ux and rows 2*uy, 2*uy + 1 of the output region; ux >= U
    // marks a spare lane; lanes past the last slot repeat it (and write nothing)
    __device__ static void at(int g, int& ux, int& uy)
    {
        g = min(g, SLOTS - 1);
        uy = g / VW;
        ux = g % VW;
    }
    __device__ static bool live(int g)
    {
        int ux, uy;
        at(g, ux, uy);
        return ux < U;
    }
    __device__ static int centre(int g)  // first row's centre in the input plane
    {
        int ux, uy;
        at(g, ux, uy);
        return (2 * uy + 2) * WIN + 2 * ux + 2;
    }
    __device__ static int out(int g)
    {
        int ux, uy;
        at(g, ux, uy);
        return 2 * uy * WOUT + 2 * ux;
    }
};

// One stage at one plane for a 2x2 block; point p = 2 f + e (row f, column e).  Each group
// of LDS reads is issued at raised wave priority: of the two waves sharing a SIMD, the one
// about to wait on LDS gets its reads out first while the other computes (-3 to -4 % per
// step at 16 x 128^3, 8 x 160^3, 24 x 96^3).
template <int S, int R>
__device__ __forceinline__ void quad_step(const double* __restrict__ P, const StarW<double>& w, double (&acc)[4][4],
                                          double (&done)[4], d2 (&cen)[2])
{
    constexpr int W = QuadMap<S>::WIN;
    constexpr int sm2 = (R + 2) & 3, sm1 = (R + 3) & 3, s0 = R & 3, sp1 = (R + 1) & 3;
    const double* c = P + QuadMap<S>::centre((int)threadIdx.x);
    double ax[4];
#pragma unroll
    for (int f = 0; f < 2; ++f) {
        // centres and x neighbours of row f, and the z sums
        __builtin_amdgcn_s_setprio(1);
        const d2 C = *(const d2*)(c + f * W), Lx = *(const d2*)(c + f * W - 2), Rx = *(const d2*)(c + f * W + 2);
        __builtin_amdgcn_s_setprio(0);
        cen[f] = C;
        const double u[2] = {C.x, C.y};
        const double xm2[2] = {Lx.x, Lx.y}, xm1[2] = {Lx.y, C.x}, xp1[2] = {C.y, Rx.x}, xp2[2] = {Rx.x, Rx.y};
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int p = 2 * f + e;
            done[p] = fmaT(w.zp2, u[e], acc[p][sm2]);
            acc[p][sm2] = w.zm2 * u[e];
            acc[p][sm1] = fmaT(w.zp1, u[e], acc[p][sm1]);
            acc[p][sp1] = fmaT(w.zm1, u[e], acc[p][sp1]);
            double a = fmaT(w.c, u[e], acc[p][s0]);
            a = fmaT(w.xm2, xm2[e], a);
            a = fmaT(w.xm1, xm1[e], a);
            a = fmaT(w.xp1, xp1[e], a);
            ax[p] = fmaT(w.xp2, xp2[e], a);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    {
        // y neighbours: rows y-2, y-1, y+2, y+3 and the two centre rows
        __builtin_amdgcn_s_setprio(1);
        const d2 Ym2 = *(const d2*)(c - 2 * W), Ym1 = *(const d2*)(c - W);
        const d2 Yp2 = *(const d2*)(c + 2 * W), Yp3 = *(const d2*)(c + 3 * W);
        __builtin_amdgcn_s_setprio(0);
        const double r[6][2] = {{Ym2.x, Ym2.y}, {Ym1.x, Ym1.y}, {cen[0].x, cen[0].y},
                                {cen[1].x, cen[1].y}, {Yp2.x, Yp2.y}, {Yp3.x, Yp3.y}};
#pragma unroll
        for (int f = 0; f < 2; ++f)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                // row y+f: neighbours at rows (y+f) -2, -1, +1, +2 = r[f], r[f+1], r[f+3], r[f+4]
                double a = w.ym2 * r[f][e];
                a = fmaT(w.ym1, r[f + 1][e], a);
                a = fmaT(w.yp1, r[f + 3][e], a);
                a = fmaT(w.yp2, r[f + 4][e], a);
                acc[2 * f + e][s0] = ax[2 * f + e] + a;
            }
    }
    __builtin_amdgcn_sched_barrier(0);
}

// Global byte offsets of the two rows of slot g's block for stage S (periodic images).
template <int S>
__device__ __forceinline__ void quad_offsets(int (&q)[2], int x0, int y0, int ni, int nj, int sj)
{
    using G = Geo<QuadCfg::TX, QuadCfg::TY, S>;
    int ux, uy;
    QuadMap<S>::at((int)threadIdx.x, ux, uy);
    const int gx = pmod(x0 - G::HO + 2 * ux, ni);
#pragma unroll
    for (int f = 0; f < 2; ++f) q[f] = (gx + pmod(y0 - G::HO + 2 * uy + f, nj) * sj) * 8;
}

__device__ __forceinline__ void rk4_quad_body(const FusedArgs& a, const int bid)
{
    constexpr int NT = QuadCfg::NT, TX = QuadCfg::TX, TY = QuadCfg::TY;
    using G1 = Geo<TX, TY, 1>;
    using G4 = Geo<TX, TY, 4>;
    using PM4 = PairMap<TX, TY, 4, NT>;
    static_assert(PM4::M == 1 && PM4::VW == 16, "stage 4: one pass, one row per 16-lane group");
    constexpr int H0 = G1::HI;                         // rows of the q plane incl. halo
    constexpr int SLOTS0 = QuadCfg::VW0 * H0;          // loader slots (cols >= WI/2 fill padding)
    constexpr int M0 = (SLOTS0 + NT - 1) / NT;
    static_assert(QuadCfg::VW0 * 2 == QuadCfg::W0, "loader slots fill the q plane's rows contiguously");
    constexpr int W0 = QuadCfg::W0, W1 = QuadCfg::W1, W2 = QuadCfg::W2, W3 = QuadCfg::W3, WQ = QuadCfg::WQ;

    // q planes of even and odd phase (chunks start on even planes): two arrays, so the
    // compiler sees that stage 1's reads of one never alias the load into the other
    __shared__ __attribute__((aligned(16))) double P0a[H0 * W0];
    __shared__ __attribute__((aligned(16))) double P0b[H0 * W0];
    __shared__ __attribute__((aligned(16))) double P1[2][Geo<TX, TY, 2>::HI * W1];
    __shared__ __attribute__((aligned(16))) double P2[2][Geo<TX, TY, 3>::HI * W2];
    __shared__ __attribute__((aligned(16))) double P3[2][G4::HI * W3];
    __shared__ __attribute__((aligned(16))) double Q4[2][QuadCfg::HQ * WQ];

    const int tid = threadIdx.x;
    const int ni = a.ni, nj = a.nj, nk = a.nk;
    const int tile = bid % a.tiles;
    const int chunk = (bid / a.tiles) % a.chunks;
    const int slice = bid / (a.tiles * a.chunks);
    const int x0 = (tile % a.tiles_x) * TX, y0 = (tile / a.tiles_x) * TY;
    const int z0 = chunk * a.kchunk;
    const int zend = min(z0 + a.kchunk, nk);
    const __amdgpu_buffer_rsrc_t rin = make_rsrc((const double*)a.in.p + slice * a.in.slice, a.in_bytes);
    const __amdgpu_buffer_rsrc_t rout = make_rsrc((double*)a.out.p + slice * a.out.slice, a.out_bytes);
    const int iskb = (int)a.in.sk * 8, oskb = (int)a.out.sk * 8;
    const int isj = (int)a.in.sj, osj = (int)a.out.sj;
    const StarW<double> w = a.cdev ? load_w<double>(a.cdev[slice]) : load_w<double>(a.c);

    // q plane loader: VW0 lanes per row; lanes past the row's 24 pairs load the periodic
    // continuation into the row's padding (never read), which keeps the writes conflict-free
    // slot g of the loader is P0 element 2 g: wave w's m-th load fills slots
    // m NT + 64 w .. + 63, one contiguous kilobyte
    const int wave = tid >> 6;
    int off0[M0], dst0[M0];
#pragma unroll
    for (int m = 0; m < M0; ++m) {
        const int g = min(tid + m * NT, SLOTS0 - 1);
        const int col = g % QuadCfg::VW0, row = g / QuadCfg::VW0;
        off0[m] = (pmod(x0 - 8 + 2 * col, ni) + pmod(y0 - 8 + row, nj) * isj) * 8;
        dst0[m] = row * W0 + 2 * col;
    }
    const auto wr0 = [&](int m) { return m < M0 - 1 || tid + m * NT < SLOTS0; };  // loader tail: no write
    int q2[2], q3[2];
    quad_offsets<2>(q2, x0, y0, ni, nj, isj);
    quad_offsets<3>(q3, x0, y0, ni, nj, isj);
    const int o1 = QuadMap<1>::out(tid), o2 = QuadMap<2>::out(tid), o3 = QuadMap<3>::out(tid);
    static_assert(QuadMap<1>::U == QuadCfg::VW, "stage 1 has no spare lanes");
    // lanes past a stage's last slot repeat it and write nothing (several lanes writing one
    // LDS address serialise; reads of one address are a broadcast)
    const bool live1 = tid < QuadMap<1>::SLOTS;
    const bool live2 = tid < QuadMap<2>::SLOTS && QuadMap<2>::live(tid);
    const bool live3 = tid < QuadMap<3>::SLOTS && QuadMap<3>::live(tid);
    int wq;  // stage 3's block in Q4 (stage 4's region); blocks outside it go to spare cells
    {
        int ux, uy;
        QuadMap<3>::at(tid, ux, uy);
        int qx = 2 * ux - 2, qy = 2 * uy - 2;
        if (qx < 0) qx = G4::WO + 2;  // spare columns G4::WO .. WQ-1
        if (qy < 0) qy = G4::HOo;     // spare rows G4::HOo .. HQ-1
        wq = qy * WQ + qx;
    }
    int o4, r4;
    {
        int ax, by;
        PM4::at(tid, ax, by);
        o4 = (pmod(x0 + ax, ni) + pmod(y0 + by, nj) * osj) * 8;  // periodic image for partial tiles
        r4 = by * WQ + ax;
    }

    double acc1[4][4], acc2[4][4], acc3[4][4], acc4[1][2][4];
    double d1[4], d2v[4], d3[4], d4[1][2];
    d2 c1[2], c2[2], c3[2], c4[1];
    d2 qv2[2], qv3[2];
    d2 r1a[2], r1b[2], r4a, r4b;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc1[p][j] = acc2[p][j] = acc3[p][j] = 0.;
#pragma unroll
    for (int e = 0; e < 2; ++e) acc4[0][e][0] = acc4[0][e][1] = acc4[0][e][2] = acc4[0][e][3] = 0.;
    r1a[0] = r1a[1] = r1b[0] = r1b[1] = r4a = r4b = d2{0., 0.};

    // Skewed schedule of the pair kernel: iteration t, stage 1: q t -> v1 t-2; stage 2:
    // v1 t-3 -> v2 t-5; stage 3: v2 t-6 -> v3 t-8; stage 4: v3 t-9 -> q' t-11.
    const int tq0 = z0 - 8, tq1 = zend + 8;
    const int tend = zend + 11;
    {
        const int zo = wrapz(tq0, nk) * iskb;
#pragma unroll
        for (int m = 0; m < M0; ++m)
            if (wr0(m)) *(d2*)&P0a[dst0[m]] = ld2(rin, off0[m], zo);
        const int z2 = wrapz(z0 - 8, nk) * iskb, z3 = wrapz(z0 - 6, nk) * iskb;
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            qv2[f] = ld2(rin, q2[f], z2);
            qv3[f] = ld2(rin, q3[f], z3);
        }
    }
    __syncthreads();

    // One plane at phase R = (t - tq0) & 3.  Every plane issues the same global loads at the
    // same points, whichever stages are active (a stage's q load fetches the plane it needs
    // at the next plane, so loads ahead of its first active plane fetch the same values;
    // past tq1 they fetch valid wrapped planes that no stage reads): the compiler's memory
    // waits then count them exactly instead of waiting for every load and store of the
    // previous plane.  Only the output store depends on the plane.
    auto plane = [&](auto Rc, int t) {
        constexpr int R = decltype(Rc)::value;
        const int cur = t & 1, prv = cur ^ 1;
        double* const P0cur = (R & 1) ? P0b : P0a;  // q(t) (t - tq0 = R mod 4, tq0 even)
        double* const P0nxt = (R & 1) ? P0a : P0b;
        __builtin_amdgcn_sched_barrier(0);
        {
            // q plane t+1 into P0nxt (read by stage 1 at the next plane; the wait for it
            // comes before this plane's barrier)
            const int zo = wrapz(t + 1, nk) * iskb;
#pragma unroll
            for (int m = 0; m < M0; ++m)
                if (wr0(m)) ld2_to_lds(rin, &P0nxt[2 * (m * NT + wave * 64)], off0[m], zo);
        }
        __builtin_amdgcn_sched_barrier(0);
        if (t < tq1) {
            // centres of plane t are q(t); r1b = q(t-2) is the update's q
            quad_step<1, R & 3>(P0cur, w, acc1, d1, c1);
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                if (live1)
                    *(d2*)&P1[cur][o1 + f * W1] =
                        d2{fmaT(w.f1, d1[2 * f], r1b[f].x), fmaT(w.f1, d1[2 * f + 1], r1b[f].y)};
                r1b[f] = r1a[f];
                r1a[f] = c1[f];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (t >= tq0 + 5 && t < tq1 + 1) {
            quad_step<2, (R + 1) & 3>(P1[prv], w, acc2, d2v, c2);
            if (live2) {
#pragma unroll
                for (int f = 0; f < 2; ++f)
                    *(d2*)&P2[cur][o2 + f * W2] =
                        d2{fmaT(w.f2, d2v[2 * f], qv2[f].x), fmaT(w.f2, d2v[2 * f + 1], qv2[f].y)};
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        {
            const int zo = wrapz(t - 4, nk) * iskb;  // q(t-4): stage 2's q at the next plane
#pragma unroll
            for (int f = 0; f < 2; ++f) qv2[f] = ld2(rin, q2[f], zo);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (t >= tq0 + 10 && t < tq1 + 2) {
            quad_step<3, (R + 2) & 3>(P2[prv], w, acc3, d3, c3);
            if (live3) {
#pragma unroll
                for (int f = 0; f < 2; ++f) {
                    *(d2*)&P3[cur][o3 + f * W3] =
                        d2{fmaT(w.f3, d3[2 * f], qv3[f].x), fmaT(w.f3, d3[2 * f + 1], qv3[f].y)};
                    *(d2*)&Q4[cur][wq + f * WQ] = qv3[f];  // q(t-8) for stage 4
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        {
            const int zo = wrapz(t - 7, nk) * iskb;  // q(t-7): stage 3's q at the next plane
#pragma unroll
            for (int f = 0; f < 2; ++f) qv3[f] = ld2_last(rin, q3[f], zo);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (t >= tq0 + 15) {
            pair_step<TX, TY, 4, NT, 1, (R + 3) & 3, W3>(P3[prv], w, acc4, d4, c4);
            const int zp = t - 11;
            if (zp >= z0 && zp < zend)
                st2_stream(rout, o4, zp * oskb, d2{fmaT(w.f4, d4[0][0], r4b.x), fmaT(w.f4, d4[0][1], r4b.y)});
            r4b = r4a;
            r4a = *(const d2*)&Q4[prv][r4];  // q(t-9)
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
    };
    for (int t = tq0; t < tend; t += 4) {
        plane(std::integral_constant<int, 0>{}, t);
        if (t + 1 < tend) plane(std::integral_constant<int, 1>{}, t + 1);
        if (t + 2 < tend) plane(std::integral_constant<int, 2>{}, t + 2);
        if (t + 3 < tend) plane(std::integral_constant<int, 3>{}, t + 3);
    }
}

// One launch over one batch of slices.
template <typename SI, typename SO, typename T, int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) k_rk4_fused(FusedArgs a)
{
    rk4_scalar_body<SI, SO, T, TX, TY, NT>(a, xcd_block(blockIdx.x, gridDim.x));
}

template <typename SI, typename SO, typename T, int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) k_rk4_fused_pair(FusedArgs a)
{
    rk4_pair_body<SI, SO, T, TX, TY, NT>(a, xcd_block(blockIdx.x, gridDim.x));
}

__global__ void __launch_bounds__(QuadCfg::NT) k_rk4_fused_quad(FusedArgs a)
{
    rk4_quad_body(a, xcd_block(blockIdx.x, gridDim.x));
}

template <int TX, int TY, int NT, bool FIRST>
__global__ void __launch_bounds__(NT) k_rk4_half_pair(HalfArgs h)
{
    rk4_half_pair_body<TX, TY, NT, FIRST>(h, xcd_block(blockIdx.x, gridDim.x));
}

// One launch over several independent batches (problems of different sizes, steps and
// coefficients): entry e owns blocks [first[e], first[e+1]); a block finds its entry by a
// scalar scan (entries are few) and runs the same body on that entry's arguments.
__device__ __forceinline__ int group_entry(const GroupArgs& g, int bid)
{
    int e = 0;
    while (e + 1 < g.n && g.first[e + 1] <= bid) ++e;
    return e;
}

// The records are read through the constant address space: with a block-uniform index
// the loads are scalar, so the arguments live in scalar registers as kernel arguments do.
typedef __attribute__((address_space(4))) const FusedArgs ConstArgs;

template <typename SI, typename SO, typename T, int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) k_rk4_fused_grouped(GroupArgs g)
{
    const int bid = xcd_block(blockIdx.x, gridDim.x);
    const int e = group_entry(g, bid);
    ConstArgs* rec = (ConstArgs*)g.args + e;
    rk4_scalar_body<SI, SO, T, TX, TY, NT>(*(const FusedArgs*)rec, bid - g.first[e]);
}

template <typename SI, typename SO, typename T, int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) k_rk4_fused_pair_grouped(GroupArgs g)
{
    const int bid = xcd_block(blockIdx.x, gridDim.x);
    const int e = group_entry(g, bid);
    ConstArgs* rec = (ConstArgs*)g.args + e;
    rk4_pair_body<SI, SO, T, TX, TY, NT>(*(const FusedArgs*)rec, bid - g.first[e]);
}

__global__ void __launch_bounds__(QuadCfg::NT) k_rk4_fused_quad_grouped(GroupArgs g)
{
    const int bid = xcd_block(blockIdx.x, gridDim.x);
    const int e = group_entry(g, bid);
    ConstArgs* rec = (ConstArgs*)g.args + e;
    rk4_quad_body(*(const FusedArgs*)rec, bid - g.first[e]);
}

}  // namespace

Coefs make_coefs(double nu, double cx, double cy, double cz, double dx, double h)
{
    const Coef13 k = make_coef13(nu, cx, cy, cz, dx);
    Coefs c;
    const double w[13] = {k.x[0], k.x[1], k.x[2], k.x[3], k.y[0], k.y[1], k.y[2],
                          k.y[3], k.z[0], k.z[1], k.z[2], k.z[3], k.c};
    const double f[4] = {0.25 * h, h / 3., 0.5 * h, h};
    for (int i = 0; i < 13; ++i) {
        c.wd[i] = w[i];
        c.wf[i] = (float)w[i];
    }
    for (int i = 0; i < 4; ++i) {
        c.fd[i] = f[i];
        c.ff[i] = (float)f[i];
    }
    return c;
}

size_t coefs_size() { return sizeof(Coefs); }

void fill_coefs(void* dst, double nu, double cx, double cy, double cz, double dx, double h)
{
    *(Coefs*)dst = make_coefs(nu, cx, cy, cz, dx, h);
}

namespace {

int fill_args(FusedArgs& a, const FusedStep& st, const FusedTile& tile, View in, View out, int isz, int osz)
{
    a.in = in;
    a.out = out;
    const long long span_in = (st.ni - 1) + (st.nj - 1) * in.sj + (st.nk - 1) * in.sk + 1;
    const long long span_out = (st.ni - 1) + (st.nj - 1) * out.sj + (st.nk - 1) * out.sk + 1;
    const long long bi = span_in * isz, bo = span_out * osz;
    if (bi >= (1LL << 31) || bo >= (1LL << 31))
        return fail(SP_ERR_UNSUPPORTED, "fused RK4: one field must span less than 2 GiB");
    a.in_bytes = (unsigned)bi;
    a.out_bytes = (unsigned)bo;
    a.ni = st.ni;
    a.nj = st.nj;
    a.nk = st.nk;
    a.kchunk = tile.kchunk > 0 ? min(tile.kchunk, st.nk) : st.nk;
    // the 2x2 kernel's chunks start on even planes (its q-plane buffers alternate by plane
    // parity); the chunking does not change the results
    if (tile.variant == SP_KERNEL_QUAD && a.kchunk < st.nk && (a.kchunk & 1)) a.kchunk += 1;
    a.tiles_x = (st.ni + tile.tx - 1) / tile.tx;
    const int tiles_y = (st.nj + tile.ty - 1) / tile.ty;
    a.tiles = a.tiles_x * tiles_y;
    a.chunks = (st.nk + a.kchunk - 1) / a.kchunk;
    a.c = make_coefs(st.nu, st.cx, st.cy, st.cz, st.dx, st.dt);
    a.cdev = (const Coefs*)st.coef_dev;
    const long long nblocks = (long long)a.tiles * a.chunks * st.nslices;
    if (nblocks > 0x7fffffffLL) return fail(SP_ERR_UNSUPPORTED, "fused RK4: too many blocks");
    return SP_OK;
}

unsigned nblocks_of(const FusedArgs& a, int nslices) { return (unsigned)((long long)a.tiles * a.chunks * nslices); }

template <typename SI, typename SO, typename T>
int launch_typed(hipStream_t s, const FusedStep& st, const FusedTile& tile, View in, View out)
{
    FusedArgs a;
    SP_TRY(fill_args(a, st, tile, in, out, (int)sizeof(SI), (int)sizeof(SO)));
    const dim3 grid(nblocks_of(a, st.nslices));
#define SP_FUSED_CASE(TX_, TY_, NT_)                                                             \
    if (tile.tx == TX_ && tile.ty == TY_ && tile.nt == NT_) {                                    \
        hipLaunchKernelGGL((k_rk4_fused<SI, SO, T, TX_, TY_, NT_>), grid, dim3(NT_), 0, s, a);   \
        SP_HIP(hipGetLastError());                                                               \
        return SP_OK;                                                                            \
    }
    SP_FUSED_CASE(32, 32, 1024)
    SP_FUSED_CASE(32, 24, 1024)
    SP_FUSED_CASE(24, 24, 1024)
    SP_FUSED_CASE(32, 16, 512)
    SP_FUSED_CASE(16, 16, 256)
#undef SP_FUSED_CASE
    return fail(SP_ERR_UNSUPPORTED, "fused RK4: unsupported tile shape");
}

template <typename SI, typename SO, typename T>
int launch_pair(hipStream_t s, const FusedStep& st, const FusedTile& tile, View in, View out)
{
    FusedArgs a;
    SP_TRY(fill_args(a, st, tile, in, out, (int)sizeof(SI), (int)sizeof(SO)));
    const dim3 grid(nblocks_of(a, st.nslices));
#define SP_PAIR_CASE(TX_, TY_, NT_)                                                                  \
    if (tile.tx == TX_ && tile.ty == TY_ && tile.nt == NT_) {                                        \
        hipLaunchKernelGGL((k_rk4_fused_pair<SI, SO, T, TX_, TY_, NT_>), grid, dim3(NT_), 0, s, a); \
        SP_HIP(hipGetLastError());                                                                   \
        return SP_OK;                                                                                \
    }
    SP_PAIR_CASE(32, 32, 512)
    SP_PAIR_CASE(32, 16, 512)
    SP_PAIR_CASE(16, 16, 256)
#undef SP_PAIR_CASE
    return fail(SP_ERR_UNSUPPORTED, "fused RK4 (pair): unsupported tile shape");
}

int launch_quad(hipStream_t s, const FusedStep& st, const FusedTile& tile, View in, View out)
{
    SP_CHECK(tile.tx == QuadCfg::TX && tile.ty == QuadCfg::TY && tile.nt == QuadCfg::NT, SP_ERR_UNSUPPORTED,
             "fused RK4 (2x2): unsupported tile shape");
    FusedArgs a;
    SP_TRY(fill_args(a, st, tile, in, out, 8, 8));
    hipLaunchKernelGGL(k_rk4_fused_quad, dim3(nblocks_of(a, st.nslices)), dim3(QuadCfg::NT), 0, s, a);
    SP_HIP(hipGetLastError());
    return SP_OK;
}

// Launch A (in -> work: v2) and launch B (work, in -> out: q').  Requires st.work.
int launch_split(hipStream_t s, const FusedStep& st, const FusedTile& tile, View in, View out)
{
    SP_CHECK(st.work && st.work_bytes >= split_work_bytes(st.ni, st.nj, st.nk, st.nslices), SP_ERR_INVALID,
             "fused RK4 (split): no work field for this batch");
    View wv;
    wv.p = st.work;
    wv.sj = st.ni;
    wv.sk = (long long)st.ni * st.nj;
    wv.slice = wv.sk * st.nk;
    HalfArgs A, B;
    SP_TRY(fill_args(A.a, st, tile, in, wv, 8, 8));
    A.q = in;
    A.q_bytes = A.a.in_bytes;
    SP_TRY(fill_args(B.a, st, tile, wv, out, 8, 8));
    B.q = in;
    B.q_bytes = A.a.in_bytes;
    const dim3 grid(nblocks_of(A.a, st.nslices));
#define SP_SPLIT_CASE(TX_, TY_, NT_)                                                           \
    if (tile.tx == TX_ && tile.ty == TY_ && tile.nt == NT_) {                                  \
        hipLaunchKernelGGL((k_rk4_half_pair<TX_, TY_, NT_, true>), grid, dim3(NT_), 0, s, A);  \
        SP_HIP(hipGetLastError());                                                             \
        hipLaunchKernelGGL((k_rk4_half_pair<TX_, TY_, NT_, false>), grid, dim3(NT_), 0, s, B); \
        SP_HIP(hipGetLastError());                                                             \
        return SP_OK;                                                                          \
    }
    SP_SPLIT_CASE(32, 32, 512)
    SP_SPLIT_CASE(32, 16, 512)
    SP_SPLIT_CASE(32, 16, 256)
    SP_SPLIT_CASE(16, 16, 256)
#undef SP_SPLIT_CASE
    return fail(SP_ERR_UNSUPPORTED, "fused RK4 (split): unsupported tile shape");
}

// Pair access (two elements per lane) needs even x extent, even strides and origins aligned
// to two elements of each field's storage type (FP64: 16 bytes).
bool pair_ok(const FusedStep& st, View in, View out, int isz = 8, int osz = 8)
{
    const auto even = [](long long v) { return (v & 1) == 0; };
    return even(st.ni) && even(in.sj) && even(in.sk) && even(out.sj) && even(out.sk) && even(in.slice) &&
           even(out.slice) && ((uintptr_t)in.p % (2 * isz)) == 0 && ((uintptr_t)out.p % (2 * osz)) == 0;
}

template <typename SI, typename SO, typename T>
int launch_grouped_pair(hipStream_t s, const FusedTile& tile, GroupArgs g, long long nblocks)
{
    const dim3 grid((unsigned)nblocks);
#define SP_PAIR_CASE(TX_, TY_, NT_)                                                                            \
    if (tile.tx == TX_ && tile.ty == TY_ && tile.nt == NT_) {                                                  \
        hipLaunchKernelGGL((k_rk4_fused_pair_grouped<SI, SO, T, TX_, TY_, NT_>), grid, dim3(NT_), 0, s, g); \
        SP_HIP(hipGetLastError());                                                                             \
        return SP_OK;                                                                                          \
    }
    SP_PAIR_CASE(32, 32, 512)
    SP_PAIR_CASE(32, 16, 512)
    SP_PAIR_CASE(16, 16, 256)
#undef SP_PAIR_CASE
    return fail(SP_ERR_UNSUPPORTED, "grouped fused RK4 (pair): unsupported tile shape");
}

template <typename SI, typename SO, typename T>
int launch_grouped_typed(hipStream_t s, const FusedTile& tile, GroupArgs g, long long nblocks)
{
    const dim3 grid((unsigned)nblocks);
#define SP_FUSED_CASE(TX_, TY_, NT_)                                                                     \
    if (tile.tx == TX_ && tile.ty == TY_ && tile.nt == NT_) {                                            \
        hipLaunchKernelGGL((k_rk4_fused_grouped<SI, SO, T, TX_, TY_, NT_>), grid, dim3(NT_), 0, s, g);   \
        SP_HIP(hipGetLastError());                                                                       \
        return SP_OK;                                                                                    \
    }
    SP_FUSED_CASE(32, 32, 1024)
    SP_FUSED_CASE(32, 24, 1024)
    SP_FUSED_CASE(24, 24, 1024)
    SP_FUSED_CASE(32, 16, 512)
    SP_FUSED_CASE(16, 16, 256)
#undef SP_FUSED_CASE
    return fail(SP_ERR_UNSUPPORTED, "grouped fused RK4: unsupported tile shape");
}

}  // namespace

size_t fused_record_size() { return sizeof(FusedArgs); }

int fused_record_fill(void* rec, const FusedStep& st, const FusedTile& tile, int in_dtype, View in, int out_dtype,
                      View out, long long* nblocks)
{
    SP_CHECK(st.nk >= kFusedMinK, SP_ERR_UNSUPPORTED, "grouped fused RK4: ksize below the fused kernel's minimum");
    FusedArgs a;
    SP_TRY(fill_args(a, st, tile, in, out, dtype_size(in_dtype), dtype_size(out_dtype)));
    std::memcpy(rec, &a, sizeof(a));
    *nblocks = (long long)nblocks_of(a, st.nslices);
    return SP_OK;
}

bool fused_record_pair_ok(const FusedStep& st, View in, View out, int in_dtype, int out_dtype)
{
    return pair_ok(st, in, out, dtype_size(in_dtype), dtype_size(out_dtype));
}

int launch_rk4_grouped(hipStream_t s, const FusedTile& tile, int in_dtype, int out_dtype, const void* records,
                       const int* first, int n, long long nblocks)
{
    SP_CHECK(n >= 1 && nblocks >= 1 && nblocks <= 0x7fffffffLL, SP_ERR_INVALID, "grouped fused RK4: empty or too large");
    const GroupArgs g{(const FusedArgs*)records, first, n};
    if (tile.variant == SP_KERNEL_QUAD) {
        SP_CHECK(in_dtype == SP_F64 && out_dtype == SP_F64, SP_ERR_UNSUPPORTED, "grouped 2x2 kernel is FP64 only");
        SP_CHECK(tile.tx == QuadCfg::TX && tile.ty == QuadCfg::TY && tile.nt == QuadCfg::NT, SP_ERR_UNSUPPORTED,
                 "grouped fused RK4 (2x2): unsupported tile shape");
        hipLaunchKernelGGL(k_rk4_fused_quad_grouped, dim3((unsigned)nblocks), dim3(QuadCfg::NT), 0, s, g);
        SP_HIP(hipGetLastError());
        return SP_OK;
    }
    if (tile.variant == SP_KERNEL_PAIR) {
        if (in_dtype == SP_F64 && out_dtype == SP_F64) return launch_grouped_pair<double, double, double>(s, tile, g, nblocks);
        if (in_dtype == SP_F64 && out_dtype == SP_F32) return launch_grouped_pair<double, float, float>(s, tile, g, nblocks);
        if (in_dtype == SP_F32 && out_dtype == SP_F32) return launch_grouped_pair<float, float, float>(s, tile, g, nblocks);
        if (in_dtype == SP_F64 && out_dtype == SP_BF16) return launch_grouped_pair<double, bf16, float>(s, tile, g, nblocks);
        if (in_dtype == SP_BF16 && out_dtype == SP_BF16) return launch_grouped_pair<bf16, bf16, float>(s, tile, g, nblocks);
        return fail(SP_ERR_UNSUPPORTED, "grouped fused RK4 (pair): unsupported dtype pair");
    }
    if (in_dtype == SP_F64 && out_dtype == SP_F64) return launch_grouped_typed<double, double, double>(s, tile, g, nblocks);
    if (in_dtype == SP_F64 && out_dtype == SP_F32) return launch_grouped_typed<double, float, float>(s, tile, g, nblocks);
    if (in_dtype == SP_F32 && out_dtype == SP_F32) return launch_grouped_typed<float, float, float>(s, tile, g, nblocks);
    if (in_dtype == SP_F64 && out_dtype == SP_BF16) return launch_grouped_typed<double, bf16, float>(s, tile, g, nblocks);
    if (in_dtype == SP_BF16 && out_dtype == SP_BF16) return launch_grouped_typed<bf16, bf16, float>(s, tile, g, nblocks);
    return fail(SP_ERR_UNSUPPORTED, "grouped fused RK4: unsupported dtype pair");
}

bool fused_tile_supported(const FusedTile& t, int /*compute_f64*/)
{
    static const int scalar[][3] = {{32, 32, 1024}, {32, 24, 1024}, {24, 24, 1024}, {32, 16, 512}, {16, 16, 256}};
    static const int pair[][3] = {{32, 32, 512}, {32, 16, 512}, {16, 16, 256}};
    static const int split[][3] = {{32, 32, 512}, {32, 16, 512}, {32, 16, 256}, {16, 16, 256}};
    if (t.variant == SP_KERNEL_QUAD) return t.tx == QuadCfg::TX && t.ty == QuadCfg::TY && t.nt == QuadCfg::NT;
    if (t.variant == SP_KERNEL_SPLIT) {
        for (const auto& sh : split)
            if (t.tx == sh[0] && t.ty == sh[1] && t.nt == sh[2]) return true;
        return false;
    }
    if (t.variant == SP_KERNEL_PAIR) {
        for (const auto& sh : pair)
            if (t.tx == sh[0] && t.ty == sh[1] && t.nt == sh[2]) return true;
        return false;
    }
    for (const auto& sh : scalar)
        if (t.tx == sh[0] && t.ty == sh[1] && t.nt == sh[2]) return true;
    return false;
}

FusedTile default_tile(int ni, int nj, int nk, int nslices, int compute_f64)
{
    // Measured on MI355X (FP64; 16 x 128^3, 8 x 160^3, 24 x 96^3, 1 x 128^3): the 2x2
    // kernel (32x32 tiles of 512 threads) is fastest once the tiles alone give one block
    // per CU; below that, x-pair 32x16 tiles chunked in k fill the chip better.  FP32 / BF16
    // storage runs the x-pair kernel in the 2x2 kernel's place (16 x 128^3: 218 us against
    // 283 us for the scalar kernel's best tile; 1 x 384^3 in 64-plane chunks: 479 against
    // 574 us).  Odd x extents fall back to the scalar kernel at launch.
    const long long want = 256;  // one block per CU
    FusedTile t{32, 32, 512, 0, compute_f64 ? SP_KERNEL_QUAD : SP_KERNEL_PAIR};
    long long blocks = (long long)((ni + t.tx - 1) / t.tx) * ((nj + t.ty - 1) / t.ty) * nslices;
    // Long k columns (one 384^3 slice: 144 tiles): 64-plane chunks of 32x32 tiles give three
    // or more blocks per CU at a 19-plane pipeline fill (measured, FP64: 0.55 ms per 384^3
    // step, against 1.07 ms for whole-k 32x16 tiles, whose 288 blocks leave a ragged second wave).
    if (blocks < want && blocks * (nk / 64) >= 3 * want) {
        t.kchunk = 64;
        return t;
    }
    if (blocks < want) {
        t.ty = 16;
        t.variant = SP_KERNEL_PAIR;
        blocks = (long long)((ni + t.tx - 1) / t.tx) * ((nj + t.ty - 1) / t.ty) * nslices;
    }
    int chunks = (int)((want + blocks - 1) / blocks);
    if (chunks < 1) chunks = 1;
    int kc = (nk + chunks - 1) / chunks;
    if (kc < 16) kc = 16;
    t.kchunk = kc >= nk ? 0 : kc;
    return t;
}

void fused_candidates(int ni, int nj, int nk, int nslices, int compute_f64, std::vector<FusedTile>& out)
{
    static const int scalar[][3] = {{32, 32, 1024}, {32, 24, 1024}, {24, 24, 1024}, {32, 16, 512}, {16, 16, 256}};
    static const int pair[][3] = {{32, 32, 512}, {32, 16, 512}, {16, 16, 256}};
    static const int split[][3] = {{32, 32, 512}, {32, 16, 512}, {32, 16, 256}, {16, 16, 256}};
    std::vector<FusedTile> shapes;
    if (!compute_f64 && ni % 2 == 0)
        for (const auto& sh : pair) shapes.push_back(FusedTile{sh[0], sh[1], sh[2], 0, SP_KERNEL_PAIR});
    if (compute_f64 && ni % 2 == 0) {
        shapes.push_back(FusedTile{QuadCfg::TX, QuadCfg::TY, QuadCfg::NT, 0, SP_KERNEL_QUAD});
        for (const auto& sh : pair) shapes.push_back(FusedTile{sh[0], sh[1], sh[2], 0, SP_KERNEL_PAIR});
        // the split steps move a third field: candidates while a field of the batch stays
        // within 64 MiB (the last-level cache holds the three)
        if (split_work_bytes(ni, nj, nk, nslices) <= (64u << 20))
            for (const auto& sh : split) shapes.push_back(FusedTile{sh[0], sh[1], sh[2], 0, SP_KERNEL_SPLIT});
    }
    for (const auto& sh : scalar) shapes.push_back(FusedTile{sh[0], sh[1], sh[2], 0, SP_KERNEL_SCALAR});
    out.clear();
    for (const FusedTile& t : shapes) {
        const long long blocks = (long long)((ni + t.tx - 1) / t.tx) * ((nj + t.ty - 1) / t.ty) * nslices;
        // k chunks: the whole extent, and shorter ones (>= 16 planes, >= 8 for the split
        // kernel's 9-plane fill) only while the tiles alone give fewer than 4 blocks per CU
        const int cmin = t.variant == SP_KERNEL_SPLIT ? 8 : 16;
        std::vector<int> kcs{0};
        if (blocks < 1024)
            for (int c : {nk / 2, nk / 4, nk / 8, 64, 32, 16, 8})
                if (c >= cmin && c < nk && std::find(kcs.begin(), kcs.end(), c) == kcs.end()) kcs.push_back(c);
        for (int c : kcs) {
            FusedTile x = t;
            x.kchunk = c;
            out.push_back(x);
        }
    }
}

int launch_rk4_fused(hipStream_t s, const FusedStep& st, const FusedTile& tile_in, int in_dtype, View in,
                     int out_dtype, View out)
{
    FusedTile tile = tile_in;
    const bool f64 = in_dtype == SP_F64 && out_dtype == SP_F64;
    if (st.nk < kFusedMinK)
        return fail(SP_ERR_UNSUPPORTED, "fused RK4: ksize " + std::to_string(st.nk) + " < " +
                                            std::to_string(kFusedMinK) + " (use the staged path)");
    if (f64 && tile.variant == SP_KERNEL_QUAD && pair_ok(st, in, out)) return launch_quad(s, st, tile, in, out);
    if (f64 && tile.variant == SP_KERNEL_PAIR && pair_ok(st, in, out))
        return launch_pair<double, double, double>(s, st, tile, in, out);
    if (!f64 && tile.variant == SP_KERNEL_PAIR &&
        pair_ok(st, in, out, (int)dtype_size(in_dtype), (int)dtype_size(out_dtype))) {
        if (in_dtype == SP_F64 && out_dtype == SP_F32) return launch_pair<double, float, float>(s, st, tile, in, out);
        if (in_dtype == SP_F32 && out_dtype == SP_F32) return launch_pair<float, float, float>(s, st, tile, in, out);
        if (in_dtype == SP_F64 && out_dtype == SP_BF16) return launch_pair<double, bf16, float>(s, st, tile, in, out);
        if (in_dtype == SP_BF16 && out_dtype == SP_BF16) return launch_pair<bf16, bf16, float>(s, st, tile, in, out);
    }
    if (f64 && tile.variant == SP_KERNEL_SPLIT && pair_ok(st, in, out)) return launch_split(s, st, tile, in, out);
    if (tile.variant == SP_KERNEL_PAIR || tile.variant == SP_KERNEL_QUAD || tile.variant == SP_KERNEL_SPLIT) {
        // odd sizes, unaligned views or reduced-precision storage: the scalar kernel with
        // its fastest FP64 shape and the caller's k chunk
        tile.tx = 32;
        tile.ty = 16;
        tile.nt = 512;
        tile.variant = SP_KERNEL_SCALAR;
    }
    if (f64) return launch_typed<double, double, double>(s, st, tile, in, out);
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
