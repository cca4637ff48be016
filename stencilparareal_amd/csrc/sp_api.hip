// sp_api.hip -- the C ABI (include/stencilparareal.h): contexts, the Convection
// propagator, the periodic fill, the Parareal update and the batched single-GPU
// Parareal driver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "sp_handles.hpp"
#include "sp_internal.hpp"

namespace sp {

static thread_local std::string g_err;

void set_error(const std::string& m) { g_err = m; }
int fail(int code, const std::string& m)
{
    g_err = m;
    return code;
}

int check_layout(const sp_layout* l, const char* what)
{
    SP_CHECK(l, SP_ERR_INVALID, std::string(what) + ": null layout");
    SP_CHECK(l->ni >= 2 && l->nj >= 2 && l->nk >= 2, SP_ERR_INVALID, std::string(what) + ": sizes must be >= 2");
    SP_CHECK(l->halo >= 0, SP_ERR_INVALID, std::string(what) + ": negative halo");
    SP_CHECK(l->sj >= l->ni + 2LL * l->halo && l->sk >= l->sj * (l->nj + 2LL * l->halo), SP_ERR_INVALID,
             std::string(what) + ": strides smaller than the padded sizes");
    SP_CHECK(l->origin >= l->halo * (1 + l->sj + l->sk), SP_ERR_INVALID, std::string(what) + ": origin inside halo");
    SP_CHECK(l->dtype == SP_F64 || l->dtype == SP_F32 || l->dtype == SP_BF16, SP_ERR_INVALID,
             std::string(what) + ": unknown dtype");
    SP_CHECK(l->sk * (l->nk + 2LL * l->halo) < (1LL << 31), SP_ERR_UNSUPPORTED,
             std::string(what) + ": field exceeds 2^31 elements");
    return SP_OK;
}

}  // namespace sp

using namespace sp;

// Scratch of sp_do_rk4_noalloc: the one intermediate field of a multi-step or in-place run,
// in a file-scope device array instead of a library allocation.  Every device the library
// runs on holds its own copy from code-object load until the process ends; calls on one
// device share it (and keep its last contents), so calls that overlap on a device race.
constexpr long long kNoAllocElems = 1LL << 22;  // 4194304 points: FP64 fields up to 160^3
static __device__ __align__(256) double scratch[kNoAllocElems];

namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};  // per device: the device address of `scratch`

// With `device` current: looks up this device's copy of `scratch` once.  Host code uses only
// this address; the array's name on the host is a stand-in, not a device address.
int noalloc_scratch_setup(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices) return SP_OK;  // entry point refuses later
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return SP_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return sp::fail(SP_ERR_HIP, std::string("scratch lookup: ") + hipGetErrorString(e));
    g_noalloc_ptr[device] = ptr;
    return SP_OK;
}
}  // namespace

namespace {

double computenu(const sp_prop* p, double t) { return p->nu0 * (1. + .5 * std::sin(p->nufreq * t)); }

sp_layout compact(int ni, int nj, int nk, int dt)
{
    sp_layout l;
    l.ni = ni;
    l.nj = nj;
    l.nk = nk;
    l.halo = 0;
    l.sj = ni;
    l.sk = (long long)ni * nj;
    l.origin = 0;
    l.dtype = dt;
    return l;
}

View cview(void* p, int ni, int nj) { return View{p, ni, (long long)ni * nj, 0}; }

bool same_origin(const sp_layout* la, const void* a, const sp_layout* lb, const void* b)
{
    return (const char*)a + la->origin * dtype_size(la->dtype) == (const char*)b + lb->origin * dtype_size(lb->dtype);
}

int check_prop_layout(const sp_prop* p, const sp_layout* l, const void* ptr, const char* what)
{
    SP_TRY(check_layout(l, what));
    SP_CHECK(ptr, SP_ERR_INVALID, std::string(what) + ": null pointer");
    SP_CHECK(l->ni == p->ni && l->nj == p->nj && l->nk == p->nk, SP_ERR_INVALID,
             std::string(what) + ": size differs from the propagator's");
    return SP_OK;
}

FusedTile pick_tile(const sp_prop* p, int nslices, int f64)
{
    FusedTile t = default_tile(p->ni, p->nj, p->nk, nslices, f64);
    const auto hit = p->tuned.find({nslices, f64 ? 1 : 0});
    if (p->tile.tx == 0 && hit != p->tuned.end()) return hit->second;
    if (p->tile.tx > 0) {
        t.tx = p->tile.tx;
        t.ty = p->tile.ty;
        t.nt = p->tile.nt;
        t.variant = p->tile.variant;
    }
    if (p->tile.kchunk >= 0) t.kchunk = p->tile.kchunk;
    return t;
}

FusedStep fused_step(const sp_prop* p, int nslices, double dt, const void* coef_dev, double nu)
{
    FusedStep st;
    st.ni = p->ni;
    st.nj = p->nj;
    st.nk = p->nk;
    st.nslices = nslices;
    st.dt = dt;
    st.cx = p->cx;
    st.cy = p->cy;
    st.cz = p->cz;
    st.dx = p->dx;
    st.coef_dev = coef_dev;
    st.nu = nu;
    st.work = p->split.p;
    st.work_bytes = p->split.cap;
    return st;
}

// The work field a split tile needs for `nslices` slices; nothing for the other variants.
// Called where a tile is picked, before any launch is captured.
int ensure_split(sp_prop* p, const FusedTile& t, int nslices)
{
    if (t.variant != SP_KERNEL_SPLIT) return SP_OK;
    return p->split.ensure(split_work_bytes(p->ni, p->nj, p->nk, nslices));
}

// Autotuner: times every candidate tile / k chunk of the fused kernel for `nslices` slices
// (views in -> out, slice stride set) and records the fastest for pick_tile.  Synchronous;
// refuses to run while the stream is being captured.
// Autotuner results, process-wide (every propagator of the same sizes shares them) and,
// when SP_AUTOTUNE_CACHE names a file, across processes: one line per entry,
//   device ni nj nk nslices in_dtype out_dtype variant tx ty nthreads kchunk microseconds
struct TuneKey {
    int dev, ni, nj, nk, nslices, idt, odt;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(dev, ni, nj, nk, nslices, idt, odt) < std::tie(o.dev, o.ni, o.nj, o.nk, o.nslices, o.idt, o.odt);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, std::pair<FusedTile, double>> g_tune_memo;
bool g_tune_file_read = false;

void tune_cache_read_locked()
{
    if (g_tune_file_read) return;
    g_tune_file_read = true;
    const char* path = std::getenv("SP_AUTOTUNE_CACHE");
    if (!path || !*path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    TuneKey k;
    FusedTile t;
    double us;
    while (std::fscanf(f, "%d %d %d %d %d %d %d %d %d %d %d %d %lf", &k.dev, &k.ni, &k.nj, &k.nk, &k.nslices, &k.idt,
                       &k.odt, &t.variant, &t.tx, &t.ty, &t.nt, &t.kchunk, &us) == 13)
        if (fused_tile_supported(t, 1)) g_tune_memo[k] = {t, us};
    std::fclose(f);
}

void tune_cache_write_locked(const TuneKey& k, const FusedTile& t, double us)
{
    g_tune_memo[k] = {t, us};
    const char* path = std::getenv("SP_AUTOTUNE_CACHE");
    if (!path || !*path) return;
    if (FILE* f = std::fopen(path, "a")) {
        std::fprintf(f, "%d %d %d %d %d %d %d %d %d %d %d %d %.3f\n", k.dev, k.ni, k.nj, k.nk, k.nslices, k.idt, k.odt,
                     t.variant, t.tx, t.ty, t.nt, t.kchunk, us);
        std::fclose(f);
    }
}

// Mean launch time of tile t in microseconds: one untimed launch, then enough launches for
// about 2 ms of timed work (at least 3, at most 50; one when a launch takes over 20 ms).
int time_tile(hipStream_t s, const FusedStep& st, const FusedTile& t, int idt, View in, int odt, View out,
              hipEvent_t e0, hipEvent_t e1, double* us)
{
    float ms = 0.f;
    SP_TRY(launch_rk4_fused(s, st, t, idt, in, odt, out));
    SP_HIP(hipEventRecord(e0, s));
    SP_TRY(launch_rk4_fused(s, st, t, idt, in, odt, out));
    SP_HIP(hipEventRecord(e1, s));
    SP_HIP(hipEventSynchronize(e1));
    SP_HIP(hipEventElapsedTime(&ms, e0, e1));
    *us = ms * 1e3;
    if (ms >= 20.f) return SP_OK;
    const int reps = std::max(3, std::min(50, (int)(2000. / std::max(*us, 1.))));
    SP_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) SP_TRY(launch_rk4_fused(s, st, t, idt, in, odt, out));
    SP_HIP(hipEventRecord(e1, s));
    SP_HIP(hipEventSynchronize(e1));
    SP_HIP(hipEventElapsedTime(&ms, e0, e1));
    *us = ms * 1e3 / reps;
    return SP_OK;
}

// A result kept for these sizes and types (this process, or the cache file): sets the
// propagator's tile and returns true.
bool tune_memo_hit(sp_prop* p, int nslices, int idt, int odt, double* best_us)
{
    const TuneKey key{p->ctx->device, p->ni, p->nj, p->nk, nslices, idt, odt};
    std::lock_guard<std::mutex> lock(g_tune_mu);
    tune_cache_read_locked();
    const auto hit = g_tune_memo.find(key);
    if (hit == g_tune_memo.end()) return false;
    p->tuned[{nslices, idt == SP_F64 && odt == SP_F64 ? 1 : 0}] = hit->second.first;
    if (best_us) *best_us = hit->second.second;
    return true;
}

int tune_fused(sp_prop* p, int nslices, int idt, int odt, View in, View out, double* best_us)
{
    hipStream_t s = p->ctx->stream;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    SP_HIP(hipStreamIsCapturing(s, &cs));
    SP_CHECK(cs == hipStreamCaptureStatusNone, SP_ERR_INVALID, "autotune: stream is being captured");
    const int f64 = idt == SP_F64 && odt == SP_F64;
    const TuneKey key{p->ctx->device, p->ni, p->nj, p->nk, nslices, idt, odt};
    if (tune_memo_hit(p, nslices, idt, odt, best_us)) return SP_OK;
    std::vector<FusedTile> cand;
    fused_candidates(p->ni, p->nj, p->nk, nslices, f64, cand);
    for (const FusedTile& t : cand) SP_TRY(ensure_split(p, t, nslices));
    hipEvent_t e0, e1;
    SP_HIP(hipEventCreate(&e0));
    SP_HIP(hipEventCreate(&e1));
    const FusedStep st = fused_step(p, nslices, 1e-6, nullptr, p->nu0);
    std::vector<std::pair<double, size_t>> times;
    int r = SP_OK;
    for (size_t i = 0; i < cand.size() && r == SP_OK; ++i) {
        double us = 0.;
        r = time_tile(s, st, cand[i], idt, in, odt, out, e0, e1, &us);
        times.push_back({us, i});
    }
    // the two fastest once more, so a single noisy timing does not decide
    if (r == SP_OK && times.size() > 1) {
        std::sort(times.begin(), times.end());
        for (int j = 0; j < 2 && r == SP_OK; ++j) {
            double us = 0.;
            r = time_tile(s, st, cand[times[j].second], idt, in, odt, out, e0, e1, &us);
            times[j].first = std::min(times[j].first, us);
        }
        std::sort(times.begin(), times.begin() + 2);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    SP_TRY(r);
    const FusedTile best_t = cand[times.front().second];
    const double best = times.front().first;
    p->tuned[{nslices, f64}] = best_t;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        tune_cache_write_locked(key, best_t, best);
    }
    if (best_us) *best_us = best;
    return SP_OK;
}

// RK4 steps with the fused kernel on one field; in/out views and dtypes given.
int rk4_fused_run(sp_prop* p, int idt, View vin, int odt, View vout, bool aliased, double t0, double dt, int steps)
{
    hipStream_t s = p->ctx->stream;
    const size_t n = (size_t)p->ni * p->nj * p->nk;
    const int sdt = odt;  // intermediate storage precision = output precision
    const size_t bytes = n * dtype_size(sdt);
    SP_TRY(p->scratch[0].ensure(bytes));
    if (aliased) SP_TRY(p->scratch[1].ensure(bytes));
    const FusedTile tile = pick_tile(p, 1, idt == SP_F64 && odt == SP_F64);
    SP_TRY(ensure_split(p, tile, 1));
    const View A = cview(p->scratch[0].p, p->ni, p->nj), B = cview(p->scratch[1].p, p->ni, p->nj);
    View src = vin;
    int sdtype = idt;
    for (int i = 0; i < steps; ++i) {
        View dst;
        int ddt = sdt;
        const int left = steps - 1 - i;  // steps remaining after this one
        if (!aliased) {
            dst = (left % 2 == 0) ? vout : A;
        } else {
            if (left == 0 && steps > 1) dst = vout;
            else dst = (i % 2 == 0) ? A : B;
        }
        const double t = t0 + i * dt;  // Convection.cpp:302: tStart + (++i)*dt
        const double nu = computenu(p, t);
        SP_TRY(launch_rk4_fused(s, fused_step(p, 1, dt, nullptr, nu), tile, sdtype, src, ddt, dst));
        src = dst;
        sdtype = ddt;
    }
    if (aliased && steps == 1) SP_TRY(launch_convert(s, p->ni, p->nj, p->nk, 1, sdt, A, odt, vout));
    return SP_OK;
}

// One reference-structured RK4 step (Convection.cpp:318-370), f64, in == out allowed.
int rk4_staged_step(sp_prop* p, View vin, View vout, double t, double dt)
{
    hipStream_t s = p->ctx->stream;
    const size_t bytes = (size_t)p->ni * p->nj * p->nk * 8;
    SP_TRY(p->qI.ensure(bytes));
    SP_TRY(p->k1.ensure(bytes));
    SP_TRY(p->k2.ensure(bytes));
    SP_TRY(p->k3.ensure(bytes));
    SP_TRY(p->k4.ensure(bytes));
    const StagedParams P{p->ni, p->nj, p->nk, p->cx, p->cy, p->cz, p->dx, p->dx * p->dx};
    const View qI = cview(p->qI.p, p->ni, p->nj), k1 = cview(p->k1.p, p->ni, p->nj), k2 = cview(p->k2.p, p->ni, p->nj),
               k3 = cview(p->k3.p, p->ni, p->nj), k4 = cview(p->k4.p, p->ni, p->nj);
    const double dthalf = dt * .5;
    const double nu = computenu(p, t);
    SP_TRY(launch_rhs4(s, P, vin, k1, nu));
    SP_TRY(launch_euler_stage(s, P, vin, k1, qI, dthalf));
    SP_TRY(launch_rhs4(s, P, qI, k2, nu));
    SP_TRY(launch_euler_stage(s, P, vin, k2, qI, dthalf));
    SP_TRY(launch_rhs4(s, P, qI, k3, nu));
    SP_TRY(launch_euler_stage(s, P, vin, k3, qI, dt));
    SP_TRY(launch_rhs4(s, P, qI, k4, nu));
    SP_TRY(launch_rk_stage(s, P, vin, k1, k2, k3, k4, vout, dt));
    return SP_OK;
}

// Forward Euler steps (DoEuler), f64, with ping-pong so the fused sweep never runs in place.
int euler_run(sp_prop* p, View vin, View vout, bool aliased, double t0, double dt, int steps, int nslices,
              View scrA, View scrB)
{
    hipStream_t s = p->ctx->stream;
    const StagedParams P{p->ni, p->nj, p->nk, p->cx, p->cy, p->cz, p->dx, p->dx * p->dx};
    View src = vin;
    for (int i = 0; i < steps; ++i) {
        const int left = steps - 1 - i;
        View dst;
        if (!aliased) dst = (left % 2 == 0) ? vout : scrA;
        else if (left == 0 && steps > 1) dst = vout;
        else dst = (i % 2 == 0) ? scrA : scrB;
        const double nu = computenu(p, t0 + i * dt);
        SP_TRY(launch_euler_fused(s, P, src, dst, nu, dt, p->upwind, nslices));
        src = dst;
    }
    if (aliased && steps == 1) SP_TRY(launch_convert(s, p->ni, p->nj, p->nk, nslices, SP_F64, scrA, SP_F64, vout));
    return SP_OK;
}

}  // namespace

// ======================================================================================
extern "C" {

const char* sp_version(void) { return "stencilparareal-gfx950 0.1"; }
const char* sp_last_error(void) { return g_err.c_str(); }

int sp_ctx_create(int device, sp_ctx** out)
{
    SP_CHECK(out, SP_ERR_INVALID, "sp_ctx_create: null out");
    int n = 0;
    SP_HIP(hipGetDeviceCount(&n));
    SP_CHECK(device >= 0 && device < n, SP_ERR_INVALID, "sp_ctx_create: no such device");
    SP_HIP(hipSetDevice(device));
    SP_TRY(noalloc_scratch_setup(device));
    sp_ctx* c = new sp_ctx;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(SP_ERR_HIP, "hipStreamCreate failed");
    }
    c->stream = c->own;
    *out = c;
    return SP_OK;
}

static void ctx_release(sp_ctx* c)
{
    if (!c || --c->refs > 0) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->red) (void)hipFree(c->red);
    if (c->red_host) (void)hipHostFree(c->red_host);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
}

int sp_ctx_destroy(sp_ctx* c)
{
    ctx_release(c);
    return SP_OK;
}

int sp_ctx_set_stream(sp_ctx* c, void* s)
{
    SP_CHECK(c, SP_ERR_INVALID, "sp_ctx_set_stream: null context");
    c->stream = (hipStream_t)s;  // NULL: HIP's null stream
    return SP_OK;
}

int sp_ctx_get_stream(sp_ctx* c, void** s)
{
    SP_CHECK(c && s, SP_ERR_INVALID, "sp_ctx_get_stream: null argument");
    *s = (void*)c->stream;
    return SP_OK;
}

int sp_ctx_sync(sp_ctx* c)
{
    SP_CHECK(c, SP_ERR_INVALID, "sp_ctx_sync: null context");
    SP_HIP(hipStreamSynchronize(c->stream));
    return SP_OK;
}

int sp_malloc(sp_ctx* c, long long bytes, void** out)
{
    SP_CHECK(c && out && bytes > 0, SP_ERR_INVALID, "sp_malloc: bad argument");
    SP_HIP(hipSetDevice(c->device));
    if (hipMalloc(out, (size_t)bytes) != hipSuccess) return fail(SP_ERR_NOMEM, "sp_malloc: out of device memory");
    return SP_OK;
}

int sp_free(sp_ctx* c, void* p)
{
    SP_CHECK(c, SP_ERR_INVALID, "sp_free: null context");
    if (p) SP_HIP(hipFree(p));
    return SP_OK;
}

int sp_memcpy_h2d(sp_ctx* c, void* dst, const void* src, long long bytes)
{
    SP_CHECK(c && dst && src && bytes >= 0, SP_ERR_INVALID, "sp_memcpy_h2d: bad argument");
    SP_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, c->stream));
    SP_HIP(hipStreamSynchronize(c->stream));
    return SP_OK;
}

int sp_memcpy_d2h(sp_ctx* c, void* dst, const void* src, long long bytes)
{
    SP_CHECK(c && dst && src && bytes >= 0, SP_ERR_INVALID, "sp_memcpy_d2h: bad argument");
    SP_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, c->stream));
    SP_HIP(hipStreamSynchronize(c->stream));
    return SP_OK;
}

int sp_layout_compact(int ni, int nj, int nk, int dtype, sp_layout* out)
{
    SP_CHECK(out, SP_ERR_INVALID, "sp_layout_compact: null out");
    *out = compact(ni, nj, nk, dtype);
    return check_layout(out, "sp_layout_compact");
}

int sp_layout_stella_cuda(int ni, int nj, int nk, sp_layout* out)
{
    SP_CHECK(out, SP_ERR_INVALID, "sp_layout_stella_cuda: null out");
    const int h = 2;
    const long long pi = ((ni + 2 * h + 31) / 32) * 32LL;
    out->ni = ni;
    out->nj = nj;
    out->nk = nk;
    out->halo = h;
    out->sj = pi;
    out->sk = pi * (nj + 2 * h);
    out->origin = h + h * out->sj + h * out->sk;
    out->dtype = SP_F64;
    return check_layout(out, "sp_layout_stella_cuda");
}

long long sp_layout_storage_elems(const sp_layout* l)
{
    if (!l) return -1;
    return l->origin + (l->ni + l->halo) + (long long)(l->nj - 1 + l->halo) * l->sj +
           (long long)(l->nk - 1 + l->halo) * l->sk;
}

int sp_prop_create(sp_ctx* c, int isize, int jsize, int ksize, double dx, double nu0, double nufreq, double cx,
                   double cy, double cz, sp_prop** out)
{
    SP_CHECK(c && out, SP_ERR_INVALID, "sp_prop_create: null argument");
    SP_CHECK(isize >= 2 && jsize >= 2 && ksize >= 2, SP_ERR_INVALID, "sp_prop_create: sizes must be >= 2");
    SP_CHECK(dx > 0 && std::isfinite(dx), SP_ERR_INVALID, "sp_prop_create: dx must be positive");
    sp_prop* p = new sp_prop;
    p->ctx = c;
    c->refs++;
    p->ni = isize;
    p->nj = jsize;
    p->nk = ksize;
    p->dx = dx;
    p->nu0 = nu0;
    p->nufreq = nufreq;
    p->cx = cx;
    p->cy = cy;
    p->cz = cz;
    // the fused kernel's k stream needs nk >= kFusedMinK; thinner grids run the
    // reference-order sweeps
    if (ksize < kFusedMinK) p->path = SP_PATH_STAGED;
    *out = p;
    return SP_OK;
}

static void prop_release(sp_prop* p)
{
    if (!p || --p->refs > 0) return;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    for (Buf* b : {&p->scratch[0], &p->scratch[1], &p->split, &p->qI, &p->k1, &p->k2, &p->k3, &p->k4, &p->nu_tab})
        b->release();
    sp_ctx* c = p->ctx;
    delete p;
    ctx_release(c);
}

int sp_prop_destroy(sp_prop* p)
{
    prop_release(p);
    return SP_OK;
}

int sp_prop_set_upwind(sp_prop* p, int upwind)
{
    SP_CHECK(p, SP_ERR_INVALID, "sp_prop_set_upwind: null propagator");
    p->upwind = upwind ? 1 : 0;
    return SP_OK;
}

int sp_prop_set_path(sp_prop* p, int path)
{
    SP_CHECK(p, SP_ERR_INVALID, "sp_prop_set_path: null propagator");
    SP_CHECK(path == SP_PATH_FUSED || path == SP_PATH_STAGED, SP_ERR_INVALID, "sp_prop_set_path: unknown path");
    SP_CHECK(path == SP_PATH_STAGED || p->nk >= kFusedMinK, SP_ERR_UNSUPPORTED,
             "sp_prop_set_path: the fused path needs ksize >= " + std::to_string(kFusedMinK));
    p->path = path;
    return SP_OK;
}

int sp_prop_set_tile(sp_prop* p, int variant, int tx, int ty, int nthreads, int kchunk)
{
    SP_CHECK(p, SP_ERR_INVALID, "sp_prop_set_tile: null propagator");
    if (variant == SP_KERNEL_AUTO) {
        p->tile = FusedTile{0, 0, 0, kchunk > 0 ? kchunk : -1, SP_KERNEL_AUTO};
        return SP_OK;
    }
    SP_CHECK(variant == SP_KERNEL_SCALAR || variant == SP_KERNEL_PAIR || variant == SP_KERNEL_QUAD ||
                 variant == SP_KERNEL_SPLIT,
             SP_ERR_INVALID, "sp_prop_set_tile: unknown variant");
    const FusedTile t{tx, ty, nthreads, kchunk, variant};
    SP_CHECK(fused_tile_supported(t, 1), SP_ERR_UNSUPPORTED, "sp_prop_set_tile: tile shape not compiled");
    SP_CHECK(kchunk >= 0, SP_ERR_INVALID, "sp_prop_set_tile: negative kchunk");
    p->tile = t;
    return SP_OK;
}

int sp_prop_get_tile(const sp_prop* p, int nslices, int f64, int* variant, int* tx, int* ty, int* nthreads,
                     int* kchunk)
{
    SP_CHECK(p, SP_ERR_INVALID, "sp_prop_get_tile: null propagator");
    SP_CHECK(nslices >= 1, SP_ERR_INVALID, "sp_prop_get_tile: nslices < 1");
    FusedTile t = pick_tile(p, nslices, f64 ? 1 : 0);
    // the pair kernel needs an even x extent (as every layout then has even strides); the
    // 2x2 and split kernels also FP64
    if (t.variant != SP_KERNEL_SCALAR && (p->ni % 2 != 0 || (!f64 && t.variant != SP_KERNEL_PAIR)))
        t = FusedTile{32, 16, 512, t.kchunk, SP_KERNEL_SCALAR};
    if (variant) *variant = t.variant;
    if (tx) *tx = t.tx;
    if (ty) *ty = t.ty;
    if (nthreads) *nthreads = t.nt;
    if (kchunk) *kchunk = t.kchunk;
    return SP_OK;
}

int sp_prop_autotune(sp_prop* p, int nslices, int in_dtype, int out_dtype, double* best_us)
{
    SP_CHECK(p, SP_ERR_INVALID, "sp_prop_autotune: null propagator");
    SP_CHECK(nslices >= 1, SP_ERR_INVALID, "sp_prop_autotune: nslices < 1");
    SP_CHECK(in_dtype >= SP_F64 && in_dtype <= SP_BF16 && out_dtype >= SP_F64 && out_dtype <= SP_BF16,
             SP_ERR_INVALID, "sp_prop_autotune: dtype");
    SP_CHECK(p->nk >= kFusedMinK, SP_ERR_UNSUPPORTED, "sp_prop_autotune: ksize below the fused kernel's minimum");
    SP_HIP(hipSetDevice(p->ctx->device));
    if (tune_memo_hit(p, nslices, in_dtype, out_dtype, best_us)) return SP_OK;  // no buffers needed
    const size_t n = (size_t)p->ni * p->nj * p->nk;
    Buf a, b;
    int r = a.ensure(n * nslices * dtype_size(in_dtype));
    if (r == SP_OK) r = b.ensure(n * nslices * dtype_size(out_dtype));
    if (r == SP_OK && hipMemsetAsync(a.p, 0, a.cap, p->ctx->stream) != hipSuccess)
        r = fail(SP_ERR_HIP, "sp_prop_autotune: clear");
    if (r == SP_OK) {
        View vi = cview(a.p, p->ni, p->nj), vo = cview(b.p, p->ni, p->nj);
        vi.slice = vo.slice = (long long)n;
        r = tune_fused(p, nslices, in_dtype, out_dtype, vi, vo, best_us);
    }
    (void)hipStreamSynchronize(p->ctx->stream);
    a.release();
    b.release();
    return r;
}

static int rk4_common(sp_prop* p, const sp_layout* lin, const void* in, const sp_layout* lout, void* out,
                      double tStart, double dt, int steps, bool f64_only, const char* name)
{
    SP_CHECK(p, SP_ERR_INVALID, std::string(name) + ": null propagator");
    SP_TRY(check_prop_layout(p, lin, in, name));
    SP_TRY(check_prop_layout(p, lout, out, name));
    SP_CHECK(steps >= 0, SP_ERR_INVALID, std::string(name) + ": steps < 0");
    SP_CHECK(std::isfinite(dt) && std::isfinite(tStart), SP_ERR_INVALID, std::string(name) + ": non-finite time");
    if (f64_only)
        SP_CHECK(lin->dtype == SP_F64 && lout->dtype == SP_F64, SP_ERR_INVALID,
                 std::string(name) + ": fields must be SP_F64");
    if (steps == 0) return SP_OK;  // Convection.cpp:296: out untouched
    SP_HIP(hipSetDevice(p->ctx->device));
    const View vin = view_of(*lin, in), vout = view_of(*lout, out);
    const bool aliased = same_origin(lin, in, lout, out);
    if (p->path == SP_PATH_STAGED && lin->dtype == SP_F64 && lout->dtype == SP_F64) {
        SP_TRY(rk4_staged_step(p, vin, vout, tStart, dt));
        for (int i = 1; i < steps; ++i) SP_TRY(rk4_staged_step(p, vout, vout, tStart + i * dt, dt));
        return SP_OK;
    }
    return rk4_fused_run(p, lin->dtype, vin, lout->dtype, vout, aliased, tStart, dt, steps);
}

int sp_do_rk4(sp_prop* p, const sp_layout* lin, const void* in, const sp_layout* lout, void* out, double tStart,
              double dt, int steps)
{
    return rk4_common(p, lin, in, lout, out, tStart, dt, steps, true, "sp_do_rk4");
}

// sp_do_rk4 without allocation: the fused kernels on FP64 fields, the intermediate field in
// this device's `scratch` (so the launches are those of sp_do_rk4: bit-for-bit its result).
int sp_do_rk4_noalloc(sp_prop* p, const sp_layout* lin, const void* in, const sp_layout* lout, void* out,
                      double tStart, double dt, int steps)
{
    const char* name = "sp_do_rk4_noalloc";
    SP_CHECK(p, SP_ERR_INVALID, std::string(name) + ": null propagator");
    SP_TRY(check_prop_layout(p, lin, in, name));
    SP_TRY(check_prop_layout(p, lout, out, name));
    SP_CHECK(steps >= 0, SP_ERR_INVALID, std::string(name) + ": steps < 0");
    SP_CHECK(std::isfinite(dt) && std::isfinite(tStart), SP_ERR_INVALID, std::string(name) + ": non-finite time");
    SP_CHECK(lin->dtype == SP_F64 && lout->dtype == SP_F64, SP_ERR_INVALID,
             std::string(name) + ": fields must be SP_F64");
    SP_CHECK(p->path == SP_PATH_FUSED && p->nk >= kFusedMinK, SP_ERR_UNSUPPORTED,
             std::string(name) + ": runs the fused kernels only (the reference-order path and k extents below " +
                 std::to_string(kFusedMinK) + " need stage fields; use sp_do_rk4)");
    const long long n = (long long)p->ni * p->nj * p->nk;
    SP_CHECK(n <= kNoAllocElems, SP_ERR_UNSUPPORTED,
             std::string(name) + ": a field of " + std::to_string(n) + " points exceeds the scratch of " +
                 std::to_string(kNoAllocElems) + " points (use sp_do_rk4)");
    const int dev = p->ctx->device;
    SP_CHECK(dev >= 0 && dev < kNoAllocMaxDevices && g_noalloc_ptr[dev], SP_ERR_UNSUPPORTED,
             std::string(name) + ": no scratch on device " + std::to_string(dev));
    if (steps == 0) return SP_OK;  // Convection.cpp:296: out untouched
    SP_HIP(hipSetDevice(dev));
    hipStream_t s = p->ctx->stream;
    const View vin = view_of(*lin, in), vout = view_of(*lout, out);
    const bool aliased = same_origin(lin, in, lout, out);
    const View A = cview(g_noalloc_ptr[dev], p->ni, p->nj);
    FusedTile tile = pick_tile(p, 1, 1);
    // the split steps need a second intermediate field: the built-in one-launch tile instead
    // (the same result bit for bit)
    if (tile.variant == SP_KERNEL_SPLIT) tile = default_tile(p->ni, p->nj, p->nk, 1, 1);
    // out and A alternate so the last step lands in out; in place (in == out), the first
    // step must not write in, so an odd count ends in A and is copied to out
    View src = vin;
    for (int i = 0; i < steps; ++i) {
        const int left = steps - 1 - i;
        const View dst = aliased ? ((i % 2 == 0) ? A : vout) : ((left % 2 == 0) ? vout : A);
        const double t = tStart + i * dt;
        SP_TRY(launch_rk4_fused(s, fused_step(p, 1, dt, nullptr, computenu(p, t)), tile, SP_F64, src, SP_F64, dst));
        src = dst;
    }
    if (aliased && steps % 2 == 1) SP_TRY(launch_convert(s, p->ni, p->nj, p->nk, 1, SP_F64, A, SP_F64, vout));
    return SP_OK;
}

int sp_do_rk4_ex(sp_prop* p, const sp_layout* lin, const void* in, const sp_layout* lout, void* out, double tStart,
                 double dt, int steps)
{
    if (p && lin && lout && !(lin->dtype == SP_F64 && lout->dtype == SP_F64)) {
        // mixed precision always runs the fused kernel
        const int keep = p->path;
        p->path = SP_PATH_FUSED;
        const int r = rk4_common(p, lin, in, lout, out, tStart, dt, steps, false, "sp_do_rk4_ex");
        p->path = keep;
        return r;
    }
    return rk4_common(p, lin, in, lout, out, tStart, dt, steps, false, "sp_do_rk4_ex");
}

int sp_do_rk4_timestep(sp_prop* p, const sp_layout* lin, const void* in, const sp_layout* lout, void* out, double t,
                       double dt)
{
    return rk4_common(p, lin, in, lout, out, t, dt, 1, true, "sp_do_rk4_timestep");
}

int sp_do_rk4_batched(sp_prop* p, int nslices, const sp_layout* lin, const void* in, const sp_layout* lout, void* out,
                      long long slice_stride, double tStart, double dt, int steps)
{
    SP_CHECK(p, SP_ERR_INVALID, "sp_do_rk4_batched: null propagator");
    SP_TRY(check_prop_layout(p, lin, in, "sp_do_rk4_batched"));
    SP_TRY(check_prop_layout(p, lout, out, "sp_do_rk4_batched"));
    SP_CHECK(nslices >= 1, SP_ERR_INVALID, "sp_do_rk4_batched: nslices < 1");
    SP_CHECK(steps >= 0, SP_ERR_INVALID, "sp_do_rk4_batched: steps < 0");
    SP_CHECK(lin->dtype == lout->dtype || lin->dtype == SP_F64, SP_ERR_UNSUPPORTED,
             "sp_do_rk4_batched: dtype pair");
    SP_CHECK(slice_stride >= sp_layout_storage_elems(lin) && slice_stride >= sp_layout_storage_elems(lout),
             SP_ERR_INVALID, "sp_do_rk4_batched: slices overlap");
    {
        // A block reads the columns around its tile while other blocks write theirs: any
        // overlap of the two batches' storage, not only equal origins, races.
        const char* ib = (const char*)in;
        const char* ob = (const char*)out;
        const long long ie = ((long long)(nslices - 1) * slice_stride + sp_layout_storage_elems(lin)) *
                             dtype_size(lin->dtype);
        const long long oe = ((long long)(nslices - 1) * slice_stride + sp_layout_storage_elems(lout)) *
                             dtype_size(lout->dtype);
        SP_CHECK(ib + ie <= ob || ob + oe <= ib, SP_ERR_INVALID, "sp_do_rk4_batched: in and out overlap");
    }
    if (steps == 0) return SP_OK;
    SP_HIP(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    View vin = view_of(*lin, in), vout = view_of(*lout, out);
    vin.slice = slice_stride;
    vout.slice = slice_stride;
    // ping-pong between out and a batched scratch so the last step lands in out
    const size_t bytes = (size_t)slice_stride * nslices * dtype_size(lout->dtype);
    View A = vout;
    if (steps > 1) {
        SP_TRY(p->scratch[0].ensure(bytes));
        A.p = (char*)p->scratch[0].p + lout->origin * dtype_size(lout->dtype);
    }
    const FusedTile tile = pick_tile(p, nslices, lin->dtype == SP_F64 && lout->dtype == SP_F64);
    SP_TRY(ensure_split(p, tile, nslices));
    View cur = vin;
    int cdt = lin->dtype;
    for (int i = 0; i < steps; ++i) {
        const View d = ((steps - 1 - i) % 2 == 0) ? vout : A;
        SP_TRY(launch_rk4_fused(s, fused_step(p, nslices, dt, nullptr, computenu(p, tStart + i * dt)), tile, cdt, cur,
                                lout->dtype, d));
        cur = d;
        cdt = lout->dtype;
    }
    return SP_OK;
}

int sp_do_euler(sp_prop* p, const sp_layout* lin, const void* in, const sp_layout* lout, void* out, double tStart,
                double dt, int steps)
{
    SP_CHECK(p, SP_ERR_INVALID, "sp_do_euler: null propagator");
    SP_TRY(check_prop_layout(p, lin, in, "sp_do_euler"));
    SP_TRY(check_prop_layout(p, lout, out, "sp_do_euler"));
    SP_CHECK(lin->dtype == SP_F64 && lout->dtype == SP_F64, SP_ERR_INVALID, "sp_do_euler: fields must be SP_F64");
    SP_CHECK(steps >= 0, SP_ERR_INVALID, "sp_do_euler: steps < 0");
    if (steps == 0) return SP_OK;
    SP_HIP(hipSetDevice(p->ctx->device));
    const size_t bytes = (size_t)p->ni * p->nj * p->nk * 8;
    const bool aliased = same_origin(lin, in, lout, out);
    SP_TRY(p->scratch[0].ensure(bytes));
    if (aliased) SP_TRY(p->scratch[1].ensure(bytes));
    return euler_run(p, view_of(*lin, in), view_of(*lout, out), aliased, tStart, dt, steps, 1,
                     cview(p->scratch[0].p, p->ni, p->nj), cview(p->scratch[1].p, p->ni, p->nj));
}

int sp_do_euler_timestep(sp_prop* p, const sp_layout* lin, const void* in, const sp_layout* lout, void* out, double t,
                         double dt)
{
    return sp_do_euler(p, lin, in, lout, out, t, dt, 1);
}

int sp_periodic_fill(sp_ctx* c, const sp_layout* l, void* base)
{
    SP_CHECK(c && base, SP_ERR_INVALID, "sp_periodic_fill: null argument");
    SP_TRY(check_layout(l, "sp_periodic_fill"));
    SP_CHECK(l->dtype == SP_F64, SP_ERR_INVALID, "sp_periodic_fill: SP_F64 only");
    SP_CHECK(l->halo >= 1, SP_ERR_INVALID, "sp_periodic_fill: layout has no halo");
    SP_CHECK(l->ni >= l->halo && l->nj >= l->halo && l->nk >= l->halo, SP_ERR_INVALID,
             "sp_periodic_fill: halo wider than the domain");
    SP_HIP(hipSetDevice(c->device));
    return launch_periodic_fill(c->stream, *l, base);
}

int sp_parareal_update(sp_ctx* c, const sp_layout* l, void* q, const void* fine, const void* cnew, const void* cold)
{
    SP_CHECK(c && q && fine && cnew && cold, SP_ERR_INVALID, "sp_parareal_update: null argument");
    SP_TRY(check_layout(l, "sp_parareal_update"));
    SP_CHECK(l->dtype == SP_F64, SP_ERR_INVALID, "sp_parareal_update: SP_F64 only");
    SP_HIP(hipSetDevice(c->device));
    return launch_update(c->stream, l->ni, l->nj, l->nk, view_of(*l, q), view_of(*l, fine), view_of(*l, cnew),
                         view_of(*l, cold));
}

static int ensure_red(sp_ctx* c, int n)
{
    if (n <= c->red_cap) return SP_OK;
    if (c->red) (void)hipFree(c->red);
    if (c->red_host) (void)hipHostFree(c->red_host);
    c->red = nullptr;
    c->red_host = nullptr;
    c->red_cap = 0;
    if (hipMalloc(&c->red, n * sizeof(double)) != hipSuccess) return fail(SP_ERR_NOMEM, "reduction buffer");
    if (hipHostMalloc(&c->red_host, n * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return fail(SP_ERR_NOMEM, "pinned reduction buffer");
    c->red_cap = n;
    return SP_OK;
}

int sp_norms(sp_ctx* c, const sp_layout* la, const void* a, const sp_layout* lb, const void* b, double out[4])
{
    SP_CHECK(c && a && b && out, SP_ERR_INVALID, "sp_norms: null argument");
    SP_TRY(check_layout(la, "sp_norms a"));
    SP_TRY(check_layout(lb, "sp_norms b"));
    SP_CHECK(la->dtype == SP_F64 && lb->dtype == SP_F64, SP_ERR_INVALID, "sp_norms: SP_F64 only");
    SP_CHECK(la->ni == lb->ni && la->nj == lb->nj && la->nk == lb->nk, SP_ERR_INVALID, "sp_norms: size mismatch");
    SP_HIP(hipSetDevice(c->device));
    SP_TRY(ensure_red(c, kReduceBlocks * 4));
    SP_TRY(launch_norms(c->stream, la->ni, la->nj, la->nk, view_of(*la, a), view_of(*lb, b), c->red, kReduceBlocks));
    SP_HIP(hipMemcpyAsync(c->red_host, c->red, kReduceBlocks * 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SP_HIP(hipStreamSynchronize(c->stream));
    double r[4] = {0, 0, 0, 0};
    for (int bI = 0; bI < kReduceBlocks; ++bI) {
        r[0] = std::max(r[0], c->red_host[bI * 4 + 0]);
        r[1] = std::max(r[1], c->red_host[bI * 4 + 1]);
        r[2] += c->red_host[bI * 4 + 2];
        r[3] += c->red_host[bI * 4 + 3];
    }
    for (int q = 0; q < 4; ++q) out[q] = r[q];
    return SP_OK;
}

int sp_fill_exact(sp_ctx* c, const sp_layout* l, void* base, double nu0, double nufreq, double cx, double cy,
                  double cz, double t)
{
    SP_CHECK(c && base, SP_ERR_INVALID, "sp_fill_exact: null argument");
    SP_TRY(check_layout(l, "sp_fill_exact"));
    SP_CHECK(l->dtype == SP_F64, SP_ERR_INVALID, "sp_fill_exact: SP_F64 only");
    SP_HIP(hipSetDevice(c->device));
    return launch_fill_exact(c->stream, *l, base, nu0, nufreq, cx, cy, cz, t);
}

int sp_fill_gaussians(sp_ctx* c, const sp_layout* l, void* base, int ng, const double* xc, const double* yc,
                      const double* zc, const double* sigma, const double* amp)
{
    SP_CHECK(c && base && ng > 0 && xc && yc && zc && sigma && amp, SP_ERR_INVALID, "sp_fill_gaussians: bad argument");
    SP_TRY(check_layout(l, "sp_fill_gaussians"));
    SP_CHECK(l->dtype == SP_F64, SP_ERR_INVALID, "sp_fill_gaussians: SP_F64 only");
    SP_HIP(hipSetDevice(c->device));
    std::vector<double> g(5 * (size_t)ng);
    for (int q = 0; q < ng; ++q) {
        g[q] = xc[q];
        g[ng + q] = yc[q];
        g[2 * ng + q] = zc[q];
        g[3 * ng + q] = sigma[q];
        g[4 * ng + q] = amp[q];
    }
    double* gd = nullptr;
    SP_HIP(hipMalloc(&gd, g.size() * sizeof(double)));
    int r = SP_OK;
    if (hipMemcpyAsync(gd, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        r = fail(SP_ERR_HIP, "sp_fill_gaussians: copy failed");
    if (r == SP_OK) r = launch_fill_gaussians(c->stream, *l, base, ng, gd);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(gd);
    return r;
}

}  // extern "C"

// ======================================================================================
// Batched Parareal.
static int ensure_red(sp_ctx* c, int n);
struct sp_parareal {
    sp_prop* p = nullptr;
    sp_parareal_cfg cfg{};
    int S = 0, tsF = 0, tsG = 0;  // per slice (tsG counts sub-steps)
    double dtF = 0, dtG = 0;      // dtG is the sub-step
    double T = 0;
    int gdt = SP_F64;
    size_t n = 0;                 // points per slice
    Buf U, F, FS, Gold, Gw0, Gw1, Q, coefF, coefG, part;
    Buf D;  // difference form: U_n^{k+1} - U_n^k of the chain's current slice
    Buf W;  // work field of split-tile steps (its own: captured graphs keep its address)
    SpectralPlan* spec = nullptr;  // spectral coarse kinds: FFT plans and coefficient buffer
    Buf nu_spec;  // ... and its per-sub-step nu: [0, S*tsG) from t = 0, then [i*S + r] per slice
    std::vector<double> nuF_host, nuG_host;
    bool nu_const = true;
    bool begun = false;
    int iters_done = 0;
    double u0_norm2 = 0.;  // sum of u0^2, for the blow-up check
    FusedTile tileF{}, tileG1{}, tileGS{};
    FusedTile tileG1_64{};  // one-slice FP64 tile of an FP64 initial coarse run
    // With use_graphs an iteration is two graphs, the fine sweep and the coarse chain (its
    // coarse runs and corrections), launched back to back with events between them.
    hipGraphExec_t gexecF = nullptr, gexecC = nullptr;
    hipGraph_t graphF = nullptr, graphC = nullptr;
    std::vector<double> part_host;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};  // before the sweep, after it, after the chain
    // without graphs: events around every correction launch of the chain (2 per slice)
    std::vector<hipEvent_t> evU;
    double fine_ms = 0., chain_ms = 0., update_ms = 0.;
};

namespace {

double slice_start(const sp_parareal* pr, int r) { return (pr->T / pr->S) * r; }  // main.cpp:157

View batch_view(void* p, const sp_parareal* pr, int dt, int slice0 = 0)
{
    View v = cview((char*)p + (size_t)slice0 * pr->n * dtype_size(dt), pr->p->ni, pr->p->nj);
    v.slice = (long long)pr->n;
    return v;
}

// A fused step of this Parareal's propagator with the Parareal's own split work field.
FusedStep pr_step(const sp_parareal* pr, int nslices, double dt, const void* nud, double nu)
{
    FusedStep st = fused_step(pr->p, nslices, dt, nud, nu);
    st.work = pr->W.p;
    st.work_bytes = pr->W.cap;
    return st;
}

// The coarse kinds that apply each segment as its Fourier multiplier (sp_spectral.hip).
bool spectral_coarse(int kind) { return kind == SP_COARSE_RK4_SPECTRAL || kind == SP_COARSE_EULER_SPECTRAL; }

// Coarse propagation of `nslices` batched slices, tsG sub-steps from per-slice start
// times; step i of slice r runs at t0[r] + i*dtG (or at (first + i)*dtG from t = 0 for
// the initial chain, first_index >= 0).  Input f64 (U) or gdt; output gdt in `dst`.
int coarse_run(sp_parareal* pr, int nslices, int slice0, int idt, View src, View dst, View scr, const void* coef_base,
               long long first_index, int steps, int odt = -1)
{
    sp_prop* p = pr->p;
    hipStream_t s = p->ctx->stream;
    const int gdt = odt < 0 ? pr->gdt : odt;  // the chain's storage type (FP64 for an FP64 initial run)
    if (spectral_coarse(pr->cfg.coarse_kind)) {
        // each slice's segment as one Fourier multiplier (compact FP64 fields)
        SP_CHECK(idt == SP_F64 && gdt == SP_F64 && src.sj == p->ni && dst.sj == p->ni &&
                     src.sk == (long long)p->ni * p->nj && dst.sk == (long long)p->ni * p->nj,
                 SP_ERR_INVALID, "spectral coarse propagator: compact FP64 fields only");
        const long long per = (long long)pr->S * pr->tsG;
        for (int q = 0; q < nslices; ++q) {
            SpectralStep st{p->dx, p->cx, p->cy, p->cz, pr->dtG, steps, p->nu0, nullptr, 1};
            if (pr->cfg.coarse_kind == SP_COARSE_EULER_SPECTRAL)
                st.kind = p->upwind ? kSpectralEulerUpwind : kSpectralEulerCentral;
            if (!pr->nu_const) {
                const double* tab = (const double*)pr->nu_spec.p;
                if (first_index >= 0) {
                    st.nu_tab = tab + first_index;
                } else {
                    st.nu_tab = tab + per + slice0 + q;
                    st.nu_stride = pr->S;
                }
            }
            SP_TRY(spectral_apply(pr->spec, s, (const double*)src.p + (size_t)q * src.slice,
                                  (double*)dst.p + (size_t)q * dst.slice, st));
        }
        return SP_OK;
    }
    if (p->path == SP_PATH_STAGED && pr->cfg.coarse_kind == SP_COARSE_RK4 && gdt == SP_F64 && idt == SP_F64) {
        // reference-order sweeps, slice by slice (src -> dst, then dst -> dst as DoRK4 does)
        for (int q = 0; q < nslices; ++q) {
            View a = src, b = dst;
            a.p = (char*)src.p + (size_t)q * src.slice * 8;
            b.p = (char*)dst.p + (size_t)q * dst.slice * 8;
            a.slice = b.slice = 0;
            for (int i = 0; i < steps; ++i) {
                const double t = first_index >= 0 ? 0. + (double)(first_index + i) * pr->dtG
                                                  : slice_start(pr, slice0 + q) + i * pr->dtG;
                SP_TRY(rk4_staged_step(p, i == 0 ? a : b, b, t, pr->dtG));
            }
        }
        return SP_OK;
    }
    View cur = src;
    int cdt = idt;
    for (int i = 0; i < steps; ++i) {
        const int left = steps - 1 - i;
        const View d = (left % 2 == 0) ? dst : scr;
        double nu;
        const void* nud = nullptr;
        if (first_index >= 0) {
            nu = computenu(p, 0. + (double)(first_index + i) * pr->dtG);
        } else if (pr->nu_const) {
            nu = p->nu0;
        } else if (nslices == 1) {
            nu = pr->nuG_host[(size_t)i * pr->S + slice0];
        } else {
            nu = p->nu0;
            nud = (const char*)coef_base + ((size_t)i * pr->S + slice0) * coefs_size();
        }
        if (pr->cfg.coarse_kind == SP_COARSE_RK4) {
            const FusedTile& tile = gdt != pr->gdt ? pr->tileG1_64 : nslices == 1 ? pr->tileG1 : pr->tileGS;
            SP_TRY(launch_rk4_fused(s, pr_step(pr, nslices, pr->dtG, nud, nu), tile, cdt, cur, gdt, d));
        } else {
            if (nud) return fail(SP_ERR_UNSUPPORTED, "batched Euler coarse with time-dependent nu");
            const StagedParams P{p->ni, p->nj, p->nk, p->cx, p->cy, p->cz, p->dx, p->dx * p->dx};
            SP_TRY(launch_euler_fused(s, P, cur, d, nu, pr->dtG, p->upwind, nslices));
        }
        cur = d;
        cdt = gdt;
    }
    return SP_OK;
}

double fine_time(const sp_parareal* pr, int r, int i)
{
    const double t0 = slice_start(pr, r);
    // Parareal.h:119-128: the async schedule runs step 0 at tStart, then from tStart + dtF
    if (pr->cfg.async_schedule && r > 0 && i > 0) return (t0 + pr->dtF) + (i - 1) * pr->dtF;
    return t0 + i * pr->dtF;
}

int fine_sweep(sp_parareal* pr)
{
    sp_prop* p = pr->p;
    hipStream_t s = p->ctx->stream;
    if (p->path == SP_PATH_STAGED) {
        // reference-order sweeps, slice by slice: DoRK4(U_r -> F_r, tStart_r, dtF, tsF)
        for (int r = 0; r < pr->S; ++r) {
            const View Ur = cview((char*)pr->U.p + (size_t)r * pr->n * 8, p->ni, p->nj);
            const View Fr = cview((char*)pr->F.p + (size_t)r * pr->n * 8, p->ni, p->nj);
            for (int i = 0; i < pr->tsF; ++i) SP_TRY(rk4_staged_step(p, i == 0 ? Ur : Fr, Fr, fine_time(pr, r, i), pr->dtF));
        }
        return SP_OK;
    }
    const View U = batch_view(pr->U.p, pr, SP_F64), F = batch_view(pr->F.p, pr, SP_F64),
               FS = batch_view(pr->FS.p, pr, SP_F64);
    View cur = U;
    for (int i = 0; i < pr->tsF; ++i) {
        const int left = pr->tsF - 1 - i;
        const View d = (left % 2 == 0) ? F : FS;
        const void* nud = pr->nu_const ? nullptr : (const char*)pr->coefF.p + (size_t)i * pr->S * coefs_size();
        SP_TRY(launch_rk4_fused(s, pr_step(pr, pr->S, pr->dtF, nud, p->nu0), pr->tileF, SP_F64, cur, SP_F64, d));
        cur = d;
    }
    return SP_OK;
}

// Events around the correction of slice r (timed = true only outside graph capture).
int mark_update(sp_parareal* pr, int r, int which, bool timed)
{
    if (timed) SP_HIP(hipEventRecord(pr->evU[2 * r + which], pr->p->ctx->stream));
    return SP_OK;
}

int coarse_chain(sp_parareal* pr, bool timed)
{
    sp_prop* p = pr->p;
    hipStream_t s = p->ctx->stream;
    const int gs = dtype_size(pr->gdt);
    for (int r = 0; r < pr->S; ++r) {
        View cold = cview((char*)pr->Gold.p + (size_t)r * pr->n * gs, p->ni, p->nj);
        View cnew;
        if (r == 0) {
            // Rank 0 never receives: G(U_0) is recomputed from the same input and times, so it
            // equals its initial coarse value bit for bit.
            cnew = cold;
        } else {
            const View Ur = cview((char*)pr->U.p + (size_t)r * pr->n * 8, p->ni, p->nj);
            const View A = cview(pr->Gw0.p, p->ni, p->nj), B = cview(pr->Gw1.p, p->ni, p->nj);
            SP_TRY(coarse_run(pr, 1, r, SP_F64, Ur, A, B, pr->coefG.p, -1, pr->tsG));
            cnew = A;
        }
        const View fine = cview((char*)pr->F.p + (size_t)r * pr->n * 8, p->ni, p->nj);
        const View dst = (r < pr->S - 1) ? cview((char*)pr->U.p + (size_t)(r + 1) * pr->n * 8, p->ni, p->nj)
                                         : cview(pr->Q.p, p->ni, p->nj);
        SP_TRY(mark_update(pr, r, 0, timed));
        SP_TRY(launch_correction(s, p->ni, p->nj, p->nk, pr->gdt, cnew, cold, fine, dst,
                                 (double*)pr->part.p + (size_t)r * kReduceBlocks * 2, kReduceBlocks));
        SP_TRY(mark_update(pr, r, 1, timed));
    }
    return SP_OK;
}

// The chain in difference form (delta_correction): U_{r+1} = F(U_r^k) + G(U_r^{k+1} - U_r^k),
// equal to the reference's G(U_r^{k+1}) + F(U_r^k) - G(U_r^k) for the linear G.  D carries
// the difference from one slice to the next; slice 0's is zero (its start state is u0).
int coarse_chain_delta(sp_parareal* pr, bool timed)
{
    sp_prop* p = pr->p;
    hipStream_t s = p->ctx->stream;
    const View D = cview(pr->D.p, p->ni, p->nj);
    for (int r = 0; r < pr->S; ++r) {
        View gd = cview(nullptr, p->ni, p->nj);
        if (r > 0) {
            const View A = cview(pr->Gw0.p, p->ni, p->nj), B = cview(pr->Gw1.p, p->ni, p->nj);
            SP_TRY(coarse_run(pr, 1, r, SP_F64, D, A, B, pr->coefG.p, -1, pr->tsG));
            gd = A;
        }
        const View fine = cview((char*)pr->F.p + (size_t)r * pr->n * 8, p->ni, p->nj);
        const bool last = r == pr->S - 1;
        const View dst = last ? cview(pr->Q.p, p->ni, p->nj)
                              : cview((char*)pr->U.p + (size_t)(r + 1) * pr->n * 8, p->ni, p->nj);
        SP_TRY(mark_update(pr, r, 0, timed));
        SP_TRY(launch_correction_delta(s, p->ni, p->nj, p->nk, pr->gdt, gd, fine, dst,
                                       last ? cview(nullptr, p->ni, p->nj) : D,
                                       (double*)pr->part.p + (size_t)r * kReduceBlocks * 2, kReduceBlocks));
        SP_TRY(mark_update(pr, r, 1, timed));
    }
    return SP_OK;
}

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

int sp_parareal_create(sp_prop* p, const sp_parareal_cfg* cfg, sp_parareal** out)
{
    SP_CHECK(p && cfg && out, SP_ERR_INVALID, "sp_parareal_create: null argument");
    SP_CHECK(cfg->slices >= 1, SP_ERR_INVALID, "sp_parareal_create: slices < 1");
    SP_CHECK(cfg->kmax >= 0 && cfg->kmax <= SP_MAX_ITERS, SP_ERR_INVALID, "sp_parareal_create: kmax out of range");
    SP_CHECK(cfg->end_time > 0 && std::isfinite(cfg->end_time), SP_ERR_INVALID, "sp_parareal_create: end_time");
    SP_CHECK(cfg->steps_fine >= 1 && cfg->steps_coarse >= 1, SP_ERR_INVALID, "sp_parareal_create: step counts");
    SP_CHECK(cfg->coarse_kind == SP_COARSE_EULER || cfg->coarse_kind == SP_COARSE_RK4 ||
                 spectral_coarse(cfg->coarse_kind),
             SP_ERR_INVALID, "sp_parareal_create: coarse_kind");
    SP_CHECK(!spectral_coarse(cfg->coarse_kind) || cfg->coarse_dtype == SP_F64, SP_ERR_UNSUPPORTED,
             "sp_parareal_create: the spectral coarse propagator runs in FP64");
    SP_CHECK(cfg->coarse_dtype == SP_F64 || cfg->coarse_dtype == SP_F32 || cfg->coarse_dtype == SP_BF16,
             SP_ERR_INVALID, "sp_parareal_create: coarse_dtype");
    SP_CHECK(cfg->coarse_kind == SP_COARSE_RK4 || cfg->coarse_dtype == SP_F64, SP_ERR_UNSUPPORTED,
             "sp_parareal_create: Euler coarse propagator is SP_F64 only");
    SP_CHECK(cfg->coarse_substeps >= 1, SP_ERR_INVALID, "sp_parareal_create: coarse_substeps < 1");
    SP_CHECK(cfg->tol >= 0, SP_ERR_INVALID, "sp_parareal_create: tol < 0");
    SP_CHECK(cfg->initial_fp64 == 0 || cfg->initial_fp64 == 1, SP_ERR_INVALID, "sp_parareal_create: initial_fp64");
    SP_CHECK(!cfg->initial_fp64 || cfg->delta_correction, SP_ERR_UNSUPPORTED,
             "sp_parareal_create: initial_fp64 needs the difference-form correction (the reference's form "
             "subtracts coarse values that must come from one coarse propagator)");
    SP_CHECK(cfg->delta_correction == 0 || cfg->delta_correction == 1, SP_ERR_INVALID,
             "sp_parareal_create: delta_correction must be 0 or 1");
    SP_HIP(hipSetDevice(p->ctx->device));
    sp_parareal* pr = new sp_parareal;
    pr->p = p;
    p->refs++;
    pr->cfg = *cfg;
    pr->S = cfg->slices;
    pr->T = cfg->end_time;
    // RuntimeConfiguration.cpp:133-151: totals rounded up to multiples of the slice count
    int sf = cfg->steps_fine, sc = cfg->steps_coarse;
    if (sf % pr->S != 0) sf = (sf / pr->S + 1) * pr->S;
    if (sc % pr->S != 0) sc = (sc / pr->S + 1) * pr->S;
    pr->dtF = pr->T / sf;  // RuntimeConfiguration.h:62-73
    pr->tsF = sf / pr->S;
    pr->dtG = (pr->T / sc) / cfg->coarse_substeps;
    pr->tsG = (sc / pr->S) * cfg->coarse_substeps;
    pr->gdt = cfg->coarse_dtype;
    pr->n = (size_t)p->ni * p->nj * p->nk;
    pr->nu_const = (p->nufreq == 0.);
    const size_t nb = pr->n * pr->S;
    int r = SP_OK;
    auto chk = [&](int x) {
        if (r == SP_OK && x != SP_OK) r = x;
    };
    chk(pr->U.ensure(nb * 8));
    chk(pr->F.ensure(nb * 8));
    chk(pr->FS.ensure(nb * 8));
    // the difference form keeps no coarse values: one slice of Gold serves as scratch
    chk(pr->Gold.ensure((cfg->delta_correction ? pr->n : nb) * dtype_size(pr->gdt)));
    if (cfg->delta_correction) chk(pr->D.ensure(pr->n * 8));
    chk(pr->Gw0.ensure(pr->n * dtype_size(pr->gdt)));
    chk(pr->Gw1.ensure(pr->n * dtype_size(pr->gdt)));
    chk(pr->Q.ensure(pr->n * 8));
    chk(pr->part.ensure((size_t)pr->S * kReduceBlocks * 2 * sizeof(double)));
    if (r != SP_OK) {
        sp_parareal_destroy(pr);
        return r;
    }
    // Per-slice nu tables (computenu at every step start, Convection.cpp:321).
    if (!pr->nu_const) {
        pr->nuF_host.resize((size_t)pr->tsF * pr->S);
        pr->nuG_host.resize((size_t)pr->tsG * pr->S);
        for (int rr = 0; rr < pr->S; ++rr) {
            const double t0 = slice_start(pr, rr);
            for (int i = 0; i < pr->tsF; ++i) pr->nuF_host[(size_t)i * pr->S + rr] = computenu(p, fine_time(pr, rr, i));
            for (int i = 0; i < pr->tsG; ++i) pr->nuG_host[(size_t)i * pr->S + rr] = computenu(p, t0 + i * pr->dtG);
        }
        // per-step, per-slice coefficient records of the fused kernel
        const size_t cs = coefs_size();
        std::vector<char> cf(pr->nuF_host.size() * cs), cg(pr->nuG_host.size() * cs);
        for (size_t i = 0; i < pr->nuF_host.size(); ++i)
            fill_coefs(cf.data() + i * cs, pr->nuF_host[i], p->cx, p->cy, p->cz, p->dx, pr->dtF);
        for (size_t i = 0; i < pr->nuG_host.size(); ++i)
            fill_coefs(cg.data() + i * cs, pr->nuG_host[i], p->cx, p->cy, p->cz, p->dx, pr->dtG);
        chk(pr->coefF.ensure(cf.size()));
        chk(pr->coefG.ensure(cg.size()));
        if (r == SP_OK && (hipMemcpy(pr->coefF.p, cf.data(), cf.size(), hipMemcpyHostToDevice) != hipSuccess ||
                           hipMemcpy(pr->coefG.p, cg.data(), cg.size(), hipMemcpyHostToDevice) != hipSuccess))
            r = fail(SP_ERR_HIP, "sp_parareal_create: coefficient table upload");
        if (r != SP_OK) {
            sp_parareal_destroy(pr);
            return r;
        }
    }
    if (spectral_coarse(cfg->coarse_kind)) {
        r = spectral_create(&pr->spec, p->ni, p->nj, p->nk);
        if (r == SP_OK && !pr->nu_const) {
            // nu of every coarse sub-step: from t = 0 (the initial run), then per slice start
            const size_t per = (size_t)pr->S * pr->tsG;
            std::vector<double> tab(2 * per);
            for (size_t j = 0; j < per; ++j) tab[j] = computenu(p, 0. + (double)j * pr->dtG);
            std::copy(pr->nuG_host.begin(), pr->nuG_host.end(), tab.begin() + per);
            chk(pr->nu_spec.ensure(tab.size() * sizeof(double)));
            if (r == SP_OK && hipMemcpy(pr->nu_spec.p, tab.data(), tab.size() * sizeof(double),
                                        hipMemcpyHostToDevice) != hipSuccess)
                r = fail(SP_ERR_HIP, "sp_parareal_create: nu table upload");
        }
        if (r != SP_OK) {
            sp_parareal_destroy(pr);
            return r;
        }
    }
    if (p->path == SP_PATH_STAGED) {
        // staged temporaries exist before any graph capture
        const size_t bytes = pr->n * 8;
        for (Buf* b : {&p->qI, &p->k1, &p->k2, &p->k3, &p->k4}) chk(b->ensure(bytes));
        if (r != SP_OK) {
            sp_parareal_destroy(pr);
            return r;
        }
    }
    if (cfg->autotune && p->path == SP_PATH_FUSED && p->nk >= kFusedMinK && p->tile.tx == 0) {
        // tune on the Parareal's own (cleared) slice arrays, before any capture
        hipStream_t s = p->ctx->stream;
        if (hipMemsetAsync(pr->U.p, 0, pr->U.cap, s) != hipSuccess ||
            hipMemsetAsync(pr->F.p, 0, pr->F.cap, s) != hipSuccess ||
            hipMemsetAsync(pr->Gw0.p, 0, pr->Gw0.cap, s) != hipSuccess)
            r = fail(SP_ERR_HIP, "sp_parareal_create: clear");
        View u = cview(pr->U.p, p->ni, p->nj), f = cview(pr->FS.p, p->ni, p->nj);
        u.slice = f.slice = (long long)pr->n;
        if (r == SP_OK) r = tune_fused(p, pr->S, SP_F64, SP_F64, u, f, nullptr);
        // DoSerial steps one slice: its FP64 tile too
        if (r == SP_OK && pr->S > 1)
            r = tune_fused(p, 1, SP_F64, SP_F64, cview(pr->U.p, p->ni, p->nj), cview(pr->FS.p, p->ni, p->nj), nullptr);
        if (r == SP_OK && cfg->coarse_kind == SP_COARSE_RK4) {
            View gi = cview(pr->F.p, p->ni, p->nj), go = cview(pr->Gold.p, p->ni, p->nj);
            gi.slice = go.slice = (long long)pr->n;
            if (pr->S > 1 && !cfg->delta_correction) r = tune_fused(p, pr->S, pr->gdt, pr->gdt, gi, go, nullptr);
            if (r == SP_OK)
                r = tune_fused(p, 1, pr->gdt, pr->gdt, cview(pr->Gw0.p, p->ni, p->nj), cview(pr->Gw1.p, p->ni, p->nj),
                               nullptr);
        }
        if (r != SP_OK) {
            sp_parareal_destroy(pr);
            return r;
        }
    }
    pr->tileF = pick_tile(p, pr->S, 1);
    pr->tileG1 = pick_tile(p, 1, pr->gdt == SP_F64);
    pr->tileGS = pick_tile(p, pr->S, pr->gdt == SP_F64);
    pr->tileG1_64 = pick_tile(p, 1, 1);
    {
        size_t wb = 0;
        if (pr->tileG1_64.variant == SP_KERNEL_SPLIT) wb = split_work_bytes(p->ni, p->nj, p->nk, 1);
        if (pr->tileF.variant == SP_KERNEL_SPLIT) wb = std::max(wb, split_work_bytes(p->ni, p->nj, p->nk, pr->S));
        if (pr->gdt == SP_F64 && pr->tileG1.variant == SP_KERNEL_SPLIT)
            wb = std::max(wb, split_work_bytes(p->ni, p->nj, p->nk, 1));
        if (pr->gdt == SP_F64 && pr->tileGS.variant == SP_KERNEL_SPLIT)
            wb = std::max(wb, split_work_bytes(p->ni, p->nj, p->nk, pr->S));
        const int r = wb ? pr->W.ensure(wb) : SP_OK;
        if (r != SP_OK) {
            sp_parareal_destroy(pr);
            return r;
        }
    }
    pr->part_host.resize((size_t)pr->S * kReduceBlocks * 2);
    *out = pr;
    return SP_OK;
}

int sp_parareal_destroy(sp_parareal* pr)
{
    if (!pr) return SP_OK;
    (void)hipStreamSynchronize(pr->p->ctx->stream);
    for (hipGraphExec_t g : {pr->gexecF, pr->gexecC})
        if (g) (void)hipGraphExecDestroy(g);
    for (hipGraph_t g : {pr->graphF, pr->graphC})
        if (g) (void)hipGraphDestroy(g);
    for (hipEvent_t e : pr->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : pr->evU)
        if (e) (void)hipEventDestroy(e);
    for (Buf* b : {&pr->U, &pr->F, &pr->FS, &pr->Gold, &pr->Gw0, &pr->Gw1, &pr->Q, &pr->coefF, &pr->coefG, &pr->part,
                   &pr->D, &pr->W, &pr->nu_spec})
        b->release();
    spectral_destroy(pr->spec);
    sp_prop* p = pr->p;
    delete pr;
    prop_release(p);
    return SP_OK;
}

int sp_parareal_begin(sp_parareal* pr, const sp_layout* l0, const void* u0)
{
    SP_CHECK(pr, SP_ERR_INVALID, "sp_parareal_begin: null handle");
    sp_prop* p = pr->p;
    SP_TRY(check_prop_layout(p, l0, u0, "sp_parareal_begin"));
    SP_CHECK(l0->dtype == SP_F64, SP_ERR_INVALID, "sp_parareal_begin: u0 must be SP_F64");
    SP_HIP(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    const int gs = dtype_size(pr->gdt);
    // U_0 = u0
    SP_TRY(launch_convert(s, p->ni, p->nj, p->nk, 1, SP_F64, view_of(*l0, u0), SP_F64, cview(pr->U.p, p->ni, p->nj)));
    // U_r = G^{r*tsG}(u0), one continuous coarse run from t = 0 (Parareal.h:113): rank r's own
    // run visits the same states at the same times (step j at t = 0 + j*dtG), so its value is
    // the snapshot after r*tsG steps.  The chain state stays in coarse precision and
    // alternates between Gw0 and Gw1; slice 0 of FS is the sub-step scratch.
    // Gold_r = G(U_r) (Parareal.h:114) is the chain's next snapshot whenever nu is constant:
    // the same coarse steps on the same values (the kernel variants agree bitwise), only the
    // times differ, and they enter only through nu.  The chain then writes Gold itself and
    // runs one more segment for Gold_{S-1}; otherwise Gold is one batched run below.
    const bool gold_from_chain = !pr->cfg.delta_correction && pr->nu_const;
    // FP64 initial run (initial_fp64): each segment from U_{r-1} straight into U_r; slice 0 of
    // FS is the sub-step scratch.  (The difference form needs no coarse values.)
    const bool init64 = pr->cfg.initial_fp64 && pr->cfg.delta_correction && pr->gdt != SP_F64;
    if (init64) {
        const View Z = cview(pr->FS.p, p->ni, p->nj);
        for (int r = 1; r < pr->S; ++r)
            SP_TRY(coarse_run(pr, 1, r - 1, SP_F64, cview((char*)pr->U.p + (size_t)(r - 1) * pr->n * 8, p->ni, p->nj),
                              cview((char*)pr->U.p + (size_t)r * pr->n * 8, p->ni, p->nj), Z, nullptr,
                              (long long)(r - 1) * pr->tsG, pr->tsG, SP_F64));
    } else {
        const View bufs[2] = {cview(pr->Gw0.p, p->ni, p->nj), cview(pr->Gw1.p, p->ni, p->nj)};
        const View Z = cview(pr->FS.p, p->ni, p->nj);
        const auto gold = [&](int r) { return cview((char*)pr->Gold.p + (size_t)r * pr->n * gs, p->ni, p->nj); };
        int cur = -1;
        for (int r = 1; r < pr->S; ++r) {
            const int nxt = cur < 0 ? 0 : 1 - cur;
            const View src = cur < 0 ? cview(pr->U.p, p->ni, p->nj) : bufs[cur];
            SP_TRY(coarse_run(pr, 1, r - 1, cur < 0 ? SP_F64 : pr->gdt, src, bufs[nxt], Z, nullptr,
                              (long long)(r - 1) * pr->tsG, pr->tsG));
            cur = nxt;
            SP_TRY(launch_convert(s, p->ni, p->nj, p->nk, 1, pr->gdt, bufs[cur], SP_F64,
                                  cview((char*)pr->U.p + (size_t)r * pr->n * 8, p->ni, p->nj)));
            if (gold_from_chain) SP_TRY(launch_convert(s, p->ni, p->nj, p->nk, 1, pr->gdt, bufs[cur], pr->gdt, gold(r - 1)));
        }
        if (gold_from_chain) {
            const View src = cur < 0 ? cview(pr->U.p, p->ni, p->nj) : bufs[cur];
            SP_TRY(coarse_run(pr, 1, pr->S - 1, cur < 0 ? SP_F64 : pr->gdt, src, gold(pr->S - 1), Z, nullptr,
                              (long long)(pr->S - 1) * pr->tsG, pr->tsG));
        }
    }
    // Otherwise Gold_r = G(U_r) from tStart_r, all slices in one batch; FS is scratch.
    // (The difference form needs no coarse values.)
    if (!pr->cfg.delta_correction && !gold_from_chain) {
        View src = batch_view(pr->U.p, pr, SP_F64);
        View dst = batch_view(pr->Gold.p, pr, pr->gdt);
        View scr = batch_view(pr->FS.p, pr, pr->gdt);
        if (pr->cfg.coarse_kind == SP_COARSE_EULER && !pr->nu_const) {
            for (int r = 0; r < pr->S; ++r) {
                View sr = cview((char*)pr->U.p + (size_t)r * pr->n * 8, p->ni, p->nj);
                View dr = cview((char*)pr->Gold.p + (size_t)r * pr->n * gs, p->ni, p->nj);
                SP_TRY(coarse_run(pr, 1, r, SP_F64, sr, dr, cview(pr->Gw0.p, p->ni, p->nj), nullptr, -1, pr->tsG));
            }
        } else {
            SP_TRY(coarse_run(pr, pr->S, 0, SP_F64, src, dst, scr, pr->coefG.p, -1, pr->tsG));
        }
    }
    // ||u0||^2: advection-diffusion never increases the L2 norm, so slice states far above
    // it mean an unstable propagator.
    {
        const View u = cview(pr->U.p, p->ni, p->nj);
        SP_TRY(ensure_red(p->ctx, kReduceBlocks * 4));
        SP_TRY(launch_norms(s, p->ni, p->nj, p->nk, u, u, p->ctx->red, kReduceBlocks));
        SP_HIP(hipMemcpyAsync(p->ctx->red_host, p->ctx->red, kReduceBlocks * 4 * sizeof(double),
                              hipMemcpyDeviceToHost, s));
        SP_HIP(hipStreamSynchronize(s));
        double n2 = 0.;
        for (int bI = 0; bI < kReduceBlocks; ++bI) n2 += p->ctx->red_host[bI * 4 + 3];
        pr->u0_norm2 = n2;
    }
    // The coarse prediction of the final state is the last slice's coarse value.
    if (init64) {
        SP_TRY(coarse_run(pr, 1, pr->S - 1, SP_F64, cview((char*)pr->U.p + (size_t)(pr->S - 1) * pr->n * 8, p->ni, p->nj),
                          cview(pr->Q.p, p->ni, p->nj), cview(pr->FS.p, p->ni, p->nj), pr->coefG.p, -1, pr->tsG,
                          SP_F64));
    } else if (pr->cfg.delta_correction) {
        const View A = cview(pr->Gw0.p, p->ni, p->nj), B = cview(pr->Gw1.p, p->ni, p->nj);
        SP_TRY(coarse_run(pr, 1, pr->S - 1, SP_F64, cview((char*)pr->U.p + (size_t)(pr->S - 1) * pr->n * 8, p->ni, p->nj),
                          A, B, pr->coefG.p, -1, pr->tsG));
        SP_TRY(launch_convert(s, p->ni, p->nj, p->nk, 1, pr->gdt, A, SP_F64, cview(pr->Q.p, p->ni, p->nj)));
    } else {
        SP_TRY(launch_convert(s, p->ni, p->nj, p->nk, 1, pr->gdt,
                              cview((char*)pr->Gold.p + (size_t)(pr->S - 1) * pr->n * gs, p->ni, p->nj), SP_F64,
                              cview(pr->Q.p, p->ni, p->nj)));
    }
    SP_HIP(hipStreamSynchronize(s));
    pr->begun = true;
    pr->iters_done = 0;
    return SP_OK;
}

// The context's stream set to s for a scope (work launched beside a captured graph).
struct StreamScope {
    sp_ctx* c;
    hipStream_t keep;
    StreamScope(sp_ctx* c_, hipStream_t s) : c(c_), keep(c_->stream) { c->stream = s; }
    ~StreamScope() { c->stream = keep; }
};

// Captures `body` (enqueued on stream s) as a graph and instantiates it.
static int capture(sp_parareal* pr, hipStream_t s, hipGraph_t* graph, hipGraphExec_t* exec,
                   const std::function<int()>& body)
{
    sp_prop* p = pr->p;
    const hipStream_t user = p->ctx->stream;
    // begin the capture first: on failure the context still follows the caller's stream
    SP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    p->ctx->stream = s;
    const int r = body();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    p->ctx->stream = user;
    if (r != SP_OK) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    SP_HIP(e);
    *graph = g;
    SP_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return SP_OK;
}

static int iterate_impl(sp_parareal* pr, double* incr, double* tfine, double* tcoarse)
{
    sp_prop* p = pr->p;
    hipStream_t s = p->ctx->stream;
    for (hipEvent_t& e : pr->ev)
        if (!e) SP_HIP(hipEventCreate(&e));
    const bool graphs = pr->cfg.use_graphs != 0;
    if (graphs) {
        // Capture needs a created stream: the null stream is first drained, then the graphs
        // run on the context's own stream.
        if (s == nullptr) {
            SP_HIP(hipStreamSynchronize(nullptr));
            s = p->ctx->own;
        }
        // The spectral coarse chain (S FFT pairs) is launched, not captured: its FFT library
        // calls stay out of the graph.
        const bool chain_graph = !spectral_coarse(pr->cfg.coarse_kind);
        if (!pr->gexecF) {
            SP_TRY(capture(pr, s, &pr->graphF, &pr->gexecF, [&] { return fine_sweep(pr); }));
            if (chain_graph)
                SP_TRY(capture(pr, s, &pr->graphC, &pr->gexecC, [&] {
                    return pr->cfg.delta_correction ? coarse_chain_delta(pr, false) : coarse_chain(pr, false);
                }));
        }
        SP_HIP(hipEventRecord(pr->ev[0], s));
        SP_HIP(hipGraphLaunch(pr->gexecF, s));
        SP_HIP(hipEventRecord(pr->ev[1], s));
        if (chain_graph) {
            SP_HIP(hipGraphLaunch(pr->gexecC, s));
        } else {
            const StreamScope scope(p->ctx, s);
            SP_TRY((pr->cfg.delta_correction ? coarse_chain_delta(pr, false) : coarse_chain(pr, false)));
        }
        SP_HIP(hipEventRecord(pr->ev[2], s));
    } else {
        if (pr->evU.empty()) {
            pr->evU.assign((size_t)2 * pr->S, nullptr);
            for (hipEvent_t& e : pr->evU) SP_HIP(hipEventCreate(&e));
        }
        SP_HIP(hipEventRecord(pr->ev[0], s));
        SP_TRY(fine_sweep(pr));
        SP_HIP(hipEventRecord(pr->ev[1], s));
        SP_TRY((pr->cfg.delta_correction ? coarse_chain_delta(pr, true) : coarse_chain(pr, true)));
        SP_HIP(hipEventRecord(pr->ev[2], s));
    }
    SP_HIP(hipStreamSynchronize(s));
    // phase times from the events on the stream: the fine sweep, and the coarse chain with
    // its corrections (the corrections alone when each was bracketed)
    float f_ms = 0.f, c_ms = 0.f;
    SP_HIP(hipEventElapsedTime(&f_ms, pr->ev[0], pr->ev[1]));
    SP_HIP(hipEventElapsedTime(&c_ms, pr->ev[1], pr->ev[2]));
    pr->fine_ms = f_ms;
    pr->chain_ms = c_ms;
    pr->update_ms = 0.;
    if (!graphs)
        for (int r = 0; r < pr->S; ++r) {
            float u = 0.f;
            SP_HIP(hipEventElapsedTime(&u, pr->evU[2 * r], pr->evU[2 * r + 1]));
            pr->update_ms += u;
        }
    if (tfine) *tfine = f_ms * 1e-3;
    if (tcoarse) *tcoarse = c_ms * 1e-3;
    SP_HIP(hipMemcpy(pr->part_host.data(), pr->part.p, pr->part_host.size() * sizeof(double), hipMemcpyDeviceToHost));
    double d2 = 0., q2 = 0.;
    for (size_t i = 0; i < pr->part_host.size(); i += 2) {
        d2 += pr->part_host[i];
        q2 += pr->part_host[i + 1];
    }
    // No change at all (an all-zero solution stays zero) is a zero increment, not 0/0.
    const double inc = d2 == 0. ? 0. : std::sqrt(d2) / std::sqrt(q2);
    pr->iters_done++;
    if (incr) *incr = inc;
    // A non-finite increment, or slice states whose L2 norm exceeds the initial data's by
    // orders of magnitude (the exact solution's norm never grows), mean an unstable
    // propagator: report it instead of returning the data.
    const bool blown = !(q2 <= 1e8 * pr->S * pr->u0_norm2);
    SP_CHECK(std::isfinite(inc) && !blown, SP_ERR_DIVERGED,
             "Parareal diverged (increment " + std::to_string(inc) + "): coarse or fine step unstable");
    return SP_OK;
}

int sp_parareal_iterate(sp_parareal* pr, double* increment)
{
    SP_CHECK(pr, SP_ERR_INVALID, "sp_parareal_iterate: null handle");
    SP_CHECK(pr->begun, SP_ERR_INVALID, "sp_parareal_iterate: call sp_parareal_begin first");
    SP_HIP(hipSetDevice(pr->p->ctx->device));
    return iterate_impl(pr, increment, nullptr, nullptr);
}

int sp_parareal_iteration_times(const sp_parareal* pr, double* fine_s, double* chain_s, double* update_s)
{
    SP_CHECK(pr, SP_ERR_INVALID, "sp_parareal_iteration_times: null handle");
    SP_CHECK(pr->iters_done > 0, SP_ERR_INVALID, "sp_parareal_iteration_times: no iteration has run");
    if (fine_s) *fine_s = pr->fine_ms * 1e-3;
    if (chain_s) *chain_s = pr->chain_ms * 1e-3;
    if (update_s) *update_s = pr->update_ms * 1e-3;
    return SP_OK;
}

int sp_parareal_solution(sp_parareal* pr, const sp_layout* lq, void* q)
{
    SP_CHECK(pr, SP_ERR_INVALID, "sp_parareal_solution: null handle");
    sp_prop* p = pr->p;
    SP_TRY(check_prop_layout(p, lq, q, "sp_parareal_solution"));
    SP_CHECK(lq->dtype == SP_F64, SP_ERR_INVALID, "sp_parareal_solution: q must be SP_F64");
    SP_CHECK(pr->begun, SP_ERR_INVALID, "sp_parareal_solution: call sp_parareal_begin first");
    SP_HIP(hipSetDevice(p->ctx->device));
    return launch_convert(p->ctx->stream, p->ni, p->nj, p->nk, 1, SP_F64, cview(pr->Q.p, p->ni, p->nj), SP_F64,
                          view_of(*lq, q));
}

int sp_parareal_run(sp_parareal* pr, const sp_layout* l0, const void* u0, const sp_layout* lq, void* q,
                    sp_parareal_report* rep)
{
    SP_CHECK(pr, SP_ERR_INVALID, "sp_parareal_run: null handle");
    if (rep) {
        std::memset(rep, 0, sizeof(*rep));
        rep->steps_fine_per_slice = pr->tsF;
        rep->steps_coarse_per_slice = pr->tsG;
        rep->dt_fine = pr->dtF;
        rep->dt_coarse = pr->dtG;
    }
    const double t0 = now();
    SP_TRY(sp_parareal_begin(pr, l0, u0));
    const double t1 = now();
    if (rep) rep->t_init = t1 - t0;
    int k = 0;
    for (; k < pr->cfg.kmax; ++k) {
        double inc = 0., tf = 0., tc = 0.;
        SP_TRY(iterate_impl(pr, &inc, &tf, &tc));
        if (rep) {
            rep->increment[k] = inc;
            rep->t_fine[k] = tf;
            rep->t_coarse[k] = tc;
            rep->ms_fine_events[k] = pr->fine_ms;
            rep->t_update[k] = pr->update_ms * 1e-3;
            rep->iterations = k + 1;
        }
        if (pr->cfg.tol > 0 && inc <= pr->cfg.tol) {
            ++k;
            break;
        }
    }
    if (q) SP_TRY(sp_parareal_solution(pr, lq, q));
    SP_HIP(hipStreamSynchronize(pr->p->ctx->stream));
    if (rep) rep->t_total = now() - t0;
    return SP_OK;
}

int sp_parareal_serial(sp_parareal* pr, const sp_layout* l0, const void* u0, const sp_layout* lq, void* q)
{
    SP_CHECK(pr, SP_ERR_INVALID, "sp_parareal_serial: null handle");
    // Parareal.h:250-253: DoRK4(qinitial, q, 0, dtFine, tsFine*S)
    return sp_do_rk4(pr->p, l0, u0, lq, q, 0., pr->dtF, pr->tsF * pr->S);
}

}  // extern "C"
