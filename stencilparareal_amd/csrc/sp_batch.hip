// sp_batch.hip -- batches of independent Parareal problems on one GPU (sp_batch_* of the C
// ABI).  Each problem has its own grid, slice count, coefficients and horizon; every launch
// covers all of them:
//   * the fine sweep of all problems' slices is one grouped fused-RK4 launch per step (per
//     tile shape in use), each problem an entry with its own sizes, views and coefficients;
//   * position j of the coarse chains is one grouped launch per coarse sub-step over the
//     problems that have a slice j, followed by one grouped correction launch;
//   * the correction is the difference form U_{j+1} = F(U_j^k) + G(U_j^{k+1} - U_j^k), equal to
//     the reference's update (src/Parareal/Parareal.h:19-35,143-164) for this linear G.
// The launch records are built once at creation, so one iteration can be captured as a
// hipGraph and replayed.  Problems use constant nu (nufreq = 0).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "sp_handles.hpp"
#include "sp_internal.hpp"

using namespace sp;

namespace {

constexpr int kCorrBlocks = 64;  // blocks (and partial sums) per slice correction

struct Prob {
    sp_problem d{};
    int S = 0, tsF = 0, tsG = 0;  // steps per slice (tsG counts coarse sub-steps)
    double dtF = 0, dtG = 0;
    size_t n3 = 0;
    Buf U, F, FS, Gw0, Gw1, D, Q;
    sp_prop* prop = nullptr;  // this problem's propagator (serial solution, tuning)
    FusedTile tileF{}, tileG{};
};

// One grouped fused-RK4 launch: records and first-block table at arena offsets.
struct GLaunch {
    FusedTile tile{};
    int idt = SP_F64, odt = SP_F64;
    size_t rec = 0, first = 0;
    int n = 0;
    long long blocks = 0;
};

struct Entry {
    FusedStep st;
    FusedTile tile;
    View in, out;
};

View cv(const void* p, int n) { return View{(void*)p, n, (long long)n * n, 0}; }
View slice_of(const Buf& b, const Prob& q, int j, int dt = SP_F64)
{
    return cv((const char*)b.p + (size_t)j * q.n3 * dtype_size(dt), q.d.n);
}
View batch_of(const Buf& b, const Prob& q)
{
    View v = cv(b.p, q.d.n);
    v.slice = (long long)q.n3;
    return v;
}

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct sp_batch {
    sp_ctx* ctx = nullptr;
    sp_batch_cfg cfg{};
    int gdt = SP_F64;
    std::vector<Prob> pr;
    int maxS = 0;
    Buf part;                         // partial sums: per problem, per slice, kCorrBlocks x 2
    std::vector<size_t> part_off;     // first partial of each problem (doubles)
    std::vector<double> part_host;
    std::vector<char> arena_host;     // launch records, uploaded to arena
    Buf arena;
    std::vector<std::vector<GLaunch>> groups;  // distinct launch groups
    std::vector<int> fine_plan;                // fine step i -> group
    std::vector<std::vector<int>> chain_plan;  // position j -> one group per coarse sub-step
    std::vector<size_t> corr_rec;              // position j -> correction records
    std::vector<int> corr_n;
    std::vector<std::vector<int>> init_plan;   // initial chain position j (1..maxS) -> groups
    std::vector<std::vector<double>> inc;  // per problem, per iteration
    hipGraphExec_t gexec = nullptr;   // the fine sweep of one iteration
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexecC = nullptr;  // the coarse chains and corrections of one iteration
    hipGraph_t graphC = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    std::vector<double> t_fine, t_chain;  // per iteration of the last run, milliseconds
    int iterations = 0;
};

namespace {

size_t arena_put(sp_batch* b, const void* data, size_t bytes)
{
    size_t off = (b->arena_host.size() + 255) & ~(size_t)255;
    b->arena_host.resize(off + bytes);
    std::memcpy(b->arena_host.data() + off, data, bytes);
    return off;
}

// Groups entries by tile shape (pair / 2x2 kernels need even, aligned fields, the 2x2 kernel
// FP64 storage: other entries take the scalar 32x16x512 kernel) and appends the launches'
// records; returns the group index.
int add_group(sp_batch* b, std::vector<Entry> ents, int idt, int odt, std::map<std::string, int>& memo,
              const std::string& key)
{
    const auto hit = memo.find(key);
    if (hit != memo.end()) return hit->second;
    std::map<std::vector<int>, std::vector<Entry>> by_shape;
    for (Entry& e : ents) {
        // the x-pair kernel takes any storage types, the 2x2 one FP64 only
        const bool f64 = idt == SP_F64 && odt == SP_F64;
        if (e.tile.variant != SP_KERNEL_SCALAR &&
            !((f64 || e.tile.variant == SP_KERNEL_PAIR) && fused_record_pair_ok(e.st, e.in, e.out, idt, odt)))
            e.tile = FusedTile{32, 16, 512, e.tile.kchunk, SP_KERNEL_SCALAR};
        by_shape[{e.tile.variant, e.tile.tx, e.tile.ty, e.tile.nt}].push_back(e);
    }
    std::vector<GLaunch> out;
    const size_t rs = fused_record_size();
    for (auto& kv : by_shape) {
        std::vector<char> recs(rs * kv.second.size());
        std::vector<int> first(kv.second.size());
        long long total = 0;
        for (size_t i = 0; i < kv.second.size(); ++i) {
            const Entry& e = kv.second[i];
            long long nb = 0;
            const int r = fused_record_fill(recs.data() + i * rs, e.st, e.tile, idt, e.in, odt, e.out, &nb);
            if (r != SP_OK) return -r;
            first[i] = (int)total;
            total += nb;
        }
        GLaunch g;
        g.tile = kv.second.front().tile;
        g.idt = idt;
        g.odt = odt;
        g.rec = arena_put(b, recs.data(), recs.size());
        g.first = arena_put(b, first.data(), first.size() * sizeof(int));
        g.n = (int)kv.second.size();
        g.blocks = total;
        out.push_back(g);
    }
    b->groups.push_back(out);
    const int id = (int)b->groups.size() - 1;
    memo[key] = id;
    return id;
}

FusedStep step_of(const Prob& q, int nslices, double dt)
{
    FusedStep st;
    st.ni = st.nj = st.nk = q.d.n;
    st.nslices = nslices;
    st.dt = dt;
    st.cx = q.d.cx;
    st.cy = q.d.cy;
    st.cz = q.d.cz;
    st.dx = 1.0 / q.d.n;
    st.coef_dev = nullptr;
    st.nu = q.d.nu;
    return st;
}

int run_groups(sp_batch* b, hipStream_t s, int gi)
{
    for (const GLaunch& g : b->groups[gi])
        SP_TRY(launch_rk4_grouped(s, g.tile, g.idt, g.odt, (const char*)b->arena.p + g.rec,
                                  (const int*)((const char*)b->arena.p + g.first), g.n, g.blocks));
    return SP_OK;
}

int build_plans(sp_batch* b)
{
    const int gdt = b->gdt;
    const int P = (int)b->pr.size();
    std::map<std::string, int> memo;
    auto mask = [&](auto pred) {
        std::string m(P, '0');
        for (int p = 0; p < P; ++p)
            if (pred(b->pr[p])) m[p] = '1';
        return m;
    };
    // fine sweep: step i over every problem with tsF > i; the last step lands in F
    int maxF = 0, maxG = 0;
    for (const Prob& q : b->pr) {
        maxF = std::max(maxF, q.tsF);
        maxG = std::max(maxG, q.tsG);
    }
    for (int i = 0; i < maxF; ++i) {
        std::vector<Entry> ents;
        for (const Prob& q : b->pr) {
            if (q.tsF <= i) continue;
            const auto dst = [&](int k) { return ((q.tsF - 1 - k) % 2 == 0) ? batch_of(q.F, q) : batch_of(q.FS, q); };
            ents.push_back(Entry{step_of(q, q.S, q.dtF), q.tileF, i == 0 ? batch_of(q.U, q) : dst(i - 1), dst(i)});
        }
        const std::string key = "F" + std::to_string(i == 0) + std::to_string(i & 1) + mask([&](const Prob& q) {
                                    return q.tsF > i;
                                });
        const int g = add_group(b, ents, SP_F64, SP_F64, memo, key);
        if (g < 0) return -g;
        b->fine_plan.push_back(g);
    }
    // coarse chains, difference form: at position j >= 1, G(D) over tsG sub-steps into Gw0
    b->chain_plan.assign(b->maxS, {});
    for (int j = 1; j < b->maxS; ++j) {
        for (int i = 0; i < maxG; ++i) {
            std::vector<Entry> ents;
            for (const Prob& q : b->pr) {
                if (q.S <= j || q.tsG <= i) continue;
                const auto dst = [&](int k) { return ((q.tsG - 1 - k) % 2 == 0) ? cv(q.Gw0.p, q.d.n) : cv(q.Gw1.p, q.d.n); };
                ents.push_back(Entry{step_of(q, 1, q.dtG), q.tileG, i == 0 ? cv(q.D.p, q.d.n) : dst(i - 1), dst(i)});
            }
            if (ents.empty()) continue;
            const std::string key = "G" + std::to_string(i == 0) + std::to_string(i & 1) + mask([&](const Prob& q) {
                                        return q.S > j && q.tsG > i;
                                    });
            const int g = add_group(b, ents, i == 0 ? SP_F64 : gdt, gdt, memo, key);
            if (g < 0) return -g;
            b->chain_plan[j].push_back(g);
        }
    }
    // corrections at position j: U_{j+1} (or the final state) = F_j + G(d_j); d_{j+1} -> D
    b->corr_rec.assign(b->maxS, 0);
    b->corr_n.assign(b->maxS, 0);
    for (int j = 0; j < b->maxS; ++j) {
        std::vector<CorrRecord> recs;
        for (int p = 0; p < P; ++p) {
            const Prob& q = b->pr[p];
            if (q.S <= j) continue;
            const bool last = j == q.S - 1;
            CorrRecord r;
            r.gd = j == 0 ? cv(nullptr, q.d.n) : cv(q.Gw0.p, q.d.n);
            r.fine = slice_of(q.F, q, j);
            r.dst = last ? cv(q.Q.p, q.d.n) : slice_of(q.U, q, j + 1);
            r.dnext = last ? cv(nullptr, q.d.n) : cv(q.D.p, q.d.n);
            r.ni = r.nj = r.nk = q.d.n;
            r.partials = (double*)b->part.p + b->part_off[p] + (size_t)j * kCorrBlocks * 2;
            recs.push_back(r);
        }
        b->corr_rec[j] = arena_put(b, recs.data(), recs.size() * sizeof(CorrRecord));
        b->corr_n[j] = (int)recs.size();
    }
    // initial chains: position j (1..maxS) advances each problem with S >= j by one slice of
    // coarse steps, from u0 (j = 1) or the previous position's state, alternating Gw0 / Gw1
    // with D as the sub-step scratch; the state is then U_j (j < S) or the final state Q.
    b->init_plan.assign(b->maxS + 1, {});
    for (int j = 1; j <= b->maxS; ++j) {
        for (int i = 0; i < maxG; ++i) {
            std::vector<Entry> ents;
            for (const Prob& q : b->pr) {
                if (q.S < j || q.tsG <= i) continue;
                const View out = ((j - 1) % 2 == 0) ? cv(q.Gw0.p, q.d.n) : cv(q.Gw1.p, q.d.n);
                const View prev = ((j - 2) % 2 == 0) ? cv(q.Gw0.p, q.d.n) : cv(q.Gw1.p, q.d.n);
                const View scr = cv(q.D.p, q.d.n);
                const auto dst = [&](int k) { return ((q.tsG - 1 - k) % 2 == 0) ? out : scr; };
                const View src = i > 0 ? dst(i - 1) : (j == 1 ? slice_of(q.U, q, 0) : prev);
                ents.push_back(Entry{step_of(q, 1, q.dtG), q.tileG, src, dst(i)});
            }
            if (ents.empty()) continue;
            const int idt = (i == 0 && j == 1) ? SP_F64 : gdt;
            const std::string key = "I" + std::to_string(j) + "_" + std::to_string(i) + mask([&](const Prob& q) {
                                        return q.S >= j && q.tsG > i;
                                    });
            const int g = add_group(b, ents, idt, gdt, memo, key);
            if (g < 0) return -g;
            b->init_plan[j].push_back(g);
        }
    }
    SP_TRY(b->arena.ensure(b->arena_host.size()));
    SP_HIP(hipMemcpy(b->arena.p, b->arena_host.data(), b->arena_host.size(), hipMemcpyHostToDevice));
    return SP_OK;
}

// One iteration is the fine sweep (every problem's slices, one grouped launch per fine step
// and tile shape), then the coarse chains with the corrections, position by position.
int fine_sweep(sp_batch* b, hipStream_t s)
{
    for (int g : b->fine_plan) SP_TRY(run_groups(b, s, g));
    return SP_OK;
}

int chains(sp_batch* b, hipStream_t s)
{
    for (int j = 0; j < b->maxS; ++j) {
        for (int g : b->chain_plan[j]) SP_TRY(run_groups(b, s, g));
        SP_TRY(launch_correction_delta_grouped(s, b->gdt, (const CorrRecord*)((const char*)b->arena.p + b->corr_rec[j]),
                                               b->corr_n[j], kCorrBlocks));
    }
    return SP_OK;
}

// Records what fn launches on stream s into an instantiated graph.
template <typename F>
int capture(hipStream_t s, hipGraph_t* graph, hipGraphExec_t* exec, F fn)
{
    SP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rr = fn();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rr != SP_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rr;
    }
    SP_HIP(e);
    *graph = g;
    SP_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return SP_OK;
}

void release(sp_batch* b)
{
    if (!b) return;
    if (b->ctx) {
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
    }
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->gexecC) (void)hipGraphExecDestroy(b->gexecC);
    if (b->graphC) (void)hipGraphDestroy(b->graphC);
    for (hipEvent_t e : b->ev)
        if (e) (void)hipEventDestroy(e);
    for (Prob& q : b->pr) {
        for (Buf* x : {&q.U, &q.F, &q.FS, &q.Gw0, &q.Gw1, &q.D, &q.Q}) x->release();
        if (q.prop) sp_prop_destroy(q.prop);
    }
    b->part.release();
    b->arena.release();
    if (b->ctx) sp_ctx_destroy(b->ctx);  // drops this batch's reference
    delete b;
}

}  // namespace

extern "C" {

int sp_batch_create(sp_ctx* ctx, int nprob, const sp_problem* probs, const sp_batch_cfg* cfg, sp_batch** out)
{
    SP_CHECK(ctx && probs && cfg && out, SP_ERR_INVALID, "sp_batch_create: null argument");
    SP_CHECK(nprob >= 1, SP_ERR_INVALID, "sp_batch_create: no problems");
    SP_CHECK(cfg->kmax >= 0 && cfg->kmax <= SP_MAX_ITERS, SP_ERR_INVALID, "sp_batch_create: kmax out of range");
    SP_CHECK(cfg->coarse_dtype == SP_F64 || cfg->coarse_dtype == SP_F32 || cfg->coarse_dtype == SP_BF16,
             SP_ERR_INVALID, "sp_batch_create: coarse_dtype");
    SP_CHECK(cfg->coarse_substeps >= 1 && cfg->tol >= 0, SP_ERR_INVALID, "sp_batch_create: substeps or tol");
    for (int p = 0; p < nprob; ++p) {
        const sp_problem& d = probs[p];
        const std::string w = "sp_batch_create: problem " + std::to_string(p) + ": ";
        SP_CHECK(d.n >= kFusedMinK && d.n % 2 == 0, SP_ERR_INVALID, w + "n must be even and >= 8");
        SP_CHECK(d.slices >= 1 && d.steps_fine >= 1 && d.steps_coarse >= 1, SP_ERR_INVALID, w + "slices / steps");
        SP_CHECK(d.end_time > 0 && std::isfinite(d.end_time) && std::isfinite(d.nu), SP_ERR_INVALID, w + "end_time / nu");
    }
    SP_HIP(hipSetDevice(ctx->device));
    sp_batch* b = new sp_batch;
    b->ctx = ctx;
    ctx->refs++;
    b->cfg = *cfg;
    b->gdt = cfg->coarse_dtype;
    b->pr.resize(nprob);
    int r = SP_OK;
    auto chk = [&](int x) {
        if (r == SP_OK && x != SP_OK) r = x;
    };
    size_t part_total = 0;
    std::map<std::vector<int>, FusedTile> tuned;  // (n, slices, f64) -> tile, one tuning per shape
    // The fine sweep is one grouped launch per tile shape in use, and a problem's own choice
    // is made for the problem alone (k chunks and narrow tiles fill the device for a small
    // problem; inside a grouped launch they only add pipeline fill and split the step into
    // several launches).  Once whole-k 2x2 tiles over all problems give several blocks per
    // CU, every problem's fine step uses them.
    bool shared_fine = false;
    {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus < 1)
            cus = 256;
        long long blocks = 0;
        for (int p = 0; p < nprob; ++p) {
            const long long t = (probs[p].n + 31) / 32;
            blocks += t * t * probs[p].slices;
        }
        shared_fine = blocks >= 4LL * cus;
    }
    for (int p = 0; p < nprob && r == SP_OK; ++p) {
        Prob& q = b->pr[p];
        q.d = probs[p];
        q.S = q.d.slices;
        // totals rounded up to multiples of the slice count (RuntimeConfiguration.cpp:133-151)
        int sf = q.d.steps_fine, sc = q.d.steps_coarse;
        if (sf % q.S) sf = (sf / q.S + 1) * q.S;
        if (sc % q.S) sc = (sc / q.S + 1) * q.S;
        q.dtF = q.d.end_time / sf;
        q.tsF = sf / q.S;
        q.dtG = (q.d.end_time / sc) / cfg->coarse_substeps;
        q.tsG = (sc / q.S) * cfg->coarse_substeps;
        q.n3 = (size_t)q.d.n * q.d.n * q.d.n;
        b->maxS = std::max(b->maxS, q.S);
        const size_t nb = q.n3 * q.S;
        chk(q.U.ensure(nb * 8));
        chk(q.F.ensure(nb * 8));
        chk(q.FS.ensure(nb * 8));
        chk(q.Gw0.ensure(q.n3 * dtype_size(b->gdt)));
        chk(q.Gw1.ensure(q.n3 * dtype_size(b->gdt)));
        chk(q.D.ensure(q.n3 * 8));
        chk(q.Q.ensure(q.n3 * 8));
        chk(sp_prop_create(ctx, q.d.n, q.d.n, q.d.n, 1.0 / q.d.n, q.d.nu, 0.0, q.d.cx, q.d.cy, q.d.cz, &q.prop));
        if (r != SP_OK) break;
        b->part_off.push_back(part_total);
        part_total += (size_t)q.S * kCorrBlocks * 2;
        // tiles: autotuned once per (n, slices), else the built-in choice
        const auto pick = [&](int nslices, int dt, FusedTile& t) {
            const int f64 = dt == SP_F64;
            const std::vector<int> key{q.d.n, nslices, f64};
            const auto hit = tuned.find(key);
            if (hit != tuned.end()) {
                t = hit->second;
                return;
            }
            if (cfg->autotune) chk(sp_prop_autotune(q.prop, nslices, dt, dt, nullptr));
            int v = 0;
            chk(sp_prop_get_tile(q.prop, nslices, f64, &v, &t.tx, &t.ty, &t.nt, &t.kchunk));
            t.variant = v;
            // grouped launches have no split form: the built-in one-launch tile
            if (t.variant == SP_KERNEL_SPLIT) t = default_tile(q.d.n, q.d.n, q.d.n, nslices, f64);
            tuned[key] = t;
        };
        if (shared_fine)
            q.tileF = FusedTile{32, 32, 512, 0, SP_KERNEL_QUAD};
        else
            pick(q.S, SP_F64, q.tileF);
        pick(1, b->gdt, q.tileG);
    }
    if (r == SP_OK) r = b->part.ensure(part_total * sizeof(double));
    if (r == SP_OK) {
        b->part_host.resize(part_total);
        b->inc.assign(nprob, {});
        r = build_plans(b);
    }
    if (r != SP_OK) {
        release(b);
        return r;
    }
    *out = b;
    return SP_OK;
}

int sp_batch_destroy(sp_batch* b)
{
    release(b);
    return SP_OK;
}

int sp_batch_run(sp_batch* b, void* const* u0, void* const* q, sp_batch_report* rep)
{
    SP_CHECK(b && u0 && q, SP_ERR_INVALID, "sp_batch_run: null argument");
    const int P = (int)b->pr.size();
    for (int p = 0; p < P; ++p) SP_CHECK(u0[p] && q[p], SP_ERR_INVALID, "sp_batch_run: null field");
    SP_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    if (rep) std::memset(rep, 0, sizeof(*rep));
    const double t0 = now();
    // initial coarse chains: U_0 = u0, U_j = G(U_{j-1}) continued from t = 0, final state G(U_{S-1})
    for (int p = 0; p < P; ++p)
        SP_HIP(hipMemcpyAsync(b->pr[p].U.p, u0[p], b->pr[p].n3 * 8, hipMemcpyDeviceToDevice, s));
    for (int j = 1; j <= b->maxS; ++j) {
        for (int g : b->init_plan[j]) SP_TRY(run_groups(b, s, g));
        for (const Prob& x : b->pr) {
            if (x.S < j) continue;
            const View st = ((j - 1) % 2 == 0) ? cv(x.Gw0.p, x.d.n) : cv(x.Gw1.p, x.d.n);
            const View dst = j < x.S ? slice_of(x.U, x, j) : cv(x.Q.p, x.d.n);
            SP_TRY(launch_convert(s, x.d.n, x.d.n, x.d.n, 1, b->gdt, st, SP_F64, dst));
        }
    }
    SP_HIP(hipStreamSynchronize(s));
    const double t1 = now();
    if (rep) rep->t_init = t1 - t0;
    for (auto& v : b->inc) v.clear();
    int k = 0;
    double lup = 0.;
    for (const Prob& x : b->pr) lup += (double)x.n3 * x.S * x.tsG;  // the initial chains
    // the stream the iterations run on: a graph needs a created stream, so with the null
    // stream as the context's stream it is replayed on the context's own one
    const hipStream_t ws = (b->cfg.use_graphs && !s) ? b->ctx->own : s;
    for (hipEvent_t& e : b->ev)
        if (!e) SP_HIP(hipEventCreate(&e));
    b->t_fine.clear();
    b->t_chain.clear();
    for (; k < b->cfg.kmax; ++k) {
        // the fine sweep and the chains between HIP events (two graphs when captured)
        SP_HIP(hipEventRecord(b->ev[0], ws));
        if (b->cfg.use_graphs) {
            if (!b->gexec) {
                SP_TRY(capture(ws, &b->graph, &b->gexec, [&] { return fine_sweep(b, ws); }));
                SP_TRY(capture(ws, &b->graphC, &b->gexecC, [&] { return chains(b, ws); }));
            }
            SP_HIP(hipGraphLaunch(b->gexec, ws));
            SP_HIP(hipEventRecord(b->ev[1], ws));
            SP_HIP(hipGraphLaunch(b->gexecC, ws));
        } else {
            SP_TRY(fine_sweep(b, ws));
            SP_HIP(hipEventRecord(b->ev[1], ws));
            SP_TRY(chains(b, ws));
        }
        SP_HIP(hipEventRecord(b->ev[2], ws));
        SP_HIP(hipEventSynchronize(b->ev[2]));
        float f_ms = 0.f, c_ms = 0.f;
        SP_HIP(hipEventElapsedTime(&f_ms, b->ev[0], b->ev[1]));
        SP_HIP(hipEventElapsedTime(&c_ms, b->ev[1], b->ev[2]));
        b->t_fine.push_back(f_ms);
        b->t_chain.push_back(c_ms);
        SP_HIP(hipMemcpyAsync(b->part_host.data(), b->part.p, b->part_host.size() * sizeof(double),
                              hipMemcpyDeviceToHost, ws));
        SP_HIP(hipStreamSynchronize(ws));
        double worst = 0.;
        for (int p = 0; p < P; ++p) {
            const Prob& x = b->pr[p];
            double d2 = 0., q2 = 0.;
            const double* h = b->part_host.data() + b->part_off[p];
            for (size_t i = 0; i < (size_t)x.S * kCorrBlocks; ++i) {
                d2 += h[2 * i];
                q2 += h[2 * i + 1];
            }
            // no change at all (an all-zero problem stays zero) is a zero increment, not 0/0
            const double inc = d2 == 0. ? 0. : std::sqrt(d2) / std::sqrt(q2);
            SP_CHECK(std::isfinite(inc), SP_ERR_DIVERGED,
                     "sp_batch_run: problem " + std::to_string(p) + " diverged (non-finite increment)");
            b->inc[p].push_back(inc);
            worst = std::max(worst, inc);
            lup += (double)x.n3 * x.S * x.tsF + (double)x.n3 * (x.S - 1) * x.tsG;
        }
        if (rep) {
            rep->increment[k] = worst;
            rep->iterations = k + 1;
        }
        if (b->cfg.tol > 0 && worst <= b->cfg.tol) {
            ++k;
            break;
        }
    }
    b->iterations = k;
    for (int p = 0; p < P; ++p)
        SP_HIP(hipMemcpyAsync(q[p], b->pr[p].Q.p, b->pr[p].n3 * 8, hipMemcpyDeviceToDevice, s));
    SP_HIP(hipStreamSynchronize(s));
    if (rep) {
        rep->t_total = now() - t0;
        rep->lattice_updates = lup;
    }
    return SP_OK;
}

int sp_batch_iteration_times(const sp_batch* b, int k, double* fine_s, double* chain_s)
{
    SP_CHECK(b, SP_ERR_INVALID, "sp_batch_iteration_times: null handle");
    SP_CHECK(k >= 0 && k < (int)b->t_fine.size(), SP_ERR_INVALID,
             "sp_batch_iteration_times: iteration " + std::to_string(k) + " did not run in the last sp_batch_run");
    if (fine_s) *fine_s = b->t_fine[k] * 1e-3;
    if (chain_s) *chain_s = b->t_chain[k] * 1e-3;
    return SP_OK;
}

int sp_batch_increments(const sp_batch* b, int p, double* inc)
{
    SP_CHECK(b && inc, SP_ERR_INVALID, "sp_batch_increments: null argument");
    SP_CHECK(p >= 0 && p < (int)b->pr.size(), SP_ERR_INVALID, "sp_batch_increments: problem index");
    for (size_t k = 0; k < b->inc[p].size(); ++k) inc[k] = b->inc[p][k];
    return SP_OK;
}

int sp_batch_problem_info(const sp_batch* b, int p, sp_batch_info* info)
{
    SP_CHECK(b && info, SP_ERR_INVALID, "sp_batch_problem_info: null argument");
    SP_CHECK(p >= 0 && p < (int)b->pr.size(), SP_ERR_INVALID, "sp_batch_problem_info: problem index");
    const Prob& q = b->pr[p];
    info->steps_fine_per_slice = q.tsF;
    info->steps_coarse_per_slice = q.tsG;
    info->dt_fine = q.dtF;
    info->dt_coarse = q.dtG;
    info->fine_variant = q.tileF.variant;
    info->fine_tx = q.tileF.tx;
    info->fine_ty = q.tileF.ty;
    info->fine_nthreads = q.tileF.nt;
    info->fine_kchunk = q.tileF.kchunk;
    return SP_OK;
}

int sp_batch_serial(sp_batch* b, int p, const void* u0, void* q)
{
    SP_CHECK(b && u0 && q, SP_ERR_INVALID, "sp_batch_serial: null argument");
    SP_CHECK(p >= 0 && p < (int)b->pr.size(), SP_ERR_INVALID, "sp_batch_serial: problem index");
    const Prob& x = b->pr[p];
    sp_layout l;
    SP_TRY(sp_layout_compact(x.d.n, x.d.n, x.d.n, SP_F64, &l));
    return sp_do_rk4(x.prop, &l, u0, &l, q, 0.0, x.dtF, x.tsF * x.S);
}

}  // extern "C"
