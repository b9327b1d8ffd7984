// asv_batch.cpp -- several clips running their Lloyd loops together on one
// GPU: the batched form of Render::run (GPU-AS/Render.cpp:254-319) for many
// clips of uneven sizes.
//
// Each clip keeps its own context (uploads, prepare, downloads go through
// the clip API).  Per Lloyd iteration the batch enqueues every clip's seed
// stage, then for each JFA pass index p one launch of the marching kernel
// per distinct step K among the clips that have a pass p (a (clip, column)
// work queue in one persistent launch), then every clip's reduction and
// finalisation.  Because every iteration starts in label buffer 0, pass p
// always reads buffer p & 1, so an iteration's launches are the same in
// every iteration of one kind; each kind (first, middle of either parity,
// last) is captured once as a HIP graph and replayed.
//
// Before the first capture an autotuner picks, per batched launch, the
// persistent block count and the work-item granularity (work items per
// resident block, which sets the slices per segment) by timing candidates
// with HIP events on the batch's own first iteration.
#include "../../include/asv.h"
#include "asv_ctx.h"
#include "asv_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace asv;

namespace {

thread_local std::string g_batch_err;

int bfail(int code, const char* msg, const char* detail = nullptr)
{
    g_batch_err = detail ? std::string(msg) + ": " + detail : std::string(msg);
    return code;
}

#define BHIP(expr)                                                                 \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) return bfail(ASV_ERR_HIP, #expr, hipGetErrorString(e_)); \
    } while (0)

// One launch of pass index p: either the marching kernel over `clips`
// (all with step K), or one clip's own pass (steps above 16, scan-order tie,
// or more ids than the marching table holds).
struct PassGroup {
    int p = 0, K = 0;
    bool batched = false;
    std::vector<int> clips;
    std::vector<JfaClip> desc;     // host copy (seg, nseg, item_begin planned)
    size_t dev_off = 0;            // offset of desc in the device array
    int grid = 0, nitems = 0, dyn = 0;
    int per_block = 0, grid_limit = 0, seg_cap = 0;   // per_block 0: planned by predicted time
    bool split = false;            // tuned: the clips' own launches are faster
};

JfaClip descriptor(const asv_ctx* c, int p)
{
    JfaClip d{};
    d.in = c->labels[p & 1];
    d.out = c->labels[1 - (p & 1)];
    d.lab = c->lab;
    d.rec = c->rec;
    d.W = c->d.W;
    d.H = c->d.H;
    d.D = c->d.D;
    d.A = c->d.A;
    d.tiles_x = (c->d.W + 15) / 16;
    d.tiles_y = (c->d.H + 15) / 16;
    d.ratio = c->mult[(size_t)p] * c->ratio;
    d.unsafe = c->unsafe;
    return d;
}

}  // namespace

struct asv_batch {
    std::vector<asv_ctx*> clips;
    int iters = 0;
    int pmax = 0;
    std::vector<PassGroup> groups;          // in pass order
    JfaClip* dev_desc = nullptr;
    size_t n_desc = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraphExec_t graph[2][2][2] = {};     // [first][last][iteration parity]
    bool tuned = false;
    bool autotune = true;
    float tune_ms = 0.0f;

    ~asv_batch()
    {
        for (auto& a : graph)
            for (auto& b : a)
                for (auto& g : b)
                    if (g) (void)hipGraphExecDestroy(g);
        if (dev_desc) (void)hipFree(dev_desc);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (st) (void)hipStreamDestroy(st);
        for (asv_ctx* c : clips) asv_destroy(c);
    }
};

namespace {

int upload_group(asv_batch* b, PassGroup& g)
{
    g.grid = march_plan(g.K, g.desc.data(), (int)g.desc.size(), g.per_block, g.grid_limit, g.seg_cap, &g.nitems,
                        &g.dyn);
    if (g.grid <= 0) return bfail(ASV_ERR_ARG, "no marching kernel for this step");
    BHIP(hipMemcpy(b->dev_desc + g.dev_off, g.desc.data(), sizeof(JfaClip) * g.desc.size(), hipMemcpyHostToDevice));
    return ASV_OK;
}

int launch_group(asv_batch* b, const PassGroup& g, hipStream_t st)
{
    if (g.batched && !g.split) {
        BHIP(launch_march_batch(st, g.K, b->dev_desc + g.dev_off, (int)g.desc.size(), g.grid, g.nitems, g.dyn));
        return ASV_OK;
    }
    for (int i : g.clips) {
        const asv_ctx* c = b->clips[(size_t)i];
        const float ratio = c->mult[(size_t)g.p] * c->ratio;
        const int32_t* in = c->labels[g.p & 1];
        int32_t* out = c->labels[1 - (g.p & 1)];
        switch (c->pass_kind((size_t)g.p)) {       // every iteration starts with the seed stage
        case 1:
            BHIP(launch_jfa_seeds(st, c->d, in, out, c->lab, c->rec, c->pos, g.K, ratio, c->p.tie_rule,
                                  c->first_pass_fill()));
            break;
        case 2:
            BHIP(launch_jfa_seeds2(st, c->d, in, out, c->lab, c->rec, c->pos, c->steps[0], g.K, ratio, c->p.tie_rule,
                                   c->keys));
            break;
        default: BHIP(launch_jfa(st, c->d, in, out, c->lab, c->rec, g.K, ratio, c->p.tie_rule, c->unsafe)); break;
        }
    }
    return ASV_OK;
}

// Seed stages, the JFA passes group by group, reductions and finalisation
// of iteration `iter` for every clip, enqueued on st.
int enqueue_iteration(asv_batch* b, int iter, hipStream_t st)
{
    int rc;
    for (asv_ctx* c : b->clips)
        if ((rc = asv_seed_stage(c, iter, st))) return rc;
    for (const PassGroup& g : b->groups)
        if ((rc = launch_group(b, g, st))) return rc;
    for (asv_ctx* c : b->clips) {
        c->fresh = false;
        c->cur = (int)(c->steps.size() & 1);
        if ((rc = asv_centroids(c, iter, st))) return rc;
    }
    return ASV_OK;
}

// Times each batched group's candidate launch shapes (grid limit x work
// items per block x slices per segment, and the clips' own launches) on
// iteration 1's own labels (every pass reads the previous pass's output),
// keeps the fastest, and leaves the group planned with it.
int autotune(asv_batch* b)
{
    hipStream_t st = b->st;
    int rc;
    for (asv_ctx* c : b->clips)
        if ((rc = asv_seed_stage(c, 1, st))) return rc;
    hipEvent_t e0, e1;
    BHIP(hipEventCreate(&e0));
    BHIP(hipEventCreate(&e1));
    float total = 0.0f;
    for (PassGroup& g : b->groups) {
        if (g.batched) {
            const int gmax = march_max_grid(g.K, g.dyn);
            struct Shape {
                int gl, pb, cap;
                bool split;
            };
            std::vector<Shape> cand;
            for (int gl : {0, (gmax * 3 / 4 / 8) * 8})
                for (int pb : {0, 2, 4, 8})
                    for (int cap : {0, 32, 12}) cand.push_back({gl, pb, cap, false});
            if (g.clips.size() > 1) cand.push_back({0, 0, 0, true});
            auto time_shape = [&](const Shape& sh, float* ms) -> int {
                g.grid_limit = sh.gl;
                g.per_block = sh.pb;
                g.seg_cap = sh.cap;
                g.split = sh.split;
                int r;
                if (!sh.split && (r = upload_group(b, g))) return r;
                if ((r = launch_group(b, g, st))) return r;             // warm
                BHIP(hipEventRecord(e0, st));
                for (int k = 0; k < 3; ++k)
                    if ((r = launch_group(b, g, st))) return r;
                BHIP(hipEventRecord(e1, st));
                BHIP(hipEventSynchronize(e1));
                BHIP(hipEventElapsedTime(ms, e0, e1));
                return ASV_OK;
            };
            float best = 1e30f;
            Shape bs = cand[0];
            for (const Shape& sh : cand) {
                float ms = 0.0f;
                if ((rc = time_shape(sh, &ms))) return rc;
                total += ms;
                if (ms < best) {
                    best = ms;
                    bs = sh;
                }
            }
            g.grid_limit = bs.gl;
            g.per_block = bs.pb;
            g.seg_cap = bs.cap;
            g.split = bs.split;
            if ((rc = upload_group(b, g))) return rc;
        }
        if ((rc = launch_group(b, g, st))) return rc;                   // the pass's real output
    }
    BHIP(hipStreamSynchronize(st));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->tune_ms = total;
    b->tuned = true;
    return ASV_OK;
}

}  // namespace

extern "C" {

int asv_batch_create(const asv_params* params, int32_t n, asv_batch** out)
{
    if (!out) return bfail(ASV_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!params || n < 1) return bfail(ASV_ERR_ARG, "need at least one clip");
    for (int i = 1; i < n; ++i)
        if (params[i].lloyd_iters != params[0].lloyd_iters)
            return bfail(ASV_ERR_ARG, "clips of one batch share the Lloyd iteration count");
    asv_batch* b = new asv_batch();
    b->iters = params[0].lloyd_iters;
    for (int i = 0; i < n; ++i) {
        asv_ctx* c = nullptr;
        int rc = asv_create(&params[i], &c);
        if (rc) {
            delete b;
            return bfail(rc, "clip context", asv_last_error());
        }
        b->clips.push_back(c);
        b->pmax = std::max(b->pmax, (int)c->steps.size());
    }
    // pass groups: batched per (p, K) where the marching kernel applies
    for (int p = 0; p < b->pmax; ++p) {
        std::vector<PassGroup> row;
        for (int i = 0; i < n; ++i) {
            const asv_ctx* c = b->clips[(size_t)i];
            if (p >= (int)c->steps.size()) continue;
            const int K = c->steps[(size_t)p];
            // passes over the seeds' neighbourhoods alone run per clip
            const bool march = c->pass_kind((size_t)p) == 0 && K <= 16 && c->p.tie_rule == ASV_TIE_LOWEST_ID &&
                               c->d.A <= kMarchMaxIds;
            PassGroup* g = nullptr;
            if (march)
                for (PassGroup& q : row)
                    if (q.batched && q.K == K) g = &q;
            if (!g) {
                row.emplace_back();
                g = &row.back();
                g->p = p;
                g->K = K;
                g->batched = march;
            }
            g->clips.push_back(i);
            if (march) g->desc.push_back(descriptor(c, p));
        }
        for (PassGroup& g : row) {
            g.dev_off = b->n_desc;
            b->n_desc += g.desc.size();
            b->groups.push_back(std::move(g));
        }
    }
    hipError_t e = hipStreamCreate(&b->st);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming);
    if (e == hipSuccess && b->n_desc) e = asv::dev_malloc((void**)&b->dev_desc, sizeof(JfaClip) * b->n_desc);
    if (e != hipSuccess) {
        delete b;
        return bfail(ASV_ERR_HIP, "batch set-up", hipGetErrorString(e));
    }
    for (PassGroup& g : b->groups)
        if (g.batched) {
            int rc = upload_group(b, g);
            if (rc) {
                delete b;
                return rc;
            }
        }
    *out = b;
    return ASV_OK;
}

int asv_batch_destroy(asv_batch* b)
{
    delete b;
    return ASV_OK;
}

int asv_batch_size(const asv_batch* b) { return b ? (int)b->clips.size() : -1; }

int asv_batch_clip(asv_batch* b, int32_t i, asv_ctx** ctx)
{
    if (!b || !ctx || i < 0 || i >= (int)b->clips.size()) return bfail(ASV_ERR_ARG, "no such clip");
    *ctx = b->clips[(size_t)i];
    return ASV_OK;
}

int asv_batch_set_autotune(asv_batch* b, int32_t on)
{
    if (!b) return bfail(ASV_ERR_ARG, "null batch");
    b->autotune = on != 0;
    return ASV_OK;
}

int asv_batch_launch_count(const asv_batch* b, int32_t* per_iteration)
{
    if (!b || !per_iteration) return bfail(ASV_ERR_ARG, "null argument");
    *per_iteration = (int32_t)b->groups.size();
    return ASV_OK;
}

int asv_batch_run(asv_batch* b, void* stream)
{
    if (!b) return bfail(ASV_ERR_ARG, "null batch");
    hipStream_t user = reinterpret_cast<hipStream_t>(stream);
    BHIP(hipEventRecord(b->ev_in, user));
    BHIP(hipStreamWaitEvent(b->st, b->ev_in, 0));
    int rc;
    for (asv_ctx* c : b->clips)                  // LAB + metric, as asv_run does
        if ((rc = asv_prepare(c, b->st))) return bfail(rc, "prepare", asv_last_error());
    if (b->autotune && !b->tuned && (rc = autotune(b))) return rc;
    for (int iter = 1; iter <= b->iters; ++iter) {
        const int f = iter == 1, l = iter == b->iters, par = iter & 1;
        hipGraphExec_t& ex = b->graph[f][l][par];
        if (!ex) {
            hipGraph_t g = nullptr;
            BHIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
            rc = enqueue_iteration(b, iter, b->st);
            hipError_t e = hipStreamEndCapture(b->st, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return bfail(rc, "capture", asv_last_error());
            }
            if (e != hipSuccess) return bfail(ASV_ERR_HIP, "hipStreamEndCapture", hipGetErrorString(e));
            e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                ex = nullptr;
                return bfail(ASV_ERR_HIP, "hipGraphInstantiate", hipGetErrorString(e));
            }
        }
        BHIP(hipGraphLaunch(ex, b->st));
        for (asv_ctx* c : b->clips) {            // host state the stage calls keep
            c->last_iter = iter;
            c->cur = (int)(c->steps.size() & 1);
        }
    }
    BHIP(hipEventRecord(b->ev_out, b->st));
    BHIP(hipStreamWaitEvent(user, b->ev_out, 0));
    return ASV_OK;
}

int asv_batch_tuning(const asv_batch* b, int32_t* shapes, int32_t cap, float* tune_ms)
{
    if (!b) return bfail(ASV_ERR_ARG, "null batch");
    int n = 0;
    for (const PassGroup& g : b->groups) {
        if (shapes && n < cap) {
            shapes[5 * n + 0] = g.p;
            shapes[5 * n + 1] = g.K;
            shapes[5 * n + 2] = g.batched && !g.split ? (int)g.desc.size() : -(int)g.clips.size();
            shapes[5 * n + 3] = g.grid;
            shapes[5 * n + 4] = g.per_block;
        }
        ++n;
    }
    if (tune_ms) *tune_ms = b->tune_ms;
    return n;
}

const char* asv_batch_last_error(void) { return g_batch_err.c_str(); }

}  // extern "C"
