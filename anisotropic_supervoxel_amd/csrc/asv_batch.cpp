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
#include "asv_host_internal.h"
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
    bool ht = false;               // a clip has more seeds than the marching table has slots
    bool counted = false;          // a clip keeps marching-step counters
    bool dense = false;            // few voxels per seed over the clips (march_dense)
    int lt = 0;                    // the label widths of the group's clips (launch_jfa)
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
    const int tw = march_tile_w(c->steps[(size_t)p]);
    d.tiles_x = (c->d.W + tw - 1) / tw;
    d.tiles_y = (c->d.H + kMarchTileH - 1) / kMarchTileH;
    d.ratio = c->mult[(size_t)p] * c->ratio;
    d.unsafe = c->unsafe;
    d.stats = c->march_stats;
    return d;
}

// Pass groups of a set of contexts, in pass order: for each pass index p,
// one batched group per step K over the contexts that march pass p (the
// marching kernel applies), and one group per remaining context (passes
// over the seeds' neighbourhoods, steps above 16, the scan-order tie).
// Descriptors are numbered consecutively through the groups (dev_off).
void build_groups(const std::vector<asv_ctx*>& clips, std::vector<PassGroup>* groups, size_t* n_desc)
{
    int pmax = 0;
    for (const asv_ctx* c : clips) pmax = std::max(pmax, (int)c->steps.size());
    groups->clear();
    *n_desc = 0;
    for (int p = 0; p < pmax; ++p) {
        std::vector<PassGroup> row;
        for (int i = 0; i < (int)clips.size(); ++i) {
            const asv_ctx* c = clips[(size_t)i];
            if (p >= (int)c->steps.size()) continue;
            const int K = c->steps[(size_t)p];
            // passes over the seeds' neighbourhoods alone run per clip
            const bool march = c->pass_kind((size_t)p) == 0 && K <= 16 && c->p.tie_rule == ASV_TIE_LOWEST_ID &&
                               c->d.A <= kMarchMaxIds;
            const int lt = c->label_widths((size_t)p, c->steps.size(), true);
            PassGroup* g = nullptr;
            if (march)
                for (PassGroup& q : row)
                    if (q.batched && q.K == K && q.lt == lt) g = &q;
            if (!g) {
                row.emplace_back();
                g = &row.back();
                g->p = p;
                g->K = K;
                g->batched = march;
                g->lt = lt;
            }
            g->clips.push_back(i);
            if (march) {
                g->desc.push_back(descriptor(c, p));
                g->ht = g->ht || c->d.A > kMarchTableIds;
                g->counted = g->counted || c->march_stats != nullptr;
            }
        }
        for (PassGroup& g : row) {
            if (g.batched) g.dense = march_dense(g.desc.data(), (int)g.desc.size());
            g.dev_off = *n_desc;
            *n_desc += g.desc.size();
            groups->push_back(std::move(g));
        }
    }
}

// One group's launches on st: the marching kernel over the group's
// descriptors at dev_desc + dev_off, or each context's own pass.
int launch_group(const std::vector<asv_ctx*>& clips, const JfaClip* dev_desc, const PassGroup& g, hipStream_t st)
{
    if (g.batched && !g.split) {
        BHIP(launch_march_batch(st, g.K, dev_desc + g.dev_off, (int)g.desc.size(), g.grid, g.nitems, g.dyn, g.ht,
                                g.counted, g.dense, g.lt));
        return ASV_OK;
    }
    for (int i : g.clips) {
        const asv_ctx* c = clips[(size_t)i];
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
        case 3:                                     // both first passes, into the buffer pass 2 reads
            BHIP(launch_seed_bricks(st, c->d, c->pos, c->rec, c->lab, c->labels[0], c->steps[0], c->steps[1], ratio,
                                    c->mult[1] * c->ratio, c->bricks, c->writes_label16(1, c->steps.size(), true)));
            break;
        case 4:                                     // done by pass 0
            break;
        default:
            BHIP(launch_jfa(st, c->d, in, out, c->lab, c->rec, g.K, ratio, c->p.tie_rule, c->unsafe, c->march_stats,
                            c->label_widths((size_t)g.p, c->steps.size(), true)));
            break;
        }
    }
    return ASV_OK;
}

// Seed stages, the JFA passes group by group, reductions and finalisation
// of iteration `iter` for every context, enqueued on st.
int enqueue_iteration(const std::vector<asv_ctx*>& clips, const std::vector<PassGroup>& groups,
                      const JfaClip* dev_desc, int iter, hipStream_t st)
{
    int rc;
    for (asv_ctx* c : clips)
        if ((rc = asv_seed_stage(c, iter, st))) return rc;
    for (const PassGroup& g : groups)
        if ((rc = launch_group(clips, dev_desc, g, st))) return rc;
    for (asv_ctx* c : clips) {
        c->fresh = false;
        c->cur = (int)(c->steps.size() & 1);
        if ((rc = asv_centroids(c, iter, st))) return rc;
    }
    return ASV_OK;
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
                        &g.dyn, g.lt);
    if (g.grid <= 0) return bfail(ASV_ERR_ARG, "no marching kernel for this step");
    BHIP(hipMemcpy(b->dev_desc + g.dev_off, g.desc.data(), sizeof(JfaClip) * g.desc.size(), hipMemcpyHostToDevice));
    return ASV_OK;
}

int launch_group(asv_batch* b, const PassGroup& g, hipStream_t st)
{
    return launch_group(b->clips, b->dev_desc, g, st);
}

int enqueue_iteration(asv_batch* b, int iter, hipStream_t st)
{
    return enqueue_iteration(b->clips, b->groups, b->dev_desc, iter, st);
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
            const int gmax = march_max_grid(g.K, g.dyn, g.ht, g.dense, g.lt);
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
    build_groups(b->clips, &b->groups, &b->n_desc);
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

// ---------------------------------------------------------------------------
// WindowBatcher: the Lloyd loops of contexts owned elsewhere (the served
// loop's streams) run together, with no autotuning and no graphs (the set
// of contexts changes from one call to the next).  Each context's own work
// -- prepare, seed stages, its own passes (those over the seeds'
// neighbourhoods), reduction and finalisation -- goes on its own stream, so
// that work of different contexts overlaps on the device as it would with
// one stream per context; the marching passes of all contexts sharing a
// pass index and step go as one launch on the batch stream.  Events join a
// context's stream and the batch stream where its work moves between them,
// so every context's work stays in order and ends on its own stream.  The
// marching descriptors go from a page-locked buffer to the device on the
// batch stream; the buffers form a ring of four, each reused only after
// the event recorded behind the last launch that read it.
// ---------------------------------------------------------------------------
struct asv::WindowBatcher::Impl {
    struct Slot {
        JfaClip* host = nullptr;      // page-locked staging of the descriptors
        JfaClip* dev = nullptr;
        size_t cap = 0;
        hipEvent_t done = nullptr;    // recorded after the call's last batched launch
        bool used = false;
    };
    static constexpr int kSlots = 4;
    Slot ring[kSlots];
    int next = 0;
    std::vector<hipEvent_t> ev;       // per context: its stream's work so far
    hipEvent_t ev_batch = nullptr;    // the batch stream's work so far

    ~Impl()
    {
        for (Slot& s : ring) {
            if (s.done) {
                if (s.used) (void)hipEventSynchronize(s.done);
                (void)hipEventDestroy(s.done);
            }
            if (s.host) (void)hipHostFree(s.host);
            if (s.dev) (void)hipFree(s.dev);
        }
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        if (ev_batch) (void)hipEventDestroy(ev_batch);
    }
};

asv::WindowBatcher::WindowBatcher() : impl_(new Impl) {}

asv::WindowBatcher::~WindowBatcher() { delete impl_; }

int asv::WindowBatcher::run(asv_ctx* const* ctxs, void* const* streams, int n, void* batch_stream)
{
    if (!ctxs || !streams || n < 1) return bfail(ASV_ERR_ARG, "no contexts to run");
    Impl& im = *impl_;
    hipStream_t bst = reinterpret_cast<hipStream_t>(batch_stream);
    std::vector<asv_ctx*> clips(ctxs, ctxs + n);
    std::vector<hipStream_t> own(n);
    for (int i = 0; i < n; ++i) own[(size_t)i] = reinterpret_cast<hipStream_t>(streams[i]);
    const int iters = clips[0]->p.lloyd_iters;
    for (const asv_ctx* c : clips)
        if (!c || c->p.lloyd_iters != iters)
            return bfail(ASV_ERR_ARG, "contexts of one batched run share the Lloyd iteration count");
    while ((int)im.ev.size() < n) {
        hipEvent_t e = nullptr;
        BHIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        im.ev.push_back(e);
    }
    if (!im.ev_batch) BHIP(hipEventCreateWithFlags(&im.ev_batch, hipEventDisableTiming));
    int rc;
    for (int i = 0; i < n; ++i)                   // LAB, as asv_run does
        if ((rc = asv_prepare(clips[(size_t)i], own[(size_t)i]))) return bfail(rc, "prepare", asv_last_error());
    std::vector<PassGroup> groups;
    size_t n_desc = 0;
    build_groups(clips, &groups, &n_desc);
    for (PassGroup& g : groups)
        if (g.batched) {
            g.grid = march_plan(g.K, g.desc.data(), (int)g.desc.size(), 0, 0, 0, &g.nitems, &g.dyn, g.lt);
            if (g.grid <= 0) g.split = true;      // each context's own launch
        }
    Impl::Slot& sl = im.ring[im.next];
    im.next = (im.next + 1) % Impl::kSlots;
    if (sl.used) BHIP(hipEventSynchronize(sl.done));
    sl.used = false;
    if (n_desc > sl.cap) {
        if (sl.host) (void)hipHostFree(sl.host);
        if (sl.dev) (void)hipFree(sl.dev);
        sl.host = nullptr;
        sl.dev = nullptr;
        sl.cap = 0;
        const size_t cap = 2 * n_desc;
        BHIP(hipHostMalloc((void**)&sl.host, sizeof(JfaClip) * cap, hipHostMallocDefault));
        BHIP(dev_malloc((void**)&sl.dev, sizeof(JfaClip) * cap));
        sl.cap = cap;
    }
    if (!sl.done) BHIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    size_t k = 0;                                 // descriptors in dev_off order
    for (const PassGroup& g : groups)
        for (const JfaClip& d : g.desc) sl.host[k++] = d;
    if (n_desc) BHIP(hipMemcpyAsync(sl.dev, sl.host, sizeof(JfaClip) * n_desc, hipMemcpyHostToDevice, bst));

    // where each context's latest work is (0 its stream, 1 the batch
    // stream), and whether the batch stream has work past ev_batch
    std::vector<char> on_batch((size_t)n, 0);
    bool batch_ahead = false;
    auto to_own = [&](int i) -> int {
        if (!on_batch[(size_t)i]) return ASV_OK;
        if (batch_ahead) {
            BHIP(hipEventRecord(im.ev_batch, bst));
            batch_ahead = false;
        }
        BHIP(hipStreamWaitEvent(own[(size_t)i], im.ev_batch, 0));
        on_batch[(size_t)i] = 0;
        return ASV_OK;
    };
    auto to_batch = [&](int i) -> int {
        if (on_batch[(size_t)i]) return ASV_OK;
        BHIP(hipEventRecord(im.ev[(size_t)i], own[(size_t)i]));
        BHIP(hipStreamWaitEvent(bst, im.ev[(size_t)i], 0));
        on_batch[(size_t)i] = 1;
        return ASV_OK;
    };
    for (int iter = 1; iter <= iters && !rc; ++iter) {
        for (int i = 0; i < n && !rc; ++i)
            if (!(rc = to_own(i))) rc = asv_seed_stage(clips[(size_t)i], iter, own[(size_t)i]);
        for (const PassGroup& g : groups) {
            if (rc) break;
            if (g.batched && !g.split) {
                for (int i : g.clips)
                    if (!rc) rc = to_batch(i);
                if (!rc) {
                    const hipError_t e = launch_march_batch(bst, g.K, sl.dev + g.dev_off, (int)g.desc.size(), g.grid,
                                                            g.nitems, g.dyn, g.ht, g.counted, g.dense, g.lt);
                    if (e != hipSuccess) rc = bfail(ASV_ERR_HIP, "marching launch", hipGetErrorString(e));
                    batch_ahead = true;
                }
            } else {
                for (int i : g.clips) {
                    if (rc) break;
                    PassGroup one = g;
                    one.clips.assign(1, i);
                    one.split = true;
                    if (!(rc = to_own(i))) rc = launch_group(clips, sl.dev, one, own[(size_t)i]);
                }
            }
        }
        for (int i = 0; i < n && !rc; ++i) {
            asv_ctx* c = clips[(size_t)i];
            if ((rc = to_own(i))) break;
            c->fresh = false;
            c->cur = (int)(c->steps.size() & 1);
            rc = asv_centroids(c, iter, own[(size_t)i]);
        }
    }
    BHIP(hipEventRecord(sl.done, bst));           // also behind a partial enqueue
    sl.used = true;
    for (int i = 0; i < n; ++i)                   // every context's work ends on its stream
        if (!rc) rc = to_own(i);
    return rc;
}
