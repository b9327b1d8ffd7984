// asv_serve.cpp -- the served streaming form: many concurrent video
// streams, each processed as sliding windows (a temporal-block run with
// seeds carried across windows, as the streamGPU-AS program does per block,
// streamGPU-AS/Render.cpp:249-277), with continuous batching on one GPU.
//
// A dispatcher admits requests in arrival order as a stream slot and device
// memory (the stream's resident window footprint) free up; a request that
// does not fit waits while later, smaller ones may go first.  Each admitted
// stream runs on a worker thread with its own HIP stream.  With
// batch_windows, a worker sets up its window (seeds, frames) on its stream
// and hands the window's context and stream to a batcher thread, which
// enqueues the Lloyd loops of every window handed over since its last call
// together: each window's own kernels on the window's stream, the marching
// passes of all of them as one launch per pass index and step on the
// batcher's stream, joined by events; the worker then reads back on its
// stream.  Windows that become ready while a batch is on the device join
// the next one, and the workers' set-up and read-back overlap it.
// Without it each worker runs its window's loop on its own stream.  Per
// window the loop records when it became ready (frames present and the
// previous window done), when its GPU work was enqueued and when its labels
// were back.
#include "../../include/asv.h"
#include "asv_host_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

namespace {

using clock_t_ = std::chrono::steady_clock;

double ms_since(clock_t_::time_point t0)
{
    return std::chrono::duration<double, std::milli>(clock_t_::now() - t0).count();
}

// A request the loop can run at all: sizes, a finite positive density, and
// either all three host arrays or none.
bool well_formed(const asv_serve_request* r)
{
    const int have = (r->rgb != nullptr) + (r->fwd != nullptr) + (r->bwd != nullptr);
    return r->width >= 2 && r->height >= 2 && r->frames >= 1 && r->density > 0.0f &&
           r->density < 1e6f && r->n_ellipses >= 0 && (have == 0 || have == 3);
}

asv_stream_params stream_params_of(const asv_serve_params* sp, const asv_serve_request* r)
{
    asv_stream_params p;
    asv_stream_default_params(&p, r->width, r->height, r->frames, 1);
    // density d supervoxels per 1000 pixels per frame = d*W*H/10 per 100 frames
    p.seeds_per_100_frames = (int32_t)((double)r->density * r->width * r->height / 10.0);
    if (p.seeds_per_100_frames < 1) p.seeds_per_100_frames = 1;
    p.block_size = sp->block_size;
    p.overlap = sp->overlap;
    p.lloyd_iters = sp->lloyd_iters;
    p.compactness = sp->compactness;
    p.angle_gate = sp->angle_gate;
    return p;
}

}  // namespace

extern "C" {

void asv_serve_default_params(asv_serve_params* p)
{
    std::memset(p, 0, sizeof(*p));
    p->memory_budget = (int64_t)64 << 30;
    p->max_active = 8;
    p->block_size = 10;
    p->overlap = 5;
    p->lloyd_iters = 5;
    p->compactness = 6.0f;
    p->angle_gate = 0;
    p->honor_arrivals = 1;
    p->batch_windows = 0;      // measured slower than per-stream launches on the served workload
}

int64_t asv_serve_footprint(const asv_serve_params* sp, const asv_serve_request* r)
{
    if (!sp || !r || !well_formed(r)) return -1;
    asv_stream_params p = stream_params_of(sp, r);
    int32_t iv[2 * 1024];
    int nb = asv_stream_plan(&p, iv, 1024);
    if (nb < 1) return -1;
    int depth = 0;
    for (int b = 0; b < nb && b < 1024; ++b) depth = std::max(depth, iv[2 * b + 1] - iv[2 * b]);
    const int64_t V = (int64_t)r->width * r->height * depth;
    const int64_t seeds = 4 * ((int64_t)p.seeds_per_100_frames * p.block_size / 100) + 64;
    // frames, flows, LAB, two label volumes; seeds; the brick lists of the
    // fused first passes where the window's seeds are sparse enough for them
    const int64_t bricks = seeds * 729 * 3 <= V ? asv::seed_brick_bytes(r->width, r->height, depth, seeds, 2 + 16) : 0;
    return V * (3 + 16 + 4 + 8) + seeds * 240 + bricks + (1 << 20);
}

int asv_serve_run(const asv_serve_params* sp, const asv_serve_request* reqs, int32_t n, asv_serve_result* results,
                  asv_serve_window* windows, int32_t window_cap, int32_t* n_windows, double* wall_ms)
{
    if (!sp || (!reqs && n > 0) || !results || sp->max_active < 1 || sp->memory_budget <= 0) return ASV_ERR_ARG;
    int dev = 0;
    (void)hipGetDevice(&dev);     // without a device every admitted stream fails with a HIP error

    std::vector<int> order((size_t)n);
    for (int i = 0; i < n; ++i) order[(size_t)i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return reqs[a].arrival_ms < reqs[b].arrival_ms; });
    for (int i = 0; i < n; ++i) {
        std::memset(&results[i], 0, sizeof(results[i]));
        results[i].stream_id = reqs[i].stream_id;
        results[i].status = ASV_ERR_STATE;
    }

    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> work;            // admitted request indices
    int active = 0;
    int64_t active_bytes = 0;
    // Idle resident contexts left by finished streams, oldest first.  They
    // count against the memory budget and are freed when admission needs
    // the room or more than 4 x max_active of them wait (hipFree waits for
    // the whole device, so the loop frees as little as it can).
    struct Idle {
        asv_ctx* c;
        int64_t bytes;
    };
    std::deque<Idle> idle;
    int64_t idle_bytes = 0;
    bool closing = false;
    int nwin = 0;
    const auto t0 = clock_t_::now();

    auto record = [&](const asv_serve_window& w) {
        std::lock_guard<std::mutex> g(mu);
        if (windows && nwin < window_cap) windows[nwin] = w;
        ++nwin;
    };

    // A new stream adopts an idle context of its frame size and settings
    // when one is there (no device allocation on admission).
    // Each worker owns page-locked buffers sized once for the largest window
    // of the run: the window's finished labels (one transfer out) and the
    // generated frames and flows of a window (direct transfers in).  No
    // worker allocates or frees page-locked memory while others copy.
    size_t max_window = 0;
    for (int i = 0; i < n; ++i) {
        if (!well_formed(&reqs[i])) continue;
        const asv_stream_params q = stream_params_of(sp, &reqs[i]);
        const int nbq = asv_stream_plan(&q, nullptr, 0);
        if (nbq < 1) continue;
        std::vector<int32_t> ivq((size_t)nbq * 2);
        asv_stream_plan(&q, ivq.data(), nbq);
        for (int b = 0; b < nbq; ++b)
            max_window = std::max(max_window, (size_t)(ivq[2 * (size_t)b + 1] - ivq[2 * (size_t)b]) *
                                                  (size_t)reqs[i].width * (size_t)reqs[i].height);
    }
    struct WindowOut {
        int32_t* p = nullptr;
        uint8_t* rgb = nullptr;
        float* fwd = nullptr;
        float* bwd = nullptr;
        size_t cap = 0;
        bool ok = false;
        void alloc(size_t voxels)
        {
            cap = voxels;
            ok = hipHostMalloc((void**)&p, voxels * sizeof(int32_t), hipHostMallocDefault) == hipSuccess &&
                 hipHostMalloc((void**)&rgb, voxels * 3, hipHostMallocDefault) == hipSuccess &&
                 hipHostMalloc((void**)&fwd, voxels * 2 * sizeof(float), hipHostMallocDefault) == hipSuccess &&
                 hipHostMalloc((void**)&bwd, voxels * 2 * sizeof(float), hipHostMallocDefault) == hipSuccess;
        }
        void release()
        {
            void* ptrs[] = {p, rgb, fwd, bwd};
            for (void* q : ptrs)
                if (q) (void)hipHostFree(q);
            p = nullptr;
            rgb = nullptr;
            fwd = bwd = nullptr;
            cap = 0;
            ok = false;
        }
    };

    // Windows handed to the batcher: the context whose loop to run, the
    // worker's stream, and the outcome of the enqueue.
    struct Ticket {
        asv_ctx* ctx;
        hipStream_t st;
        int rc;
        bool finished;
    };
    std::mutex bmu;
    std::condition_variable bcv;
    std::deque<Ticket*> bqueue;
    bool bstop = false;
    auto batcher = [&]() {
        (void)hipSetDevice(dev);
        hipStream_t bst = nullptr;
        const hipError_t se = hipStreamCreate(&bst);
        asv::WindowBatcher wb;
        std::vector<Ticket*> take;
        std::vector<asv_ctx*> ctxs;
        std::vector<void*> streams;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(bmu);
                bcv.wait(lk, [&] { return bstop || !bqueue.empty(); });
                if (bqueue.empty()) break;
                take.assign(bqueue.begin(), bqueue.end());
                bqueue.clear();
            }
            int rc = se == hipSuccess ? ASV_OK : ASV_ERR_HIP;
            ctxs.clear();
            streams.clear();
            for (Ticket* t : take) {
                ctxs.push_back(t->ctx);
                streams.push_back(t->st);
            }
            if (!rc) rc = wb.run(ctxs.data(), streams.data(), (int)ctxs.size(), bst);
            {
                std::lock_guard<std::mutex> g(bmu);
                for (Ticket* t : take) {
                    t->rc = rc;
                    t->finished = true;
                }
            }
            bcv.notify_all();
        }
        if (bst) {
            (void)hipStreamSynchronize(bst);
            (void)hipStreamDestroy(bst);
        }
    };
    // One window's Lloyd loop through the batcher, enqueued (behind the
    // window's set-up) on the worker's stream and the batcher's.
    auto run_batched = [&](asv_ctx* c, hipStream_t st) -> int {
        Ticket t{c, st, ASV_OK, false};
        {
            std::lock_guard<std::mutex> g(bmu);
            bqueue.push_back(&t);
        }
        bcv.notify_all();
        {
            std::unique_lock<std::mutex> lk(bmu);
            bcv.wait(lk, [&] { return t.finished; });
        }
        return t.rc;
    };
    auto process = [&](int ri, hipStream_t st, WindowOut& wo) -> int {
        const asv_serve_request& r = reqs[ri];
        asv_stream_params p = stream_params_of(sp, &r);
        asv_stream* s = nullptr;
        int rc = asv_stream_create(&p, &s);
        if (rc) return rc;
        {
            std::lock_guard<std::mutex> g(mu);
            for (auto it = idle.begin(); it != idle.end(); ++it)
                if (asv::stream_adopt_ctx(s, it->c)) {
                    idle_bytes -= it->bytes;
                    idle.erase(it);
                    break;
                }
        }
        const int nb = asv_stream_block_count(s);
        std::vector<int32_t> iv((size_t)nb * 2);
        asv_stream_plan(&p, iv.data(), nb);
        const size_t plane = (size_t)r.width * r.height;
        size_t max_nf = 0;
        for (int b = 0; b < nb; ++b) max_nf = std::max(max_nf, (size_t)(iv[2 * (size_t)b + 1] - iv[2 * (size_t)b]));
        if (!wo.ok || max_nf * plane > wo.cap) {
            asv_stream_destroy(s);
            return ASV_ERR_NOMEM;
        }
        uint64_t vox_iters = 0;
        double ready = ms_since(t0);
        for (int b = 0; b < nb && !rc; ++b) {
            // a cancelled request stops between windows; its context stays
            // consistent (the last window completed) and goes back to the
            // idle pool like a finished stream's
            if ((r.cancel && *r.cancel) || (r.cancel_after_windows > 0 && b >= r.cancel_after_windows)) {
                rc = ASV_ERR_CANCELLED;
                break;
            }
            const int f0 = iv[2 * (size_t)b], f1 = iv[2 * (size_t)b + 1];
            const size_t nf = (size_t)(f1 - f0);
            const uint8_t* rgb;
            const float *fwd, *bwd;
            if (r.rgb && r.fwd && r.bwd) {
                rgb = r.rgb + (size_t)f0 * plane * 3;
                fwd = r.fwd + (size_t)f0 * plane * 2;
                bwd = r.bwd + (size_t)f0 * plane * 2;
            } else {
                rc = asv_synth_video_range(r.video_seed, r.width, r.height, r.frames, r.n_ellipses, r.vmax,
                                           r.noise_sigma, f0, f1, wo.rgb, wo.fwd, wo.bwd, nullptr);
                if (rc) break;
                rgb = wo.rgb;
                fwd = wo.fwd;
                bwd = wo.bwd;
            }
            // the window is ready once its frames are present and the previous window is done
            ready = std::max(ready, ms_since(t0));
            asv_serve_window w;
            w.stream_id = r.stream_id;
            w.window = b;
            w.frames = (int32_t)nf;
            w.ready_ms = ready;
            w.start_ms = ms_since(t0);
            int32_t done = 0;
            if (sp->batch_windows) {
                rc = asv::stream_block_begin(s, b, rgb, fwd, bwd, st);
                if (!rc) rc = run_batched(asv::stream_ctx(s), st);
                if (!rc) rc = asv::stream_block_end(s, wo.p, &done, st);
            } else {
                rc = asv_stream_block_into(s, b, rgb, fwd, bwd, wo.p, &done, st);
            }
            w.end_ms = ms_since(t0);
            int32_t info[6];
            asv_stream_block_info(s, info);
            w.seeds = info[1];
            record(w);
            ready = w.end_ms;
            vox_iters += (uint64_t)plane * nf * (uint64_t)p.lloyd_iters;
        }
        if (!rc) {
            uint64_t ck = 0;
            asv_stream_checksum(s, &ck);
            results[ri].labels_hash = ck;
            results[ri].voxel_iters = (int64_t)vox_iters;
        }
        // after an error the context may hold a partial window: it is freed
        // with the stream instead of being offered to later streams
        asv_ctx* left = (rc == ASV_OK || rc == ASV_ERR_CANCELLED) ? asv::stream_take_ctx(s) : nullptr;
        asv_stream_destroy(s);
        if (left) {
            asv_ctx* evict = nullptr;
            {
                std::lock_guard<std::mutex> g(mu);
                idle.push_back(Idle{left, asv_device_bytes(left)});
                idle_bytes += idle.back().bytes;
                if ((int)idle.size() > 4 * sp->max_active) {
                    evict = idle.front().c;
                    idle_bytes -= idle.front().bytes;
                    idle.pop_front();
                }
            }
            if (evict) asv_destroy(evict);
        }
        return rc;
    };

    auto worker = [&]() {
        (void)hipSetDevice(dev);
        hipStream_t st = nullptr;
        (void)hipStreamCreate(&st);
        WindowOut wo;
        wo.alloc(max_window);
        for (;;) {
            int ri;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return closing || !work.empty(); });
                if (work.empty()) break;
                ri = work.front();
                work.pop_front();
            }
            int rc = process(ri, st, wo);
            {
                std::lock_guard<std::mutex> g(mu);
                results[ri].status = rc;
                results[ri].done_ms = ms_since(t0);
                --active;
                active_bytes -= results[ri].footprint;
            }
            cv.notify_all();
        }
        wo.release();
        if (st) (void)hipStreamDestroy(st);
    };

    std::thread batch_thread;
    if (sp->batch_windows) batch_thread = std::thread(batcher);
    std::vector<std::thread> pool;
    for (int i = 0; i < sp->max_active; ++i) pool.emplace_back(worker);

    std::deque<int> pending(order.begin(), order.end());
    {
        std::unique_lock<std::mutex> lk(mu);
        while (!pending.empty()) {
            const double now = ms_since(t0);
            double next_arrival = -1.0;
            for (auto it = pending.begin(); it != pending.end();) {
                const int ri = *it;
                const asv_serve_request& r = reqs[ri];
                if (sp->honor_arrivals && r.arrival_ms > now) {
                    if (next_arrival < 0.0 || r.arrival_ms < next_arrival) next_arrival = r.arrival_ms;
                    ++it;
                    continue;
                }
                const int64_t fp = asv_serve_footprint(sp, &r);
                if (fp < 0 || fp > sp->memory_budget) {          // malformed, or can never be admitted
                    results[ri].status = fp < 0 ? ASV_ERR_ARG : ASV_ERR_ADMIT;
                    results[ri].footprint = fp;
                    it = pending.erase(it);
                    continue;
                }
                if (active < sp->max_active && active_bytes + fp <= sp->memory_budget &&
                    active_bytes + idle_bytes + fp > sp->memory_budget && !idle.empty()) {
                    // room is held by idle contexts: free the oldest, then rescan
                    asv_ctx* evict = idle.front().c;
                    idle_bytes -= idle.front().bytes;
                    idle.pop_front();
                    lk.unlock();
                    asv_destroy(evict);
                    lk.lock();
                    it = pending.begin();
                    continue;
                }
                if (active < sp->max_active && active_bytes + idle_bytes + fp <= sp->memory_budget) {
                    ++active;
                    active_bytes += fp;
                    results[ri].footprint = fp;
                    results[ri].admitted_ms = now;
                    work.push_back(ri);
                    it = pending.erase(it);
                    cv.notify_all();
                    continue;
                }
                ++it;
            }
            if (pending.empty()) break;
            if (next_arrival >= 0.0) {
                const double wait = std::max(0.0, next_arrival - ms_since(t0));
                cv.wait_for(lk, std::chrono::microseconds((int64_t)(wait * 1000.0) + 1));
            } else {
                cv.wait(lk);
            }
        }
        closing = true;
    }
    cv.notify_all();
    for (auto& th : pool) th.join();
    if (batch_thread.joinable()) {
        {
            std::lock_guard<std::mutex> g(bmu);
            bstop = true;
        }
        bcv.notify_all();
        batch_thread.join();
    }
    for (const Idle& e : idle) asv_destroy(e.c);
    if (n_windows) *n_windows = nwin;
    if (wall_ms) *wall_ms = ms_since(t0);
    // each request's outcome is in its result; one request's failure ends
    // only that request
    return ASV_OK;
}

}  // extern "C"
