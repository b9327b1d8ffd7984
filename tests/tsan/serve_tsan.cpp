// serve_tsan.cpp -- CPU harness for the served loop's host logic under
// ThreadSanitizer: the library's real asv_serve.cpp, asv_stream.cpp and
// asv_synth.cpp, linked against host stand-ins for the device-side entry
// points and the six HIP runtime calls they make.  The stand-ins keep the
// contracts the loop relies on (bounds and capacity checks, sums consistent
// with the labels, the stitched checksum) and sleep briefly so that worker
// threads interleave; every "device" buffer is host memory that only the
// context's owner may touch, so a context used by two threads at once, an
// idle context handed out twice, or a buffer freed while another thread
// copies into it shows up as a data race or a failed check.
//
// Test infrastructure: built and run by tests/test_serve_tsan.py, never part
// of the library.
#include "../../include/asv.h"
#include "../../anisotropic_supervoxel_amd/csrc/asv_host_internal.h"

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

std::atomic<int64_t> g_fail_from{0};   // stand-in allocations of at least this many bytes fail

void nap(int us) { std::this_thread::sleep_for(std::chrono::microseconds(us)); }

void must(bool ok, const char* what)
{
    if (!ok) {
        std::fprintf(stderr, "harness check failed: %s\n", what);
        std::abort();
    }
}

uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

}  // namespace

struct asv_ctx {
    static constexpr uint32_t kMagic = 0x5ca1ab1e;
    uint32_t magic = kMagic;
    asv_params p{};
    int depth_cap = 0, seed_cap = 0, D = 0, A = 0, s = 1;
    std::vector<int32_t> xyz;
    std::vector<uint8_t> frames;   // depth_cap frames of W x H x 3
    std::vector<int32_t> labels;   // depth_cap frames of W x H
    int64_t bytes = 0;
    bool ran = false;
};

static void check_ctx(const asv_ctx* c) { must(c && c->magic == asv_ctx::kMagic, "live context"); }

// ---- HIP runtime calls of the served loop --------------------------------
extern "C" {
hipError_t hipGetDevice(int* d)
{
    *d = 0;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipStreamCreate(hipStream_t* s)
{
    *s = reinterpret_cast<hipStream_t>(new int(0));
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s)
{
    delete reinterpret_cast<int*>(s);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned int)
{
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void* p)
{
    std::free(p);
    return hipSuccess;
}
// Events and stream waits: the stand-in "device" work runs synchronously on
// the calling thread, so recording and waiting are bookkeeping only.
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned)
{
    *e = reinterpret_cast<hipEvent_t>(new int(0));
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e)
{
    delete reinterpret_cast<int*>(e);
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t)
{
    must(e != nullptr, "event exists");
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t e, unsigned)
{
    must(e != nullptr, "event exists");
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }

// ---- device-side entry points ---------------------------------------------
void asv_default_params(asv_params* p, int32_t w, int32_t h, int32_t d, int32_t k)
{
    std::memset(p, 0, sizeof(*p));
    p->width = w;
    p->height = h;
    p->depth = d;
    p->seeds_num = k;
    p->lloyd_iters = 20;
    p->compactness = 8.0f;
    p->angle_gate = 1;
}

// a regular placement: seed i at (i mod W, (i / W) mod H, i mod D), step 2
int asv_plan(const asv_params* p, int32_t* seeds_placed, int32_t* grid_step, int32_t*, float*, int32_t)
{
    if (!p || p->width < 2 || p->height < 2 || p->depth < 1 || p->seeds_num < 1) return ASV_ERR_ARG;
    if (seeds_placed) *seeds_placed = p->seeds_num;
    if (grid_step) *grid_step = 2;
    return 6;
}

int asv_plan_seeds(const asv_params* p, int32_t* xyz, int32_t cap)
{
    const int n = p->seeds_num;
    if (cap < 3 * n) return ASV_ERR_ARG;
    for (int i = 0; i < n; ++i) {
        xyz[3 * i] = i % p->width;
        xyz[3 * i + 1] = (i / p->width) % p->height;
        xyz[3 * i + 2] = i % p->depth;
    }
    return n;
}

int asv_create_seeded(const asv_params* p, const int32_t* xyz, int32_t n, int32_t grid_step, asv_ctx** out)
{
    if (!p || !xyz || n < 1 || !out) return ASV_ERR_ARG;
    const int64_t V = (int64_t)p->width * p->height * p->depth;
    const int cap = n > p->seeds_num ? n : p->seeds_num;
    const int64_t bytes = V * 31 + (int64_t)cap * 240;
    const int64_t from = g_fail_from.load();
    if (from > 0 && bytes >= from) return ASV_ERR_NOMEM;
    asv_ctx* c = new asv_ctx();
    c->p = *p;
    c->depth_cap = p->depth;
    c->seed_cap = cap;
    c->D = p->depth;
    c->A = n;
    c->s = grid_step;
    c->xyz.assign(xyz, xyz + 3 * (size_t)n);
    c->frames.assign((size_t)V * 3, 0);
    c->labels.assign((size_t)V, -1);
    c->bytes = bytes;
    *out = c;
    return ASV_OK;
}

int asv_destroy(asv_ctx* c)
{
    if (!c) return ASV_OK;
    check_ctx(c);
    c->magic = 0;
    delete c;
    return ASV_OK;
}

int64_t asv_device_bytes(const asv_ctx* c) { return c ? c->bytes : -1; }

int asv_reseed(asv_ctx* c, const int32_t* xyz, int32_t n, int32_t grid_step, int32_t depth, void*)
{
    check_ctx(c);
    if (n < 1 || n > c->seed_cap || depth < 1 || depth > c->depth_cap || grid_step < 1) return ASV_ERR_ARG;
    for (int i = 0; i < n; ++i)
        must(xyz[3 * i] >= 0 && xyz[3 * i] < c->p.width && xyz[3 * i + 1] >= 0 && xyz[3 * i + 1] < c->p.height &&
                 xyz[3 * i + 2] >= 0 && xyz[3 * i + 2] < depth,
             "seed inside the window");
    c->xyz.assign(xyz, xyz + 3 * (size_t)n);
    c->A = n;
    c->D = depth;
    c->s = grid_step;
    c->ran = false;
    return ASV_OK;
}

int asv_shift_frames(asv_ctx* c, int32_t src, int32_t n, void*)
{
    check_ctx(c);
    if (src < 0 || n < 0 || src + n > c->depth_cap || n > src) return ASV_ERR_ARG;
    const size_t plane = (size_t)c->p.width * c->p.height * 3;
    std::memmove(c->frames.data(), c->frames.data() + (size_t)src * plane, (size_t)n * plane);
    return ASV_OK;
}

int asv_upload_frames_range(asv_ctx* c, int32_t first, int32_t count, const uint8_t* rgb, const float* fwd,
                            const float* bwd, void*)
{
    check_ctx(c);
    if (!rgb || !fwd || !bwd || first < 0 || count < 0 || first + count > c->depth_cap) return ASV_ERR_ARG;
    const size_t plane = (size_t)c->p.width * c->p.height;
    std::memcpy(c->frames.data() + (size_t)first * plane * 3, rgb, (size_t)count * plane * 3);
    float acc = 0.0f;                     // the flows are read too
    for (size_t i = 0; i < (size_t)count * plane * 2; i += 97) acc += fwd[i] - bwd[i];
    (void)acc;
    nap(50);
    return ASV_OK;
}

// "Lloyd loop": voxel (x, y, z) takes the seed with the nearest x (ties to
// the lower id), weighted by the frame's first colour byte so the labels
// depend on the uploaded frames.
int asv_run(asv_ctx* c, void*)
{
    check_ctx(c);
    const int W = c->p.width, H = c->p.height;
    for (int z = 0; z < c->D; ++z)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const size_t v = ((size_t)z * H + y) * W + x;
                const int shift = c->frames[v * 3] & 1;
                int best = 0, bd = 1 << 30;
                for (int i = 0; i < c->A; ++i) {
                    const int d = std::abs(c->xyz[3 * i] - x) * 2 + std::abs(c->xyz[3 * i + 2] - z) + shift * (i & 1);
                    if (d < bd) { bd = d; best = i; }
                }
                c->labels[v] = best;
            }
    c->ran = true;
    nap(200);
    return ASV_OK;
}

int asv_download_sums(asv_ctx* c, int64_t* sums, void*)
{
    check_ctx(c);
    if (!c->ran) return ASV_ERR_STATE;
    const int W = c->p.width, H = c->p.height;
    std::memset(sums, 0, sizeof(int64_t) * 8 * (size_t)(c->A + 1));
    for (int z = 0; z < c->D; ++z)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const int l = c->labels[((size_t)z * H + y) * W + x];
                int64_t* q = sums + 8 * (size_t)(l < 0 ? c->A : l);
                q[0] += 2 * x + 1; q[1] += 2 * y + 1; q[2] += 2 * z + 1; q[3] += 1;
            }
    return ASV_OK;
}

int asv_postprocess_device(asv_ctx* c, const int64_t* sums, int32_t* dead, int32_t* nfin, void*)
{
    check_ctx(c);
    int d = 0;
    for (int i = 0; i < c->A; ++i) d += sums[8 * (size_t)i + 3] == 0;
    if (dead) *dead = d;
    if (nfin) *nfin = c->A - d;
    return ASV_OK;
}

int asv_download_labels_mapped(asv_ctx* c, const int32_t* table, int32_t n_table, int32_t n_frames, int32_t* labels,
                               int64_t index0, uint64_t* checksum, void*)
{
    check_ctx(c);
    if (!table || !labels || n_table < 1 || n_frames < 0 || n_frames > c->D) return ASV_ERR_ARG;
    const int64_t n = (int64_t)n_frames * c->p.width * c->p.height;
    uint64_t ck = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int l = c->labels[(size_t)i];
        const int32_t g = l >= 0 && l < n_table ? table[l] : 0;
        labels[i] = g;
        ck += splitmix64(((uint64_t)(index0 + i) << 32) | (uint32_t)g);
    }
    if (checksum) *checksum = ck;
    nap(30);
    return ASV_OK;
}

}  // extern "C"

// The batched Lloyd loops: each context's stand-in run, on the batcher's
// thread (a context handed over while its worker still touched it would
// race here).
struct asv::WindowBatcher::Impl {
    int calls = 0;
};
asv::WindowBatcher::WindowBatcher() : impl_(new Impl) {}
asv::WindowBatcher::~WindowBatcher() { delete impl_; }
int asv::WindowBatcher::run(asv_ctx* const* ctxs, void* const* streams, int n, void*)
{
    ++impl_->calls;
    for (int i = 0; i < n; ++i) {
        const int rc = asv_run(ctxs[i], streams[i]);
        if (rc) return rc;
    }
    return ASV_OK;
}

bool asv::ctx_reusable(const asv_ctx* c, const asv_params* p, int32_t seeds)
{
    check_ctx(c);
    return c->p.width == p->width && c->p.height == p->height && c->depth_cap >= p->depth && c->seed_cap >= seeds &&
           c->p.lloyd_iters == p->lloyd_iters && c->p.compactness == p->compactness;
}

int32_t asv::ctx_seed_cap(const asv_ctx* c)
{
    check_ctx(c);
    return c->seed_cap;
}

int64_t asv::seed_brick_bytes(int W, int H, int D, int64_t seeds, int)
{
    return (int64_t)((W + 31) / 32) * ((H + 15) / 16) * ((D + 15) / 16) * 12 + 4 + seeds * 48 * 64;
}

// ---- the harness ----------------------------------------------------------
namespace {

std::vector<asv_serve_request> requests(std::vector<int32_t>& cancel_flags)
{
    const int spec[][4] = {{24, 20, 33, 15}, {32, 24, 21, 20}, {24, 16, 47, 30}, {40, 24, 26, 10}, {16, 16, 40, 40},
                           {24, 24, 15, 25}, {32, 24, 28, 20}, {24, 20, 19, 15}, {40, 32, 36, 12}, {24, 20, 50, 30},
                           {32, 24, 22, 20}, {24, 16, 31, 25}};
    std::vector<asv_serve_request> r;
    cancel_flags.assign(16, 0);
    for (int i = 0; i < 12; ++i) {
        asv_serve_request q;
        std::memset(&q, 0, sizeof(q));
        q.stream_id = i;
        q.width = spec[i][0];
        q.height = spec[i][1];
        q.frames = spec[i][2];
        q.density = (float)spec[i][3] / 10.0f;
        q.n_ellipses = 3;
        q.video_seed = 500 + (uint64_t)i;
        q.vmax = 2.0f;
        q.noise_sigma = 4.0f;
        q.arrival_ms = 0.5 * i;
        q.cancel = &cancel_flags[(size_t)i];
        r.push_back(q);
    }
    return r;
}

int g_batch = 1;                                  // batch_windows of the runs below

int run(const std::vector<asv_serve_request>& reqs, int max_active, int64_t budget, std::vector<asv_serve_result>& res)
{
    asv_serve_params sp;
    asv_serve_default_params(&sp);
    sp.batch_windows = g_batch;
    sp.max_active = max_active;
    sp.memory_budget = budget;
    sp.honor_arrivals = 1;
    res.assign(reqs.size(), asv_serve_result{});
    std::vector<asv_serve_window> w(4096);
    int32_t nw = 0;
    double wall = 0.0;
    return asv_serve_run(&sp, reqs.data(), (int32_t)reqs.size(), res.data(), w.data(), (int32_t)w.size(), &nw, &wall);
}

}  // namespace

int main()
{
    std::vector<int32_t> flags;
    std::vector<asv_serve_request> reqs = requests(flags);
    // every stream alone, one at a time: the reference checksums
    std::vector<uint64_t> alone(reqs.size());
    for (size_t i = 0; i < reqs.size(); ++i) {
        std::vector<asv_serve_result> r1;
        must(run({reqs[i]}, 1, (int64_t)1 << 30, r1) == ASV_OK, "alone run");
        must(r1[0].status == ASV_OK, "alone status");
        alone[i] = r1[0].labels_hash;
    }
    // the mix: four workers, a budget of about three streams (idle contexts
    // evicted and adopted), one malformed request, one cancelled after two
    // windows, one cancelled by its flag before it starts, and one whose
    // context allocation fails
    int64_t fp_max = 0;
    asv_serve_params sp;
    asv_serve_default_params(&sp);
    for (const auto& q : reqs) fp_max = std::max(fp_max, asv_serve_footprint(&sp, &q));
    std::vector<asv_serve_request> mix = reqs;
    mix[3].width = 1;                                   // malformed
    mix[9].cancel_after_windows = 2;                   // 50 frames: nine windows
    flags[7] = 1;                                       // cancelled before its first window
    mix[8].width = 64;                                  // the only context of at least 64 x 32 x 10 voxels
    mix[8].height = 32;
    g_fail_from.store((int64_t)64 * 32 * 10 * 31);
    for (int rep = 0; rep < 4; ++rep) {
        g_batch = rep & 1;                              // both forms of the loop
        std::vector<asv_serve_result> res;
        must(run(mix, 4, 3 * fp_max, res) == ASV_OK, "mix run");
        for (size_t i = 0; i < mix.size(); ++i) {
            const int st = res[i].status;
            if (i == 3) must(st == ASV_ERR_ARG, "malformed request refused as such");
            else if (i == 9 || i == 7) must(st == ASV_ERR_CANCELLED, "cancelled request");
            else if (i == 8) must(st == ASV_ERR_NOMEM, "failed allocation ends its own request");
            else {
                must(st == ASV_OK, "other requests complete");
                must(res[i].labels_hash == alone[i], "checksum in the mix equals alone");
            }
        }
    }
    g_fail_from.store(0);
    std::printf("serve harness ok\n");
    return 0;
}
