// asv_stream.cpp -- temporal-block (streaming) form of the run, the
// reference's streamGPU-AS program (streamGPU-AS/Render.cpp:100-569): the
// video is cut into overlapping blocks, each block is a plain run on the
// GPU from kept seeds plus new grid seeds, and labels are stitched through
// per-block maps from block seed ids to global ids.  Host logic over the C
// ABI; the device work of a block is asv_create_seeded + asv_run.
#include "../../include/asv.h"
#include "asv_host_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

thread_local std::string g_stream_err;

int sfail(int code, const char* msg)
{
    g_stream_err = msg;
    return code;
}

// Block intervals, streamGPU-AS/Render.cpp:198-241 with block_size B,
// overlap O (one_third) and stride S = B - O (two_third).
std::vector<int> plan_blocks(int D, int B, int O)
{
    std::vector<int> iv;
    const int S = B - O;
    int last_end = B;
    if (D < B + O) {
        iv.push_back(0);
        iv.push_back(D);
        return iv;
    }
    iv.push_back(0);
    iv.push_back(last_end);
    bool done = false;
    while (last_end + S < D) {
        if (D - last_end - S < O) {
            iv.push_back(last_end - O);
            iv.push_back(D);
            done = true;
        } else {
            iv.push_back(last_end - O);
            iv.push_back(last_end + S);
        }
        last_end += S;
    }
    if (!done) {
        iv.push_back(last_end - O);
        iv.push_back(D);
    }
    return iv;
}

// Grid placement of Image::initialization_grid (streamGPU-AS/Image.cpp:26-119),
// the same rule as the plain init_seeds: returns A (and s) or -1.
int grid_place(int W, int H, int D, int K, std::vector<int>* xyz, int* s_out)
{
    asv_params p;
    asv_default_params(&p, W, H, D, K);
    int32_t A = 0, s = 0;
    if (asv_plan(&p, &A, &s, nullptr, nullptr, 0) < 0) return -1;
    xyz->assign((size_t)A * 3, 0);
    if (asv_plan_seeds(&p, xyz->data(), A * 3) != A) return -1;
    if (s_out) *s_out = s;
    return A;
}

struct KeptSeed {
    float x, y, z;   // centre coordinates (voxel + .5), z already shifted
    int old_id;      // id in the block it comes from
};

}  // namespace

struct asv_stream {
    asv_stream_params p{};
    std::vector<int> intervals;
    int seeds100 = 0;
    int placed = 0;                       // seeds_num_placed
    int next_block = 0;
    std::vector<KeptSeed> kept;
    std::vector<int> prev_map;            // block2all of the previous block (dense: ids 0..A-1)
    int last_A = 0, last_dead = 0, last_budget = 0;
    // resident device window: one context reused across blocks
    asv_ctx* ctx = nullptr;
    int depth_cap = 0, seed_cap = 0, prev_depth = 0;
    uint64_t checksum = 0;                // of the frames finished so far
    // the block set up by stream_block_begin and not yet finished
    int pending = -1;
    int pending_A = 0, pending_budget = 0, pending_nkept = 0;
    std::vector<int> pending_map;         // its block2all
    ~asv_stream()
    {
        if (ctx) asv_destroy(ctx);
    }
};

namespace {

// Parameters of the stream's resident context (all blocks share it).
asv_params ctx_params(const asv_stream* s, int seeds)
{
    asv_params cp;
    asv_default_params(&cp, s->p.width, s->p.height, s->depth_cap, seeds);
    cp.lloyd_iters = s->p.lloyd_iters;
    cp.compactness = s->p.compactness;
    cp.angle_gate = s->p.angle_gate;
    cp.tie_rule = s->p.tie_rule;
    cp.frames_prepared = s->p.frames_prepared;
    return cp;
}

// Seed capacity a new resident context gets: four blocks' budget.
int wanted_seed_cap(const asv_stream* s) { return 4 * (s->seeds100 * s->p.block_size / 100) + 64; }

}  // namespace

asv_ctx* asv::stream_take_ctx(asv_stream* s)
{
    if (!s) return nullptr;
    asv_ctx* c = s->ctx;
    s->ctx = nullptr;
    s->seed_cap = 0;
    return c;
}

bool asv::stream_adopt_ctx(asv_stream* s, asv_ctx* c)
{
    if (!c || !s || s->ctx || s->next_block != 0) return false;
    const asv_params cp = ctx_params(s, wanted_seed_cap(s));
    if (!asv::ctx_reusable(c, &cp, wanted_seed_cap(s))) return false;
    s->ctx = c;
    s->seed_cap = asv::ctx_seed_cap(c);
    return true;
}

extern "C" {

void asv_stream_default_params(asv_stream_params* p, int32_t width, int32_t height, int32_t total_frames,
                               int32_t seeds_num)
{
    std::memset(p, 0, sizeof(*p));
    p->width = width;
    p->height = height;
    p->total_frames = total_frames;
    p->seeds_num = seeds_num;
    p->seeds_per_100_frames = 0;     // derived: seeds_num * 100 / total_frames
    p->block_size = 30;              // streamGPU-AS/Render.cpp:199
    p->overlap = 10;                 // one_third, :200
    p->lloyd_iters = 20;             // :252
    p->compactness = 6.0f;           // streamGPU-AS/Image.cpp:968
    p->angle_gate = 0;               // streamGPU-AS/Image.cpp:1384-1400 has no gate
    p->tie_rule = ASV_TIE_LOWEST_ID;
    p->frames_prepared = 0;
}

int asv_stream_plan(const asv_stream_params* p, int32_t* intervals, int32_t cap)
{
    if (!p || p->total_frames < 1 || p->block_size < 2 || p->overlap < 0 || p->overlap >= p->block_size)
        return sfail(ASV_ERR_ARG, "invalid stream parameters");
    std::vector<int> iv = plan_blocks(p->total_frames, p->block_size, p->overlap);
    int n = (int)iv.size() / 2;
    for (int i = 0; i < 2 * n && i < 2 * cap; ++i) intervals[i] = iv[i];
    return n;
}

int asv_stream_create(const asv_stream_params* p, asv_stream** out)
{
    if (!out) return sfail(ASV_ERR_ARG, "null output pointer");
    *out = nullptr;
    int32_t dummy[2];
    if (asv_stream_plan(p, dummy, 0) < 0) return ASV_ERR_ARG;
    if (p->width < 2 || p->height < 2 || p->lloyd_iters < 1 || !(p->compactness > 0.0f))
        return sfail(ASV_ERR_ARG, "invalid stream parameters");
    asv_stream* s = new asv_stream();
    s->p = *p;
    s->intervals = plan_blocks(p->total_frames, p->block_size, p->overlap);
    for (size_t i = 0; i + 1 < s->intervals.size(); i += 2)
        s->depth_cap = std::max(s->depth_cap, s->intervals[i + 1] - s->intervals[i]);
    // seeds_num_100 = K * 100 / depth_all (streamGPU-AS/Render.cpp:156-158)
    s->seeds100 = p->seeds_per_100_frames > 0 ? p->seeds_per_100_frames
                                              : (int)((int64_t)p->seeds_num * 100 / p->total_frames);
    *out = s;
    return ASV_OK;
}

int asv_stream_destroy(asv_stream* s)
{
    delete s;
    return ASV_OK;
}

int asv_stream_block_count(const asv_stream* s) { return s ? (int)s->intervals.size() / 2 : -1; }
int asv_stream_seeds_placed(const asv_stream* s) { return s ? s->placed : -1; }

int asv_stream_block_info(const asv_stream* s, int32_t* info)
{
    if (!s || !info) return sfail(ASV_ERR_ARG, "null argument");
    info[0] = s->next_block;
    info[1] = s->last_A;
    info[2] = (int)s->kept.size();
    info[3] = s->last_dead;
    info[4] = s->last_budget;
    info[5] = s->placed;
    return ASV_OK;
}

// One block, in order (prepare_data_seeds + block_segment + post_and_next,
// streamGPU-AS/Render.cpp:348-569).  rgb / fwd / bwd hold the block's
// frames [start, end) only; the finished frames go to labels_all (global
// ids, W x H x total_frames, GL rows).  Returns the block's first frame and
// frame count written through *frames_done when not null.
int asv_stream_block(asv_stream* s, int32_t block, const uint8_t* rgb, const float* fwd, const float* bwd,
                     int32_t* labels_all, int32_t* frames_done, void* stream)
{
    if (!s || !labels_all) return sfail(ASV_ERR_ARG, "null argument");
    if (block < 0 || 2 * (size_t)block >= s->intervals.size()) return sfail(ASV_ERR_STATE, "no such block");
    const size_t start = (size_t)s->intervals[2 * (size_t)block];
    return asv_stream_block_into(s, block, rgb, fwd, bwd, labels_all + start * s->p.width * s->p.height,
                                 frames_done, stream);
}

}  // extern "C"

// First half of a block (prepare_data_seeds, streamGPU-AS/Render.cpp:348-470):
// the new-seed budget and grid, kept ++ new seeds and the block's map to
// global ids, the resident context reseeded, the overlap frames moved on the
// device and the new frames uploaded, all on `stream`.  The Lloyd loop is
// the caller's (asv_run, or a batched run over several streams' contexts).
int asv::stream_block_begin(asv_stream* s, int32_t block, const uint8_t* rgb, const float* fwd, const float* bwd,
                            void* stream)
{
    if (!s || !rgb || !fwd || !bwd) return sfail(ASV_ERR_ARG, "null argument");
    const int nblocks = (int)s->intervals.size() / 2;
    if (s->pending >= 0) return sfail(ASV_ERR_STATE, "the previous block is not finished");
    if (block != s->next_block || block >= nblocks) return sfail(ASV_ERR_STATE, "blocks must be processed in order");
    const int W = s->p.width, H = s->p.height, O = s->p.overlap;
    const int start = s->intervals[2 * block], end = s->intervals[2 * block + 1];
    const int Db = end - start;

    // new-seed budget (:354-356, :394)
    int budget = (block == 0 && nblocks == 1) ? s->seeds100 * Db / 100 : s->seeds100 * (Db - O) / 100;
    if (budget < 1) budget = 1;
    std::vector<int> newxyz;
    int sgrid = 0;
    int nnew = grid_place(W, H, Db, budget, &newxyz, &sgrid);
    if (nnew <= 0) return sfail(ASV_ERR_ARG, "grid seeding places no seed in this block");

    // kept ++ new seeds, and the block's map to global ids (:434-460)
    const int nkept = (int)s->kept.size();
    std::vector<int> xyz;
    xyz.reserve((size_t)(nkept + nnew) * 3);
    std::vector<int> block2all((size_t)nkept + (size_t)nnew, 0);
    for (int i = 0; i < nkept; ++i) {
        const KeptSeed& k = s->kept[(size_t)i];
        xyz.push_back((int)std::floor(k.x));
        xyz.push_back((int)std::floor(k.y));
        xyz.push_back((int)std::floor(k.z));
        block2all[(size_t)i] = block == 0 ? i : s->prev_map[(size_t)k.old_id];
    }
    int added = s->placed;
    for (int j = 0; j < nnew; ++j) {
        xyz.push_back(newxyz[3 * j]);
        xyz.push_back(newxyz[3 * j + 1]);
        xyz.push_back(newxyz[3 * j + 2]);
        block2all[(size_t)(nkept + j)] = block == 0 ? j : added++;
    }
    const int A = nkept + nnew;

    // The block's frames on the device: the first O frames are the previous
    // block's last O, moved on the device; only the new frames are uploaded.
    const size_t plane = (size_t)W * H;
    int rc = ASV_OK;
    int resident = 0;
    if (s->ctx && A <= s->seed_cap) {
        if (O > 0 && s->prev_depth >= 2 * O) {
            rc = asv_shift_frames(s->ctx, s->prev_depth - O, O, stream);
            resident = O;
        }
        if (!rc) rc = asv_reseed(s->ctx, xyz.data(), A, sgrid, Db, stream);
    } else {
        if (s->ctx) asv_destroy(s->ctx);
        s->ctx = nullptr;
        s->seed_cap = wanted_seed_cap(s);
        if (s->seed_cap < 2 * A) s->seed_cap = 2 * A;
        asv_params cp = ctx_params(s, s->seed_cap);
        rc = asv_create_seeded(&cp, xyz.data(), A, sgrid, &s->ctx);
        if (!rc) rc = asv_reseed(s->ctx, xyz.data(), A, sgrid, Db, stream);
    }
    if (rc) return rc;
    rc = asv_upload_frames_range(s->ctx, resident, Db - resident, rgb + (size_t)resident * plane * 3,
                                 fwd + (size_t)resident * plane * 2, bwd + (size_t)resident * plane * 2, stream);
    if (rc) return rc;
    s->pending = block;
    s->pending_A = A;
    s->pending_budget = budget;
    s->pending_nkept = nkept;
    s->pending_map.swap(block2all);
    return ASV_OK;
}

asv_ctx* asv::stream_ctx(asv_stream* s) { return s ? s->ctx : nullptr; }

// Second half (block_segment's readback and post_and_next,
// streamGPU-AS/Render.cpp:470-569), after the block's Lloyd loop on
// `stream`: check_dead_seeds (streamGPU-AS/Image.cpp:1534-1579) on the
// device, the block2all stitch (:559-563, an id absent from the map reads as
// 0 as std::map operator[] gives) with only the finished frames copied to
// `out`, and the seeds kept for the next block.
int asv::stream_block_end(asv_stream* s, int32_t* out, int32_t* frames_done, void* stream)
{
    if (!s || !out) return sfail(ASV_ERR_ARG, "null argument");
    if (s->pending < 0) return sfail(ASV_ERR_STATE, "no block has begun");
    const int block = s->pending;
    s->pending = -1;
    const int nblocks = (int)s->intervals.size() / 2;
    const int W = s->p.width, H = s->p.height, B = s->p.block_size, O = s->p.overlap, S = B - O;
    const int start = s->intervals[2 * block], end = s->intervals[2 * block + 1];
    const int Db = end - start;
    const bool last = block == nblocks - 1;
    const int A = s->pending_A, nkept = s->pending_nkept;
    const size_t plane = (size_t)W * H;
    asv_ctx* ctx = s->ctx;
    std::vector<int64_t> sums((size_t)(A + 1) * 8);
    int rc = asv_download_sums(ctx, sums.data(), stream);
    if (rc) return rc;
    s->prev_depth = Db;
    const int done_frames = last ? Db : Db - O;
    int dead = 0, nfin = 0;
    rc = asv_postprocess_device(ctx, sums.data(), &dead, &nfin, stream);
    uint64_t ck = 0;
    if (!rc)
        rc = asv_download_labels_mapped(ctx, s->pending_map.data(), A, done_frames, out, (int64_t)start * (int64_t)plane,
                                        &ck, stream);
    if (rc) return rc;
    std::vector<KeptSeed> alive;
    for (int i = 0; i < A; ++i) {
        const int64_t* q = &sums[(size_t)i * 8];
        if (q[3] < 1) continue;
        KeptSeed k;
        k.x = (float)(q[0] / (2 * q[3])) + 0.5f;
        k.y = (float)(q[1] / (2 * q[3])) + 0.5f;
        k.z = (float)(q[2] / (2 * q[3])) + 0.5f;
        k.old_id = i;
        alive.push_back(k);
    }
    s->placed += (int)alive.size() - nkept;                         // :517
    std::vector<KeptSeed> next;
    if (!last) {
        for (const KeptSeed& k : alive) {                          // :524-540
            if (k.z > (float)(B - O)) {
                KeptSeed n = k;
                n.z = k.z - (float)S;
                if (n.z < 1.0f) n.z = 1.5f;
                next.push_back(n);
            }
        }
    }
    s->kept.swap(next);
    s->prev_map.swap(s->pending_map);
    s->checksum += ck;
    s->last_A = A;
    s->last_dead = dead;
    s->last_budget = s->pending_budget;
    s->next_block = block + 1;
    if (frames_done) *frames_done = done_frames;
    return ASV_OK;
}

extern "C" {

int asv_stream_block_into(asv_stream* s, int32_t block, const uint8_t* rgb, const float* fwd, const float* bwd,
                          int32_t* out, int32_t* frames_done, void* stream)
{
    if (!s || !rgb || !fwd || !bwd || !out) return sfail(ASV_ERR_ARG, "null argument");
    int rc = asv::stream_block_begin(s, block, rgb, fwd, bwd, stream);
    if (!rc) rc = asv_run(s->ctx, stream);
    if (!rc) rc = asv::stream_block_end(s, out, frames_done, stream);
    else s->pending = -1;
    return rc;
}

// Whole video through the blocks in order (Render::run of the streaming
// program).  Arrays cover all total_frames frames (image rows).
int asv_stream_run_video(const asv_stream_params* p, const uint8_t* rgb, const float* fwd, const float* bwd,
                         int32_t* labels_all, int32_t* seeds_placed, void* stream)
{
    asv_stream* s = nullptr;
    int rc = asv_stream_create(p, &s);
    if (rc) return rc;
    const size_t plane = (size_t)p->width * p->height;
    const int nb = asv_stream_block_count(s);
    for (int b = 0; b < nb && !rc; ++b) {
        const size_t f0 = (size_t)s->intervals[2 * b];
        rc = asv_stream_block(s, b, rgb + f0 * plane * 3, fwd + f0 * plane * 2, bwd + f0 * plane * 2, labels_all,
                              nullptr, stream);
    }
    if (!rc && seeds_placed) *seeds_placed = s->placed;
    asv_stream_destroy(s);
    return rc;
}

int asv_stream_checksum(const asv_stream* s, uint64_t* out)
{
    if (!s || !out) return sfail(ASV_ERR_ARG, "null argument");
    *out = s->checksum;
    return ASV_OK;
}

const char* asv_stream_last_error(void) { return g_stream_err.c_str(); }

}  // extern "C"
