// asv_capi.cpp -- C ABI of libasv: clip contexts, device buffers and the
// stage sequence of one plain run (GPU-AS/Render.cpp:254-293), with no host
// round trip inside the Lloyd loop.
#include "../../include/asv.h"
#include "asv_internal.h"
#include "asv_host_internal.h"
#include "asv_ctx.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace asv;

namespace {

thread_local std::string g_err;

// asv_debug_fail_alloc's threshold: allocations of at least this many bytes
// fail (0: none).
std::atomic<int64_t> g_fail_alloc_from{0};

int fail(int code, const char* fmt, const char* detail = nullptr)
{
    char buf[512];
    if (detail) std::snprintf(buf, sizeof(buf), fmt, detail);
    else std::snprintf(buf, sizeof(buf), "%s", fmt);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) return fail(ASV_ERR_HIP, #expr ": %s", hipGetErrorString(e_)); \
    } while (0)

// Grid seeding, GPU-AS/Image.cpp:20-92 (integer divisions, double pow,
// float sqrt and truncations as written there).  Returns A or -1.
int grid_seeds(int W, int H, int D, int K, std::vector<int>* xyz, int* s_out, int* dims = nullptr)
{
    int V = W * H * D;
    double voxel_size = (double)(V / K);
    int s = (int)std::floor(std::pow(voxel_size, 1.0 / 3));
    int dNum = (int)std::pow((float)(K * D * D / W / H), 1.0 / 3.0);
    if (dNum == 0) dNum = 1;
    int frame_seeds = K / dNum;
    int wNum = (int)std::sqrt((float)(frame_seeds * W / H));
    if (wNum <= 0) return -1;
    int hNum = frame_seeds / wNum;
    if (hNum <= 0) return -1;
    int stepD = D / dNum, stepH = H / hNum, stepW = W / wNum;
    int d_rem = D - stepD * dNum, h_rem = H - stepH * hNum, w_rem = W - stepW * wNum;
    int td = d_rem < 1 ? 0 : 1;
    int count = 0;
    for (int k = 0; k < dNum; ++k) {
        int th = 1;
        for (int i = 0; i < hNum; ++i) {
            int tw = 1;
            for (int j = 0; j < wNum; ++j) {
                int sx = (int)(j * stepW + 0.5 * stepW + tw);
                int sy = (int)(i * stepH + 0.5 * stepH + th);
                int sz = (int)(k * stepD + 0.5 * stepD + td);
                sx = sx > W - 1 ? W - 1 : sx;
                sy = sy > H - 1 ? H - 1 : sy;
                sz = sz > D - 1 ? D - 1 : sz;
                if (tw < w_rem) ++tw;
                if (xyz) { xyz->push_back(sx); xyz->push_back(sy); xyz->push_back(sz); }
                ++count;
            }
            if (th < h_rem) ++th;
        }
        if (td < d_rem) ++td;
    }
    if (s_out) *s_out = s;
    if (dims) { dims[0] = dNum; dims[1] = hNum; dims[2] = wNum; }
    return count;
}

// JFA pass list, GPU-AS/Image.cpp:1055-1181 (mode 0) or plain halving from
// W/2 plus one step-1 pass (mode 1).
void schedule_list(int mode, int s, int W, std::vector<int>& steps, std::vector<float>& mult)
{
    steps.clear();
    mult.clear();
    if (mode == ASV_SCHED_REFERENCE) {
        steps.push_back(1); mult.push_back(1.0f);
        float k = std::floor(0.5f * (float)s);
        while (k > 0.0f) {
            steps.push_back((int)k); mult.push_back(1.0f);
            if ((int)k == 1) k = 0.0f;
            k = std::floor(0.5f * k + 0.5f);
        }
        k = 2.0f;
        while (k > 0.0f) {
            steps.push_back((int)k); mult.push_back(2.0f);
            if ((int)k == 1) k = 0.0f;
            k = std::floor(0.5f * k + 0.5f);
        }
    } else {
        float k = std::floor(0.5f * (float)W);
        while (k > 0.0f) {
            steps.push_back((int)k); mult.push_back(1.0f);
            if ((int)k == 1) k = 0.0f;
            k = std::floor(0.5f * k + 0.5f);
        }
        steps.push_back(1); mult.push_back(1.0f);
    }
}

// sRGB linearisation table (RGB2LAB.fs:27-40): double pow rounded to float.
void lin_table(float* lut)
{
    for (int c = 0; c < 256; ++c) {
        float u = (float)c / 255.0f;
        if (u <= 0.04045f) lut[c] = u / 12.92f;
        else {
            float t = (u + 0.055f) / 1.055f;
            lut[c] = (float)std::pow((double)t, 2.4);
        }
    }
}

}  // namespace

hipError_t asv::dev_malloc(void** p, size_t bytes)
{
    const int64_t from = g_fail_alloc_from.load(std::memory_order_relaxed);
    if (from > 0 && (int64_t)bytes >= from) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

namespace {

template <typename T>
hipError_t dalloc(T** p, size_t count, int64_t* total)
{
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 16;
    *total += (int64_t)bytes;
    return dev_malloc((void**)p, bytes);
}

bool valid_params(const asv_params* p)
{
    if (!p) return false;
    if (p->width < 2 || p->height < 2 || p->depth < 1) return false;
    if ((int64_t)p->width * p->height * p->depth > (int64_t)1 << 30) return false;
    if (p->seeds_num < 1 || p->lloyd_iters < 1) return false;
    if (!(p->compactness > 0.0f) || !std::isfinite(p->compactness)) return false;
    if (p->tie_rule != ASV_TIE_LOWEST_ID && p->tie_rule != ASV_TIE_SCAN_ORDER) return false;
    if (p->schedule != ASV_SCHED_REFERENCE && p->schedule != ASV_SCHED_HALVING_PLUS_ONE) return false;
    if (p->seed_init != ASV_INIT_GRID && p->seed_init != ASV_INIT_DENSITY) return false;
    return true;
}

}  // namespace


namespace {

void free_ctx(asv_ctx* c)
{
    void* ptrs[] = {c->rgb, c->fwd, c->bwd, c->lut, c->lab, c->seed_own, c->seed_need,
                    c->labels[0], c->labels[1], c->grid, c->pos, c->slab, c->rec, c->sums[0], c->sums[1],
                    c->trace, c->map_table, c->post_scratch, c->ck_word, c->keys, c->unsafe, c->march_stats};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    free_seed_bricks(&c->bricks);
    if (c->sums_host) (void)hipHostFree(c->sums_host);
}

// Whether a context keeps the brick lists of the fused first two passes
// (launch_seed_bricks): the lowest-id rule, a schedule of two passes or
// more starting with a step of at most kSeedFuseMaxK1, coordinates that fit
// 16 bits, and A seeds sparse enough for the second pass over the seeds'
// neighbourhoods.  Whether a run takes them is decided per run
// (asv_ctx::fused_seed_passes).
bool wants_bricks(const asv_params* p, const std::vector<int>& steps, int64_t A)
{
    const int64_t V = (int64_t)p->width * p->height * p->depth;
    return p->tie_rule == ASV_TIE_LOWEST_ID && steps.size() > 1 && steps[0] <= kSeedFuseMaxK1 &&
           seed_pass_kind(steps, A, V, 1) == 2 && p->width < 32768 && p->height < 32768 && p->depth < 32768;
}

// Whether a context keeps the keys of the scatter-form second pass: the
// lowest-id rule, a sparse second pass that the fused form does not take
// (a first step above kSeedFuseMaxK1), coordinates that fit 16 bits.
bool wants_keys(const asv_params* p, const std::vector<int>& steps, int64_t A)
{
    const int64_t V = (int64_t)p->width * p->height * p->depth;
    return p->tie_rule == ASV_TIE_LOWEST_ID && seed_pass_kind(steps, A, V, 1) == 2 && !wants_bricks(p, steps, A) &&
           p->width < 32768 && p->height < 32768 && p->depth < 32768;
}

int64_t footprint_of(int64_t V, int64_t A, int iters, bool trace, bool keys, int64_t bricks)
{
    int64_t b = bricks;
    if (keys) b += V * 8;        // scatter keys of the sparse second pass
    b += V * 3;                  // rgb
    b += V * 16;                 // fwd + bwd
    b += V * 4;                  // lab
    b += V * 8;                  // labels x2
    b += A * (16 * 3 + 48 + 36); // grid, pos, slab, rec, seed_own, seed_need
    b += (A + 1) * 64 * 2;       // sums x2
    if (trace) b += (int64_t)iters * A * 48;
    return b + 4096;
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

int check_ready(asv_ctx* c)
{
    if (!c) return fail(ASV_ERR_ARG, "null context");
    if (!c->have_frames) return fail(ASV_ERR_STATE, "frames not uploaded");
    if (!c->have_flow) return fail(ASV_ERR_STATE, "flow not uploaded");
    return ASV_OK;
}

}  // namespace

extern "C" {

void asv_default_params(asv_params* p, int32_t width, int32_t height, int32_t depth, int32_t seeds_num)
{
    std::memset(p, 0, sizeof(*p));
    p->width = width;
    p->height = height;
    p->depth = depth;
    p->seeds_num = seeds_num;
    p->lloyd_iters = 20;             // GPU-AS/Render.cpp:11
    p->compactness = 8.0f;           // GPU-AS/Image.cpp:1050
    p->tie_rule = ASV_TIE_LOWEST_ID;
    p->angle_gate = 1;
    p->schedule = ASV_SCHED_REFERENCE;
    p->frames_prepared = 0;
}

const char* asv_last_error(void) { return g_err.c_str(); }

int asv_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int64_t asv_footprint(const asv_params* p)
{
    if (!valid_params(p)) return -1;
    int s = 0;
    int A = grid_seeds(p->width, p->height, p->depth, p->seeds_num, nullptr, &s);
    if (A <= 0) return -1;
    std::vector<int> st;
    std::vector<float> mu;
    schedule_list(p->schedule, s < 1 ? 1 : s, p->width, st, mu);
    return footprint_of((int64_t)p->width * p->height * p->depth, A, p->lloyd_iters, false, wants_keys(p, st, A),
                        wants_bricks(p, st, A) ? seed_brick_bytes(p->width, p->height, p->depth, A, st[0] + st[1]) : 0);
}

int asv_plan(const asv_params* p, int32_t* seeds_placed, int32_t* grid_step, int32_t* steps, float* mult,
             int32_t cap)
{
    if (!valid_params(p)) return fail(ASV_ERR_ARG, "invalid parameters");
    int s = 0;
    int A = grid_seeds(p->width, p->height, p->depth, p->seeds_num, nullptr, &s);
    if (A <= 0) return fail(ASV_ERR_ARG, "grid seeding places no seed for these sizes");
    if (s < 1) s = 1;
    std::vector<int> st;
    std::vector<float> mu;
    schedule_list(p->schedule, s, p->width, st, mu);
    if (seeds_placed) *seeds_placed = A;
    if (grid_step) *grid_step = s;
    for (size_t i = 0; i < st.size() && (int)i < cap; ++i) {
        if (steps) steps[i] = st[i];
        if (mult) mult[i] = mu[i];
    }
    return (int)st.size();
}

}  // extern "C"

namespace {

// Allocates a context for p whose initial seeds are xyz (A x 3) with grid
// step s (ratio and schedule follow from s).
int create_ctx(const asv_params* p, std::vector<int>&& xyz, int s, asv_ctx** out, int seed_cap = 0)
{
    const int A = (int)(xyz.size() / 3);
    const int Acap = seed_cap > A ? seed_cap : A;
    asv_ctx* c = new asv_ctx();
    c->p = *p;
    c->grid_host = std::move(xyz);
    if (s < 1) s = 1;
    c->s = s;
    c->d = Dims{p->width, p->height, p->depth, A};
    c->seed_cap = Acap;
    c->depth_cap = p->depth;
    c->ratio = (p->compactness / (float)s) * (p->compactness / (float)s);   // Image.cpp:1051
    schedule_list(p->schedule, s, p->width, c->steps, c->mult);

    const int64_t V = c->V();
    hipError_t e = hipSuccess;
    int64_t& tb = c->bytes;
    if (e == hipSuccess) e = dalloc(&c->rgb, (size_t)V * 3, &tb);
    if (e == hipSuccess) e = dalloc(&c->fwd, (size_t)V, &tb);
    if (e == hipSuccess) e = dalloc(&c->bwd, (size_t)V, &tb);
    if (e == hipSuccess) e = dalloc(&c->lut, 256, &tb);
    if (e == hipSuccess) e = dalloc(&c->lab, (size_t)V, &tb);
    if (e == hipSuccess) e = dalloc(&c->labels[0], (size_t)V, &tb);
    if (e == hipSuccess) e = dalloc(&c->labels[1], (size_t)V, &tb);
    if (e == hipSuccess) e = dalloc(&c->grid, (size_t)Acap, &tb);
    if (e == hipSuccess) e = dalloc(&c->pos, (size_t)Acap, &tb);
    if (e == hipSuccess) e = dalloc(&c->slab, (size_t)Acap, &tb);
    if (e == hipSuccess) e = dalloc(&c->rec, (size_t)Acap, &tb);
    if (e == hipSuccess) e = dalloc(&c->seed_own, (size_t)Acap * 2, &tb);
    if (e == hipSuccess) e = dalloc(&c->seed_need, (size_t)Acap, &tb);
    if (e == hipSuccess) e = dalloc(&c->sums[0], (size_t)(Acap + 1) * kSumWords, &tb);
    if (e == hipSuccess) e = dalloc(&c->sums[1], (size_t)(Acap + 1) * kSumWords, &tb);
    if (e == hipSuccess) e = dalloc(&c->unsafe, 1, &tb);
    if (e == hipSuccess) e = launch_zero_i32(nullptr, c->unsafe, 1);
    // page-locked landing buffers of the small device-to-host copies (sums,
    // checksum): explicit device-to-host transfers into memory the context
    // owns, whatever buffer the caller passes
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&c->sums_host, sizeof(unsigned long long) * ((size_t)(Acap + 1) * kSumWords + 1),
                          hipHostMallocDefault);
    if (e == hipSuccess && wants_keys(p, c->steps, A)) {
        e = dalloc(&c->keys, (size_t)V, &tb);
        if (e == hipSuccess) e = launch_fill_keys(nullptr, c->keys, V);
    }
    if (e == hipSuccess && wants_bricks(p, c->steps, A))
        e = alloc_seed_bricks(&c->bricks, p->width, p->height, p->depth, Acap, c->steps[0] + c->steps[1], &tb);
    if (e != hipSuccess) {
        free_ctx(c);
        delete c;
        return fail(ASV_ERR_NOMEM, "device allocation failed: %s", hipGetErrorString(e));
    }
    float lut[256];
    lin_table(lut);
    std::vector<int4> g((size_t)A);
    for (int i = 0; i < A; ++i)
        g[i] = make_int4(c->grid_host[3 * i], c->grid_host[3 * i + 1], c->grid_host[3 * i + 2], 0);
    e = hipMemcpy(c->lut, lut, sizeof(lut), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->grid, g.data(), sizeof(int4) * (size_t)A, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_ctx(c);
        delete c;
        return fail(ASV_ERR_HIP, "upload of tables failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return ASV_OK;
}

// Density-based re-placement of the reference grid (Image::initialization,
// GPU-AS/Image.cpp:350-620): the device sums each grid seed's colour-gradient
// density, the host walks the density prefix sums (float, as the reference's
// host code) and re-places the A seeds; the result replaces the context's
// initial seeds.  Always starts from the reference grid, so a second
// prepare of the same context gives the same seeds.
int density_reseed(asv_ctx* c, hipStream_t st)
{
    const int W = c->d.W, H = c->d.H, D = c->d.D, A = c->d.A;
    const int dN = c->gdims[0], hN = c->gdims[1], wN = c->gdims[2];
    const std::vector<int>& g = c->grid_ref;
    if (dN * hN * wN != A || (int)g.size() != 3 * A) return fail(ASV_ERR_STATE, "density init needs the reference grid");
    // box of grid index i along an axis of n seeds (Image.cpp:360-433)
    auto box = [](const std::vector<int>& p, int n, int N, int i, int* lo, int* hi) {
        *lo = i > 0 ? (int)std::floor(0.5 * (((double)p[i - 1] + 0.5) + ((double)p[i] + 0.5))) : 0;
        *hi = i < n - 1 ? (int)std::floor(0.5 * (((double)p[i + 1] + 0.5) + ((double)p[i] + 0.5))) : N - 1;
    };
    std::vector<int> px((size_t)wN), py((size_t)hN), pz((size_t)dN);
    for (int i = 0; i < wN; ++i) px[i] = g[3 * i];
    for (int j = 0; j < hN; ++j) py[j] = g[3 * (j * wN) + 1];
    for (int t = 0; t < dN; ++t) pz[t] = g[3 * (t * wN * hN) + 2];
    // grid index of every coordinate; a later box overwrites (Image.cpp:436-458)
    std::vector<int> maps((size_t)(W + H + D), -1);
    int* mx = maps.data();
    int* my = mx + W;
    int* mz = my + H;
    int lo, hi;
    for (int i = 0; i < wN; ++i) { box(px, wN, W, i, &lo, &hi); for (int x = std::max(lo, 0); x <= std::min(hi, W - 1); ++x) mx[x] = i; }
    for (int j = 0; j < hN; ++j) { box(py, hN, H, j, &lo, &hi); for (int y = std::max(lo, 0); y <= std::min(hi, H - 1); ++y) my[y] = j; }
    for (int t = 0; t < dN; ++t) { box(pz, dN, D, t, &lo, &hi); for (int z = std::max(lo, 0); z <= std::min(hi, D - 1); ++z) mz[z] = t; }
    // Ids pass through R16F, so the highest slot a voxel can address is the
    // float16 rounding of A - 1, which may lie below A - 1 (A = 2146: 2145
    // rounds to 2144).  The reference reads all A texels (Image.cpp:524-552);
    // a slot no voxel addresses holds only the clear colour, so the host
    // table covers max(nslot, A) slots, zero beyond the device's nslot.
    const int nslot = (int)(float)(_Float16)(float)(A - 1) + 1;
    int* d_maps = nullptr;
    unsigned long long* d_acc = nullptr;
    std::vector<unsigned long long> acc((size_t)std::max(nslot, A), 0ull);
    hipError_t e = asv::dev_malloc((void**)&d_maps, sizeof(int) * maps.size());
    if (e == hipSuccess) e = asv::dev_malloc((void**)&d_acc, sizeof(unsigned long long) * (size_t)nslot);
    if (e == hipSuccess) e = hipMemcpyAsync(d_maps, maps.data(), sizeof(int) * maps.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_zero_u64(st, reinterpret_cast<uint64_t*>(d_acc), nslot);
    if (e == hipSuccess)
        e = launch_density(st, c->d, c->rgb, c->p.frames_prepared, d_maps, d_maps + W, d_maps + W + H, wN, hN, nslot,
                           d_acc);
    if (e == hipSuccess)
        e = hipMemcpyAsync(acc.data(), d_acc, sizeof(unsigned long long) * (size_t)nslot, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (d_maps) (void)hipFree(d_maps);
    if (d_acc) (void)hipFree(d_acc);
    if (e != hipSuccess) return fail(ASV_ERR_HIP, "density init: %s", hipGetErrorString(e));
    // densities, prefix sums and the walk (Image.cpp:545-616), in float
    std::vector<float> len((size_t)A);
    for (int i = 0; i < A; ++i) {
        const float res = (float)(3.0 + (double)acc[(size_t)i] / 4096.0);   // clear colour 3 + the blended sum
        const float d = (float)(0.001 * (double)res);
        len[(size_t)i] = i == 0 ? d : d + len[(size_t)i - 1];
    }
    const float interval = len[(size_t)A - 1] / (float)(A + 1);
    std::vector<int> xyz((size_t)A * 3);
    int j1 = 0, placed = 0;
    for (int i = 0; i < A;) {
        const float seed_in = (float)(i + 1) * interval;
        if (seed_in < len[(size_t)j1]) {
            box(px, wN, W, j1 % wN, &lo, &hi);
            const float frac = j1 == 0 ? seed_in / len[0]
                                       : (seed_in - len[(size_t)j1 - 1]) / (len[(size_t)j1] - len[(size_t)j1 - 1]);
            xyz[3 * (size_t)i] = (int)std::floor((float)lo + (float)(hi - lo) * frac);
            xyz[3 * (size_t)i + 1] = g[3 * (size_t)j1 + 1];
            xyz[3 * (size_t)i + 2] = g[3 * (size_t)j1 + 2];
            ++i;
            ++placed;
        } else if (++j1 == A) {
            break;
        }
    }
    if (placed != A) return fail(ASV_ERR_STATE, "density init placed fewer seeds than the grid");
    c->grid_host = xyz;
    c->grid_staging.resize((size_t)A);
    for (int i = 0; i < A; ++i)
        c->grid_staging[(size_t)i] = make_int4(xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2], 0);
    HIP_TRY(hipMemcpyAsync(c->grid, c->grid_staging.data(), sizeof(int4) * (size_t)A, hipMemcpyHostToDevice, st));
    return ASV_OK;
}

}  // namespace

extern "C" {

int asv_create(const asv_params* p, asv_ctx** out)
{
    if (!out) return fail(ASV_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!valid_params(p)) return fail(ASV_ERR_ARG, "invalid parameters");
    int s = 0, dims[3] = {0, 0, 0};
    std::vector<int> xyz;
    int A = grid_seeds(p->width, p->height, p->depth, p->seeds_num, &xyz, &s, dims);
    if (A <= 0) return fail(ASV_ERR_ARG, "grid seeding places no seed for these sizes");
    std::vector<int> ref = xyz;
    int r = create_ctx(p, std::move(xyz), s, out);
    if (r == ASV_OK) {
        (*out)->grid_ref = std::move(ref);
        for (int i = 0; i < 3; ++i) (*out)->gdims[i] = dims[i];
    }
    return r;
}

int asv_create_seeded(const asv_params* p, const int32_t* xyz, int32_t n, int32_t grid_step, asv_ctx** out)
{
    if (!out) return fail(ASV_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!valid_params(p)) return fail(ASV_ERR_ARG, "invalid parameters");
    if (!xyz || n < 1 || grid_step < 1) return fail(ASV_ERR_ARG, "need at least one seed and a grid step >= 1");
    std::vector<int> v((size_t)n * 3);
    for (int i = 0; i < n; ++i) {
        int x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        if (x < 0 || x >= p->width || y < 0 || y >= p->height || z < 0 || z >= p->depth)
            return fail(ASV_ERR_ARG, "seed voxel outside the clip");
        v[3 * i] = x; v[3 * i + 1] = y; v[3 * i + 2] = z;
    }
    return create_ctx(p, std::move(v), grid_step, out, p->seeds_num);
}

int asv_destroy(asv_ctx* c)
{
    if (!c) return ASV_OK;
    free_ctx(c);
    delete c;
    return ASV_OK;
}

}  // extern "C"

bool asv::ctx_reusable(const asv_ctx* c, const asv_params* p, int32_t seeds)
{
    return c && p && c->p.width == p->width && c->p.height == p->height && c->depth_cap >= p->depth &&
           c->seed_cap >= seeds && c->p.lloyd_iters == p->lloyd_iters && c->p.compactness == p->compactness &&
           c->p.tie_rule == p->tie_rule && c->p.angle_gate == p->angle_gate && c->p.schedule == p->schedule &&
           c->p.frames_prepared == p->frames_prepared && c->p.seed_init == p->seed_init && !c->trace_on;
}

int32_t asv::ctx_seed_cap(const asv_ctx* c) { return c ? c->seed_cap : 0; }



extern "C" {

int asv_seeds_placed(const asv_ctx* c) { return c ? c->d.A : fail(ASV_ERR_ARG, "null context"); }
int asv_grid_step(const asv_ctx* c) { return c ? c->s : fail(ASV_ERR_ARG, "null context"); }
int64_t asv_device_bytes(const asv_ctx* c) { return c ? c->bytes : -1; }

int asv_schedule(const asv_ctx* c, int32_t* steps, float* mult, int32_t cap)
{
    if (!c) return fail(ASV_ERR_ARG, "null context");
    int n = (int)c->steps.size();
    for (int i = 0; i < n && i < cap; ++i) {
        if (steps) steps[i] = c->steps[i];
        if (mult) mult[i] = c->mult[i];
    }
    return n;
}

int asv_plan_seeds(const asv_params* p, int32_t* xyz, int32_t cap)
{
    if (!valid_params(p)) return fail(ASV_ERR_ARG, "invalid parameters");
    std::vector<int> v;
    int A = grid_seeds(p->width, p->height, p->depth, p->seeds_num, &v, nullptr);
    if (A <= 0) return fail(ASV_ERR_ARG, "grid seeding places no seed for these sizes");
    for (int i = 0; i < A * 3 && i < cap; ++i) xyz[i] = v[(size_t)i];
    return A;
}

int asv_grid_seeds(const asv_ctx* c, int32_t* xyz, int32_t cap)
{
    if (!c) return fail(ASV_ERR_ARG, "null context");
    int n = c->d.A;
    for (int i = 0; i < n * 3 && i < cap; ++i) xyz[i] = c->grid_host[i];
    return n;
}

int asv_upload_frames(asv_ctx* c, const uint8_t* rgb, void* stream)
{
    if (!c || !rgb) return fail(ASV_ERR_ARG, "null argument");
    HIP_TRY(hipMemcpyAsync(c->rgb, rgb, (size_t)c->V() * 3, hipMemcpyDefault, S(stream)));
    c->have_frames = true;
    c->prepared = false;
    c->lab_frames = c->lab_kept = 0;
    return ASV_OK;
}

int asv_reseed(asv_ctx* c, const int32_t* xyz, int32_t n, int32_t grid_step, int32_t depth, void* stream)
{
    if (!c || !xyz || n < 1 || grid_step < 1) return fail(ASV_ERR_ARG, "need a context, seeds and a grid step");
    if (n > c->seed_cap) return fail(ASV_ERR_ARG, "more seeds than the context holds");
    if (depth < 1 || depth > c->depth_cap) return fail(ASV_ERR_ARG, "depth outside the context's capacity");
    for (int i = 0; i < n; ++i)
        if (xyz[3 * i] < 0 || xyz[3 * i] >= c->d.W || xyz[3 * i + 1] < 0 || xyz[3 * i + 1] >= c->d.H ||
            xyz[3 * i + 2] < 0 || xyz[3 * i + 2] >= depth)
            return fail(ASV_ERR_ARG, "seed voxel outside the clip");
    c->p.depth = depth;
    c->d.D = depth;
    c->d.A = n;
    c->s = grid_step;
    c->ratio = (c->p.compactness / (float)grid_step) * (c->p.compactness / (float)grid_step);
    schedule_list(c->p.schedule, grid_step, c->d.W, c->steps, c->mult);
    c->grid_host.assign(xyz, xyz + 3 * (size_t)n);
    c->grid_ref.clear();                       // no longer the reference grid
    c->gdims[0] = c->gdims[1] = c->gdims[2] = 0;
    c->grid_staging.resize((size_t)n);
    for (int i = 0; i < n; ++i) c->grid_staging[(size_t)i] = make_int4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0);
    HIP_TRY(hipMemcpyAsync(c->grid, c->grid_staging.data(), sizeof(int4) * (size_t)n, hipMemcpyHostToDevice, S(stream)));
    c->prepared = false;
    c->last_iter = 0;
    if (c->trace) { (void)hipFree(c->trace); c->trace = nullptr; c->trace_on = false; }
    return ASV_OK;
}

int asv_shift_frames(asv_ctx* c, int32_t src_frame, int32_t n_frames, void* stream)
{
    if (!c || src_frame < 0 || n_frames < 0 || src_frame + n_frames > c->depth_cap || n_frames > src_frame)
        return fail(ASV_ERR_ARG, "frame ranges must lie in the context and not overlap");
    if (n_frames == 0) return ASV_OK;
    const size_t plane = (size_t)c->d.W * c->d.H;
    const size_t f0 = (size_t)src_frame * plane, n = (size_t)n_frames * plane;
    HIP_TRY(hipMemcpyAsync(c->rgb, c->rgb + f0 * 3, n * 3, hipMemcpyDeviceToDevice, S(stream)));
    HIP_TRY(hipMemcpyAsync(c->fwd, c->fwd + f0, n * 8, hipMemcpyDeviceToDevice, S(stream)));
    HIP_TRY(hipMemcpyAsync(c->bwd, c->bwd + f0, n * 8, hipMemcpyDeviceToDevice, S(stream)));
    // the moved frames' LAB moves with them when it is current (4 bytes a
    // voxel copied instead of the blur and conversion of those frames)
    if (c->lab_frames >= src_frame + n_frames) {
        HIP_TRY(hipMemcpyAsync(c->lab, c->lab + f0, n * 4, hipMemcpyDeviceToDevice, S(stream)));
        c->lab_frames = c->lab_kept = n_frames;
    } else {
        c->lab_frames = c->lab_kept = 0;
    }
    c->prepared = false;
    return ASV_OK;
}

int asv_upload_frames_range(asv_ctx* c, int32_t first, int32_t count, const uint8_t* rgb, const float* fwd,
                            const float* bwd, void* stream)
{
    if (!c || !rgb || !fwd || !bwd || first < 0 || count < 0 || first + count > c->depth_cap)
        return fail(ASV_ERR_ARG, "frame range outside the context");
    const size_t plane = (size_t)c->d.W * c->d.H;
    const size_t f0 = (size_t)first * plane, n = (size_t)count * plane;
    HIP_TRY(hipMemcpyAsync(c->rgb + f0 * 3, rgb, n * 3, hipMemcpyDefault, S(stream)));
    HIP_TRY(hipMemcpyAsync(c->fwd + f0, fwd, n * 8, hipMemcpyDefault, S(stream)));
    HIP_TRY(hipMemcpyAsync(c->bwd + f0, bwd, n * 8, hipMemcpyDefault, S(stream)));
    if (count > 0 && c->lab_frames > first) c->lab_frames = first;
    if (c->lab_kept > c->lab_frames) c->lab_kept = c->lab_frames;
    c->have_frames = true;
    c->have_flow = true;
    c->prepared = false;
    return ASV_OK;
}

int asv_upload_flow(asv_ctx* c, const float* fwd, const float* bwd, void* stream)
{
    if (!c || !fwd || !bwd) return fail(ASV_ERR_ARG, "null argument");
    HIP_TRY(hipMemcpyAsync(c->fwd, fwd, (size_t)c->V() * 8, hipMemcpyDefault, S(stream)));
    HIP_TRY(hipMemcpyAsync(c->bwd, bwd, (size_t)c->V() * 8, hipMemcpyDefault, S(stream)));
    c->have_flow = true;
    c->prepared = false;
    return ASV_OK;
}

int asv_prepare(asv_ctx* c, void* stream)
{
    int r = check_ready(c);
    if (r) return r;
    hipStream_t st = S(stream);
    // frames [z0, D): all but the frames moved with their LAB (per frame,
    // so the launch covers a suffix of the volume)
    const int z0 = std::min(c->lab_kept, c->d.D);
    if (z0 < c->d.D) {
        const size_t plane = (size_t)c->d.W * c->d.H;
        Dims rest = c->d;
        rest.D = c->d.D - z0;
        HIP_TRY(launch_prepare(st, rest, c->rgb + (size_t)z0 * plane * 3, c->p.frames_prepared, c->lut,
                               c->lab + (size_t)z0 * plane));
    }
    c->lab_frames = c->d.D;
    c->lab_kept = 0;
    // the seed stages of the run raise it when a metric needs the exact clamp
    HIP_TRY(launch_fill(st, c->unsafe, 1, c->force_exact ? 1 : 0));
    if (c->p.seed_init == ASV_INIT_DENSITY && c->gdims[0] > 0 && (r = density_reseed(c, st))) return r;
    c->prepared = true;
    return ASV_OK;
}

int asv_seed_stage(asv_ctx* c, int32_t iter, void* stream)
{
    int r = check_ready(c);
    if (r) return r;
    if (!c->prepared) return fail(ASV_ERR_STATE, "asv_prepare has not run on these inputs");
    if (iter < 1 || iter > c->p.lloyd_iters) return fail(ASV_ERR_ARG, "iteration out of range");
    if (iter != 1 && c->last_iter != iter - 1) return fail(ASV_ERR_STATE, "seed stage out of order");
    hipStream_t st = S(stream);
    if (iter == 1) {
        HIP_TRY(launch_seed_init(st, c->d, c->grid, c->lab, c->pos, c->slab));
        HIP_TRY(launch_zero_u64(st, reinterpret_cast<uint64_t*>(c->sums[1]), (int64_t)(c->d.A + 1) * kSumWords));
    }
    // the fill rewrites the whole volume, so every iteration can start in
    // buffer 0 (pass p then reads buffer p & 1 in every iteration)
    c->cur = 0;
    float* tr = c->trace_on && c->trace ? c->trace + (size_t)(iter - 1) * c->d.A * 12 : nullptr;
    int gate = (iter >= 2 && c->p.angle_gate) ? 1 : 0;
    HIP_TRY(launch_seed_metric(st, c->d, c->pos, c->slab, c->fwd, c->bwd, gate, c->rec, c->seed_own, c->seed_need,
                               tr, c->unsafe));
    // the fused first two passes read the seed positions, not a volume
    if (c->pass_kind(0) != 3) {
        HIP_TRY(launch_fill(st, c->labels[c->cur], c->V(), -1));
        HIP_TRY(launch_place(st, c->d, c->pos, c->labels[c->cur]));
    }
    c->fresh = true;
    c->last_iter = iter;
    return ASV_OK;
}

namespace {

// Pass i of the schedule from labels[cur] into labels[1 - cur].  Right after
// the seed stage the first pass sees only the scattered seeds and the second
// only their neighbourhoods, so those evaluate just the voxels that can
// carry a label (asv_ctx::pass_kind).
int pass_kind_now(const asv_ctx* c, size_t i) { return c->fresh ? c->pass_kind(i) : 0; }

// (n: the passes this flood runs, whose last writes 32-bit labels)
hipError_t launch_pass(asv_ctx* c, size_t i, size_t n, hipStream_t st)
{
    const float ratio = c->mult[i] * c->ratio;
    const int32_t* in = c->labels[c->cur];
    int32_t* out = c->labels[1 - c->cur];
    switch (pass_kind_now(c, i)) {
    case 1:
        return launch_jfa_seeds(st, c->d, in, out, c->lab, c->rec, c->pos, c->steps[0], ratio, c->p.tie_rule,
                                c->first_pass_fill());
    case 2:
        return launch_jfa_seeds2(st, c->d, in, out, c->lab, c->rec, c->pos, c->steps[0], c->steps[1], ratio,
                                 c->p.tie_rule, c->keys);
    case 3:
        // both passes: the second's output lands where it would after two
        // passes, in the buffer this pass reads
        return launch_seed_bricks(st, c->d, c->pos, c->rec, c->lab, c->labels[c->cur], c->steps[0], c->steps[1],
                                  ratio, c->mult[1] * c->ratio, c->bricks, c->writes_label16(1, n, true));
    case 4:
        return hipSuccess;
    default:
        return launch_jfa(st, c->d, in, out, c->lab, c->rec, c->steps[i], ratio, c->p.tie_rule, c->unsafe,
                          c->march_stats, c->label_widths(i, n, c->fresh));
    }
}

}  // namespace

int asv_flood(asv_ctx* c, int32_t iter, void* stream)
{
    if (!c) return fail(ASV_ERR_ARG, "null context");
    if (iter != c->last_iter) return fail(ASV_ERR_STATE, "flood before the seed stage of this iteration");
    hipStream_t st = S(stream);
    for (size_t i = 0; i < c->steps.size(); ++i) {
        HIP_TRY(launch_pass(c, i, c->steps.size(), st));
        c->cur = 1 - c->cur;
    }
    c->fresh = false;
    return ASV_OK;
}

int asv_flood_passes(asv_ctx* c, int32_t iter, int32_t npasses, void* stream)
{
    if (!c) return fail(ASV_ERR_ARG, "null context");
    if (iter != c->last_iter || !c->fresh)
        return fail(ASV_ERR_STATE, "partial flood only right after the seed stage of this iteration");
    if (npasses < 1 || npasses > (int)c->steps.size()) return fail(ASV_ERR_ARG, "pass count outside the schedule");
    if (npasses == 1 && pass_kind_now(c, 0) == 3)
        return fail(ASV_ERR_ARG, "the first two passes run fused: stop after the second");
    hipStream_t st = S(stream);
    for (int i = 0; i < npasses; ++i) {
        HIP_TRY(launch_pass(c, (size_t)i, (size_t)npasses, st));
        c->cur = 1 - c->cur;
    }
    c->fresh = false;
    return ASV_OK;
}

int asv_centroids(asv_ctx* c, int32_t iter, void* stream)
{
    if (!c) return fail(ASV_ERR_ARG, "null context");
    if (iter != c->last_iter) return fail(ASV_ERR_STATE, "centroids before the seed stage of this iteration");
    hipStream_t st = S(stream);
    unsigned long long* s = c->sums[iter & 1];
    HIP_TRY(launch_reduce(st, c->d, c->labels[c->cur], c->lab, s));
    if (iter != c->p.lloyd_iters)                                  // GPU-AS/Image.cpp:1269
        HIP_TRY(launch_finalize(st, c->d, s, c->sums[(iter + 1) & 1], c->lab, c->pos, c->slab));
    return ASV_OK;
}

int asv_run_loop(asv_ctx* c, void* stream)
{
    int r = check_ready(c);
    if (r) return r;
    if (!c->prepared) return fail(ASV_ERR_STATE, "asv_prepare has not run on these inputs");
    for (int it = 1; it <= c->p.lloyd_iters; ++it) {
        if ((r = asv_seed_stage(c, it, stream))) return r;
        if ((r = asv_flood(c, it, stream))) return r;
        if ((r = asv_centroids(c, it, stream))) return r;
    }
    return ASV_OK;
}

int asv_run(asv_ctx* c, void* stream)
{
    int r = asv_prepare(c, stream);
    if (r) return r;
    return asv_run_loop(c, stream);
}

int asv_run_timed(asv_ctx* c, void* stream, float* ms, int32_t* passes)
{
    int r = check_ready(c);
    if (r) return r;
    if (!ms) return fail(ASV_ERR_ARG, "null timing output");
    hipStream_t st = S(stream);
    // Two events per launch group; recorded on the launch stream itself.
    std::vector<hipEvent_t> ev;
    auto mark = [&](void) -> hipEvent_t {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ev.push_back(e);
        (void)hipEventRecord(e, st);
        return e;
    };
    struct Span { hipEvent_t a, b; int cls; };
    std::vector<Span> spans;
    hipEvent_t t0 = mark();
    hipEvent_t a = mark();
    if ((r = asv_prepare(c, stream))) return r;
    spans.push_back({a, mark(), 1});
    int np = 0;
    for (int it = 1; it <= c->p.lloyd_iters; ++it) {
        a = mark();
        if ((r = asv_seed_stage(c, it, stream))) return r;
        spans.push_back({a, mark(), 4});
        for (size_t i = 0; i < c->steps.size(); ++i) {
            a = mark();
            const bool first = pass_kind_now(c, i) != 0;
            HIP_TRY(launch_pass(c, i, c->steps.size(), st));
            spans.push_back({a, mark(), first ? 5 : 2});
            c->cur = 1 - c->cur;
            if (!first) ++np;
        }
        c->fresh = false;
        unsigned long long* s = c->sums[it & 1];
        a = mark();
        HIP_TRY(launch_reduce(st, c->d, c->labels[c->cur], c->lab, s));
        spans.push_back({a, mark(), 3});
        if (it != c->p.lloyd_iters) {
            a = mark();
            HIP_TRY(launch_finalize(st, c->d, s, c->sums[(it + 1) & 1], c->lab, c->pos, c->slab));
            spans.push_back({a, mark(), 4});
        }
    }
    hipEvent_t t1 = mark();
    HIP_TRY(hipEventSynchronize(t1));
    for (int i = 0; i < 6; ++i) ms[i] = 0.0f;
    float f = 0.0f;
    HIP_TRY(hipEventElapsedTime(&f, t0, t1));
    ms[0] = f;
    for (auto& sp : spans) {
        HIP_TRY(hipEventElapsedTime(&f, sp.a, sp.b));
        ms[sp.cls] += f;
    }
    for (auto e : ev) (void)hipEventDestroy(e);
    if (passes) *passes = np;
    return ASV_OK;
}

int asv_download_labels(asv_ctx* c, int32_t* labels, void* stream)
{
    if (!c || !labels) return fail(ASV_ERR_ARG, "null argument");
    if (c->last_iter < 1) return fail(ASV_ERR_STATE, "no run has completed");
    HIP_TRY(hipMemcpyAsync(labels, c->labels[c->cur], (size_t)c->V() * 4, hipMemcpyDefault, S(stream)));
    HIP_TRY(hipStreamSynchronize(S(stream)));
    return ASV_OK;
}

int asv_download_labels_mapped(asv_ctx* c, const int32_t* table, int32_t n_table, int32_t n_frames,
                               int32_t* labels, int64_t index0, uint64_t* checksum, void* stream)
{
    if (!c || !table || !labels || n_table < 1 || n_frames < 0 || n_frames > c->d.D)
        return fail(ASV_ERR_ARG, "bad label map arguments");
    if (c->last_iter < 1) return fail(ASV_ERR_STATE, "no run has completed");
    hipStream_t st = S(stream);
    if (n_table > c->map_cap) {
        if (c->map_table) (void)hipFree(c->map_table);
        c->map_table = nullptr;
        c->map_cap = 0;
        HIP_TRY(asv::dev_malloc((void**)&c->map_table, sizeof(int32_t) * (size_t)n_table));
        c->map_cap = n_table;
    }
    HIP_TRY(hipMemcpyAsync(c->map_table, table, sizeof(int32_t) * (size_t)n_table, hipMemcpyDefault, st));
    const int64_t n = (int64_t)n_frames * c->d.W * c->d.H;
    int32_t* scratch = c->labels[1 - c->cur];          // free after the run
    HIP_TRY(launch_map_labels(st, c->labels[c->cur], scratch, n, c->map_table, n_table));
    HIP_TRY(hipMemcpyAsync(labels, scratch, (size_t)n * 4, hipMemcpyDefault, st));
    if (checksum) {
        if (!c->ck_word) HIP_TRY(asv::dev_malloc((void**)&c->ck_word, sizeof(unsigned long long)));
        unsigned long long* ck = c->ck_word;
        unsigned long long* host = c->sums_host + (size_t)(c->seed_cap + 1) * kSumWords;
        HIP_TRY(launch_zero_u64(st, reinterpret_cast<uint64_t*>(ck), 1));
        HIP_TRY(launch_checksum(st, scratch, n, index0, ck));
        HIP_TRY(hipMemcpyAsync(host, ck, sizeof(*host), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        *checksum = *host;
        return ASV_OK;
    }
    HIP_TRY(hipStreamSynchronize(st));
    return ASV_OK;
}

}  // extern "C"

namespace {

// The decision of check_dead_seeds (GPU-AS/Image.cpp:1579-1720) from the
// per-slot counts: which rewrite the labels need and its table.
//   mode -1: none (no dead seed, no unlabelled voxel);
//   mode  0: label_replace composition for up to three dead seeds (the
//            current last id takes the dead id, once per dead seed, in
//            order), then, if unlabelled voxels exist, the raster fill;
//   mode  1: more than three dead seeds: alive ids to their rank.
struct PostPlan {
    int mode = -1;
    std::vector<int32_t> table;
    int dead = 0, valid = 0;
};

PostPlan post_plan(const int64_t* sums, int A)
{
    PostPlan pl;
    const bool invalid = sums[(size_t)A * 8 + 3] > 0;
    std::vector<int> dead;
    std::vector<int32_t> rank((size_t)A, 0);
    for (int i = 0; i < A; ++i) {
        if (sums[(size_t)i * 8 + 3] < 1) dead.push_back(i);
        else rank[(size_t)i] = pl.valid++;
    }
    pl.dead = (int)dead.size();
    if (pl.dead == 0) {
        if (invalid) pl.mode = 0;
    } else if (pl.dead <= 3) {
        pl.mode = 0;
        pl.table.resize((size_t)A);
        for (int l = 0; l < A; ++l) {
            int v = l, last = A;
            for (int d : dead) {
                last -= 1;
                if (v == last) v = d;
            }
            pl.table[(size_t)l] = v;
        }
    } else {
        pl.mode = 1;
        pl.table.swap(rank);
    }
    return pl;
}

}  // namespace

extern "C" {

int asv_postprocess_device(asv_ctx* c, const int64_t* sums, int32_t* dead_out, int32_t* final_label_num,
                           void* stream)
{
    if (!c || !sums) return fail(ASV_ERR_ARG, "null argument");
    if (c->last_iter < 1) return fail(ASV_ERR_STATE, "no run has completed");
    const int A = c->d.A;
    PostPlan pl = post_plan(sums, A);
    if (dead_out) *dead_out = pl.dead;
    if (final_label_num) *final_label_num = pl.valid;
    if (pl.mode < 0) return ASV_OK;
    hipStream_t st = S(stream);
    const int64_t V = c->V();
    const int64_t words = post_scratch_words(V);
    if (words > c->post_cap) {
        if (c->post_scratch) (void)hipFree(c->post_scratch);
        c->post_scratch = nullptr;
        c->post_cap = 0;
        HIP_TRY(asv::dev_malloc((void**)&c->post_scratch, sizeof(long long) * (size_t)words));
        c->post_cap = words;
    }
    const int n_table = (int)pl.table.size();
    if (n_table > c->map_cap) {
        if (c->map_table) (void)hipFree(c->map_table);
        c->map_table = nullptr;
        c->map_cap = 0;
        HIP_TRY(asv::dev_malloc((void**)&c->map_table, sizeof(int32_t) * (size_t)n_table));
        c->map_cap = n_table;
    }
    if (n_table > 0)
        HIP_TRY(hipMemcpyAsync(c->map_table, pl.table.data(), sizeof(int32_t) * (size_t)n_table, hipMemcpyDefault,
                               st));
    HIP_TRY(launch_postprocess(st, c->labels[c->cur], c->labels[1 - c->cur], V, pl.mode, c->map_table, n_table,
                               c->post_scratch));
    c->cur = 1 - c->cur;
    HIP_TRY(hipStreamSynchronize(st));          // the table's host copy is local
    return ASV_OK;
}

int asv_postprocess_labels(const int64_t* sums, int32_t A, const int32_t* in, int32_t* out, int64_t V,
                           int32_t* dead_out, int32_t* final_label_num, void* stream)
{
    if (!sums || !in || !out || A < 1 || V < 1 || in == out) return fail(ASV_ERR_ARG, "bad post-process arguments");
    PostPlan pl = post_plan(sums, A);
    if (dead_out) *dead_out = pl.dead;
    if (final_label_num) *final_label_num = pl.valid;
    hipStream_t st = S(stream);
    if (pl.mode < 0) {
        HIP_TRY(hipMemcpyAsync(out, in, sizeof(int32_t) * (size_t)V, hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));
        return ASV_OK;
    }
    const int64_t words = post_scratch_words(V);
    long long* scratch = nullptr;
    int32_t* table = nullptr;
    const int n_table = (int)pl.table.size();
    HIP_TRY(asv::dev_malloc((void**)&scratch, sizeof(long long) * (size_t)words));
    hipError_t e = hipSuccess;
    if (n_table > 0) {
        e = asv::dev_malloc((void**)&table, sizeof(int32_t) * (size_t)n_table);
        if (e == hipSuccess)
            e = hipMemcpyAsync(table, pl.table.data(), sizeof(int32_t) * (size_t)n_table, hipMemcpyDefault, st);
    }
    if (e == hipSuccess) e = launch_postprocess(st, in, out, V, pl.mode, table, n_table, scratch);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(scratch);
    if (table) (void)hipFree(table);
    if (e != hipSuccess) return fail(ASV_ERR_HIP, "post-process: %s", hipGetErrorString(e));
    return ASV_OK;
}

}  // extern "C"

namespace {

// Scratch of asv_postprocess_labels_noalloc: one copy of this array per
// device, held until the process ends.  Words [0, kNaChunkWords) take the
// scan's chunk maxima (one 64-bit word per 4096 voxels), the rest the label
// table (one int32 per seed).  Host code uses only the device address that
// hipGetSymbolAddress returns for the current device, never the array's
// name (a host-side stand-in).
constexpr int64_t kNaChunkWords = 131072;         // volumes up to 131072 x 4096 = 2^29 voxels
constexpr int64_t kNaTableInts = 65536;           // labels 0 .. 65535
constexpr int64_t kNaWords = kNaChunkWords + kNaTableInts / 2;
static __device__ __align__(256) unsigned long long scratch[kNaWords];

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
void* g_na_ptr[kNaMaxDevices] = {};

// This device's address of `scratch`, looked up on the first call on the
// device (with the device current) and kept.
hipError_t noalloc_scratch(void** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kNaMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_na_mu);
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_na_ptr[dev] = p;
    }
    *out = g_na_ptr[dev];
    return hipSuccess;
}

}  // namespace

extern "C" {

int asv_postprocess_labels_noalloc(const int64_t* sums, int32_t A, const int32_t* in, int32_t* out, int64_t V,
                                   int32_t* dead_out, int32_t* final_label_num, void* stream)
{
    if (!sums || !in || !out || A < 1 || V < 1 || in == out) return fail(ASV_ERR_ARG, "bad post-process arguments");
    if (post_scratch_words(V) > kNaChunkWords)
        return fail(ASV_ERR_ARG, "volume exceeds the no-allocation scratch (at most 2^29 voxels)");
    if (A > kNaTableInts) return fail(ASV_ERR_ARG, "more seeds than the no-allocation scratch holds (65536)");
    PostPlan pl = post_plan(sums, A);
    if (dead_out) *dead_out = pl.dead;
    if (final_label_num) *final_label_num = pl.valid;
    hipStream_t st = S(stream);
    if (pl.mode < 0) {
        HIP_TRY(hipMemcpyAsync(out, in, sizeof(int32_t) * (size_t)V, hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));
        return ASV_OK;
    }
    void* base = nullptr;
    HIP_TRY(noalloc_scratch(&base));
    long long* chunk = static_cast<long long*>(base);
    int32_t* table = reinterpret_cast<int32_t*>(static_cast<unsigned long long*>(base) + kNaChunkWords);
    const int n_table = (int)pl.table.size();
    if (n_table > 0)
        HIP_TRY(hipMemcpyAsync(table, pl.table.data(), sizeof(int32_t) * (size_t)n_table, hipMemcpyHostToDevice, st));
    HIP_TRY(launch_postprocess(st, in, out, V, pl.mode, table, n_table, chunk));
    HIP_TRY(hipStreamSynchronize(st));          // the table's host copy is local
    return ASV_OK;
}

int asv_initialization(asv_ctx* c, int32_t seed_init)
{
    if (!c) return fail(ASV_ERR_ARG, "null context");
    if (seed_init != ASV_INIT_GRID && seed_init != ASV_INIT_DENSITY) return fail(ASV_ERR_ARG, "unknown seed placement");
    if (seed_init == ASV_INIT_DENSITY && c->gdims[0] == 0)
        return fail(ASV_ERR_STATE, "density placement needs the reference grid (context made from explicit seeds)");
    if (c->p.seed_init != seed_init) {
        c->p.seed_init = seed_init;
        if (seed_init == ASV_INIT_GRID) {               // back to the plain grid
            c->grid_host = c->grid_ref;
            const int A = c->d.A;
            c->grid_staging.resize((size_t)A);
            for (int i = 0; i < A; ++i)
                c->grid_staging[(size_t)i] = make_int4(c->grid_ref[3 * (size_t)i], c->grid_ref[3 * (size_t)i + 1],
                                                       c->grid_ref[3 * (size_t)i + 2], 0);
            HIP_TRY(hipMemcpy(c->grid, c->grid_staging.data(), sizeof(int4) * (size_t)A, hipMemcpyHostToDevice));
        }
        c->prepared = false;
    }
    return ASV_OK;
}

int asv_connectivity(asv_ctx* c, int32_t supsz, int32_t* num_labels, void* stream)
{
    int r = check_ready(c);
    if (r) return r;
    if (c->last_iter < 1) return fail(ASV_ERR_STATE, "connectivity before any Lloyd iteration");
    const int64_t V = c->V();
    const int sz = supsz > 0 ? supsz : (int)(V / (c->p.seeds_num > 0 ? c->p.seeds_num : 1));   // Image.cpp:1833
    int32_t n = 0;
    hipError_t e = connect_labels(S(stream), c->labels[c->cur], c->labels[1 - c->cur], c->d.W, c->d.H, c->d.D, sz, &n);
    if (e != hipSuccess) return fail(ASV_ERR_HIP, "connectivity: %s", hipGetErrorString(e));
    c->cur = 1 - c->cur;
    if (num_labels) *num_labels = n;
    return ASV_OK;
}

int asv_connectivity_labels(const int32_t* in, int32_t* out, int32_t width, int32_t height, int32_t depth,
                            int32_t supsz, int32_t* num_labels, void* stream)
{
    if (!in || !out || in == out || width < 1 || height < 1 || depth < 1 || supsz < 0)
        return fail(ASV_ERR_ARG, "bad connectivity arguments");
    if ((int64_t)width * height * depth >= 0x7fffffffLL) return fail(ASV_ERR_ARG, "volume too large for int32 indices");
    int32_t n = 0;
    hipError_t e = connect_labels(S(stream), in, out, width, height, depth, supsz, &n);
    if (e != hipSuccess) return fail(ASV_ERR_HIP, "connectivity: %s", hipGetErrorString(e));
    if (num_labels) *num_labels = n;
    return ASV_OK;
}

int asv_host_alloc(int64_t bytes, void** out)
{
    if (!out || bytes <= 0) return fail(ASV_ERR_ARG, "bad host allocation request");
    *out = nullptr;
    hipError_t e = hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault);
    if (e != hipSuccess) return fail(ASV_ERR_NOMEM, "pinned host allocation failed: %s", hipGetErrorString(e));
    return ASV_OK;
}

int asv_host_free(void* p)
{
    if (p) (void)hipHostFree(p);
    return ASV_OK;
}

int asv_download_sums(asv_ctx* c, int64_t* sums, void* stream)
{
    if (!c || !sums) return fail(ASV_ERR_ARG, "null argument");
    if (c->last_iter < 1) return fail(ASV_ERR_STATE, "no run has completed");
    const size_t bytes = (size_t)(c->d.A + 1) * kSumWords * 8;
    HIP_TRY(hipMemcpyAsync(c->sums_host, c->sums[c->last_iter & 1], bytes, hipMemcpyDeviceToHost, S(stream)));
    HIP_TRY(hipStreamSynchronize(S(stream)));
    std::memcpy(sums, c->sums_host, bytes);
    return ASV_OK;
}

int asv_debug_fail_alloc(int64_t min_bytes)
{
    if (min_bytes < 0) return fail(ASV_ERR_ARG, "negative allocation threshold");
    g_fail_alloc_from.store(min_bytes, std::memory_order_relaxed);
    return ASV_OK;
}

int asv_force_exact_clamp(asv_ctx* c, int32_t on)
{
    if (!c) return fail(ASV_ERR_ARG, "null context");
    c->force_exact = on != 0;
    return ASV_OK;
}

// Marching-step counters (asv_enable_march_stats): a row of 16 words (13
// used) per step size 1..16.
constexpr int kMarchStatWords = 16 * 16;

int asv_enable_march_stats(asv_ctx* c, int32_t on)
{
    if (!c) return fail(ASV_ERR_ARG, "null context");
    if (on && !c->march_stats) {
        hipError_t e = asv::dev_malloc((void**)&c->march_stats, kMarchStatWords * sizeof(unsigned long long));
        if (e != hipSuccess) { c->march_stats = nullptr; return fail(ASV_ERR_NOMEM, "counter allocation failed"); }
        c->bytes += kMarchStatWords * 8;
    }
    if (on) {
        HIP_TRY(hipMemsetAsync(c->march_stats, 0, kMarchStatWords * sizeof(unsigned long long), nullptr));
        HIP_TRY(hipStreamSynchronize(nullptr));
    } else if (c->march_stats) {
        HIP_TRY(hipDeviceSynchronize());
        (void)hipFree(c->march_stats);
        c->march_stats = nullptr;
        c->bytes -= kMarchStatWords * 8;
    }
    return ASV_OK;
}

int asv_download_march_stats(asv_ctx* c, uint64_t* out, void* stream)
{
    if (!c || !out) return fail(ASV_ERR_ARG, "null argument");
    if (!c->march_stats) return fail(ASV_ERR_STATE, "marching-step counters not enabled");
    HIP_TRY(hipMemcpyAsync(out, c->march_stats, kMarchStatWords * sizeof(unsigned long long), hipMemcpyDeviceToHost, S(stream)));
    HIP_TRY(hipStreamSynchronize(S(stream)));
    return ASV_OK;
}

int asv_enable_trace(asv_ctx* c, int32_t on)
{
    if (!c) return fail(ASV_ERR_ARG, "null context");
    if (on && !c->trace) {
        size_t n = (size_t)c->p.lloyd_iters * c->d.A * 12;
        hipError_t e = asv::dev_malloc((void**)&c->trace, n * sizeof(float));
        if (e != hipSuccess) { c->trace = nullptr; return fail(ASV_ERR_NOMEM, "trace allocation failed"); }
        c->bytes += (int64_t)(n * sizeof(float));
    }
    c->trace_on = on != 0;
    return ASV_OK;
}

int asv_download_trace(asv_ctx* c, float* trace, void* stream)
{
    if (!c || !trace) return fail(ASV_ERR_ARG, "null argument");
    if (!c->trace) return fail(ASV_ERR_STATE, "trace not enabled");
    size_t n = (size_t)c->p.lloyd_iters * c->d.A * 12;
    HIP_TRY(hipMemcpyAsync(trace, c->trace, n * sizeof(float), hipMemcpyDefault, S(stream)));
    HIP_TRY(hipStreamSynchronize(S(stream)));
    return ASV_OK;
}

int asv_download_lab(asv_ctx* c, uint32_t* lab, void* stream)
{
    if (!c || !lab) return fail(ASV_ERR_ARG, "null argument");
    if (!c->prepared) return fail(ASV_ERR_STATE, "asv_prepare has not run");
    HIP_TRY(hipMemcpyAsync(lab, c->lab, (size_t)c->V() * 4, hipMemcpyDefault, S(stream)));
    HIP_TRY(hipStreamSynchronize(S(stream)));
    return ASV_OK;
}

int asv_download_metric(asv_ctx* c, float* m8, void* stream)
{
    if (!c || !m8) return fail(ASV_ERR_ARG, "null argument");
    if (!c->prepared) return fail(ASV_ERR_STATE, "asv_prepare has not run");
    hipStream_t st = S(stream);
    float4* field = nullptr;
    HIP_TRY(asv::dev_malloc((void**)&field, (size_t)c->V() * 32));
    hipError_t e = launch_metric_field(st, c->d, c->fwd, c->bwd, field);
    if (e == hipSuccess) e = hipMemcpyAsync(m8, field, (size_t)c->V() * 32, hipMemcpyDefault, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    const hipError_t ef = hipFree(field);
    if (e == hipSuccess) e = ef;
    if (e != hipSuccess) return fail(ASV_ERR_HIP, "metric field: %s", hipGetErrorString(e));
    return ASV_OK;
}

}  // extern "C"
