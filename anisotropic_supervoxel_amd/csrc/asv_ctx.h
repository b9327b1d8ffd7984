// asv_ctx.h -- the clip context behind the C ABI's opaque asv_ctx, shared
// by the library's own translation units (asv_capi.cpp, asv_batch.cpp).
#pragma once

#include "../../include/asv.h"
#include "asv_internal.h"

#include <vector>

struct asv_ctx {
    asv_params p{};
    asv::Dims d{};
    int s = 0;
    float ratio = 0.0f;
    std::vector<int> steps;
    std::vector<float> mult;
    std::vector<int> grid_host;
    std::vector<int> grid_ref;      // the reference grid (asv_create), empty for seeded contexts
    int gdims[3] = {0, 0, 0};       // its dNum, hNum, wNum
    std::vector<int4> grid_staging;
    int64_t bytes = 0;
    bool have_frames = false, have_flow = false, prepared = false;
    // Leading frames whose LAB matches the frames now on the device (a
    // frame's LAB depends on that frame alone), and of them the frames that
    // asv_shift_frames moved with their LAB since the last asv_prepare,
    // which converts only the frames after those (every other prepare
    // converts the whole volume)
    int lab_frames = 0;
    int lab_kept = 0;
    bool trace_on = false;
    int cur = 0;            // index of the label buffer holding the latest labels
    bool fresh = false;     // labels[cur] holds only the seed stage's scattered seeds
    int last_iter = 0;

    uint8_t* rgb = nullptr;
    float2* fwd = nullptr;
    float2* bwd = nullptr;
    float* lut = nullptr;
    uint32_t* lab = nullptr;
    int32_t* labels[2] = {nullptr, nullptr};
    int4* grid = nullptr;
    int4* pos = nullptr;
    float4* slab = nullptr;
    asv::SeedRec* rec = nullptr;
    float4* seed_own = nullptr;     // a seed's own (non-PSD) metric for the ring fallback
    int* seed_need = nullptr;       // 1 where the ring fallback runs
    unsigned long long* sums[2] = {nullptr, nullptr};
    float* trace = nullptr;
    int32_t* map_table = nullptr;
    int map_cap = 0;
    long long* post_scratch = nullptr;
    int64_t post_cap = 0;
    unsigned long long* ck_word = nullptr;
    unsigned long long* keys = nullptr;     // V keys for the scatter form of the sparse second pass
    asv::SeedBricks bricks;                 // brick lists of the fused first two passes (launch_seed_bricks)
    int* unsafe = nullptr;                  // raised when a seed metric needs the exact clamp (JfaClip::unsafe)
    bool force_exact = false;               // asv_force_exact_clamp: raise it at every prepare
    unsigned long long* march_stats = nullptr;   // asv_enable_march_stats: JfaClip::stats
    // page-locked host words: (seed_cap + 1) x kSumWords sums, then the
    // checksum word; every device-to-host copy of sums and checksums lands
    // here first (one explicit transfer into memory the context owns)
    unsigned long long* sums_host = nullptr;

    int seed_cap = 0;       // seeds the buffers hold (asv_reseed)
    int depth_cap = 0;      // frames the volume buffers hold
    int64_t V() const { return (int64_t)d.W * d.H * d.D; }

    // How pass i of an iteration that began with the seed stage runs: 1 the
    // first pass over the scattered seeds alone, 2 the second pass over the
    // first's output when it is sparse (about 729 voxels per seed against
    // a third of the volume), 0 the full marching/tiled pass; 3 and 4 where
    // the two sparse passes run fused from the seed positions (the first
    // pass launches both and writes the second's output, the second
    // launches nothing).
    int pass_kind(size_t i) const
    {
        if (i < 2 && fused_seed_passes()) return i == 0 ? 3 : 4;
        return asv::seed_pass_kind(steps, d.A, V(), i);
    }

    // Whether the first two passes run as launch_seed_bricks: the lowest-id
    // rule, a first step of at most kSeedFuseMaxK1, a reach the brick lists
    // were sized for and a sparse second pass.
    bool fused_seed_passes() const
    {
        return bricks.entries != nullptr && p.tie_rule == ASV_TIE_LOWEST_ID && steps.size() > 1 &&
               steps[0] <= asv::kSeedFuseMaxK1 && steps[0] + steps[1] <= bricks.r_cap &&
               asv::seed_pass_kind(steps, d.A, V(), 1) == 2;
    }

    // 16-bit labels between marching passes (launch_jfa's lt): pass i of a
    // flood of n passes writes them where the next pass of the same flood
    // marches in a form that reads them, so only the flood's last pass
    // writes 32-bit labels.  Passes of step 1-5 read 16-bit labels (the
    // schedules' steps after the second pass; the forms are not built for
    // larger steps).  `fresh_iter`: the passes of an iteration that began
    // with the seed stage (pass_kind), else every pass is a full one.
    bool pass_label16(size_t i, bool fresh_iter) const
    {
        return (fresh_iter ? pass_kind(i) : 0) == 0 &&
               asv::march_label16(d, steps[i], p.tie_rule, march_stats != nullptr);
    }
    // Whether pass i's output is 16-bit: a marching pass followed by one that
    // reads 16-bit labels, or the fused first passes' second (written by the
    // first's launch, pass_kind 4) followed by one.
    bool writes_label16(size_t i, size_t n, bool fresh_iter) const
    {
        if (i + 1 >= n || steps[i + 1] > 5 || !pass_label16(i + 1, fresh_iter)) return false;
        const int k = fresh_iter ? pass_kind(i) : 0;
        return k == 4 || (k == 0 && pass_label16(i, fresh_iter));
    }
    int label_widths(size_t i, size_t n, bool fresh_iter) const
    {
        const bool out16 = writes_label16(i, n, fresh_iter);
        const bool in16 = i > 0 && i < n && writes_label16(i - 1, n, fresh_iter);
        return in16 ? (out16 ? 2 : 3) : (out16 ? 1 : 0);
    }

    // Whether the first pass must fill its output with -1: not when the
    // second pass is the scatter form, which reads the first pass's output
    // only at the voxels the first pass writes (the seeds' step-1 sets).
    bool first_pass_fill() const
    {
        return !(pass_kind(1) == 2 && p.tie_rule == ASV_TIE_LOWEST_ID && keys != nullptr);
    }
};
