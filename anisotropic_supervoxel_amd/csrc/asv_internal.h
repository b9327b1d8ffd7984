// asv_internal.h -- device data layout and kernel launchers shared by the
// C ABI (asv_capi.cpp) and the kernels (asv_kernels.hip).  Not installed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace asv {

// Every device allocation of the library goes through here: hipMalloc,
// unless a test has asked allocations of at least some size to fail
// (asv_debug_fail_alloc), in which case it returns hipErrorOutOfMemory.
hipError_t dev_malloc(void** p, size_t bytes);

// One seed as the flood reads it: 48 bytes, three 16-byte loads.
//   a = (x, y, z, m0)   voxel coordinates as exact floats
//   b = (m1, m2, m3, m4)
//   c = (m5, L, a, b)
struct alignas(16) SeedRec {
    float4 a, b, c;
};

// One clip's share of a JFA pass launch (the marching kernel covers one or
// several clips; work items item_begin .. of this clip are its 16 x 16
// columns x K residues x nseg segments of up to seg slices).
struct JfaClip {
    const int32_t* in;
    int32_t* out;
    const uint32_t* lab;
    const SeedRec* rec;
    int W, H, D, A;
    int tiles_x, tiles_y, seg, nseg;
    int item_begin;
    float ratio;
    // Raised by the seed stage when some seed's metric could make a
    // quadratic form NaN (launch_seed_metric); the candidate rounds clamp
    // with one integer maximum, exact only without NaN, so the clip's
    // columns are then evaluated tap by tap with the exact clamp.
    const int* unsafe;
    // Optional step counters (null: none kept), summed over the launch's
    // blocks: steps, wide steps, chunks of wide steps, per-tap steps.
    unsigned long long* stats;
};

// Per-slot centroid sums: sum(2x+1), sum(2y+1), sum(2z+1), n, sum L, a, b, pad.
constexpr int kSumWords = 8;

// Geometry of one clip.
struct Dims {
    int W, H, D;
    int A;                  // seeds placed
};

// Launchers (all asynchronous on `st`).  Return hipSuccess or the launch error.
hipError_t launch_prepare(hipStream_t st, const Dims& d, const uint8_t* rgb, int prepared, const float* lin_lut,
                          uint32_t* lab);
// The metric of every voxel into field (2 float4 per voxel: m0..m3, m4, m5,
// psd flag, 0), for tests and tools.
hipError_t launch_metric_field(hipStream_t st, const Dims& d, const float2* fwd, const float2* bwd, float4* field);
// Colour-gradient density of every grid seed's box (GPU-AS/Image.cpp:350-560):
// acc[slot] += 4096 * g over the voxels whose float16 grid id addresses
// slot (< nslot); mx / my / mz map x / y / z to the grid index along each
// axis.
hipError_t launch_density(hipStream_t st, const Dims& d, const uint8_t* rgb, int prepared, const int* mx,
                          const int* my, const int* mz, int wnum, int hnum, int nslot, unsigned long long* acc);
hipError_t launch_seed_init(hipStream_t st, const Dims& d, const int4* grid, const uint32_t* lab,
                            int4* pos, float4* slab);
// Seed metrics and records: the metric at each seed's voxel, ring fallbacks
// for the seeds where it is not PSD (own: 2 float4 per seed, need: 1 int
// per seed, scratch of the launch).  A metric with an entry that is not
// finite or exceeds 2^60 in magnitude sets *unsafe to 1 (the marching pass
// then evaluates tap by tap, see JfaClip::unsafe); the caller clears it.
hipError_t launch_seed_metric(hipStream_t st, const Dims& d, const int4* pos, const float4* slab, const float2* fwd,
                              const float2* bwd, int gate, SeedRec* rec, float4* own, int* need, float* trace_iter,
                              int* unsafe);
hipError_t launch_fill(hipStream_t st, int32_t* p, int64_t n, int32_t value);
hipError_t launch_zero_u64(hipStream_t st, uint64_t* p, int64_t n);
hipError_t launch_zero_i32(hipStream_t st, int* p, int64_t n);
hipError_t launch_place(hipStream_t st, const Dims& d, const int4* pos, int32_t* labels);
// lt: the label widths of a marching pass (0: 32-bit in and out; 1 32-bit
// in, 16-bit out; 2 16-bit in and out; 3 16-bit in, 32-bit out; 16-bit
// labels are 0xFFFF for -1), only where march_label16 allows them, and
// 16-bit input only at steps 1-5.
hipError_t launch_jfa(hipStream_t st, const Dims& d, const int32_t* in, int32_t* out,
                      const uint32_t* lab, const SeedRec* rec, int k, float ratio, int tie_rule, const int* unsafe,
                      unsigned long long* stats = nullptr, int lt = 0);
// Whether launch_jfa runs a pass of step k over d as a marching kernel
// with 16-bit label forms: the lowest-id rule, the direct id table (at most
// kMarchTableIds seeds, so every id fits 16 bits), two slices per step, an
// uncounted launch, and a clip small enough that its one-clip plan holds.
bool march_label16(const Dims& d, int k, int tie_rule, bool counted);
// The first pass of an iteration when `in` holds only the seeds scattered
// by the seed stage at `pos` (every other voxel -1): fills `out` with -1
// and evaluates only the voxels that have a seed among their taps, at most
// max(3, k + 1)^3 per seed; used for steps k <= kSeedPassMaxStep.
constexpr int kSeedPassMaxStep = 16;
// The second pass (step k2) over the first pass's output (step k1 from the
// scattered seeds): evaluates only the voxels that can have a labelled tap,
// at most (3 (k2 + 1))^3 per seed (9^3 inside), each seed's own voxel among
// them.  `out` must hold -1 everywhere but at the seed voxels: the seed
// stage's volume, which is where the schedule's second pass writes.
// With `keys` (V 64-bit words, all at the initial key, as
// launch_fill_keys leaves them and this pass returns them) the lowest-id
// rule runs by scatter: one atomic minimum per (labelled tap, voxel) pair.
hipError_t launch_jfa_seeds2(hipStream_t st, const Dims& d, const int32_t* in, int32_t* out, const uint32_t* lab,
                             const SeedRec* rec, const int4* pos, int k1, int k2, float ratio, int tie_rule,
                             unsigned long long* keys);
hipError_t launch_fill_keys(hipStream_t st, unsigned long long* keys, int64_t n);

// The first two passes of an iteration (steps k1 <= kSeedFuseMaxK1 and
// k2 <= kSeedPassMaxStep, lowest-id rule) straight from the seed positions,
// over bricks of 32 x 16 x kSeedBrickZ voxels (asv_seedpass.hip): `out`
// receives the second pass's labels for every voxel, so neither the seed
// stage's volume nor the first pass's output is needed.  The brick buffers
// (alloc_seed_bricks, for clips of up to W x H x D voxels, `seeds` seeds
// and a reach k1 + k2 of at most R): per-brick counts (zero between uses),
// offsets and cursors, and the entries of every brick's seed list.
constexpr int kSeedFuseMaxK1 = 2;
constexpr int kSeedBrickZ = 16;
struct SeedBricks {
    int* count = nullptr;
    int* off = nullptr;
    int* cursor = nullptr;
    void* entries = nullptr;
    int64_t nb_cap = 0;
    int seed_cap = 0;
    int r_cap = 0;
};
void seed_brick_grid(int W, int H, int D, int* nbx, int* nby, int* nbz);
int64_t seed_brick_bytes(int W, int H, int D, int64_t seeds, int R);
hipError_t alloc_seed_bricks(SeedBricks* bb, int W, int H, int D, int seeds, int R, int64_t* total);
void free_seed_bricks(SeedBricks* bb);
// (out16: 16-bit output labels, 0xFFFF for -1, for a marching pass that
// reads them; at most kMarchTableIds seeds)
hipError_t launch_seed_bricks(hipStream_t st, const Dims& d, const int4* pos, const SeedRec* rec, const uint32_t* lab,
                              int32_t* out, int k1, int k2, float ratio1, float ratio2, const SeedBricks& bb,
                              bool out16 = false);

// How pass i of an iteration that began with the seed stage runs, for a
// schedule `steps`, A seeds and V voxels: 1 the first pass over the
// scattered seeds alone, when their neighbourhoods (27 voxels per seed at
// step 1) are at most a sixteenth of the volume (denser seeds make the full
// pass cheaper: 10-frame windows at 250 voxels per seed, 640x480 31 + 6 us
// fill against 23 us, 1280x720 78 + 7 against 67); 2 the second pass over
// the first's output when that is sparse (about 729 voxels per seed against
// a third of the volume); 0 the full marching/tiled pass.
inline int seed_pass_kind(const std::vector<int>& steps, int64_t A, int64_t V, size_t i)
{
    if (steps.empty() || steps[0] > kSeedPassMaxStep || A * 27 * 16 > V) return 0;
    if (i == 0) return 1;
    if (i == 1 && steps.size() > 1 && steps[1] <= kSeedPassMaxStep && A * 729 * 3 <= V) return 2;
    return 0;
}
hipError_t launch_jfa_seeds(hipStream_t st, const Dims& d, const int32_t* in, int32_t* out, const uint32_t* lab,
                            const SeedRec* rec, const int4* pos, int k, float ratio, int tie_rule, bool fill = true);
// The marching pass over several clips at once (step k <= 16, lowest-id
// tie; more than kMarchTableIds seeds share table slots, see k_jfa_march).  march_plan fills seg, nseg and
// item_begin of the n host descriptors (per_block > 0: about per_block
// items per resident block; 0: the segment count with the least predicted
// time; grid_limit > 0 caps the grid, seg_cap > 0 the slices per segment),
// sets the launch's item count and dynamic LDS bytes, and returns the grid
// size (0: the marching pass does not apply); the descriptors then go to
// device memory for launch_march_batch.  march_max_grid is the resident
// block count for `dyn` dynamic LDS bytes.
constexpr int kMarchTableIds = 16384;       // byte-table slots (dynamic LDS <= 16 KiB)
// Columns of the marching pass at step k: 16 rows of march_tile_w(k) voxels,
// one thread per voxel of a slice.  From step 4 the columns are 32 wide,
// which stages fewer halo cells per voxel (the halo is k on each side):
// measured at 640x480x200 and 320x240x85, step 4 145 -> 139 us, step 5
// 393 -> 379, step 7 210 -> 181, while step 2 went 319 -> 350 and step 3
// 472 -> 504 (longer candidate lists per column) and step 1 barely moved.
constexpr int march_tile_w(int k) { return k >= 4 ? 32 : 16; }
constexpr int kMarchTileH = 16;
constexpr int kMarchMaxIds = (1 << 30) - 1;  // ids below the resolved-index tag
// ht: some clip of the launch has more than kMarchTableIds seeds (the
// kernel whose table slots are shared and checked); counted: some clip has
// step counters (JfaClip::stats; the counting kernel).
// dense: the launch's clips hold few voxels per seed (march_dense), which
// selects the wider candidate masks (march_mc).
bool march_dense(const JfaClip* c, int n);
int march_mc(int k, bool dense = false);
int march_max_grid(int k, int dyn, bool ht = false, bool dense = false, int lt = 0);
int march_plan(int k, JfaClip* c, int n, int per_block, int grid_limit, int seg_cap, int* nitems, int* dyn,
               int lt = 0);
hipError_t launch_march_batch(hipStream_t st, int k, const JfaClip* dev, int n, int grid, int nitems, int dyn,
                              bool ht, bool counted = false, bool dense = false, int lt = 0);

hipError_t launch_reduce(hipStream_t st, const Dims& d, const int32_t* labels, const uint32_t* lab,
                         unsigned long long* sums);
hipError_t launch_finalize(hipStream_t st, const Dims& d, const unsigned long long* sums,
                           unsigned long long* zero_next, const uint32_t* lab, int4* pos,
                           float4* slab);

hipError_t launch_map_labels(hipStream_t st, const int32_t* in, int32_t* out, int64_t n, const int32_t* table,
                             int n_table);

// *out += sum over i < n of splitmix64(((index0 + i) << 32) | (uint32)labels[i]).
hipError_t launch_checksum(hipStream_t st, const int32_t* labels, int64_t n, int64_t index0, unsigned long long* out);

// Device post-processing: mode 0 fills unlabelled voxels from the previous
// labelled voxel after mapping through `table` (label_replace composition),
// mode 1 is the dead-seed rank remap.  `scratch` holds post_scratch_words(V)
// 64-bit words.  out must not alias in.
int64_t post_scratch_words(int64_t V);
hipError_t launch_postprocess(hipStream_t st, const int32_t* in, int32_t* out, int64_t V, int mode,
                              const int32_t* table, int n_table, long long* scratch);

// 3D connectivity enforcement (GPU-AS/Image.cpp:1938-2062) from V = W*H*D
// labels `in` into `out` (device, in != out): components of the same-label
// 24-neighbourhood smaller than supsz / 4 voxels take the adjacent label,
// the others are numbered from 0 in raster order of their first voxel.
// Allocates its own scratch, synchronises st; *num_labels = the count.
hipError_t connect_labels(hipStream_t st, const int32_t* in, int32_t* out, int W, int H, int D, int supsz,
                          int32_t* num_labels);

}  // namespace asv
