/*
 * asv.h -- C ABI of the MI355X supervoxel hot path (libasv.so).
 *
 * Plain C types only: host pointers, sizes, an opaque context and an opaque
 * stream handle (a hipStream_t passed as void*, NULL for the null stream).
 * Every entry point returns an int status (0 = ASV_OK, negative = error) and
 * never throws; asv_last_error() gives the message of the calling thread's
 * last failure.  One context serves one clip (or one stream window) on the
 * current HIP device; a context must not be used by two threads at once.
 *
 * The reference drives the same stages through OpenGL from class Image
 * (GPU-AS/Image.h:71-88), called in order by Render::run
 * (GPU-AS/Render.cpp:254-319).  Each entry point names the reference method
 * it replaces.
 */
#ifndef ASV_H
#define ASV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ASV_OK = 0,
    ASV_ERR_ARG = -1,      /* bad argument or size mismatch */
    ASV_ERR_HIP = -2,      /* a HIP runtime call failed */
    ASV_ERR_STATE = -3,    /* call out of order (e.g. run before upload) */
    ASV_ERR_NOMEM = -4,    /* device allocation failed */
    ASV_ERR_ADMIT = -5,    /* serving: footprint exceeds the memory budget */
    ASV_ERR_CANCELLED = -6 /* serving: the request was cancelled between windows */
};

enum { ASV_TIE_LOWEST_ID = 0, ASV_TIE_SCAN_ORDER = 1 };
enum { ASV_SCHED_REFERENCE = 0, ASV_SCHED_HALVING_PLUS_ONE = 1 };
enum { ASV_INIT_GRID = 0, ASV_INIT_DENSITY = 1 };

typedef struct asv_params {
    int32_t width, height, depth;   /* W, H, D (frames) */
    int32_t seeds_num;              /* K requested; grid places A <= K */
    int32_t lloyd_iters;            /* Lloyd iterations (reference default 20) */
    float   compactness;            /* 8 (GPU-AS/Image.cpp:1050) or 6 (streamGPU-AS/Image.cpp:968) */
    int32_t tie_rule;               /* ASV_TIE_* */
    int32_t angle_gate;             /* 1: flow-angle gate on ring fallbacks from iteration 2
                                       (GPU-AS/Image.cpp:1404-1421); 0: none (streamGPU-AS) */
    int32_t schedule;               /* ASV_SCHED_* */
    int32_t frames_prepared;        /* 0: raw frames, image row order (row 0 = top);
                                       1: already blurred and in GL row order */
    int32_t seed_init;              /* ASV_INIT_GRID: the regular grid (GPU-AS/Image.cpp:20-92);
                                       ASV_INIT_DENSITY: the grid re-placed by colour-gradient
                                       density (Image::initialization, :350-620), which the
                                       reference's Render::run always applies (Render.cpp:258) */
} asv_params;

typedef struct asv_ctx asv_ctx;

/* Fills p with the reference's defaults for W x H x D and K seeds. */
void asv_default_params(asv_params* p, int32_t width, int32_t height, int32_t depth, int32_t seeds_num);

/* Image::init_gpu + Image::init_seeds + init_variables
 * (GPU-AS/Image.cpp:320-345, :20-158, :1047-1053): validates p, places the
 * grid seeds and allocates every device buffer of the clip. */
int asv_create(const asv_params* p, asv_ctx** out);
int asv_destroy(asv_ctx* ctx);

/* Number of seeds placed (ac_seeds_num, GPU-AS/Image.cpp:45) and grid step s
 * (GPU-AS/Image.cpp:32). */
int asv_seeds_placed(const asv_ctx* ctx);
int asv_grid_step(const asv_ctx* ctx);
/* Resident device bytes held by the context. */
int64_t asv_device_bytes(const asv_ctx* ctx);
/* Device bytes asv_create would allocate for p (admission control). */
int64_t asv_footprint(const asv_params* p);

/* JFA pass list (GPU-AS/Image.cpp:1061-1181): writes up to cap step lengths
 * and ratio multipliers, returns the number of passes P. */
int asv_schedule(const asv_ctx* ctx, int32_t* steps, float* mult, int32_t cap);

/* Host-only planning (no device needed): seeds placed A, grid step s and the
 * JFA pass list for p, from the same rules asv_create applies.  Returns the
 * pass count P, or a negative status. */
int asv_plan(const asv_params* p, int32_t* seeds_placed, int32_t* grid_step,
             int32_t* steps, float* mult, int32_t cap);

/* Host-only: the grid seed voxels asv_create would place for p (A x 3
 * int32), GPU-AS/Image.cpp:20-92; returns A. */
int asv_plan_seeds(const asv_params* p, int32_t* xyz, int32_t cap);

/* A context whose initial seeds are given (n x 3 voxel coordinates, GL
 * orientation, ids 0..n-1 in that order) with grid step grid_step setting
 * ratio and the JFA schedule: the streaming blocks' kept ++ new seeds
 * (streamGPU-AS/Render.cpp:434-441, :470-504). */
int asv_create_seeded(const asv_params* p, const int32_t* xyz, int32_t n, int32_t grid_step, asv_ctx** out);

/* Reuse a context for a new window without reallocating: new initial seeds
 * (n <= the seed capacity, which is max(n, p->seeds_num) at creation by
 * asv_create_seeded), grid step and depth (<= the depth it was created
 * with).  No counterpart in the reference, which rebuilds its GL objects per
 * block (streamGPU-AS/Render.cpp:429, :566); serves the resident windows of
 * the streaming and served forms. */
int asv_reseed(asv_ctx* ctx, const int32_t* xyz, int32_t n, int32_t grid_step, int32_t depth, void* stream);
/* Device-side move of frames [src_frame, src_frame + n_frames) to [0, n_frames)
 * of the frame and flow buffers (the overlap kept between blocks,
 * streamGPU-AS/Render.cpp:545-556); ranges must not overlap.  When the
 * moved frames' LAB is current it moves too, and the next asv_prepare
 * converts only the frames after them. */
int asv_shift_frames(asv_ctx* ctx, int32_t src_frame, int32_t n_frames, void* stream);
/* Upload frames [first, first + count) of frames and both flows. */
int asv_upload_frames_range(asv_ctx* ctx, int32_t first, int32_t count, const uint8_t* rgb, const float* fwd,
                            const float* bwd, void* stream);

/* Grid seed voxels (A x 3 int32, x y z in GL orientation). */
int asv_grid_seeds(const asv_ctx* ctx, int32_t* xyz, int32_t cap);

/* Frame volume: D x H x W x 3 uint8 RGB (host or device pointer), replacing
 * the texture upload in Image::init_image_texture (GPU-AS/Image.cpp:198-219). */
int asv_upload_frames(asv_ctx* ctx, const uint8_t* rgb, void* stream);

/* Ground-truth optical flow, D x H x W x 2 float32 each, image row order,
 * (u, v) in pixels per frame with v pointing down; replaces the Farneback
 * fields of Image::forward_backward_flow (GPU-AS/Image.cpp:643-686).
 * Frame D-1 of fwd and frame 0 of bwd are ignored, as in the reference. */
int asv_upload_flow(asv_ctx* ctx, const float* fwd, const float* bwd, void* stream);

/* Image::RGB2LAB (GPU-AS/Image.cpp:296-315) with the frame preprocessing of
 * Render::cvMat2TexureInput (GPU-AS/Render.cpp:354-359): builds the LAB
 * volume.  The per-voxel matrix_slover metric (GPU-AS/Image.cpp:874-1009) is
 * evaluated by asv_seed_stage at the seeds and their rings, as the
 * reference does, so no per-voxel field is kept. */
int asv_prepare(asv_ctx* ctx, void* stream);

/* Image::Ld_first_init (GPU-AS/Image.cpp:1017-1045) for iteration 1 and
 * Image::Ld_next_seeds_matrix (GPU-AS/Image.cpp:1300-1575) afterwards: seed
 * metric selection (seed_matrix / matrix_slover at the seed voxel, then the
 * ring fallbacks, GPU-AS/Image.cpp:689-1009) and scatter into the label
 * volume. */
int asv_seed_stage(asv_ctx* ctx, int32_t iter, void* stream);

/* Image::flood (GPU-AS/Image.cpp:1061-1190): the P JFA passes. */
int asv_flood(asv_ctx* ctx, int32_t iter, void* stream);
/* The first npasses passes of iteration iter's schedule only, right after
 * its seed stage; asv_download_labels then returns pass npasses' output.
 * Where the first two passes run fused (lowest-id rule, sparse seeds),
 * npasses = 1 is refused.  Tests and tools; no counterpart in the
 * reference. */
int asv_flood_passes(asv_ctx* ctx, int32_t iter, int32_t npasses, void* stream);

/* Image::centroids_next_seeds (GPU-AS/Image.cpp:1238-1297): reduction, then
 * centroid finalisation unless iter is the last iteration. */
int asv_centroids(asv_ctx* ctx, int32_t iter, void* stream);

/* Whole run: asv_prepare, then for iter = 1..lloyd_iters the three stages
 * above (GPU-AS/Render.cpp:274-293), all enqueued on stream. */
int asv_run(asv_ctx* ctx, void* stream);
/* The Lloyd loop only (asv_prepare must have run on the current inputs). */
int asv_run_loop(asv_ctx* ctx, void* stream);

/* Final label volume (D x H x W int32, GL row order, -1 = unlabelled),
 * before any post-processing; replaces the label pass + glGetTexImage of
 * Image::check_dead_seeds (GPU-AS/Image.cpp:1625-1643).  Synchronises stream. */
int asv_download_labels(asv_ctx* ctx, int32_t* labels, void* stream);

/* Final reduction slots: (A+1) x 8 int64 = sum(2x+1), sum(2y+1), sum(2z+1),
 * count, sum L, sum a, sum b, 0; slot A collects label -1.  These replace the
 * slot texels read back at GPU-AS/Image.cpp:1585-1588. */
int asv_download_sums(asv_ctx* ctx, int64_t* sums, void* stream);

/* The first n_frames frames of the final labels mapped through table
 * (label l -> table[l], labels outside [0, n_table) -> 0) on the device,
 * then copied out: the block2all stitch of the streaming program
 * (streamGPU-AS/Render.cpp:559-563) without a full label download.  When
 * checksum is not NULL it receives the sum, modulo 2^64, over the copied
 * voxels i of splitmix64(((index0 + i) << 32) | (uint32_t)label), computed
 * on the device (an order-free fingerprint; index0 is the first voxel's
 * index in the whole video).  Synchronises stream. */
int asv_download_labels_mapped(asv_ctx* ctx, const int32_t* table, int32_t n_table, int32_t n_frames,
                               int32_t* labels, int64_t index0, uint64_t* checksum, void* stream);

/* check_dead_seeds (GPU-AS/Image.cpp:1579-1720) on the device: from the
 * final counts in sums ((A+1) x 8 as asv_download_sums writes them), the
 * labels of the context are rewritten as the reference's label_replace
 * (label_replace.gs:18-34) and sequence_label (:1727-1778) leave them;
 * later downloads read the result.  Outputs the dead-seed count and the
 * final label count (final_label_num).  Synchronises stream. */
int asv_postprocess_device(asv_ctx* ctx, const int64_t* sums, int32_t* dead_out, int32_t* final_label_num,
                           void* stream);

/* The same rewrite between two device buffers of V labels (in != out),
 * GL raster order, for a label volume produced elsewhere.  Synchronises. */
int asv_postprocess_labels(const int64_t* sums, int32_t A, const int32_t* in, int32_t* out, int64_t V,
                           int32_t* dead_out, int32_t* final_label_num, void* stream);
/* No counterpart in the reference: asv_postprocess_labels for callers that
 * cannot allocate device memory (a serving loop that must not call
 * hipMalloc, harnesses that forbid it).  Same arguments, kernels and
 * results; its scratch is a fixed 1.25 MiB array per device, held until the
 * process ends, so V is at most 2^29 voxels and A at most 65536 (larger
 * sizes return ASV_ERR_ARG).  Calls on one device share that array: calls
 * that overlap on one device (other threads, other streams) race.
 * Synchronises stream. */
int asv_postprocess_labels_noalloc(const int64_t* sums, int32_t A, const int32_t* in, int32_t* out, int64_t V,
                                   int32_t* dead_out, int32_t* final_label_num, void* stream);

/* Image::initialization (GPU-AS/Image.cpp:350-620): selects the initial
 * seed placement, ASV_INIT_GRID or ASV_INIT_DENSITY (the grid re-placed by
 * colour-gradient density), for the next asv_prepare; the same as creating
 * the context with p->seed_init.  Contexts created from explicit seeds
 * (asv_create_seeded, asv_reseed) have no reference grid and refuse
 * ASV_INIT_DENSITY. */
int asv_initialization(asv_ctx* ctx, int32_t seed_init);

/* 3D connectivity enforcement (Image::enforce_connectivity_3d,
 * GPU-AS/Image.cpp:1808-1837, over EnforceLabelConnectivity3D :1938-2062)
 * on the context's current labels (after a run: the raw labels of the last
 * JFA pass, which is what the reference reads; after
 * asv_postprocess_device: the post-processed ones).  Segments are the
 * connected components of equal labels under the reference's 24-voxel
 * neighbourhood; one of fewer than supsz / 4 voxels takes the adjacent
 * label, the others are numbered from 0 in raster order.  supsz <= 0 uses
 * the reference's V / K (K the requested supervoxel count).  The labels of
 * the context are rewritten (later downloads read the result);
 * *num_labels receives the number of segments kept.  Synchronises stream. */
int asv_connectivity(asv_ctx* ctx, int32_t supsz, int32_t* num_labels, void* stream);

/* The same between two device buffers of W x H x D labels (in != out), GL
 * raster order, for a label volume produced elsewhere.  Synchronises. */
int asv_connectivity_labels(const int32_t* in, int32_t* out, int32_t width, int32_t height, int32_t depth,
                            int32_t supsz, int32_t* num_labels, void* stream);

/* Page-locked host memory for frame and label buffers (faster transfers).
 * No counterpart in the reference; for callers that stage frames. */
int asv_host_alloc(int64_t bytes, void** out);
int asv_host_free(void* p);

/* The JFA distance clamps a negative quadratic form to 0 and rejects a NaN
 * one, as flood.gs:90-95 leaves a NaN to lose `curr_dist<min_dist`.  The
 * marching pass's candidate rounds clamp with one integer maximum, exact
 * while no form can be NaN; when a seed metric has an entry that is not
 * finite or exceeds 2^60 in magnitude, the seed stage switches the run's
 * marching passes to an exact per-tap path.  on = 1 takes that path in
 * every later run of the context (for tests and tools); no counterpart in
 * the reference. */
int asv_force_exact_clamp(asv_ctx* ctx, int32_t on);
/* Counters of the marching full passes, kept while enabled (on = 1 zeroes
 * them; on = 0 drops them), summed over every later run of the context.
 * out holds 16 rows of 16 words, row k - 1 for JFA step k: [0] steps (one
 * step: two slices of one 16 x 16 column of voxels), [1] wide steps (more
 * seeds in reach than one candidate mask holds, evaluated in chunks), [2]
 * chunks of the wide steps, [3] steps evaluated per tap; of the steps
 * within one mask: [4] candidate evaluations (voxel-slices x candidates),
 * [5] wave rounds (per 64-voxel wave and slice, its busiest voxel's
 * candidates), [6] wave-slices; of the steps that extend a table
 * generation, shader-clock cycles the waves spent in [7] claims, [8] index
 * conversion and record staging, [9] masks, [10] next-step load issue,
 * [11] evaluation and stores, and [12] those wave-steps; [13..15] 0.  Runs
 * with counters enabled use a counting form of the marching kernel (the
 * cycle counts are of that form).  Tests and tools only; no counterpart in
 * the reference. */
int asv_enable_march_stats(asv_ctx* ctx, int32_t on);
int asv_download_march_stats(asv_ctx* ctx, uint64_t* out, void* stream);
/* Fault injection for tests of failure handling: from now on every device
 * allocation of at least min_bytes fails as hipErrorOutOfMemory would
 * (0 restores normal allocation).  Process-wide; no counterpart in the
 * reference. */
int asv_debug_fail_alloc(int64_t min_bytes);
/* Per-iteration seed records, when enabled before the run:
 * lloyd_iters x A x 12 float (x, y, z, L, a, b, m0..m5). */
int asv_enable_trace(asv_ctx* ctx, int32_t on);
int asv_download_trace(asv_ctx* ctx, float* trace, void* stream);

/* LAB volume (D x H x W packed L | a<<8 | b<<16) and metric field
 * (D x H x W x 8 float: m0..m5, psd flag, 0, computed on request into a
 * temporary device buffer), for tests and tools. */
int asv_download_lab(asv_ctx* ctx, uint32_t* lab, void* stream);
int asv_download_metric(asv_ctx* ctx, float* m8, void* stream);

/* Timed run: same work as asv_run, with HIP events around every kernel
 * launch on stream.  ms[0] total, ms[1] prepare, ms[2] sum of the full JFA
 * passes, ms[3] sum of reductions, ms[4] seed stages + finalisation, ms[5]
 * sum of the passes that evaluate only the neighbourhoods of the seeds
 * (each iteration's first, and its second when that is sparse); passes
 * receives the count of full passes. */
int asv_run_timed(asv_ctx* ctx, void* stream, float* ms, int32_t* passes);

/* Synthetic video of the benchmark generator: background bilinear gradient,
 * n_ellipses moving ellipses, Gaussian noise, ground-truth forward and
 * backward flow and region ids (0 = background, e+1 = ellipse e).  Arrays are
 * D x H x W (x3 / x2), image row order.  Any output may be NULL. */
int asv_synth_video(uint64_t seed, int32_t width, int32_t height, int32_t depth,
                    int32_t n_ellipses, float vmax, float noise_sigma,
                    uint8_t* rgb, float* fwd, float* bwd, int32_t* region);

/* Host-only post-processing, Image::check_dead_seeds
 * (GPU-AS/Image.cpp:1579-1720) with label_replace.gs and sequence_label
 * (:1727-1778): sums as written by asv_download_sums, labels (V int32)
 * rewritten in place; outputs the dead-seed count and final_label_num. */
int asv_postprocess(const int64_t* sums, int32_t seeds_placed, int32_t* labels, int64_t voxels,
                    int32_t* dead_seeds, int32_t* final_label_num);

/* Image::print_sv_txt (GPU-AS/Image.cpp:1780-1806): the per-voxel label
 * text file (top image row first, "%-4d " per label, two blank lines per
 * frame). */
int asv_write_label_txt(const int32_t* labels, int32_t width, int32_t height, int32_t depth,
                        const char* path);

/* Image::print_render (GPU-AS/Image.cpp:1883-1935): pseudo-colour frames as
 * binary PPM files <prefix>NNNNN.ppm (top row first), rand() % 256 colours
 * after srand(1), black for -1. */
int asv_write_render_ppm(const int32_t* labels, int32_t width, int32_t height, int32_t depth,
                         int32_t n_labels, const char* prefix);

/* No counterpart in the reference (its evaluation used the external LIBSVX
 * tools, README.md:28): 3D undersegmentation error and 3D boundary recall
 * of labels against ground-truth regions of the same shape and row order. */
int asv_eval_metrics(const int32_t* labels, const int32_t* regions, int32_t width, int32_t height,
                     int32_t depth, double* undersegmentation, double* boundary_recall);

/* ---- Temporal blocks (the streamGPU-AS program) ---------------------- */

typedef struct asv_stream_params {
    int32_t width, height;
    int32_t total_frames;           /* frames of the whole video (depth_all) */
    int32_t seeds_num;              /* K for the whole video */
    int32_t seeds_per_100_frames;   /* 0: K*100/total_frames (streamGPU-AS/Render.cpp:156-158) */
    int32_t block_size;             /* 30 (streamGPU-AS/Render.cpp:199) */
    int32_t overlap;                /* 10, frames shared with the next block (:200) */
    int32_t lloyd_iters;            /* 20 (:252) */
    float   compactness;            /* 6 (streamGPU-AS/Image.cpp:968) */
    int32_t angle_gate;             /* 0: the streaming program has no gate */
    int32_t tie_rule;
    int32_t frames_prepared;
} asv_stream_params;

typedef struct asv_stream asv_stream;

void asv_stream_default_params(asv_stream_params* p, int32_t width, int32_t height, int32_t total_frames,
                               int32_t seeds_num);
/* Host-only: block intervals [start, end) (streamGPU-AS/Render.cpp:198-241);
 * returns the block count. */
int asv_stream_plan(const asv_stream_params* p, int32_t* intervals, int32_t cap);
int asv_stream_create(const asv_stream_params* p, asv_stream** out);
int asv_stream_destroy(asv_stream* s);
int asv_stream_block_count(const asv_stream* s);
int asv_stream_seeds_placed(const asv_stream* s);
/* After a block: {next block, A, kept seeds carried out, dead seeds,
 * new-seed budget, seeds placed so far}. */
int asv_stream_block_info(const asv_stream* s, int32_t* info6);
/* Block `block` (in order): prepare_data_seeds + block_segment +
 * post_and_next (streamGPU-AS/Render.cpp:348-569).  rgb/fwd/bwd hold the
 * block's frames only (image rows); its finished frames are written to
 * labels_all (W x H x total_frames, GL rows, global ids). */
int asv_stream_block(asv_stream* s, int32_t block, const uint8_t* rgb, const float* fwd, const float* bwd,
                     int32_t* labels_all, int32_t* frames_done, void* stream);
/* The same block, with its finished frames written from `out` onward
 * (frames_done x W x H labels), for callers that deliver results window by
 * window (page-locked `out` makes the copy one transfer). */
int asv_stream_block_into(asv_stream* s, int32_t block, const uint8_t* rgb, const float* fwd, const float* bwd,
                          int32_t* out, int32_t* frames_done, void* stream);
/* Checksum of every frame finished so far (stitched global labels, voxel
 * index over the whole video), as asv_download_labels_mapped defines it. */
int asv_stream_checksum(const asv_stream* s, uint64_t* out);
/* All blocks of a whole video (arrays cover total_frames frames). */
int asv_stream_run_video(const asv_stream_params* p, const uint8_t* rgb, const float* fwd, const float* bwd,
                         int32_t* labels_all, int32_t* seeds_placed, void* stream);
const char* asv_stream_last_error(void);

/* ---- Served streaming ------------------------------------------------- */
/* Batched clips: several clips (uneven sizes, one Lloyd iteration count)
 * run their loops together, the batched form of Render::run
 * (GPU-AS/Render.cpp:254-319) for many clips.  Every JFA pass index is one
 * persistent launch over the (clip, column) work items of all clips sharing
 * its step; each Lloyd iteration is captured once as a HIP graph and
 * replayed; an autotuner picks each launch's persistent block count and
 * work-item granularity on the first run.  Labels and sums of every clip are
 * bit-identical to a single-clip asv_run.  No counterpart in the reference
 * (one clip per process there); serves callers with many clips. */
typedef struct asv_batch asv_batch;
/* n clips with parameters params[0..n-1] (equal lloyd_iters); the batch
 * owns one context per clip. */
int asv_batch_create(const asv_params* params, int32_t n, asv_batch** out);
int asv_batch_destroy(asv_batch* b);
int asv_batch_size(const asv_batch* b);
/* Context of clip i, for asv_upload_frames / asv_upload_flow before a run and
 * asv_download_labels / asv_download_sums after it (do not destroy it). */
int asv_batch_clip(asv_batch* b, int32_t i, asv_ctx** ctx);
/* LAB + metric for every clip, then all Lloyd iterations, ordered after
 * prior work on stream; returns without waiting. */
int asv_batch_run(asv_batch* b, void* stream);
/* 0 turns the autotuner off (default shapes); before the first run. */
int asv_batch_set_autotune(asv_batch* b, int32_t on);
/* Launch groups per iteration (JFA launches: batched and per-clip). */
int asv_batch_launch_count(const asv_batch* b, int32_t* per_iteration);
/* Launch shapes in use: per group {pass index, step K, clips in the one
 * launch (or minus the number of clips launched one by one), grid, work
 * items per block}; returns the group count; *tune_ms = GPU time the
 * autotuner spent. */
int asv_batch_tuning(const asv_batch* b, int32_t* shapes, int32_t cap, float* tune_ms);
const char* asv_batch_last_error(void);

/* No single counterpart in the reference (its streaming program handles one
 * video, streamGPU-AS/Render.cpp:100-300); the served loop runs many streams'
 * temporal blocks concurrently on one GPU with admission by footprint. */

typedef struct asv_serve_params {
    int64_t memory_budget;      /* device bytes the admitted streams' windows may hold */
    int32_t max_active;         /* streams in flight (worker threads, one HIP stream each) */
    int32_t block_size;         /* window frames (10) */
    int32_t overlap;            /* frames shared by consecutive windows (5: stride 5) */
    int32_t lloyd_iters;        /* per window (5) */
    float   compactness;        /* 6 */
    int32_t angle_gate;         /* 0 */
    int32_t honor_arrivals;     /* 1: a request is eligible from arrival_ms of wall time */
    int32_t batch_windows;      /* 1: the windows of concurrent streams run their Lloyd loops
                                   together, one marching launch per pass index and step over all
                                   of them; 0 (default): each stream runs its own launches on its
                                   own HIP stream (measured faster on the served workload) */
} asv_serve_params;

typedef struct asv_serve_request {
    int32_t stream_id;
    int32_t width, height, frames;
    float   density;            /* supervoxels per 1000 pixels per frame */
    int32_t n_ellipses;
    uint64_t video_seed;        /* synthetic source of the stream's frames */
    float   vmax, noise_sigma;
    double  arrival_ms;
    /* the stream's frames and flows in host memory (frames x H x W x 3 / x 2,
     * image rows); when null they are generated from video_seed per window */
    const uint8_t* rgb;
    const float* fwd;
    const float* bwd;
    /* cancellation, checked before each window: the request ends with
     * ASV_ERR_CANCELLED once *cancel is non-zero (a flag the caller may set
     * from any thread; NULL: none) or once cancel_after_windows windows are
     * done (0: never) */
    const volatile int32_t* cancel;
    int32_t cancel_after_windows;
} asv_serve_request;

typedef struct asv_serve_window {
    int32_t stream_id, window, frames, seeds;
    double  ready_ms, start_ms, end_ms;     /* since the run started */
} asv_serve_window;

typedef struct asv_serve_result {
    int32_t stream_id;
    int32_t status;             /* ASV_OK, ASV_ERR_ADMIT (never fits the budget), ASV_ERR_ARG
                                   (malformed request), ASV_ERR_CANCELLED or the error that
                                   ended the stream; the other streams are unaffected */
    uint64_t labels_hash;       /* asv_stream_checksum of the stitched global labels */
    int64_t footprint;          /* device bytes admitted for the stream */
    int64_t voxel_iters;
    double  admitted_ms, done_ms;
} asv_serve_result;

void asv_serve_default_params(asv_serve_params* p);
int64_t asv_serve_footprint(const asv_serve_params* p, const asv_serve_request* r);
/* Runs the requests to completion.  Returns ASV_ERR_ARG for unusable
 * arguments, else ASV_OK: every request's own outcome is in its result (a
 * malformed, refused, cancelled or failing request ends alone, the others
 * run on), and the windows (up to window_cap) are recorded in the order
 * they finish. */
int asv_serve_run(const asv_serve_params* p, const asv_serve_request* reqs, int32_t n, asv_serve_result* results,
                  asv_serve_window* windows, int32_t window_cap, int32_t* n_windows, double* wall_ms);

/* Frames [t0, t1) of the synthetic video (same content as asv_synth_video). */
int asv_synth_video_range(uint64_t seed, int32_t width, int32_t height, int32_t depth, int32_t n_ellipses,
                          float vmax, float noise_sigma, int32_t t0, int32_t t1, uint8_t* rgb, float* fwd,
                          float* bwd, int32_t* region);

const char* asv_last_error(void);
int asv_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* ASV_H */
