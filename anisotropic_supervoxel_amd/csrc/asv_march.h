// asv_march.h -- the marching persistent JFA pass (flood.gs:26-182 for
// steps 1..16 under the lowest-id rule), a template over the step, the
// candidate capacity, the slices per step, the table form and counting.
// Its instances are compiled in four translation units (asv_march.hip,
// steps 1-4, 5-8, 9-12, 13-16) so that the build runs them in parallel.
#pragma once

#include "asv_device.h"

namespace asv {

// LDS id table of the marching pass: one byte per seed id, sized at launch
// for the largest seed count of the launch (dynamic shared memory); at most
// kListCap ids are listed per table generation.
constexpr int kListCap = 512;
constexpr unsigned kSlotEmpty = 0xFFu;
constexpr unsigned kSlotClaimed = 0xFEu;

// Candidate masks: MC compact indices per table generation, bit MC marks a
// tap without a seed (or past the table).
template <int MC> struct MarchMask;
template <> struct MarchMask<31> { using T = uint32_t; };
template <> struct MarchMask<63> { using T = unsigned long long; };

// A staged id after the claim: the compact index it resolved to, tagged
// with this bit (ids of the marching pass stay below 2^30).
constexpr int kResolved = 0x40000000;
// Table bytes of one generation: compact indices 0 .. kIdxCap - 1, and
// kIdxCap for every id listed past them.
constexpr unsigned kIdxCap = 0xFDu;
// The index byte of a cell without an id while a rebuild's list is wider
// than one mask (any value past kIdxCap).
constexpr unsigned kIdxNone = 0xFFu;

// Slices of residue r (z = r, r + K, ...) in segment g of a clip.
template <int K>
__device__ __forceinline__ int march_len(const JfaClip& c, int g, int r)
{
    if (r >= c.D) return 0;
    const int rest = (c.D - r + K - 1) / K - g * c.seg;
    return rest <= 0 ? 0 : (rest < c.seg ? rest : c.seg);
}

// Marching persistent pass for the lowest-id rule (K <= 16), also the
// batched form: one launch covers one clip
// (`one`, clips == nullptr) or the `nclips` descriptors of `clips`.
//
// Work items: a clip's 16 x 16 columns (tx, ty) x residues r = z mod K x
// segments g of up to `seg` slices z = r + K (g seg), r + K (g seg + 1), ...,
// ordered (g, r, ty, tx), tx fastest.  The items are split into 8
// contiguous ranges, one per XCD (block b runs on XCD b mod 8), and the
// blocks of one XCD take its items round-robin, so concurrently marching
// columns of an XCD are neighbours at about the same z and share halo rows
// in its L2.  The host picks the segment count and the grid so that every
// block of an XCD gets the same number of items (rounds).
//
// Moving from z to z + K keeps two of the three slabs a voxel's taps read
// (z and z + K become the new z - K and z), so each step stages, claims and
// converts only the new slab clamp(z + 2K):
//  - the direct-mapped byte table slot[id] and the candidate list persist
//    along a column (an id claimed earlier keeps its compact index); when a
//    step would take the list past MC candidates, the table is released and
//    rebuilt from the three current slabs;
//  - every thread keeps, in registers, the 9-tap candidate masks of its
//    voxel position in the three slabs, and computes only the new slab's;
//    the voxel's 27-tap mask is the OR of the three;
//  - the candidates' records sit in LDS as structure-of-arrays, with the
//    three dz terms of the distance (dz 2m2, dz 2m4, dz^2 m5) recomputed
//    once per slice for the whole block;
//  - two barriers per step (claims done; new slab and records staged).
// With S = 2 a step evaluates slices z and z + K from four slab masks and
// stages the two new slabs z + 2K and z + 3K, so the barriers, claims and
// record staging are paid once per two slices.
// A rebuild that still leaves more than MC candidates (a wide step) keeps
// their full compact indices in the index bytes and evaluates the step in
// chunks of MC list entries: each chunk stages its records, builds the slab
// masks of the indices it holds and lowers every voxel's running minimum,
// so each candidate is still evaluated once per voxel and slice.  A list
// past kIdxCap entries evaluates the step per tap from global labels and
// records.
//
// The byte table has kMarchTableIds slots.  HT (a launch with a clip of
// more seeds than slots): id i takes slot i mod kMarchTableIds, so a clip
// with more seeds than slots can have two ids of one slot in a block's
// region; for such a clip every looked-up index is checked against the
// candidate list after the claims, ids are not resolved during the claim,
// and a mismatch makes the step evaluate per tap and the next step rebuild
// (so no later mask holds a cell whose id was dropped).
//
// ST: the counting form, which adds each item's steps, wide steps, their
// chunks and per-tap steps to c.stats (when set); the kernels of ordinary
// launches carry none of that code.
//
// I16 / O16: the input / output labels are 16-bit (0xFFFF for -1), for
// clips of at most kLabel16Ids seeds between two marching passes: the
// passes move 2 bytes per label instead of 4 (the slabs with their halos,
// and the output), and only the last pass of a flood writes 32-bit labels.
template <int K, int MC, int S, bool HT, bool ST, bool I16, bool O16>
__global__ __launch_bounds__(512) void k_jfa_march(JfaClip one, const JfaClip* __restrict__ clips, int nclips,
                                                   int nitems, int slot_words)
{
    constexpr int kInB = I16 ? 2 : 4;         // bytes per input label
    static_assert(S == 1 || S == 2, "one or two slices of a chain per step");
    static_assert(kMarchTableIds <= 32768, "16-bit labels (direct table only) read sign-extended");
    using MaskT = typename MarchMask<MC>::T;
    constexpr int TW = march_tile_w(K);       // column width (and kMarchTileH rows)
    constexpr int NT = TW * kMarchTileH;      // threads: one voxel of the slice each
    constexpr int R = TW + 2 * K;             // staged columns per slab row
    constexpr int RA = R * (kMarchTileH + 2 * K);
    constexpr int NRS = (RA + NT - 1) / NT;   // staged elements per thread and slab
    constexpr MaskT kBadBit = (MaskT)1 << MC;
    // LAB of the next slice loaded during this one; at K = 2 the extra
    // register costs a block per CU, more than the latency it hides
    constexpr bool kLabAhead = K != 2;
    // Claims that also resolve the ids already listed (see claim below):
    // fewer table reads, four to eight more registers, which cost a wave per
    // SIMD above K = 3 (measured at 640x480x200, K = 1 327 vs 340 us, K = 3
    // 529 vs 548, K = 5 510 vs 444)
    constexpr bool kResolveInClaim = K <= 3;
    // Straight-line prefetch (every step issues its slab and LAB loads
    // unconditionally, clamped), so the compiler waits for just the loads a
    // value needs instead of all of them: K = 2 373 -> 355 us, K = 3 508 ->
    // 496; above K = 3 its registers cost a wave per SIMD (K = 5 436 -> 498)
    constexpr bool kStraight = K <= 3;
    // Records of newly listed ids requested before the new slabs are
    // converted (their latency overlaps the conversion); above K = 3 the
    // registers held across the conversion cost a wave per SIMD
    constexpr bool kStageEarly = K <= 3;
    extern __shared__ uint32_t s_slot[];      // slot_words words: one byte per seed id
    __shared__ uint8_t s_idx[(S + 2) * RA];
    __shared__ float4 s_p[MC];                // (x, y, m0, 2 m1)
    __shared__ float4 s_u[S][MC];             // (dz 2m2, dz 2m4, dz^2 m5, m3) at the step's slices
    __shared__ float4 s_c[MC];                // (id bits, L, a, b)
    __shared__ float4 s_z[MC];                // (z, 2 m2, 2 m4, m5)
    __shared__ int s_list[kListCap];
    __shared__ int s_count;
    __shared__ int s_coll;                    // the last step (gstep) that met a slot shared by two ids
    // ST: the item's wide steps, their chunks, per-tap steps, candidate
    // evaluations of the one-mask path, its wave rounds and wave-slices
    __shared__ unsigned s_cnt[6];
    __shared__ unsigned s_wmax[NT / 64];
    // ST: cycles the waves spent in the phases of steps that extend a table
    // generation (claims, conversion and staging, masks, prefetch issue,
    // evaluation), and those wave-steps
    __shared__ unsigned long long s_tcyc[6];

    const int tid = threadIdx.x;
    for (int i = tid; i < slot_words; i += NT) s_slot[i] = 0xFFFFFFFFu;
    if (tid == 0) {
        s_count = 0;
        s_coll = -1;
        for (int i = 0; i < 6; ++i) {
            s_cnt[i] = 0;
            s_tcyc[i] = 0;
        }
    }
    __syncthreads();
    int gstep = 0;                            // steps of this block so far (uniform)

    const int nb = gridDim.x;                 // a multiple of 8
    const int xcd = blockIdx.x & 7, slot_in_xcd = blockIdx.x >> 3, per_xcd = nb >> 3;
    const int qx = nitems >> 3, rx = nitems & 7;
    const int i_beg = xcd * qx + min(xcd, rx);
    const int i_end = i_beg + qx + (xcd < rx ? 1 : 0);

    const int wave = tid >> 6, lane = tid & 63;
    // a wave covers four rows of 16 voxels (64-byte label rows, against
    // 32-byte rows of an 8 x 8 wave: 318 -> 308 us per pass at step 1,
    // 419 -> 411 at step 5, 640x480x200), TW / 16 of them across a row
    const int lx = (wave % (TW / 16)) * 16 + (lane & 15);
    const int ly = (wave / (TW / 16)) * 4 + (lane >> 4);
    const int mbase = (ly + K) * R + (lx + K);

    int known = 0;                            // s_count as of the last barrier
    // a wide step keeps its voxels' running keys in the table's words: the
    // next release clears the whole table (uniform)
    bool scrubbed = false;
    // key words of a wide step: two per thread and slice
    constexpr int kKeyWords = 2 * NT * S;
    // Wide steps (chunked lists) where their registers keep the occupancy:
    // at K = 2, and in the hashed form above K = 3, they cost a wave per
    // SIMD, and those steps evaluate per tap instead
    constexpr bool kWide = K != 2 && (!HT || K <= 3);
    const bool wide_ok = kWide && slot_words >= kKeyWords;
    // a block's items increase, so its clip index only moves forward: the
    // descriptor is reloaded only when the item leaves the current clip
    int ci = -1, ci_end = 0;
    JfaClip c = one;
    for (int item = i_beg + slot_in_xcd; item < i_end; item += per_xcd) {
        if (clips && item >= ci_end) {
            if (ci < 0) {
                int lo = 0, hi = nclips - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (clips[mid].item_begin <= item) lo = mid;
                    else hi = mid - 1;
                }
                ci = lo;
            }
            while (ci + 1 < nclips && clips[ci + 1].item_begin <= item) ++ci;
            c = clips[ci];
            ci_end = ci + 1 < nclips ? clips[ci + 1].item_begin : nitems;
        }
        const int T = c.tiles_x * c.tiles_y;
        const int local = item - c.item_begin;
        const int t = local % T, gr = local / T;
        const int r = gr % K, g = gr / K;
        const int n = march_len<K>(c, g, r);
        if (n <= 0) continue;
        const int W = c.W, H = c.H, D = c.D;
        // the descriptor's pointers are global memory (flat accesses would
        // share the LDS wait counter)
        const gptr<const char> g_in = (gptr<const char>)c.in;
        const gptr<char> g_out = (gptr<char>)c.out;
        const gptr<const uint32_t> g_lab = (gptr<const uint32_t>)c.lab;
        const gptr<const float> g_rec = (gptr<const float>)c.rec;
        const float ratio = c.ratio;
        const int tx = t % c.tiles_x, ty = t / c.tiles_x;
        const int x0 = tx * TW, y0 = ty * kMarchTileH;
        const bool hashed = HT && c.A > kMarchTableIds;   // uniform per item
        // a slab is staged in raster order over the grown tile: thread tid
        // holds elements tid, tid + NT, ... (offsets within a frame, -1 for
        // none), so every lane of every wave stages
        // (byte offsets within a frame: a load takes the frame's base in
        // scalar registers and this offset)
        int soff[NRS];
#pragma unroll
        for (int i = 0; i < NRS; ++i) {
            const int f = tid + NT * i;
            const int rr = f / R, cc = f - (f / R) * R;
            soff[i] = f < RA ? kInB * (clampi(y0 - K + rr, 0, H - 1) * W + clampi(x0 - K + cc, 0, W - 1)) : -1;
        }
        const int x = x0 + lx, y = y0 + ly;
        const bool mine = x < W && y < H;
        // this thread's voxel within a frame (clamped into it), in bytes
        const uint32_t voff = 4u * (uint32_t)(min(y, H - 1) * W + min(x, W - 1));
        const size_t plane = (size_t)H * W;
        const float fx = (float)x, fy = (float)y;
        // A clip whose seed stage raised `unsafe` (a metric that could make a
        // quadratic form NaN) takes the per-tap path at every step, which
        // clamps with cand_dist's exact rule (the flag is uniform per item)
        const bool exact = c.unsafe && *(gptr<const int>)c.unsafe;

        // a label at a frame base plus a byte offset, and a label stored at
        // this thread's voxel of frame zk (a 16-bit label is read sign-
        // extended: ids stay below 2^15, and 0xFFFF becomes -1 with no
        // compare; 351 against 355 us per full pass at 640x480x200)
        auto ld_label = [&](gptr<const char> base, uint32_t off) -> int {
            if (I16) return (int)ld_off((gptr<const int16_t>)base, off);
            return ld_off((gptr<const int32_t>)base, off);
        };
        auto st_label = [&](int zk, int v) {
            if (O16)
                st_off((gptr<uint16_t>)(g_out + (size_t)zk * plane * 2), voff >> 1, (uint16_t)v);
            else
                st_off((gptr<int32_t>)(g_out + (size_t)zk * plane * 4), voff, (int32_t)v);
        };
        auto load_slab = [&](int gz, int* ids) {
            const gptr<const char> slab = g_in + (size_t)gz * plane * kInB;
            // kStraight: cells past the grown tile read cell 0 and drop it
#pragma unroll
            for (int i = 0; i < NRS; ++i) {
                if (kStraight) {
                    const int v = ld_label(slab, (uint32_t)max(soff[i], 0));
                    ids[i] = soff[i] >= 0 ? v : -1;
                } else {
                    ids[i] = soff[i] >= 0 ? ld_label(slab, (uint32_t)soff[i]) : -1;
                }
            }
        };
        // The first lane of each run of equal ids claims an unclaimed id
        // (list position, published index byte).  With kResolveInClaim
        // every staged id reads its table byte here, and an id already
        // listed resolves to its compact index at once (kept in ids[i] with
        // the kResolved tag); the rest are looked up after the barrier.
        auto claim = [&](int* ids) {
#pragma unroll
            for (int i = 0; i < NRS; ++i) {
                const int id = ids[i];
                const int left = __builtin_amdgcn_update_dpp(-2, id, 0x111, 0xF, 0xF, false);
                if (id >= 0 && (kResolveInClaim || id != left)) {
                    const int sid = HT ? id & (kMarchTableIds - 1) : id;
                    const unsigned sh = (unsigned)(sid & 3) * 8u;
                    uint32_t old = s_slot[sid >> 2];
                    unsigned byte = (old >> sh) & 0xFFu;
                    if (byte == kSlotEmpty && id != left) {
                        while (((old >> sh) & 0xFFu) == kSlotEmpty) {
                            const uint32_t nw = (old & ~(0xFFu << sh)) | (kSlotClaimed << sh);
                            const uint32_t prev = atomicCAS(&s_slot[sid >> 2], old, nw);
                            if (prev == old) {
                                const int k = atomicAdd(&s_count, 1);
                                if (k < kListCap) s_list[k] = id;
                                byte = k < (int)kIdxCap ? (unsigned)k : kIdxCap;
                                // one atomic turns the claimed byte into the
                                // index: lanes reading the word meanwhile see
                                // either the claim or the index
                                atomicXor(&s_slot[sid >> 2], (kSlotClaimed ^ byte) << sh);
                                break;
                            }
                            old = prev;
                            byte = (old >> sh) & 0xFFu;
                        }
                    }
                    if (kResolveInClaim && !hashed && byte < kSlotClaimed) ids[i] = kResolved | (int)byte;
                }
            }
        };
        // index bytes of a staged slab: a cell without an id gets `none` (MC,
        // the mask's bad bit, or kIdxNone in a wide step)
        auto convert = [&](const int* ids, int slot, unsigned none) {
#pragma unroll
            for (int i = 0; i < NRS; ++i) {
                const int f = tid + NT * i;
                const int v = ids[i];
                if (f < RA) {
                    unsigned cix = none;
                    if (v >= kResolved) {
                        cix = (unsigned)v & 0xFFu;
                    } else if (v >= 0) {
                        const int sid = HT ? v & (kMarchTableIds - 1) : v;
                        cix = (s_slot[sid >> 2] >> ((unsigned)(sid & 3) * 8u)) & 0xFFu;
                        if (hashed && cix < kIdxCap && s_list[cix] != v) {
                            cix = none;                    // the slot holds another id
                            s_coll = gstep;
                        }
                    }
                    s_idx[slot * RA + f] = (uint8_t)cix;
                }
            }
        };
        // records of list entries [from, to), entry e's by thread e (MC <
        // 256): with kStageEarly requested by stage_load before the new slabs
        // are converted, else loaded in stage; stage writes them to LDS and
        // sets the dz terms of every entry [0, to) for the step's slices
        // z, z + K, ...
        static_assert(MC < 256, "one list entry per thread");
        struct Staged { float4 a, b, c; int id; };
        auto load_rec = [&](int e) {
            Staged r;
            r.id = s_list[e];
            const gptr<const float> q = g_rec + (size_t)r.id * 12;
            r.a = ld_f4(q);
            r.b = ld_f4(q + 4);
            r.c = ld_f4(q + 8);
            return r;
        };
        auto stage_load = [&](int from, int to) {
            Staged r;                         // read by stage only where loaded
            if (tid >= from && tid < to) r = load_rec(tid);
            return r;
        };
        auto stage = [&](const Staged& pre, int from, int to, int z) {
            const int e = tid;
            if (e < to) {
                float sz, m2d, m4d, m5, m3;
                if (e >= from) {
                    const Staged r = kStageEarly ? pre : load_rec(e);
                    s_p[e] = make_float4(r.a.x, r.a.y, r.a.w, r.b.x);
                    s_z[e] = make_float4(r.a.z, r.b.y, r.b.w, r.c.x);
                    s_c[e] = make_float4(__int_as_float(r.id), r.c.y, r.c.z, r.c.w);
                    sz = r.a.z; m2d = r.b.y; m4d = r.b.w; m5 = r.c.x; m3 = r.b.z;
                } else {
                    const float4 zz = s_z[e];
                    sz = zz.x; m2d = zz.y; m4d = zz.z; m5 = zz.w;
                    m3 = s_u[0][e].w;
                }
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    const float dz = (float)(z + k * K) - sz;
                    s_u[k][e] = make_float4(dz * m2d, dz * m4d, (dz * dz) * m5, m3);
                }
            }
        };
        // a wide step's chunk: list entries [b, b + cnt) staged at 0 .. cnt - 1
        auto stage_chunk = [&](int b, int cnt, int z) {
            if (tid < cnt) {
                const Staged r = load_rec(b + tid);
                s_p[tid] = make_float4(r.a.x, r.a.y, r.a.w, r.b.x);
                s_c[tid] = make_float4(__int_as_float(r.id), r.c.y, r.c.z, r.c.w);
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    const float dz = (float)(z + k * K) - r.a.z;
                    s_u[k][tid] = make_float4(dz * r.b.y, dz * r.b.w, (dz * dz) * r.c.x, r.b.z);
                }
            }
        };
        // index bytes are read one by one (lds_byte): at K = 1 the 16-bit
        // reads the compiler would merge them into took the pass from 335
        // to 595 us at 640x480x200
        auto slab_mask = [&](int slot) {
            MaskT m = 0;
            lds_byte* p = (lds_byte*)(s_idx + slot * RA + mbase);
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) m |= (MaskT)1 << p[dy * K * R + dx * K];
            return m;
        };
        // the same over the indices [b, b + cnt) of a wide step's chunk (bit
        // index - b; the bad bit for the rest and for cells without an id)
        auto slab_mask_chunk = [&](int slot, unsigned b, unsigned cnt) {
            MaskT m = 0;
            lds_byte* p = (lds_byte*)(s_idx + slot * RA + mbase);
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) {
                    const unsigned i = (unsigned)p[dy * K * R + dx * K] - b;
                    m |= (MaskT)1 << (i < cnt ? i : (unsigned)MC);
                }
            return m;
        };
        // release every claimed byte and empty the list (callers barrier after)
        auto release = [&]() {
            if (!scrubbed && known <= kListCap) {
                for (int e = tid; e < known; e += NT) {
                    const int sid = HT ? s_list[e] & (kMarchTableIds - 1) : s_list[e];
                    atomicOr(&s_slot[sid >> 2], 0xFFu << ((unsigned)(sid & 3) * 8u));
                }
            } else {
                // after a wide step every lane has read its last key words
                // before any are cleared (scrubbed is uniform)
                if (scrubbed) __syncthreads();
                for (int i = tid; i < slot_words; i += NT) s_slot[i] = 0xFFFFFFFFu;
            }
        };

        // masks of slabs z - K, z, ..., z + S K at this thread's column;
        // slices z, z + K, ... (S of them) are evaluated per step
        MaskT m[S + 2];
#pragma unroll
        for (int j = 0; j < S + 2; ++j) m[j] = 0;
        int ids_new[S][NRS];
        int z = r + K * g * c.seg;
        uint32_t pk_next[S];
#pragma unroll
        for (int k = 0; k < S; ++k) {
            pk_next[k] = 0u;
            if (kLabAhead && mine && k < n) pk_next[k] = ld_off(g_lab + (size_t)(z + k * K) * plane, voff);
        }
        bool full = true;
        // ST: the phase clock of this wave (uniform)
        unsigned long long t_mark = 0;
        bool timed = false;
        auto tick = [&](int ph) {
            if (ST && c.stats && timed) {
                const unsigned long long t = __builtin_amdgcn_s_memtime();
                if (ph >= 0 && lane == 0) atomicAdd(&s_tcyc[ph], t - t_mark);
                t_mark = t;
            }
        };
        for (int step = 0; step < n;) {
            bool overflow = false;
            timed = ST && !full;
            int nchunk = 1;                   // MC-entry chunks of the step's list (uniform)
            // without the look-ahead, this step's LAB is requested before the
            // claims and barriers, so its latency hides behind them
            uint32_t pk_cur[S];
#pragma unroll
            for (int k = 0; k < S; ++k) {
                pk_cur[k] = 0u;
                if (!kLabAhead)
                    pk_cur[k] = ld_off(g_lab + (size_t)min(z + k * K, D - 1) * plane, voff);
            }
            if (full) {
                release();
                scrubbed = false;
                __syncthreads();
                if (tid == 0) s_count = 0;
                int ids_all[S + 2][NRS];
#pragma unroll
                for (int j = 0; j < S + 2; ++j) load_slab(clampi(z + (j - 1) * K, 0, D - 1), ids_all[j]);
                __syncthreads();
#pragma unroll
                for (int j = 0; j < S + 2; ++j) claim(ids_all[j]);
                __syncthreads();
                known = __builtin_amdgcn_readfirstlane(s_count);
                const bool wide = known > MC && wide_ok;
                overflow = known > (wide_ok ? (int)kIdxCap : MC) || exact;
                if (!overflow) {
                    const int c0 = wide ? MC : known;
                    Staged rs;
                    if (kStageEarly) rs = stage_load(0, c0);
#pragma unroll
                    for (int j = 0; j < S + 2; ++j) convert(ids_all[j], j, wide ? kIdxNone : (unsigned)MC);
                    stage(rs, 0, c0, z);
                }
                __syncthreads();
                overflow = overflow || (hashed && __builtin_amdgcn_readfirstlane(s_coll) == gstep);
                if (!overflow) {
                    if (wide) {
#pragma unroll
                        for (int j = 0; j < S + 2; ++j) m[j] = slab_mask_chunk(j, 0u, (unsigned)MC);
                        nchunk = (known + MC - 1) / MC;
                    } else {
#pragma unroll
                        for (int j = 0; j < S + 2; ++j) m[j] = slab_mask(j);
                    }
                }
                full = overflow || wide;      // rebuild again at the next step
                if (ST && c.stats && tid == 0) {
                    if (nchunk > 1) {
                        ++s_cnt[0];
                        s_cnt[1] += nchunk;
                    }
                    if (overflow) ++s_cnt[2];
                }
            } else {
                tick(-1);
#pragma unroll
                for (int k = 0; k < S; ++k) claim(ids_new[k]);
                __syncthreads();
                tick(0);
                const int cnt = __builtin_amdgcn_readfirstlane(s_count);
                if (cnt > MC) {               // uniform: rebuild this step
                    known = cnt;
                    full = true;
                    continue;
                }
                Staged rs;
                if (kStageEarly) rs = stage_load(known, cnt);
#pragma unroll
                for (int k = 0; k < S; ++k) convert(ids_new[k], k, (unsigned)MC);
                stage(rs, known, cnt, z);
                known = cnt;
                __syncthreads();
                tick(1);
                if (hashed && __builtin_amdgcn_readfirstlane(s_coll) == gstep) {
                    overflow = true;          // this step per tap, the next one rebuilds
                    full = true;
                    if (ST && c.stats && tid == 0) ++s_cnt[2];
                } else {
                    m[0] = m[S];
                    m[1] = m[S + 1];
#pragma unroll
                    for (int k = 0; k < S; ++k) m[2 + k] = slab_mask(k);
                }
                tick(2);
            }
            const int ns = min(S, n - step);  // slices of this step
            // the next step's new slabs (and LAB), loaded while this step is evaluated
            uint32_t pk_here[S];
#pragma unroll
            for (int k = 0; k < S; ++k) pk_here[k] = kLabAhead ? pk_next[k] : pk_cur[k];
            if (kStraight || step + S < n) {
                // kStraight: issued at the last step too (clamped, unused)
                const int zn = z + S * K;
#pragma unroll
                for (int k = 0; k < S; ++k) load_slab(clampi(zn + (k + 1) * K, 0, D - 1), ids_new[k]);
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    if (kLabAhead && kStraight)
                        pk_next[k] = ld_off(g_lab + (size_t)min(zn + k * K, D - 1) * plane, voff);
                    else if (kLabAhead && mine && step + S + k < n)
                        pk_next[k] = ld_off(g_lab + (size_t)(zn + k * K) * plane, voff);
                }
            }

            // one slice's candidates of the mask path, lowering `best`
            auto eval_masks = [&](int k, float l0, float l1, float l2, unsigned long long best) {
                MaskT mask = (m[k] | m[k + 1] | m[k + 2]) & ~kBadBit;
                while (mask) {
                    const int e = MC == 31 ? __builtin_ctz((uint32_t)mask) : __builtin_ctzll((unsigned long long)mask);
                    mask &= mask - 1;
                    const float4 p = s_p[e], u = s_u[k][e], cc = s_c[e];
                    const float dx = fx - p.x, dy = fy - p.y;
                    const float t1 = __fmaf_rn(dx, p.z, __fmaf_rn(dy, p.w, u.x));
                    const float t2 = __fmaf_rn(dy, u.w, u.y);
                    const float q = nonneg(__fmaf_rn(dx, t1, __fmaf_rn(dy, t2, u.z)));
                    const float c0 = l0 - cc.y, c1 = l1 - cc.z, c2 = l2 - cc.w;
                    const float d = __fmaf_rn(ratio, q, __fmaf_rn(c2, c2, __fmaf_rn(c1, c1, c0 * c0)));
                    const unsigned long long kk =
                        ((unsigned long long)__float_as_uint(d) << 32) | __float_as_uint(cc.x);
                    best = key_min(kk, best);
                }
                return best;
            };
            auto lab_of = [&](uint32_t pk, float& l0, float& l1, float& l2) {
                l0 = (float)(pk & 255u);
                l1 = (float)((pk >> 8) & 255u);
                l2 = (float)((pk >> 16) & 255u);
            };

            tick(3);
            if (mine && nchunk == 1) {
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    if (k >= ns) break;
                    const int zk = z + k * K;
                    float l0, l1, l2;
                    lab_of(pk_here[k], l0, l1, l2);
                    unsigned long long best = ASV_KEY_INIT;
                    if (!overflow) {
                        if (ST && c.stats) {
                            // this lane's candidates, the wave's rounds (its
                            // busiest lane) and one wave-slice
                            const MaskT mk = (m[k] | m[k + 1] | m[k + 2]) & ~kBadBit;
                            const unsigned pc = MC == 31 ? __builtin_popcount((uint32_t)mk)
                                                         : __builtin_popcountll((unsigned long long)mk);
                            const unsigned long long act = __ballot(1);
                            const bool leader = lane == __builtin_ctzll(act);
                            // every active lane resets the slot, so the reset
                            // is one store ahead of the one atomic for all
                            s_wmax[wave] = 0u;
                            atomicMax(&s_wmax[wave], pc);
                            atomicAdd(&s_cnt[3], pc);
                            if (leader) {
                                atomicAdd(&s_cnt[4], s_wmax[wave]);
                                atomicAdd(&s_cnt[5], 1u);
                            }
                        }
                        best = eval_masks(k, l0, l1, l2, best);
                    } else {
                        // per tap from global labels and records, the tap
                        // coordinates formed as the loop goes (fewer live
                        // registers than three precomputed lists)
                        const float fz = (float)zk;
                        int seen0 = -1, seen1 = -1;
#pragma unroll 1
                        for (int sl = 0; sl < 3; ++sl) {
                            const int zz = sl == 0 ? zk : clampi(zk + (sl == 1 ? -K : K), 0, D - 1);
#pragma unroll 1
                            for (int rr = -1; rr <= 1; ++rr) {
                                const gptr<const char> row =
                                    g_in + ((size_t)zz * H + clampi(y + rr * K, 0, H - 1)) * W * kInB;
#pragma unroll 1
                                for (int cc = -1; cc <= 1; ++cc) {
                                    const int id = ld_label(row, (uint32_t)(clampi(x + cc * K, 0, W - 1) * kInB));
                                    if (id < 0 || id == seen0 || id == seen1) continue;
                                    seen1 = seen0;
                                    seen0 = id;
                                    const gptr<const float> q = g_rec + (size_t)id * 12;
                                    const float d = cand_dist(ld_f4(q), ld_f4(q + 4), ld_f4(q + 8), fx, fy, fz, l0,
                                                              l1, l2, ratio);
                                    const unsigned long long kk = dkey(d, id);
                                    best = kk < best ? kk : best;
                                }
                            }
                        }
                    }
                    st_label(zk, best == ASV_KEY_INIT ? -1 : (int32_t)(unsigned)(best & 0xffffffffull));
                }
            }
            if (kWide && nchunk > 1) {
                // a wide step: each slice's best key lowered chunk by chunk,
                // kept between chunks in the table's words (this thread's
                // own: word 2 k NT + tid low, 2 k NT + NT + tid high)
                for (int ch = 0;;) {
                    const bool last = ch + 1 == nchunk;
                    if (mine) {
#pragma unroll
                        for (int k = 0; k < S; ++k) {
                            if (k >= ns) break;
                            float l0, l1, l2;
                            lab_of(pk_here[k], l0, l1, l2);
                            uint32_t* kw = s_slot + k * 2 * NT + tid;
                            unsigned long long best = ASV_KEY_INIT;
                            if (ch > 0) best = ((unsigned long long)kw[NT] << 32) | kw[0];
                            best = eval_masks(k, l0, l1, l2, best);
                            if (!last) {
                                kw[0] = (uint32_t)best;
                                kw[NT] = (uint32_t)(best >> 32);
                            } else {
                                st_label(z + k * K,
                                         best == ASV_KEY_INIT ? -1 : (int32_t)(unsigned)(best & 0xffffffffull));
                            }
                        }
                    }
                    if (last) break;
                    ++ch;
                    // every lane is done with this chunk's records before
                    // the next chunk's replace them
                    const int b = ch * MC, cnt = min(MC, known - b);
                    __syncthreads();
                    stage_chunk(b, cnt, z);
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < S + 2; ++j) m[j] = slab_mask_chunk(j, (unsigned)b, (unsigned)cnt);
                }
                scrubbed = true;
            }
            tick(4);
            if (ST && c.stats && timed && lane == 0) atomicAdd(&s_tcyc[5], 1ull);
            z += S * K;
            step += S;
            ++gstep;
        }
        if (ST && c.stats) __syncthreads();   // every wave's counts of the item are in (uniform)
        if (ST && c.stats && tid == 0) {
            unsigned long long* row = c.stats + (K - 1) * 16;   // this step size's row
            atomicAdd(row, (unsigned long long)((n + S - 1) / S));
            for (int i = 0; i < 6; ++i) {
                atomicAdd(row + 1 + i, (unsigned long long)s_cnt[i]);
                s_cnt[i] = 0;
                atomicAdd(row + 7 + i, s_tcyc[i]);
                s_tcyc[i] = 0;
            }
        }
    }
}

using MarchFn = void (*)(JfaClip, const JfaClip*, int, int, int);

template <int K, bool HT, bool ST>
static MarchFn march_fn(int mc, int sl)
{
    if (sl == 2) return mc == 31 ? k_jfa_march<K, 31, 2, HT, ST, false, false> : k_jfa_march<K, 63, 2, HT, ST, false, false>;
    return mc == 31 ? k_jfa_march<K, 31, 1, HT, ST, false, false> : k_jfa_march<K, 63, 1, HT, ST, false, false>;
}

// The 16-bit label forms exist for two slices per step, the direct table
// and ordinary launches (lt: 1 32-bit in, 16-bit out; 2 both 16-bit; 3
// 16-bit in, 32-bit out), with 16-bit input only up to step 5.
template <int K>
static MarchFn march_fn16(int mc, int lt)
{
    if (lt == 1)
        return mc == 31 ? k_jfa_march<K, 31, 2, false, false, false, true> : k_jfa_march<K, 63, 2, false, false, false, true>;
    if constexpr (K <= 5) {
        if (lt == 2)
            return mc == 31 ? k_jfa_march<K, 31, 2, false, false, true, true> : k_jfa_march<K, 63, 2, false, false, true, true>;
        if (lt == 3)
            return mc == 31 ? k_jfa_march<K, 31, 2, false, false, true, false> : k_jfa_march<K, 63, 2, false, false, true, false>;
    }
    return nullptr;
}

// Instances of steps 4P+1 .. 4P+4 (defined by asv_march.hip built with
// ASV_MARCH_PART = P): the kernel for step k, capacity mc, slices per step
// sl; ht selects the hashed-table form, counted the counting form, lt the
// label widths (0: 32-bit in and out; see march_fn16), null where no such
// instance exists.
template <int P>
MarchFn march_kernel_part(int k, int mc, int sl, bool ht, bool counted, int lt);

}  // namespace asv
