// asv_seedpass.hip -- the first two JFA passes of an iteration, resolved
// brick by brick in LDS (lowest-id rule).
//
// Right after the seed stage (draw_seeds / draw_next_seeds: every voxel -1
// but the seed voxels, the higher id winning a shared voxel) the reference
// runs the schedule's first two passes like every other
// (GPU-AS/Image.cpp:1047-1190, shaders/flood.gs:26-182).  Their outputs are
// sparse: after the first pass (step k1) only voxels with a seed among their
// 27 clamped taps carry a label, and after the second (step k2) only voxels
// with such a voxel among theirs.  Here both passes run from the seed
// positions:
//  - the seeds are binned by the bricks (32 x 16 x kSeedBrickZ voxels) that
//    their reach, k1 + k2 voxels per axis, meets: counts, one exclusive
//    scan, then each seed's entry (id, voxel, record) copied into the
//    exact list of every such brick;
//  - one block per brick stages its entries in LDS (in chunks when there
//    are more than LDS holds), finds for each seed the other seeds within
//    2 k1 (the only ones that can share a first-pass voxel with it) and
//    whether a higher id hides it on its voxel (the seed stage's scatter
//    keeps the higher id); then one thread per (seed, voxel t of the seed's
//    step-k1 sets) that is a tap of the brick evaluates the first pass at t
//    from those seeds and, where t gets a label, offers it to every brick
//    voxel that has t among its step-k2 taps by a 64-bit LDS atomic minimum
//    on (distance bits, id): the lowest-id rule's full-pass winner, since a
//    voxel's candidates are exactly the labels at its taps and repeats
//    cannot change a minimum;
//  - the brick is written out whole, -1 where no candidate arrived.
// The seed stage's -1 fill and scatter and the first pass's output volume
// are not needed.
#include "asv_device.h"

namespace asv {

namespace {

constexpr int kBrickX = 32, kBrickY = 16;
constexpr int kBrickThreads = 512;
constexpr int kBrickLds = 96;      // list entries staged in LDS at a time
constexpr int kNear = 4;           // seeds within 2 k1 of a seed kept per entry

// One seed in a brick's list: id, voxel (x | y << 16, z) and its record.
struct alignas(16) BrickEntry {
    int4 p;                        // (id, x | y << 16, z, 0)
    float4 a, b, c;
};

// Bricks [lo, hi] along one axis that coordinates p - R .. p + R (clamped to
// the clip) meet.
__device__ __forceinline__ void brick_span(int p, int R, int n, int B, int* lo, int* hi)
{
    *lo = max(p - R, 0) / B;
    *hi = min(p + R, n - 1) / B;
}

// Thread (seed i, u) takes the u-th brick of the at most mx x my x mz that
// seed i's reach meets: with entries == nullptr it adds one to count[b]
// (and the first thread zeroes the running total k_brick_alloc takes the
// lists' segments from), else it writes the seed's entry at cursor[b] (and
// advances it).
__global__ __launch_bounds__(256) void k_brick_bin(int W, int H, int D, int A, int R, int bz_dim, int nbx, int nby,
                                                   int mx, int my, int mz, const int4* __restrict__ pos,
                                                   const SeedRec* __restrict__ rec, int* __restrict__ count,
                                                   int* __restrict__ cursor, BrickEntry* __restrict__ entries,
                                                   int* __restrict__ total)
{
    if (!entries && blockIdx.x == 0 && threadIdx.x == 0) *total = 0;
    const int per = mx * my * mz;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int i = (int)(g / per), u = (int)(g - (int64_t)i * per);
    if (i >= A) return;
    const int4 p = pos[i];
    int x0, x1, y0, y1, z0, z1;
    brick_span(p.x, R, W, kBrickX, &x0, &x1);
    brick_span(p.y, R, H, kBrickY, &y0, &y1);
    brick_span(p.z, R, D, bz_dim, &z0, &z1);
    const int bx = x0 + u % mx, by = y0 + (u / mx) % my, bz = z0 + u / (mx * my);
    if (bx > x1 || by > y1 || bz > z1) return;
    const int b = (bz * nby + by) * nbx + bx;
    if (!entries) {
        atomicAdd(&count[b], 1);
        return;
    }
    const SeedRec r = rec[i];
    BrickEntry e;
    e.p = make_int4(i, p.x | (p.y << 16), p.z, 0);
    e.a = r.a;
    e.b = r.b;
    e.c = r.c;
    entries[atomicAdd(&cursor[b], 1)] = e;
}

// Thread b: brick b's list segment [off[b], off[b] + count[b]) of the
// entries buffer, taken from the running total with one atomic per wave
// (the segments' order in the buffer is arbitrary, as the entries' order
// within a list is: the brick kernel's minima depend on neither); cursor[b]
// starts at off[b], so after the filling pass it ends the list, and count[b]
// returns to zero for the next binning.
__global__ __launch_bounds__(256) void k_brick_alloc(int nb, int* __restrict__ count, int* __restrict__ off,
                                                      int* __restrict__ cursor, int* __restrict__ total)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int c = b < nb ? count[b] : 0;
    int incl = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d);
        if (lane >= d) incl += v;
    }
    int base = 0;
    if (lane == 63) base = atomicAdd(total, incl);
    base = __shfl(base, 63);
    if (b < nb) {
        off[b] = base + incl - c;
        cursor[b] = base + incl - c;
        count[b] = 0;
    }
}

// The coordinates c of [lo, hi] (inside [0, n - 1]) that have t among their
// step-k taps (c - k, c, c + k clamped to [0, n - 1]), as first + stride i
// for i < count, packed first | stride << 16 | count << 21 (0: none).  An
// inner t is the tap of t - k, t and t + k only; t = 0 is the clamped tap of
// every c <= k, t = n - 1 of every c >= n - 1 - k.
__device__ __forceinline__ int pair_axis(int t, int n, int k, int lo, int hi)
{
    int first, stride, last;
    if (t == 0 || t == n - 1) {
        first = max(t == 0 ? 0 : n - 1 - k, lo);
        last = min(t == 0 ? k : n - 1, hi);
        stride = 1;
    } else {
        int j0 = 0, j1 = 2;
        while (j0 <= 2 && t + (j0 - 1) * k < lo) ++j0;
        while (j1 >= 0 && t + (j1 - 1) * k > hi) --j1;
        first = t + (j0 - 1) * k;
        last = t + (j1 - 1) * k;
        stride = k;
    }
    if (last < first) return 0;
    const int count = (last - first) / stride + 1;
    return first | (stride << 16) | (count << 21);
}

__device__ __forceinline__ int ax_first(int a) { return a & 0xffff; }
__device__ __forceinline__ int ax_stride(int a) { return (a >> 16) & 31; }
__device__ __forceinline__ int ax_count(int a) { return a >> 21; }

// Brick kernel: BZ frames deep, kBrickX x kBrickY voxels per frame; O16:
// 16-bit output labels (0xFFFF for -1) for a marching pass that reads them.
template <int BZ, bool O16>
__global__ __launch_bounds__(kBrickThreads) void k_seed_bricks(int W, int H, int D, int k1, int k2, float ratio1,
                                                               float ratio2, int nbx, int nby,
                                                               const uint32_t* __restrict__ lab,
                                                               const int* __restrict__ off,
                                                               const int* __restrict__ end,
                                                               const BrickEntry* __restrict__ entries,
                                                               int32_t* __restrict__ out)
{
    constexpr int kVox = kBrickX * kBrickY * BZ;
    constexpr int kPairBatch = 4;             // pairs whose loads are issued together
    __shared__ unsigned long long s_key[kVox];
    __shared__ int4 s_p[kBrickLds];           // staged entries: (id, x | y << 16, z | visible << 31, 0)
    __shared__ float4 s_ra[kBrickLds], s_rb[kBrickLds], s_rc[kBrickLds];
    __shared__ int s_near[kBrickLds][kNear];  // list indices within 2 k1 (itself included)
    __shared__ int s_nnear[kBrickLds];        // how many (more than kNear: scan every entry)
    // per staged seed and axis: its step-k1 coordinates t that are taps of
    // the brick's voxels, with pair_axis of each (at most 3 for k1 <= 2)
    __shared__ short s_t[3][3][kBrickLds];
    __shared__ int s_pa[3][3][kBrickLds];
    __shared__ unsigned char s_nt[3][kBrickLds];

    const int tid = threadIdx.x;
    const int b = blockIdx.x;
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const int x0 = bx * kBrickX, y0 = by * kBrickY, z0 = bz * BZ;
    const int x1 = min(x0 + kBrickX, W) - 1, y1 = min(y0 + kBrickY, H) - 1, z1 = min(z0 + BZ, D) - 1;
    const int o0 = off[b], n = end[b] - o0;
    const BrickEntry* __restrict__ g_e = entries + o0;
#pragma unroll
    for (int i = 0; i < kVox / kBrickThreads; ++i) s_key[i * kBrickThreads + tid] = ASV_KEY_INIT;

    for (int c0 = 0; c0 < n; c0 += kBrickLds) {
        const int cn = min(n - c0, kBrickLds);
        __syncthreads();                      // the previous chunk's items are done
        if (tid < cn) {
            const BrickEntry e = g_e[c0 + tid];
            s_p[tid] = e.p;
            s_ra[tid] = e.a;
            s_rb[tid] = e.b;
            s_rc[tid] = e.c;
        }
        __syncthreads();
        // list entry q: staged when it lies in this chunk
        auto pos_of = [&](int q) -> int4 {
            return q - c0 >= 0 && q - c0 < cn ? s_p[q - c0] : g_e[q].p;
        };
        // the chunk's entries: visibility, the entries within 2 k1, and the
        // per-axis first-pass coordinates that are taps of the brick
        if (tid < cn) {
            const int4 me = s_p[tid];
            bool hid = false;
            int nn = 0;
            for (int q = 0; q < n; ++q) {
                const int4 o = pos_of(q);
                const int oz = o.z & 0x7fffffff;
                hid = hid || (o.y == me.y && oz == me.z && o.x > me.x);
                if (abs((o.y & 0xffff) - (me.y & 0xffff)) <= 2 * k1 && abs((o.y >> 16) - (me.y >> 16)) <= 2 * k1 &&
                    abs(oz - me.z) <= 2 * k1) {
                    if (nn < kNear) s_near[tid][nn] = q;
                    ++nn;
                }
            }
            s_nnear[tid] = nn;
            const int pc[3] = {me.y & 0xffff, me.y >> 16, me.z};
            const int dim[3] = {W, H, D}, lo[3] = {x0, y0, z0}, hi[3] = {x1, y1, z1};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                int m = 0;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    int t;
                    if (!seed_axis(pc[a], dim[a], k1, i, &t) || t < lo[a] - k2 || t > hi[a] + k2) continue;
                    const int pa = pair_axis(t, dim[a], k2, lo[a], hi[a]);
                    if (pa == 0) continue;
                    s_t[a][m][tid] = (short)t;
                    s_pa[a][m][tid] = pa;
                    ++m;
                }
                s_nt[a][tid] = (unsigned char)m;
            }
            if (!hid) atomicOr(&s_p[tid].z, (int)0x80000000);
        }
        __syncthreads();
        // whether entry q is a candidate (a higher id on its voxel hides it)
        auto visible = [&](int q, int4 o) {
            if (q - c0 >= 0 && q - c0 < cn) return s_p[q - c0].z < 0;
            for (int r = 0; r < n; ++r) {
                const int4 x = pos_of(r);
                if (x.y == o.y && (x.z & 0x7fffffff) == o.z && x.x > o.x) return false;
            }
            return true;
        };

        // one item per (staged seed, voxel t of its step-k1 sets)
        for (int item = tid; item < cn * 27; item += kBrickThreads) {
            const int j = item / 27, ti = item - j * 27;
            const int ix = ti % 3, iy = (ti / 3) % 3, iz = ti / 9;
            if (s_p[j].z >= 0 || ix >= s_nt[0][j] || iy >= s_nt[1][j] || iz >= s_nt[2][j]) continue;
            const int tx = s_t[0][ix][j], ty = s_t[1][iy][j], tz = s_t[2][iz][j];
            const int ax = s_pa[0][ix][j], ay = s_pa[1][iy][j], az = s_pa[2][iz][j];
            const int cx = ax_count(ax), cy = ax_count(ay), cz = ax_count(az);
            const int np = cx * cy * cz;
            // the brick voxels that tap t, walked x fastest; the first ones'
            // LAB requested with t's
            int px = 0, py = 0, pz = 0;
            int vx[kPairBatch], vy[kPairBatch], vz[kPairBatch];
            uint32_t vk[kPairBatch];
            auto load_pairs = [&](int p0) {
#pragma unroll
                for (int u = 0; u < kPairBatch; ++u) {
                    vx[u] = ax_first(ax) + px * ax_stride(ax);
                    vy[u] = ax_first(ay) + py * ax_stride(ay);
                    vz[u] = ax_first(az) + pz * ax_stride(az);
                    vk[u] = lab[((size_t)vz[u] * H + vy[u]) * W + vx[u]];
                    if (p0 + u + 1 < np) {            // the next pair (the last one repeats)
                        if (++px == cx) {
                            px = 0;
                            if (++py == cy) {
                                py = 0;
                                ++pz;
                            }
                        }
                    }
                }
            };
            const uint32_t pk = lab[((size_t)tz * H + ty) * W + tx];
            load_pairs(0);
            const float l0 = (float)(pk & 255u), l1 = (float)((pk >> 8) & 255u), l2 = (float)((pk >> 16) & 255u);
            // first pass at t: the visible seeds among t's step-k1 taps, all
            // within 2 k1 of this seed
            unsigned long long best = ASV_KEY_INIT;
            float4 ba = make_float4(0.f, 0.f, 0.f, 0.f), bb = ba, bc = ba;
            auto offer = [&](int q) {
                const int4 o = pos_of(q);
                const int oz = o.z & 0x7fffffff;
                if (!axis_hit(tx, o.y & 0xffff, W, k1) || !axis_hit(ty, o.y >> 16, H, k1) || !axis_hit(tz, oz, D, k1) ||
                    !visible(q, make_int4(o.x, o.y, oz, 0)))
                    return;
                float4 ra, rb, rc;
                if (q - c0 >= 0 && q - c0 < cn) {
                    ra = s_ra[q - c0];
                    rb = s_rb[q - c0];
                    rc = s_rc[q - c0];
                } else {
                    ra = g_e[q].a;
                    rb = g_e[q].b;
                    rc = g_e[q].c;
                }
                const float d = cand_dist(ra, rb, rc, (float)tx, (float)ty, (float)tz, l0, l1, l2, ratio1);
                const unsigned long long kk = dkey(d, o.x);
                if (kk < best) {
                    best = kk;
                    ba = ra;
                    bb = rb;
                    bc = rc;
                }
            };
            const int nn = s_nnear[j];
            if (nn <= kNear) {
                for (int u = 0; u < nn; ++u) offer(s_near[j][u]);
            } else {
                for (int q = 0; q < n; ++q) offer(q);
            }
            if (best >= ASV_KEY_INIT) continue;
            const int L = (int)(unsigned)(best & 0xffffffffull);
            // second pass: t's label offered to the brick voxels that tap t
            for (int p0 = 0;;) {
#pragma unroll
                for (int u = 0; u < kPairBatch; ++u) {
                    if (p0 + u >= np) break;
                    const uint32_t q = vk[u];
                    const float d = cand_dist(ba, bb, bc, (float)vx[u], (float)vy[u], (float)vz[u],
                                              (float)(q & 255u), (float)((q >> 8) & 255u), (float)((q >> 16) & 255u),
                                              ratio2);
                    const unsigned long long kk = dkey(d, L);
                    if (kk < ASV_KEY_INIT)
                        atomicMin(&s_key[((vz[u] - z0) * kBrickY + (vy[u] - y0)) * kBrickX + (vx[u] - x0)], kk);
                }
                p0 += kPairBatch;
                if (p0 >= np) break;
                load_pairs(p0);
            }
        }
    }
    __syncthreads();
    // kBrickThreads voxels (whole rows of whole slices) per round
    constexpr int kRows = kBrickThreads / kBrickX;
    const int x = x0 + tid % kBrickX, yz = tid / kBrickX;
    const int y = y0 + yz % kBrickY, zs = yz / kBrickY;
    if (x <= x1 && y <= y1) {
        const size_t v0 = ((size_t)(z0 + zs) * H + y) * W + x;
        const size_t step = (size_t)(kRows / kBrickY) * H * W;
#pragma unroll
        for (int i = 0; i < kVox / kBrickThreads; ++i) {
            if (z0 + zs + i * (kRows / kBrickY) <= z1) {
                const unsigned long long kk = s_key[i * kBrickThreads + tid];
                const int32_t l = kk < ASV_KEY_INIT ? (int32_t)(unsigned)(kk & 0xffffffffull) : -1;
                if (O16)
                    reinterpret_cast<uint16_t*>(out)[v0 + i * step] = (uint16_t)l;
                else
                    out[v0 + i * step] = l;
            }
        }
    }
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// Bricks along an axis of brick side B that a reach of R meets at most.
inline int64_t reach_bricks(int R, int B) { return 2 * R / B + 2; }

}  // namespace

void seed_brick_grid(int W, int H, int D, int* nbx, int* nby, int* nbz)
{
    *nbx = (int)cdiv(W, kBrickX);
    *nby = (int)cdiv(H, kBrickY);
    *nbz = (int)cdiv(D, kSeedBrickZ);
}

int64_t seed_brick_bytes(int W, int H, int D, int64_t seeds, int R)
{
    int nbx, nby, nbz;
    seed_brick_grid(W, H, D, &nbx, &nby, &nbz);
    const int64_t per_seed = reach_bricks(R, kBrickX) * reach_bricks(R, kBrickY) * reach_bricks(R, kSeedBrickZ);
    return (int64_t)nbx * nby * nbz * 12 + 4 + seeds * per_seed * (int64_t)sizeof(BrickEntry);
}

hipError_t alloc_seed_bricks(SeedBricks* bb, int W, int H, int D, int seeds, int R, int64_t* total)
{
    int nbx, nby, nbz;
    seed_brick_grid(W, H, D, &nbx, &nby, &nbz);
    bb->nb_cap = (int64_t)nbx * nby * nbz;
    bb->seed_cap = seeds;
    bb->r_cap = R;
    const int64_t cap = seeds * reach_bricks(R, kBrickX) * reach_bricks(R, kBrickY) * reach_bricks(R, kSeedBrickZ);
    hipError_t e = dev_malloc((void**)&bb->count, sizeof(int) * (size_t)bb->nb_cap);
    if (e == hipSuccess) e = dev_malloc((void**)&bb->off, sizeof(int) * (size_t)(bb->nb_cap + 1));
    if (e == hipSuccess) e = dev_malloc((void**)&bb->cursor, sizeof(int) * (size_t)bb->nb_cap);
    if (e == hipSuccess) e = dev_malloc(&bb->entries, sizeof(BrickEntry) * (size_t)cap);
    if (e == hipSuccess) e = hipMemsetAsync(bb->count, 0, sizeof(int) * (size_t)bb->nb_cap, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e == hipSuccess && total) *total += seed_brick_bytes(W, H, D, seeds, R);
    return e;
}

void free_seed_bricks(SeedBricks* bb)
{
    void* ptrs[] = {bb->count, bb->off, bb->cursor, bb->entries};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    *bb = SeedBricks{};
}

hipError_t launch_seed_bricks(hipStream_t st, const Dims& d, const int4* pos, const SeedRec* rec, const uint32_t* lab,
                              int32_t* out, int k1, int k2, float ratio1, float ratio2, const SeedBricks& bb, bool out16)
{
    if (k1 < 1 || k1 > kSeedFuseMaxK1 || k2 < 1 || k2 > kSeedPassMaxStep || d.W >= 32768 || d.H >= 32768 ||
        d.D >= 32768)
        return hipErrorInvalidValue;
    int nbx, nby, nbz;
    seed_brick_grid(d.W, d.H, d.D, &nbx, &nby, &nbz);
    const int64_t nb = (int64_t)nbx * nby * nbz;
    const int R = k1 + k2;
    if (nb > bb.nb_cap || d.A > bb.seed_cap || R > bb.r_cap || nb >= 0x7fffffff) return hipErrorInvalidValue;
    BrickEntry* entries = (BrickEntry*)bb.entries;
    const int mx = (int)reach_bricks(R, kBrickX), my = (int)reach_bricks(R, kBrickY),
              mz = (int)reach_bricks(R, kSeedBrickZ);
    const unsigned g = cdiv((int64_t)d.A * mx * my * mz, 256);
    // the running total of the lists' segments (zeroed by the counting
    // pass; without seeds every list is empty and no entry is read)
    int* total = bb.off + bb.nb_cap;
    if (d.A > 0)
        hipLaunchKernelGGL(k_brick_bin, dim3(g), dim3(256), 0, st, d.W, d.H, d.D, d.A, R, kSeedBrickZ, nbx, nby, mx,
                           my, mz, pos, rec, bb.count, bb.cursor, (BrickEntry*)nullptr, total);
    hipLaunchKernelGGL(k_brick_alloc, dim3(cdiv(nb, 256)), dim3(256), 0, st, (int)nb, bb.count, bb.off, bb.cursor,
                       total);
    if (d.A > 0)
        hipLaunchKernelGGL(k_brick_bin, dim3(g), dim3(256), 0, st, d.W, d.H, d.D, d.A, R, kSeedBrickZ, nbx, nby, mx,
                           my, mz, pos, rec, bb.count, bb.cursor, entries, total);
    if (out16 && d.A > kMarchTableIds) return hipErrorInvalidValue;   // ids must fit 16 bits
    if (out16)
        hipLaunchKernelGGL((k_seed_bricks<kSeedBrickZ, true>), dim3((unsigned)nb), dim3(kBrickThreads), 0, st, d.W, d.H,
                           d.D, k1, k2, ratio1, ratio2, nbx, nby, lab, bb.off, bb.cursor, entries, out);
    else
        hipLaunchKernelGGL((k_seed_bricks<kSeedBrickZ, false>), dim3((unsigned)nb), dim3(kBrickThreads), 0, st, d.W,
                           d.H, d.D, k1, k2, ratio1, ratio2, nbx, nby, lab, bb.off, bb.cursor, entries, out);
    return hipGetLastError();
}

}  // namespace asv
