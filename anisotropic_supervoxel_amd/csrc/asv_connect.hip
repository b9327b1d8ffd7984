// asv_connect.hip -- 3D connectivity enforcement on the device
// (EnforceLabelConnectivity3D, GPU-AS/Image.cpp:1938-2062).
//
// The reference visits voxels in raster order and grows each unvisited
// voxel's segment breadth-first over the 24 neighbours (dz in -1..1 times
// the 8 in-plane offsets; (0, 0, +-1) is not among them) with the same input
// label; a segment smaller than supsz / 4 takes the "adjacent label" (the
// output label of the last already-visited neighbour of its first voxel, in
// scan order, or the previous segment's value when there is none) and every
// other segment gets the next number.  The same result, in parallel:
//  1. connected components of the same-label 24-neighbourhood graph by
//     lock-free union-find (hooking the larger root under the smaller, the
//     scheme of ECL-CC), so every component's root is its smallest voxel
//     index, i.e. the voxel where the reference starts its segment;
//  2. the roots in raster order by an ordered compaction, component sizes;
//  3. for each root, the rank of the component of the last neighbour (scan
//     order of the reference) whose component starts earlier -- exactly the
//     neighbours the reference finds already visited;
//  4. on the host, one pass over the components in order assigns numbers
//     and adjacent labels (a chain through earlier components only);
//  5. every voxel takes its component's label.
#include "asv_internal.h"

#include <vector>

namespace asv {

namespace {

constexpr int kCcChunk = 4096;                 // voxels per compaction block (256 threads x 16)

__device__ __forceinline__ int ld_parent(const int* p, int i)
{
    return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of v; links always point to smaller indices, so the walk ends where
// parent[x] == x.  Shortens the path as it goes (path halving).
__device__ int cc_root(int* parent, int v)
{
    int cur = ld_parent(parent, v);
    if (cur != v) {
        int prev = v, next;
        while (cur > (next = ld_parent(parent, cur))) {
            __hip_atomic_store(parent + prev, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            prev = cur;
            cur = next;
        }
    }
    return cur;
}

// Joins the components of a and b: the larger root is hooked under the
// smaller one by compare-and-swap, retried from the new roots when another
// thread hooked it first.
__device__ void cc_hook(int* parent, int a, int b)
{
    int ra = cc_root(parent, a), rb = cc_root(parent, b);
    while (ra != rb) {
        if (ra < rb) {
            const int t = ra;
            ra = rb;
            rb = t;
        }
        const int old = atomicCAS(parent + ra, ra, rb);
        if (old == ra) return;
        ra = cc_root(parent, old);
        rb = cc_root(parent, rb);
    }
}

__global__ __launch_bounds__(256) void k_cc_init(int64_t V, int* __restrict__ parent)
{
    for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < V; v += (int64_t)gridDim.x * 256)
        parent[v] = (int)v;
}

// In-plane offsets in the reference's scan order (Image.cpp:1944-1945).
__device__ __constant__ int c_ox[8] = {-1, -1, 0, 1, 1, 1, 0, -1};
__device__ __constant__ int c_oy[8] = {0, -1, -1, -1, 0, 1, 1, 1};

// Union over the 12 neighbours that precede v in raster order (the slice
// below, and the four earlier in-plane offsets), same label only.
__global__ __launch_bounds__(256) void k_cc_union(int W, int H, int D, const int32_t* __restrict__ labels,
                                                  int* parent)
{
    const int64_t V = (int64_t)W * H * D;
    for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < V; v += (int64_t)gridDim.x * 256) {
        const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((int64_t)W * H));
        const int32_t l = labels[v];
        for (int dz = -1; dz <= 0; ++dz) {
            const int zz = z + dz;
            if (zz < 0) continue;
            for (int n = 0; n < 8; ++n) {
                const int xx = x + c_ox[n], yy = y + c_oy[n];
                if (xx < 0 || xx >= W || yy < 0 || yy >= H) continue;
                const int64_t u = ((int64_t)zz * H + yy) * W + xx;
                if (u >= v) continue;                  // later neighbours hook from their own side
                if (labels[u] == l) cc_hook(parent, (int)v, (int)u);
            }
        }
    }
}

// Every voxel points at its root.  The walk only reads: a path-halving
// store of one thread could otherwise overwrite, with a mere ancestor, the
// root another thread has just written.
__global__ __launch_bounds__(256) void k_cc_flatten(int64_t V, int* parent)
{
    for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < V; v += (int64_t)gridDim.x * 256) {
        int cur = (int)v, next;
        while ((next = ld_parent(parent, cur)) != cur) cur = next;
        __hip_atomic_store(parent + v, cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Roots per chunk of kCcChunk voxels.
__global__ __launch_bounds__(256) void k_cc_count(int64_t V, const int* __restrict__ parent, int* __restrict__ counts)
{
    __shared__ int s_n;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kCcChunk;
    int n = 0;
    for (int i = threadIdx.x; i < kCcChunk; i += 256) {
        const int64_t v = base + i;
        if (v < V && parent[v] == (int)v) ++n;
    }
    atomicAdd(&s_n, n);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s_n;
}

// Exclusive prefix sum of n counts in one block; out[n] = total.
__global__ __launch_bounds__(1024) void k_cc_scan(int n, const int* __restrict__ counts, int* __restrict__ out)
{
    __shared__ int s_part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b = t * per, e = min(n, b + per);
    int sum = 0;
    for (int i = b; i < e; ++i) sum += counts[i];
    s_part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int add = t >= off ? s_part[t - off] : 0;
        __syncthreads();
        s_part[t] += add;
        __syncthreads();
    }
    int run = s_part[t] - sum;
    for (int i = b; i < e; ++i) {
        out[i] = run;
        run += counts[i];
    }
    if (t == 1023) out[n] = s_part[1023];
}

// Roots in raster order: rank_at[root] = its rank, roots[rank] = root.
// Thread t of a chunk owns voxels base + 16t .. base + 16t + 15.
__global__ __launch_bounds__(256) void k_cc_place(int64_t V, const int* __restrict__ parent,
                                                  const int* __restrict__ offsets, int* __restrict__ rank_at,
                                                  int* __restrict__ roots)
{
    __shared__ int s_pre[256];
    const int t = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * kCcChunk + (int64_t)t * 16;
    int n = 0;
    for (int i = 0; i < 16; ++i) {
        const int64_t v = v0 + i;
        if (v < V && parent[v] == (int)v) ++n;
    }
    s_pre[t] = n;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int add = t >= off ? s_pre[t - off] : 0;
        __syncthreads();
        s_pre[t] += add;
        __syncthreads();
    }
    int r = offsets[blockIdx.x] + s_pre[t] - n;
    for (int i = 0; i < 16; ++i) {
        const int64_t v = v0 + i;
        if (v < V && parent[v] == (int)v) {
            rank_at[v] = r;
            roots[r] = (int)v;
            ++r;
        }
    }
}

// Component sizes: each thread walks 16 consecutive voxels and adds one
// count per run of equal roots.
__global__ __launch_bounds__(256) void k_cc_sizes(int64_t V, const int* __restrict__ parent,
                                                  const int* __restrict__ rank_at, int* __restrict__ size)
{
    const int64_t v0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 16;
    if (v0 >= V) return;
    const int64_t v1 = v0 + 16 < V ? v0 + 16 : V;
    int cur = parent[v0], run = 0;
    for (int64_t v = v0; v < v1; ++v) {
        const int p = parent[v];
        if (p != cur) {
            atomicAdd(size + rank_at[cur], run);
            cur = p;
            run = 0;
        }
        ++run;
    }
    atomicAdd(size + rank_at[cur], run);
}

// For each component (rank r, first voxel roots[r]), the rank of the last
// neighbour's component, in the reference's scan order (dz -1, 0, 1; the 8
// in-plane offsets), among components that start earlier; -1 when none.
__global__ __launch_bounds__(256) void k_cc_adjacent(int W, int H, int D, int n, const int* __restrict__ roots,
                                                     const int* __restrict__ parent, const int* __restrict__ rank_at,
                                                     int* __restrict__ adj)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    const int64_t v = roots[r];
    const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((int64_t)W * H));
    int a = -1;
    for (int dz = -1; dz <= 1; ++dz) {
        const int zz = z + dz;
        if (zz < 0 || zz >= D) continue;
        for (int k = 0; k < 8; ++k) {
            const int xx = x + c_ox[k], yy = y + c_oy[k];
            if (xx < 0 || xx >= W || yy < 0 || yy >= H) continue;
            const int p = parent[((int64_t)zz * H + yy) * W + xx];
            if (p < v) a = rank_at[p];
        }
    }
    adj[r] = a;
}

__global__ __launch_bounds__(256) void k_cc_apply(int64_t V, const int* __restrict__ parent,
                                                  const int* __restrict__ rank_at, const int32_t* __restrict__ final_label,
                                                  int32_t* __restrict__ out)
{
    for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < V; v += (int64_t)gridDim.x * 256)
        out[v] = final_label[rank_at[parent[v]]];
}

unsigned grid_for(int64_t V)
{
    const int64_t b = (V + 255) / 256;
    return (unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

}  // namespace

hipError_t connect_labels(hipStream_t st, const int32_t* in, int32_t* out, int W, int H, int D, int supsz,
                          int32_t* num_labels)
{
    const int64_t V = (int64_t)W * H * D;
    if (W < 1 || H < 1 || D < 1 || V >= 0x7fffffffLL || !in || !out) return hipErrorInvalidValue;
    const int64_t nchunk = (V + kCcChunk - 1) / kCcChunk;
    int *parent = nullptr, *rank_at = nullptr, *roots = nullptr, *counts = nullptr, *offsets = nullptr;
    int *size = nullptr, *adj = nullptr, *fin = nullptr;
    hipError_t e = asv::dev_malloc((void**)&parent, sizeof(int) * (size_t)V);
    if (e == hipSuccess) e = asv::dev_malloc((void**)&rank_at, sizeof(int) * (size_t)V);
    if (e == hipSuccess) e = asv::dev_malloc((void**)&roots, sizeof(int) * (size_t)V);
    if (e == hipSuccess) e = asv::dev_malloc((void**)&counts, sizeof(int) * (size_t)nchunk);
    if (e == hipSuccess) e = asv::dev_malloc((void**)&offsets, sizeof(int) * (size_t)(nchunk + 1));
    int n = 0;
    std::vector<int> h_size, h_adj, h_fin;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_cc_init, dim3(grid_for(V)), dim3(256), 0, st, V, parent);
        hipLaunchKernelGGL(k_cc_union, dim3(grid_for(V)), dim3(256), 0, st, W, H, D, in, parent);
        hipLaunchKernelGGL(k_cc_flatten, dim3(grid_for(V)), dim3(256), 0, st, V, parent);
        hipLaunchKernelGGL(k_cc_count, dim3((unsigned)nchunk), dim3(256), 0, st, V, parent, counts);
        hipLaunchKernelGGL(k_cc_scan, dim3(1), dim3(1024), 0, st, (int)nchunk, counts, offsets);
        hipLaunchKernelGGL(k_cc_place, dim3((unsigned)nchunk), dim3(256), 0, st, V, parent, offsets, rank_at, roots);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&n, offsets + nchunk, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && n > 0) {
        e = asv::dev_malloc((void**)&size, sizeof(int) * (size_t)n);
        if (e == hipSuccess) e = asv::dev_malloc((void**)&adj, sizeof(int) * (size_t)n);
        if (e == hipSuccess) e = asv::dev_malloc((void**)&fin, sizeof(int) * (size_t)n);
        if (e == hipSuccess) e = hipMemsetAsync(size, 0, sizeof(int) * (size_t)n, st);
        if (e == hipSuccess) {
            const int64_t th = (V + 15) / 16;
            hipLaunchKernelGGL(k_cc_sizes, dim3((unsigned)((th + 255) / 256)), dim3(256), 0, st, V, parent, rank_at,
                               size);
            hipLaunchKernelGGL(k_cc_adjacent, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, W, H, D, n, roots,
                               parent, rank_at, adj);
            e = hipGetLastError();
        }
        h_size.resize((size_t)n);
        h_adj.resize((size_t)n);
        h_fin.resize((size_t)n);
        if (e == hipSuccess) e = hipMemcpyAsync(h_size.data(), size, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipMemcpyAsync(h_adj.data(), adj, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        int next = 0;
        if (e == hipSuccess) {
            // the reference's sequence of segments (Image.cpp:1962-2050)
            int adjl = 0;
            for (int r = 0; r < n; ++r) {
                if (h_adj[(size_t)r] >= 0) adjl = h_fin[(size_t)h_adj[(size_t)r]];
                h_fin[(size_t)r] = h_size[(size_t)r] < supsz / 4 ? adjl : next++;
            }
            e = hipMemcpyAsync(fin, h_fin.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, st);
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_cc_apply, dim3(grid_for(V)), dim3(256), 0, st, V, parent, rank_at, fin, out);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (num_labels) *num_labels = next;
    }
    for (int* p : {parent, rank_at, roots, counts, offsets, size, adj, fin})
        if (p) (void)hipFree(p);
    return e;
}

}  // namespace asv
