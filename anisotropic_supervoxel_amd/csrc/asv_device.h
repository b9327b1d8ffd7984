// asv_device.h -- device helpers shared by the kernel translation units:
// global and LDS pointer forms, the JFA distance of the arithmetic
// specification (flood.gs:26-182) and the selection key.
#pragma once

#include "asv_internal.h"

namespace asv {

// A pointer into global memory (address space 1): loads and stores through
// it compile to global_* instructions rather than flat ones.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

// A byte of LDS read on its own: taps of the JFA passes that are adjacent
// bytes (step 1) would otherwise be merged by the compiler into 16-bit reads
// at odd LDS addresses, which are misaligned and slow.
using lds_byte = const volatile __attribute__((address_space(3))) uint8_t;

// One 16-byte record field through a global pointer.
__device__ __forceinline__ float4 ld_f4(gptr<const float> p)
{
    return make_float4(p[0], p[1], p[2], p[3]);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// A load or store at a wave-uniform base plus a 32-bit byte offset per lane
// (the base formed once per wave in scalar registers, one 64-bit add per
// access instead of the full index arithmetic).
template <typename T>
__device__ __forceinline__ T ld_off(gptr<const T> base, uint32_t byte_off)
{
    return *(gptr<const T>)((gptr<const char>)base + byte_off);
}
template <typename T>
__device__ __forceinline__ void st_off(gptr<T> base, uint32_t byte_off, T v)
{
    *(gptr<T>)((gptr<char>)base + byte_off) = v;
}

// Side of the square column of voxels a JFA block works on.
constexpr int kTile = 16;

// ---------------------------------------------------------------------------
// JFA distance and selection key (flood.gs:26-182)
// ---------------------------------------------------------------------------

// The non-negative part of a quadratic form value that is not NaN: a value
// with its sign bit set (negative, -0) becomes +0.  One integer maximum; the
// candidate rounds of the marching pass use it where the seed stage has
// checked that no quadratic form can be NaN (metric_safe).
__device__ __forceinline__ float nonneg(float q) { return __int_as_float(max(__float_as_int(q), 0)); }

// The distance every other path uses.  flood.gs:90-95 clamps a negative
// quadratic form to 0 (`if(curr_pos_dist<0)`), which leaves a NaN as it
// is, and the NaN distance then loses `curr_dist<min_dist`: here a NaN
// quadratic form (of either sign) gives a +inf distance, which every
// selection rejects as it rejects any distance >= 1e8.
constexpr uint32_t kRejectBits = 0x7F800000u;   // +inf

// Distance of voxel (fx, fy, fz) with colour (l0, l1, l2) to one seed record
// (a = (x, y, z, m0), b = (2m1, 2m2, m3, 2m4), c = (m5, L, a, b)), in the
// fixed order of the specification:
//   u1 = dz*2m2,  u2 = dz*2m4,  u3 = (dz*dz)*m5     (rounded products)
//   t1 = fma(dx, m0, fma(dy, 2m1, u1)),  t2 = fma(dy, m3, u2)
//   q  = fma(dx, t1, fma(dy, t2, u3)): NaN -> +inf distance, else nonneg
//   d  = fma(ratio, q, fma(c2, c2, fma(c1, c1, c0*c0)))
// The u terms depend on the seed and the slice only, so a block working on
// one slice can compute them once per seed.
__device__ __forceinline__ float cand_dist(const float4 a, const float4 b, const float4 c, float fx, float fy,
                                           float fz, float l0, float l1, float l2, float ratio)
{
    const float dx = fx - a.x, dy = fy - a.y, dz = fz - a.z;
    const float u1 = dz * b.y, u2 = dz * b.w, u3 = (dz * dz) * c.x;
    const float t1 = __fmaf_rn(dx, a.w, __fmaf_rn(dy, b.x, u1));
    const float t2 = __fmaf_rn(dy, b.z, u2);
    const float qr = __fmaf_rn(dx, t1, __fmaf_rn(dy, t2, u3));
    // a select, not a branch: +inf makes d = +inf (ratio > 0)
    const float q = qr != qr ? __uint_as_float(kRejectBits) : nonneg(qr);
    const float c0 = l0 - c.y, c1 = l1 - c.z, c2 = l2 - c.w;
    const float cd = __fmaf_rn(c2, c2, __fmaf_rn(c1, c1, c0 * c0));
    return __fmaf_rn(ratio, q, cd);
}

// Selection key for the lowest-id rule: distances are >= +0 (or NaN), so the
// float bits order like the values; (d, id) lexicographic minimum = u64 min.
// The initial key (1e8, 0) rejects d >= 1e8 exactly as `d < min_dist` with
// min_dist = 1e8 does (flood.gs:63, :95).
__device__ __forceinline__ unsigned long long dkey(float d, int id)
{
    return ((unsigned long long)__float_as_uint(d) << 32) | (unsigned)id;
}

// Minimum of two selection keys in one v_min_f64 (instead of a 64-bit
// compare and two selects).  With ratio >= 0 a distance is +0 .. +inf or the
// canonical NaN 0x7FC00000, so a key's high word is below 0x7FF00000: read as
// a double it is a non-negative number, never NaN, and non-negative doubles
// (denormals included; the kernels keep fp64 denormals) order exactly like
// their bit patterns.  The double minimum is therefore the u64 minimum.
__device__ __forceinline__ unsigned long long key_min(unsigned long long a, unsigned long long b)
{
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(__longlong_as_double((long long)a)),
        "v"(__longlong_as_double((long long)b)));
    return (unsigned long long)__double_as_longlong(r);
}
// Whether coordinate c has coordinate a among its taps c - k, c, c + k
// (each clamped to [0, n - 1]) along one axis.
__device__ __forceinline__ bool axis_hit(int c, int a, int n, int k)
{
    return clampi(c - k, 0, n - 1) == a || c == a || clampi(c + k, 0, n - 1) == a;
}

// The idx-th coordinate v in [0, n) (ascending) whose taps v - k, v, v + k,
// clamped to [0, n - 1], include p; false when there are fewer.  Such v lie
// within k of p: v = p - k o exactly, or the clamped tap at an edge p = 0
// (v <= k) or p = n - 1 (v >= n - 1 - k).
__device__ __forceinline__ bool seed_axis(int p, int n, int k, int idx, int* v)
{
    const int lo = max(0, p - k), hi = min(n - 1, p + k);
    for (int c = lo; c <= hi; ++c) {
        const bool hit = clampi(c - k, 0, n - 1) == p || c == p || clampi(c + k, 0, n - 1) == p;
        if (hit && idx-- == 0) {
            *v = c;
            return true;
        }
    }
    return false;
}

#define ASV_KEY_INIT (((unsigned long long)0x4CBEBC20u) << 32)   /* bits of 1e8f */

}  // namespace asv
