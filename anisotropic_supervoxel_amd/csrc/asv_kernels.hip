// asv_kernels.hip -- gfx950 kernels of the supervoxel hot path.
//
// Arithmetic follows one fixed specification shared with the CPU oracle:
// no contraction (built with -ffp-contract=off), IEEE division and square
// root, exact integer centroid sums, lowest-seed-id tie breaking (or the
// reference's scan order).  Every kernel is a streaming or gather kernel over
// int32 labels and packed u8 LAB: none of it is a dense contraction, so there
// is no MFMA work here; the bound is HBM bandwidth.
//
// Reference passes replaced (paths relative to the reference root):
//   k_prepare             shaders/RGB2LAB.fs:11-79 (blur + flip of Render.cpp:354-359)
//   k_seed_init           shaders/draw_seeds.vs:20-34 (colour at the seed voxel)
//   k_seed_metric(_rings) Image.cpp:689-872 and :1351-1534, with matrix_slover
//                         (Image.cpp:874-1009) evaluated at the seeds and rings
//   k_fill + k_place      draw_seeds / draw_next_seeds (clear -1, later id wins)
//   k_jfa                 shaders/flood.gs:26-182
//   k_reduce_rows         shaders/centroids.vs:19-53 with additive blending
//   k_finalize            shaders/aver_exception.vs:22-81

#include "asv_device.h"
#include "asv_march.h"

#include <cstdlib>


namespace asv {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------


__device__ __forceinline__ int reflect101(int i, int n)
{
    if (n == 1) return 0;
    if (i < 0) return -i;
    if (i >= n) return 2 * n - 2 - i;
    return i;
}

// ---------------------------------------------------------------------------
// LAB (RGB2LAB.fs:11-79)
// ---------------------------------------------------------------------------

// Cube root of the white-normalised X, Y or Z: the same float as the
// specification's five Halley steps in double from 0.45 + 0.55 r, rounded to
// float, for every ratio the LAB kernel meets, at a fifth of the double
// work: an inverse cube root z ~ r^(-1/3)
// from the exponent bits, two Newton steps z += z (1 - r z^3) / 3 in float
// and two in double, then r z^2 rounded to float.  Checked equal, bit for
// bit, on all 3 x 2^24 ratios of the 2^24 RGB inputs (and the LAB volume of
// every RGB triple is compared with the oracle's).
__device__ __forceinline__ float cbrt_fast(float r)
{
    float z = __uint_as_float(0x54a2fa8cu - __float_as_uint(r) / 3u);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const float z3 = z * z * z;
        z = z + z * ((1.0f - r * z3) * 0.333333343f);
    }
    double zd = z;
    const double rd = r;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double z3 = zd * zd * zd;
        zd = zd + zd * ((1.0 - rd * z3) * (1.0 / 3.0));
    }
    return (float)(rd * zd * zd);
}

__device__ __forceinline__ float lab_f(float r)
{
    if (r > 0.008856f) return cbrt_fast(r);
    return (903.3f * r + 16.0f) / 116.0f;
}

__device__ __forceinline__ uint32_t round_u8(float v)
{
    float r = floorf(v + 0.5f);
    r = (r >= 0.0f) ? r : 0.0f;
    r = (r > 255.0f) ? 255.0f : r;
    return (uint32_t)r;
}

__device__ __forceinline__ uint32_t lab_pack(uint32_t r, uint32_t g, uint32_t b, const float* lin)
{
    float tr = lin[r], tg = lin[g], tb = lin[b];
    float X = (tr * 0.412453f + tg * 0.357580f) + tb * 0.180423f;
    float Y = (tr * 0.212671f + tg * 0.715160f) + tb * 0.072169f;
    float Z = (tr * 0.019334f + tg * 0.119193f) + tb * 0.950227f;
    float fx = lab_f(X / 0.950456f);
    float fy = lab_f(Y / 1.0f);
    float fz = lab_f(Z / 1.088754f);
    float L = 116.0f * fy - 16.0f;
    float A = 500.0f * (fx - fy);
    float B = 200.0f * (fy - fz);
    return round_u8(L * 2.55f) | (round_u8(A + 128.0f) << 8) | (round_u8(B + 128.0f) << 16);
}

// ---------------------------------------------------------------------------
// per-voxel metric (Image.cpp:874-1009)
// ---------------------------------------------------------------------------

__device__ __constant__ int c_expected_b[9] = {3, 2, 3, 2, 1, 2, 3, 2, 3};

// Upper triangle of N = A^T A (21 entries, row-major i<=j) and c = A^T b,
// accumulated row by row in the reference's neighbour order (offset_y outer,
// offset_x inner, forward row before backward row).  Entries are exact
// integers.  Returns the row count.
__device__ __forceinline__ int normal_equations(const float2* __restrict__ fwd, const float2* __restrict__ bwd,
                                                int W, int H, int D, int x, int y, int t,
                                                double N[21], double c[6])
{
#pragma unroll
    for (int i = 0; i < 21; ++i) N[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) c[i] = 0.0;
    int rows = 0;
    const size_t frame = (size_t)t * H * W;
    // every flow tap is loaded before the first row is accumulated, so the
    // 18 loads are in flight together
    float2 ff[9], fb[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int nx = x + k % 3 - 1, ny = y + k / 3 - 1;
        const bool in = nx >= 0 && nx < W && ny >= 0 && ny < H;
        const size_t pix = frame + (size_t)(H - 1 - ny) * W + nx;   // image row of GL row ny
        ff[k] = in && t != D - 1 ? fwd[pix] : make_float2(0.0f, 0.0f);
        fb[k] = in && t != 0 ? bwd[pix] : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int m = -1; m <= 1; ++m) {
#pragma unroll
        for (int n = -1; n <= 1; ++n) {
            int nx = x + n, ny = y + m;
            if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
            double b = (double)c_expected_b[(m + 1) * 3 + (n + 1)];
#pragma unroll
            for (int dir = 0; dir < 2; ++dir) {
                if (dir == 0 && t == D - 1) continue;
                if (dir == 1 && t == 0) continue;
                float2 f = dir == 0 ? ff[(m + 1) * 3 + (n + 1)] : fb[(m + 1) * 3 + (n + 1)];
                float fu = f.x, fv = -f.y;
                float dx = (float)n + floorf(fu + 0.5f);
                float dy = (float)m + floorf(fv + 0.5f);
                double sgn = dir == 0 ? 2.0 : -2.0;
                double r[6];
                r[0] = (double)dx * (double)dx;
                r[1] = 2.0 * (double)dx * (double)dy;
                r[2] = sgn * (double)dx;
                r[3] = (double)dy * (double)dy;
                r[4] = sgn * (double)dy;
                r[5] = 1.0;
                int q = 0;
#pragma unroll
                for (int i = 0; i < 6; ++i) {
#pragma unroll
                    for (int j = i; j < 6; ++j) { N[q] += r[i] * r[j]; ++q; }
                    c[i] += r[i] * b;
                }
                ++rows;
            }
        }
    }
    return rows;
}

__device__ __forceinline__ int tri(int i, int j) { return i <= j ? i * 6 - (i * (i - 1)) / 2 + (j - i) : j * 6 - (j * (j - 1)) / 2 + (i - j); }

// LDL^T without pivoting; returns false when a pivot collapses (rank
// deficiency), in which case the pseudo-inverse path must solve.
__device__ __forceinline__ bool ldl_solve(const double N[21], const double c[6], double m[6])
{
    double L[6][6];
    double d[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double s = N[tri(j, j)];
#pragma unroll
        for (int k = 0; k < j; ++k) s = s - (L[j][k] * L[j][k]) * d[k];
        if (!(s > 1e-9 * N[tri(j, j)])) return false;
        d[j] = s;
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
            double t = N[tri(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) t = t - (L[i][k] * L[j][k]) * d[k];
            L[i][j] = t / s;
        }
    }
    double y[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double t = c[i];
#pragma unroll
        for (int k = 0; k < i; ++k) t = t - L[i][k] * y[k];
        y[i] = t;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) y[i] = y[i] / d[i];
#pragma unroll
    for (int i = 5; i >= 0; --i) {
        double t = y[i];
#pragma unroll
        for (int k = i + 1; k < 6; ++k) t = t - L[k][i] * m[k];
        m[i] = t;
    }
    return true;
}

// Cyclic Jacobi on a symmetric n x n matrix (double); eigenvalues end on the
// diagonal of a, eigenvectors in the columns of v.
template <int n>
__device__ void jacobi_sym(double* a, double* v, int sweeps)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) v[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int sw = 0; sw < sweeps; ++sw) {
        int any = 0;
        for (int p = 0; p < n - 1; ++p) {
            for (int q = p + 1; q < n; ++q) {
                double apq = a[p * n + q];
                if (apq == 0.0) continue;
                any = 1;
                double app = a[p * n + p], aqq = a[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t;
                if (fabs(theta) > 1e150) t = 0.5 / theta;
                else {
                    double den = fabs(theta) + sqrt(theta * theta + 1.0);
                    t = 1.0 / den;
                    if (theta < 0.0) t = -t;
                }
                double cs = 1.0 / sqrt(t * t + 1.0);
                double sn = t * cs;
                for (int k = 0; k < n; ++k) {
                    double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = cs * akp - sn * akq;
                    a[k * n + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = cs * apk - sn * aqk;
                    a[q * n + k] = sn * apk + cs * aqk;
                }
                a[p * n + q] = 0.0;
                a[q * n + p] = 0.0;
                for (int k = 0; k < n; ++k) {
                    double vkp = v[k * n + p], vkq = v[k * n + q];
                    v[k * n + p] = cs * vkp - sn * vkq;
                    v[k * n + q] = sn * vkp + cs * vkq;
                }
            }
        }
        if (!any) break;
    }
}

__device__ __forceinline__ int psd3(float m0, float m1, float m2, float m3, float m4, float m5)
{
    double a = m0, b = m1, c = m2, d = m3, e = m4, f = m5;
    if (a < 0.0 || d < 0.0 || f < 0.0) return 0;
    if (a * d - b * b < 0.0) return 0;
    if (a * f - c * c < 0.0) return 0;
    if (d * f - e * e < 0.0) return 0;
    double det = a * (d * f - e * e) - b * (b * f - e * c) + c * (b * e - d * c);
    if (det < 0.0) return 0;
    return 1;
}

// Minimum-norm solution of voxel (x, y, t)'s rank-deficient system through
// the eigendecomposition of N (pseudo-inverse, eigenvalues below 1e-9 of
// the largest dropped).  Kept out of line, and it rebuilds the normal
// equations itself: only voxels whose LDL^T pivot collapses take it, and
// neither its 6 x 6 work arrays nor the address of the caller's system may
// weigh on every caller (an address passed down would keep the caller's
// accumulators in scratch memory).
__device__ __noinline__ void pinv_solve(const float2* __restrict__ fwd, const float2* __restrict__ bwd, int W, int H,
                                        int D, int x, int y, int t, double* m)
{
    double N[21], c[6];
    normal_equations(fwd, bwd, W, H, D, x, y, t, N, c);
    double a[36], ev[36];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) a[i * 6 + j] = N[tri(i, j)];
    jacobi_sym<6>(a, ev, 30);
    double lmax = 0.0;
    for (int i = 0; i < 6; ++i) lmax = a[i * 6 + i] > lmax ? a[i * 6 + i] : lmax;
    for (int k = 0; k < 6; ++k) m[k] = 0.0;
    for (int e = 0; e < 6; ++e) {
        double lam = a[e * 6 + e];
        if (!(lam > 1e-9 * lmax)) continue;
        double proj = 0.0;
        for (int k = 0; k < 6; ++k) proj = proj + ev[k * 6 + e] * c[k];
        double coef = proj / lam;
        for (int k = 0; k < 6; ++k) m[k] = m[k] + ev[k * 6 + e] * coef;
    }
}

// The metric of voxel (x, y, t) (matrix_slover, GPU-AS/Image.cpp:874-1009):
// the least-squares 6-vector rounded to float into m, and whether the 3 x 3
// matrix it spells is positive semi-definite.  No rows (a 1 x 1 frame of a
// one-frame clip) gives the zero matrix.  The reference evaluates it only at
// the seeds and their rings, every iteration; so does this build (the seed
// kernels call it), instead of keeping a field for every voxel.
__device__ __forceinline__ int voxel_metric(const float2* __restrict__ fwd, const float2* __restrict__ bwd, int W,
                                            int H, int D, int x, int y, int t, float m[6])
{
    double N[21], c[6], md[6];
    const int rows = normal_equations(fwd, bwd, W, H, D, x, y, t, N, c);
    if (rows == 0) {
#pragma unroll
        for (int i = 0; i < 6; ++i) md[i] = 0.0;
    } else if (!ldl_solve(N, c, md)) {
        double mp[6];
        pinv_solve(fwd, bwd, W, H, D, x, y, t, mp);
#pragma unroll
        for (int i = 0; i < 6; ++i) md[i] = mp[i];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) m[i] = (float)md[i];
    return psd3(m[0], m[1], m[2], m[3], m[4], m[5]);
}

// (R, G, B) of voxel (x, y, t), GL row order, as the reference's frame
// texture holds it (Render::cvMat2TexureInput, GPU-AS/Render.cpp:354-359):
// GaussianBlur 3x3 [1,2,1]^2/16 with reflect-101 borders and round half up,
// on the image row that becomes GL row y after the vertical flip; frames
// already prepared are read as they are.
__device__ __forceinline__ uint3 frame_rgb(const uint8_t* __restrict__ rgb, int prepared, int W, int H, int x, int y,
                                           int t)
{
    if (prepared) {
        const uint8_t* p = rgb + (((size_t)t * H + y) * W + x) * 3;
        return make_uint3(p[0], p[1], p[2]);
    }
    const uint8_t* fr = rgb + (size_t)t * H * W * 3;
    const int yi = H - 1 - y;
    uint32_t ar = 0, ag = 0, ab = 0;
#pragma unroll
    for (int j = -1; j <= 1; ++j) {
        const int yy = reflect101(yi + j, H);
        const uint32_t wj = j == 0 ? 2u : 1u;
#pragma unroll
        for (int i = -1; i <= 1; ++i) {
            const int xx = reflect101(x + i, W);
            const uint32_t w = wj * (i == 0 ? 2u : 1u);
            const uint8_t* p = fr + ((size_t)yy * W + xx) * 3;
            ar += w * p[0]; ag += w * p[1]; ab += w * p[2];
        }
    }
    return make_uint3((ar + 8u) >> 4, (ag + 8u) >> 4, (ab + 8u) >> 4);
}

// Kernel (a): blur + flip + LAB for one voxel per thread, 64 x 4 voxels of
// one frame per block.  Raw frames: the six image rows the block's blur
// reads (its four rows and one above and below, reflect-101 applied) are
// staged in LDS over 66 columns as packed R | G << 8 | B << 16, once per
// block; each thread then sums its nine taps with R and B side by side in
// one word (a weighted sum of 9 bytes stays below 2^16).  The integer
// result equals frame_rgb's.
__global__ __launch_bounds__(256) void k_prepare(int W, int H, int D, const uint8_t* __restrict__ rgb,
                                                 int prepared, const float* __restrict__ lin_lut,
                                                 uint32_t* __restrict__ lab)
{
    __shared__ float lin[256];
    __shared__ uint32_t s_px[6][66];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = ty * 64 + tx;
    lin[tid] = lin_lut[tid];
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 4, t = blockIdx.z;
    if (!prepared) {
        // staged row j is image row H - y0 - j; GL row y0 + ty is image row
        // H - y0 - (ty + 1), so its taps are staged rows ty .. ty + 2.  The
        // clamps only matter for rows and columns no valid voxel reads.
        const uint8_t* fr = rgb + (size_t)t * H * W * 3;
        for (int e = tid; e < 6 * 66; e += 256) {
            const int j = e / 66, c = e - (e / 66) * 66;
            const int yy = clampi(reflect101(H - y0 - j, H), 0, H - 1);
            const int xx = clampi(reflect101(x0 - 1 + c, W), 0, W - 1);
            const uint8_t* q = fr + ((size_t)yy * W + xx) * 3;
            s_px[j][c] = (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16);
        }
    }
    __syncthreads();

    const int x = x0 + tx, y = y0 + ty;
    if (x >= W || y >= H) return;
    size_t v = ((size_t)t * H + y) * W + x;
    uint32_t r, g, b;
    if (prepared) {
        const uint8_t* q = rgb + v * 3;
        r = q[0]; g = q[1]; b = q[2];
    } else {
        uint32_t rb = 0u, gg = 0u;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const uint32_t w = (j == 1 ? 2u : 1u) * (i == 1 ? 2u : 1u);
                const uint32_t px = s_px[ty + j][tx + i];
                rb += w * (px & 0x00FF00FFu);
                gg += w * ((px >> 8) & 0xFFu);
            }
        r = ((rb & 0xFFFFu) + 8u) >> 4;
        g = (gg + 8u) >> 4;
        b = ((rb >> 16) + 8u) >> 4;
    }
    lab[v] = lab_pack(r, g, b, lin);
}

// The metric of every voxel (m0..m5, psd flag, 0 per voxel), for tests and
// tools; the Lloyd loop itself evaluates the metric only where the seed
// kernels need it.
__global__ __launch_bounds__(256) void k_metric_field(int W, int H, int D, const float2* __restrict__ fwd,
                                                      const float2* __restrict__ bwd, float4* __restrict__ field)
{
    const int64_t V = (int64_t)W * H * D;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(v % W), y = (int)((v / W) % H), t = (int)(v / ((int64_t)W * H));
        float m[6];
        const int p = voxel_metric(fwd, bwd, W, H, D, x, y, t, m);
        field[2 * v] = make_float4(m[0], m[1], m[2], m[3]);
        field[2 * v + 1] = make_float4(m[4], m[5], p ? 1.0f : 0.0f, 0.0f);
    }
}

// ---------------------------------------------------------------------------
// density-based seed re-placement (GPU-AS/Image.cpp:350-620,
// shaders/derivative_t.gs:20-53, shaders/density_grid.vs:13-42)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float f16_round(float x) { return (float)(_Float16)x; }

__device__ __forceinline__ uint32_t abs_diff_sum(uint3 a, uint3 b)
{
    return (a.x > b.x ? a.x - b.x : b.x - a.x) + (a.y > b.y ? a.y - b.y : b.y - a.y) +
           (a.z > b.z ? a.z - b.z : b.z - a.z);
}

// Per voxel: g = 0.33 sum|c - c_left| + 0.33 sum|c - c_up| + 0.33
// sum|c - c_prev| (GL rows, clamped taps; each product and sum rounded to
// float, stored as float16 like the RGB16F target), added as an exact
// multiple of 2^-12 into the slot that the voxel's float16 grid-seed id
// addresses through density_grid.vs's float arithmetic (a point landing
// outside the row is dropped, as the rasteriser would).  The grid id is the
// separable box map (mz, my, mx).  Each thread walks 16 voxels of a row and
// adds one atomic per run of equal slots.
__global__ __launch_bounds__(256) void k_density(int W, int H, int D, const uint8_t* __restrict__ rgb, int prepared,
                                                 const int* __restrict__ mx, const int* __restrict__ my,
                                                 const int* __restrict__ mz, int wnum, int hnum, int nslot, float wr,
                                                 unsigned long long* __restrict__ acc)
{
    const int segs = (W + 15) / 16;
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= (int64_t)segs * H * D) return;
    const int seg = (int)(gid % segs);
    const int y = (int)((gid / segs) % H);
    const int t = (int)(gid / ((int64_t)segs * H));
    const int x0 = seg * 16, x1 = min(W, x0 + 16);
    if (mz[t] < 0 || my[y] < 0) return;
    const int row_id = (mz[t] * hnum + my[y]) * wnum;
    int cur = -1;
    unsigned long long run = 0ull;
    for (int x = x0; x < x1; ++x) {
        const uint3 c = frame_rgb(rgb, prepared, W, H, x, y, t);
        const uint32_t s1 = abs_diff_sum(c, frame_rgb(rgb, prepared, W, H, x > 0 ? x - 1 : 0, y, t));
        const uint32_t s2 = abs_diff_sum(c, frame_rgb(rgb, prepared, W, H, x, y > 0 ? y - 1 : 0, t));
        const uint32_t s3 = abs_diff_sum(c, frame_rgb(rgb, prepared, W, H, x, y, t > 0 ? t - 1 : 0));
        float g = __fmul_rn(0.33f, (float)s1);
        g = __fadd_rn(g, __fmul_rn(0.33f, (float)s2));
        g = __fadd_rn(g, __fmul_rn(0.33f, (float)s3));
        const float pg = f16_round(g);
        const float kr = f16_round((float)(row_id + mx[x]));
        const float ty = floorf(__fmul_rn(kr, wr));
        const float tx = floorf(__fsub_rn(kr, __fdiv_rn(ty, wr)));
        int slot = -1;
        if (mx[x] >= 0 && tx >= 0.0f && tx < (float)W) {
            slot = (int)ty * W + (int)tx;
            if (slot >= nslot) slot = -1;
        }
        if (slot != cur) {
            if (cur >= 0 && run) atomicAdd(acc + cur, run);
            cur = slot;
            run = 0ull;
        }
        run += (unsigned long long)__fmul_rn(pg, 4096.0f);
    }
    if (cur >= 0 && run) atomicAdd(acc + cur, run);
}

// ---------------------------------------------------------------------------
// seeds
// ---------------------------------------------------------------------------

// Iteration 1: grid positions, colour sampled at the seed voxel.
__global__ void k_seed_init(int W, int H, int A, const int4* __restrict__ grid, const uint32_t* __restrict__ lab,
                            int4* __restrict__ pos, float4* __restrict__ slab)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A) return;
    int4 g = grid[i];
    uint32_t pk = lab[((size_t)g.z * H + g.y) * W + g.x];
    pos[i] = g;
    slab[i] = make_float4((float)(pk & 255u), (float)((pk >> 8) & 255u), (float)((pk >> 16) & 255u), 0.0f);
}

__device__ __constant__ int c_ring1x[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
__device__ __constant__ int c_ring1y[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
__device__ __constant__ int c_ring2x[16] = {-2, -1, 0, 1, 2, -2, 2, -2, 2, -2, 2, -2, -1, 0, 1, 2};
__device__ __constant__ int c_ring2y[16] = {-2, -2, -2, -2, -2, -1, -1, 0, 0, 1, 1, 2, 2, 2, 2, 2};

// acos(cos)*180/PI < 90 under a correctly rounded float acos holds exactly
// when cos exceeds this float (bit pattern 0x328885a3).
#define ASV_GATE_COS_MIN 1.5893254712295857e-08f

__device__ __forceinline__ bool angle_gate(float su, float sv, float nu, float nv)
{
    float num = su * nu + sv * nv;
    double den = sqrt((double)su * (double)su + (double)sv * (double)sv) *
                 sqrt((double)nu * (double)nu + (double)nv * (double)nv);
    float cs = (float)((double)num / den);
    return cs > ASV_GATE_COS_MIN && cs <= 1.0f;
}

__device__ __forceinline__ float2 flow_gl(const float2* __restrict__ F, int W, int H, int x, int y, int t)
{
    float2 f = F[((size_t)t * H + (H - 1 - y)) * W + x];
    return make_float2(f.x, -f.y);
}

__device__ void nearest_psd(const float in[6], float out[6])
{
    double a[9] = {in[0], in[1], in[2], in[1], in[3], in[4], in[2], in[4], in[5]};
    double v[9];
    jacobi_sym<3>(a, v, 30);
    double lam[3];
    for (int i = 0; i < 3; ++i) lam[i] = a[i * 3 + i] < 0.0 ? 0.0 : a[i * 3 + i];
    double r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int e = 0; e < 3; ++e) s = s + (v[i * 3 + e] * lam[e]) * v[j * 3 + e];
            r[i * 3 + j] = s;
        }
    out[0] = (float)r[0]; out[1] = (float)r[1]; out[2] = (float)r[2];
    out[3] = (float)r[4]; out[4] = (float)r[5]; out[5] = (float)r[8];
}

// Whether no quadratic form with this metric can be NaN: every entry finite
// and at most 2^60 in magnitude.  Voxel offsets stay below 2^29 (volumes
// hold at most 2^30 voxels), so every product and sum of the distance then
// stays below 2^124, finite, and the rounds' one-instruction clamp (nonneg)
// is exact.
__device__ __forceinline__ bool metric_safe(const float m[6])
{
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) ok = ok && fabsf(m[k]) <= 1.152921504606846976e18f;   // NaN fails
    return ok;
}

// Writes seed i's record; a metric that fails metric_safe raises *unsafe,
// which sends the run's marching passes down the exact per-tap path.
__device__ __forceinline__ void write_seed(int i, int4 p, float4 c, const float m[6], SeedRec* __restrict__ rec,
                                           float* __restrict__ trace_iter, int* __restrict__ unsafe)
{
    if (!metric_safe(m)) atomicOr(unsafe, 1);
    SeedRec r;
    r.a = make_float4((float)p.x, (float)p.y, (float)p.z, m[0]);
    r.b = make_float4(2.0f * m[1], 2.0f * m[2], m[3], 2.0f * m[4]);
    r.c = make_float4(m[5], c.x, c.y, c.z);
    rec[i] = r;
    if (trace_iter) {
        float* tr = trace_iter + (size_t)i * 12;
        tr[0] = (float)p.x; tr[1] = (float)p.y; tr[2] = (float)p.z;
        tr[3] = c.x; tr[4] = c.y; tr[5] = c.z;
        for (int k = 0; k < 6; ++k) tr[6 + k] = m[k];
    }
}

// Metric of every seed, first part: the metric at the seed's voxel.  A PSD
// one is the seed's metric and its record is written here; otherwise the
// voxel's metric is parked in `own` (two float4 per seed) and `need` asks
// k_seed_metric_rings for the ring fallbacks.
__global__ __launch_bounds__(64) void k_seed_metric(int W, int H, int D, int A, const int4* __restrict__ pos,
                                                    const float4* __restrict__ slab, const float2* __restrict__ fwd,
                                                    const float2* __restrict__ bwd, SeedRec* __restrict__ rec,
                                                    float4* __restrict__ own, int* __restrict__ need,
                                                    float* __restrict__ trace_iter, int* __restrict__ unsafe)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A) return;
    const int4 p = pos[i];
    float m[6];
    const int ok = voxel_metric(fwd, bwd, W, H, D, p.x, p.y, p.z, m);
    need[i] = ok ? 0 : 1;
    if (ok) {
        write_seed(i, p, slab[i], m, rec, trace_iter, unsafe);
    } else {
        own[2 * i] = make_float4(m[0], m[1], m[2], m[3]);
        own[2 * i + 1] = make_float4(m[4], m[5], 0.0f, 0.0f);
    }
}

// Second part, one wave per seed, only for seeds whose own metric is not
// PSD: lanes 0..7 evaluate the 1-ring and lanes 8..23 the 2-ring on the
// seed's frame together (in bounds and, when `gate`, through the flow-angle
// gate); lane 0 then takes the mean of the PSD 1-ring metrics, else of the
// PSD 2-ring ones, each summed in ring order, else the nearest PSD matrix
// of the seed's own metric.
__global__ __launch_bounds__(64) void k_seed_metric_rings(int W, int H, int D, const int4* __restrict__ pos,
                                                          const float4* __restrict__ slab,
                                                          const float2* __restrict__ fwd,
                                                          const float2* __restrict__ bwd, int gate,
                                                          SeedRec* __restrict__ rec, const float4* __restrict__ own,
                                                          const int* __restrict__ need, float* __restrict__ trace_iter,
                                                          int* __restrict__ unsafe)
{
    const int i = blockIdx.x;
    if (!need[i]) return;
    __shared__ float s_m[24][6];
    __shared__ int s_ok[24];
    const int lane = threadIdx.x;
    const int4 p = pos[i];
    if (lane < 24) {
        const int nx = p.x + (lane < 8 ? c_ring1x[lane] : c_ring2x[lane - 8]);
        const int ny = p.y + (lane < 8 ? c_ring1y[lane] : c_ring2y[lane - 8]);
        int ok = nx >= 0 && nx < W && ny >= 0 && ny < H;
        if (ok && gate) {
            float2 sf = make_float2(0.0f, 0.0f), nf = make_float2(0.0f, 0.0f);
            if (p.z != D - 1) {
                sf = flow_gl(fwd, W, H, p.x, p.y, p.z);
                nf = flow_gl(fwd, W, H, nx, ny, p.z);
            }
            ok = angle_gate(sf.x, sf.y, nf.x, nf.y);
        }
        if (ok) {
            float m[6];
            ok = voxel_metric(fwd, bwd, W, H, D, nx, ny, p.z, m);
            for (int k = 0; k < 6; ++k) s_m[lane][k] = m[k];
        }
        s_ok[lane] = ok;
    }
    __syncthreads();
    if (lane != 0) return;
    float sum[6] = {0, 0, 0, 0, 0, 0};
    int cnt = 0;
    for (int ring = 0; ring < 2 && cnt == 0; ++ring) {
        const int b = ring == 0 ? 0 : 8, e = ring == 0 ? 8 : 24;
        for (int j = b; j < e; ++j) {
            if (!s_ok[j]) continue;
            for (int k = 0; k < 6; ++k) sum[k] = sum[k] + s_m[j][k];
            ++cnt;
        }
    }
    float m[6];
    if (cnt > 0) {
        for (int k = 0; k < 6; ++k) m[k] = sum[k] / (float)cnt;
    } else {
        const float4 o0 = own[2 * i], o1 = own[2 * i + 1];
        const float self[6] = {o0.x, o0.y, o0.z, o0.w, o1.x, o1.y};
        nearest_psd(self, m);
    }
    write_seed(i, p, slab[i], m, rec, trace_iter, unsafe);
}

// ---------------------------------------------------------------------------
// scatter
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_fill(int4* __restrict__ p4, int64_t n4, int32_t* __restrict__ tail,
                                              int64_t ntail, int32_t value)
{
    int4 val = make_int4(value, value, value, value);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x)
        p4[i] = val;
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ntail) tail[g] = value;
}

__global__ void k_zero_u64(uint64_t* __restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

__global__ void k_zero_i32(int* __restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Seed scatter: on a voxel collision the higher id wins, as the later
// primitive does under GL's in-order rasterisation.
__global__ void k_place(int W, int H, int A, const int4* __restrict__ pos, int32_t* __restrict__ labels)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A) return;
    int4 p = pos[i];
    atomicMax(&labels[((size_t)p.z * H + p.y) * W + p.x], i);
}

// ---------------------------------------------------------------------------
// JFA pass (flood.gs:26-182)
// ---------------------------------------------------------------------------


// One voxel of a pass, taps read from global memory in the reference's
// scan order (slab z, z-k, z+k; rows y-k, y, y+k; columns x-k, x, x+k).  A
// tap repeating one of the last two evaluated ids is skipped: its distance
// is identical, so neither tie rule can change the winner.
template <int TIE>
__device__ __forceinline__ int32_t jfa_voxel(int W, int H, int D, const int32_t* __restrict__ in,
                                             const uint32_t* __restrict__ lab, const SeedRec* __restrict__ rec,
                                             int k, float ratio, int x, int y, int z)
{
    const size_t plane = (size_t)W * H;
    const uint32_t pk = lab[(size_t)z * plane + (size_t)y * W + x];
    const float l0 = (float)(pk & 255u), l1 = (float)((pk >> 8) & 255u), l2 = (float)((pk >> 16) & 255u);
    const float fx = (float)x, fy = (float)y, fz = (float)z;
    const int xs[3] = {clampi(x - k, 0, W - 1), x, clampi(x + k, 0, W - 1)};
    const int ys[3] = {clampi(y - k, 0, H - 1), y, clampi(y + k, 0, H - 1)};
    const int zs[3] = {z, clampi(z - k, 0, D - 1), clampi(z + k, 0, D - 1)};

    unsigned long long best = ASV_KEY_INIT;
    float bestd = 100000000.0f;
    int32_t best_id = -1;
    int32_t seen0 = -1, seen1 = -1;
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) {
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int32_t* row = in + (size_t)zs[sl] * plane + (size_t)ys[r] * W;
#pragma unroll
            for (int cidx = 0; cidx < 3; ++cidx) {
                const int32_t id = row[xs[cidx]];
                if (id < 0 || id == seen0 || id == seen1) continue;
                seen1 = seen0;
                seen0 = id;
                const SeedRec sr = rec[id];
                const float d = cand_dist(sr.a, sr.b, sr.c, fx, fy, fz, l0, l1, l2, ratio);
                if (TIE == 0) {
                    const unsigned long long kk = dkey(d, id);
                    best = kk < best ? kk : best;
                } else if (d < bestd) {
                    bestd = d;
                    best_id = id;
                }
            }
        }
    }
    if (TIE == 0) best_id = best == ASV_KEY_INIT ? -1 : (int32_t)(unsigned)(best & 0xffffffffull);
    return best_id;
}

// Generic pass: one voxel per thread, 64 x 4 voxels per block.  Serves any
// step and the scan-order tie rule.
template <int TIE>
__global__ __launch_bounds__(256) void k_jfa(int W, int H, int D, const int32_t* __restrict__ in,
                                             int32_t* __restrict__ out, const uint32_t* __restrict__ lab,
                                             const SeedRec* __restrict__ rec, int k, float ratio)
{
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= W || y >= H) return;
    out[((size_t)z * H + y) * W + x] = jfa_voxel<TIE>(W, H, D, in, lab, rec, k, ratio, x, y, z);
}

// First pass of an iteration, over a volume holding only the scattered
// seeds (every other voxel -1, as the seed stage leaves it): a voxel finds a
// label among its 27 clamped taps only if some seed voxel p is one of them,
// which fixes each coordinate of the voxel to the few values seed_axis
// enumerates (at most max(3, k + 1) per axis).  One thread per (seed, such
// voxel) evaluates that voxel exactly as the full pass does; every other
// voxel keeps the -1 the output was filled with.  A voxel near several seeds
// is written by each with the same value.
template <int TIE>
__global__ __launch_bounds__(256) void k_jfa_seeds(int W, int H, int D, int A, int k, int m,
                                                   const int4* __restrict__ pos, const int32_t* __restrict__ in,
                                                   int32_t* __restrict__ out, const uint32_t* __restrict__ lab,
                                                   const SeedRec* __restrict__ rec, float ratio)
{
    const int per = m * m * m;
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t i = gid / per;
    if (i >= A) return;
    const int j = (int)(gid - i * per);
    const int4 p = pos[i];
    int x, y, z;
    if (!seed_axis(p.x, W, k, j % m, &x) || !seed_axis(p.y, H, k, (j / m) % m, &y) ||
        !seed_axis(p.z, D, k, j / (m * m), &z))
        return;
    out[((size_t)z * H + y) * W + x] = jfa_voxel<TIE>(W, H, D, in, lab, rec, k, ratio, x, y, z);
}

// Second pass of an iteration, over the first pass's output (step k1 from
// the seeds alone): a voxel there carries a label only if every coordinate
// lies in seed_axis's set of its seed (step k1), so a voxel of this pass
// (step k2) has a labelled tap only if every coordinate lies in the union,
// over that set, of seed_axis's sets for step k2 -- at most 3 (k2 + 1)
// values per axis around each seed, 9 inside the volume.  One block per seed
// lists the three axis sets in LDS and evaluates their product exactly as
// the full pass does; every other voxel keeps the -1 the output was filled
// with.  Voxels shared by neighbouring seeds are written by each with the
// same value.
constexpr int kSeed2Span = 2 * 2 * kSeedPassMaxStep + 1;   // coordinates within k1 + k2 of the seed


template <int TIE>
__global__ __launch_bounds__(256) void k_jfa_seeds2(int W, int H, int D, int k1, int k2,
                                                    const int4* __restrict__ pos, const int32_t* __restrict__ in,
                                                    int32_t* __restrict__ out, const uint32_t* __restrict__ lab,
                                                    const SeedRec* __restrict__ rec, float ratio)
{
    __shared__ int s_v[3][kSeed2Span];
    __shared__ int s_n[3];
    const int4 p = pos[blockIdx.x];
    const int r = k1 + k2;
    if (threadIdx.x < 3) {
        const int ax = threadIdx.x;
        const int pa = ax == 0 ? p.x : (ax == 1 ? p.y : p.z);
        const int n = ax == 0 ? W : (ax == 1 ? H : D);
        int cnt = 0;
        for (int c = max(0, pa - r); c <= min(n - 1, pa + r); ++c) {
            // c reaches, at step k2, some a that reaches the seed at step k1
            bool in_set = false;
            for (int a = max(0, pa - k1); a <= min(n - 1, pa + k1) && !in_set; ++a)
                in_set = axis_hit(a, pa, n, k1) && axis_hit(c, a, n, k2);
            if (in_set) s_v[ax][cnt++] = c;
        }
        s_n[ax] = cnt;
    }
    __syncthreads();
    const int nx = s_n[0], ny = s_n[1], nz = s_n[2];
    const int total = nx * ny * nz;
    for (int j = threadIdx.x; j < total; j += 256) {
        const int x = s_v[0][j % nx], y = s_v[1][(j / nx) % ny], z = s_v[2][j / (nx * ny)];
        out[((size_t)z * H + y) * W + x] = jfa_voxel<TIE>(W, H, D, in, lab, rec, k2, ratio, x, y, z);
    }
}

// The same second pass for the lowest-id rule, by scatter instead of gather:
// every label the first pass can have written sits at a voxel t whose
// coordinates lie in seed_axis's step-k1 sets of a seed, and a voxel v has
// t among its step-k2 taps exactly when each coordinate of v lies in
// seed_axis's step-k2 set of t's.  One block per seed lists, per axis, the
// (t, v) coordinate pairs, and their product gives every (labelled tap,
// voxel) pair of the seed's neighbourhood.  mode 0: each pair offers its
// candidate to v's 64-bit key (distance bits, id) by atomic minimum --
// the minimum over v's labelled taps is the full pass's lowest-id winner,
// and repeats (a tap reached twice, or from a neighbouring seed's block)
// cannot change a minimum; mode 1: the block of the winning seed w (the id
// in v's key) writes w into v and puts the key back to its initial value
// for the next use.  Every other block that meets v skips it, whether it
// reads the key before or after the reset, so each key is resolved and
// reset exactly once with plain loads and stores.  The winner's block does
// meet v: w's label sits at a tap t of v that lies in w's step-k1 set,
// since the first pass gave t the label of a seed among its own taps.  A
// voxel without a candidate keeps the seed stage's -1 (a seed voxel always
// has one: its own first-pass label).  A v that occurs in several pairs
// of its owner's block (small steps, clamped edges) is written with the
// same value by whichever pair sees the key first.
//
// Layout of the work: waves 0-2 list their axis's step-k1 set (the
// t-list, at most 2 k1 + 1 <= 33 candidates, one ballot) and then its
// (t-list index, v) pairs, 64 candidates per ballot, in ascending (t, v)
// order; in mode 0 wave 3 then reads the tap cells (the product of the
// three t-lists, 27 for k1 = 1) and their seeds' records into LDS when
// they are at most kSeed2Taps, so a pair reads its label and record from
// LDS and only the voxel's LAB from global memory.  Each thread takes
// three pairs at a time and issues their loads before it uses them.
constexpr int kSeed2Pairs = (kSeedPassMaxStep + 1) * (kSeedPassMaxStep + 1);
constexpr int kSeed2Axis = 2 * kSeedPassMaxStep + 1;      // t-list candidates per axis
constexpr int kSeed2Taps = 64;
constexpr int kSeed2Batch = 3;

__global__ __launch_bounds__(256) void k_jfa_seeds2_scatter(int W, int H, int D, int k1, int k2, int mode,
                                                            const int4* __restrict__ pos,
                                                            const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                            const uint32_t* __restrict__ lab,
                                                            const SeedRec* __restrict__ rec, float ratio,
                                                            unsigned long long* __restrict__ keys)
{
    __shared__ short s_t[3][kSeed2Pairs];     // pair: index into the axis's t-list
    __shared__ short s_v[3][kSeed2Pairs];     // pair: voxel coordinate
    __shared__ short s_a[3][kSeed2Axis];      // t-list: tap coordinates
    __shared__ int s_n[3], s_m[3];            // pairs and t-list length per axis
    __shared__ int s_tap[kSeed2Taps];         // staged tap labels
    __shared__ float4 s_r[3][kSeed2Taps];     // and their seeds' records
    const int4 p = pos[blockIdx.x];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int ax = wave < 3 ? wave : 0;
    const int pa = ax == 0 ? p.x : (ax == 1 ? p.y : p.z);
    const int na = ax == 0 ? W : (ax == 1 ? H : D);
    if (wave < 3) {
        const int a = pa - k1 + lane;
        const bool hit = lane <= 2 * k1 && a >= 0 && a < na && axis_hit(a, pa, na, k1);
        const unsigned long long b = __ballot(hit);
        if (hit) s_a[ax][__popcll(b & below)] = (short)a;
        if (lane == 0) s_m[ax] = __popcll(b);
    }
    __syncthreads();
    const int mx = s_m[0], my = s_m[1], mz = s_m[2];
    const bool staged = mode == 0 && mx * my * mz <= kSeed2Taps;
    if (wave < 3) {
        const int span = 2 * k2 + 1, cand = s_m[ax] * span;
        int cnt = 0;
        for (int q0 = 0; q0 < cand; q0 += 64) {
            const int q = q0 + lane;
            const int i = q / span;
            bool hit = q < cand;
            int c = 0;
            if (hit) {
                const int a = s_a[ax][i];
                c = a - k2 + (q - i * span);
                hit = c >= 0 && c < na && axis_hit(c, a, na, k2);
            }
            const unsigned long long b = __ballot(hit);
            if (hit) {
                const int w = cnt + __popcll(b & below);
                s_t[ax][w] = (short)i;
                s_v[ax][w] = (short)c;
            }
            cnt += __popcll(b);
        }
        if (lane == 0) s_n[ax] = cnt;
    } else if (staged && lane < mx * my * mz) {
        const int x = s_a[0][lane % mx], y = s_a[1][(lane / mx) % my], z = s_a[2][lane / (mx * my)];
        const int id = in[((size_t)z * H + y) * W + x];
        s_tap[lane] = id;
        if (id >= 0) {
            const SeedRec r = rec[id];
            s_r[0][lane] = r.a;
            s_r[1][lane] = r.b;
            s_r[2][lane] = r.c;
        }
    }
    __syncthreads();
    const int nx = s_n[0], ny = s_n[1], nz = s_n[2];
    const int total = nx * ny * nz;
    for (int j0 = 0; j0 < total; j0 += 256 * kSeed2Batch) {
        size_t v[kSeed2Batch];
        int ix[kSeed2Batch], iy[kSeed2Batch], iz[kSeed2Batch];
        bool ok[kSeed2Batch];
#pragma unroll
        for (int u = 0; u < kSeed2Batch; ++u) {
            const int j = j0 + u * 256 + tid;
            ok[u] = j < total;
            const int jj = ok[u] ? j : 0;
            ix[u] = jj % nx;
            iy[u] = (jj / nx) % ny;
            iz[u] = jj / (nx * ny);
            v[u] = ((size_t)s_v[2][iz[u]] * H + s_v[1][iy[u]]) * W + s_v[0][ix[u]];
        }
        if (mode == 0) {
            int id[kSeed2Batch], tap[kSeed2Batch];
            uint32_t pk[kSeed2Batch];
#pragma unroll
            for (int u = 0; u < kSeed2Batch; ++u) {
                const int tx = s_t[0][ix[u]], ty = s_t[1][iy[u]], tz = s_t[2][iz[u]];
                tap[u] = tx + mx * (ty + my * tz);
                id[u] = staged ? s_tap[tap[u]]
                               : in[((size_t)s_a[2][tz] * H + s_a[1][ty]) * W + s_a[0][tx]];
                if (!ok[u]) id[u] = -1;
                pk[u] = lab[id[u] >= 0 ? v[u] : 0];
            }
#pragma unroll
            for (int u = 0; u < kSeed2Batch; ++u) {
                if (id[u] < 0) continue;
                float4 ra, rb, rc;
                if (staged) {
                    ra = s_r[0][tap[u]];
                    rb = s_r[1][tap[u]];
                    rc = s_r[2][tap[u]];
                } else {
                    const SeedRec r = rec[id[u]];
                    ra = r.a;
                    rb = r.b;
                    rc = r.c;
                }
                const float d = cand_dist(ra, rb, rc, (float)s_v[0][ix[u]], (float)s_v[1][iy[u]],
                                          (float)s_v[2][iz[u]], (float)(pk[u] & 255u), (float)((pk[u] >> 8) & 255u),
                                          (float)((pk[u] >> 16) & 255u), ratio);
                const unsigned long long kk = dkey(d, id[u]);
                if (kk < ASV_KEY_INIT) atomicMin(keys + v[u], kk);
            }
        } else {
            unsigned long long kk[kSeed2Batch];
#pragma unroll
            for (int u = 0; u < kSeed2Batch; ++u) kk[u] = keys[ok[u] ? v[u] : 0];
#pragma unroll
            for (int u = 0; u < kSeed2Batch; ++u) {
                if (ok[u] && kk[u] < ASV_KEY_INIT && (unsigned)(kk[u] & 0xffffffffull) == (unsigned)blockIdx.x) {
                    out[v[u]] = (int32_t)blockIdx.x;
                    keys[v[u]] = ASV_KEY_INIT;
                }
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* __restrict__ p, int64_t n,
                                                  unsigned long long value)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = value;
}

// Tiled pass for the lowest-id rule and step K (template), 16 x 16 voxels of
// one slice per block, four 8 x 8 waves.
//  1. The labels of the three slabs (z, z-K, z+K) over the tile grown by K
//     on each side are staged in LDS with clamped coordinates, so every tap
//     is a fixed LDS offset and CLAMP_TO_EDGE is already applied.
//  2. The distinct seed ids of that region are hashed into at most 63
//     compact indices (LDS open addressing), their records copied to LDS.
//  3. Each lane ORs the compact index of its 27 taps into a 64-bit mask, then
//     evaluates each set bit once: exact per-voxel de-duplication, and the
//     wave runs as many rounds as its busiest lane has distinct candidates.
// A block whose region holds more than 63 ids takes the per-tap path over
// the staged labels.  Blocks are mapped so that each XCD (blocks b with the
// same b % 8) sweeps one contiguous range of tiles (z-major), keeping the
// slabs it re-reads in its own L2.
constexpr int kHashSlots = 128;
constexpr int kMaxCand = 63;
constexpr int kEmpty = -2;

__device__ __forceinline__ unsigned hash_slot(int id) { return ((unsigned)id * 2654435761u) >> 25; }

template <int K>
__global__ __launch_bounds__(256) void k_jfa_tile(int W, int H, int D, int tiles_x, int tiles_y, int ntiles,
                                                  const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                  const uint32_t* __restrict__ lab, const SeedRec* __restrict__ rec,
                                                  float ratio)
{
    constexpr int R = kTile + 2 * K;
    constexpr int RA = R * R;
    constexpr int NV = 3 * RA;
    constexpr int CH = (NV + 255) / 256;
    __shared__ int s_ids[NV];
    __shared__ uint8_t s_idx[NV];
    __shared__ int s_hkey[kHashSlots];
    __shared__ uint8_t s_hidx[kHashSlots];
    __shared__ float4 s_rec[kMaxCand * 3];
    __shared__ int s_rid[kMaxCand];
    __shared__ int s_wcount[2];

    const int tid = threadIdx.x;
    // XCD-aware tile order: hardware sends block b to XCD b % 8.
    const int b = blockIdx.x;
    const int q = ntiles >> 3, rm = ntiles & 7;
    const int xcd = b & 7;
    const int tile = xcd * q + min(xcd, rm) + (b >> 3);
    const int tx = tile % tiles_x;
    const int ty = (tile / tiles_x) % tiles_y;
    const int z = tile / (tiles_x * tiles_y);
    const int x0 = tx * kTile, y0 = ty * kTile;

    if (tid < kHashSlots) s_hkey[tid] = kEmpty;

    // 1. stage the region
    const int zs0 = z, zs1 = clampi(z - K, 0, D - 1), zs2 = clampi(z + K, 0, D - 1);
    for (int i = tid; i < NV; i += 256) {
        int sl = i / RA;
        int r = i - sl * RA;
        int ry = r / R, rx = r - ry * R;
        int gz = sl == 0 ? zs0 : (sl == 1 ? zs1 : zs2);
        int gy = clampi(y0 - K + ry, 0, H - 1);
        int gx = clampi(x0 - K + rx, 0, W - 1);
        s_ids[i] = in[((size_t)gz * H + gy) * W + gx];
    }
    __syncthreads();

    // 2. hash the distinct ids; each thread scans a contiguous chunk and
    //    inserts only where the id changes
    {
        int prev = -1;
        const int beg = tid * CH, end = min(beg + CH, NV);
        for (int i = beg; i < end; ++i) {
            int id = s_ids[i];
            if (id < 0 || id == prev) continue;
            prev = id;
            unsigned h = hash_slot(id);
            for (int p = 0; p < kHashSlots; ++p) {
                int old = atomicCAS(&s_hkey[h], kEmpty, id);
                if (old == kEmpty || old == id) break;
                h = (h + 1) & (kHashSlots - 1);
            }
        }
    }
    __syncthreads();
    // compact indices of the occupied slots (waves 0 and 1 cover 128 slots)
    int occ = 0, rank = 0;
    if (tid < kHashSlots) {
        occ = s_hkey[tid] != kEmpty;
        unsigned long long m = __ballot(occ);
        rank = __popcll(m & ((1ull << (tid & 63)) - 1ull));
        if ((tid & 63) == 0) s_wcount[tid >> 6] = __popcll(m);
    }
    __syncthreads();
    const int ncand = s_wcount[0] + s_wcount[1];
    const bool overflow = ncand > kMaxCand;
    if (!overflow) {
        if (tid < kHashSlots) {
            int ci = rank + ((tid >> 6) ? s_wcount[0] : 0);
            s_hidx[tid] = occ ? (uint8_t)ci : (uint8_t)kMaxCand;
            if (occ) {
                int id = s_hkey[tid];
                SeedRec r = rec[id];
                s_rec[ci * 3 + 0] = r.a;
                s_rec[ci * 3 + 1] = r.b;
                s_rec[ci * 3 + 2] = r.c;
                s_rid[ci] = id;
            }
        }
        __syncthreads();
        // compact index of every staged label (63 = no label)
        int prev = -3;
        uint8_t pidx = (uint8_t)kMaxCand;
        const int beg = tid * CH, end = min(beg + CH, NV);
        for (int i = beg; i < end; ++i) {
            int id = s_ids[i];
            if (id != prev) {
                prev = id;
                pidx = (uint8_t)kMaxCand;
                if (id >= 0) {
                    unsigned h = hash_slot(id);
                    for (int p = 0; p < kHashSlots; ++p) {
                        if (s_hkey[h] == id) { pidx = s_hidx[h]; break; }
                        h = (h + 1) & (kHashSlots - 1);
                    }
                }
            }
            s_idx[i] = pidx;
        }
    }
    __syncthreads();

    // 3. per-voxel evaluation
    const int wave = tid >> 6, lane = tid & 63;
    const int lx = (wave & 1) * 8 + (lane & 7);
    const int ly = (wave >> 1) * 8 + (lane >> 3);
    const int x = x0 + lx, y = y0 + ly;
    if (x >= W || y >= H) return;
    const size_t v = ((size_t)z * H + y) * W + x;
    const uint32_t pk = lab[v];
    const float l0 = (float)(pk & 255u), l1 = (float)((pk >> 8) & 255u), l2 = (float)((pk >> 16) & 255u);
    const float fx = (float)x, fy = (float)y, fz = (float)z;
    const int base = (ly + K) * R + (lx + K);
    unsigned long long best = ASV_KEY_INIT;

    if (!overflow) {
        unsigned long long mask = 0ull;
        lds_byte* pidx = (lds_byte*)(s_idx + base);
#pragma unroll
        for (int sl = 0; sl < 3; ++sl)
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx)
                    mask |= 1ull << pidx[sl * RA + dy * K * R + dx * K];
        mask &= ~(1ull << kMaxCand);
        while (mask) {
            int j = __ffsll((long long)mask) - 1;
            mask &= mask - 1ull;
            float d = cand_dist(s_rec[j * 3 + 0], s_rec[j * 3 + 1], s_rec[j * 3 + 2], fx, fy, fz, l0, l1, l2, ratio);
            unsigned long long kk = dkey(d, s_rid[j]);
            best = kk < best ? kk : best;
        }
    } else {
        int seen0 = -1, seen1 = -1;
#pragma unroll
        for (int sl = 0; sl < 3; ++sl)
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) {
                    int id = s_ids[sl * RA + base + dy * K * R + dx * K];
                    if (id < 0 || id == seen0 || id == seen1) continue;
                    seen1 = seen0;
                    seen0 = id;
                    SeedRec s = rec[id];
                    float d = cand_dist(s.a, s.b, s.c, fx, fy, fz, l0, l1, l2, ratio);
                    unsigned long long kk = dkey(d, id);
                    best = kk < best ? kk : best;
                }
    }
    out[v] = best == ASV_KEY_INIT ? -1 : (int32_t)(unsigned)(best & 0xffffffffull);
}

// ---------------------------------------------------------------------------
// centroid reduction (centroids.vs:19-53)
// ---------------------------------------------------------------------------

// Block-local slot table of the reduction: open addressing by label, at most
// kRedProbes probes, else the run goes straight to its global slot.
// Integer sums are exact, so no order of these additions changes a result.
constexpr int kRedSlots = 256;
constexpr int kRedProbes = 16;

// Inclusive prefix sum over the 64 lanes of a wave (DPP: row shifts, then
// the row broadcasts of lanes 15 and 31).
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false);   // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false);   // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);   // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);   // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);   // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);   // row_bcast:31
    return x;
}

// Rows as wave-wide runs: a wave reads 64 consecutive voxels of a row (one
// per lane, coalesced), finds the runs of equal labels with a ballot of run
// heads, and takes each run's LAB sums as differences of two wave prefix
// sums (L and a packed in one word: each field's sum over 64 voxels stays
// below 2^16).  Only a run's last lane adds it into the block's LDS table,
// as 32-bit sums (a brick of 64 x kRowBrick x zc voxels, zc <= 8, keeps
// every slot's sums below 2^32 for W, H, D < 2^18); the table goes to the
// global u64 slots at the end.  Position sums are closed forms: a run of n
// voxels from x = a to x = e on row (y, z) adds
//   sum(2x+1) = (e+1)^2 - a^2,  sum(2y+1) = n(2y+1),  sum(2z+1) = n(2z+1).
constexpr int kRowBrick = 16;

__global__ __launch_bounds__(256) void k_reduce_rows(int W, int H, int D, int A, int tiles_x, int tiles_y, int zc,
                                                     const int32_t* __restrict__ labels,
                                                     const uint32_t* __restrict__ lab,
                                                     unsigned long long* __restrict__ sums)
{
    __shared__ int s_key[kRedSlots];
    __shared__ uint32_t s_acc[kRedSlots][9];  // rows of 9 words: the slots of a wave's run tails fall in different banks
    const int tid = threadIdx.x;
    s_key[tid] = kEmpty;
#pragma unroll
    for (int i = 0; i < 8; ++i) s_acc[tid][i] = 0u;
    __syncthreads();

    const int b = blockIdx.x;
    const int tx = b % tiles_x, ty = (b / tiles_x) % tiles_y, tz = b / (tiles_x * tiles_y);
    const int lane = tid & 63, wave = tid >> 6;
    const int x0 = tx * 64, x = x0 + lane;
    const int nvalid = min(64, W - x0);
    const bool valid = lane < nvalid;
    const int z0 = tz * zc, z1 = min(D, z0 + zc);
    constexpr int RPW = kRowBrick / 4;           // rows per wave and frame
    for (int z = z0; z < z1; ++z) {
        int lv[RPW];
        uint32_t cv[RPW];
#pragma unroll
        for (int j = 0; j < RPW; ++j) {
            const int y = ty * kRowBrick + wave + 4 * j;
            const size_t v = ((size_t)z * H + y) * W + x;
            const bool ok = valid && y < H;
            lv[j] = ok ? labels[v] : 0;
            cv[j] = ok ? lab[v] : 0u;
        }
#pragma unroll
        for (int j = 0; j < RPW; ++j) {
            const int y = ty * kRowBrick + wave + 4 * j;
            if (y >= H) break;                   // uniform
            const int l = lv[j];
            const uint32_t c = cv[j];
            const int left = __shfl_up(l, 1);
            const bool head = valid && (lane == 0 || l != left);
            const unsigned long long heads = __ballot(head);
            const bool tail = valid && (lane == nvalid - 1 || ((heads >> (lane + 1)) & 1ull));
            const uint32_t pla = wave_scan_incl((c & 255u) | (((c >> 8) & 255u) << 16));
            const uint32_t pb = wave_scan_incl((c >> 16) & 255u);
            const unsigned long long upto = lane == 63 ? heads : heads & ((2ull << lane) - 1ull);
            const int h = upto ? 63 - __clzll((long long)upto) : 0;
            const int src = h > 0 ? h - 1 : 0;
            const uint32_t qla = __shfl(pla, src), qb = __shfl(pb, src);
            if (tail) {
                const uint32_t sla = h > 0 ? pla - qla : pla;
                const uint32_t sb = h > 0 ? pb - qb : pb;
                const uint32_t n = (uint32_t)(lane - h + 1);
                const uint32_t xa = (uint32_t)(x0 + h), xe1 = (uint32_t)(x + 1);
                uint32_t vals[7];
                vals[0] = n * (xa + xe1);          // (e+1)^2 - a^2 = n (a + e + 1)
                vals[1] = n * (uint32_t)(2 * y + 1);
                vals[2] = n * (uint32_t)(2 * z + 1);
                vals[3] = n;
                vals[4] = sla & 0xFFFFu;
                vals[5] = sla >> 16;
                vals[6] = sb;
                const int key = l < 0 ? A : l;
                unsigned hs = (hash_slot(key) << 1) & (kRedSlots - 1);
                bool done = false;
                for (int p = 0; p < kRedProbes; ++p) {
                    const int old = atomicCAS(&s_key[hs], kEmpty, key);
                    if (old == kEmpty || old == key) {
#pragma unroll
                        for (int i = 0; i < 7; ++i) atomicAdd(&s_acc[hs][i], vals[i]);
                        done = true;
                        break;
                    }
                    hs = (hs + 1) & (kRedSlots - 1);
                }
                if (!done) {
                    unsigned long long* o = sums + (size_t)key * kSumWords;
#pragma unroll
                    for (int i = 0; i < 7; ++i) atomicAdd(o + i, (unsigned long long)vals[i]);
                }
            }
        }
    }
    __syncthreads();
    const int key = s_key[tid];
    if (key != kEmpty) {
        unsigned long long* o = sums + (size_t)key * kSumWords;
#pragma unroll
        for (int i = 0; i < 7; ++i) atomicAdd(o + i, (unsigned long long)s_acc[tid][i]);
    }
}

// ---------------------------------------------------------------------------
// centroid finalize (aver_exception.vs:22-81)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float slot_mean(const unsigned long long* sums, int A, int j, int axis)
{
    if (j > A) return 0.0f;
    const unsigned long long* s = sums + (size_t)j * kSumWords;
    unsigned long long n = s[3];
    if (n == 0ull) return 0.0f;
    return (float)((double)s[axis] / (2.0 * (double)n));
}

__global__ void k_finalize(int W, int H, int D, int A, const unsigned long long* __restrict__ sums,
                           unsigned long long* __restrict__ zero_next, const uint32_t* __restrict__ lab,
                           int4* __restrict__ pos, float4* __restrict__ slab)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= A) {
#pragma unroll
        for (int w = 0; w < kSumWords; ++w) zero_next[(size_t)i * kSumWords + w] = 0ull;
    }
    if (i >= A) return;
    const unsigned long long* s = sums + (size_t)i * kSumWords;
    unsigned long long n = s[3];
    int p[3];
    float4 c;
    if (n == 0ull) {
        int jm = i == 0 ? W - 1 : i - 1;
        int jp = i + 1;
        int dims[3] = {W, H, D};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float cf = floorf(0.5f * (slot_mean(sums, A, jm, a) + slot_mean(sums, A, jp, a)));
            p[a] = clampi((int)cf, 0, dims[a] - 1);
        }
        uint32_t pk = lab[((size_t)p[2] * H + p[1]) * W + p[0]];
        c = make_float4((float)(pk & 255u), (float)((pk >> 8) & 255u), (float)((pk >> 16) & 255u), 0.0f);
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) p[a] = (int)(s[a] / (2ull * n));
        c = make_float4((float)((double)s[4] / (double)n), (float)((double)s[5] / (double)n),
                        (float)((double)s[6] / (double)n), 0.0f);
    }
    pos[i] = make_int4(p[0], p[1], p[2], 0);
    slab[i] = c;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

hipError_t launch_prepare(hipStream_t st, const Dims& d, const uint8_t* rgb, int prepared, const float* lin_lut,
                          uint32_t* lab)
{
    dim3 block(64, 4, 1);
    dim3 grid(cdiv(d.W, 64), cdiv(d.H, 4), (unsigned)d.D);
    hipLaunchKernelGGL(k_prepare, grid, block, 0, st, d.W, d.H, d.D, rgb, prepared, lin_lut, lab);
    return hipGetLastError();
}

hipError_t launch_metric_field(hipStream_t st, const Dims& d, const float2* fwd, const float2* bwd, float4* field)
{
    const int64_t V = (int64_t)d.W * d.H * d.D;
    unsigned blocks = V > 0 ? cdiv(V, 256) : 1;
    if (blocks > 65536u) blocks = 65536u;
    hipLaunchKernelGGL(k_metric_field, dim3(blocks), dim3(256), 0, st, d.W, d.H, d.D, fwd, bwd, field);
    return hipGetLastError();
}

hipError_t launch_density(hipStream_t st, const Dims& d, const uint8_t* rgb, int prepared, const int* mx,
                          const int* my, const int* mz, int wnum, int hnum, int nslot, unsigned long long* acc)
{
    const int64_t threads = (int64_t)((d.W + 15) / 16) * d.H * d.D;
    const int64_t blocks = (threads + 255) / 256;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    const float wr = 1.0f / (float)d.W;      // width_ratio (GPU-AS/Image.cpp)
    hipLaunchKernelGGL(k_density, dim3((unsigned)blocks), dim3(256), 0, st, d.W, d.H, d.D, rgb, prepared, mx, my, mz,
                       wnum, hnum, nslot, wr, acc);
    return hipGetLastError();
}

hipError_t launch_seed_init(hipStream_t st, const Dims& d, const int4* grid, const uint32_t* lab,
                            int4* pos, float4* slab)
{
    hipLaunchKernelGGL(k_seed_init, dim3(cdiv(d.A, 256)), dim3(256), 0, st, d.W, d.H, d.A, grid, lab, pos, slab);
    return hipGetLastError();
}

hipError_t launch_seed_metric(hipStream_t st, const Dims& d, const int4* pos, const float4* slab, const float2* fwd,
                              const float2* bwd, int gate, SeedRec* rec, float4* own, int* need, float* trace_iter,
                              int* unsafe)
{
    if (d.A <= 0) return hipSuccess;
    if (!unsafe) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_seed_metric, dim3(cdiv(d.A, 64)), dim3(64), 0, st, d.W, d.H, d.D, d.A, pos, slab, fwd, bwd,
                       rec, own, need, trace_iter, unsafe);
    hipLaunchKernelGGL(k_seed_metric_rings, dim3((unsigned)d.A), dim3(64), 0, st, d.W, d.H, d.D, pos, slab, fwd,
                       bwd, gate, rec, own, need, trace_iter, unsafe);
    return hipGetLastError();
}

hipError_t launch_fill(hipStream_t st, int32_t* p, int64_t n, int32_t value)
{
    int64_t n4 = n / 4;
    int64_t ntail = n - n4 * 4;
    unsigned blocks = n4 > 0 ? cdiv(n4, 256) : 1;
    if (blocks > 65536u) blocks = 65536u;
    hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, st, reinterpret_cast<int4*>(p), n4, p + n4 * 4,
                       ntail, value);
    return hipGetLastError();
}

hipError_t launch_zero_u64(hipStream_t st, uint64_t* p, int64_t n)
{
    unsigned blocks = n > 0 ? cdiv(n, 256) : 1;
    if (blocks > 4096u) blocks = 4096u;
    hipLaunchKernelGGL(k_zero_u64, dim3(blocks), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_zero_i32(hipStream_t st, int* p, int64_t n)
{
    unsigned blocks = n > 0 ? cdiv(n, 256) : 1;
    if (blocks > 4096u) blocks = 4096u;
    hipLaunchKernelGGL(k_zero_i32, dim3(blocks), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_place(hipStream_t st, const Dims& d, const int4* pos, int32_t* labels)
{
    hipLaunchKernelGGL(k_place, dim3(cdiv(d.A, 256)), dim3(256), 0, st, d.W, d.H, d.A, pos, labels);
    return hipGetLastError();
}

// A launch whose clips hold fewer than this many voxels per seed (grid
// step below about 11) is dense: a column segment meets more seeds than 31
// candidates hold, and the 64-bit masks avoid the rebuilds and per-tap steps.
constexpr int64_t kDenseVoxelsPerSeed = 1500;

bool march_dense(const JfaClip* c, int n)
{
    int64_t v = 0, a = 0;
    for (int i = 0; i < n; ++i) {
        v += (int64_t)c[i].W * c[i].H * c[i].D;
        a += c[i].A;
    }
    return a > 0 && v < kDenseVoxelsPerSeed * a;
}

// Candidate capacity of the marching pass at step k: 32-bit masks (31
// candidates per table generation) or 64-bit masks (63).  Small steps meet
// few distinct seeds along a column unless the seeds are dense; the
// environment variable ASV_MARCH_MC (31 or 63) overrides the choice for
// every step.
int march_mc(int k, bool dense)
{
    static const int forced = [] {          // read once, thread-safe initialisation
        const char* e = getenv("ASV_MARCH_MC");
        const int v = e ? atoi(e) : 0;
        return (v == 31 || v == 63) ? v : 0;
    }();
    if (forced) return forced;
    // measured at 640x480x200 (31 vs 63): K = 1 313 vs 354 us, K = 2 332 vs
    // 371, K = 3 480 vs 513, K = 5 430 vs 425 (two runs each, within 0.3 %);
    // at 640x480x10 with 12288 seeds (250 voxels per seed): K = 1 43.8 vs
    // 30.5 us, K = 2 132.6 vs 51.5, K = 3 38.2 vs 37.9
    return k <= 4 && !dense ? 31 : 63;
}

// Slices of a chain evaluated per step of the marching pass at step k: two
// halve the barriers, table round trips and step bookkeeping per voxel and
// need a fourth slab mask per thread.  The environment variable ASV_MARCH_S
// (1 or 2) overrides the choice for every step.
int march_s(int k)
{
    static const int forced = [] {
        const char* e = getenv("ASV_MARCH_S");
        const int v = e ? atoi(e) : 0;
        return (v == 1 || v == 2) ? v : 0;
    }();
    if (forced) return forced;
    // measured at 640x480x200: K = 1 362 vs 335 us, K = 2 420 vs 383,
    // K = 3 622 vs 566, K = 5 573 vs 511 (one vs two slices per step)
    (void)k;
    return 2;
}

// The marching kernel for step k, candidate capacity mc, slices per step sl;
// ht: some clip of the launch has more seeds than the table has slots;
// counted: the counting form.
static MarchFn march_kernel(int k, int mc, int sl, bool ht = false, bool counted = false, int lt = 0)
{
    switch ((k - 1) >> 2) {
    case 0: return march_kernel_part<0>(k, mc, sl, ht, counted, lt);
    case 1: return march_kernel_part<1>(k, mc, sl, ht, counted, lt);
    case 2: return march_kernel_part<2>(k, mc, sl, ht, counted, lt);
    case 3: return march_kernel_part<3>(k, mc, sl, ht, counted, lt);
    default: return nullptr;
    }
}

// Dynamic LDS of the marching pass at step k: the byte-per-id table for
// ids < amax, in whole KiB, and at least what a wide step keeps its keys in
// (two words per thread and slice: 4 KiB for 256 threads, 8 KiB for 512).
static int march_dyn_bytes(int k, int amax)
{
    const int bytes = ((amax > 0 ? amax : 1) + 3) / 4 * 4;
    const int kib = (bytes + 1023) / 1024 * 1024;
    const int keys = 2 * 2 * march_tile_w(k) * kMarchTileH * 4;
    return kib > keys ? kib : keys;
}

// Resident blocks of the marching pass (CUs x blocks per CU, a multiple of
// 8), cached per step, capacity and table size.
int march_max_grid(int k, int dyn, bool ht, bool dense, int lt)
{
    static int cache[17][2][2][2][4][kMarchTableIds / 1024 + 1];
    if (k < 1 || k > 16 || dyn < 0 || dyn > kMarchTableIds || lt < 0 || lt > 3) return 0;
    const int mc = march_mc(k, dense), sl = march_s(k);
    if (lt != 0 && !march_kernel(k, mc, sl, ht, false, lt)) return 0;
    int& slot = cache[k][mc == 63][sl == 2][ht][lt][dyn / 1024];
    int g = __atomic_load_n(&slot, __ATOMIC_RELAXED);
    if (g == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)march_kernel(k, mc, sl, ht, false, lt), march_tile_w(k) * kMarchTileH,
                                                           (size_t)dyn);
        g = cus * (per_cu > 0 ? per_cu : 1);
        g = (g / 8) * 8;
        if (g < 8) g = 8;
        __atomic_store_n(&slot, g, __ATOMIC_RELAXED);
    }
    return g;
}

// Plans one marching launch over n clips: the segments per column (want,
// the same for every clip as far as its slices allow; seg_cap > 0 caps the
// slices per segment) and each clip's first item, the launch's item count
// and its dynamic LDS bytes.  The grid is the fewest blocks per XCD that
// still take each XCD's items in the fewest rounds of the resident blocks
// (grid_limit > 0 caps them).  per_block > 0 fixes want at about per_block
// items per resident block; per_block = 0 searches want in 1..32 for the
// least predicted time, rounds x (slices per item + 2), the 2 standing for
// the three-slab rebuild that starts every item.  Returns the grid size, 0
// when the marching pass does not apply.
static int64_t plan_items(int k, JfaClip* c, int n, int64_t want, int seg_cap, int64_t* slices)
{
    int64_t items = 0, sl = 0;
    for (int i = 0; i < n; ++i) {
        const int steps = (c[i].D + k - 1) / k;
        int nseg = (int)(want < steps ? want : steps);
        if (nseg < 1) nseg = 1;
        if (seg_cap > 0 && (steps + nseg - 1) / nseg > seg_cap) nseg = (steps + seg_cap - 1) / seg_cap;
        c[i].seg = (steps + nseg - 1) / nseg;
        c[i].nseg = (steps + c[i].seg - 1) / c[i].seg;
        c[i].item_begin = (int)(items < 0x7fffffff ? items : 0x7fffffff);
        items += (int64_t)c[i].tiles_x * c[i].tiles_y * k * c[i].nseg;
        sl += (int64_t)c[i].tiles_x * c[i].tiles_y * c[i].D;
    }
    if (slices) *slices = sl;
    return items;
}

int march_plan(int k, JfaClip* c, int n, int per_block, int grid_limit, int seg_cap, int* nitems, int* dyn, int lt)
{
    if (k < 1 || k > 16 || n < 1) return 0;
    int amax = 1;
    for (int i = 0; i < n; ++i) amax = c[i].A > amax ? c[i].A : amax;
    if (amax > kMarchMaxIds) return 0;
    const int lds = march_dyn_bytes(k, amax < kMarchTableIds ? amax : kMarchTableIds);
    int gmax = march_max_grid(k, lds, amax > kMarchTableIds, march_dense(c, n), lt);
    if (gmax <= 0) return 0;
    if (grid_limit > 0 && grid_limit < gmax) gmax = grid_limit < 8 ? 8 : (grid_limit / 8) * 8;
    const int64_t pmax = gmax / 8;                       // resident blocks per XCD
    int64_t columns = 0;
    for (int i = 0; i < n; ++i) columns += (int64_t)c[i].tiles_x * c[i].tiles_y * k;
    int64_t want = 1;
    if (per_block > 0) {
        const int64_t target = (int64_t)per_block * gmax;
        want = columns >= target ? 1 : (target + columns - 1) / columns;
    } else {
        double best = 1e300;
        for (int64_t w = 1; w <= 32; ++w) {
            int64_t slices = 0;
            const int64_t items = plan_items(k, c, n, w, seg_cap, &slices);
            const int64_t per = (items + 7) / 8;
            const int64_t rounds = (per + pmax - 1) / pmax;
            const double t = (double)rounds * ((double)slices / (double)items + 2.0);
            if (t < best * 0.999) {
                best = t;
                want = w;
            }
        }
    }
    const int64_t items = plan_items(k, c, n, want, seg_cap, nullptr);
    if (items >= (int64_t)0x7fffffff) return 0;
    const int64_t per = (items + 7) / 8;
    const int64_t rounds = (per + pmax - 1) / pmax;
    int64_t p = (per + rounds - 1) / rounds;
    if (p < 1) p = 1;
    if (nitems) *nitems = (int)items;
    if (dyn) *dyn = lds;
    return (int)(8 * p);
}

hipError_t launch_march(hipStream_t st, int k, const JfaClip& one, const JfaClip* dev, int n, int grid, int nitems,
                        int dyn, bool ht, bool counted, bool dense, int lt)
{
    MarchFn fn = march_kernel(k, march_mc(k, dense), march_s(k), ht, counted, lt);
    if (!fn || grid < 8 || (grid & 7) || dyn < 4) return hipErrorInvalidValue;
    JfaClip a0 = one;
    const JfaClip* a1 = dev;
    int a2 = n, a3 = nitems, a4 = dyn / 4;
    void* args[] = {(void*)&a0, (void*)&a1, (void*)&a2, (void*)&a3, (void*)&a4};
    return hipLaunchKernel((const void*)fn, dim3((unsigned)grid), dim3((unsigned)(march_tile_w(k) * kMarchTileH)), args,
                           (size_t)dyn, st);
}

hipError_t launch_march_batch(hipStream_t st, int k, const JfaClip* dev, int n, int grid, int nitems, int dyn,
                              bool ht, bool counted, bool dense, int lt)
{
    JfaClip none{};
    return launch_march(st, k, none, dev, n, grid, nitems, dyn, ht, counted, dense, lt);
}

template <int K>
static hipError_t launch_tile(hipStream_t st, const Dims& d, const int32_t* in, int32_t* out, const uint32_t* lab,
                              const SeedRec* rec, float ratio, const int* unsafe, unsigned long long* stats, int lt)
{
    if (d.A <= kMarchMaxIds) {
        const int mx = (d.W + march_tile_w(K) - 1) / march_tile_w(K), my = (d.H + kMarchTileH - 1) / kMarchTileH;
        JfaClip c{in, out, lab, rec, d.W, d.H, d.D, d.A, mx, my, 1, 1, 0, ratio, unsafe, stats};
        int nitems = 0, dyn = 0;
        const int g = march_plan(K, &c, 1, 0, 0, 0, &nitems, &dyn, lt);
        if (g > 0)
            return launch_march(st, K, c, nullptr, 1, g, nitems, dyn, d.A > kMarchTableIds, stats != nullptr,
                                march_dense(&c, 1), lt);
    }
    if (lt != 0) return hipErrorInvalidValue;   // 16-bit labels only between marching passes
    const int tx = (d.W + kTile - 1) / kTile, ty = (d.H + kTile - 1) / kTile;
    int nt = tx * ty * d.D;
    hipLaunchKernelGGL(k_jfa_tile<K>, dim3((unsigned)nt), dim3(256), 0, st, d.W, d.H, d.D, tx, ty, nt, in, out, lab,
                       rec, ratio);
    return hipGetLastError();
}

hipError_t launch_jfa_seeds(hipStream_t st, const Dims& d, const int32_t* in, int32_t* out, const uint32_t* lab,
                            const SeedRec* rec, const int4* pos, int k, float ratio, int tie_rule, bool fill)
{
    // without the fill, `out` keeps stale labels outside the seeds' sets:
    // only for a next pass that reads nothing else (the scatter form)
    hipError_t e = fill ? launch_fill(st, out, (int64_t)d.W * d.H * d.D, -1) : hipSuccess;
    if (e != hipSuccess || d.A < 1) return e;
    const int m = k + 1 > 3 ? k + 1 : 3;
    const int64_t threads = (int64_t)d.A * m * m * m;
    const int64_t blocks = (threads + 255) / 256;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    if (tie_rule == 0)
        hipLaunchKernelGGL(k_jfa_seeds<0>, dim3((unsigned)blocks), dim3(256), 0, st, d.W, d.H, d.D, d.A, k, m, pos, in,
                           out, lab, rec, ratio);
    else
        hipLaunchKernelGGL(k_jfa_seeds<1>, dim3((unsigned)blocks), dim3(256), 0, st, d.W, d.H, d.D, d.A, k, m, pos, in,
                           out, lab, rec, ratio);
    return hipGetLastError();
}

hipError_t launch_jfa_seeds2(hipStream_t st, const Dims& d, const int32_t* in, int32_t* out, const uint32_t* lab,
                             const SeedRec* rec, const int4* pos, int k1, int k2, float ratio, int tie_rule,
                             unsigned long long* keys)
{
    if (k1 < 1 || k2 < 1 || k1 > kSeedPassMaxStep || k2 > kSeedPassMaxStep) return hipErrorInvalidValue;
    if (d.A < 1) return hipSuccess;
    // `out` is the seed stage's volume: -1 except at the seed voxels, and each
    // seed voxel lies in its own block's product set, so no fill is needed
    if (tie_rule == 0 && keys) {
        for (int mode = 0; mode < 2; ++mode)
            hipLaunchKernelGGL(k_jfa_seeds2_scatter, dim3((unsigned)d.A), dim3(256), 0, st, d.W, d.H, d.D, k1, k2, mode,
                               pos, in, out, lab, rec, ratio, keys);
    } else if (tie_rule == 0)
        hipLaunchKernelGGL(k_jfa_seeds2<0>, dim3((unsigned)d.A), dim3(256), 0, st, d.W, d.H, d.D, k1, k2, pos, in, out,
                           lab, rec, ratio);
    else
        hipLaunchKernelGGL(k_jfa_seeds2<1>, dim3((unsigned)d.A), dim3(256), 0, st, d.W, d.H, d.D, k1, k2, pos, in, out,
                           lab, rec, ratio);
    return hipGetLastError();
}

hipError_t launch_fill_keys(hipStream_t st, unsigned long long* keys, int64_t n)
{
    const int64_t b = (n + 255) / 256;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536)), dim3(256), 0, st, keys, n,
                       ASV_KEY_INIT);
    return hipGetLastError();
}

bool march_label16(const Dims& d, int k, int tie_rule, bool counted)
{
    // ASV_LABEL16=0 keeps 32-bit labels between passes (tests and measurement)
    static const bool off = [] {
        const char* e = getenv("ASV_LABEL16");
        return e && e[0] == '0';
    }();
    if (off) return false;
    const int64_t tiles = (int64_t)((d.W + kTile - 1) / kTile) * ((d.H + kTile - 1) / kTile);
    // launch_tile's conditions for the marching kernel; the item bound is
    // the largest a one-clip plan reaches (32 segments per column)
    return tie_rule == 0 && k >= 1 && k <= 16 && tiles * d.D < (1ll << 31) && d.A <= kMarchTableIds &&
           !counted && march_s(k) == 2 && tiles * k * 32 < 0x7fffffffll;
}

hipError_t launch_jfa(hipStream_t st, const Dims& d, const int32_t* in, int32_t* out,
                      const uint32_t* lab, const SeedRec* rec, int k, float ratio, int tie_rule, const int* unsafe,
                      unsigned long long* stats, int lt)
{
    if (lt != 0 && !march_label16(d, k, tie_rule, stats != nullptr)) return hipErrorInvalidValue;
    int64_t ntiles = (int64_t)((d.W + kTile - 1) / kTile) * ((d.H + kTile - 1) / kTile) * d.D;
    if (tie_rule == 0 && ntiles < (1ll << 31)) {
        switch (k) {
        case 1: return launch_tile<1>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 2: return launch_tile<2>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 3: return launch_tile<3>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 4: return launch_tile<4>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 5: return launch_tile<5>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 6: return launch_tile<6>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 7: return launch_tile<7>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 8: return launch_tile<8>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 9: return launch_tile<9>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 10: return launch_tile<10>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 11: return launch_tile<11>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 12: return launch_tile<12>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 13: return launch_tile<13>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 14: return launch_tile<14>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 15: return launch_tile<15>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        case 16: return launch_tile<16>(st, d, in, out, lab, rec, ratio, unsafe, stats, lt);
        default: break;
        }
    }
    dim3 block(64, 4, 1);
    dim3 grid(cdiv(d.W, 64), cdiv(d.H, 4), (unsigned)d.D);
    if (tie_rule == 0)
        hipLaunchKernelGGL(k_jfa<0>, grid, block, 0, st, d.W, d.H, d.D, in, out, lab, rec, k, ratio);
    else
        hipLaunchKernelGGL(k_jfa<1>, grid, block, 0, st, d.W, d.H, d.D, in, out, lab, rec, k, ratio);
    return hipGetLastError();
}

hipError_t launch_reduce(hipStream_t st, const Dims& d, const int32_t* labels, const uint32_t* lab,
                         unsigned long long* sums)
{
    // 32-bit brick sums: a slot of a 64 x 16 x 8 brick sums at most 8192
    // voxels' 2c+1 < 2^19 each
    if (d.W >= (1 << 18) || d.H >= (1 << 18) || d.D >= (1 << 18)) return hipErrorInvalidValue;
    // bricks of 64 columns x kRowBrick rows x zc frames: as deep as the
    // 32-bit brick sums allow (8), split evenly over the clip's frames.
    // Shallower bricks on short clips (more blocks) were slower: each block
    // flushes its table to the global slots, and a label then spans more
    // blocks (640x480x10 at 250 voxels per seed: 56 us with one-frame
    // bricks, 33 us with 8 + 2)
    const int tiles_x = (int)cdiv(d.W, 64);
    const int tiles_y = (int)cdiv(d.H, kRowBrick);
    const int64_t per_frame = (int64_t)tiles_x * tiles_y;
    const int zchunks = (d.D + 7) / 8;
    const int zc = (d.D + zchunks - 1) / zchunks;
    const int64_t nb = per_frame * cdiv(d.D, zc);
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_reduce_rows, dim3((unsigned)nb), dim3(256), 0, st, d.W, d.H, d.D, d.A, tiles_x, tiles_y, zc,
                       labels, lab, sums);
    return hipGetLastError();
}

hipError_t launch_finalize(hipStream_t st, const Dims& d, const unsigned long long* sums,
                           unsigned long long* zero_next, const uint32_t* lab, int4* pos, float4* slab)
{
    hipLaunchKernelGGL(k_finalize, dim3(cdiv(d.A + 1, 64)), dim3(64), 0, st, d.W, d.H, d.D, d.A, sums,
                       zero_next, lab, pos, slab);
    return hipGetLastError();
}

// Label remap through a table (the streaming blocks' block2all map):
// out[i] = table[in[i]], and 0 for a label outside the table (as an absent
// std::map key reads in the reference, streamGPU-AS/Render.cpp:559-563).
__global__ __launch_bounds__(256) void k_map_labels(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                    int64_t n, const int32_t* __restrict__ table, int n_table)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int l = in[i];
        out[i] = (l >= 0 && l < n_table) ? table[l] : 0;
    }
}

hipError_t launch_map_labels(hipStream_t st, const int32_t* in, int32_t* out, int64_t n, const int32_t* table,
                             int n_table)
{
    unsigned blocks = n > 0 ? cdiv(n, 256) : 1;
    if (blocks > 16384u) blocks = 16384u;
    hipLaunchKernelGGL(k_map_labels, dim3(blocks), dim3(256), 0, st, in, out, n, table, n_table);
    return hipGetLastError();
}

// Order-free checksum of a label run: sum over voxels of
// splitmix64((global index << 32) | (uint32)label), modulo 2^64, added into
// *out.  Integer sums, so the value is exact and launch-shape independent.
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// One 64-bit atomic per block: every block's sum goes to the same word, so
// the grid is kept to a few blocks per CU (with one atomic per wave over
// 4096 blocks the adds serialised on that word: 155 us for a 1280x720x5
// window's frames, against a few us of arithmetic).
__global__ __launch_bounds__(256) void k_checksum(const int32_t* __restrict__ labels, int64_t n, int64_t index0,
                                                  unsigned long long* __restrict__ out)
{
    __shared__ unsigned long long s_w[4];
    unsigned long long acc = 0ull;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += splitmix64(((unsigned long long)(index0 + i) << 32) | (unsigned)labels[i]);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, s_w[0] + s_w[1] + s_w[2] + s_w[3]);
}

hipError_t launch_checksum(hipStream_t st, const int32_t* labels, int64_t n, int64_t index0, unsigned long long* out)
{
    unsigned blocks = n > 0 ? cdiv(n, 256 * 16) : 1;
    if (blocks > 512u) blocks = 512u;
    hipLaunchKernelGGL(k_checksum, dim3(blocks), dim3(256), 0, st, labels, n, index0, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// post-processing of the final labels on the device
// (check_dead_seeds / label_replace.gs / sequence_label, GPU-AS/Image.cpp:1579-1778)
// ---------------------------------------------------------------------------
//
// Both raster-order rules of sequence_label carry "the last voxel so far
// that satisfies a predicate", i.e. a prefix maximum of
// j(i) = max{ j <= i : pred(j) } over the linear volume:
//   kPostFill  (no dead-seed rewrite): pred(j) = in[j] != -1; out[i] is
//     m(in[j(i)]) with m the label_replace composition table (identity when
//     no seed died); a voxel with no such j reads 0, except voxel 0, which
//     keeps its -1 (:1750-1776);
//   kPostRemap (more than three dead seeds): pred(j) = j == 0 or in[j] is
//     neither in[0] nor -1; out[i] = seq(in[j(i)]) with seq the alive rank
//     (0 for any other id, as std::map::operator[] gives) (:1727-1748).
// Three kernels: per-chunk maximum, one-block exclusive scan over chunks,
// and the apply pass that finishes the scan inside each chunk.
constexpr int kPostPer = 16;                       // voxels per thread
constexpr int kPostChunk = 256 * kPostPer;         // voxels per block

__device__ __forceinline__ bool post_pred(int mode, int64_t i, int l, int l0)
{
    return mode == 0 ? l != -1 : (i == 0 || (l != l0 && l != -1));
}

__device__ __forceinline__ int post_value(int mode, int l, const int32_t* __restrict__ tab, int n_table)
{
    if (l >= 0 && l < n_table) return tab[l];
    return mode == 0 ? l : 0;
}

__global__ __launch_bounds__(256) void k_post_chunk_max(const int32_t* __restrict__ in, int64_t V, int mode,
                                                        long long* __restrict__ chunk_max)
{
    __shared__ long long s_w[4];
    const int l0 = in[0];
    const int64_t base = (int64_t)blockIdx.x * kPostChunk + (int64_t)threadIdx.x * kPostPer;
    long long m = -1;
    for (int e = 0; e < kPostPer; ++e) {
        const int64_t i = base + e;
        if (i < V && post_pred(mode, i, in[i], l0)) m = i;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const long long o = __shfl_xor(m, off);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long r = s_w[0];
        for (int w = 1; w < 4; ++w) r = s_w[w] > r ? s_w[w] : r;
        chunk_max[blockIdx.x] = r;
    }
}

// exclusive prefix maximum over n chunk maxima, in place, one block
__global__ __launch_bounds__(1024) void k_post_scan(long long* __restrict__ v, int64_t n)
{
    __shared__ long long s_t[1024];
    const int t = threadIdx.x;
    const int64_t per = (n + 1023) / 1024;
    const int64_t b = (int64_t)t * per, e = b + per < n ? b + per : n;
    long long m = -1;
    for (int64_t i = b; i < e; ++i) m = v[i] > m ? v[i] : m;
    s_t[t] = m;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const long long o = t >= off ? s_t[t - off] : -1;
        __syncthreads();
        s_t[t] = o > s_t[t] ? o : s_t[t];
        __syncthreads();
    }
    long long run = t > 0 ? s_t[t - 1] : -1;
    for (int64_t i = b; i < e; ++i) {
        const long long x = v[i];
        v[i] = run;
        run = x > run ? x : run;
    }
}

__global__ __launch_bounds__(256) void k_post_apply(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                    int64_t V, int mode, const int32_t* __restrict__ tab,
                                                    int n_table, const long long* __restrict__ chunk_prefix)
{
    __shared__ long long s_w[4];
    const int l0 = in[0];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t base = (int64_t)blockIdx.x * kPostChunk + (int64_t)t * kPostPer;
    int l[kPostPer];
    long long m = -1;
#pragma unroll
    for (int e = 0; e < kPostPer; ++e) {
        const int64_t i = base + e;
        l[e] = i < V ? in[i] : -1;
        if (i < V && post_pred(mode, i, l[e], l0)) m = i;
    }
    // exclusive prefix maximum of the per-thread maxima over the block
    long long inc = m;
    for (int off = 1; off < 64; off <<= 1) {
        const long long o = __shfl_up(inc, off);
        if (lane >= off) inc = o > inc ? o : inc;
    }
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    long long carry = chunk_prefix[blockIdx.x];
    for (int w = 0; w < wave; ++w) carry = s_w[w] > carry ? s_w[w] : carry;
    long long excl = __shfl_up(inc, 1);
    if (lane == 0) excl = -1;
    long long run = excl > carry ? excl : carry;
#pragma unroll
    for (int e = 0; e < kPostPer; ++e) {
        const int64_t i = base + e;
        if (i >= V) break;
        int r;
        if (post_pred(mode, i, l[e], l0)) {
            run = i;
            r = post_value(mode, l[e], tab, n_table);
        } else if (run >= 0) {
            r = post_value(mode, in[run], tab, n_table);
        } else {
            r = i == 0 ? l[e] : 0;
        }
        out[i] = r;
    }
}

int64_t post_scratch_words(int64_t V) { return V > 0 ? (V + kPostChunk - 1) / kPostChunk : 1; }

hipError_t launch_postprocess(hipStream_t st, const int32_t* in, int32_t* out, int64_t V, int mode,
                              const int32_t* table, int n_table, long long* scratch)
{
    const int64_t nch = post_scratch_words(V);
    if (nch > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_post_chunk_max, dim3((unsigned)nch), dim3(256), 0, st, in, V, mode, scratch);
    hipLaunchKernelGGL(k_post_scan, dim3(1), dim3(1024), 0, st, scratch, nch);
    hipLaunchKernelGGL(k_post_apply, dim3((unsigned)nch), dim3(256), 0, st, in, out, V, mode, table, n_table,
                       (const long long*)scratch);
    return hipGetLastError();
}

}  // namespace asv
