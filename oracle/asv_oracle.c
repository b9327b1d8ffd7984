/*
 * asv_oracle.c -- CPU FP32 restatement of the anisotropic-supervoxel hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke()
 * entry and bench.py's cpu_baseline leg may load it; the product library never
 * links, loads or calls it, and nothing here is a fallback for the GPU path.
 *
 * It restates, in plain C, what the reference GLSL passes and host code do
 * (citations are relative to the reference repository root), under one fixed
 * arithmetic specification that the HIP kernels implement independently:
 *   - no implicit FMA contraction (built with -ffp-contract=off); the JFA
 *     distance uses explicit fmaf in a fixed order; IEEE float/double
 *     division and square root;
 *   - sRGB linearisation through a 256-entry table (double pow, rounded to
 *     float), cube root by a fixed Halley iteration in double;
 *   - per-voxel least squares from the exact integer normal equations,
 *     solved in double (LDL^T, or a cyclic-Jacobi pseudo-inverse when a pivot
 *     collapses), rounded to float;
 *   - PSD test by principal minors in double;
 *   - centroid sums as exact integers (sum of 2x+1, of L, a, b, and counts);
 *   - JFA ties broken by lowest seed id (or by the reference's scan order).
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   orc_lab_of        closed forms white->(255,128,128), black->(0,128,128),
 *                     primaries against an independent numpy CIE formula.
 *   orc_blur_flip     brute-force 3x3 [1,2,1]^2/16 with reflect-101 borders
 *                     (OpenCV 4.0 CV_8U behaviour is assumed: parity with
 *                     OpenCV itself is unpinned, OpenCV is absent).
 *   orc_metric_at     zero flow -> identity; uniform integer flow (u,v) ->
 *                     (1,0,-u,1,-v,1+u^2+v^2); underdetermined corners ->
 *                     numpy.linalg.lstsq min-norm solution.
 *   orc_grid_seeds    placed-seed counts 468 / 1955 / 9720 for the three
 *                     plain configurations and perfect-cube truncation.
 *   orc_schedule      step lists 1,5,3,2,1,2,1 / 1,7,4,2,1,2,1 / 1,9,5,3,2,1,2,1.
 *   seed_metric       constructed fields (tests/test_oracle_pins.py): own PSD
 *                     metric; one PSD 1-ring neighbour exactly; float means
 *                     summed in the ring order of Image.cpp:701-702 and
 *                     :766-767 (values whose sum depends on the order);
 *                     empty 1-ring -> 2-ring; out-of-volume neighbours; both
 *                     rings empty -> nearest PSD; the angle gate on and off,
 *                     zero flow and the last frame.
 *   angle_gate        the reference expression (float numerator, double
 *                     norms, float acos correctly rounded, PI = 3.141592653,
 *                     theta < 90) over every float cosine near the threshold
 *                     and random flows.
 *   orc_jfa_pass      brute force over the 27 clamped taps in numpy on tiny
 *                     volumes, both tie rules (the scan-order rule with many
 *                     exact ties); one-seed flood; invariants; NaN and
 *                     negative quadratic forms.
 *   orc_run           exact-sum identities (sum of counts = V), constructed
 *                     dead seed, two seeds on one voxel (the higher id wins,
 *                     the loser takes the neighbour-slot rule).
 * Eigen's SVD/eigenvalue rounding and the GL driver's pow/unorm rounding are
 * not reproducible here; those differences are inside the reference's own
 * float noise and are not pinned.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the OpenMP loops (the CPU baseline times all cores and one). */
ORC_API void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* a2: RGB -> LAB  (GPU-AS/shaders/RGB2LAB.fs:11-79)                          */
/* ------------------------------------------------------------------------ */

static float g_lin[256];
static int g_lin_ready = 0;

static void lin_table(void)
{
    if (g_lin_ready) return;
    for (int c = 0; c < 256; ++c) {
        float u = (float)c / 255.0f;                      /* GL unorm8 -> float */
        if (u <= 0.04045f) g_lin[c] = u / 12.92f;          /* RGB2LAB.fs:27-28 */
        else {
            float t = (u + 0.055f) / 1.055f;               /* RGB2LAB.fs:30 */
            g_lin[c] = (float)pow((double)t, 2.4);
        }
    }
    g_lin_ready = 1;
}

ORC_API void orc_lin_table(float* out)
{
    lin_table();
    memcpy(out, g_lin, sizeof(g_lin));
}

/* Cube root in double by Halley's iteration from a linear start; valid on
 * (0, 1.1], which covers every XYZ/white ratio above the CIE epsilon. */
static double cbrt_halley(double x)
{
    double y = 0.45 + 0.55 * x;
    for (int i = 0; i < 5; ++i) {
        double y3 = y * y * y;
        y = y * ((y3 + 2.0 * x) / (2.0 * y3 + x));
    }
    return y;
}

static float lab_f(float r)
{
    /* RGB2LAB.fs:50-70: r > eps ? r^(1/3) : (kappa r + 16)/116 */
    if (r > 0.008856f) return (float)cbrt_halley((double)r);
    return (903.3f * r + 16.0f) / 116.0f;
}

static uint8_t round_u8(float v)
{
    float r = floorf(v + 0.5f);                            /* RGB2LAB.fs:77 */
    if (!(r >= 0.0f)) r = 0.0f;
    if (r > 255.0f) r = 255.0f;
    return (uint8_t)r;
}

ORC_API void orc_lab_of(int r, int g, int b, uint8_t out[3])
{
    lin_table();
    float tr = g_lin[r], tg = g_lin[g], tb = g_lin[b];
    /* RGB2LAB.fs:19-21, :42-44 (dot products, left to right) */
    float X = (tr * 0.412453f + tg * 0.357580f) + tb * 0.180423f;
    float Y = (tr * 0.212671f + tg * 0.715160f) + tb * 0.072169f;
    float Z = (tr * 0.019334f + tg * 0.119193f) + tb * 0.950227f;
    /* RGB2LAB.fs:54-55 (D65 white) */
    float fx = lab_f(X / 0.950456f);
    float fy = lab_f(Y / 1.0f);
    float fz = lab_f(Z / 1.088754f);
    /* RGB2LAB.fs:72-77 */
    float L = 116.0f * fy - 16.0f;
    float A = 500.0f * (fx - fy);
    float B = 200.0f * (fy - fz);
    out[0] = round_u8(L * 2.55f);
    out[1] = round_u8(A + 128.0f);
    out[2] = round_u8(B + 128.0f);
}

/* ------------------------------------------------------------------------ */
/* a1: frame preprocessing (GPU-AS/Render.cpp:354-359)                       */
/* GaussianBlur 3x3 (sigma from ksize: kernel [1,2,1]/4 per axis, reflect-101 */
/* border, CV_8U rounding half up), then vertical flip to GL rows.            */
/* ------------------------------------------------------------------------ */

static int reflect101(int i, int n)
{
    if (n == 1) return 0;
    if (i < 0) return -i;
    if (i >= n) return 2 * n - 2 - i;
    return i;
}

/* raw: D frames of H rows (row 0 = top) x W x 3 u8.
 * out: D frames of H rows in GL order (row 0 = bottom) x W x 3, blurred. */
ORC_API void orc_blur_flip(const uint8_t* raw, int W, int H, int D, uint8_t* out)
{
    static const int w3[3] = {1, 2, 1};
    for (int t = 0; t < D; ++t) {
        const uint8_t* fr = raw + (size_t)t * W * H * 3;
        uint8_t* fo = out + (size_t)t * W * H * 3;
        for (int yi = 0; yi < H; ++yi) {
            int ygl = H - 1 - yi;
            for (int x = 0; x < W; ++x) {
                for (int c = 0; c < 3; ++c) {
                    int acc = 0;
                    for (int j = -1; j <= 1; ++j) {
                        int yy = reflect101(yi + j, H);
                        for (int i = -1; i <= 1; ++i) {
                            int xx = reflect101(x + i, W);
                            acc += w3[j + 1] * w3[i + 1] * fr[((size_t)yy * W + xx) * 3 + c];
                        }
                    }
                    fo[((size_t)ygl * W + x) * 3 + c] = (uint8_t)((acc + 8) >> 4);
                }
            }
        }
    }
}

/* prepared (GL rows, blurred) RGB -> packed LAB (L | a<<8 | b<<16) */
ORC_API void orc_lab_volume(const uint8_t* rgb, int64_t n, uint32_t* lab)
{
    for (int64_t i = 0; i < n; ++i) {
        uint8_t o[3];
        orc_lab_of(rgb[i * 3], rgb[i * 3 + 1], rgb[i * 3 + 2], o);
        lab[i] = (uint32_t)o[0] | ((uint32_t)o[1] << 8) | ((uint32_t)o[2] << 16);
    }
}

/* ------------------------------------------------------------------------ */
/* a3: per-voxel anisotropic metric (GPU-AS/Image.cpp:874-1009)               */
/* ------------------------------------------------------------------------ */

/* Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix in double.
 * a is overwritten (eigenvalues on the diagonal), v receives eigenvectors
 * as columns.  Fixed pivot order p<q row-major, at most `sweeps` sweeps,
 * stopping early once every off-diagonal entry is exactly zero. */
static void jacobi_sym(double* a, double* v, int n, int sweeps)
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
                double c = 1.0 / sqrt(t * t + 1.0);
                double s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
                a[p * n + q] = 0.0;
                a[q * n + p] = 0.0;
                for (int k = 0; k < n; ++k) {
                    double vkp = v[k * n + p], vkq = v[k * n + q];
                    v[k * n + p] = c * vkp - s * vkq;
                    v[k * n + q] = s * vkp + c * vkq;
                }
            }
        }
        if (!any) break;
    }
}

/* Minimum-norm least squares for the 6 unknowns from the normal equations
 * N m = c (N = A^T A, c = A^T b, exact integers held in double).  Stands in
 * for Eigen's bdcSvd(ThinU|ThinV).solve at GPU-AS/Image.cpp:961. */
static void lsq6(const double N_in[36], const double c_in[6], double m[6])
{
    double maxd = 0.0;
    for (int i = 0; i < 6; ++i) if (N_in[i * 6 + i] > maxd) maxd = N_in[i * 6 + i];
    if (maxd == 0.0) { for (int i = 0; i < 6; ++i) m[i] = 0.0; return; }
    /* LDL^T without pivoting */
    double L[36], d[6];
    int ok = 1;
    for (int j = 0; j < 6 && ok; ++j) {
        double s = N_in[j * 6 + j];
        for (int k = 0; k < j; ++k) s = s - (L[j * 6 + k] * L[j * 6 + k]) * d[k];
        if (!(s > 1e-9 * N_in[j * 6 + j])) { ok = 0; break; }
        d[j] = s;
        L[j * 6 + j] = 1.0;
        for (int i = j + 1; i < 6; ++i) {
            double t = N_in[i * 6 + j];
            for (int k = 0; k < j; ++k) t = t - (L[i * 6 + k] * L[j * 6 + k]) * d[k];
            L[i * 6 + j] = t / s;
        }
    }
    if (ok) {
        double y[6];
        for (int i = 0; i < 6; ++i) {
            double t = c_in[i];
            for (int k = 0; k < i; ++k) t = t - L[i * 6 + k] * y[k];
            y[i] = t;
        }
        for (int i = 0; i < 6; ++i) y[i] = y[i] / d[i];
        for (int i = 5; i >= 0; --i) {
            double t = y[i];
            for (int k = i + 1; k < 6; ++k) t = t - L[k * 6 + i] * m[k];
            m[i] = t;
        }
        return;
    }
    /* rank-deficient: pseudo-inverse through the eigen-decomposition */
    double a[36], v[36];
    memcpy(a, N_in, sizeof(a));
    jacobi_sym(a, v, 6, 30);
    double lmax = 0.0;
    for (int i = 0; i < 6; ++i) if (a[i * 6 + i] > lmax) lmax = a[i * 6 + i];
    for (int i = 0; i < 6; ++i) m[i] = 0.0;
    for (int e = 0; e < 6; ++e) {
        double lam = a[e * 6 + e];
        if (!(lam > 1e-9 * lmax)) continue;
        double proj = 0.0;
        for (int k = 0; k < 6; ++k) proj = proj + v[k * 6 + e] * c_in[k];
        double coef = proj / lam;
        for (int k = 0; k < 6; ++k) m[k] = m[k] + v[k * 6 + e] * coef;
    }
}

/* PSD test: no eigenvalue below zero (GPU-AS/Image.cpp:983-1006), decided
 * by the signs of all principal minors evaluated in double. */
static int psd3(const float m[6])
{
    double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5];
    if (a < 0.0 || d < 0.0 || f < 0.0) return 0;
    if (a * d - b * b < 0.0) return 0;
    if (a * f - c * c < 0.0) return 0;
    if (d * f - e * e < 0.0) return 0;
    double det = a * (d * f - e * e) - b * (b * f - e * c) + c * (b * e - d * c);
    if (det < 0.0) return 0;
    return 1;
}

ORC_API int orc_psd3(const float* m) { return psd3(m); }

/* GL-oriented flow at (x, y_gl) of frame t from an image-oriented field
 * (row 0 = top): the reference computes flow on vertically flipped frames
 * (GPU-AS/Render.cpp:238-242), so rows flip and the v component negates. */
static void flow_gl(const float* F, int W, int H, int x, int y, int t, float* u, float* v)
{
    int yi = H - 1 - y;
    const float* p = F + (((size_t)t * H + yi) * W + x) * 2;
    *u = p[0];
    *v = -p[1];
}

static const int k_expected_b[9] = {3, 2, 3, 2, 1, 2, 3, 2, 3};   /* Image.cpp:884 */

/* matrix_slover (GPU-AS/Image.cpp:874-1009) at GL pixel (x, y) of frame t. */
ORC_API void orc_metric_at(const float* fwd, const float* bwd, int W, int H, int D,
                           int x, int y, int t, float out[6], int* psd)
{
    double N[36], c[6];
    memset(N, 0, sizeof(N));
    memset(c, 0, sizeof(c));
    int rows = 0;
    for (int m = -1; m <= 1; ++m) {            /* offset_y outer, Image.cpp:887 */
        for (int n = -1; n <= 1; ++n) {        /* offset_x inner, Image.cpp:889 */
            int nx = x + n, ny = y + m;
            if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
            double b = (double)k_expected_b[(m + 1) * 3 + (n + 1)];
            for (int dir = 0; dir < 2; ++dir) {
                if (dir == 0 && t == D - 1) continue;       /* Image.cpp:898 */
                if (dir == 1 && t == 0) continue;           /* Image.cpp:915 */
                float fu, fv;
                flow_gl(dir == 0 ? fwd : bwd, W, H, nx, ny, t, &fu, &fv);
                float dx = (float)n + floorf(fu + 0.5f);    /* Image.cpp:901-903 */
                float dy = (float)m + floorf(fv + 0.5f);
                double sgn = dir == 0 ? 2.0 : -2.0;         /* Image.cpp:908 vs :925 */
                double r[6];
                r[0] = (double)dx * (double)dx;
                r[1] = 2.0 * (double)dx * (double)dy;
                r[2] = sgn * (double)dx;
                r[3] = (double)dy * (double)dy;
                r[4] = sgn * (double)dy;
                r[5] = 1.0;
                for (int i = 0; i < 6; ++i) {
                    for (int j = 0; j < 6; ++j) N[i * 6 + j] += r[i] * r[j];
                    c[i] += r[i] * b;
                }
                ++rows;
            }
        }
    }
    double md[6];
    if (rows == 0) { for (int i = 0; i < 6; ++i) md[i] = 0.0; }
    else lsq6(N, c, md);
    for (int i = 0; i < 6; ++i) out[i] = (float)md[i];
    *psd = psd3(out);
}

ORC_API void orc_metric_field(const float* fwd, const float* bwd, int W, int H, int D,
                              float* m6, uint8_t* psd)
{
    #pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < D; ++t)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                size_t v = ((size_t)t * H + y) * W + x;
                int p;
                orc_metric_at(fwd, bwd, W, H, D, x, y, t, m6 + v * 6, &p);
                psd[v] = (uint8_t)p;
            }
}

/* Nearest PSD matrix (GPU-AS/Image.cpp:787-835): eigen-decompose, clamp the
 * negative eigenvalues to zero, rebuild Q D+ Q^T. */
ORC_API void orc_nearest_psd(const float in[6], float out[6])
{
    double a[9] = {in[0], in[1], in[2], in[1], in[3], in[4], in[2], in[4], in[5]};
    double v[9];
    jacobi_sym(a, v, 3, 30);
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

/* ------------------------------------------------------------------------ */
/* a4: seed metric selection with ring fallbacks                            */
/* (GPU-AS/Image.cpp:689-872 first iteration; :1351-1534 later iterations)  */
/* ------------------------------------------------------------------------ */

static const int k_ring1x[8] = {-1, 0, 1, -1, 1, -1, 0, 1};        /* Image.cpp:701 */
static const int k_ring1y[8] = {-1, -1, -1, 0, 0, 1, 1, 1};        /* Image.cpp:702 */
static const int k_ring2x[16] = {-2, -1, 0, 1, 2, -2, 2, -2, 2, -2, 2, -2, -1, 0, 1, 2};
static const int k_ring2y[16] = {-2, -2, -2, -2, -2, -1, -1, 0, 0, 1, 1, 2, 2, 2, 2, 2};

/* acos(cos) * 180/PI < 90 (Image.cpp:1407-1409) for a correctly rounded
 * float acos holds exactly when cos exceeds this float (bits 0x328885a3). */
#define ORC_GATE_COS_MIN 1.5893254712295857e-08f

static int angle_gate(float su, float sv, float nu, float nv)
{
    float num = su * nu + sv * nv;
    double den = sqrt((double)su * (double)su + (double)sv * (double)sv) *
                 sqrt((double)nu * (double)nu + (double)nv * (double)nv);
    float cs = (float)((double)num / den);
    return cs > ORC_GATE_COS_MIN && cs <= 1.0f;
}

typedef struct {
    int W, H, D;
    const float* m6;      /* V x 6 metric field */
    const uint8_t* psd;   /* V flags */
    const float* fwd;     /* image-oriented forward flow, V x 2 */
} orc_field;

static void seed_metric(const orc_field* F, int x, int y, int z, int gate, float out[6])
{
    size_t v = ((size_t)z * F->H + y) * F->W + x;
    if (F->psd[v]) { memcpy(out, F->m6 + v * 6, 6 * sizeof(float)); return; }
    float su = 0.0f, sv = 0.0f;
    if (gate && z != F->D - 1) flow_gl(F->fwd, F->W, F->H, x, y, z, &su, &sv);
    float sum[6] = {0, 0, 0, 0, 0, 0};
    int cnt = 0;
    for (int ring = 0; ring < 2 && cnt == 0; ++ring) {
        int nr = ring == 0 ? 8 : 16;
        for (int t = 0; t < nr; ++t) {
            int nx = x + (ring == 0 ? k_ring1x[t] : k_ring2x[t]);
            int ny = y + (ring == 0 ? k_ring1y[t] : k_ring2y[t]);
            if (nx < 0 || nx >= F->W || ny < 0 || ny >= F->H) continue;
            if (gate) {
                float nu = 0.0f, nv = 0.0f;
                if (z != F->D - 1) flow_gl(F->fwd, F->W, F->H, nx, ny, z, &nu, &nv);
                if (!angle_gate(su, sv, nu, nv)) continue;
            }
            size_t nvx = ((size_t)z * F->H + ny) * F->W + nx;
            if (!F->psd[nvx]) continue;
            for (int k = 0; k < 6; ++k) sum[k] = sum[k] + F->m6[nvx * 6 + k];
            ++cnt;
        }
    }
    if (cnt > 0) {
        for (int k = 0; k < 6; ++k) out[k] = sum[k] / (float)cnt;
        return;
    }
    orc_nearest_psd(F->m6 + v * 6, out);
}

/* seed_metric over a given field (m6: V x 6 metric, psd: V flags, fwd:
 * image-oriented forward flow for the gate), for constructed cases */
ORC_API void orc_seed_metric(const float* m6, const uint8_t* psd, const float* fwd, int W, int H, int D,
                             int x, int y, int z, int gate, float* out)
{
    orc_field F = {W, H, D, m6, psd, fwd};
    seed_metric(&F, x, y, z, gate, out);
}

ORC_API int orc_angle_gate(float su, float sv, float nu, float nv) { return angle_gate(su, sv, nu, nv); }

/* ------------------------------------------------------------------------ */
/* a5: grid seeding (GPU-AS/Image.cpp:20-92)                                 */
/* ------------------------------------------------------------------------ */

/* Writes A x (x,y,z) voxel coordinates into xyz (capacity K*3) and the
 * grid step s into *s_out; returns A, the number of seeds placed. */
ORC_API int orc_grid_seeds(int W, int H, int D, int K, int* xyz, int* s_out)
{
    int V = W * H * D;
    double voxel_size = (double)(V / K);                       /* Image.cpp:24 */
    int s = (int)floor(pow(voxel_size, 1.0 / 3));              /* Image.cpp:32 */
    int dNum = (int)pow((float)(K * D * D / W / H), 1.0 / 3.0);/* Image.cpp:37 */
    if (dNum == 0) dNum = 1;
    int frame_seeds = K / dNum;
    int wNum = (int)sqrtf((float)(frame_seeds * W / H));       /* Image.cpp:43 */
    if (wNum == 0) return -1;
    int hNum = frame_seeds / wNum;
    if (hNum == 0) return -1;
    int stepD = D / dNum, stepH = H / hNum, stepW = W / wNum;
    int d_rem = D - stepD * dNum, h_rem = H - stepH * hNum, w_rem = W - stepW * wNum;
    int td = d_rem < 1 ? 0 : 1;
    int th, tw;
    int count = 0;
    for (int k = 0; k < dNum; ++k) {
        th = 1;
        for (int i = 0; i < hNum; ++i) {
            tw = 1;
            for (int j = 0; j < wNum; ++j) {
                int sx = (int)(j * stepW + 0.5 * stepW + tw);
                int sy = (int)(i * stepH + 0.5 * stepH + th);
                int sz = (int)(k * stepD + 0.5 * stepD + td);
                if (sx > W - 1) sx = W - 1;
                if (sy > H - 1) sy = H - 1;
                if (sz > D - 1) sz = D - 1;
                if (tw < w_rem) ++tw;
                if (xyz) { xyz[count * 3] = sx; xyz[count * 3 + 1] = sy; xyz[count * 3 + 2] = sz; }
                ++count;
            }
            if (th < h_rem) ++th;
        }
        if (td < d_rem) ++td;
    }
    if (s_out) *s_out = s;
    return count;
}

/* ------------------------------------------------------------------------ */
/* a7: JFA schedule (GPU-AS/Image.cpp:1055-1181)                              */
/* ------------------------------------------------------------------------ */

/* mode 0: reference 1 + [floor(s/2) .. 1] + [2, 1], ratio x1 then x2 on the
 * last two passes.  mode 1: k = floor(W/2) halving to 1, then one extra
 * step-1 pass, ratio x1 throughout.  Returns the number of passes. */
ORC_API int orc_schedule(int mode, int s, int W, int* steps, float* mult, int cap)
{
    int n = 0;
    if (mode == 0) {
        if (n < cap) { steps[n] = 1; mult[n] = 1.0f; } ++n;
        float k = floorf(0.5f * (float)s);
        while (k > 0.0f) {
            if (n < cap) { steps[n] = (int)k; mult[n] = 1.0f; } ++n;
            if ((int)k == 1) k = 0.0f;
            k = floorf(0.5f * k + 0.5f);
        }
        k = 2.0f;
        while (k > 0.0f) {
            if (n < cap) { steps[n] = (int)k; mult[n] = 2.0f; } ++n;
            if ((int)k == 1) k = 0.0f;
            k = floorf(0.5f * k + 0.5f);
        }
    } else {
        float k = floorf(0.5f * (float)W);
        while (k > 0.0f) {
            if (n < cap) { steps[n] = (int)k; mult[n] = 1.0f; } ++n;
            if ((int)k == 1) k = 0.0f;
            k = floorf(0.5f * k + 0.5f);
        }
        if (n < cap) { steps[n] = 1; mult[n] = 1.0f; } ++n;
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* a7: one JFA pass (GPU-AS/shaders/flood.gs:26-182)                          */
/* ------------------------------------------------------------------------ */

typedef struct {
    int x, y, z;          /* seed voxel */
    float lab[3];         /* seed colour */
    float m[6];           /* seed metric */
} orc_seed;

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Candidate distance, flood.gs:85-94.  The reference evaluates
 * dot((dx^2, 2dxdy, 2dxdz), m0..m2) + dot((dy^2, 2dydz, dz^2), m3..m5) and
 * clamps negatives to 0, leaving contraction and rounding to the GL driver;
 * this fixes one order of fused multiply-adds for the same quadratic form
 * dx(m0dx + 2m1dy + 2m2dz) + dy(m3dy + 2m4dz) + m5dz^2:
 *   u1 = dz*2m2,  u2 = dz*2m4,  u3 = (dz*dz)*m5          (rounded products)
 *   t1 = fma(dx, m0, fma(dy, 2m1, u1)),  t2 = fma(dy, m3, u2)
 *   q  = fma(dx, t1, fma(dy, t2, u3)), then +0 if its sign bit is set
 *   cd = fma(c2, c2, fma(c1, c1, c0*c0)),    d = fma(ratio, q, cd)
 * The u terms depend on the seed and the slice only. */
static inline float cand_dist(const orc_seed* s, int x, int y, int z, const float lab[3], float ratio)
{
    float dx = (float)(x - s->x), dy = (float)(y - s->y), dz = (float)(z - s->z);
    float m1d = 2.0f * s->m[1], m2d = 2.0f * s->m[2], m4d = 2.0f * s->m[4];
    float u1 = dz * m2d, u2 = dz * m4d, u3 = (dz * dz) * s->m[5];
    float t1 = fmaf(dx, s->m[0], fmaf(dy, m1d, u1));
    float t2 = fmaf(dy, s->m[3], u2);
    float q = fmaf(dx, t1, fmaf(dy, t2, u3));
    /* flood.gs:90-93 clamps a negative value to 0 and leaves a NaN, whose
     * distance then fails `curr_dist<min_dist` (:95): a NaN of either sign
     * is rejected here through a +inf distance; -0 becomes +0 */
    if (isnan(q)) return INFINITY;
    if (!(q > 0.0f)) q = 0.0f;
    float c0 = lab[0] - s->lab[0], c1 = lab[1] - s->lab[1], c2 = lab[2] - s->lab[2];
    float cd = fmaf(c2, c2, fmaf(c1, c1, c0 * c0));           /* flood.gs:89 */
    return fmaf(ratio, q, cd);                                /* flood.gs:94 */
}

static const int k_off9x[9] = {-1, 0, 1, -1, 0, 1, -1, 0, 1};   /* flood.gs:36-46 */
static const int k_off9y[9] = {-1, -1, -1, 0, 0, 0, 1, 1, 1};

/* tie_rule 0: lowest seed id among equal distances; 1: first in the
 * reference's scan order (slab z, z-k, z+k; 9 offsets each; strict <). */
ORC_API void orc_jfa_pass(const int32_t* in, int32_t* out, const uint32_t* lab,
                          const orc_seed* seeds, int W, int H, int D, int k,
                          float ratio, int tie_rule)
{
    #pragma omp parallel for schedule(static)
    for (int z = 0; z < D; ++z) {
        int zs[3] = {z, clampi(z - k, 0, D - 1), clampi(z + k, 0, D - 1)};
        for (int y = 0; y < H; ++y) {
            for (int x = 0; x < W; ++x) {
                size_t v = ((size_t)z * H + y) * W + x;
                uint32_t pk = lab[v];
                float lv[3] = {(float)(pk & 255u), (float)((pk >> 8) & 255u), (float)((pk >> 16) & 255u)};
                float best = 100000000.0f;                    /* flood.gs:63 */
                int32_t best_id = -1;
                for (int sl = 0; sl < 3; ++sl) {
                    for (int o = 0; o < 9; ++o) {
                        int xx = clampi(x + k_off9x[o] * k, 0, W - 1);   /* CLAMP_TO_EDGE */
                        int yy = clampi(y + k_off9y[o] * k, 0, H - 1);
                        int32_t id = in[((size_t)zs[sl] * H + yy) * W + xx];
                        if (id < 0) continue;                             /* flood.gs:77 */
                        float d = cand_dist(&seeds[id], x, y, z, lv, ratio);
                        if (d < best || (tie_rule == 0 && d == best && id < best_id)) {
                            best = d;
                            best_id = id;
                        }
                    }
                }
                out[v] = best_id;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* a8: centroid reduction (GPU-AS/shaders/centroids.vs:19-53)                */
/* sums: (A+1) x 7 int64: sum(2x+1), sum(2y+1), sum(2z+1), n, sum L, a, b;    */
/* label -1 accumulates in slot A.                                          */
/* ------------------------------------------------------------------------ */

ORC_API void orc_reduce(const int32_t* labels, const uint32_t* lab, int W, int H, int D,
                        int A, int64_t* sums)
{
    memset(sums, 0, sizeof(int64_t) * 7 * (size_t)(A + 1));
    for (int z = 0; z < D; ++z)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                size_t v = ((size_t)z * H + y) * W + x;
                int l = labels[v];
                int slot = l < 0 ? A : l;
                int64_t* s = sums + (size_t)slot * 7;
                uint32_t pk = lab[v];
                s[0] += 2 * x + 1;
                s[1] += 2 * y + 1;
                s[2] += 2 * z + 1;
                s[3] += 1;
                s[4] += pk & 255u;
                s[5] += (pk >> 8) & 255u;
                s[6] += (pk >> 16) & 255u;
            }
}

/* ------------------------------------------------------------------------ */
/* a9: centroid finalize with the dead-seed rule                             */
/* (GPU-AS/shaders/aver_exception.vs:22-81)                                  */
/* ------------------------------------------------------------------------ */

/* mean position of slot j along one axis: the sums as float division by
 * the count when the count is non-zero, else the raw (zero) sum */
static float slot_mean(const int64_t* sums, int A, int j, int axis)
{
    if (j > A) return 0.0f;                       /* cleared texels past slot A */
    const int64_t* s = sums + (size_t)j * 7;
    if (s[3] == 0) return 0.0f;
    return (float)((double)s[axis] / (2.0 * (double)s[3]));
}

ORC_API void orc_finalize(const int64_t* sums, const uint32_t* lab, int W, int H, int D,
                          int A, orc_seed* seeds)
{
    int dims[3] = {W, H, D};
    for (int i = 0; i < A; ++i) {
        const int64_t* s = sums + (size_t)i * 7;
        int p[3];
        if (s[3] == 0) {
            /* aver_exception.vs:25-62: neighbours by id; slot -1 reads the
             * clamped texel (W-1, row 0), i.e. slot W-1 */
            int jm = i == 0 ? W - 1 : i - 1;
            int jp = i + 1;
            for (int a = 0; a < 3; ++a) {
                float c = floorf(0.5f * (slot_mean(sums, A, jm, a) + slot_mean(sums, A, jp, a)));
                p[a] = clampi((int)c, 0, dims[a] - 1);
            }
            uint32_t pk = lab[((size_t)p[2] * H + p[1]) * W + p[0]];
            seeds[i].lab[0] = (float)(pk & 255u);
            seeds[i].lab[1] = (float)((pk >> 8) & 255u);
            seeds[i].lab[2] = (float)((pk >> 16) & 255u);
        } else {
            /* aver_exception.vs:63-70 */
            for (int a = 0; a < 3; ++a) p[a] = (int)(s[a] / (2 * s[3]));
            for (int c = 0; c < 3; ++c)
                seeds[i].lab[c] = (float)((double)s[4 + c] / (double)s[3]);
        }
        seeds[i].x = p[0];
        seeds[i].y = p[1];
        seeds[i].z = p[2];
    }
}

/* ------------------------------------------------------------------------ */
/* Whole plain run (GPU-AS/Render.cpp:256-293)                                */
/* ------------------------------------------------------------------------ */

typedef struct {
    int W, H, D, K, iters;
    float compactness;
    int tie_rule;        /* 0 lowest id, 1 reference scan order */
    int angle_gate;      /* 1: GPU-AS flow-angle gate on ring fallbacks, iterations >= 2 */
    int schedule;        /* 0 reference, 1 W/2..1 + 1 */
    int prepared;        /* 1: rgb already blurred and in GL row order */
} orc_params;

/* Outputs (any may be NULL):
 *   labels  V int32, final JFA labels (before post-processing)
 *   sums    (A+1) x 7 int64, final reduction
 *   trace   iters x A x 12 float: per iteration the seed record used by the
 *           flood (x, y, z, L, a, b, m0..m5)
 * Returns A (seeds placed) or a negative value on bad input. */
/* Plain run from explicit initial seed voxels (gxyz, A x 3) and grid step s
 * (which sets ratio and the schedule): the seeded form the streaming
 * blocks use (streamGPU-AS/Render.cpp:434-441 builds kept ++ new seeds).
 * Outputs as orc_run. */
ORC_API int orc_run_seeded(const orc_params* P, const uint8_t* rgb, const float* fwd, const float* bwd,
                           const int* gxyz, int A, int s, int32_t* labels_out, int64_t* sums_out, float* trace)
{
    int W = P->W, H = P->H, D = P->D;
    size_t V = (size_t)W * H * D;
    if (W < 1 || H < 1 || D < 1 || A < 1 || P->iters < 1) return -1;
    for (int i = 0; i < A; ++i)
        if (gxyz[i * 3] < 0 || gxyz[i * 3] >= W || gxyz[i * 3 + 1] < 0 || gxyz[i * 3 + 1] >= H ||
            gxyz[i * 3 + 2] < 0 || gxyz[i * 3 + 2] >= D)
            return -3;
    uint8_t* prep = NULL;
    const uint8_t* img = rgb;
    if (!P->prepared) {
        prep = (uint8_t*)malloc(V * 3);
        orc_blur_flip(rgb, W, H, D, prep);
        img = prep;
    }
    uint32_t* lab = (uint32_t*)malloc(V * 4);
    orc_lab_volume(img, (int64_t)V, lab);
    float* m6 = (float*)malloc(V * 6 * sizeof(float));
    uint8_t* psd = (uint8_t*)malloc(V);
    orc_metric_field(fwd, bwd, W, H, D, m6, psd);
    orc_field F = {W, H, D, m6, psd, fwd};

    if (s < 1) s = 1;
    float ratio = (P->compactness / (float)s) * (P->compactness / (float)s);   /* Image.cpp:1051 */
    int steps[64];
    float mult[64];
    int np = orc_schedule(P->schedule, s, W, steps, mult, 64);

    orc_seed* seeds = (orc_seed*)calloc((size_t)A, sizeof(orc_seed));
    int32_t* L0 = (int32_t*)malloc(V * 4);
    int32_t* L1 = (int32_t*)malloc(V * 4);
    int64_t* sums = (int64_t*)malloc(sizeof(int64_t) * 7 * (size_t)(A + 1));

    for (int i = 0; i < A; ++i) {
        seeds[i].x = gxyz[i * 3];
        seeds[i].y = gxyz[i * 3 + 1];
        seeds[i].z = gxyz[i * 3 + 2];
        /* draw_seeds.vs:24: first-iteration colour is the LAB at the seed voxel */
        uint32_t pk = lab[((size_t)seeds[i].z * H + seeds[i].y) * W + seeds[i].x];
        seeds[i].lab[0] = (float)(pk & 255u);
        seeds[i].lab[1] = (float)((pk >> 8) & 255u);
        seeds[i].lab[2] = (float)((pk >> 16) & 255u);
    }
    int32_t* cur = L0;
    int32_t* nxt = L1;
    for (int it = 1; it <= P->iters; ++it) {
        /* seed_matrix (Image.cpp:689) for iteration 1, Ld_next_seeds_matrix
         * (Image.cpp:1300) afterwards: the gate applies from iteration 2 */
        int gate = (it >= 2) && P->angle_gate;
        for (int i = 0; i < A; ++i)
            seed_metric(&F, seeds[i].x, seeds[i].y, seeds[i].z, gate, seeds[i].m);
        if (trace) {
            for (int i = 0; i < A; ++i) {
                float* tr = trace + ((size_t)(it - 1) * A + i) * 12;
                tr[0] = (float)seeds[i].x; tr[1] = (float)seeds[i].y; tr[2] = (float)seeds[i].z;
                tr[3] = seeds[i].lab[0]; tr[4] = seeds[i].lab[1]; tr[5] = seeds[i].lab[2];
                for (int k = 0; k < 6; ++k) tr[6 + k] = seeds[i].m[k];
            }
        }
        /* draw_seeds / draw_next_seeds: clear to -1, later primitive wins */
        for (size_t v = 0; v < V; ++v) cur[v] = -1;
        for (int i = 0; i < A; ++i)
            cur[((size_t)seeds[i].z * H + seeds[i].y) * W + seeds[i].x] = i;
        for (int p = 0; p < np; ++p) {
            orc_jfa_pass(cur, nxt, lab, seeds, W, H, D, steps[p], mult[p] * ratio, P->tie_rule);
            int32_t* t = cur; cur = nxt; nxt = t;
        }
        orc_reduce(cur, lab, W, H, D, A, sums);
        if (it != P->iters)                                   /* Image.cpp:1269 */
            orc_finalize(sums, lab, W, H, D, A, seeds);
    }
    if (labels_out) memcpy(labels_out, cur, V * 4);
    if (sums_out) memcpy(sums_out, sums, sizeof(int64_t) * 7 * (size_t)(A + 1));

    free(prep); free(lab); free(m6); free(psd); free(seeds);
    free(L0); free(L1); free(sums);
    return A;
}

ORC_API int orc_run(const orc_params* P, const uint8_t* rgb, const float* fwd, const float* bwd,
                    int32_t* labels_out, int64_t* sums_out, float* trace)
{
    if (P->W < 1 || P->H < 1 || P->D < 1 || P->K < 1 || P->iters < 1) return -1;
    int s = 0;
    int A = orc_grid_seeds(P->W, P->H, P->D, P->K, NULL, &s);
    if (A <= 0) return -2;
    int* gxyz = (int*)malloc(sizeof(int) * 3 * (size_t)A);
    orc_grid_seeds(P->W, P->H, P->D, P->K, gxyz, &s);
    int r = orc_run_seeded(P, rgb, fwd, bwd, gxyz, A, s, labels_out, sums_out, trace);
    free(gxyz);
    return r;
}

ORC_API int orc_seed_record_floats(void) { return (int)(sizeof(orc_seed) / sizeof(float)); }

/* 3D connectivity enforcement, GPU-AS/Image.cpp:1938-2062
 * (EnforceLabelConnectivity3D).  Voxels are visited in raster order (x
 * fastest, then y, then z); an unvisited voxel starts a new segment, grown
 * breadth-first over the 24 neighbours (dz in -1..1 times the 8 in-plane
 * offsets; the pure z neighbours (0, 0, +-1) are not among them) that carry
 * the same input label.  Before growing, the segment's first voxel scans
 * those neighbours (dz -1, 0, 1; in-plane order (-1,0), (-1,-1), (0,-1),
 * (1,-1), (1,0), (1,1), (0,1), (-1,1)) and keeps the output label of the
 * last one already visited as the adjacent label; the variable persists
 * from one segment to the next and starts at 0.  A segment of fewer than
 * supsz / 4 voxels takes the adjacent label and gives its number back;
 * every other segment gets the next number from 0.  Returns the count of
 * numbers given. */
ORC_API int orc_connectivity(const int32_t* labels, int W, int H, int D, int supsz, int32_t* out)
{
    static const int ox[8] = {-1, -1, 0, 1, 1, 1, 0, -1};
    static const int oy[8] = {0, -1, -1, -1, 0, 1, 1, 1};
    const size_t V = (size_t)W * H * D;
    int* queue = (int*)malloc(V * sizeof(int));
    if (!queue) return -1;
    for (size_t i = 0; i < V; ++i) out[i] = -1;
    int next = 0, adj = 0;
    for (size_t v0 = 0; v0 < V; ++v0) {
        if (out[v0] >= 0) continue;
        const int x0 = (int)(v0 % (size_t)W), y0 = (int)((v0 / (size_t)W) % (size_t)H), z0 = (int)(v0 / ((size_t)W * H));
        out[v0] = next;
        for (int dz = -1; dz <= 1; ++dz) {
            const int z = z0 + dz;
            if (z < 0 || z >= D) continue;
            for (int n = 0; n < 8; ++n) {
                const int x = x0 + ox[n], y = y0 + oy[n];
                if (x < 0 || x >= W || y < 0 || y >= H) continue;
                const int32_t l = out[((size_t)z * H + y) * W + x];
                if (l >= 0) adj = l;
            }
        }
        size_t count = 1;
        queue[0] = (int)v0;
        for (size_t c = 0; c < count; ++c) {
            const size_t q = (size_t)queue[c];
            const int qx = (int)(q % (size_t)W), qy = (int)((q / (size_t)W) % (size_t)H), qz = (int)(q / ((size_t)W * H));
            for (int dz = -1; dz <= 1; ++dz) {
                const int z = qz + dz;
                if (z < 0 || z >= D) continue;
                for (int n = 0; n < 8; ++n) {
                    const int x = qx + ox[n], y = qy + oy[n];
                    if (x < 0 || x >= W || y < 0 || y >= H) continue;
                    const size_t u = ((size_t)z * H + y) * W + x;
                    if (out[u] < 0 && labels[u] == labels[v0]) {
                        out[u] = next;
                        queue[count++] = (int)u;
                    }
                }
            }
        }
        if ((long long)count < (long long)(supsz / 4)) {
            for (size_t c = 0; c < count; ++c) out[queue[c]] = adj;
        } else {
            ++next;
        }
    }
    free(queue);
    return next;
}

/* Nearest float16 value of x (0 <= x < 65520), ties to even: the rounding
 * of a float written to the reference's RGB16F / R16F textures. */
static float half_round(float x)
{
    if (x == 0.0f) return 0.0f;
    int e;
    frexpf(x, &e);                        /* x = f * 2^e, 0.5 <= f < 1 */
    int q = e - 11;                       /* 11 significant bits */
    if (q < -24) q = -24;                 /* float16 subnormal spacing */
    float scaled = ldexpf(x, -q);
    return ldexpf(nearbyintf(scaled), q);
}
ORC_API float orc_half_round(float x) { return half_round(x); }

/* dNum, hNum, wNum of the grid (GPU-AS/Image.cpp:37-45). */
static int grid_dims(int W, int H, int D, int K, int* dn, int* hn, int* wn)
{
    int dNum = (int)pow((float)(K * D * D / W / H), 1.0 / 3.0);
    if (dNum == 0) dNum = 1;
    int fs = K / dNum;
    int wNum = (int)sqrt((float)(fs * W / H));
    if (wNum <= 0) return -1;
    int hNum = fs / wNum;
    if (hNum <= 0) return -1;
    *dn = dNum; *hn = hNum; *wn = wNum;
    return dNum * hNum * wNum;
}

/* Box of grid index i along one axis of n seeds at (integer) positions p,
 * GPU-AS/Image.cpp:360-433: from floor(0.5 (p[i-1]+.5 + p[i]+.5)) (0 for the
 * first) to floor(0.5 (p[i+1]+.5 + p[i]+.5)) (N-1 for the last). */
static void axis_box(const int* p, int n, int N, int i, int* lo, int* hi)
{
    *lo = i > 0 ? (int)floor(0.5 * (((double)p[i - 1] + 0.5) + ((double)p[i] + 0.5))) : 0;
    *hi = i < n - 1 ? (int)floor(0.5 * (((double)p[i + 1] + 0.5) + ((double)p[i] + 0.5))) : N - 1;
}

/* Density-based re-placement of the grid seeds, GPU-AS/Image.cpp:350-620
 * (Image::initialization) with derivative_t.gs and density_grid.vs:
 *  1. every grid seed owns the box between the midpoints to its grid
 *     neighbours; a voxel in two boxes belongs to the later seed (the grid
 *     volume is painted in seed order), stored as float16 (R16F);
 *  2. per voxel g = 0.33 sum|c - c_left| + 0.33 sum|c - c_up| + 0.33
 *     sum|c - c_prev| over R, G, B of the blurred frames (GL rows: up is
 *     y - 1, prev is t - 1; clamped taps at the edges), each product and sum
 *     rounded to float, the result stored as float16 (RGB16F);
 *  3. density of slot k = 3 (the clear colour) + the sum of g over voxels
 *     whose float16 seed id addresses texel k of slice 0 by density_grid.vs's
 *     float arithmetic; the reference adds in blend order, this sums exactly
 *     and rounds once to float;
 *  4. densities 0.001 * d (double product, stored as float), float prefix
 *     sums, A seeds at (i + 1) * total / (A + 1) along them; each takes the
 *     y and z of the grid seed whose interval holds it and an x interpolated
 *     across that seed's x box.
 * img: D x H x W x 3 frames in GL row order (blurred).  Writes A x 3 voxel
 * positions (x, y, z) and, if dens is not NULL, the A slot densities
 * (step 3).  Returns A, or a negative value on bad input. */
ORC_API int orc_density_seeds(const uint8_t* img, int W, int H, int D, int K, int* xyz, float* dens)
{
    int dN, hN, wN;
    const int A = grid_dims(W, H, D, K, &dN, &hN, &wN);
    if (A <= 0) return -1;
    int* g = (int*)malloc(sizeof(int) * 3 * (size_t)A);
    int s = 0;
    if (orc_grid_seeds(W, H, D, K, g, &s) != A) { free(g); return -2; }
    int* px = (int*)malloc(sizeof(int) * (size_t)wN);
    int* py = (int*)malloc(sizeof(int) * (size_t)hN);
    int* pz = (int*)malloc(sizeof(int) * (size_t)dN);
    for (int i = 0; i < wN; ++i) px[i] = g[3 * i];
    for (int j = 0; j < hN; ++j) py[j] = g[3 * (j * wN) + 1];
    for (int t = 0; t < dN; ++t) pz[t] = g[3 * (t * wN * hN) + 2];
    int* mx = (int*)malloc(sizeof(int) * (size_t)W);
    int* my = (int*)malloc(sizeof(int) * (size_t)H);
    int* mz = (int*)malloc(sizeof(int) * (size_t)D);
    for (int x = 0; x < W; ++x) mx[x] = -1;
    for (int y = 0; y < H; ++y) my[y] = -1;
    for (int z = 0; z < D; ++z) mz[z] = -1;
    for (int i = 0; i < wN; ++i) { int lo, hi; axis_box(px, wN, W, i, &lo, &hi); for (int x = lo; x <= hi; ++x) mx[x] = i; }
    for (int j = 0; j < hN; ++j) { int lo, hi; axis_box(py, hN, H, j, &lo, &hi); for (int y = lo; y <= hi; ++y) my[y] = j; }
    for (int t = 0; t < dN; ++t) { int lo, hi; axis_box(pz, dN, D, t, &lo, &hi); for (int z = lo; z <= hi; ++z) mz[z] = t; }

    /* the highest slot a float16 id can address may lie below A - 1 (A =
     * 2146: 2145 rounds to 2144); the reference still reads all A texels
     * (Image.cpp:524-552), and a slot nothing addresses keeps the clear
     * colour, so the table covers max(nslot, A) slots */
    const int nslot = (int)half_round((float)(A - 1)) + 1;
    int64_t* acc = (int64_t*)calloc((size_t)(nslot > A ? nslot : A), sizeof(int64_t));
    const float wr = 1.0f / (float)W;
    for (int z = 0; z < D; ++z)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const uint8_t* c = img + (((size_t)z * H + y) * W + x) * 3;
                const uint8_t* nb[3] = {img + (((size_t)z * H + y) * W + (x > 0 ? x - 1 : 0)) * 3,
                                        img + (((size_t)z * H + (y > 0 ? y - 1 : 0)) * W + x) * 3,
                                        img + (((size_t)(z > 0 ? z - 1 : 0) * H + y) * W + x) * 3};
                float grad = 0.0f;
                for (int q = 0; q < 3; ++q) {
                    int sum = abs((int)nb[q][0] - (int)c[0]) + abs((int)nb[q][1] - (int)c[1]) +
                              abs((int)nb[q][2] - (int)c[2]);
                    grad = grad + 0.33f * (float)sum;
                }
                const float pg = half_round(grad);
                if (mx[x] < 0 || my[y] < 0 || mz[z] < 0) continue;
                const int k = (mz[z] * hN + my[y]) * wN + mx[x];
                const float kr = half_round((float)k);
                const float ty = floorf(kr * wr);
                const float tx = floorf(kr - ty / wr);
                if (tx < 0.0f || tx >= (float)W) continue;
                const int slot = (int)ty * W + (int)tx;
                if (slot < 0 || slot >= nslot) continue;
                acc[slot] += (int64_t)(pg * 4096.0f);
            }
    float* len = (float*)malloc(sizeof(float) * (size_t)A);
    for (int i = 0; i < A; ++i) {
        const float res = (float)(3.0 + (double)acc[i] / 4096.0);
        if (dens) dens[i] = res;
        const float d = (float)(0.001 * (double)res);
        len[i] = i == 0 ? d : d + len[i - 1];
    }
    const float interval = len[A - 1] / (float)(A + 1);
    int j1 = 0, placed = 0;
    for (int i = 0; i < A;) {
        const float sin_ = (float)(i + 1) * interval;
        if (sin_ < len[j1]) {
            int lo, hi;
            axis_box(px, wN, W, j1 % wN, &lo, &hi);
            const float frac = j1 == 0 ? sin_ / len[j1] : (sin_ - len[j1 - 1]) / (len[j1] - len[j1 - 1]);
            const float ax = floorf((float)lo + (float)(hi - lo) * frac);
            xyz[3 * i] = (int)ax;
            xyz[3 * i + 1] = g[3 * j1 + 1];
            xyz[3 * i + 2] = g[3 * j1 + 2];
            ++i;
            ++placed;
        } else if (++j1 == A) {
            break;
        }
    }
    free(g); free(px); free(py); free(pz); free(mx); free(my); free(mz); free(acc); free(len);
    return placed == A ? A : -3;
}
