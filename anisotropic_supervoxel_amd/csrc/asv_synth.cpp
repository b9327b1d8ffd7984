// asv_synth.cpp -- the seeded synthetic video of the benchmark workloads.
//
// The reference evaluates on SegTrackv2 / BuffaloXiph videos (README.md:28),
// which are not available offline; the benchmark configurations define this
// generator instead.  Specification (all randomness from SplitMix64):
//   stream G = SplitMix64(seed):
//     4 corner colours, order top-left, top-right, bottom-left, bottom-right,
//       each r, g, b = 255 * U[0,1)
//     per ellipse e = 0 .. n-1, in this order:
//       semi-axes a, b = min(W,H) * (0.05 + 0.20 * U)
//       orientation theta = pi * U
//       centre cx = W * U, cy = H * U   (at frame 0)
//       colour r, g, b = floor(256 * U) each
//       velocity vx, vy = vmax * (2U - 1) each (pixels per frame)
//   ellipse e is drawn after (on top of) ellipses 0 .. e-1, so the random
//   draw order is the random depth order.
//   frame t, pixel (px, py) of the image (row 0 = top), centre (px+.5, py+.5):
//     inside e  <=>  (u/a)^2 + (w/b)^2 <= 1, with d = centre - (cx + vx t, cy + vy t),
//                   u = d.x cos(theta) + d.y sin(theta), w = -d.x sin(theta) + d.y cos(theta)
//     top = largest e containing the pixel, or none
//     base colour = top colour, or the bilinear background at
//                   (px / (W-1), py / (H-1)) between the corners
//     noise: stream N_t = SplitMix64(seed ^ (0x9E3779B97F4A7C15 * (t + 1))),
//            Box-Muller pairs in raster order, channel fastest
//            (z0 = sqrt(-2 ln u1) cos(2 pi u2), z1 = ... sin, u1 in (0,1])
//     rgb = clamp(floor(base + sigma z + 0.5), 0, 255)
//     forward flow = (vx, vy) of top, 0 on background; backward flow = -(vx, vy)
//     region id = top + 1, or 0 on background
#include "../../include/asv.h"

#include <atomic>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }          // [0,1)
    double uniform_open0() { return (double)((next() >> 11) + 1) * (1.0 / 9007199254740992.0); }  // (0,1]
};

struct Ellipse {
    double a, b, ct, st, cx, cy, vx, vy;
    int col[3];
};

}  // namespace

extern "C" int asv_synth_video_range(uint64_t seed, int32_t W, int32_t H, int32_t D, int32_t n_ell, float vmax,
                                     float sigma, int32_t t0, int32_t t1, uint8_t* rgb, float* fwd, float* bwd,
                                     int32_t* region)
{
    if (W < 2 || H < 2 || D < 1 || n_ell < 0 || t0 < 0 || t1 > D || t0 >= t1) return ASV_ERR_ARG;
    const double kPi = 3.14159265358979323846;
    SplitMix64 g(seed);
    double corner[4][3];
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 3; ++k) corner[c][k] = 255.0 * g.uniform();
    std::vector<Ellipse> el((size_t)n_ell);
    const double mn = (double)(W < H ? W : H);
    for (int e = 0; e < n_ell; ++e) {
        Ellipse& E = el[(size_t)e];
        E.a = mn * (0.05 + 0.20 * g.uniform());
        E.b = mn * (0.05 + 0.20 * g.uniform());
        double th = kPi * g.uniform();
        E.ct = std::cos(th);
        E.st = std::sin(th);
        E.cx = (double)W * g.uniform();
        E.cy = (double)H * g.uniform();
        for (int k = 0; k < 3; ++k) {
            int v = (int)std::floor(256.0 * g.uniform());
            E.col[k] = v > 255 ? 255 : v;
        }
        E.vx = (double)vmax * (2.0 * g.uniform() - 1.0);
        E.vy = (double)vmax * (2.0 * g.uniform() - 1.0);
    }

    std::atomic<int> next_frame(t0);
    auto worker = [&]() {
    for (int t = next_frame++; t < t1; t = next_frame++) {
        SplitMix64 nz(seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(t + 1)));
        double spare = 0.0;
        bool have_spare = false;
        for (int py = 0; py < H; ++py) {
            for (int px = 0; px < W; ++px) {
                size_t v = ((size_t)(t - t0) * H + py) * W + px;
                double cxp = px + 0.5, cyp = py + 0.5;
                int top = -1;
                for (int e = n_ell - 1; e >= 0; --e) {
                    const Ellipse& E = el[(size_t)e];
                    double dx = cxp - (E.cx + E.vx * t), dy = cyp - (E.cy + E.vy * t);
                    double u = dx * E.ct + dy * E.st;
                    double w = -dx * E.st + dy * E.ct;
                    double q = (u / E.a) * (u / E.a) + (w / E.b) * (w / E.b);
                    if (q <= 1.0) { top = e; break; }
                }
                double base[3];
                if (top >= 0) {
                    for (int k = 0; k < 3; ++k) base[k] = (double)el[(size_t)top].col[k];
                } else {
                    double fx = (double)px / (double)(W - 1), fy = (double)py / (double)(H - 1);
                    for (int k = 0; k < 3; ++k)
                        base[k] = (1.0 - fx) * (1.0 - fy) * corner[0][k] + fx * (1.0 - fy) * corner[1][k] +
                                  (1.0 - fx) * fy * corner[2][k] + fx * fy * corner[3][k];
                }
                for (int k = 0; k < 3; ++k) {
                    double z;
                    if (have_spare) { z = spare; have_spare = false; }
                    else {
                        double u1 = nz.uniform_open0(), u2 = nz.uniform();
                        double r = std::sqrt(-2.0 * std::log(u1));
                        z = r * std::cos(2.0 * kPi * u2);
                        spare = r * std::sin(2.0 * kPi * u2);
                        have_spare = true;
                    }
                    double val = std::floor(base[k] + (double)sigma * z + 0.5);
                    val = val < 0.0 ? 0.0 : (val > 255.0 ? 255.0 : val);
                    if (rgb) rgb[v * 3 + k] = (uint8_t)val;
                }
                float fu = 0.0f, fv = 0.0f;
                if (top >= 0) { fu = (float)el[(size_t)top].vx; fv = (float)el[(size_t)top].vy; }
                if (fwd) { fwd[v * 2] = fu; fwd[v * 2 + 1] = fv; }
                if (bwd) { bwd[v * 2] = -fu; bwd[v * 2 + 1] = -fv; }
                if (region) region[v] = top + 1;
            }
        }
    }
    };
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt == 0 ? 1 : (nt > 16 ? 16 : nt);
    if ((int)nt > t1 - t0) nt = (unsigned)(t1 - t0);
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < nt; ++i) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    return ASV_OK;
}

extern "C" int asv_synth_video(uint64_t seed, int32_t W, int32_t H, int32_t D, int32_t n_ell, float vmax,
                               float sigma, uint8_t* rgb, float* fwd, float* bwd, int32_t* region)
{
    return asv_synth_video_range(seed, W, H, D, n_ell, vmax, sigma, 0, D, rgb, fwd, bwd, region);
}
