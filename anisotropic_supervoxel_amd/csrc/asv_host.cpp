// asv_host.cpp -- host-side stages around the GPU loop: post-processing of
// the final labels (dead seeds, unlabelled voxels), the on-disk formats
// (label text, pseudo-colour frames) and the segmentation-quality metrics.
// Pure C++ on host memory; no device calls.
#include "../../include/asv.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

// Image::sequence_label (GPU-AS/Image.cpp:1727-1778).
//  dead_seed: remap alive ids to 0..n-1 in raster order.  The reference
//    compares against labels[0]'s original value, which it never updates
//    (:1733-1746); a voxel equal to that value keeps the previous voxel's
//    mapped label, and -1 voxels do too.  std::map::operator[] on an id that
//    is not alive yields 0, as there.
//  !dead_seed: an unlabelled voxel (-1) takes the previous voxel's label in
//    raster order; a leading -1 stays -1 (:1750-1776).
int64_t sequence_label(int32_t* labels, int64_t V, std::map<int, int>& seq, bool dead_seed)
{
    int64_t changed = 0;
    if (V <= 0) return 0;
    if (dead_seed) {
        int last_label = labels[0];
        int last_label_map = seq[last_label];
        labels[0] = last_label_map;
        for (int64_t i = 1; i < V; ++i) {
            if (labels[i] != last_label) {
                if (labels[i] != -1) last_label_map = seq[labels[i]];
            }
            labels[i] = last_label_map;
        }
        return V;
    }
    int last_label = 0;
    if (labels[0] < 0) ++changed;
    else last_label = labels[0];
    for (int64_t i = 1; i < V; ++i) {
        if (labels[i] != last_label) {
            if (labels[i] == -1) {
                labels[i] = last_label;
                ++changed;
            }
            last_label = labels[i];
        }
    }
    return changed;
}

}  // namespace

extern "C" {

// Image::check_dead_seeds (GPU-AS/Image.cpp:1579-1720) on host memory:
// sums is the final reduction, (A+1) x 8 int64 as asv_download_sums writes
// it; labels (V int32, GL raster order) is rewritten in place.  Outputs the
// number of dead seeds and the final label count (final_label_num).
int asv_postprocess(const int64_t* sums, int32_t A, int32_t* labels, int64_t V, int32_t* dead_out,
                    int32_t* final_label_num)
{
    if (!sums || !labels || A < 1 || V < 1) return ASV_ERR_ARG;
    const bool invalid_label_exist = sums[(size_t)A * 8 + 3] > 0;          // :1590
    std::vector<int> seed_invalid;
    std::map<int, int> seq;
    int valid = 0;
    for (int i = 0; i < A; ++i) {
        if (sums[(size_t)i * 8 + 3] < 1) seed_invalid.push_back(i);        // :1605-1610
        else seq.insert(std::pair<int, int>(i, valid++));
    }
    const int dead = (int)seed_invalid.size();
    if (dead == 0) {
        if (invalid_label_exist) sequence_label(labels, V, seq, false);
    } else if (dead <= 3) {
        // label_replace.gs:18-34, one pass per dead seed: the current last
        // id takes the dead seed's id (:1652-1673)
        int last_seed_index = A;
        for (int dsi : seed_invalid) {
            last_seed_index -= 1;
            for (int64_t v = 0; v < V; ++v)
                if (labels[v] == last_seed_index) labels[v] = dsi;
        }
        if (invalid_label_exist) sequence_label(labels, V, seq, false);
    } else {
        sequence_label(labels, V, seq, true);
    }
    if (dead_out) *dead_out = dead;
    if (final_label_num) *final_label_num = valid;
    return ASV_OK;
}

// Image::print_sv_txt (GPU-AS/Image.cpp:1780-1806): per frame, rows from the
// top of the image (GL row H-1) down, each label left-aligned in width 4 and
// followed by a space, a newline per row and two blank lines per frame.
int asv_write_label_txt(const int32_t* labels, int32_t W, int32_t H, int32_t D, const char* path)
{
    if (!labels || !path || W < 1 || H < 1 || D < 1) return ASV_ERR_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return ASV_ERR_ARG;
    std::vector<char> line((size_t)W * 16 + 4);
    for (int t = 0; t < D; ++t) {
        for (int i = H - 1; i >= 0; --i) {
            const int32_t* row = labels + ((size_t)t * H + i) * W;
            size_t n = 0;
            for (int j = 0; j < W; ++j) n += (size_t)std::snprintf(line.data() + n, 16, "%-4d ", row[j]);
            line[n++] = '\n';
            std::fwrite(line.data(), 1, n, f);
        }
        std::fputs("\n\n", f);
    }
    std::fclose(f);
    return ASV_OK;
}

// Image::print_render (GPU-AS/Image.cpp:1883-1935): one random colour per
// final label from rand() % 256 (generator state as after srand(1), the C
// default), black for -1; frames are written as binary PPM (P6), top row
// first, to <prefix>NNNNN.ppm.  The reference shows them with imshow.
int asv_write_render_ppm(const int32_t* labels, int32_t W, int32_t H, int32_t D, int32_t nlabels,
                         const char* prefix)
{
    if (!labels || !prefix || W < 1 || H < 1 || D < 1 || nlabels < 0) return ASV_ERR_ARG;
    std::srand(1);
    std::vector<unsigned char> color((size_t)nlabels * 3);
    for (auto& c : color) c = (unsigned char)(std::rand() % 256);
    std::vector<unsigned char> img((size_t)W * H * 3);
    for (int t = 0; t < D; ++t) {
        for (int yi = 0; yi < H; ++yi) {
            const int32_t* row = labels + ((size_t)t * H + (H - 1 - yi)) * W;
            for (int x = 0; x < W; ++x) {
                int l = row[x];
                unsigned char* px = &img[((size_t)yi * W + x) * 3];
                if (l < 0 || l >= nlabels) { px[0] = px[1] = px[2] = 0; }
                else { px[0] = color[3 * (size_t)l]; px[1] = color[3 * (size_t)l + 1]; px[2] = color[3 * (size_t)l + 2]; }
            }
        }
        std::string fn = std::string(prefix);
        char num[32];
        std::snprintf(num, sizeof(num), "%05d.ppm", t);
        fn += num;
        FILE* f = std::fopen(fn.c_str(), "wb");
        if (!f) return ASV_ERR_ARG;
        std::fprintf(f, "P6\n%d %d\n255\n", W, H);
        std::fwrite(img.data(), 1, img.size(), f);
        std::fclose(f);
    }
    return ASV_OK;
}

// Segmentation quality against ground-truth regions (both D x H x W, the
// same row order):
//  3D undersegmentation error: mean over ground-truth regions g of
//    (sum of |s| over supervoxels s meeting g  -  |g|) / |g|;
//  3D boundary recall: fraction of ground-truth boundary voxels (a 6-
//    neighbour in x, y or t with another region) that have a supervoxel
//    boundary voxel within Chebyshev distance 1 in the same frame.
// Labels < 0 count as one extra supervoxel.
int asv_eval_metrics(const int32_t* labels, const int32_t* regions, int32_t W, int32_t H, int32_t D,
                     double* undersegmentation, double* boundary_recall)
{
    if (!labels || !regions || W < 1 || H < 1 || D < 1) return ASV_ERR_ARG;
    const int64_t V = (int64_t)W * H * D;
    std::map<int, int64_t> sv_size, gt_size;
    std::map<std::pair<int, int>, int64_t> overlap;
    for (int64_t v = 0; v < V; ++v) {
        sv_size[labels[v]]++;
        gt_size[regions[v]]++;
        overlap[std::make_pair(regions[v], labels[v])]++;
    }
    double ue = 0.0;
    std::map<int, int64_t> leak;
    for (auto& o : overlap) leak[o.first.first] += sv_size[o.first.second];
    for (auto& g : gt_size) ue += (double)(leak[g.first] - g.second) / (double)g.second;
    ue /= (double)gt_size.size();

    auto idx = [&](int t, int y, int x) { return ((int64_t)t * H + y) * W + x; };
    std::vector<unsigned char> svb((size_t)V, 0);
    for (int t = 0; t < D; ++t)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                int l = labels[idx(t, y, x)];
                bool b = (x + 1 < W && labels[idx(t, y, x + 1)] != l) || (x > 0 && labels[idx(t, y, x - 1)] != l) ||
                         (y + 1 < H && labels[idx(t, y + 1, x)] != l) || (y > 0 && labels[idx(t, y - 1, x)] != l) ||
                         (t + 1 < D && labels[idx(t + 1, y, x)] != l) || (t > 0 && labels[idx(t - 1, y, x)] != l);
                svb[(size_t)idx(t, y, x)] = b;
            }
    int64_t gtb = 0, hit = 0;
    for (int t = 0; t < D; ++t)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                int g = regions[idx(t, y, x)];
                bool b = (x + 1 < W && regions[idx(t, y, x + 1)] != g) || (x > 0 && regions[idx(t, y, x - 1)] != g) ||
                         (y + 1 < H && regions[idx(t, y + 1, x)] != g) || (y > 0 && regions[idx(t, y - 1, x)] != g) ||
                         (t + 1 < D && regions[idx(t + 1, y, x)] != g) || (t > 0 && regions[idx(t - 1, y, x)] != g);
                if (!b) continue;
                ++gtb;
                bool found = false;
                for (int dy = -1; dy <= 1 && !found; ++dy)
                    for (int dx = -1; dx <= 1 && !found; ++dx) {
                        int yy = y + dy, xx = x + dx;
                        if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
                        found = svb[(size_t)idx(t, yy, xx)] != 0;
                    }
                hit += found;
            }
    if (undersegmentation) *undersegmentation = ue;
    if (boundary_recall) *boundary_recall = gtb ? (double)hit / (double)gtb : 1.0;
    return ASV_OK;
}

}  // extern "C"
