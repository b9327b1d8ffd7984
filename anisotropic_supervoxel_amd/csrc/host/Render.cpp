// Render.cpp -- see Render.h.  Timers are wall-clock (steady_clock) around
// synchronised stages, where the reference printed clock() CPU time.
#include "Render.h"

#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <fstream>

#include "Image.h"

namespace {

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool read_file(const std::string& path, void* dst, size_t bytes)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    size_t n = std::fread(dst, 1, bytes, f);
    std::fclose(f);
    return n == bytes;
}

bool write_file(const std::string& path, const void* src, size_t bytes)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    size_t n = std::fwrite(src, 1, bytes, f);
    std::fclose(f);
    return n == bytes;
}

void mkdirs(const std::string& path)
{
    for (size_t i = 1; i <= path.size(); ++i)
        if (i == path.size() || path[i] == '/') mkdir(path.substr(0, i).c_str(), 0755);
}

}  // namespace

Render::Render()
    : Lloyd_ends(false), Lloyd_curr_iter(0), Lloyd_max_iter(20), width(0), height(0), depth(0),
      output_label(true), output_render(true), connect(false), density_init(false), ms_total(0), ms_loop(0)
{
}

Render::~Render() {}

bool Render::write_clip(const std::string& dir, const uint8_t* rgb, const float* fwd, const float* bwd, int W, int H,
                        int D)
{
    mkdirs(dir);
    std::ofstream sz(dir + "/size.txt");
    sz << W << " " << H << " " << D << "\n";
    sz.close();
    size_t V = (size_t)W * H * D;
    return write_file(dir + "/frames.rgb", rgb, V * 3) && write_file(dir + "/flow_fwd.f32", fwd, V * 8) &&
           write_file(dir + "/flow_bwd.f32", bwd, V * 8);
}

int Render::run(std::string& dataset_path_, std::string& save_path_, std::string& dbname_, std::string& vdname_,
                int seeds_num_, float compactness_)
{
    std::string video_path = dataset_path_ + "/" + vdname_;
    std::string save_path = save_path_ + dbname_ + "/AS/" + vdname_ + "/";       // GPU-AS/Render.cpp:137-148
    mkdirs(save_path);
    std::printf("%s,%d:\n", vdname_.c_str(), seeds_num_);
    int W = 0, H = 0, D = 0;
    {
        std::ifstream sz(video_path + "/size.txt");
        if (!(sz >> W >> H >> D) || W < 2 || H < 2 || D < 1) {
            std::printf("Error open %s/size.txt\n", video_path.c_str());
            return -1;
        }
    }
    size_t V = (size_t)W * H * D;
    std::vector<uint8_t> rgb(V * 3);
    std::vector<float> fwd(V * 2), bwd(V * 2);
    double t0 = now_ms();
    if (!read_file(video_path + "/frames.rgb", rgb.data(), rgb.size()) ||
        !read_file(video_path + "/flow_fwd.f32", fwd.data(), fwd.size() * 4) ||
        !read_file(video_path + "/flow_bwd.f32", bwd.data(), bwd.size() * 4)) {
        std::printf("Error reading clip files in %s\n", video_path.c_str());
        return -1;
    }
    std::printf("load data, time_ms:%.1f\n", now_ms() - t0);
    std::string prefix = save_path + std::to_string(seeds_num_);
    return run_arrays(rgb.data(), fwd.data(), bwd.data(), W, H, D, seeds_num_, compactness_, prefix, nullptr);
}

int Render::run_arrays(const uint8_t* rgb, const float* fwd, const float* bwd, int W, int H, int D, int seeds_num,
                       float compactness, const std::string& out_prefix, std::vector<int>* labels_out)
{
    width = W;
    height = H;
    depth = D;
    std::string db = "synthetic";
    double t0 = now_ms();
    Image image;
    // GPU-AS/Render.cpp:256-274
    if (!image.init_gpu(db, const_cast<unsigned char*>(rgb), W, H, D, seeds_num, Lloyd_max_iter, compactness))
        return -1;
    if (density_init && !image.initialization(true)) return -1;          // GPU-AS/Render.cpp:258
    if (!image.forward_backward_flow(fwd, bwd)) return -1;
    if (!image.seed_matrix() || !image.RGB2LAB()) return -1;
    // GPU-AS/Render.cpp:277-293
    double tl = now_ms();
    Lloyd_curr_iter = 0;
    Lloyd_ends = false;
    while (!Lloyd_ends) {
        Lloyd_curr_iter++;
        bool good = Lloyd_curr_iter == 1 ? image.Ld_first_init() : image.Ld_next_seeds_matrix(Lloyd_curr_iter);
        if (!good || !image.flood(Lloyd_curr_iter)) return -1;
        if (Lloyd_curr_iter == Lloyd_max_iter) Lloyd_ends = true;
        if (!image.centroids_next_seeds(Lloyd_curr_iter)) return -1;
    }
    std::vector<int> labels((size_t)W * H * D);
    if (connect) {                                                // GPU-AS/Render.cpp:324-341
        int label_num = 0;
        if (!image.enforce_connectivity_3d(labels.data(), label_num)) return -1;
        std::printf("label_num=%d\n", label_num);
    } else if (!image.check_dead_seeds(labels.data())) {          // synchronises the stream
        return -1;
    }
    ms_loop = now_ms() - tl;
    std::printf(connect ? "Lloyd loop + connectivity, time_ms:%.2f\n" : "Lloyd loop + check_dead_seeds, time_ms:%.2f\n",
                ms_loop);
    if (!out_prefix.empty()) {
        double ts = now_ms();
        if (!image.print_sv_txt(labels.data(), out_prefix, output_label, connect)) return -1;
        if (output_render) {
            std::string rdir = out_prefix + "_render/";
            mkdirs(rdir);
            if (!image.print_render(labels.data(), rdir)) return -1;
        }
        std::printf("save txt, time_ms:%.1f\n", now_ms() - ts);
    }
    ms_total = now_ms() - t0;
    std::printf("video segmentation done, total time(ms):%.1f\n", ms_total);
    if (labels_out) labels_out->swap(labels);
    return 0;
}
