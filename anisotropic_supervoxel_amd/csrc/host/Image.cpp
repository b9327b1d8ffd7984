// Image.cpp -- stage methods over the C ABI; see Image.h.
#include "Image.h"

#include <cstdio>
#include <vector>

Image::Image() {}

Image::~Image()
{
    if (ctx) asv_destroy(ctx);
}

bool Image::ok(int rc, const char* what)
{
    if (rc == ASV_OK) return true;
    std::printf("%s failed (%d): %s\n", what, rc, asv_last_error());
    return false;
}

bool Image::init_gpu(std::string& dbname, unsigned char* data, int width_, int height_, int depth_, int seeds_num,
                     int Lloyd_max_iter, float compactness, bool data_prepared)
{
    (void)dbname;
    width = width_;
    height = height_;
    depth = depth_;
    asv_params p;
    asv_default_params(&p, width, height, depth, seeds_num);
    p.lloyd_iters = Lloyd_max_iter;
    p.compactness = compactness;
    p.frames_prepared = data_prepared ? 1 : 0;
    if (ctx) asv_destroy(ctx);
    ctx = nullptr;
    if (!ok(asv_create(&p, &ctx), "init_gpu")) return false;
    ac_seeds_num = asv_seeds_placed(ctx);
    std::printf("seed num placed:%d\n", ac_seeds_num);                      // Image.cpp:92
    return ok(asv_upload_frames(ctx, data, nullptr), "init_gpu upload");
}

bool Image::forward_backward_flow(const float* flow_forward, const float* flow_backward)
{
    prepared = false;
    return ok(asv_upload_flow(ctx, flow_forward, flow_backward, nullptr), "forward_backward_flow");
}

bool Image::seed_matrix()
{
    // The metric field is built together with LAB in one streaming kernel.
    if (prepared) return true;
    prepared = ok(asv_prepare(ctx, nullptr), "seed_matrix");
    return prepared;
}

bool Image::RGB2LAB() { return seed_matrix(); }

bool Image::Ld_first_init() { return ok(asv_seed_stage(ctx, 1, nullptr), "Ld_first_init"); }

bool Image::flood(int it)
{
    std::printf("Lloyd iteration: %d\n", it);                                // Image.cpp:1063
    return ok(asv_flood(ctx, it, nullptr), "flood");
}

bool Image::centroids_next_seeds(int it) { return ok(asv_centroids(ctx, it, nullptr), "centroids_next_seeds"); }

bool Image::Ld_next_seeds_matrix(int it) { return ok(asv_seed_stage(ctx, it, nullptr), "Ld_next_seeds_matrix"); }

bool Image::check_dead_seeds(int* labels)
{
    const int64_t V = (int64_t)width * height * depth;
    std::vector<int64_t> sums((size_t)(ac_seeds_num + 1) * 8);
    if (!ok(asv_download_sums(ctx, sums.data(), nullptr), "check_dead_seeds sums")) return false;
    if (!ok(asv_download_labels(ctx, labels, nullptr), "check_dead_seeds labels")) return false;
    if (sums[(size_t)ac_seeds_num * 8 + 3] > 0)
        std::printf("!!!-1 label exist,num:%lld\n", (long long)sums[(size_t)ac_seeds_num * 8 + 3]);
    int dead = 0, nfin = 0;
    if (!ok(asv_postprocess(sums.data(), ac_seeds_num, labels, V, &dead, &nfin), "check_dead_seeds")) return false;
    dead_seed_num = dead;
    final_label_num = nfin;
    std::printf("!!!dead seeds num:%d\nfinal label num:%d\n", dead, nfin);  // Image.cpp:1621-1623
    return true;
}

bool Image::initialization(bool density)
{
    return ok(asv_initialization(ctx, density ? ASV_INIT_DENSITY : ASV_INIT_GRID), "initialization");
}

bool Image::enforce_connectivity_3d(int* labels, int& label_num)
{
    int32_t n = 0;
    if (!ok(asv_connectivity(ctx, 0, &n, nullptr), "enforce_connectivity_3d")) return false;   // SUPSZ = V / K
    if (!ok(asv_download_labels(ctx, labels, nullptr), "enforce_connectivity_3d labels")) return false;
    final_label_num = n;
    label_num = n;
    return true;
}

bool Image::print_sv_txt(int* labels, std::string dirName, bool output_label, bool connected)
{
    if (!output_label) return true;
    std::string txt = dirName + ".txt";
    if (!ok(asv_write_label_txt(labels, width, height, depth, txt.c_str()), "print_sv_txt")) return false;
    std::printf(connected ? "save connected label!!!\n" : "save not connected label!!!\n");
    return true;
}

bool Image::print_render(int* labels, std::string prefix)
{
    return ok(asv_write_render_ppm(labels, width, height, depth, final_label_num, prefix.c_str()), "print_render");
}
