// Image.h -- the reference's per-clip algorithm object (GPU-AS/Image.h:27-108)
// kept as the host surface, with its state in HIP device buffers behind the
// C ABI (include/asv.h) instead of GL textures.  Method names, argument
// meaning and call order follow the reference; methods return false and
// print the library's message where the reference printed and returned.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/asv.h"

class Image {
public:
    Image();
    ~Image();

    // GPU-AS/Image.cpp:320-345.  data: D frames of H x W x 3 RGB u8, image
    // row order (the blur and vertical flip of Render::cvMat2TexureInput run
    // on the device); data_prepared = true when they were already applied.
    bool init_gpu(std::string& dbname, unsigned char* data, int width, int height, int depth, int seeds_num,
                  int Lloyd_max_iter, float compactness, bool data_prepared = false);
    // GPU-AS/Image.cpp:350-620: density-based re-placement of the grid seeds
    // (applied when seed_matrix / RGB2LAB prepare the clip); false keeps the
    // plain grid.
    bool initialization(bool density = true);
    // GPU-AS/Image.cpp:643-686: the flow fields, here ground truth (D x H x W
    // x 2 float, image rows) in place of Farneback.
    bool forward_backward_flow(const float* flow_forward, const float* flow_backward);
    // GPU-AS/Image.cpp:689-872 and :296-315: per-voxel metric field and LAB.
    bool seed_matrix();
    bool RGB2LAB();
    // GPU-AS/Image.cpp:1017-1045, :1061-1190, :1238-1297, :1300-1575.
    bool Ld_first_init();
    bool flood(int Lloyd_curr_iter);
    bool centroids_next_seeds(int Lloyd_curr_iter);
    bool Ld_next_seeds_matrix(int Lloyd_curr_iter);
    // GPU-AS/Image.cpp:1579-1720: labels is caller-allocated int[W*H*D].
    bool check_dead_seeds(int* labels);
    // GPU-AS/Image.cpp:1808-1837 with EnforceLabelConnectivity3D
    // (:1938-2062), on the raw labels of the last pass: labels is
    // caller-allocated int[W*H*D]; label_num receives the segments kept.
    bool enforce_connectivity_3d(int* labels, int& label_num);
    // GPU-AS/Image.cpp:1780-1806 and :1883-1935.
    bool print_sv_txt(int* labels, std::string dirName, bool output_label, bool connected);
    bool print_render(int* labels, std::string prefix);

    int seeds_placed() const { return ac_seeds_num; }
    int final_labels() const { return final_label_num; }
    int dead_seeds() const { return dead_seed_num; }
    void set_sync(bool on) { sync_each_stage = on; }

private:
    bool ok(int rc, const char* what);

    asv_ctx* ctx = nullptr;
    int width = 0, height = 0, depth = 0;
    int ac_seeds_num = 0;
    int final_label_num = 0;
    int dead_seed_num = 0;
    bool prepared = false;
    bool sync_each_stage = false;
};
