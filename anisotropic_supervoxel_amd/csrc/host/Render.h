// Render.h -- the reference's pipeline driver (GPU-AS/Render.h:26-50): load a
// frame volume and its flow, set K and the Lloyd iteration count, run, and
// emit the per-voxel label text and pseudo-colour frames.  The GL context,
// PNG decoding and Farneback flow of the reference are replaced by raw
// clip files (or in-memory arrays) and ground-truth flow.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

class Render {
public:
    Render();
    ~Render();

    // GPU-AS/Render.cpp:100-350.  Reads <dataset_path>/<vdname>/ as written
    // by write_clip (size.txt "W H D", frames.rgb, flow_fwd.f32,
    // flow_bwd.f32) and writes <save_path><dbname>/AS/<vdname>/<K>.txt and
    // render/NNNNN.ppm.  Returns 0, or -1 on failure.
    int run(std::string& dataset_path_, std::string& save_path_, std::string& dbname_, std::string& vdname_,
            int seeds_num_, float compactness_);
    // Same pipeline on arrays already in memory (image rows).  labels_out may
    // be null; out_prefix empty skips the files.
    int run_arrays(const uint8_t* rgb, const float* fwd, const float* bwd, int W, int H, int D, int seeds_num,
                   float compactness, const std::string& out_prefix, std::vector<int>* labels_out);

    static bool write_clip(const std::string& dir, const uint8_t* rgb, const float* fwd, const float* bwd, int W,
                           int H, int D);

    bool Lloyd_ends;
    int Lloyd_curr_iter;
    int Lloyd_max_iter;      // 20 as in GPU-AS/Render.cpp:11
    int width, height, depth;
    bool output_label;       // write the label text (the reference hard-codes false)
    bool output_render;      // write the pseudo-colour frames
    bool connect;            // 3D connectivity in place of check_dead_seeds (the reference's step 6.2)
    bool density_init;       // density-based seed re-placement (the reference's step 2, Render.cpp:258)
    double ms_total, ms_loop;
};
