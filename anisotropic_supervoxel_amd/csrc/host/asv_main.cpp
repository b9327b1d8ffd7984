// asv_main.cpp -- command-line front end of the Render host (the reference's
// main() hard-codes its dataset and K list, GPU-AS/main.cpp:49-86).
//
//   asv --synth SEED,W,H,D,ELLIPSES,VMAX,SIGMA [--write-clip DIR]
//   asv --dataset DIR --video NAME
//   common: --K 500 --iters 20 --compactness 8 --save DIR --db NAME
//           --no-txt --no-render --eval --connect (3D connectivity instead of
//           check_dead_seeds) --density (density-based seed placement)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/asv.h"
#include "Render.h"

static void usage()
{
    std::printf("usage: asv (--synth SEED,W,H,D,E,VMAX,SIGMA | --dataset DIR --video NAME) [--K n] [--iters n]\n"
                "           [--compactness c] [--save DIR] [--db NAME] [--write-clip DIR] [--no-txt] [--no-render]\n"
                "           [--eval] [--connect] [--density]\n");
}

int main(int argc, char** argv)
{
    std::string synth, dataset, video, save = "results/", db = "synthetic", write_clip;
    int K = 500, iters = 20;
    float compactness = 8.0f;
    bool txt = true, render = true, eval = false, connect = false, density = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) { usage(); std::exit(2); }
            return argv[++i];
        };
        if (a == "--synth") synth = next();
        else if (a == "--dataset") dataset = next();
        else if (a == "--video") video = next();
        else if (a == "--save") save = next();
        else if (a == "--db") db = next();
        else if (a == "--write-clip") write_clip = next();
        else if (a == "--K") K = std::atoi(next().c_str());
        else if (a == "--iters") iters = std::atoi(next().c_str());
        else if (a == "--compactness") compactness = (float)std::atof(next().c_str());
        else if (a == "--no-txt") txt = false;
        else if (a == "--no-render") render = false;
        else if (a == "--eval") eval = true;
        else if (a == "--connect") connect = true;
        else if (a == "--density") density = true;
        else { usage(); return 2; }
    }
    if (!save.empty() && save.back() != '/') save += "/";
    Render r;
    r.Lloyd_max_iter = iters;
    r.output_label = txt;
    r.output_render = render;
    r.connect = connect;
    r.density_init = density;
    if (!synth.empty()) {
        unsigned long long seed;
        int W, H, D, E;
        float vmax, sigma;
        if (std::sscanf(synth.c_str(), "%llu,%d,%d,%d,%d,%f,%f", &seed, &W, &H, &D, &E, &vmax, &sigma) != 7) {
            usage();
            return 2;
        }
        size_t V = (size_t)W * H * D;
        std::vector<uint8_t> rgb(V * 3);
        std::vector<float> fwd(V * 2), bwd(V * 2);
        std::vector<int32_t> reg(V);
        if (asv_synth_video(seed, W, H, D, E, vmax, sigma, rgb.data(), fwd.data(), bwd.data(), reg.data()) != ASV_OK) {
            std::printf("bad synthetic parameters\n");
            return 2;
        }
        if (!write_clip.empty()) {
            if (!Render::write_clip(write_clip, rgb.data(), fwd.data(), bwd.data(), W, H, D)) return 1;
            std::FILE* f = std::fopen((write_clip + "/regions.i32").c_str(), "wb");
            if (f) { std::fwrite(reg.data(), 4, V, f); std::fclose(f); }
            std::printf("wrote clip %s\n", write_clip.c_str());
            return 0;
        }
        std::string vdname = "synth" + std::to_string(seed);
        std::string out = save + db + "/AS/" + vdname + "/";
        std::string cmd = "mkdir -p " + out;
        if (std::system(cmd.c_str()) != 0) return 1;
        std::vector<int> labels;
        if (r.run_arrays(rgb.data(), fwd.data(), bwd.data(), W, H, D, K, compactness,
                         (txt || render) ? out + std::to_string(K) : std::string(), &labels))
            return 1;
        if (eval) {
            // labels are in GL rows; regions in image rows
            std::vector<int32_t> flipped(V);
            for (int t = 0; t < D; ++t)
                for (int y = 0; y < H; ++y)
                    std::memcpy(&flipped[((size_t)t * H + y) * W], &labels[((size_t)t * H + (H - 1 - y)) * W],
                                (size_t)W * 4);
            double ue = 0, br = 0;
            asv_eval_metrics(flipped.data(), reg.data(), W, H, D, &ue, &br);
            std::printf("3D undersegmentation error: %.6f\n3D boundary recall: %.6f\n", ue, br);
        }
        return 0;
    }
    if (dataset.empty() || video.empty()) { usage(); return 2; }
    return r.run(dataset, save, db, video, K, compactness) == 0 ? 0 : 1;
}
