// pt_render — offscreen renderer driving the C ABI, the counterpart of the
// reference application's OFFSCREENRENDER main loop (src/pathtracer.cpp:3965-4093):
// load a scene JSON, dispatch `spf` samples per frame until `spp` samples
// accumulate, print progress with samples per second, save an 8-bit PPM
// (SaveRender, src/pathtracer.cpp:3491-3518).
//
//   pt_render scene.json [--spp N] [--spf K] [--path-length L] [--shot S]
//                        [--width W] [--height H] [--mode ref|ext]
//                        [--tonemap 0..3] [--out render.ppm] [--xyz accum.bin]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "pathtracer.h"

namespace {

struct Args {
    std::string scene, out = "render.ppm", xyz;
    int spp = 1000, spf = 1, path_length = 5, shot = 1, width = 1280, height = 720, tonemap = 3;
    pt_mode mode = PT_MODE_REF;
};

[[noreturn]] void usage() {
    std::fprintf(stderr,
                 "usage: pt_render scene.json [--spp N] [--spf K] [--path-length L] [--shot S] [--width W]\n"
                 "                 [--height H] [--mode ref|ext] [--tonemap 0..3] [--out file.ppm] [--xyz file.bin]\n");
    std::exit(2);
}

Args parse(int argc, char** argv) {
    Args a;
    for (int i = 1; i < argc; ++i) {
        std::string k = argv[i];
        auto val = [&]() -> const char* {
            if (i + 1 >= argc) usage();
            return argv[++i];
        };
        if (k == "--spp") a.spp = std::atoi(val());
        else if (k == "--spf") a.spf = std::atoi(val());
        else if (k == "--path-length") a.path_length = std::atoi(val());
        else if (k == "--shot") a.shot = std::atoi(val());
        else if (k == "--width") a.width = std::atoi(val());
        else if (k == "--height") a.height = std::atoi(val());
        else if (k == "--tonemap") a.tonemap = std::atoi(val());
        else if (k == "--out") a.out = val();
        else if (k == "--xyz") a.xyz = val();
        else if (k == "--mode") {
            std::string m = val();
            if (m == "ref") a.mode = PT_MODE_REF;
            else if (m == "ext") a.mode = PT_MODE_EXT;
            else usage();
        } else if (!k.empty() && k[0] != '-' && a.scene.empty()) a.scene = k;
        else usage();
    }
    if (a.scene.empty() || a.spp <= 0 || a.spf <= 0 || a.width <= 0 || a.height <= 0) usage();
    return a;
}

int fail(const char* what, pt_ctx* ctx) {
    std::fprintf(stderr, "%s: %s\n", what, ctx ? pt_last_error(ctx) : pt_last_error_global());
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    Args a = parse(argc, argv);
    std::ifstream f(a.scene, std::ios::binary);
    if (!f) {
        std::fprintf(stderr, "cannot open scene %s\n", a.scene.c_str());
        return 1;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string json = ss.str();

    pt_ctx* ctx = nullptr;
    if (pt_create(&ctx, 0, nullptr) != PT_OK) return fail("pt_create", nullptr);
    pt_scene* scene = nullptr;
    if (pt_scene_from_json(ctx, json.data(), json.size(), &scene) != PT_OK) return fail("scene", ctx);
    pt_accum* acc = nullptr;
    if (pt_accum_create(ctx, a.width, a.height, &acc) != PT_OK) return fail("accumulator", ctx);

    // Push constants with the application's defaults (src/pathtracer.cpp:1163-1174).
    pt_push_constants pc;
    std::memset(&pc, 0, sizeof pc);
    pc.resolution[0] = a.width;
    pc.resolution[1] = a.height;
    pc.samplesPerFrame = a.spf;
    pc.FPS = 1.0f / 0.0166f;
    pc.persistence = 0.0625f;
    pc.pathLength = a.path_length;
    pc.tonemap = a.tonemap;
    if (pt_scene_camera(scene, a.shot, &pc) != PT_OK) return fail("camera shot", nullptr);

    // MainLoop bookkeeping: frame counts every sample dispatched, currentSamples
    // the samples in the running mean (src/pathtracer.cpp:4042-4048).
    int frame = 0, current = 0;
    auto t0 = std::chrono::steady_clock::now();
    auto prev = t0;
    while (current < a.spp) {
        frame += a.spf;
        current += a.spf;
        pc.frame = frame;
        pc.currentSamples = current;
        if (pt_dispatch(ctx, scene, acc, &pc, a.mode) != PT_OK) return fail("dispatch", ctx);
        if (pt_sync(ctx) != PT_OK) return fail("sync", ctx);
        auto now = std::chrono::steady_clock::now();
        double dt = std::chrono::duration<double>(now - prev).count();
        double elapsed = std::chrono::duration<double>(now - t0).count();
        prev = now;
        double progress = (double)current / a.spp;
        std::printf("Rendering: %3d%% %d/%d [%0.1fs|%0.1fs, %0.3fSPP/s]\r", (int)(100.0 * progress), current, a.spp,
                    elapsed, elapsed * (1.0 / progress - 1.0), a.spf / dt);
        std::fflush(stdout);
    }
    double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    pt_counters c;
    pt_counters_get(ctx, &c);
    std::printf("\nRendering Completed In %0.3fs. %llu segments (%0.3f G segments/s), %llu shadow rays.\n", total,
                (unsigned long long)c.segments, c.segments / total * 1e-9, (unsigned long long)c.shadow_segments);

    std::vector<uint8_t> rgb((size_t)a.width * a.height * 3);
    if (pt_resolve_srgb8(ctx, acc, a.tonemap, rgb.data(), rgb.size()) != PT_OK) return fail("resolve", ctx);
    std::ofstream out(a.out, std::ios::binary);
    out << "P6\n" << a.width << "\n" << a.height << "\n255\n";
    out.write(reinterpret_cast<const char*>(rgb.data()), (std::streamsize)rgb.size());
    if (!out) {
        std::fprintf(stderr, "could not write %s\n", a.out.c_str());
        return 1;
    }
    if (!a.xyz.empty()) {
        // The FP32 accumulator plus the sample counters: enough to resume,
        // since the random numbers are keyed on the sample index.
        std::vector<float> tex((size_t)a.width * a.height * 4);
        if (pt_accum_read_xyz(acc, tex.data(), tex.size()) != PT_OK) return fail("read", ctx);
        std::ofstream x(a.xyz, std::ios::binary);
        int hdr[4] = {a.width, a.height, frame, current};
        x.write(reinterpret_cast<const char*>(hdr), sizeof hdr);
        x.write(reinterpret_cast<const char*>(tex.data()), (std::streamsize)(tex.size() * 4));
    }
    std::printf("Saved %s\n", a.out.c_str());
    pt_accum_release(acc);
    pt_scene_release(scene);
    pt_destroy(ctx);
    return 0;
}
