// bmfr_frame_loop: the reference's frame loop (opencl/bmfr.cpp:417-485) as a
// C++ client of the C ABI, the integration a maintainer of the reference's
// host program would write.  Per frame: the four input uploads
// (bmfr.cpp:420-427), the five stages with the reference's arguments
// (camera_matrices[max(frame - 1, 0)], pixel_offsets[frame], the frame
// number; bmfr.cpp:429-476), the result readback (bmfr.cpp:478-480) and the
// double-buffer swap (bmfr.cpp:482-484); then the reference's "Total time in
// all kernels" over frames 1..N-1 (bmfr.cpp:488-517).  The inputs come from
// the library's procedural scene standing in for the dataset's EXR files.
//
//   bmfr_frame_loop WIDTH HEIGHT FRAMES OUT.f32 [SEED] [FUSED]
// writes the last frame's result (W*H*3 float32) to OUT.f32 and prints one
// line of timing.  FUSED (0, 1, 2) runs bmfr_run_frame at that fuse level
// instead of the five separate calls.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bmfr_hip.h"

#define CHECK(expr)                                                                               \
    do {                                                                                          \
        const int32_t st_ = (expr);                                                               \
        if (st_ != BMFR_OK) {                                                                     \
            std::fprintf(stderr, "%s failed: %s (%s)\n", #expr, bmfr_status_string(st_),          \
                         bmfr_last_error());                                                      \
            return 1;                                                                             \
        }                                                                                         \
    } while (0)

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s WIDTH HEIGHT FRAMES OUT.f32 [SEED] [FUSED]\n", argv[0]);
        return 2;
    }
    const int w = std::atoi(argv[1]), h = std::atoi(argv[2]), frames = std::atoi(argv[3]);
    const unsigned long long seed = argc > 5 ? std::strtoull(argv[5], nullptr, 0) : 0xB3F40000ull;
    const int fused = argc > 6 ? std::atoi(argv[6]) : -1;
    if (frames < 1) return 2;

    bmfr_params p;
    CHECK(bmfr_default_params(w, h, &p));   // bmfr.cpp:56-118
    bmfr_stream s;
    CHECK(bmfr_stream_open(&p, &s));        // bmfr.cpp:179-384
    bmfr_scene sc;
    CHECK(bmfr_scene_create(seed, 8, 1e-3f, w, h, 0, &sc));
    const size_t n = (size_t)w * h * 3;
    std::vector<float> albedo(n), normals(n), positions(n), noisy(n), out(n);
    double total_ms = 0.0;
    CHECK(bmfr_set_timing(s, 1));
    if (fused >= 0) CHECK(bmfr_set_fused(s, fused));
    for (int frame = 0; frame < frames; ++frame) {
        CHECK(bmfr_scene_render_host(sc, frame, albedo.data(), normals.data(), positions.data(), noisy.data()));
        CHECK(bmfr_upload_frame(s, albedo.data(), normals.data(), positions.data(), noisy.data()));
        float matrix[16], offset[2], unused_matrix[16], unused_offset[2];
        const int matrix_index = frame == 0 ? 0 : frame - 1;   // bmfr.cpp:440
        CHECK(bmfr_scene_camera(sc, matrix_index, matrix, unused_offset));
        CHECK(bmfr_scene_camera(sc, frame, unused_matrix, offset));
        float stage_ms[6];
        if (fused >= 0) {
            CHECK(bmfr_run_frame(s, frame, matrix, offset));
        } else {
            CHECK(bmfr_accumulate_noisy_data(s, frame, matrix, offset));
            CHECK(bmfr_fitter(s, frame));
            CHECK(bmfr_weighted_sum(s, frame));
            CHECK(bmfr_accumulate_filtered_data(s, frame));
            CHECK(bmfr_taa(s, frame));
            CHECK(bmfr_end_frame(s));
        }
        CHECK(bmfr_download_result(s, out.data()));
        CHECK(bmfr_stage_times(s, stage_ms));
        if (frame > 0) total_ms += stage_ms[5];
    }
    FILE *f = std::fopen(argv[4], "wb");
    if (!f || std::fwrite(out.data(), sizeof(float), n, f) != n) {
        std::fprintf(stderr, "cannot write %s\n", argv[4]);
        return 1;
    }
    std::fclose(f);
    std::printf("frames %d %dx%d: total time in all kernels %.4f ms/frame (frames 1..%d)\n", frames, w, h,
                frames > 1 ? total_ms / (frames - 1) : 0.0, frames - 1);
    CHECK(bmfr_scene_destroy(sc));
    CHECK(bmfr_stream_close(s));
    return 0;
}
