// Interface between the C-ABI host layer and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bmfr {

constexpr int kMaxFeatures = 16;         // longest configurable feature set
constexpr int kMaxColumns = kMaxFeatures + 3;

// A configurable feature set (opencl/bmfr.cpp:63-77): feature i of a pixel
// is factor[i][0] * factor[i][1] over {1, normal.xyz, world_position.xyz}
// (factor ids 0..6); features [not_scaled, count) are scaled per block.
struct FeatureSet {
    int32_t count;
    int32_t not_scaled;
    uint8_t factor[kMaxFeatures][2];
};

// Per-launch scalars of one stream (the reference's -D values plus the
// per-frame kernel arguments of opencl/bmfr.cpp:429-476).
struct StageArgs {
    int32_t width, height;
    int32_t margin_width, margin_height;
    int32_t blocks_x, blocks;
    float position_limit_squared, normal_limit_squared;
    float blend_alpha, second_blend_alpha, taa_blend_alpha;
    double noise_amount;
    int32_t frame;
    int32_t storage_f16;   // G-buffers and histories stored as half
    int32_t tmp_f16;       // tmp_data stored as half (reference default)
    float matrix[16];      // previous frame's view-projection, row-vector
    float pixel_offset_x, pixel_offset_y;
    int32_t qr_mode;       // bmfr_qr_mode
    float refactor_threshold;
    FeatureSet features;   // read by the kernels of a configured feature set only
    int32_t surface_format;   // bmfr_surface_format of StageBuffers::surface
    int64_t surface_pitch;    // its row pitch in bytes
};

// Device pointers of one stream's buffer set for the current frame (the
// double buffers already resolved to current/previous).
struct StageBuffers {
    void *albedo;
    void *normals_cur, *normals_prev;
    void *positions_cur, *positions_prev;
    void *noisy_in, *noisy_acc_cur, *noisy_acc_prev;
    void *spp_cur, *spp_prev;
    void *prev_pixel, *accept;
    void *tmp_data, *weights, *mins_maxs;
    void *filtered, *tone_mapped;
    void *out_cur, *out_prev;
    void *result_cur, *result_prev;
    void *noise;          // float[(N-1)*1024]: the fitter's feature noise of one frame
    void *flags;          // int32[blocks]: FP64 re-factorisation flags
    void *surface;        // caller's output surface (null: none)
};

// Fitter launch shape, chosen per stream (or by the autotuner).
struct FitterConfig {
    int32_t rows_per_lane;    // 4, 8 or 16: 1024 rows over 256, 128 or 64 lanes
    int32_t blocks_per_group; // 32x32 blocks handled by one work-group
};

// ---- batched launches: many streams per kernel -------------------------
// Device-resident description of one stream in a batch: its constants and
// both halves of every double buffer.
struct BatchStream {
    StageArgs args;  // constants; frame/matrix/offset come from BatchStep
    void *prev_pixel, *accept, *tmp_data, *weights, *mins_maxs;
    void *filtered, *tone_mapped, *noise, *flags;
    void *noisy_acc[2], *spp[2], *out[2], *result[2];
};

// Per-step parameters of one stream (written by the host before each step):
// the frame's arguments, the double-buffer parity, and the frame's input
// G-buffers (the stream's own or caller-bound ones) and output surface.
struct BatchStep {
    int32_t frame;
    int32_t cur;  // index of the current half of the double buffers
    float matrix[16];
    float pixel_offset_x, pixel_offset_y;
    const void *albedo, *normals_cur, *normals_prev, *positions_cur, *positions_prev, *noisy_in;
    void *surface;
    int32_t surface_format;
    int64_t surface_pitch;
};

enum { kStartsPadded = 0, kStartsImage = 1, kStartsBlocks = 2, kStartKinds = 3 };
constexpr int kMaxBatch = 64;

// Kernel argument of the batched launches.  starts[k][s] is the first work
// unit of stream s for unit kind k (padded-grid 32x8 tiles, image 64x4
// tiles, 32x32 blocks); starts[k][n] is the total.
struct BatchLaunch {
    const BatchStream *streams;
    const BatchStep *steps;
    const int32_t *starts;  // [kStartKinds][kMaxBatch + 1]
    int32_t n;
};

hipError_t launch_batch_accumulate_noisy_data(const BatchLaunch &L, int total, bool g16, bool t16,
                                              hipStream_t s);
hipError_t launch_batch_fitter(const BatchLaunch &L, int total_blocks, bool t16,
                               const FitterConfig &cfg, int grid_cap, int32_t qr_mode, hipStream_t s);
// FP64 least-squares weights of every block (diagnostic: error against FP64).
hipError_t launch_fitter_fp64_all(const StageArgs &a, const StageBuffers &b, float *out, hipStream_t s);
hipError_t launch_batch_weighted_sum(const BatchLaunch &L, int total, bool g16, hipStream_t s);
hipError_t launch_batch_accumulate_filtered_data(const BatchLaunch &L, int total, bool g16,
                                                 hipStream_t s);
hipError_t launch_batch_taa(const BatchLaunch &L, int total, bool g16, hipStream_t s);
// stages 3+4+5 fused, one work-group per 32x32 block (total = block count)
hipError_t launch_batch_filter_taa(const BatchLaunch &L, int total, bool g16, hipStream_t s);

// Up to four independent byte copies in one launch (16-byte aligned bases).
struct CopyJob {
    const void *src[4];
    void *dst[4];
    int64_t bytes[4];
};
hipError_t launch_copy4(const CopyJob &job, int count, hipStream_t s);
// LDS read bandwidth probe: `groups` work-groups of 256, `steps` x 64 bytes
// read per work-item; out holds groups * 256 uint4.
hipError_t launch_lds_read(uint4 *out, int groups, int steps, hipStream_t s);

// write_noise: stage 1 also fills the frame's noise table (spare threads).
hipError_t launch_accumulate_noisy_data(const StageArgs &a, const StageBuffers &b, bool write_noise,
                                        hipStream_t s);
hipError_t launch_noise_table(const StageArgs &a, const StageBuffers &b, hipStream_t s);
hipError_t launch_fitter(const StageArgs &a, const StageBuffers &b, const FitterConfig &cfg,
                         int32_t qr_mode, hipStream_t s);
hipError_t launch_weighted_sum(const StageArgs &a, const StageBuffers &b, hipStream_t s);
hipError_t launch_accumulate_filtered_data(const StageArgs &a, const StageBuffers &b, hipStream_t s);
hipError_t launch_taa(const StageArgs &a, const StageBuffers &b, hipStream_t s);
hipError_t launch_filter_taa(const StageArgs &a, const StageBuffers &b, hipStream_t s);
// Stages 3 and 4 in one kernel (the filtered colour is not written).
hipError_t launch_weighted_sum_filtered(const StageArgs &a, const StageBuffers &b, hipStream_t s);
hipError_t launch_batch_weighted_sum_filtered(const BatchLaunch &L, int total, bool g16, hipStream_t s);

}  // namespace bmfr
