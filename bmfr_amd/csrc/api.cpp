// C-ABI host layer of the gfx950 BMFR library (include/bmfr_hip.h).
//
// A bmfr_stream owns one render stream's device buffers (the reference's
// cl::Buffer set, opencl/bmfr.cpp:315-347, sized at run time from W and H)
// in a single allocation, its own HIP stream (the reference's in-order
// queue, bmfr.cpp:191) and the double-buffer parity (Double_buffer,
// bmfr.cpp:120-135).  Stage calls resolve the current/previous buffers and
// launch the matching kernel.  Nothing here throws or exits: failures come
// back as bmfr_status codes with a thread-local message.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "bmfr_hip.h"
#include "host_error.h"
#include "kernels.h"
#include "scene.h"
#include "scene_launch.h"
#include "surface.h"

namespace bmfr_host {

namespace {
thread_local std::string g_last_error;
}

int32_t fail(int32_t code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

const char *last_error() { return g_last_error.c_str(); }

}  // namespace bmfr_host

namespace {

using bmfr_host::fail;

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(BMFR_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));       \
    } while (0)

constexpr int kStages = 5;
constexpr int kMinImageSize = 47;
constexpr int kMarkers = 8;

// Physical buffers; double buffers take two slots.
enum Slot {
    S_ALBEDO,
    S_NORMALS_A, S_NORMALS_B,
    S_POSITIONS_A, S_POSITIONS_B,
    S_NOISY_IN,
    S_NOISY_ACC_A, S_NOISY_ACC_B,
    S_SPP_A, S_SPP_B,
    S_PREV_PIXEL, S_ACCEPT, S_TMP, S_WEIGHTS, S_MINS_MAXS, S_FILTERED, S_TONE,
    S_OUT_A, S_OUT_B,
    S_RESULT_A, S_RESULT_B,
    S_NOISE,   // the fitter's per-frame feature-noise table (not a reference buffer)
    S_FLAGS,   // per-block FP64 re-factorisation flags
    S_W64,     // FP64 least-squares weights (diagnostic)
    S_COUNT
};

int32_t validate(const bmfr_params *p) {
    if (!p) return fail(BMFR_ERR_INVALID_ARGUMENT, "params is null");
    // Stage 1 mirrors padded-grid coordinates up to workset + 29 back into
    // the image (opencl/bmfr.cl:209-216, 312-316); that stays inside the
    // image only when workset + 30 <= 2 * size, i.e. size >= 47.  (The
    // reference's comment says "bigger than BLOCK_EDGE_LENGTH"; for sizes
    // 33..46 it reads outside the image.)
    if (p->width < kMinImageSize || p->height < kMinImageSize || p->width > 16384 ||
        p->height > 16384)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "image size %dx%d outside %d..16384", p->width,
                    p->height, kMinImageSize);
    if (p->storage != BMFR_STORAGE_F32 && p->storage != BMFR_STORAGE_F16)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "unknown storage %d", p->storage);
    if (p->tmp_storage != BMFR_STORAGE_F32 && p->tmp_storage != BMFR_STORAGE_F16)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "unknown tmp_storage %d", p->tmp_storage);
    if (p->qr_mode < BMFR_QR_FP32 || p->qr_mode > BMFR_QR_KAHAN_FP64)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "unknown qr_mode %d", p->qr_mode);
    if (!(p->blend_alpha >= 0.f && p->blend_alpha <= 1.f) ||
        !(p->second_blend_alpha >= 0.f && p->second_blend_alpha <= 1.f) ||
        !(p->taa_blend_alpha >= 0.f && p->taa_blend_alpha <= 1.f))
        return fail(BMFR_ERR_INVALID_ARGUMENT, "blend factors must lie in [0, 1]");
    if (p->feature_count < 0 || p->feature_count > BMFR_MAX_FEATURES)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "feature_count %d outside 0..%d", p->feature_count,
                    BMFR_MAX_FEATURES);
    if (p->feature_count > 0) {
        if (p->features_not_scaled < 0 || p->features_not_scaled > p->feature_count)
            return fail(BMFR_ERR_INVALID_ARGUMENT, "features_not_scaled %d outside 0..%d", p->features_not_scaled,
                        p->feature_count);
        for (int i = 0; i < p->feature_count; ++i)
            for (int k = 0; k < 2; ++k)
                if (p->feature_factors[i][k] > BMFR_FACTOR_POSITION_Z)
                    return fail(BMFR_ERR_INVALID_ARGUMENT, "feature %d: unknown factor %d", i,
                                (int)p->feature_factors[i][k]);
        if (p->qr_mode != BMFR_QR_FP32)
            return fail(BMFR_ERR_INVALID_ARGUMENT, "a configured feature set runs with qr_mode BMFR_QR_FP32");
    }
    return BMFR_OK;
}

// Features N and scaled features S of params (the reference's set: 10, 6).
int features_of(const bmfr_params &p) { return p.feature_count ? p.feature_count : 10; }
int scaled_of(const bmfr_params &p) { return p.feature_count ? p.feature_count - p.features_not_scaled : 6; }

void sizes_of(const bmfr_params &p, bmfr_sizes &s, size_t *slot_bytes) {
    // opencl/bmfr.cpp:107-118
    s.workset_width = 32 * ((p.width + 31) / 32);
    s.workset_height = 32 * ((p.height + 31) / 32);
    s.margin_width = s.workset_width + 32;
    s.margin_height = s.workset_height + 32;
    s.blocks_x = s.margin_width / 32;
    s.blocks_y = s.margin_height / 32;
    s.blocks = s.blocks_x * s.blocks_y;
    s.pixels = (int64_t)p.width * p.height;
    s.padded_pixels = (int64_t)s.margin_width * s.margin_height;
    const size_t P = (size_t)s.pixels;
    const size_t g = p.storage == BMFR_STORAGE_F16 ? 2 : 4;
    const size_t t = p.tmp_storage == BMFR_STORAGE_F16 ? 2 : 4;
    size_t b[S_COUNT];
    b[S_ALBEDO] = 12 * P;
    b[S_NORMALS_A] = b[S_NORMALS_B] = 3 * g * P;
    b[S_POSITIONS_A] = b[S_POSITIONS_B] = 3 * g * P;
    b[S_NOISY_IN] = 3 * g * P;
    b[S_NOISY_ACC_A] = b[S_NOISY_ACC_B] = 3 * g * P;
    b[S_SPP_A] = b[S_SPP_B] = P;
    b[S_PREV_PIXEL] = 8 * P;
    b[S_ACCEPT] = P;
    const size_t nf = (size_t)features_of(p), ns = (size_t)scaled_of(p);
    b[S_TMP] = (nf + 3) * t * (size_t)s.padded_pixels;
    b[S_WEIGHTS] = 3 * nf * 4 * (size_t)s.blocks;
    b[S_MINS_MAXS] = (ns ? 2 * ns : 1) * 4 * (size_t)s.blocks;
    b[S_FILTERED] = 12 * P;
    b[S_TONE] = 12 * P;
    b[S_OUT_A] = b[S_OUT_B] = 3 * g * P;
    b[S_RESULT_A] = b[S_RESULT_B] = 3 * g * P;
    b[S_NOISE] = (nf > 1 ? nf - 1 : 1) * 1024 * sizeof(float);
    b[S_FLAGS] = 4 * (size_t)s.blocks;
    b[S_W64] = 3 * nf * 4 * (size_t)s.blocks;
    size_t total = 0;
    for (int i = 0; i < S_COUNT; ++i) {
        if (slot_bytes) slot_bytes[i] = b[i];
        total += (b[i] + 255) & ~(size_t)255;
    }
    s.device_bytes = (int64_t)total;
}

// A frame submitted to a server's request loop, waiting for its step.
struct ServerFrame {
    bmfr_frame_request req;
    hipEvent_t producer = nullptr;   // recorded on req.producer_stream at submit (or null)
};

}  // namespace

struct bmfr_stream_s {
    bmfr_params p{};
    bmfr_sizes sz{};
    int device = 0;
    hipStream_t stream = nullptr;
    char *base = nullptr;
    void *slot[S_COUNT]{};
    size_t slot_bytes[S_COUNT]{};
    bool swapped = false;   // Double_buffer::swapped
    int noise_frame = -1;   // frame whose noise table S_NOISE holds
    int last_result = -1;   // slot holding the latest TAA output
    int timing = 0;         // 0 off, 1 per stage, 2 first start / last end only
    hipEvent_t ev_start[kStages]{}, ev_end[kStages]{};
    hipEvent_t markers[kMarkers]{};
    hipEvent_t sync_ev = nullptr;   // orders this stream against batch launches
    bool recorded[kStages]{};
    // one wave per block, four blocks per work-group: a 720p frame's 984
    // blocks are 246 work-groups, one per CU with one wave on each SIMD.
    // Measured in the 720p chain (kernel trace): 20.8 us against 28.4 us with
    // one or two blocks per work-group, whose single waves the dispatcher
    // stacks two to a SIMD while the previous stage drains (the kernel has
    // room for two waves per SIMD).  Every 16-row shape gives the same bits.
    bmfr::FitterConfig fitter{16, 4};
    int fuse = BMFR_FUSE_345;  // how bmfr_run_frame runs stages 3-5 (bmfr_fuse_level)
    // Renderer handoff: the current frame's caller-bound inputs (albedo,
    // normals, positions, noisy; null: the stream's own buffers), the
    // previous frame's normals and positions when they were caller-bound
    // (null: the stream's own previous buffers), and the output surface.
    const void *bind_in[4]{};
    const void *prev_in[2]{};
    void *surface = nullptr;
    int32_t surface_format = 0;
    int64_t surface_pitch = 0;
    // Ownership: batches the stream is a member of (0 or 1), the server
    // that owns it.  Guarded by the owning server's mutex: the request loop's
    // queue of submitted frames and the frames in launched steps.
    int batch_refs = 0;
    bmfr_server owner = nullptr;
    std::deque<ServerFrame> queue;
    int frames_in_flight = 0;
    bool closing = false;
    int64_t last_step_id = 0;   // the last server step the stream took part in
    int last_lane = 0;          // and the server lane it ran on
    // Ordering between server steps and the stream's own HIP stream, kept
    // lazily: ordered_step is the last step the HIP stream waits for (an API
    // call on the stream, or closing it, adds the wait for later steps);
    // user_work: work may have been enqueued on the HIP stream since the
    // last step that was ordered after it; external: the caller holds the
    // HIP stream (bmfr_stream_hip_stream), so every step is ordered both ways.
    int64_t ordered_step = 0;
    bool user_work = true;
    bool external = false;

    int cur(int a_slot) const { return swapped ? a_slot : a_slot + 1; }
    int prev(int a_slot) const { return swapped ? a_slot + 1 : a_slot; }

    int resolve(int id) const {
        switch (id) {
            case BMFR_BUF_ALBEDO: return S_ALBEDO;
            case BMFR_BUF_NORMALS_CUR: return cur(S_NORMALS_A);
            case BMFR_BUF_NORMALS_PREV: return prev(S_NORMALS_A);
            case BMFR_BUF_POSITIONS_CUR: return cur(S_POSITIONS_A);
            case BMFR_BUF_POSITIONS_PREV: return prev(S_POSITIONS_A);
            case BMFR_BUF_NOISY_IN: return S_NOISY_IN;
            case BMFR_BUF_NOISY_ACC_CUR: return cur(S_NOISY_ACC_A);
            case BMFR_BUF_NOISY_ACC_PREV: return prev(S_NOISY_ACC_A);
            case BMFR_BUF_SPP_CUR: return cur(S_SPP_A);
            case BMFR_BUF_SPP_PREV: return prev(S_SPP_A);
            case BMFR_BUF_PREV_PIXEL: return S_PREV_PIXEL;
            case BMFR_BUF_ACCEPT: return S_ACCEPT;
            case BMFR_BUF_TMP_DATA: return S_TMP;
            case BMFR_BUF_WEIGHTS: return S_WEIGHTS;
            case BMFR_BUF_MINS_MAXS: return S_MINS_MAXS;
            case BMFR_BUF_FILTERED: return S_FILTERED;
            case BMFR_BUF_TONE_MAPPED: return S_TONE;
            case BMFR_BUF_OUT_CUR: return cur(S_OUT_A);
            case BMFR_BUF_OUT_PREV: return prev(S_OUT_A);
            case BMFR_BUF_RESULT_CUR: return cur(S_RESULT_A);
            case BMFR_BUF_RESULT_PREV: return prev(S_RESULT_A);
            default: return -1;
        }
    }

    bmfr::StageBuffers buffers() const {
        bmfr::StageBuffers b;
        b.albedo = bind_in[0] ? const_cast<void *>(bind_in[0]) : slot[S_ALBEDO];
        b.normals_cur = bind_in[1] ? const_cast<void *>(bind_in[1]) : slot[cur(S_NORMALS_A)];
        b.normals_prev = prev_in[0] ? const_cast<void *>(prev_in[0]) : slot[prev(S_NORMALS_A)];
        b.positions_cur = bind_in[2] ? const_cast<void *>(bind_in[2]) : slot[cur(S_POSITIONS_A)];
        b.positions_prev = prev_in[1] ? const_cast<void *>(prev_in[1]) : slot[prev(S_POSITIONS_A)];
        b.noisy_in = bind_in[3] ? const_cast<void *>(bind_in[3]) : slot[S_NOISY_IN];
        b.noisy_acc_cur = slot[cur(S_NOISY_ACC_A)];
        b.noisy_acc_prev = slot[prev(S_NOISY_ACC_A)];
        b.spp_cur = slot[cur(S_SPP_A)];
        b.spp_prev = slot[prev(S_SPP_A)];
        b.prev_pixel = slot[S_PREV_PIXEL];
        b.accept = slot[S_ACCEPT];
        b.tmp_data = slot[S_TMP];
        b.weights = slot[S_WEIGHTS];
        b.mins_maxs = slot[S_MINS_MAXS];
        b.filtered = slot[S_FILTERED];
        b.tone_mapped = slot[S_TONE];
        b.out_cur = slot[cur(S_OUT_A)];
        b.out_prev = slot[prev(S_OUT_A)];
        b.result_cur = slot[cur(S_RESULT_A)];
        b.result_prev = slot[prev(S_RESULT_A)];
        b.noise = slot[S_NOISE];
        b.flags = slot[S_FLAGS];
        b.surface = surface;
        return b;
    }

    // End of a frame (opencl/bmfr.cpp:482-484): the double buffers swap, and
    // this frame's current normals and positions become the next frame's
    // previous ones (caller-bound or the stream's own); a binding of inputs
    // covers one frame.
    void finish_frame() {
        prev_in[0] = bind_in[1];
        prev_in[1] = bind_in[2];
        for (auto &p : bind_in) p = nullptr;
        swapped = !swapped;
    }

    // The batched form of the frame: constants and both halves of the
    // double buffers (BatchStream), and the step's parameters with this
    // frame's resolved inputs and surface (BatchStep).
    void batch_stream(bmfr::BatchStream &d) const {
        d.args = args(0);
        d.prev_pixel = slot[S_PREV_PIXEL];
        d.accept = slot[S_ACCEPT];
        d.tmp_data = slot[S_TMP];
        d.weights = slot[S_WEIGHTS];
        d.mins_maxs = slot[S_MINS_MAXS];
        d.filtered = slot[S_FILTERED];
        d.tone_mapped = slot[S_TONE];
        d.noise = slot[S_NOISE];
        d.flags = slot[S_FLAGS];
        for (int h = 0; h < 2; ++h) {
            d.noisy_acc[h] = slot[S_NOISY_ACC_A + h];
            d.spp[h] = slot[S_SPP_A + h];
            d.out[h] = slot[S_OUT_A + h];
            d.result[h] = slot[S_RESULT_A + h];
        }
    }
    void batch_step(int frame, const float *matrix, const float *offset, void *frame_surface, int32_t frame_format,
                    int64_t frame_pitch, bmfr::BatchStep &st) const {
        st.frame = frame;
        st.cur = swapped ? 0 : 1;  // Double_buffer::current
        std::memcpy(st.matrix, matrix, sizeof(st.matrix));
        st.pixel_offset_x = offset[0];
        st.pixel_offset_y = offset[1];
        const bmfr::StageBuffers b = buffers();
        st.albedo = b.albedo;
        st.normals_cur = b.normals_cur;
        st.normals_prev = b.normals_prev;
        st.positions_cur = b.positions_cur;
        st.positions_prev = b.positions_prev;
        st.noisy_in = b.noisy_in;
        st.surface = frame_surface ? frame_surface : surface;
        st.surface_format = frame_surface ? frame_format : surface_format;
        st.surface_pitch = frame_surface ? frame_pitch : surface_pitch;
    }

    bmfr::StageArgs args(int frame) const {
        bmfr::StageArgs a{};
        a.width = p.width;
        a.height = p.height;
        a.margin_width = sz.margin_width;
        a.margin_height = sz.margin_height;
        a.blocks_x = sz.blocks_x;
        a.blocks = sz.blocks;
        a.position_limit_squared = p.position_limit_squared;
        a.normal_limit_squared = p.normal_limit_squared;
        a.blend_alpha = p.blend_alpha;
        a.second_blend_alpha = p.second_blend_alpha;
        a.taa_blend_alpha = p.taa_blend_alpha;
        a.noise_amount = p.noise_amount;
        a.frame = frame;
        a.storage_f16 = p.storage == BMFR_STORAGE_F16;
        a.tmp_f16 = p.tmp_storage == BMFR_STORAGE_F16;
        a.qr_mode = p.qr_mode;
        a.refactor_threshold = p.refactor_threshold;
        if (p.feature_count) {
            a.features.count = p.feature_count;
            a.features.not_scaled = p.features_not_scaled;
            memcpy(a.features.factor, p.feature_factors, sizeof(a.features.factor));
        }
        a.surface_format = surface_format;
        a.surface_pitch = surface_pitch;
        return a;
    }
};

struct bmfr_scene_s {
    bmfr::SceneData data{};
    bmfr::SceneData *device_copy = nullptr;
};

namespace {

int32_t order_after_server_steps(bmfr_stream s, bool external);

// Every stream entry point starts here.  A server-owned stream's HIP stream
// is first ordered after the server steps it took part in, and the next step
// is ordered after what the call enqueues.
int32_t check_stream(bmfr_stream s) {
    if (!s) return fail(BMFR_ERR_INVALID_ARGUMENT, "stream is null");
    if (s->owner) return order_after_server_steps(s, false);
    return BMFR_OK;
}

int32_t check_frame(int32_t frame) {
    if (frame < 0) return fail(BMFR_ERR_INVALID_ARGUMENT, "frame_number %d < 0", frame);
    return BMFR_OK;
}

// An output surface of stream s: a known format, a pitch that holds a row
// and keeps every pixel aligned, an aligned base.
int32_t check_surface(bmfr_stream s, const void *surface, int32_t format, int64_t pitch) {
    if (format < bmfr::kSurfaceRgbF32 || format > bmfr::kSurfaceRgbaU8)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "unknown surface format %d", format);
    const int64_t row = (int64_t)s->p.width * bmfr::surface_pixel_bytes(format);
    const int al = bmfr::surface_alignment(format);
    if (pitch < row || pitch % al != 0)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "surface row pitch %lld: needs >= %lld bytes and a multiple of %d",
                    (long long)pitch, (long long)row, al);
    if (((uintptr_t)surface % (uintptr_t)al) != 0)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "surface is not %d-byte aligned", al);
    return BMFR_OK;
}

void destroy_stream(bmfr_stream s);

template <typename F>
int32_t timed_stage(bmfr_stream s, int stage, F &&launch) {
    HIP_TRY(hipSetDevice(s->device));
    // level 1: events around every stage; level 2: only before the first
    // and after the last stage, so that no event sits between kernels
    const bool rec_start = s->timing == 1 || (s->timing == 2 && stage == 0);
    const bool rec_end = s->timing == 1 || (s->timing == 2 && stage == kStages - 1);
    if (rec_start) HIP_TRY(hipEventRecord(s->ev_start[stage], s->stream));
    hipError_t e = launch();
    if (e != hipSuccess) return fail(BMFR_ERR_HIP, "stage %d launch failed: %s", stage, hipGetErrorString(e));
    if (rec_end) HIP_TRY(hipEventRecord(s->ev_end[stage], s->stream));
    if (s->timing) s->recorded[stage] = true;
    return BMFR_OK;
}

}  // namespace

extern "C" int32_t bmfr_batch_destroy(bmfr_batch b);
extern "C" int32_t bmfr_server_destroy(bmfr_server sv);

extern "C" {

int32_t bmfr_api_version(void) { return BMFR_API_VERSION; }

const char *bmfr_status_string(int32_t status) {
    switch (status) {
        case BMFR_OK: return "ok";
        case BMFR_ERR_INVALID_ARGUMENT: return "invalid argument";
        case BMFR_ERR_OUT_OF_BUDGET: return "out of device-memory budget";
        case BMFR_ERR_HIP: return "HIP runtime error";
        case BMFR_ERR_NO_DEVICE: return "no usable gfx950 device";
        case BMFR_ERR_STATE: return "invalid state";
        default: return "unknown status";
    }
}

const char *bmfr_last_error(void) { return bmfr_host::last_error(); }

int32_t bmfr_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int32_t bmfr_measure_lds_read_bandwidth(int32_t steps, double *bytes_per_second) {
    if (!bytes_per_second || steps <= 0)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "need steps > 0 and an output pointer");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(BMFR_ERR_NO_DEVICE, "no HIP device visible");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    const int groups = prop.multiProcessorCount * 8;   // eight 256-wide work-groups per CU
    constexpr int kReps = 5;
    uint4 *out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&out, (size_t)groups * 256 * sizeof(uint4));
    if (e == hipSuccess) e = hipStreamCreate(&s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = bmfr::launch_lds_read(out, groups, steps, s);   // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (int r = 0; r < kReps && e == hipSuccess; ++r) e = bmfr::launch_lds_read(out, groups, steps, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e1) (void)hipEventDestroy(e1);
    if (e0) (void)hipEventDestroy(e0);
    if (s) (void)hipStreamDestroy(s);
    if (out) (void)hipFree(out);
    if (e != hipSuccess) return fail(BMFR_ERR_HIP, "LDS bandwidth probe failed: %s", hipGetErrorString(e));
    *bytes_per_second = (double)groups * 256 * 64 * (double)steps * kReps / (ms * 1e-3);
    return BMFR_OK;
}

int32_t bmfr_default_params(int32_t width, int32_t height, bmfr_params *out) {
    if (!out) return fail(BMFR_ERR_INVALID_ARGUMENT, "out is null");
    // opencl/bmfr.cpp:58-62, 85-88; the position/normal limits come from the
    // dataset header in the reference (bmfr.cpp:226-227) and are chosen here.
    out->width = width;
    out->height = height;
    out->position_limit_squared = 0.01f;
    out->normal_limit_squared = 0.1f;
    out->blend_alpha = 0.2f;
    out->second_blend_alpha = 0.1f;
    out->taa_blend_alpha = 0.2f;
    out->noise_amount = 1e-2;
    out->storage = BMFR_STORAGE_F32;
    out->tmp_storage = BMFR_STORAGE_F16;
    out->qr_mode = BMFR_QR_FP32;
    out->refactor_threshold = 1e-6f;
    out->feature_count = 0;   // the reference's set, bmfr.cpp:65-77
    out->features_not_scaled = 0;
    memset(out->feature_factors, 0, sizeof(out->feature_factors));
    return validate(out);
}

}  // extern "C"

namespace {

// One term of the reference's feature strings: `1.f`, `normal.x`,
// `world_position.y`, or a product `a*b` of two of those.
bool parse_factor(const std::string &f, uint8_t &id) {
    static const char *const names[] = {"normal.x", "normal.y", "normal.z",
                                        "world_position.x", "world_position.y", "world_position.z"};
    if (f == "1" || f == "1.f" || f == "1.0" || f == "1.0f") {
        id = BMFR_FACTOR_ONE;
        return true;
    }
    for (int i = 0; i < 6; ++i)
        if (f == names[i]) {
            id = (uint8_t)(BMFR_FACTOR_NORMAL_X + i);
            return true;
        }
    return false;
}

// Appends the terms of one comma-separated string (spaces dropped, empty
// terms skipped) to `terms`.
int32_t parse_terms(const char *text, std::vector<std::pair<uint8_t, uint8_t>> &terms) {
    std::string cur;
    auto flush = [&]() -> int32_t {
        if (cur.empty()) return BMFR_OK;
        const size_t star = cur.find('*');
        uint8_t a = 0, b = BMFR_FACTOR_ONE;
        bool ok;
        if (star == std::string::npos) {
            ok = parse_factor(cur, a);
        } else {
            ok = cur.find('*', star + 1) == std::string::npos && parse_factor(cur.substr(0, star), a) &&
                 parse_factor(cur.substr(star + 1), b);
        }
        if (!ok)
            return fail(BMFR_ERR_INVALID_ARGUMENT,
                        "feature term '%s' is not 1.f, normal.x|y|z, world_position.x|y|z or a product of two",
                        cur.c_str());
        terms.emplace_back(a, b);
        cur.clear();
        return BMFR_OK;
    };
    for (const char *c = text; *c; ++c) {
        if (*c == ' ' || *c == '\t' || *c == '\n' || *c == '\r') continue;
        if (*c == ',') {
            if (int32_t st = flush()) return st;
        } else {
            cur.push_back(*c);
        }
    }
    return flush();
}

}  // namespace

extern "C" {

int32_t bmfr_parse_feature_buffers(const char *not_scaled, const char *scaled, bmfr_params *params) {
    if (!not_scaled || !scaled || !params) return fail(BMFR_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<std::pair<uint8_t, uint8_t>> terms;
    if (int32_t st = parse_terms(not_scaled, terms)) return st;
    const int ns = (int)terms.size();
    if (int32_t st = parse_terms(scaled, terms)) return st;
    const int n = (int)terms.size();
    if (n < 1 || n > BMFR_MAX_FEATURES)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "%d feature terms outside 1..%d", n, BMFR_MAX_FEATURES);
    // The reference's own strings (bmfr.cpp:65-77) select its set, on the
    // kernels specialised for it.
    static const uint8_t ref[10][2] = {{0, 0}, {1, 0}, {2, 0}, {3, 0}, {4, 0},
                                       {5, 0}, {6, 0}, {4, 4}, {5, 5}, {6, 6}};
    bool is_ref = n == 10 && ns == 4;
    for (int i = 0; is_ref && i < n; ++i) is_ref = terms[i].first == ref[i][0] && terms[i].second == ref[i][1];
    memset(params->feature_factors, 0, sizeof(params->feature_factors));
    if (is_ref) {
        params->feature_count = 0;
        params->features_not_scaled = 0;
        return BMFR_OK;
    }
    params->feature_count = n;
    params->features_not_scaled = ns;
    for (int i = 0; i < n; ++i) {
        params->feature_factors[i][0] = terms[i].first;
        params->feature_factors[i][1] = terms[i].second;
    }
    return BMFR_OK;
}

int32_t bmfr_compute_sizes(const bmfr_params *params, bmfr_sizes *out) {
    int32_t st = validate(params);
    if (st) return st;
    if (!out) return fail(BMFR_ERR_INVALID_ARGUMENT, "out is null");
    sizes_of(*params, *out, nullptr);
    return BMFR_OK;
}

int32_t bmfr_stream_open(const bmfr_params *params, bmfr_stream *out) {
    int32_t st = validate(params);
    if (st) return st;
    if (!out) return fail(BMFR_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(BMFR_ERR_NO_DEVICE, "no HIP device visible");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(BMFR_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950", dev,
                    prop.gcnArchName);
    bmfr_stream s = new (std::nothrow) bmfr_stream_s;
    if (!s) return fail(BMFR_ERR_INVALID_ARGUMENT, "host allocation failed");
    s->p = *params;
    s->device = dev;
    sizes_of(s->p, s->sz, s->slot_bytes);
    hipError_t e = hipMalloc((void **)&s->base, (size_t)s->sz.device_bytes);
    if (e != hipSuccess) {
        delete s;
        return fail(BMFR_ERR_OUT_OF_BUDGET, "hipMalloc of %lld bytes failed: %s",
                    (long long)s->sz.device_bytes, hipGetErrorString(e));
    }
    size_t off = 0;
    for (int i = 0; i < S_COUNT; ++i) {
        s->slot[i] = s->base + off;
        off += (s->slot_bytes[i] + 255) & ~(size_t)255;
    }
    e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemsetAsync(s->base, 0, (size_t)s->sz.device_bytes, s->stream);
    for (int i = 0; i < kStages && e == hipSuccess; ++i) {
        e = hipEventCreate(&s->ev_start[i]);
        if (e == hipSuccess) e = hipEventCreate(&s->ev_end[i]);
    }
    for (int i = 0; i < kMarkers && e == hipSuccess; ++i) e = hipEventCreate(&s->markers[i]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->sync_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        destroy_stream(s);
        return fail(BMFR_ERR_HIP, "stream set-up failed: %s", hipGetErrorString(e));
    }
    *out = s;
    return BMFR_OK;
}

int32_t bmfr_stream_close(bmfr_stream s) {
    if (!s) return BMFR_OK;
    if (s->owner) return fail(BMFR_ERR_STATE, "the stream belongs to a server: close it with bmfr_server_close_stream");
    if (s->batch_refs) return fail(BMFR_ERR_STATE, "the stream is a member of a batch: destroy the batch first");
    destroy_stream(s);
    return BMFR_OK;
}

}  // extern "C"

namespace {

void destroy_stream(bmfr_stream s) {
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (int i = 0; i < kStages; ++i) {
        if (s->ev_start[i]) (void)hipEventDestroy(s->ev_start[i]);
        if (s->ev_end[i]) (void)hipEventDestroy(s->ev_end[i]);
    }
    for (int i = 0; i < kMarkers; ++i)
        if (s->markers[i]) (void)hipEventDestroy(s->markers[i]);
    if (s->sync_ev) (void)hipEventDestroy(s->sync_ev);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    if (s->base) (void)hipFree(s->base);
    delete s;
}

}  // namespace

extern "C" {

int32_t bmfr_stream_sizes(bmfr_stream s, bmfr_sizes *out) {
    if (int32_t st = check_stream(s)) return st;
    if (!out) return fail(BMFR_ERR_INVALID_ARGUMENT, "out is null");
    *out = s->sz;
    return BMFR_OK;
}

int32_t bmfr_stream_reset(bmfr_stream s) {
    if (int32_t st = check_stream(s)) return st;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipMemsetAsync(s->base, 0, (size_t)s->sz.device_bytes, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->swapped = false;
    s->last_result = -1;
    s->noise_frame = -1;
    for (auto &p : s->bind_in) p = nullptr;
    for (auto &p : s->prev_in) p = nullptr;
    return BMFR_OK;
}

int32_t bmfr_set_fitter_config(bmfr_stream s, int32_t rows_per_lane, int32_t blocks_per_group) {
    if (int32_t st = check_stream(s)) return st;
    const int key = rows_per_lane * 10 + blocks_per_group;
    if (key != 41 && key != 42 && key != 81 && key != 82 && key != 161 && key != 162 && key != 164)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "unsupported fitter config rows_per_lane=%d blocks_per_group=%d",
                    rows_per_lane, blocks_per_group);
    s->fitter.rows_per_lane = rows_per_lane;
    s->fitter.blocks_per_group = blocks_per_group;
    return BMFR_OK;
}

int32_t bmfr_upload_frame(bmfr_stream s, const float *albedo, const float *normals,
                          const float *positions, const float *noisy) {
    if (int32_t st = check_stream(s)) return st;
    if (!albedo || !normals || !positions || !noisy)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "null input array");
    HIP_TRY(hipSetDevice(s->device));
    const size_t n = (size_t)s->sz.pixels * 3;
    HIP_TRY(hipMemcpyAsync(s->slot[S_ALBEDO], albedo, n * 4, hipMemcpyHostToDevice, s->stream));
    const float *src[3] = {normals, positions, noisy};
    const int dst[3] = {s->cur(S_NORMALS_A), s->cur(S_POSITIONS_A), S_NOISY_IN};
    if (s->p.storage == BMFR_STORAGE_F32) {
        for (int i = 0; i < 3; ++i)
            HIP_TRY(hipMemcpyAsync(s->slot[dst[i]], src[i], n * 4, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    } else {
        std::vector<__half> h(n);
        for (int i = 0; i < 3; ++i) {
            for (size_t k = 0; k < n; ++k) h[k] = __float2half_rn(src[i][k]);
            HIP_TRY(hipMemcpyAsync(s->slot[dst[i]], h.data(), n * 2, hipMemcpyHostToDevice, s->stream));
            HIP_TRY(hipStreamSynchronize(s->stream));
        }
    }
    return BMFR_OK;
}

int32_t bmfr_copy_frame_inputs_device(bmfr_stream s, const void *albedo, const void *normals,
                                      const void *positions, const void *noisy) {
    if (int32_t st = check_stream(s)) return st;
    const void *src[4] = {albedo, normals, positions, noisy};
    const int dst[4] = {S_ALBEDO, s->cur(S_NORMALS_A), s->cur(S_POSITIONS_A), S_NOISY_IN};
    bmfr::CopyJob job{};
    for (int i = 0; i < 4; ++i) {
        if (!src[i]) return fail(BMFR_ERR_INVALID_ARGUMENT, "null device input %d", i);
        if (((uintptr_t)src[i] & 15) != 0)
            return fail(BMFR_ERR_INVALID_ARGUMENT, "device input %d is not 16-byte aligned", i);
        job.src[i] = src[i];
        job.dst[i] = s->slot[dst[i]];
        job.bytes[i] = (int64_t)s->slot_bytes[dst[i]];
    }
    HIP_TRY(hipSetDevice(s->device));
    hipError_t e = bmfr::launch_copy4(job, 4, s->stream);
    if (e != hipSuccess) return fail(BMFR_ERR_HIP, "copy launch failed: %s", hipGetErrorString(e));
    return BMFR_OK;
}

int32_t bmfr_bind_device_inputs(bmfr_stream s, const void *albedo, const void *normals, const void *positions,
                                const void *noisy) {
    if (int32_t st = check_stream(s)) return st;
    const void *src[4] = {albedo, normals, positions, noisy};
    const int nulls = !albedo + !normals + !positions + !noisy;
    if (nulls != 0 && nulls != 4) return fail(BMFR_ERR_INVALID_ARGUMENT, "bind all four inputs or none");
    for (int i = 0; i < 4; ++i)
        if (((uintptr_t)src[i] & 15) != 0)
            return fail(BMFR_ERR_INVALID_ARGUMENT, "device input %d is not 16-byte aligned", i);
    for (int i = 0; i < 4; ++i) s->bind_in[i] = src[i];
    return BMFR_OK;
}

int32_t bmfr_bind_output_surface(bmfr_stream s, void *surface, int32_t format, int64_t pitch) {
    if (int32_t st = check_stream(s)) return st;
    if (surface) {
        if (int32_t st = check_surface(s, surface, format, pitch)) return st;
    }
    s->surface = surface;
    s->surface_format = surface ? format : 0;
    s->surface_pitch = surface ? pitch : 0;
    return BMFR_OK;
}

int32_t bmfr_stream_hip_stream(bmfr_stream s, void **hip_stream) {
    if (int32_t st = check_stream(s)) return st;
    if (!hip_stream) return fail(BMFR_ERR_INVALID_ARGUMENT, "hip_stream is null");
    if (s->owner) {
        if (int32_t st = order_after_server_steps(s, true)) return st;
    }
    *hip_stream = (void *)s->stream;
    return BMFR_OK;
}

int32_t bmfr_accumulate_noisy_data(bmfr_stream s, int32_t frame, const float *matrix,
                                   const float *pixel_offset) {
    if (int32_t st = check_stream(s)) return st;
    if (int32_t st = check_frame(frame)) return st;
    if (!matrix || !pixel_offset) return fail(BMFR_ERR_INVALID_ARGUMENT, "null matrix or offset");
    bmfr::StageArgs a = s->args(frame);
    std::memcpy(a.matrix, matrix, sizeof(a.matrix));
    a.pixel_offset_x = pixel_offset[0];
    a.pixel_offset_y = pixel_offset[1];
    const bmfr::StageBuffers b = s->buffers();
    int32_t st = timed_stage(s, 0, [&] { return bmfr::launch_accumulate_noisy_data(a, b, true, s->stream); });
    if (st == BMFR_OK) s->noise_frame = frame;
    return st;
}

int32_t bmfr_fitter(bmfr_stream s, int32_t frame) {
    if (int32_t st = check_stream(s)) return st;
    if (int32_t st = check_frame(frame)) return st;
    const bmfr::StageArgs a = s->args(frame);
    const bmfr::StageBuffers b = s->buffers();
    if (s->noise_frame != frame) {
        // fitter called without this frame's stage 1 (e.g. on tmp_data written
        // through bmfr_write_buffer): build the noise table first
        HIP_TRY(hipSetDevice(s->device));
        hipError_t e = bmfr::launch_noise_table(a, b, s->stream);
        if (e != hipSuccess) return fail(BMFR_ERR_HIP, "noise table launch failed: %s", hipGetErrorString(e));
        s->noise_frame = frame;
    }
    return timed_stage(s, 1, [&] { return bmfr::launch_fitter(a, b, s->fitter, s->p.qr_mode, s->stream); });
}

int32_t bmfr_fitter_fp64_weights(bmfr_stream s, float *out) {
    if (int32_t st = check_stream(s)) return st;
    if (!out) return fail(BMFR_ERR_INVALID_ARGUMENT, "out is null");
    if (s->noise_frame < 0) return fail(BMFR_ERR_STATE, "the fitter has not run yet");
    if (s->p.feature_count)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "FP64 weights are computed for the reference's feature set only");
    const bmfr::StageArgs a = s->args(s->noise_frame);
    const bmfr::StageBuffers b = s->buffers();
    HIP_TRY(hipSetDevice(s->device));
    hipError_t e = bmfr::launch_fitter_fp64_all(a, b, (float *)s->slot[S_W64], s->stream);
    if (e != hipSuccess) return fail(BMFR_ERR_HIP, "FP64 fitter launch failed: %s", hipGetErrorString(e));
    HIP_TRY(hipMemcpyAsync(out, s->slot[S_W64], s->slot_bytes[S_W64], hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return BMFR_OK;
}

int32_t bmfr_refactor_count(bmfr_stream s, int32_t *count) {
    if (int32_t st = check_stream(s)) return st;
    if (!count) return fail(BMFR_ERR_INVALID_ARGUMENT, "count is null");
    std::vector<int32_t> f((size_t)s->sz.blocks);
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipMemcpyAsync(f.data(), s->slot[S_FLAGS], f.size() * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    int32_t c = 0;
    for (int32_t v : f) c += v != 0;
    *count = c;
    return BMFR_OK;
}

int32_t bmfr_weighted_sum(bmfr_stream s, int32_t frame) {
    if (int32_t st = check_stream(s)) return st;
    if (int32_t st = check_frame(frame)) return st;
    const bmfr::StageArgs a = s->args(frame);
    const bmfr::StageBuffers b = s->buffers();
    return timed_stage(s, 2, [&] { return bmfr::launch_weighted_sum(a, b, s->stream); });
}

int32_t bmfr_accumulate_filtered_data(bmfr_stream s, int32_t frame) {
    if (int32_t st = check_stream(s)) return st;
    if (int32_t st = check_frame(frame)) return st;
    const bmfr::StageArgs a = s->args(frame);
    const bmfr::StageBuffers b = s->buffers();
    return timed_stage(s, 3, [&] { return bmfr::launch_accumulate_filtered_data(a, b, s->stream); });
}

int32_t bmfr_taa(bmfr_stream s, int32_t frame) {
    if (int32_t st = check_stream(s)) return st;
    if (int32_t st = check_frame(frame)) return st;
    const bmfr::StageArgs a = s->args(frame);
    const bmfr::StageBuffers b = s->buffers();
    int32_t st = timed_stage(s, 4, [&] { return bmfr::launch_taa(a, b, s->stream); });
    if (st == BMFR_OK) s->last_result = s->cur(S_RESULT_A);
    return st;
}

int32_t bmfr_end_frame(bmfr_stream s) {
    if (int32_t st = check_stream(s)) return st;
    s->finish_frame();  // opencl/bmfr.cpp:482-484
    return BMFR_OK;
}

int32_t bmfr_run_frame(bmfr_stream s, int32_t frame, const float *matrix, const float *pixel_offset) {
    int32_t st = bmfr_accumulate_noisy_data(s, frame, matrix, pixel_offset);
    if (!st) st = bmfr_fitter(s, frame);
    if (st) return st;
    if (s->fuse == BMFR_FUSE_NONE) {
        st = bmfr_weighted_sum(s, frame);
        if (!st) st = bmfr_accumulate_filtered_data(s, frame);
        if (!st) st = bmfr_taa(s, frame);
    } else {
        // fused stages: the events of a stage folded into a later kernel
        // bracket nothing (0 ms); stage 4's (accumulate_filtered_data) bracket
        // the fused kernel
        const bmfr::StageArgs a = s->args(frame);
        const bmfr::StageBuffers b = s->buffers();
        st = timed_stage(s, 2, [&] { return hipSuccess; });
        if (s->fuse == BMFR_FUSE_34) {
            if (!st) st = timed_stage(s, 3, [&] { return bmfr::launch_weighted_sum_filtered(a, b, s->stream); });
            if (!st) st = bmfr_taa(s, frame);
        } else {
            if (!st) st = timed_stage(s, 3, [&] { return bmfr::launch_filter_taa(a, b, s->stream); });
            if (!st) st = timed_stage(s, 4, [&] { return hipSuccess; });
            if (!st) s->last_result = s->cur(S_RESULT_A);
        }
    }
    if (!st) st = bmfr_end_frame(s);
    return st;
}

int32_t bmfr_set_fused(bmfr_stream s, int32_t level) {
    if (int32_t st = check_stream(s)) return st;
    if (level < BMFR_FUSE_NONE || level > BMFR_FUSE_345)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "fuse level %d is not one of 0, 1, 2", level);
    s->fuse = level;
    return BMFR_OK;
}

int32_t bmfr_synchronize(bmfr_stream s) {
    if (int32_t st = check_stream(s)) return st;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return BMFR_OK;
}

int32_t bmfr_download_result(bmfr_stream s, float *out) {
    if (int32_t st = check_stream(s)) return st;
    if (!out) return fail(BMFR_ERR_INVALID_ARGUMENT, "out is null");
    if (s->last_result < 0) return fail(BMFR_ERR_STATE, "no frame has completed taa yet");
    HIP_TRY(hipSetDevice(s->device));
    const size_t n = (size_t)s->sz.pixels * 3;
    if (s->p.storage == BMFR_STORAGE_F32) {
        HIP_TRY(hipMemcpyAsync(out, s->slot[s->last_result], n * 4, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    } else {
        std::vector<__half> h(n);
        HIP_TRY(hipMemcpyAsync(h.data(), s->slot[s->last_result], n * 2, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        for (size_t k = 0; k < n; ++k) out[k] = __half2float(h[k]);
    }
    return BMFR_OK;
}

int32_t bmfr_buffer_bytes(bmfr_stream s, int32_t buffer, size_t *bytes) {
    if (int32_t st = check_stream(s)) return st;
    const int sl = s->resolve(buffer);
    if (sl < 0) return fail(BMFR_ERR_INVALID_ARGUMENT, "unknown buffer %d", buffer);
    if (!bytes) return fail(BMFR_ERR_INVALID_ARGUMENT, "bytes is null");
    *bytes = s->slot_bytes[sl];
    return BMFR_OK;
}

int32_t bmfr_read_buffer(bmfr_stream s, int32_t buffer, void *host, size_t bytes) {
    if (int32_t st = check_stream(s)) return st;
    const int sl = s->resolve(buffer);
    if (sl < 0) return fail(BMFR_ERR_INVALID_ARGUMENT, "unknown buffer %d", buffer);
    if (!host || bytes != s->slot_bytes[sl])
        return fail(BMFR_ERR_INVALID_ARGUMENT, "buffer %d holds %zu bytes, %zu requested", buffer,
                    s->slot_bytes[sl], bytes);
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipMemcpyAsync(host, s->slot[sl], bytes, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return BMFR_OK;
}

int32_t bmfr_write_buffer(bmfr_stream s, int32_t buffer, const void *host, size_t bytes) {
    if (int32_t st = check_stream(s)) return st;
    const int sl = s->resolve(buffer);
    if (sl < 0) return fail(BMFR_ERR_INVALID_ARGUMENT, "unknown buffer %d", buffer);
    if (!host || bytes != s->slot_bytes[sl])
        return fail(BMFR_ERR_INVALID_ARGUMENT, "buffer %d holds %zu bytes, %zu given", buffer,
                    s->slot_bytes[sl], bytes);
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipMemcpyAsync(s->slot[sl], host, bytes, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return BMFR_OK;
}

int32_t bmfr_buffer_device_ptr(bmfr_stream s, int32_t buffer, void **dptr) {
    if (int32_t st = check_stream(s)) return st;
    const int sl = s->resolve(buffer);
    if (sl < 0) return fail(BMFR_ERR_INVALID_ARGUMENT, "unknown buffer %d", buffer);
    if (!dptr) return fail(BMFR_ERR_INVALID_ARGUMENT, "dptr is null");
    *dptr = s->slot[sl];
    return BMFR_OK;
}

int32_t bmfr_set_timing(bmfr_stream s, int32_t enable) {
    if (int32_t st = check_stream(s)) return st;
    if (enable < 0 || enable > 2) return fail(BMFR_ERR_INVALID_ARGUMENT, "timing level %d", enable);
    s->timing = enable;
    for (int i = 0; i < kStages; ++i) s->recorded[i] = false;
    return BMFR_OK;
}

int32_t bmfr_stage_times(bmfr_stream s, float *ms6) {
    if (int32_t st = check_stream(s)) return st;
    if (!ms6) return fail(BMFR_ERR_INVALID_ARGUMENT, "ms6 is null");
    for (int i = 0; i < kStages; ++i)
        if (!s->recorded[i]) return fail(BMFR_ERR_STATE, "stage %d has no timed launch", i);
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipEventSynchronize(s->ev_end[kStages - 1]));
    for (int i = 0; i < kStages; ++i) {
        if (s->timing == 2) ms6[i] = -1.f;
        else HIP_TRY(hipEventElapsedTime(&ms6[i], s->ev_start[i], s->ev_end[i]));
    }
    HIP_TRY(hipEventElapsedTime(&ms6[5], s->ev_start[0], s->ev_end[kStages - 1]));
    return BMFR_OK;
}

int32_t bmfr_record_marker(bmfr_stream s, int32_t marker) {
    if (int32_t st = check_stream(s)) return st;
    if (marker < 0 || marker >= kMarkers) return fail(BMFR_ERR_INVALID_ARGUMENT, "marker %d", marker);
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipEventRecord(s->markers[marker], s->stream));
    return BMFR_OK;
}

int32_t bmfr_marker_elapsed(bmfr_stream s, int32_t m0, int32_t m1, float *ms) {
    if (int32_t st = check_stream(s)) return st;
    if (m0 < 0 || m0 >= kMarkers || m1 < 0 || m1 >= kMarkers || !ms)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "bad marker arguments");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipEventSynchronize(s->markers[m1]));
    HIP_TRY(hipEventElapsedTime(ms, s->markers[m0], s->markers[m1]));
    return BMFR_OK;
}

}  // extern "C"

// ---- batches ---------------------------------------------------------------
namespace {

struct BatchGroup {
    bool g16 = false, t16 = false;
    int qr = 0;
    int lane = 0;                        // software-pipeline lane (0 .. lanes - 1)
    std::vector<int> members;            // indices into bmfr_batch_s::streams
    bmfr::BatchStream *d_streams = nullptr;
    bmfr::BatchStep *d_steps = nullptr;
    int32_t *d_starts = nullptr;
    int32_t starts[bmfr::kStartKinds][bmfr::kMaxBatch + 1]{};
    bmfr::BatchLaunch launch() const {
        bmfr::BatchLaunch L;
        L.streams = d_streams;
        L.steps = d_steps;
        L.starts = d_starts;
        L.n = (int32_t)members.size();
        return L;
    }
};

}  // namespace

struct bmfr_batch_s {
    std::vector<bmfr_stream> streams;
    std::vector<BatchGroup> groups;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipEvent_t ev_start[kStages]{}, ev_end[kStages]{};
    bool recorded = false;
    bool recorded_total_only = false;
    bmfr::FitterConfig fitter{16, 4};
    int grid_cap = 0;
    int fuse = BMFR_FUSE_345;
    // software-pipeline lanes: lane l runs on lane_stream[l] (lane 0 on
    // `stream`), its stage 1 starting when lane l - 1's stage 1 is done
    static constexpr int kMaxLanes = 4;
    int lanes = 2;
    hipStream_t lane_stream[kMaxLanes]{};
    hipEvent_t ev_s1[kMaxLanes]{}, ev_join[kMaxLanes]{};
    bool use_graph = false;
    hipGraphExec_t graph = nullptr;
    // pinned staging of the per-step parameters, one slot per group and ring entry
    static constexpr int kRing = 4;
    bmfr::BatchStep *h_steps = nullptr;    // [kRing][groups][kMaxBatch]
    hipEvent_t ring_ev[kRing]{};
    int ring = 0;
};

namespace {

// Stage `st` of the batch's groups in `lane` (-1: every group) on `stream`,
// one launch per group.
int32_t batch_enqueue_stage(bmfr_batch b, int st, int lane, hipStream_t stream) {
    for (auto &g : b->groups) {
        if (lane >= 0 && g.lane != lane) continue;
        const bmfr::BatchLaunch L = g.launch();
        hipError_t e = hipSuccess;
        switch (st) {
            case 0:
                e = bmfr::launch_batch_accumulate_noisy_data(L, g.starts[bmfr::kStartsPadded][L.n], g.g16, g.t16,
                                                             stream);
                break;
            case 1:
                e = bmfr::launch_batch_fitter(L, g.starts[bmfr::kStartsBlocks][L.n], g.t16, b->fitter, b->grid_cap,
                                              g.qr, stream);
                break;
            case 2:  // weighted_sum fused into stage 4 (0 ms)
                break;
            case 3:
                if (b->fuse == BMFR_FUSE_345)
                    e = bmfr::launch_batch_filter_taa(L, g.starts[bmfr::kStartsBlocks][L.n], g.g16, stream);
                else
                    e = bmfr::launch_batch_weighted_sum_filtered(L, g.starts[bmfr::kStartsPadded][L.n], g.g16,
                                                                 stream);
                break;
            default:  // taa, unless fused into stage 4 (0 ms)
                if (b->fuse != BMFR_FUSE_345)
                    e = bmfr::launch_batch_taa(L, g.starts[bmfr::kStartsImage][L.n], g.g16, stream);
                break;
        }
        if (e != hipSuccess) return fail(BMFR_ERR_HIP, "batched stage %d launch failed: %s", st, hipGetErrorString(e));
    }
    return BMFR_OK;
}

// The five stages of a step.  Timed per stage: stage by stage over every
// group on the batch's stream.  Otherwise, with L > 1 lanes, an L-deep
// software pipeline: lane 0 runs on the batch's stream, lane l on its own
// HIP stream starting when lane l - 1's stage 1 is done, so that each lane's
// fitter (VALU-bound) overlaps other lanes' HBM-streaming stages; the lanes
// join the batch's stream at the end.  The lanes' streams are disjoint.
int32_t batch_enqueue_stages(bmfr_batch b, bool timed) {
    if (timed || b->lanes == 1) {
        for (int st = 0; st < kStages; ++st) {
            if (timed) HIP_TRY(hipEventRecord(b->ev_start[st], b->stream));
            if (int32_t r = batch_enqueue_stage(b, st, -1, b->stream)) return r;
            if (timed) HIP_TRY(hipEventRecord(b->ev_end[st], b->stream));
        }
        return BMFR_OK;
    }
    const int L = b->lanes;
    for (int l = 0; l < L; ++l) {
        if (l > 0) HIP_TRY(hipStreamWaitEvent(b->lane_stream[l], b->ev_s1[l - 1], 0));
        if (int32_t r = batch_enqueue_stage(b, 0, l, b->lane_stream[l])) return r;
        if (l + 1 < L) HIP_TRY(hipEventRecord(b->ev_s1[l], b->lane_stream[l]));
    }
    for (int l = 0; l < L; ++l)
        for (int st = 1; st < kStages; ++st)
            if (int32_t r = batch_enqueue_stage(b, st, l, b->lane_stream[l])) return r;
    for (int l = 1; l < L; ++l) {
        HIP_TRY(hipEventRecord(b->ev_join[l], b->lane_stream[l]));
        HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_join[l], 0));
    }
    return BMFR_OK;
}

// Frees the groups' device tables and the pinned step staging.
void free_groups(bmfr_batch b) {
    for (auto &g : b->groups) {
        if (g.d_streams) (void)hipFree(g.d_streams);
        if (g.d_steps) (void)hipFree(g.d_steps);
        if (g.d_starts) (void)hipFree(g.d_starts);
    }
    b->groups.clear();
    if (b->h_steps) (void)hipHostFree(b->h_steps);
    b->h_steps = nullptr;
}

// Splits the members into `lanes` pipeline lanes of about equal pixel count
// (largest streams first, each to the lane with the fewest pixels) and into
// launch groups by (storage, tmp storage, QR mode, lane), with each group's
// device tables and the pinned staging of the per-step parameters.
hipError_t build_groups(bmfr_batch b, int lanes) {
    free_groups(b);
    const int count = (int)b->streams.size();
    std::vector<int> lane_of(count, 0), order(count);
    for (int i = 0; i < count; ++i) order[i] = i;
    auto px = [&](int i) { return (int64_t)b->streams[i]->p.width * b->streams[i]->p.height; };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return px(x) > px(y); });
    int64_t lane_px[bmfr_batch_s::kMaxLanes] = {};
    for (int i : order) {
        int l = 0;
        for (int k = 1; k < lanes; ++k)
            if (lane_px[k] < lane_px[l]) l = k;
        lane_of[i] = l;
        lane_px[l] += px(i);
    }
    for (int i = 0; i < count; ++i) {
        const bmfr_stream s = b->streams[i];
        const bool g16 = s->p.storage == BMFR_STORAGE_F16;
        const bool t16 = s->p.tmp_storage == BMFR_STORAGE_F16;
        const int qr = s->p.qr_mode;
        BatchGroup *grp = nullptr;
        for (auto &g : b->groups)
            if (g.g16 == g16 && g.t16 == t16 && g.qr == qr && g.lane == lane_of[i]) grp = &g;
        if (!grp) {
            b->groups.emplace_back();
            grp = &b->groups.back();
            grp->g16 = g16;
            grp->t16 = t16;
            grp->qr = qr;
            grp->lane = lane_of[i];
        }
        grp->members.push_back(i);
    }
    hipError_t e = hipHostMalloc((void **)&b->h_steps, sizeof(bmfr::BatchStep) * bmfr_batch_s::kRing *
                                                             b->groups.size() * bmfr::kMaxBatch);
    for (auto &g : b->groups) {
        if (e != hipSuccess) break;
        const int n = (int)g.members.size();
        std::vector<bmfr::BatchStream> hs(n);
        for (int k = 0; k < 3; ++k) g.starts[k][0] = 0;
        for (int m = 0; m < n; ++m) {
            bmfr_stream s = b->streams[g.members[m]];
            s->batch_stream(hs[m]);
            const bmfr_sizes &z = s->sz;
            g.starts[bmfr::kStartsPadded][m + 1] =
                g.starts[bmfr::kStartsPadded][m] + (z.margin_width / 32) * (z.margin_height / 8);
            g.starts[bmfr::kStartsImage][m + 1] =
                g.starts[bmfr::kStartsImage][m] + ((s->p.width + 63) / 64) * ((s->p.height + 3) / 4);
            g.starts[bmfr::kStartsBlocks][m + 1] = g.starts[bmfr::kStartsBlocks][m] + z.blocks;
        }
        e = hipMalloc((void **)&g.d_streams, sizeof(bmfr::BatchStream) * bmfr::kMaxBatch);
        if (e == hipSuccess) e = hipMalloc((void **)&g.d_steps, sizeof(bmfr::BatchStep) * bmfr::kMaxBatch);
        if (e == hipSuccess) e = hipMalloc((void **)&g.d_starts, sizeof(g.starts));
        if (e == hipSuccess)
            e = hipMemcpy(g.d_streams, hs.data(), sizeof(bmfr::BatchStream) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(g.d_starts, g.starts, sizeof(g.starts), hipMemcpyHostToDevice);
    }
    b->lanes = lanes;
    return e;
}

}  // namespace

extern "C" {

int32_t bmfr_batch_create(const bmfr_stream *streams, int32_t count, bmfr_batch *out) {
    if (!streams || !out) return fail(BMFR_ERR_INVALID_ARGUMENT, "null argument");
    if (count < 1 || count > bmfr::kMaxBatch)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "batch of %d streams outside 1..%d", count, bmfr::kMaxBatch);
    *out = nullptr;
    for (int i = 0; i < count; ++i) {
        if (!streams[i]) return fail(BMFR_ERR_INVALID_ARGUMENT, "stream %d is null", i);
        if (streams[i]->device != streams[0]->device)
            return fail(BMFR_ERR_INVALID_ARGUMENT, "streams of a batch must share one device");
        for (int j = 0; j < i; ++j)
            if (streams[j] == streams[i]) return fail(BMFR_ERR_INVALID_ARGUMENT, "stream %d listed twice", i);
        if (streams[i]->p.feature_count)
            return fail(BMFR_ERR_INVALID_ARGUMENT, "stream %d: batches run the reference's feature set", i);
        if (streams[i]->batch_refs)
            return fail(BMFR_ERR_STATE, "stream %d already belongs to a batch", i);
        if (streams[i]->owner)
            return fail(BMFR_ERR_STATE, "stream %d belongs to a server (server steps batch it)", i);
    }
    bmfr_batch b = new (std::nothrow) bmfr_batch_s;
    if (!b) return fail(BMFR_ERR_INVALID_ARGUMENT, "host allocation failed");
    b->streams.assign(streams, streams + count);
    for (bmfr_stream s : b->streams) ++s->batch_refs;
    b->device = streams[0]->device;
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    b->lane_stream[0] = b->stream;
    for (int l = 1; l < bmfr_batch_s::kMaxLanes && e == hipSuccess; ++l)
        e = hipStreamCreateWithFlags(&b->lane_stream[l], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
    for (int l = 0; l < bmfr_batch_s::kMaxLanes && e == hipSuccess; ++l) {
        e = hipEventCreateWithFlags(&b->ev_s1[l], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_join[l], hipEventDisableTiming);
    }
    for (int i = 0; i < kStages && e == hipSuccess; ++i) {
        e = hipEventCreate(&b->ev_start[i]);
        if (e == hipSuccess) e = hipEventCreate(&b->ev_end[i]);
    }
    for (int i = 0; i < bmfr_batch_s::kRing && e == hipSuccess; ++i)
        e = hipEventCreateWithFlags(&b->ring_ev[i], hipEventDisableTiming);
    if (e == hipSuccess) e = build_groups(b, std::min(2, count));
    if (e != hipSuccess) {
        bmfr_batch_destroy(b);
        return fail(BMFR_ERR_HIP, "batch set-up failed: %s", hipGetErrorString(e));
    }
    *out = b;
    return BMFR_OK;
}

int32_t bmfr_batch_destroy(bmfr_batch b) {
    if (!b) return BMFR_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (int l = 1; l < bmfr_batch_s::kMaxLanes; ++l)
        if (b->lane_stream[l]) (void)hipStreamSynchronize(b->lane_stream[l]);
    for (bmfr_stream s : b->streams) --s->batch_refs;
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    free_groups(b);
    for (int i = 0; i < kStages; ++i) {
        if (b->ev_start[i]) (void)hipEventDestroy(b->ev_start[i]);
        if (b->ev_end[i]) (void)hipEventDestroy(b->ev_end[i]);
    }
    for (auto &ev : b->ring_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (b->done) (void)hipEventDestroy(b->done);
    for (int l = 0; l < bmfr_batch_s::kMaxLanes; ++l) {
        if (b->ev_s1[l]) (void)hipEventDestroy(b->ev_s1[l]);
        if (b->ev_join[l]) (void)hipEventDestroy(b->ev_join[l]);
    }
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (int l = 1; l < bmfr_batch_s::kMaxLanes; ++l)
        if (b->lane_stream[l]) (void)hipStreamDestroy(b->lane_stream[l]);
    delete b;
    return BMFR_OK;
}

int32_t bmfr_batch_set_fused(bmfr_batch b, int32_t level) {
    if (!b) return fail(BMFR_ERR_INVALID_ARGUMENT, "batch is null");
    if (level != BMFR_FUSE_34 && level != BMFR_FUSE_345)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "batches run fuse level 1 or 2, not %d", level);
    b->fuse = level;
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return BMFR_OK;
}

int32_t bmfr_batch_set_lanes(bmfr_batch b, int32_t lanes) {
    if (!b) return fail(BMFR_ERR_INVALID_ARGUMENT, "batch is null");
    if (lanes < 1 || lanes > bmfr_batch_s::kMaxLanes)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "lanes %d outside 1..%d", lanes, bmfr_batch_s::kMaxLanes);
    lanes = std::min(lanes, (int)b->streams.size());
    if (lanes == b->lanes) return BMFR_OK;
    HIP_TRY(hipSetDevice(b->device));
    // the groups' tables and staging are in use until the last step ends
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    hipError_t e = build_groups(b, lanes);
    if (e != hipSuccess) return fail(BMFR_ERR_HIP, "batch regrouping failed: %s", hipGetErrorString(e));
    return BMFR_OK;
}

int32_t bmfr_batch_set_fitter_config(bmfr_batch b, int32_t rows_per_lane, int32_t blocks_per_group,
                                     int32_t grid_cap) {
    if (!b) return fail(BMFR_ERR_INVALID_ARGUMENT, "batch is null");
    const int key = rows_per_lane * 10 + blocks_per_group;
    if (key != 41 && key != 42 && key != 81 && key != 82 && key != 161 && key != 162 && key != 164)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "unsupported fitter config %d/%d", rows_per_lane,
                    blocks_per_group);
    if (grid_cap < 0) return fail(BMFR_ERR_INVALID_ARGUMENT, "grid_cap < 0");
    b->fitter = {rows_per_lane, blocks_per_group};
    b->grid_cap = grid_cap;
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return BMFR_OK;
}

int32_t bmfr_batch_set_graph(bmfr_batch b, int32_t enable) {
    if (!b) return fail(BMFR_ERR_INVALID_ARGUMENT, "batch is null");
    b->use_graph = enable != 0;
    return BMFR_OK;
}

int32_t bmfr_batch_run(bmfr_batch b, const int32_t *frames, const float *matrices, const float *offsets,
                       int32_t timed) {
    if (!b || !frames || !matrices || !offsets) return fail(BMFR_ERR_INVALID_ARGUMENT, "null argument");
    const int n = (int)b->streams.size();
    for (int i = 0; i < n; ++i)
        if (frames[i] < 0) return fail(BMFR_ERR_INVALID_ARGUMENT, "frame %d of stream %d < 0", frames[i], i);
    HIP_TRY(hipSetDevice(b->device));
    // 1. order after every member's pending work (uploads, renders)
    for (bmfr_stream s : b->streams) {
        HIP_TRY(hipEventRecord(s->sync_ev, s->stream));
        HIP_TRY(hipStreamWaitEvent(b->stream, s->sync_ev, 0));
    }
    // 2. per-step parameters through a pinned ring slot
    const int slot = b->ring;
    b->ring = (b->ring + 1) % bmfr_batch_s::kRing;
    HIP_TRY(hipEventSynchronize(b->ring_ev[slot]));
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        BatchGroup &g = b->groups[gi];
        bmfr::BatchStep *hs = b->h_steps + ((size_t)slot * b->groups.size() + gi) * bmfr::kMaxBatch;
        for (size_t m = 0; m < g.members.size(); ++m) {
            const int i = g.members[m];
            b->streams[i]->batch_step(frames[i], matrices + 16 * i, offsets + 2 * i, nullptr, 0, 0, hs[m]);
        }
        HIP_TRY(hipMemcpyAsync(g.d_steps, hs, sizeof(bmfr::BatchStep) * g.members.size(),
                               hipMemcpyHostToDevice, b->stream));
    }
    HIP_TRY(hipEventRecord(b->ring_ev[slot], b->stream));
    // 3. the five batched stages (captured once into a graph when enabled; a
    // timed replay records only the step's first start and last end)
    if (b->use_graph) {
        if (!b->graph) {
            hipGraph_t graph;
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int32_t st = batch_enqueue_stages(b, false);
            hipError_t e = hipStreamEndCapture(b->stream, &graph);
            if (st) return st;
            if (e != hipSuccess) return fail(BMFR_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return fail(BMFR_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
        }
        if (timed) HIP_TRY(hipEventRecord(b->ev_start[0], b->stream));
        HIP_TRY(hipGraphLaunch(b->graph, b->stream));
        if (timed) HIP_TRY(hipEventRecord(b->ev_end[kStages - 1], b->stream));
        b->recorded = timed != 0;
        b->recorded_total_only = true;
    } else {
        if (int32_t st = batch_enqueue_stages(b, timed != 0)) return st;
        b->recorded = timed != 0;
        b->recorded_total_only = false;
    }
    // 4. members' later work waits for the batch; swap their double buffers
    HIP_TRY(hipEventRecord(b->done, b->stream));
    for (int i = 0; i < n; ++i) {
        bmfr_stream s = b->streams[i];
        HIP_TRY(hipStreamWaitEvent(s->stream, b->done, 0));
        s->last_result = s->cur(S_RESULT_A);
        s->noise_frame = frames[i];
        s->finish_frame();
    }
    return BMFR_OK;
}

int32_t bmfr_batch_synchronize(bmfr_batch b) {
    if (!b) return fail(BMFR_ERR_INVALID_ARGUMENT, "batch is null");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return BMFR_OK;
}

int32_t bmfr_batch_stage_times(bmfr_batch b, float *ms6) {
    if (!b || !ms6) return fail(BMFR_ERR_INVALID_ARGUMENT, "null argument");
    if (!b->recorded) return fail(BMFR_ERR_STATE, "the last batch step was not timed");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipEventSynchronize(b->ev_end[kStages - 1]));
    for (int i = 0; i < kStages; ++i) {
        if (b->recorded_total_only) ms6[i] = -1.f;
        else HIP_TRY(hipEventElapsedTime(&ms6[i], b->ev_start[i], b->ev_end[i]));
    }
    HIP_TRY(hipEventElapsedTime(&ms6[5], b->ev_start[0], b->ev_end[kStages - 1]));
    return BMFR_OK;
}

int32_t bmfr_batch_autotune(bmfr_batch b, int32_t reps, char *log, size_t log_len) {
    if (!b) return fail(BMFR_ERR_INVALID_ARGUMENT, "batch is null");
    if (reps < 1) return fail(BMFR_ERR_INVALID_ARGUMENT, "reps < 1");
    HIP_TRY(hipSetDevice(b->device));
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
    struct Cand {
        int rpl, bpg, cap;
    };
    std::vector<Cand> cands;
    for (Cand c : {Cand{16, 1, 0}, Cand{16, 2, 0}, Cand{16, 4, 0}, Cand{8, 1, 0}, Cand{8, 2, 0},
                   Cand{4, 1, 0}, Cand{4, 2, 0}})
        cands.push_back(c);
    for (int mult : {4, 8}) {  // persistent grids: a few waves per SIMD
        cands.push_back({16, 1, cus * 4 * mult / 4});
        cands.push_back({16, 4, cus * mult / 4});
        cands.push_back({8, 2, cus * mult / 4});
    }
    std::string text = "fitter autotune (us per batched launch, min of reps):";
    float best = 1e30f;
    Cand bestc = cands[0];
    const bmfr::FitterConfig saved = b->fitter;
    const int saved_cap = b->grid_cap;
    for (const Cand &c : cands) {
        b->fitter = {c.rpl, c.bpg};
        b->grid_cap = c.cap;
        float tmin = 1e30f;
        for (int r = 0; r < reps + 1; ++r) {
            HIP_TRY(hipEventRecord(b->ev_start[1], b->stream));
            for (auto &g : b->groups) {
                hipError_t e = bmfr::launch_batch_fitter(g.launch(), g.starts[bmfr::kStartsBlocks][g.members.size()],
                                                         g.t16, b->fitter, b->grid_cap, g.qr, b->stream);
                if (e != hipSuccess) {
                    b->fitter = saved;
                    b->grid_cap = saved_cap;
                    return fail(BMFR_ERR_HIP, "autotune launch failed: %s", hipGetErrorString(e));
                }
            }
            HIP_TRY(hipEventRecord(b->ev_end[1], b->stream));
            HIP_TRY(hipEventSynchronize(b->ev_end[1]));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, b->ev_start[1], b->ev_end[1]));
            if (r > 0) tmin = std::min(tmin, ms);  // first run warms up
        }
        char buf[96];
        snprintf(buf, sizeof(buf), " rpl%d/bpg%d/cap%d=%.1f", c.rpl, c.bpg, c.cap, tmin * 1000.f);
        text += buf;
        if (tmin < best) {
            best = tmin;
            bestc = c;
        }
    }
    b->fitter = {bestc.rpl, bestc.bpg};
    b->grid_cap = bestc.cap;
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    b->recorded = false;
    char buf[96];
    snprintf(buf, sizeof(buf), "; chosen rpl%d/bpg%d/cap%d", bestc.rpl, bestc.bpg, bestc.cap);
    text += buf;
    if (log && log_len > 0) {
        std::strncpy(log, text.c_str(), log_len - 1);
        log[log_len - 1] = 0;
    }
    return BMFR_OK;
}

int32_t bmfr_batch_record_marker(bmfr_batch b, int32_t marker) {
    if (!b) return fail(BMFR_ERR_INVALID_ARGUMENT, "batch is null");
    if (marker < 0 || marker >= kStages) return fail(BMFR_ERR_INVALID_ARGUMENT, "marker %d", marker);
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipEventRecord(b->ev_start[marker], b->stream));
    return BMFR_OK;
}

int32_t bmfr_batch_marker_elapsed(bmfr_batch b, int32_t m0, int32_t m1, float *ms) {
    if (!b || !ms || m0 < 0 || m0 >= kStages || m1 < 0 || m1 >= kStages)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "bad marker arguments");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipEventSynchronize(b->ev_start[m1]));
    HIP_TRY(hipEventElapsedTime(ms, b->ev_start[m0], b->ev_start[m1]));
    b->recorded = false;
    return BMFR_OK;
}

}  // extern "C"


// ---- server: admission by HBM budget, batched steps, the request loop ------
namespace {

double clock_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A queued bmfr_server_request_stream.
struct AdmitRequest {
    bmfr_params params;
    uint64_t tag;
    double t_request;
    int64_t bytes;
};

// One member of a step: the stream, its frame arguments, the frame's output
// surface (null: the stream's bound one) and the producer event to wait for.
struct StepMember {
    bmfr_stream s;
    int32_t frame;
    const float *matrix, *offset;
    void *surface;
    int32_t surface_format;
    int64_t surface_pitch;
    hipEvent_t producer;
};

// A submitted frame inside a launched step.
struct StepFrame {
    bmfr_stream stream;
    int32_t frame;
    uint64_t tag;
    double t_arrival;
};

// A step launched by the request loop, until its completion is seen.
struct StepRecord {
    int64_t id;
    int slot;
    double t_launch;
    std::vector<StepFrame> frames;
};

constexpr int kMaxServerStreams = 4096;

}  // namespace

struct bmfr_server_s {
    int device = 0;
    int64_t budget = 0, used = 0;   // used counts admissions in progress too
    std::vector<bmfr_stream> open;
    // Two HIP streams ("lanes"): a step goes to the lane its members' last
    // steps ran on (ordered there for free) or, with no such member, to the
    // other lane than the last step's, so that consecutive steps of disjoint
    // streams overlap on the GPU (one's VALU-bound fitter beside the other's
    // streaming stages).  Steps that share a stream are ordered through that
    // stream's own HIP stream whichever lanes they run on.
    hipStream_t lane[2]{};
    int last_lane = 0;
    // per-step records, slot = step id % kRing: pinned staging, device tables
    // (one set per launch group: storage F32/F16 x tmp F32/F16 x qr mode)
    // and events; a slot is reused once its previous step has ended
    static constexpr int kGroups = 12;
    static constexpr int kRing = 8;
    // one launch group's tables, as staged and as the kernels read them; a
    // step's groups are packed at the front of its slot (one copy a step)
    struct GroupTables {
        bmfr::BatchStream streams[bmfr::kMaxBatch];
        bmfr::BatchStep steps[bmfr::kMaxBatch];
        int32_t starts[bmfr::kStartKinds][bmfr::kMaxBatch + 1];
    };
    GroupTables *d_tables = nullptr;         // [kRing][kGroups]
    GroupTables *h_stage = nullptr;          // [kRing][kGroups], pinned
    hipEvent_t step_start[kRing]{}, step_end[kRing]{};
    int64_t ring_id[kRing]{};                // step id each slot's record belongs to (0: none)
    int64_t next_step = 1;
    bmfr::FitterConfig fitter{16, 4};
    // one wave per block, no persistent cap: at saturation (62-stream steps of
    // 720p) 19.0-19.2k frames/s against 18.1-18.4k for a persistent grid of
    // 512 four-wave work-groups (two waves per SIMD); steps of up to 2048
    // blocks launch the same grid either way
    int grid_cap = 0;
    int fuse = BMFR_FUSE_345;
    // The request loop.  `mu` guards the fields above that steps change and
    // everything below, and the queues of the server's streams.
    std::mutex mu;
    std::condition_variable cv_work;     // the dispatcher: frames queued, or stop
    std::condition_variable cv_events;   // pollers: events to read
    std::condition_variable cv_drained;  // closers: a stream's frames have run
    std::thread dispatcher;
    bool stop = false;
    int max_in_flight = 2;
    int max_step_streams = bmfr::kMaxBatch;
    int queued_frames = 0;
    std::deque<StepRecord> in_flight;
    std::deque<bmfr_server_event> events;
    std::deque<AdmitRequest> admit_fifo;
    std::vector<hipEvent_t> event_pool;     // producer events not in use
};

namespace {

int server_group(bmfr_stream s) {
    return ((s->p.storage == BMFR_STORAGE_F16 ? 2 : 0) + (s->p.tmp_storage == BMFR_STORAGE_F16 ? 1 : 0)) * 3 +
           s->p.qr_mode;
}

// Launches one step over `m` (distinct open streams of the server) on the
// server's HIP stream; sv->mu is held.  Each member's frame is finished
// (double buffers swapped) once its launches are enqueued.
int32_t launch_step_locked(bmfr_server sv, const std::vector<StepMember> &m, int64_t *step_id) {
    HIP_TRY(hipSetDevice(sv->device));
    const int64_t id = sv->next_step++;
    const int slot = (int)(id % bmfr_server_s::kRing);
    HIP_TRY(hipEventSynchronize(sv->step_end[slot]));  // the slot's previous step has ended
    sv->ring_id[slot] = id;
    // the lane: where most members' recent steps ran, else the other one
    int score[2] = {0, 0};
    for (const StepMember &mb : m)
        if (mb.s->last_step_id > 0 && mb.s->last_step_id + bmfr_server_s::kRing > id) ++score[mb.s->last_lane];
    const int ln = score[0] || score[1] ? (score[1] > score[0] ? 1 : 0) : 1 - sv->last_lane;
    sv->last_lane = ln;
    const hipStream_t ls = sv->lane[ln];
    const int G = bmfr_server_s::kGroups;
    // launch groups in use, packed in group order at the front of the slot
    int gcount[bmfr_server_s::kGroups] = {}, gindex[bmfr_server_s::kGroups], gid[bmfr_server_s::kGroups];
    for (const StepMember &mb : m) ++gcount[server_group(mb.s)];
    int used = 0;
    for (int g = 0; g < G; ++g) {
        gindex[g] = gcount[g] ? used : -1;
        if (gcount[g]) gid[used++] = g;
        gcount[g] = 0;
    }
    bmfr_server_s::GroupTables *hs = sv->h_stage + (size_t)slot * G;
    bmfr_server_s::GroupTables *dt = sv->d_tables + (size_t)slot * G;
    for (const StepMember &mb : m) {
        bmfr_stream s = mb.s;
        const int g = server_group(s);
        bmfr_server_s::GroupTables &t = hs[gindex[g]];
        const int k = gcount[g]++;
        s->batch_stream(t.streams[k]);
        s->batch_step(mb.frame, mb.matrix, mb.offset, mb.surface, mb.surface_format, mb.surface_pitch, t.steps[k]);
        if (k == 0) t.starts[0][0] = t.starts[1][0] = t.starts[2][0] = 0;
        t.starts[bmfr::kStartsPadded][k + 1] = t.starts[bmfr::kStartsPadded][k] + (s->sz.margin_width / 32) * (s->sz.margin_height / 8);
        t.starts[bmfr::kStartsImage][k + 1] = t.starts[bmfr::kStartsImage][k] + ((s->p.width + 63) / 64) * ((s->p.height + 3) / 4);
        t.starts[bmfr::kStartsBlocks][k + 1] = t.starts[bmfr::kStartsBlocks][k] + s->sz.blocks;
        // Order after this stream's earlier work: what was enqueued on its
        // HIP stream (which an API call has ordered after the stream's last
        // step), or else its last step itself when that ran on the other
        // lane (on this lane it is ordered already; a step whose slot has
        // been reused has ended).  And after the request's producer.
        if (s->user_work || s->external || s->last_step_id == 0) {
            HIP_TRY(hipEventRecord(s->sync_ev, s->stream));
            HIP_TRY(hipStreamWaitEvent(ls, s->sync_ev, 0));
            s->user_work = false;
        } else if (s->last_lane != ln) {
            const int ps = (int)(s->last_step_id % bmfr_server_s::kRing);
            if (sv->ring_id[ps] == s->last_step_id) HIP_TRY(hipStreamWaitEvent(ls, sv->step_end[ps], 0));
        }
        if (mb.producer) HIP_TRY(hipStreamWaitEvent(ls, mb.producer, 0));
    }
    HIP_TRY(hipMemcpyAsync(dt, hs, sizeof(bmfr_server_s::GroupTables) * used, hipMemcpyHostToDevice, ls));
    HIP_TRY(hipEventRecord(sv->step_start[slot], ls));
    for (int st = 0; st < kStages; ++st) {
        for (int ci = 0; ci < used; ++ci) {
            const int g = gid[ci];
            bmfr::BatchLaunch L;
            L.streams = dt[ci].streams;
            L.steps = dt[ci].steps;
            L.starts = &dt[ci].starts[0][0];
            L.n = gcount[g];
            const int sg = g / 3;
            const bool g16 = sg >= 2, t16 = sg & 1;
            const int qr = g % 3;
            const int32_t(&sts)[bmfr::kStartKinds][bmfr::kMaxBatch + 1] = hs[ci].starts;
            hipError_t e;
            switch (st) {
                case 0: e = bmfr::launch_batch_accumulate_noisy_data(L, sts[bmfr::kStartsPadded][L.n], g16, t16, ls); break;
                case 1: e = bmfr::launch_batch_fitter(L, sts[bmfr::kStartsBlocks][L.n], t16, sv->fitter, sv->grid_cap, qr, ls); break;
                case 2: e = hipSuccess; break;  // fused into stage 4
                case 3:
                    e = sv->fuse == BMFR_FUSE_345
                            ? bmfr::launch_batch_filter_taa(L, sts[bmfr::kStartsBlocks][L.n], g16, ls)
                            : bmfr::launch_batch_weighted_sum_filtered(L, sts[bmfr::kStartsPadded][L.n], g16, ls);
                    break;
                default:
                    e = sv->fuse == BMFR_FUSE_345 ? hipSuccess
                                                  : bmfr::launch_batch_taa(L, sts[bmfr::kStartsImage][L.n], g16, ls);
                    break;
            }
            if (e != hipSuccess) return fail(BMFR_ERR_HIP, "server stage %d launch failed: %s", st, hipGetErrorString(e));
        }
    }
    HIP_TRY(hipEventRecord(sv->step_end[slot], ls));
    for (const StepMember &mb : m) {
        bmfr_stream s = mb.s;
        // a HIP stream the caller holds is ordered after the step now; the
        // others when an API call on the stream next needs it
        if (s->external) {
            HIP_TRY(hipStreamWaitEvent(s->stream, sv->step_end[slot], 0));
            s->ordered_step = id;
        }
        s->last_result = s->cur(S_RESULT_A);
        s->noise_frame = mb.frame;
        s->last_step_id = id;
        s->last_lane = ln;
        s->finish_frame();
    }
    if (step_id) *step_id = id;
    return BMFR_OK;
}

// Orders stream s's own HIP stream after the last server step s took part
// in (a step whose ring slot has been reused has ended); sv->mu is held.
int32_t order_stream_locked(bmfr_server sv, bmfr_stream s) {
    if (s->last_step_id > s->ordered_step) {
        const int ps = (int)(s->last_step_id % bmfr_server_s::kRing);
        if (sv->ring_id[ps] == s->last_step_id) {
            HIP_TRY(hipSetDevice(sv->device));
            HIP_TRY(hipStreamWaitEvent(s->stream, sv->step_end[ps], 0));
        }
        s->ordered_step = s->last_step_id;
    }
    return BMFR_OK;
}

// Declared before check_stream: a stream entry point on a server stream.
int32_t order_after_server_steps(bmfr_stream s, bool external) {
    bmfr_server sv = s->owner;
    std::lock_guard<std::mutex> lk(sv->mu);
    s->user_work = true;
    if (external) s->external = true;
    return order_stream_locked(sv, s);
}

hipEvent_t pool_event_locked(bmfr_server sv) {
    if (!sv->event_pool.empty()) {
        hipEvent_t e = sv->event_pool.back();
        sv->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

// Completions of the request loop's steps, oldest first; sv->mu is held.
void reap_locked(bmfr_server sv) {
    bool any = false;
    while (!sv->in_flight.empty()) {
        StepRecord &r = sv->in_flight.front();
        const hipError_t q = hipEventQuery(sv->step_end[r.slot]);
        if (q == hipErrorNotReady) break;
        const double t = clock_seconds();
        float ms = -1.f;
        if (q == hipSuccess) (void)hipEventElapsedTime(&ms, sv->step_start[r.slot], sv->step_end[r.slot]);
        for (const StepFrame &f : r.frames) {
            bmfr_server_event ev{};
            ev.kind = q == hipSuccess ? BMFR_EVENT_FRAME_DONE : BMFR_EVENT_FRAME_FAILED;
            ev.status = q == hipSuccess ? BMFR_OK : BMFR_ERR_HIP;
            ev.stream = f.stream;
            ev.tag = f.tag;
            ev.frame_number = f.frame;
            ev.step_streams = (int32_t)r.frames.size();
            ev.step_id = r.id;
            ev.t_arrival = f.t_arrival;
            ev.t_launch = r.t_launch;
            ev.t_done = t;
            ev.step_gpu_ms = ms;
            sv->events.push_back(ev);
            --f.stream->frames_in_flight;
        }
        sv->in_flight.pop_front();
        any = true;
    }
    if (any) {
        sv->cv_events.notify_all();
        sv->cv_drained.notify_all();
    }
}

// One step of the oldest queued frame of every stream that has one (oldest
// arrivals first, at most max_step_streams); sv->mu is held.
void launch_queued_locked(bmfr_server sv) {
    std::vector<bmfr_stream> cand;
    for (bmfr_stream s : sv->open)
        if (!s->queue.empty()) cand.push_back(s);
    if (cand.empty()) return;
    std::stable_sort(cand.begin(), cand.end(), [](bmfr_stream a, bmfr_stream b) {
        return a->queue.front().req.t_arrival < b->queue.front().req.t_arrival;
    });
    if ((int)cand.size() > sv->max_step_streams) cand.resize(sv->max_step_streams);
    std::vector<StepMember> m;
    StepRecord rec;
    rec.t_launch = clock_seconds();
    for (bmfr_stream s : cand) {
        const ServerFrame &f = s->queue.front();
        const bmfr_frame_request &r = f.req;
        if (r.albedo) {
            s->bind_in[0] = r.albedo;
            s->bind_in[1] = r.normals;
            s->bind_in[2] = r.positions;
            s->bind_in[3] = r.noisy;
        }
        m.push_back({s, r.frame_number, r.prev_frame_camera_matrix, r.pixel_offset, r.surface, r.surface_format,
                     r.surface_row_pitch, f.producer});
        rec.frames.push_back({s, r.frame_number, r.tag, r.t_arrival});
    }
    const int32_t st = launch_step_locked(sv, m, &rec.id);
    for (bmfr_stream s : cand) {
        if (s->queue.front().producer) sv->event_pool.push_back(s->queue.front().producer);
        s->queue.pop_front();
        --sv->queued_frames;
        if (st == BMFR_OK) ++s->frames_in_flight;
        else for (auto &p : s->bind_in) p = nullptr;
    }
    if (st == BMFR_OK) {
        rec.slot = (int)(rec.id % bmfr_server_s::kRing);
        sv->in_flight.push_back(std::move(rec));
    } else {
        for (const StepFrame &f : rec.frames) {
            bmfr_server_event ev{};
            ev.kind = BMFR_EVENT_FRAME_FAILED;
            ev.status = st;
            ev.stream = f.stream;
            ev.tag = f.tag;
            ev.frame_number = f.frame;
            ev.t_arrival = f.t_arrival;
            ev.t_launch = rec.t_launch;
            ev.t_done = clock_seconds();
            sv->events.push_back(ev);
        }
        sv->cv_events.notify_all();
    }
    sv->cv_drained.notify_all();
}

// The dispatcher thread: launch whenever fewer than max_in_flight steps are
// on the GPU and frames are queued; otherwise wait for a submit (nothing in
// flight) or look for completions again shortly (steps in flight).
void dispatcher_main(bmfr_server sv) {
    (void)hipSetDevice(sv->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    while (!sv->stop) {
        reap_locked(sv);
        if ((int)sv->in_flight.size() < sv->max_in_flight && sv->queued_frames > 0) {
            launch_queued_locked(sv);
            continue;
        }
        if (sv->in_flight.empty())
            sv->cv_work.wait(lk, [&] { return sv->stop || (sv->queued_frames > 0 && sv->max_in_flight > 0); });
        else
            sv->cv_work.wait_for(lk, std::chrono::microseconds(20));
    }
}

// Opens a stream owned by the server (its budget already reserved).
int32_t open_owned_locked_out(bmfr_server sv, const bmfr_params *params, bmfr_stream *out) {
    HIP_TRY(hipSetDevice(sv->device));
    bmfr_stream s = nullptr;
    if (int32_t st = bmfr_stream_open(params, &s)) return st;
    s->owner = sv;
    *out = s;
    return BMFR_OK;
}

// Admits queued requests, in order, while the budget allows (no lock held).
void admit_waiting(bmfr_server sv) {
    for (;;) {
        AdmitRequest r;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            if (sv->admit_fifo.empty()) return;
            r = sv->admit_fifo.front();
            if (sv->used + r.bytes > sv->budget || (int)sv->open.size() >= kMaxServerStreams) return;
            sv->admit_fifo.pop_front();
            sv->used += r.bytes;
        }
        bmfr_stream s = nullptr;
        const int32_t st = open_owned_locked_out(sv, &r.params, &s);
        std::lock_guard<std::mutex> lk(sv->mu);
        bmfr_server_event ev{};
        ev.kind = BMFR_EVENT_STREAM_ADMITTED;
        ev.status = st;
        ev.tag = r.tag;
        ev.t_arrival = r.t_request;
        ev.t_done = clock_seconds();
        if (st == BMFR_OK) {
            sv->open.push_back(s);
            ev.stream = s;
        } else {
            sv->used -= r.bytes;
        }
        sv->events.push_back(ev);
        sv->cv_events.notify_all();
    }
}

bool owns_locked(bmfr_server sv, bmfr_stream s) {
    return std::find(sv->open.begin(), sv->open.end(), s) != sv->open.end();
}

}  // namespace

extern "C" {

double bmfr_clock_seconds(void) { return clock_seconds(); }

int32_t bmfr_server_create(int64_t hbm_budget_bytes, bmfr_server *out) {
    if (!out) return fail(BMFR_ERR_INVALID_ARGUMENT, "out is null");
    if (hbm_budget_bytes <= 0) return fail(BMFR_ERR_INVALID_ARGUMENT, "budget must be positive");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(BMFR_ERR_NO_DEVICE, "no HIP device visible");
    bmfr_server sv = new (std::nothrow) bmfr_server_s;
    if (!sv) return fail(BMFR_ERR_INVALID_ARGUMENT, "host allocation failed");
    sv->budget = hbm_budget_bytes;
    hipError_t e = hipGetDevice(&sv->device);
    for (int l = 0; l < 2 && e == hipSuccess; ++l) e = hipStreamCreateWithFlags(&sv->lane[l], hipStreamNonBlocking);
    const size_t G = bmfr_server_s::kGroups, R = bmfr_server_s::kRing;
    if (e == hipSuccess) e = hipMalloc((void **)&sv->d_tables, sizeof(bmfr_server_s::GroupTables) * R * G);
    if (e == hipSuccess) e = hipHostMalloc((void **)&sv->h_stage, sizeof(bmfr_server_s::GroupTables) * R * G);
    for (int i = 0; i < bmfr_server_s::kRing && e == hipSuccess; ++i) {
        e = hipEventCreate(&sv->step_start[i]);
        if (e == hipSuccess) e = hipEventCreate(&sv->step_end[i]);
    }
    if (e != hipSuccess) {
        bmfr_server_destroy(sv);
        return fail(BMFR_ERR_HIP, "server set-up failed: %s", hipGetErrorString(e));
    }
    *out = sv;
    return BMFR_OK;
}

int32_t bmfr_server_destroy(bmfr_server sv) {
    if (!sv) return BMFR_OK;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    if (sv->dispatcher.joinable()) sv->dispatcher.join();
    (void)hipSetDevice(sv->device);
    for (hipStream_t l : sv->lane)
        if (l) (void)hipStreamSynchronize(l);
    for (bmfr_stream s : sv->open) {
        for (ServerFrame &f : s->queue)
            if (f.producer) sv->event_pool.push_back(f.producer);
        s->queue.clear();
        destroy_stream(s);
    }
    sv->open.clear();
    for (hipEvent_t e : sv->event_pool) (void)hipEventDestroy(e);
    for (int i = 0; i < bmfr_server_s::kRing; ++i) {
        if (sv->step_start[i]) (void)hipEventDestroy(sv->step_start[i]);
        if (sv->step_end[i]) (void)hipEventDestroy(sv->step_end[i]);
    }
    if (sv->h_stage) (void)hipHostFree(sv->h_stage);
    if (sv->d_tables) (void)hipFree(sv->d_tables);
    for (hipStream_t l : sv->lane)
        if (l) (void)hipStreamDestroy(l);
    delete sv;
    return BMFR_OK;
}

int32_t bmfr_server_usage(bmfr_server sv, int64_t *used, int64_t *budget, int32_t *streams) {
    if (!sv) return fail(BMFR_ERR_INVALID_ARGUMENT, "server is null");
    std::lock_guard<std::mutex> lk(sv->mu);
    if (used) *used = sv->used;
    if (budget) *budget = sv->budget;
    if (streams) *streams = (int32_t)sv->open.size();
    return BMFR_OK;
}

int32_t bmfr_server_open_stream(bmfr_server sv, const bmfr_params *params, bmfr_stream *out) {
    if (!sv || !out) return fail(BMFR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    bmfr_sizes z;
    if (int32_t st = bmfr_compute_sizes(params, &z)) return st;
    if (params->feature_count)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "servers run the reference's feature set");
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->used + z.device_bytes > sv->budget)
            return fail(BMFR_ERR_OUT_OF_BUDGET, "stream needs %lld bytes, %lld of %lld in use",
                        (long long)z.device_bytes, (long long)sv->used, (long long)sv->budget);
        if ((int)sv->open.size() >= kMaxServerStreams) return fail(BMFR_ERR_OUT_OF_BUDGET, "too many open streams");
        sv->used += z.device_bytes;
    }
    bmfr_stream s = nullptr;
    const int32_t st = open_owned_locked_out(sv, params, &s);
    std::lock_guard<std::mutex> lk(sv->mu);
    if (st) {
        sv->used -= z.device_bytes;
        return st;
    }
    sv->open.push_back(s);
    *out = s;
    return BMFR_OK;
}

int32_t bmfr_server_request_stream(bmfr_server sv, const bmfr_params *params, uint64_t tag, bmfr_stream *out) {
    if (!sv || !out) return fail(BMFR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    bmfr_sizes z;
    if (int32_t st = bmfr_compute_sizes(params, &z)) return st;
    if (params->feature_count)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "servers run the reference's feature set");
    if (z.device_bytes > sv->budget)
        return fail(BMFR_ERR_OUT_OF_BUDGET, "stream needs %lld bytes, more than the whole budget of %lld",
                    (long long)z.device_bytes, (long long)sv->budget);
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (!sv->admit_fifo.empty() || sv->used + z.device_bytes > sv->budget ||
            (int)sv->open.size() >= kMaxServerStreams) {
            sv->admit_fifo.push_back({*params, tag, clock_seconds(), z.device_bytes});
            return BMFR_OK;   // queued: admitted later by a close
        }
        sv->used += z.device_bytes;
    }
    bmfr_stream s = nullptr;
    const int32_t st = open_owned_locked_out(sv, params, &s);
    std::lock_guard<std::mutex> lk(sv->mu);
    if (st) {
        sv->used -= z.device_bytes;
        return st;
    }
    sv->open.push_back(s);
    *out = s;
    return BMFR_OK;
}

int32_t bmfr_server_cancel_request(bmfr_server sv, uint64_t tag, int32_t *found) {
    if (!sv) return fail(BMFR_ERR_INVALID_ARGUMENT, "server is null");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = std::find_if(sv->admit_fifo.begin(), sv->admit_fifo.end(),
                           [&](const AdmitRequest &r) { return r.tag == tag; });
    const bool f = it != sv->admit_fifo.end();
    if (f) sv->admit_fifo.erase(it);
    if (found) *found = f ? 1 : 0;
    return BMFR_OK;
}

int32_t bmfr_server_close_stream(bmfr_server sv, bmfr_stream s) {
    if (!sv || !s) return fail(BMFR_ERR_INVALID_ARGUMENT, "null argument");
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        if (!owns_locked(sv, s)) return fail(BMFR_ERR_INVALID_ARGUMENT, "stream does not belong to this server");
        if (s->closing) return fail(BMFR_ERR_STATE, "the stream is already being closed");
        s->closing = true;
        sv->cv_drained.wait(lk, [&] { return s->queue.empty() && s->frames_in_flight == 0; });
        sv->open.erase(std::find(sv->open.begin(), sv->open.end(), s));
        sv->used -= s->sz.device_bytes;
        if (int32_t st = order_stream_locked(sv, s)) return st;
    }
    // destroy_stream synchronises the stream's own HIP stream, now ordered
    // after every server step the stream took part in.
    destroy_stream(s);
    admit_waiting(sv);
    return BMFR_OK;
}

int32_t bmfr_server_step(bmfr_server sv, int32_t count, const bmfr_stream *streams, const int32_t *frames,
                         const float *matrices, const float *offsets, int64_t *step_id) {
    if (!sv || !streams || !frames || !matrices || !offsets)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "null argument");
    if (count < 1 || count > bmfr::kMaxBatch)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "step of %d streams outside 1..%d", count, bmfr::kMaxBatch);
    std::lock_guard<std::mutex> lk(sv->mu);
    std::vector<StepMember> m;
    for (int i = 0; i < count; ++i) {
        if (!streams[i] || frames[i] < 0) return fail(BMFR_ERR_INVALID_ARGUMENT, "bad stream or frame %d", i);
        if (!owns_locked(sv, streams[i]))
            return fail(BMFR_ERR_INVALID_ARGUMENT, "stream %d was not opened by this server", i);
        for (int j = 0; j < i; ++j)
            if (streams[j] == streams[i]) return fail(BMFR_ERR_INVALID_ARGUMENT, "stream %d listed twice", i);
        if (!streams[i]->queue.empty() || streams[i]->closing)
            return fail(BMFR_ERR_STATE, "stream %d has frames queued in the request loop or is being closed", i);
        m.push_back({streams[i], frames[i], matrices + 16 * i, offsets + 2 * i, nullptr, 0, 0, nullptr});
    }
    return launch_step_locked(sv, m, step_id);
}

int32_t bmfr_server_step_query(bmfr_server sv, int64_t step_id, int32_t *done, float *ms) {
    if (!sv || !done) return fail(BMFR_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    if (step_id < 1 || step_id >= sv->next_step)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "step %lld was never launched", (long long)step_id);
    const int slot = (int)(step_id % bmfr_server_s::kRing);
    if (sv->ring_id[slot] != step_id)
        return fail(BMFR_ERR_STATE, "step %lld has expired: its record holds step %lld now", (long long)step_id,
                    (long long)sv->ring_id[slot]);
    HIP_TRY(hipSetDevice(sv->device));
    const hipError_t q = hipEventQuery(sv->step_end[slot]);
    if (q == hipErrorNotReady) {
        *done = 0;
        return BMFR_OK;
    }
    if (q != hipSuccess) return fail(BMFR_ERR_HIP, "step query failed: %s", hipGetErrorString(q));
    *done = 1;
    if (ms) HIP_TRY(hipEventElapsedTime(ms, sv->step_start[slot], sv->step_end[slot]));
    return BMFR_OK;
}

int32_t bmfr_server_set_fused(bmfr_server sv, int32_t level) {
    if (!sv) return fail(BMFR_ERR_INVALID_ARGUMENT, "server is null");
    if (level != BMFR_FUSE_34 && level != BMFR_FUSE_345)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "server steps run fuse level 1 or 2, not %d", level);
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->fuse = level;
    return BMFR_OK;
}

int32_t bmfr_server_set_policy(bmfr_server sv, int32_t max_in_flight, int32_t max_streams_per_step) {
    if (!sv) return fail(BMFR_ERR_INVALID_ARGUMENT, "server is null");
    if (max_in_flight < 0 || max_in_flight > 4 || max_streams_per_step < 1 || max_streams_per_step > bmfr::kMaxBatch)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "policy %d steps in flight, %d streams per step", max_in_flight,
                    max_streams_per_step);
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->max_in_flight = max_in_flight;
    sv->max_step_streams = max_streams_per_step;
    sv->cv_work.notify_all();
    return BMFR_OK;
}

int32_t bmfr_server_submit(bmfr_server sv, const bmfr_frame_request *r) {
    if (!sv || !r || !r->stream) return fail(BMFR_ERR_INVALID_ARGUMENT, "null argument");
    if (int32_t st = check_frame(r->frame_number)) return st;
    bmfr_stream s = r->stream;
    const void *in[4] = {r->albedo, r->normals, r->positions, r->noisy};
    const int nulls = !in[0] + !in[1] + !in[2] + !in[3];
    if (nulls != 0 && nulls != 4) return fail(BMFR_ERR_INVALID_ARGUMENT, "give all four inputs or none");
    for (int i = 0; i < 4; ++i)
        if (((uintptr_t)in[i] & 15) != 0)
            return fail(BMFR_ERR_INVALID_ARGUMENT, "device input %d is not 16-byte aligned", i);
    std::unique_lock<std::mutex> lk(sv->mu);
    if (!owns_locked(sv, s)) return fail(BMFR_ERR_INVALID_ARGUMENT, "stream does not belong to this server");
    if (s->closing) return fail(BMFR_ERR_STATE, "the stream is being closed");
    if (r->surface)
        if (int32_t st = check_surface(s, r->surface, r->surface_format, r->surface_row_pitch)) return st;
    if (nulls == 4 && !s->queue.empty())
        return fail(BMFR_ERR_STATE, "the stream's own input buffers hold one frame and one is queued: "
                                    "give device inputs to queue more");
    HIP_TRY(hipSetDevice(sv->device));
    ServerFrame f;
    f.req = *r;
    if (r->producer_stream) {
        f.producer = pool_event_locked(sv);
        if (!f.producer) return fail(BMFR_ERR_HIP, "event creation failed");
        const hipError_t e = hipEventRecord(f.producer, (hipStream_t)r->producer_stream);
        if (e != hipSuccess) {
            sv->event_pool.push_back(f.producer);
            return fail(BMFR_ERR_HIP, "recording the producer event failed: %s", hipGetErrorString(e));
        }
    }
    if (!sv->dispatcher.joinable()) {
        try {
            sv->dispatcher = std::thread(dispatcher_main, sv);
        } catch (...) {
            if (f.producer) sv->event_pool.push_back(f.producer);
            return fail(BMFR_ERR_STATE, "cannot start the dispatcher thread");
        }
    }
    s->queue.push_back(f);
    ++sv->queued_frames;
    lk.unlock();
    sv->cv_work.notify_one();
    return BMFR_OK;
}

int32_t bmfr_server_poll(bmfr_server sv, bmfr_server_event *out, int32_t max_events, int32_t timeout_us,
                         int32_t *count) {
    if (!sv || !count || max_events < 0 || (max_events > 0 && !out) || timeout_us < 0)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "bad poll arguments");
    std::unique_lock<std::mutex> lk(sv->mu);
    if (sv->events.empty() && timeout_us > 0)
        sv->cv_events.wait_for(lk, std::chrono::microseconds(timeout_us), [&] { return !sv->events.empty(); });
    int n = 0;
    while (n < max_events && !sv->events.empty()) {
        out[n++] = sv->events.front();
        sv->events.pop_front();
    }
    *count = n;
    return BMFR_OK;
}

int32_t bmfr_server_synchronize(bmfr_server sv) {
    if (!sv) return fail(BMFR_ERR_INVALID_ARGUMENT, "server is null");
    {
        // the request loop's queued frames are launched first
        std::unique_lock<std::mutex> lk(sv->mu);
        sv->cv_drained.wait(lk, [&] { return sv->queued_frames == 0; });
    }
    HIP_TRY(hipSetDevice(sv->device));
    for (hipStream_t l : sv->lane) HIP_TRY(hipStreamSynchronize(l));
    return BMFR_OK;
}

}  // extern "C"

extern "C" {

// ---- scene -------------------------------------------------------------------
int32_t bmfr_scene_create(uint64_t seed, int32_t spheres, float firefly, int32_t width, int32_t height,
                          int32_t static_camera, bmfr_scene *out) {
    if (!out) return fail(BMFR_ERR_INVALID_ARGUMENT, "out is null");
    if (spheres < 0 || spheres > bmfr::kMaxSpheres)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "spheres %d outside 0..%d", spheres, bmfr::kMaxSpheres);
    if (width <= 0 || height <= 0) return fail(BMFR_ERR_INVALID_ARGUMENT, "bad scene size");
    bmfr_scene sc = new (std::nothrow) bmfr_scene_s;
    if (!sc) return fail(BMFR_ERR_INVALID_ARGUMENT, "host allocation failed");
    bmfr::scene_init(sc->data, seed, spheres, firefly, width, height, static_camera);
    *out = sc;
    return BMFR_OK;
}

int32_t bmfr_scene_destroy(bmfr_scene sc) {
    if (!sc) return BMFR_OK;
    if (sc->device_copy) (void)hipFree(sc->device_copy);
    delete sc;
    return BMFR_OK;
}

int32_t bmfr_scene_camera(bmfr_scene sc, int32_t frame, float *vp, float *offset) {
    if (!sc || !vp || !offset) return fail(BMFR_ERR_INVALID_ARGUMENT, "null argument");
    bmfr::CameraFrame c;
    bmfr::scene_camera(sc->data, frame, c);
    std::memcpy(vp, c.vp, sizeof(c.vp));
    offset[0] = c.offset[0];
    offset[1] = c.offset[1];
    return BMFR_OK;
}

int32_t bmfr_scene_render_host(bmfr_scene sc, int32_t frame, float *albedo, float *normals,
                               float *positions, float *noisy) {
    if (!sc) return fail(BMFR_ERR_INVALID_ARGUMENT, "scene is null");
    bmfr::CameraFrame c;
    bmfr::scene_camera(sc->data, frame, c);
    const int W = sc->data.width, H = sc->data.height;
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t] {
            float px[12];
            for (int y = (int)t; y < H; y += (int)nt)
                for (int x = 0; x < W; ++x) {
                    bmfr::scene_pixel(sc->data, c, frame, x, y, px);
                    const size_t lp = (size_t)y * W + x;
                    for (int k = 0; k < 3; ++k) {
                        if (albedo) albedo[lp * 3 + k] = px[k];
                        if (normals) normals[lp * 3 + k] = px[3 + k];
                        if (positions) positions[lp * 3 + k] = px[6 + k];
                        if (noisy) noisy[lp * 3 + k] = px[9 + k];
                    }
                }
        });
    }
    for (auto &t : th) t.join();
    return BMFR_OK;
}

int32_t bmfr_scene_render_to_stream(bmfr_scene sc, int32_t frame, bmfr_stream s) {
    if (!sc) return fail(BMFR_ERR_INVALID_ARGUMENT, "scene is null");
    if (int32_t st = check_stream(s)) return st;
    if (sc->data.width != s->p.width || sc->data.height != s->p.height)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "scene %dx%d does not match stream %dx%d", sc->data.width,
                    sc->data.height, s->p.width, s->p.height);
    HIP_TRY(hipSetDevice(s->device));
    if (!sc->device_copy) {
        HIP_TRY(hipMalloc((void **)&sc->device_copy, sizeof(bmfr::SceneData)));
        HIP_TRY(hipMemcpy(sc->device_copy, &sc->data, sizeof(bmfr::SceneData), hipMemcpyHostToDevice));
    }
    bmfr::CameraFrame c;
    bmfr::scene_camera(sc->data, frame, c);
    hipError_t e = bmfr::launch_render(sc->device_copy, sc->data.width, sc->data.height, c, frame,
                                       s->p.storage == BMFR_STORAGE_F16, s->slot[S_ALBEDO],
                                       s->slot[s->cur(S_NORMALS_A)], s->slot[s->cur(S_POSITIONS_A)],
                                       s->slot[S_NOISY_IN], s->stream);
    if (e != hipSuccess) return fail(BMFR_ERR_HIP, "scene render launch failed: %s", hipGetErrorString(e));
    return BMFR_OK;
}

int32_t bmfr_scene_render_device(bmfr_scene sc, int32_t frame, int32_t storage, void *albedo, void *normals,
                                 void *positions, void *noisy, void *hip_stream) {
    if (!sc) return fail(BMFR_ERR_INVALID_ARGUMENT, "scene is null");
    if (!albedo || !normals || !positions || !noisy) return fail(BMFR_ERR_INVALID_ARGUMENT, "null output buffer");
    if (storage != BMFR_STORAGE_F32 && storage != BMFR_STORAGE_F16)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "unknown storage %d", storage);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (!sc->device_copy) {
        HIP_TRY(hipMalloc((void **)&sc->device_copy, sizeof(bmfr::SceneData)));
        HIP_TRY(hipMemcpy(sc->device_copy, &sc->data, sizeof(bmfr::SceneData), hipMemcpyHostToDevice));
    }
    bmfr::CameraFrame c;
    bmfr::scene_camera(sc->data, frame, c);
    hipError_t e = bmfr::launch_render(sc->device_copy, sc->data.width, sc->data.height, c, frame,
                                       storage == BMFR_STORAGE_F16, albedo, normals, positions, noisy,
                                       (hipStream_t)hip_stream);
    if (e != hipSuccess) return fail(BMFR_ERR_HIP, "scene render launch failed: %s", hipGetErrorString(e));
    return BMFR_OK;
}

}  // extern "C"
