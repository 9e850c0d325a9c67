// bmfr_fit_blocks: the fitter over caller-owned buffers, for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  Its kernels are bmfr_fitter's;
// the one difference is where their scratch lives: the frame's feature-noise
// table and the per-block FP64 re-factorisation flags are in `scratch`, a
// file-scope device array of fixed size that each GPU holds until the process
// ends, instead of a stream's allocation.
//
// `scratch` names a host-side stand-in in host code, not a device address:
// its device address is fetched once per device with hipGetSymbolAddress
// (with that device current) and only that address reaches kernels.  Every
// call on a GPU shares that GPU's array, so calls that overlap on one GPU
// race; nothing in it needs to start at zero (the noise table and every
// block's flag are written before they are read, each call).
#include <hip/hip_runtime.h>

#include <mutex>

#include "bmfr_hip.h"
#include "device_common.h"
#include "host_error.h"
#include "kernels.h"

namespace {

using bmfr_host::fail;

// noise table: (features - 1) x 1024 floats of the reference's set; flags:
// one int32 per block, at most kMaxFitBlocks blocks (a 3840x2160 frame has
// 8,349)
constexpr int kMaxFitBlocks = 16384;
constexpr size_t kNoiseBytes = (size_t)bmfr::kNoiseTable * sizeof(float);
constexpr size_t kFlagsOffset = (kNoiseBytes + 255) & ~(size_t)255;
constexpr size_t kScratchBytes = kFlagsOffset + (size_t)kMaxFitBlocks * sizeof(int32_t);

static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, fetched once.
int32_t scratch_ptr(void **out) {
    int dev = 0;
    const hipError_t e0 = hipGetDevice(&dev);
    if (e0 != hipSuccess) return fail(BMFR_ERR_NO_DEVICE, "no current HIP device: %s", hipGetErrorString(e0));
    if (dev < 0 || dev >= kMaxDevices) return fail(BMFR_ERR_NO_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(BMFR_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return BMFR_OK;
}

}  // namespace

extern "C" int32_t bmfr_fit_blocks(const void *tmp_data, int32_t tmp_storage, int32_t blocks, int32_t frame_number,
                                   double noise_amount, int32_t qr_mode, float refactor_threshold, float *weights,
                                   float *mins_maxs, void *hip_stream) {
    if (!tmp_data || !weights || !mins_maxs) return fail(BMFR_ERR_INVALID_ARGUMENT, "null buffer");
    if (tmp_storage != BMFR_STORAGE_F32 && tmp_storage != BMFR_STORAGE_F16)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "unknown tmp_storage %d", tmp_storage);
    if (qr_mode < BMFR_QR_FP32 || qr_mode > BMFR_QR_KAHAN_FP64)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "unknown qr_mode %d", qr_mode);
    if (blocks < 1 || blocks > kMaxFitBlocks)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "%d blocks outside 1..%d (the scratch holds flags for %d blocks)",
                    blocks, kMaxFitBlocks, kMaxFitBlocks);
    if (frame_number < 0) return fail(BMFR_ERR_INVALID_ARGUMENT, "frame_number %d < 0", frame_number);
    if (((uintptr_t)tmp_data & 15) != 0) return fail(BMFR_ERR_INVALID_ARGUMENT, "tmp_data is not 16-byte aligned");
    void *base = nullptr;
    if (int32_t st = scratch_ptr(&base)) return st;
    bmfr::StageArgs a{};
    a.frame = frame_number;
    a.blocks = blocks;
    a.noise_amount = noise_amount;
    a.tmp_f16 = tmp_storage == BMFR_STORAGE_F16;
    a.qr_mode = qr_mode;
    a.refactor_threshold = refactor_threshold;
    bmfr::StageBuffers b{};
    b.tmp_data = const_cast<void *>(tmp_data);
    b.weights = weights;
    b.mins_maxs = mins_maxs;
    b.noise = base;
    b.flags = (char *)base + kFlagsOffset;
    const hipStream_t s = (hipStream_t)hip_stream;
    hipError_t e = bmfr::launch_noise_table(a, b, s);
    if (e == hipSuccess) e = bmfr::launch_fitter(a, b, bmfr::FitterConfig{16, 2}, qr_mode, s);
    if (e != hipSuccess) return fail(BMFR_ERR_HIP, "fitter launch failed: %s", hipGetErrorString(e));
    return BMFR_OK;
}
