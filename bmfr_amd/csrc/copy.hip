// Device-to-device hand-off of a frame's inputs (a renderer that already has
// its G-buffers in HBM; the reference uploads from the host instead,
// opencl/bmfr.cpp:420-427).  One launch copies up to four buffers with
// 16-byte loads and stores; it is also the plain copy kernel of known size
// used to calibrate HBM byte counters and to measure the copy peak.
#include "device_common.h"
#include "kernels.h"

namespace bmfr {

__global__ __launch_bounds__(256) void k_copy4(CopyJob job) {
    const int b = blockIdx.y;
    const uint4 *__restrict__ src = reinterpret_cast<const uint4 *>(job.src[b]);
    uint4 *__restrict__ dst = reinterpret_cast<uint4 *>(job.dst[b]);
    const int64_t n16 = job.bytes[b] >> 4;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
    // tail bytes (buffers are 256-byte aligned; sizes need not be)
    const int64_t tail = job.bytes[b] & 15;
    if (blockIdx.x == 0 && threadIdx.x < tail)
        reinterpret_cast<uint8_t *>(job.dst[b])[n16 * 16 + threadIdx.x] =
            reinterpret_cast<const uint8_t *>(job.src[b])[n16 * 16 + threadIdx.x];
}

hipError_t launch_copy4(const CopyJob &job, int count, hipStream_t s) {
    if (count <= 0 || count > 4) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_copy4, dim3(1024, count), dim3(256), 0, s, job);
    return hipGetLastError();
}

}  // namespace bmfr
