// Device-to-device hand-off of a frame's inputs (a renderer that already has
// its G-buffers in HBM; the reference uploads from the host instead,
// opencl/bmfr.cpp:420-427).  One launch copies up to four buffers with
// 16-byte loads and stores; it is also the plain copy kernel of known size
// used to calibrate HBM byte counters and to measure the copy peak.  The
// LDS read kernel below measures the other ceiling the fitter is judged
// against.
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

// LDS read bandwidth: each work-item reads 16 bytes (ds_read_b128) per step
// from a 32 KB tile, consecutive work-items on consecutive 16-byte words (no
// bank conflicts), four independent reads in flight per step; the folded
// values are written once so that no read is dead.  Bytes read per launch =
// gridDim.x * 256 * 16 * 4 * steps.
constexpr int kLdsWords = 2048;   // uint4 words: 32 KB
__global__ __launch_bounds__(256) void k_lds_read(uint4 *__restrict__ out, int steps) {
    __shared__ uint4 buf[kLdsWords];
    const int t = threadIdx.x;
    for (int i = t; i < kLdsWords; i += 256) buf[i] = make_uint4(i, 3 * i, 5 * i, 7 * i);
    __syncthreads();
    uint4 acc[4] = {};
    int idx = t;
    for (int k = 0; k < steps; ++k) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint4 v = buf[(idx + 256 * j) & (kLdsWords - 1)];
            acc[j].x ^= v.x;
            acc[j].y ^= v.y;
            acc[j].z ^= v.z;
            acc[j].w ^= v.w;
        }
        idx = (idx + 1024 + 1) & (kLdsWords - 1);
    }
    uint4 r = acc[0];
#pragma unroll
    for (int j = 1; j < 4; ++j) {
        r.x ^= acc[j].x;
        r.y ^= acc[j].y;
        r.z ^= acc[j].z;
        r.w ^= acc[j].w;
    }
    out[(int64_t)blockIdx.x * 256 + t] = r;
}

hipError_t launch_lds_read(uint4 *out, int groups, int steps, hipStream_t s) {
    if (groups <= 0 || steps <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_lds_read, dim3(groups), dim3(256), 0, s, out, steps);
    return hipGetLastError();
}

}  // namespace bmfr
