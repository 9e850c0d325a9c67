// Output surfaces (bmfr_bind_output_surface): a pixel's TAA result written
// into a caller-owned display or encoder surface, in one of the formats of
// bmfr_surface_format, by the kernel that computes the result.  The 8-bit
// conversion is shared with the host PNG writer (dataset_io.cpp).
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"

namespace bmfr {

enum SurfaceFormat { kSurfaceRgbF32 = 0, kSurfaceRgbaF32 = 1, kSurfaceRgbaF16 = 2, kSurfaceRgbaU8 = 3 };

// Bytes per pixel and the alignment a surface row and base must have.
__host__ __device__ inline int surface_pixel_bytes(int format) {
    return format == kSurfaceRgbF32 ? 12 : format == kSurfaceRgbaF32 ? 16 : format == kSurfaceRgbaF16 ? 8 : 4;
}
__host__ __device__ inline int surface_alignment(int format) {
    return format == kSurfaceRgbaF32 ? 16 : format == kSurfaceRgbaF16 ? 8 : 4;
}

// round(clamp(v, 0, 1) * 255): the 8-bit value of the reference's PNG output
// (opencl/bmfr.cpp:519-547); NaN gives 0.  Compiled with contraction off, so
// host and device round the product identically.
__host__ __device__ inline uint8_t quantize_u8(float v) {
    v = v > 0.f ? (v < 1.f ? v : 1.f) : 0.f;
    return (uint8_t)(v * 255.f + 0.5f);
}

// Pixel (x, y) of a surface.  `format` is a kernel argument (wave-uniform),
// so the switch is a scalar branch; each format is one store per pixel
// except RGB_F32 (three dword stores, the AoS float3 layout of the
// reference's result buffer).
__device__ __forceinline__ void store_surface(void *surface, int format, int64_t pitch, int x, int y, float3 v) {
    char *row = (char *)surface + (int64_t)y * pitch;
    switch (format) {
        case kSurfaceRgbF32: {
            float *p = (float *)row + 3 * (int64_t)x;
            gstore(p, 0, v.x);
            gstore(p, 1, v.y);
            gstore(p, 2, v.z);
            break;
        }
        case kSurfaceRgbaF32:
            gstore((float4 *)(row + 16 * (int64_t)x), 0, make_float4(v.x, v.y, v.z, 1.f));
            break;
        case kSurfaceRgbaF16: {
            const uint32_t lo = (uint32_t)__half_as_ushort(__float2half_rn(v.x)) |
                                ((uint32_t)__half_as_ushort(__float2half_rn(v.y)) << 16);
            const uint32_t hi = (uint32_t)__half_as_ushort(__float2half_rn(v.z)) | (0x3c00u << 16);
            gstore((uint2 *)(row + 8 * (int64_t)x), 0, make_uint2(lo, hi));
            break;
        }
        default:
            gstore((uint32_t *)(row + 4 * (int64_t)x), 0,
                   (uint32_t)quantize_u8(v.x) | ((uint32_t)quantize_u8(v.y) << 8) | ((uint32_t)quantize_u8(v.z) << 16) |
                       0xff000000u);
            break;
    }
}

}  // namespace bmfr
