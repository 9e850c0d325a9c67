// GPU renderer stand-in: writes one frame of the procedural scene straight
// into a stream's current input buffers (the zero-copy hand-off a path
// tracer would do; the reference instead uploads EXR data from the host,
// opencl/bmfr.cpp:420-427).
#include "device_common.h"
#include "scene.h"
#include "scene_launch.h"

namespace bmfr {

template <typename GT>
__global__ __launch_bounds__(256) void k_render(const SceneData *__restrict__ scene, CameraFrame cam,
                                                int frame, float *__restrict__ albedo,
                                                GT *__restrict__ normals, GT *__restrict__ positions,
                                                GT *__restrict__ noisy) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= scene->width || y >= scene->height) return;
    float px[12];
    scene_pixel(*scene, cam, frame, x, y, px);
    const int64_t lp = (int64_t)y * scene->width + x;
    albedo[lp * 3 + 0] = px[0];
    albedo[lp * 3 + 1] = px[1];
    albedo[lp * 3 + 2] = px[2];
    store3(normals, lp, make_float3(px[3], px[4], px[5]));
    store3(positions, lp, make_float3(px[6], px[7], px[8]));
    store3(noisy, lp, make_float3(px[9], px[10], px[11]));
}

hipError_t launch_render(const SceneData *d_scene, int width, int height, const CameraFrame &cam,
                         int frame, bool f16, void *albedo, void *normals, void *positions,
                         void *noisy, hipStream_t s) {
    dim3 grid((width + 63) / 64, (height + 3) / 4);
    if (f16)
        hipLaunchKernelGGL((k_render<__half>), grid, dim3(256), 0, s, d_scene, cam, frame,
                           (float *)albedo, (__half *)normals, (__half *)positions, (__half *)noisy);
    else
        hipLaunchKernelGGL((k_render<float>), grid, dim3(256), 0, s, d_scene, cam, frame,
                           (float *)albedo, (float *)normals, (float *)positions, (float *)noisy);
    return hipGetLastError();
}

}  // namespace bmfr
