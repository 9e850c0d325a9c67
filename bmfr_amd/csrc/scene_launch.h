#pragma once
#include <hip/hip_runtime.h>

#include "scene.h"

namespace bmfr {
hipError_t launch_render(const SceneData *d_scene, int width, int height, const CameraFrame &cam,
                         int frame, bool f16, void *albedo, void *normals, void *positions,
                         void *noisy, hipStream_t s);
}
