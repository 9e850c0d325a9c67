// Seeded procedural scene: ground plane y = 0 (|x|, |z| <= 64) and up to 64
// spheres, a pinhole camera and a 1-spp "noisy" shading, producing the four
// per-frame inputs the reference loads from EXR files (albedo, shading
// normal, world position, noisy colour; opencl/bmfr.cpp:145-172, 420-427)
// plus the per-frame view-projection matrix and pixel offset that the
// reference reads from the dataset's camera_matrices.h (bmfr.cpp:440-444).
//
// Projection convention (the one stage 1 inverts, opencl/bmfr.cl:343-356):
// the matrix M is used with a row vector, clip_j = sum_k [p,1]_k M[k][j],
// stored row-major (m[4k+j]); pixel (x, y) samples NDC
// ((x + ox)/W*2 - 1, (y + 1 - oy)/H*2 - 1) with row y growing with NDC y,
// so a static camera reprojects every pixel onto itself.
//
// The same code runs on the host (tests) and on the device (renderer
// stand-in writing straight into a stream's input buffers).
#pragma once

#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#define BMFR_HD __host__ __device__
#else
#define BMFR_HD
#endif

namespace bmfr {

constexpr int kMaxSpheres = 64;

struct SceneData {
    int32_t width, height;
    int32_t spheres;
    int32_t static_camera;
    float firefly_probability;
    uint32_t seed_lo, seed_hi;
    float center[kMaxSpheres][3];
    float radius[kMaxSpheres];
    float albedo[kMaxSpheres + 1][3];  // last entry: the plane
};

struct CameraFrame {
    float pos[3], right[3], up[3], fwd[3];
    float tan_half, aspect;
    float offset[2];
    float vp[16];
};

BMFR_HD inline uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Counter-based uniform in (0, 1): hash of (seed, frame, pixel, stream).
BMFR_HD inline float uniform01(uint64_t seed, uint32_t frame, uint32_t pixel, uint32_t k) {
    uint64_t s = seed ^ ((uint64_t)frame << 40) ^ ((uint64_t)pixel << 4) ^ (uint64_t)k;
    s = splitmix64(s);
    return ((float)(uint32_t)(s >> 40) + 0.5f) * (1.0f / 16777216.0f);
}

inline void scene_init(SceneData &d, uint64_t seed, int spheres, float firefly, int w, int h,
                       int static_camera) {
    d.width = w;
    d.height = h;
    d.spheres = spheres < 0 ? 0 : (spheres > kMaxSpheres ? kMaxSpheres : spheres);
    d.static_camera = static_camera;
    d.firefly_probability = firefly;
    d.seed_lo = (uint32_t)seed;
    d.seed_hi = (uint32_t)(seed >> 32);
    uint64_t s = seed;
    auto u = [&](float lo, float hi) {
        return lo + (hi - lo) * (float)((splitmix64(s) >> 40) * (1.0 / 16777216.0));
    };
    for (int i = 0; i < d.spheres; ++i) {
        d.radius[i] = u(0.5f, 2.0f);
        d.center[i][0] = u(-6.f, 6.f);
        d.center[i][1] = u(0.f, 3.f);
        d.center[i][2] = u(-6.f, 6.f);
    }
    for (int i = 0; i <= d.spheres; ++i)
        for (int c = 0; c < 3; ++c) d.albedo[i == d.spheres ? kMaxSpheres : i][c] = u(0.1f, 0.9f);
}

BMFR_HD inline float halton(int index, int base) {
    float f = 1.f, r = 0.f;
    while (index > 0) {
        f = f / (float)base;
        r = r + f * (float)(index % base);
        index = index / base;
    }
    return r;
}

BMFR_HD inline void normalize3(float v[3]) {
    const float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= l;
    v[1] /= l;
    v[2] /= l;
}

BMFR_HD inline float dot3(const float a[3], const float b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Camera of frame f: position (0.05 f, 2, -10), initially looking at the
// origin, yawed by 0.3 f degrees; 60 degree vertical FOV.
BMFR_HD inline void scene_camera(const SceneData &d, int frame, CameraFrame &c) {
    const int f = d.static_camera ? 0 : frame;
    c.pos[0] = 0.05f * (float)f;
    c.pos[1] = 2.f;
    c.pos[2] = -10.f;
    float f0[3] = {0.f, -2.f, 10.f};
    normalize3(f0);
    const float th = 0.3f * (float)f * 3.14159265358979f / 180.f;
    const float ct = cosf(th), st = sinf(th);
    c.fwd[0] = f0[0] * ct + f0[2] * st;
    c.fwd[1] = f0[1];
    c.fwd[2] = -f0[0] * st + f0[2] * ct;
    normalize3(c.fwd);
    // right = up_world x fwd, up = fwd x right
    c.right[0] = c.fwd[2];
    c.right[1] = 0.f;
    c.right[2] = -c.fwd[0];
    normalize3(c.right);
    c.up[0] = c.fwd[1] * c.right[2] - c.fwd[2] * c.right[1];
    c.up[1] = c.fwd[2] * c.right[0] - c.fwd[0] * c.right[2];
    c.up[2] = c.fwd[0] * c.right[1] - c.fwd[1] * c.right[0];
    c.tan_half = tanf(30.f * 3.14159265358979f / 180.f);
    c.aspect = (float)d.width / (float)d.height;
    if (d.static_camera) {
        c.offset[0] = 0.5f;
        c.offset[1] = 0.5f;
    } else {
        c.offset[0] = halton(frame + 1, 2);
        c.offset[1] = halton(frame + 1, 3);
    }
    const float sx = 1.f / (c.tan_half * c.aspect), sy = 1.f / c.tan_half;
    const float zn = 0.1f, zf = 1000.f;
    const float A = zf / (zf - zn), B = -zn * zf / (zf - zn);
    for (int k = 0; k < 3; ++k) {
        c.vp[k * 4 + 0] = c.right[k] * sx;
        c.vp[k * 4 + 1] = c.up[k] * sy;
        c.vp[k * 4 + 2] = c.fwd[k] * A;
        c.vp[k * 4 + 3] = c.fwd[k];
    }
    c.vp[12] = -dot3(c.pos, c.right) * sx;
    c.vp[13] = -dot3(c.pos, c.up) * sy;
    c.vp[14] = -dot3(c.pos, c.fwd) * A + B;
    c.vp[15] = -dot3(c.pos, c.fwd);
}

// One pixel: albedo, normal, position, noisy colour (12 floats).
BMFR_HD inline void scene_pixel(const SceneData &d, const CameraFrame &c, int frame, int x, int y,
                                float out[12]) {
    const float nx = ((float)x + c.offset[0]) / (float)d.width * 2.f - 1.f;
    const float ny = ((float)y + 1.f - c.offset[1]) / (float)d.height * 2.f - 1.f;
    float dir[3];
    for (int k = 0; k < 3; ++k)
        dir[k] = c.fwd[k] + nx * c.tan_half * c.aspect * c.right[k] + ny * c.tan_half * c.up[k];
    normalize3(dir);
    float best = 1e30f;
    int hit = -1;
    // ground plane y = 0, bounded to |x|, |z| <= 64
    if (dir[1] < 0.f) {
        const float t = -c.pos[1] / dir[1];
        const float hx = c.pos[0] + t * dir[0], hz = c.pos[2] + t * dir[2];
        if (t > 1e-3f && fabsf(hx) <= 64.f && fabsf(hz) <= 64.f) {
            best = t;
            hit = kMaxSpheres;
        }
    }
    for (int i = 0; i < d.spheres; ++i) {
        float oc[3] = {c.pos[0] - d.center[i][0], c.pos[1] - d.center[i][1], c.pos[2] - d.center[i][2]};
        const float b = dot3(oc, dir);
        const float cc = dot3(oc, oc) - d.radius[i] * d.radius[i];
        const float disc = b * b - cc;
        if (disc > 0.f) {
            const float sq = sqrtf(disc);
            float t = -b - sq;
            if (t <= 1e-3f) t = -b + sq;
            if (t > 1e-3f && t < best) {
                best = t;
                hit = i;
            }
        }
    }
    for (int k = 0; k < 12; ++k) out[k] = 0.f;
    if (hit < 0) return;  // miss: every buffer is 0
    float p[3], n[3];
    for (int k = 0; k < 3; ++k) p[k] = c.pos[k] + best * dir[k];
    if (hit == kMaxSpheres) {
        p[1] = 0.f;
        n[0] = 0.f;
        n[1] = 1.f;
        n[2] = 0.f;
    } else {
        for (int k = 0; k < 3; ++k) n[k] = p[k] - d.center[hit][k];
        normalize3(n);
    }
    float l[3] = {0.4f, 1.f, -0.3f};
    normalize3(l);
    const float ndl = fmaxf(0.f, dot3(n, l));
    const uint64_t seed = ((uint64_t)d.seed_hi << 32) | d.seed_lo;
    const uint32_t pix = (uint32_t)(y * d.width + x);
    const float fire = uniform01(seed, (uint32_t)frame, pix, 7) < d.firefly_probability ? 50.f : 1.f;
    for (int ch = 0; ch < 3; ++ch) {
        const float u1 = uniform01(seed, (uint32_t)frame, pix, (uint32_t)(2 * ch));
        const float u2 = uniform01(seed, (uint32_t)frame, pix, (uint32_t)(2 * ch + 1));
        const float g = sqrtf(-2.f * logf(u1)) * cosf(6.28318530717959f * u2);  // N(0,1)
        const float alb = d.albedo[hit][ch];
        out[ch] = alb;
        out[9 + ch] = alb * ndl * expf(g) * fire;  // lognormal(0, 1) shading noise
    }
    for (int k = 0; k < 3; ++k) {
        out[3 + k] = n[k];
        out[6 + k] = p[k];
    }
}

}  // namespace bmfr
