// Stage 1's per-work-item body, shared by the stage 1 kernel and the fused
// stage 1 + fitter kernel.
#pragma once
#include "device_common.h"
#include "kernels.h"

namespace bmfr {

// What stage 1 hands to the block matrix for one work-item: the pixel's
// world position and normal (the variables of the feature expressions,
// bmfr.cl:448-453) and its accumulated noisy colour (the last three
// columns).
struct Stage1Pixel {
    float3 wp, nrm, col;
};

// Stage 1: accumulate_noisy_data, opencl/bmfr.cl:287-485, for one
// work-item (gx, gy) of the padded grid: reprojection, discard, spp and the
// blend, the pixel's own outputs stored (by the work-item that owns the
// pixel), and what its block-matrix row is made of returned.
template <typename GT>
__device__ __forceinline__ Stage1Pixel stage1_pixel(const StageArgs &a, const StageBuffers &bufs, int gx, int gy) {
    const GT *__restrict__ cur_normals = (const GT *)bufs.normals_cur;
    const GT *__restrict__ prev_normals = (const GT *)bufs.normals_prev;
    const GT *__restrict__ cur_positions = (const GT *)bufs.positions_cur;
    const GT *__restrict__ prev_positions = (const GT *)bufs.positions_prev;
    const GT *__restrict__ cur_noisy = (const GT *)bufs.noisy_in;
    const GT *__restrict__ prev_noisy = (const GT *)bufs.noisy_acc_prev;
    const uint8_t *__restrict__ prev_spp = (const uint8_t *)bufs.spp_prev;
    GT *__restrict__ acc_noisy = (GT *)bufs.noisy_acc_cur;
    float2 *__restrict__ out_prev_pixel = (float2 *)bufs.prev_pixel;
    uint8_t *__restrict__ accept_bools = (uint8_t *)bufs.accept;
    uint8_t *__restrict__ cur_spp = (uint8_t *)bufs.spp_cur;
    const int W = a.width, H = a.height;
    const int2 off = block_offset(a.frame);
    const int pwx = gx - kBlock / 2 + off.x;
    const int pwy = gy - kBlock / 2 + off.y;
    const int px = mirror(pwx, W), py = mirror(pwy, H);
    const uint32_t lp = __umul24((unsigned)py, (unsigned)W) + (unsigned)px;

    const float3 wp = load3_px(cur_positions, lp);
    const float3 nrm = load3_px(cur_normals, lp);
    const float3 cur = load3_px(cur_noisy, lp);

    float ppf_x = (float)px, ppf_y = (float)py;
    uint32_t store_accept = 0;
    float blend_alpha = 1.f;
    float3 prev_col = make_float3(0.f, 0.f, 0.f);
    float sample_spp = 0.f;
    if (a.frame > 0) {
        const float *m = a.matrix;
        // row-vector matrix, dot() left to right, w = 1 (bmfr.cl:343-356)
        float uvx = ((m[0] * wp.x + m[4] * wp.y) + m[8] * wp.z) + m[12] * 1.f;
        float uvy = ((m[1] * wp.x + m[5] * wp.y) + m[9] * wp.z) + m[13] * 1.f;
        const float ww = ((m[3] * wp.x + m[7] * wp.y) + m[11] * wp.z) + m[15] * 1.f;
        const double rww = quotient_reciprocal(ww);   // uv / ww with IEEE bits
        uvx = quotient(uvx, rww);
        uvy = quotient(uvy, rww);
        uvx = uvx + 1.f;
        uvy = uvy + 1.f;
        uvx = uvx / 2.f;
        uvy = uvy / 2.f;
        ppf_x = uvx * (float)W;
        ppf_y = uvy * (float)H;
        ppf_x = ppf_x - a.pixel_offset_x;
        ppf_y = ppf_y - (1.f - a.pixel_offset_y);
        const Footprint fp = footprint(make_float2(ppf_x, ppf_y), W, H);
        const int ipx = fp.ix, ipy = fp.iy;
        const float fx = ppf_x - (float)ipx, fy = ppf_y - (float)ipy;
        const float omx = 1.f - fx, omy = 1.f - fy;
        const float wts[4] = {omx * omy, fx * omy, omx * fy, fx * fy};
        // All four taps' history is loaded up front (one memory round trip
        // instead of three dependent ones: position, then normal, then
        // colour and spp); a tap outside the image reads the texel clamped
        // into the image and is never used.  The tests and sums below keep
        // the reference's order (bmfr.cl:374-419).
        bool tap_in[4];
        uint32_t tap_ls[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            tap_in[i] = (fp.in_image >> i) & 1u;
            tap_ls[i] = fp.idx[i];
        }
        float3 tpp[4], tpn[4], tpc[4];
        float tsp[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            tpp[i] = load3_px(prev_positions, tap_ls[i]);
            tpn[i] = load3_px(prev_normals, tap_ls[i]);
            tpc[i] = load3_px(prev_noisy, tap_ls[i]);
            tsp[i] = (float)gload_px(prev_spp, tap_ls[i]);
        }
        // The tests are evaluated for every tap and the sums selected rather
        // than branched over (the same values, in the reference's order):
        // with a branch per tap the compiler moves that tap's loads into it,
        // after a wait for every load in flight, one round trip per test.
        float total_weight = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float3 pp = tpp[i];
            const float dx = pp.x - wp.x, dy = pp.y - wp.y, dz = pp.z - wp.z;
            const float pd = (dx * dx + dy * dy) + dz * dz;
            const float3 pn = tpn[i];
            const float nx = pn.x - nrm.x, ny = pn.y - nrm.y, nz = pn.z - nrm.z;
            const float nd = (nx * nx + ny * ny) + nz * nz;
            const bool ok = tap_in[i] & (pd < a.position_limit_squared) & (nd < a.normal_limit_squared);
            const float3 pc = tpc[i];
            store_accept |= ok ? 1u << i : 0u;
            sample_spp = ok ? sample_spp + wts[i] * tsp[i] : sample_spp;
            prev_col.x = ok ? prev_col.x + wts[i] * pc.x : prev_col.x;
            prev_col.y = ok ? prev_col.y + wts[i] * pc.y : prev_col.y;
            prev_col.z = ok ? prev_col.z + wts[i] * pc.z : prev_col.z;
            total_weight = ok ? total_weight + wts[i] : total_weight;
        }
        if (total_weight > 0.f) {  // bmfr.cl:421-429
            const double rw = quotient_reciprocal(total_weight);   // IEEE bits of x / total_weight
            prev_col.x = quotient(prev_col.x, rw);
            prev_col.y = quotient(prev_col.y, rw);
            prev_col.z = quotient(prev_col.z, rw);
            sample_spp = quotient(sample_spp, rw);
            blend_alpha = quotient(1.f, quotient_reciprocal(sample_spp + 1.f));
            blend_alpha = fmaxf(blend_alpha, a.blend_alpha);
        }
    }
    // bmfr.cl:433-442
    uint32_t new_spp = 1;
    if (blend_alpha < 1.f) {
        if (sample_spp > 254.f) {
            new_spp = 255;
        } else {
            float r = __builtin_rintf(sample_spp);
            r = fminf(fmaxf(r, 0.f), 255.f);
            new_spp = (uint32_t)r + 1u;
        }
    }
    const float omb = 1.f - blend_alpha;
    float3 nc;
    nc.x = blend_alpha * cur.x + omb * prev_col.x;
    nc.y = blend_alpha * cur.y + omb * prev_col.y;
    nc.z = blend_alpha * cur.z + omb * prev_col.z;

    // bmfr.cl:478-484 (and the spp of bmfr.cl:442, which mirrored
    // work-items would only rewrite with the same value)
    if (pwx >= 0 && pwx < W && pwy >= 0 && pwy < H) {
        store3_px(acc_noisy, lp, nc);
        gstore_px(out_prev_pixel, lp, make_float2(ppf_x, ppf_y));
        gstore_px(accept_bools, lp, (uint8_t)store_accept);
        gstore_px(cur_spp, lp, (uint8_t)new_spp);
    }
    return Stage1Pixel{wp, nrm, nc};
}

// The reference's 13 buffers of a pixel (bmfr.cpp:65-77 as bmfr.cl:448-453
// evaluates them): 1, normal, world position, its squares, colour.
__device__ __forceinline__ void default_features(const Stage1Pixel &q, float (&feat)[kBufferCount]) {
    const float3 wp = q.wp, nrm = q.nrm, nc = q.col;
    const float fv[kBufferCount] = {1.f,  nrm.x, nrm.y,       nrm.z,       wp.x,        wp.y, wp.z,
                                    wp.x * wp.x, wp.y * wp.y, wp.z * wp.z, nc.x, nc.y,  nc.z};
#pragma unroll
    for (int f = 0; f < kBufferCount; ++f) feat[f] = fv[f];
}


// A feature as stored in the block matrix (bmfr.cl:466-476): NaN becomes 0;
// binary16 storage clamps to +-65504 and rounds to nearest even.
template <typename TT>
__device__ __forceinline__ TT tmp_element(float v) {
    if (v != v) v = 0.f;
    if constexpr (sizeof(TT) == 2) {
        v = fmaxf(fminf(v, 65504.f), -65504.f);
        return __float2half_rn(v);
    } else {
        return v;
    }
}

}  // namespace bmfr
