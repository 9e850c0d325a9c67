// Stages 1, 3, 4 and 5 of the BMFR frame chain for gfx950: per-pixel,
// HBM-streaming kernels.  Each follows the operation order of the
// corresponding reference kernel so that, with contraction off, its results
// are reproducible bit for bit (the tone map's power aside: within one ulp
// of the correctly rounded value, see tone_curve.h).
//
// Launch shapes: stage 1 runs over the padded grid in 32x8 tiles so that a
// wave writes whole 32-pixel rows of one block of tmp_data (block-major,
// [by][bx][feature][y][x]); stages 3-5 run over the image in 64x4 tiles so
// that a wave reads 64 consecutive AoS pixels (768 contiguous bytes).
#include <type_traits>

#include "device_common.h"
#include "kernels.h"
#include "stage1.h"
#include "surface.h"
#include "tone_curve.h"

namespace bmfr {

// ---------------------------------------------------------------------------
// Stage 1: accumulate_noisy_data, opencl/bmfr.cl:287-485.
// One work-item of the padded grid (gx, gy); `lin` is its linear index,
// used by the first (N-1)*1024 work-items (N features) to also write the
// frame's noise table.  Called by all 256 work-items of a 32x8 tile
// (gx = 32 tx + t % 32, gy = 8 ty + t / 32, t = threadIdx.x): the body has a
// barrier.  CUSTOM evaluates the feature set in `a.features` (N + 3
// columns) instead of the reference's 13.
template <typename GT, typename TT, bool CUSTOM = false>
__device__ __forceinline__ void stage1_body(const StageArgs &a, const StageBuffers &bufs, int gx, int gy,
                                            int lin, bool write_noise) {
    TT *__restrict__ tmp_data = (TT *)bufs.tmp_data;
    const int nf = CUSTOM ? a.features.count : kFeatures;
    const int ncol = nf + 3;
    if constexpr (CUSTOM) {
        // up to 15 x 1024 terms; the padded grid has at least 96x96 work-items
        if (write_noise) {
            const int nt = (nf - 1) * kBlockPixels, stride = a.margin_width * a.margin_height;
            for (int i = lin; i < nt; i += stride)
                gstore((float *)bufs.noise, i, noise_random(i & (kBlockPixels - 1), 1 + (i >> 10), a.frame, ncol));
        }
    } else if (write_noise && lin < kNoiseTable) {
        // the padded grid has at least 96x96 = 9216 work-items (W, H >= 47)
        gstore((float *)bufs.noise, lin, noise_random(lin & (kBlockPixels - 1), 1 + (lin >> 10), a.frame));
    }
    const Stage1Pixel q = stage1_pixel<GT>(a, bufs, gx, gy);
    // The work-group's 32x8 tile is rows yin0..yin0+7 of one block, and
    // those rows of a feature are 256 contiguous elements of tmp_data: the
    // features go through LDS and out as 16-byte stores (1.6 per work-item
    // instead of 13 narrow ones).
    __shared__ __align__(16) TT feat_tile[CUSTOM ? kMaxColumns : kBufferCount][256];
    if constexpr (CUSTOM) {
        for (int f = 0; f < nf; ++f)
            feat_tile[f][threadIdx.x] = tmp_element<TT>(feature_value(a.features, f, q.nrm, q.wp));
        feat_tile[nf][threadIdx.x] = tmp_element<TT>(q.col.x);
        feat_tile[nf + 1][threadIdx.x] = tmp_element<TT>(q.col.y);
        feat_tile[nf + 2][threadIdx.x] = tmp_element<TT>(q.col.z);
    } else {
        float feat[kBufferCount];
        default_features(q, feat);
#pragma unroll
        for (int f = 0; f < kBufferCount; ++f) feat_tile[f][threadIdx.x] = tmp_element<TT>(feat[f]);
    }
    __syncthreads();
    {
        const int yin0 = (gy - (int)(threadIdx.x >> 5)) & (kBlock - 1);
        const int64_t blk = (int64_t)(gy / kBlock) * a.blocks_x + gx / kBlock;
        TT *dst = tmp_data + blk * (ncol * kBlockPixels) + yin0 * kBlock;
        constexpr int kVec = 16 / (int)sizeof(TT);             // elements per store
        constexpr int kPerFeature = 256 / kVec;
        for (int c = threadIdx.x; c < ncol * kPerFeature; c += 256) {
            const int f = c / kPerFeature, e = (c - f * kPerFeature) * kVec;
            gstore(reinterpret_cast<uint4 *>(dst + f * kBlockPixels + e), 0,
                   *reinterpret_cast<const uint4 *>(&feat_tile[f][e]));
        }
    }
}

template <typename GT, typename TT, bool CUSTOM>
__global__ __launch_bounds__(256) void k_accumulate_noisy_data(const StageArgs a, const StageBuffers b,
                                                               int write_noise) {
    const int tiles_x = a.margin_width / 32;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
    stage1_body<GT, TT, CUSTOM>(a, b, tx * 32 + (threadIdx.x & 31), ty * 8 + (threadIdx.x >> 5),
                                tile * 256 + threadIdx.x, write_noise != 0);
}

// 64x4 image tile of a single-stream launch, XCD-contiguous.
__device__ __forceinline__ void image_tile(const StageArgs &a, int &x, int &y) {
    const int tiles_x = (a.width + 63) / 64;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
    x = tx * 64 + (threadIdx.x & 63);
    y = ty * 4 + (threadIdx.x >> 6);
}

// ---------------------------------------------------------------------------
// Stage 3: weighted_sum, opencl/bmfr.cl:703-758.
// Launch shape: the padded grid's 32x8 tiles, four per block, so that every
// pixel of a work-group belongs to block g (uniform): its 30 weights and 12
// bounds come through scalar loads instead of per-lane vector loads.  Tile
// pixel (gx, gy) of the padded grid is image pixel
// (gx - 16 + offset.x, gy - 16 + offset.y) (bmfr.cl:707-710 inverted);
// margin positions outside the image do nothing.
__device__ __forceinline__ void block_tile(const StageArgs &a, int local, int &x, int &y, int64_t &g) {
    const int tiles_x = a.margin_width / 32;
    const int ty = local / tiles_x, tx = local - ty * tiles_x;
    const int2 off = block_offset(a.frame);
    x = tx * 32 + (threadIdx.x & 31) - kBlock / 2 + off.x;
    y = ty * 8 + (threadIdx.x >> 5) - kBlock / 2 + off.y;
    g = (int64_t)(ty >> 2) * a.blocks_x + tx;
}

// The block's feature-scaling reciprocals for the reference's feature set,
// once per wave: lane l < 6 forms the FP64 reciprocal of max - min of scaled
// feature l (quotient_reciprocal), and the six are read from those lanes
// (readlane, SGPR pairs) right there, so that a scaled feature costs three
// instructions instead of an IEEE division, with the same bits (quotient).
// Built where every lane of the wave is active and the block is the same for
// the whole wave.  The reads happen here, not at the use: a readlane inside
// the per-pixel branch would let the compiler sink the reciprocal's
// arithmetic into that branch, where lanes outside the image do not compute
// it.  Configured feature sets (a run-time feature count) keep the division.
struct ScaleRecips {
    double s[kScaled];   // wave-uniform
};

__device__ __forceinline__ double lane_value(double y, int l) {
    unsigned long long b;
    __builtin_memcpy(&b, &y, sizeof b);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    const unsigned long long r = ((unsigned long long)hi << 32) | lo;
    double d;
    __builtin_memcpy(&d, &r, sizeof d);
    return d;
}

__device__ __forceinline__ ScaleRecips scale_recips(const StageBuffers &bufs, int64_t g) {
    const float *mm = (const float *)bufs.mins_maxs + g * (kScaled * 2);
    const int l = (int)(threadIdx.x & 63);
    // the shift-only case (|max - min| <= 1, or NaN) as a division by 1:
    // quotient(x, 1.0) is x, so the feature needs no branch and no max
    float d = 1.f;
    if (l < kScaled) d = gload(mm, l * 2 + 1) - gload(mm, l * 2);
    if (!(fabsf(d) > 1.0f)) d = 1.f;
    const double y = quotient_reciprocal(d);
    ScaleRecips r;
#pragma unroll
    for (int i = 0; i < kScaled; ++i) r.s[i] = lane_value(y, i);
    return r;
}

// opencl/bmfr.cl:200-205 with the divisor's reciprocal (1 for the shift-only
// case, scale_recips): the bits of scale_feature
__device__ __forceinline__ float scale_feature_r(float v, float mn, double y) { return quotient(v - mn, y); }

// The weights and scaling bounds of one block for the reference's feature
// set, read once: a kernel that stores before it has used them all would
// otherwise read them again for every pixel (the compiler cannot tell that
// its stores leave them unchanged).
struct BlockFit {
    float w[kFeatures * 3];
    float mn[kScaled], mx[kScaled];
};

__device__ __forceinline__ BlockFit block_fit(const StageBuffers &bufs, int64_t g) {
    const float *w = (const float *)bufs.weights + g * (kFeatures * 3);
    const float *mm = (const float *)bufs.mins_maxs + g * (kScaled * 2);
    BlockFit f;
#pragma unroll
    for (int i = 0; i < kFeatures * 3; ++i) f.w[i] = gload(w, i);
#pragma unroll
    for (int i = 0; i < kScaled; ++i) {
        f.mn[i] = gload(mm, i * 2);
        f.mx[i] = gload(mm, i * 2 + 1);
    }
    return f;
}

// Filtered colour of a pixel with world position wp and normal n from its
// block's fit: sum of weights times the (scaled) features, negatives clamped
// (bmfr.cl:712-750).  With `rec` (the wave's reciprocals for the block) the
// scaling divisions become products.
__device__ __forceinline__ float3 stage3_color_fit(float3 wp, float3 n, const BlockFit &fit,
                                                   const ScaleRecips *rec) {
    const float feat[kFeatures] = {1.f,  n.x,  n.y,         n.z,         wp.x,
                                   wp.y, wp.z, wp.x * wp.x, wp.y * wp.y, wp.z * wp.z};
    float cx = 0.f, cy = 0.f, cz = 0.f;
#pragma unroll
    for (int f = 0; f < kFeatures; ++f) {
        float v = feat[f];
        if (f >= kNotScaled) {
            const int i = f - kNotScaled;
            v = rec ? scale_feature_r(v, fit.mn[i], rec->s[i]) : scale_feature(v, fit.mn[i], fit.mx[i]);
        }
        // color += weight * feature, one FMA a term: the reference's
        // expression (bmfr.cl:745) under OpenCL's default FP_CONTRACT ON
        cx = __builtin_fmaf(fit.w[f * 3 + 0], v, cx);
        cy = __builtin_fmaf(fit.w[f * 3 + 1], v, cy);
        cz = __builtin_fmaf(fit.w[f * 3 + 2], v, cz);
    }
    return make_float3(cx < 0.f ? 0.f : cx, cy < 0.f ? 0.f : cy, cz < 0.f ? 0.f : cz);
}

// Filtered colour of in-image pixel lp in block g (bmfr.cl:712-750).
// CUSTOM evaluates the feature set in `a.features`.  With `rec` (the wave's
// reciprocals for block g, reference set only) the scaling divisions become
// products.
template <typename GT, bool CUSTOM = false>
__device__ __forceinline__ float3 stage3_color_of(const StageArgs &a, const StageBuffers &bufs, float3 wp, float3 n,
                                                 int64_t g, const ScaleRecips *rec = nullptr) {
    if constexpr (CUSTOM) {
        const float *__restrict__ weights = (const float *)bufs.weights;
        const float *__restrict__ mins_maxs = (const float *)bufs.mins_maxs;
        float cx = 0.f, cy = 0.f, cz = 0.f;
        const FeatureSet &fs = a.features;
        const int nf = fs.count, ns = fs.count - fs.not_scaled;
        const float *w = weights + g * (nf * 3);
        const float *mm = mins_maxs + g * (ns * 2);
        for (int f = 0; f < nf; ++f) {
            float v = feature_value(fs, f, n, wp);
            if (f >= fs.not_scaled) {
                const int i = f - fs.not_scaled;
                v = scale_feature(v, gload(mm, i * 2), gload(mm, i * 2 + 1));
            }
            cx = __builtin_fmaf(gload(w, f * 3 + 0), v, cx);
            cy = __builtin_fmaf(gload(w, f * 3 + 1), v, cy);
            cz = __builtin_fmaf(gload(w, f * 3 + 2), v, cz);
        }
        return make_float3(cx < 0.f ? 0.f : cx, cy < 0.f ? 0.f : cy, cz < 0.f ? 0.f : cz);
    } else {
        return stage3_color_fit(wp, n, block_fit(bufs, g), rec);
    }
}

template <typename GT, bool CUSTOM = false>
__device__ __forceinline__ float3 stage3_color(const StageArgs &a, const StageBuffers &bufs, int64_t lp,
                                              int64_t g, const ScaleRecips *rec = nullptr) {
    return stage3_color_of<GT, CUSTOM>(a, bufs, load3_px((const GT *)bufs.positions_cur, (uint32_t)lp),
                                       load3_px((const GT *)bufs.normals_cur, (uint32_t)lp), g, rec);
}

template <typename GT, bool CUSTOM = false>
__device__ __forceinline__ void stage3_body(const StageArgs &a, const StageBuffers &bufs, int local) {
    int x, y;
    int64_t g;
    block_tile(a, local, x, y, g);
    ScaleRecips rec;
    if constexpr (!CUSTOM) rec = scale_recips(bufs, g);   // all lanes active; g is the work-group's
    if (x < 0 || y < 0 || x >= a.width || y >= a.height) return;
    const int64_t lp = __umul24((unsigned)y, (unsigned)a.width) + (unsigned)x;
    const float3 c = stage3_color<GT, CUSTOM>(a, bufs, lp, g, CUSTOM ? nullptr : &rec);
    float *__restrict__ output = (float *)bufs.filtered;
    store3_px(output, (uint32_t)lp, c);
}

template <typename GT, bool CUSTOM>
__global__ __launch_bounds__(256) void k_weighted_sum(const StageArgs a, const StageBuffers b) {
    stage3_body<GT, CUSTOM>(a, b, xcd_tile(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------------------
// Stage 4: accumulate_filtered_data, opencl/bmfr.cl:761-857.
// Split in two so that the history taps are in flight while stage 3's sum is
// computed: stage4_gather issues every load of pixel lp (a tap that the
// accept bits or the range check rule out reads the pixel's own texel and
// is never used), stage4_finish does the reference's arithmetic in its order.
template <typename GT>
struct Stage4Gather {
    float3 tap[4];
    float3 albedo;
    float wts[4];
    uint32_t use;    // bit i: tap i contributes
    uint32_t spp;
    float2 pf;       // prev_pixel (stage 5 reads it too)
};

// Stage 4's per-pixel inputs (the first memory round trip): stage 1's accept
// bits, spp and prev_pixel, and the albedo.
struct Stage4Inputs {
    float3 albedo;
    float2 pf;
    uint32_t acc;
    uint32_t spp;
};

__device__ __forceinline__ Stage4Inputs stage4_inputs(const StageArgs &a, const StageBuffers &bufs, int64_t lp) {
    const float2 *__restrict__ prev_pixel = (const float2 *)bufs.prev_pixel;
    const uint8_t *__restrict__ accept_bools = (const uint8_t *)bufs.accept;
    const float *__restrict__ albedo = (const float *)bufs.albedo;
    const uint8_t *__restrict__ cur_spp = (const uint8_t *)bufs.spp_cur;
    const uint32_t px = (uint32_t)lp;
    Stage4Inputs in;
    in.acc = a.frame > 0 ? (uint32_t)gload_px(accept_bools, px) : 0u;
    in.pf = gload_px(prev_pixel, px);
    in.spp = gload_px(cur_spp, px);
    in.albedo = load3_px(albedo, px);
    return in;
}

// Byte offsets of the four texels of footprint `fp` in a three-element
// buffer of T.
template <typename T>
__device__ __forceinline__ void footprint_offsets(const Footprint &fp, uint32_t (&off)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) off[i] = px3_bytes<T>(fp.idx[i]);
}

// The history taps of pixel lp (the second round trip, addressed by the
// inputs).
template <typename GT>
__device__ __forceinline__ Stage4Gather<GT> stage4_taps(const StageArgs &a, const StageBuffers &bufs, int64_t lp,
                                                        const Stage4Inputs &in, const Footprint &fp) {
    const GT *__restrict__ acc_prev = (const GT *)bufs.out_prev;
    Stage4Gather<GT> g;
    const float2 pf = in.pf;
    g.pf = pf;
    g.spp = in.spp;
    g.albedo = in.albedo;
    const float fx = pf.x - (float)fp.ix, fy = pf.y - (float)fp.iy;
    const float omx = 1.f - fx, omy = 1.f - fy;
    g.wts[0] = omx * omy;
    g.wts[1] = fx * omy;
    g.wts[2] = omx * fy;
    g.wts[3] = fx * fy;
    // Stage 1 sets a bit only for an in-image tap at the same floor();
    // requiring the texel to be in the image keeps inconsistent inputs (e.g.
    // written through the debug API) from reading outside.
    g.use = in.acc & fp.in_image;
    uint32_t off[4];
    footprint_offsets<GT>(fp, off);
#pragma unroll
    for (int i = 0; i < 4; ++i) g.tap[i] = load3_at(acc_prev, off[i]);
    return g;
}

template <typename GT>
__device__ __forceinline__ Stage4Gather<GT> stage4_gather(const StageArgs &a, const StageBuffers &bufs, int64_t lp) {
    const Stage4Inputs in = stage4_inputs(a, bufs, lp);
    return stage4_taps<GT>(a, bufs, lp, in, footprint(in.pf, a.width, a.height));
}

// The accumulated colour of a pixel from its filtered colour `fc` and its
// gathered history (bmfr.cl:801-849).
template <typename GT>
__device__ __forceinline__ float3 stage4_blend(const StageArgs &a, float3 fc, const Stage4Gather<GT> &g) {
    float3 pc = make_float3(0.f, 0.f, 0.f);
    float blend_alpha = 1.f;
    if (g.use) {
        float total_weight = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (g.use & (1u << i)) {
                total_weight = total_weight + g.wts[i];
                pc.x = pc.x + g.wts[i] * g.tap[i].x;
                pc.y = pc.y + g.wts[i] * g.tap[i].y;
                pc.z = pc.z + g.wts[i] * g.tap[i].z;
            }
        }
        if (total_weight > 0.f) {
            blend_alpha = quotient(1.f, quotient_reciprocal((float)g.spp));   // 1.f / spp
            blend_alpha = fmaxf(blend_alpha, a.second_blend_alpha);
            const double rw = quotient_reciprocal(total_weight);   // pc / total_weight, IEEE bits
            pc.x = quotient(pc.x, rw);
            pc.y = quotient(pc.y, rw);
            pc.z = quotient(pc.z, rw);
        }
    }
    const float omb = 1.f - blend_alpha;
    float3 ac;
    ac.x = blend_alpha * fc.x + omb * pc.x;
    ac.y = blend_alpha * fc.y + omb * pc.y;
    ac.z = blend_alpha * fc.z + omb * pc.z;
    return ac;
}

// clamp(pow(max(0, albedo * colour), 0.454545f), 0, 1), bmfr.cl:851-856
template <typename Tables = ToneTables>
__device__ __forceinline__ float3 tone_map(float3 albedo, float3 ac, const Tables &tab = Tables{}) {
    return make_float3(tone_curve(albedo.x * ac.x, tab), tone_curve(albedo.y * ac.y, tab),
                       tone_curve(albedo.z * ac.z, tab));
}

template <typename GT>
__device__ __forceinline__ void stage4_finish(const StageArgs &a, const StageBuffers &bufs, int64_t lp, float3 fc,
                                              const Stage4Gather<GT> &g) {
    float *__restrict__ tone_mapped = (float *)bufs.tone_mapped;
    const float3 ac = stage4_blend<GT>(a, fc, g);
    store3_px((GT *)bufs.out_cur, (uint32_t)lp, ac);
    store3_px(tone_mapped, (uint32_t)lp, tone_map(g.albedo, ac));
}

template <typename GT>
__device__ __forceinline__ void stage4_body(const StageArgs &a, const StageBuffers &bufs, int x, int y) {
    if (x >= a.width || y >= a.height) return;
    const int64_t lp = __umul24((unsigned)y, (unsigned)a.width) + (unsigned)x;
    const float *__restrict__ filtered = (const float *)bufs.filtered;
    const Stage4Gather<GT> gth = stage4_gather<GT>(a, bufs, lp);
    stage4_finish<GT>(a, bufs, lp,
                      load3_px(filtered, (uint32_t)lp),
                      gth);
}

// Stages 3 and 4 fused for the chained paths, on stage 3's launch shape: the
// filtered colour stays in registers (the `filtered` buffer is not
// written); same arithmetic.
template <typename GT, bool CUSTOM = false>
__device__ __forceinline__ void stage34_body(const StageArgs &a, const StageBuffers &bufs, int local) {
    int x, y;
    int64_t g;
    block_tile(a, local, x, y, g);
    ScaleRecips rec;
    if constexpr (!CUSTOM) rec = scale_recips(bufs, g);   // all lanes active; g is the work-group's
    if (x < 0 || y < 0 || x >= a.width || y >= a.height) return;
    const int64_t lp = __umul24((unsigned)y, (unsigned)a.width) + (unsigned)x;
    const Stage4Gather<GT> gth = stage4_gather<GT>(a, bufs, lp);
    stage4_finish<GT>(a, bufs, lp, stage3_color<GT, CUSTOM>(a, bufs, lp, g, CUSTOM ? nullptr : &rec), gth);
}

template <typename GT, bool CUSTOM>
__global__ __launch_bounds__(256) void k_weighted_sum_filtered(const StageArgs a, const StageBuffers b) {
    stage34_body<GT, CUSTOM>(a, b, xcd_tile(blockIdx.x, gridDim.x));
}

template <typename GT>
__global__ __launch_bounds__(256) void k_accumulate_filtered_data(const StageArgs a, const StageBuffers b) {
    int x, y;
    image_tile(a, x, y);
    stage4_body<GT>(a, b, x, y);
}

// ---------------------------------------------------------------------------
// Stage 5: taa, opencl/bmfr.cl:860-974.
__device__ __forceinline__ float3 rgb_to_ycocg(float3 c) {
    return make_float3((c.x * 1.f + c.y * 2.f) + c.z * 1.f, (c.x * 2.f + c.y * 0.f) + c.z * -2.f,
                       (c.x * -1.f + c.y * 2.f) + c.z * -1.f);
}
__device__ __forceinline__ float3 ycocg_to_rgb(float3 c) {
    return make_float3((c.x * 0.25f + c.y * 0.25f) + c.z * -0.25f,
                       (c.x * 0.25f + c.y * 0.f) + c.z * 0.25f,
                       (c.x * 0.25f + c.y * -0.25f) + c.z * -0.25f);
}

// The 64x4 work-group tile's 3x3 neighbourhoods are read from LDS: the tile
// plus a one-pixel halo (66 x 6 texels) is loaded once and converted to
// YCoCg once per texel (the same operations as per neighbour, so the same
// bits).  Texels outside the image hold NaN: fminf/fmaxf return the other
// operand for a NaN, which is exactly the reference's skipping of
// out-of-image neighbours (bmfr.cl:898-918), and the centre texel is always
// in the image.  Min and max are exact, so their order does not matter.
constexpr int kTaaTileW = 66, kTaaTileH = 6;

// The bilinear previous result at prev_pixel (bmfr.cl:879-890, 922-928):
// its four taps are issued as soon as prev_pixel is known; a tap that the
// tests of bmfr.cl:929-962 rule out reads the pixel's own texel and is never
// used.
template <typename GT>
struct TaaHistory {
    float3 tap[4];
    bool use[4];
    bool history;   // frame > 0 and floor(prev_pixel) in [-1, W) x [-1, H)
    int ix, iy;
};

template <typename GT>
__device__ __forceinline__ TaaHistory<GT> taa_history(const StageArgs &a, const StageBuffers &bufs, int64_t lp,
                                                      const Footprint &fp) {
    const GT *__restrict__ prev_result = (const GT *)bufs.result_prev;
    TaaHistory<GT> h;
    h.ix = fp.ix;
    h.iy = fp.iy;
    // floor(prev_pixel) in [-1, W) x [-1, H) exactly when a texel of the
    // footprint is in the image; each tap is used when its texel is
    const uint32_t use = a.frame != 0 ? fp.in_image : 0u;
    h.history = use != 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) h.use[i] = (use >> i) & 1u;
    uint32_t off[4];
    footprint_offsets<GT>(fp, off);
#pragma unroll
    for (int i = 0; i < 4; ++i) h.tap[i] = load3_at(prev_result, off[i]);
    return h;
}

// Minimum and maximum of three floats as v_min3_f32 / v_max3_f32 (IEEE mode:
// a quiet NaN operand is skipped, -0 orders below +0, so the grouping does
// not change the result).  Written out because the compiler first
// canonicalises every operand of fminf/fmaxf it cannot prove canonical (one
// v_max_f32 x, x, x per texel read from LDS); the tile holds arithmetic
// results and the quiet NaN of out-of-image texels, which are canonical.
__device__ __forceinline__ float vmin3(float a, float b, float c) {
    float r;
    asm("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ float vmax3(float a, float b, float c) {
    float r;
    asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// The bilinear history colour of a pixel with history, in YCoCg
// (bmfr.cl:922-963).
template <typename GT>
__device__ __forceinline__ float3 taa_history_ycocg(float2 pf, const TaaHistory<GT> &h) {
    float3 pc = make_float3(0.f, 0.f, 0.f);
    float tw = 0.f;
    const float fx = pf.x - (float)h.ix, fy = pf.y - (float)h.iy;
    const float omx = 1.f - fx, omy = 1.f - fy;
    const float wts[4] = {omx * omy, fx * omy, omx * fy, fx * fy};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (h.use[i]) {
            pc.x = pc.x + wts[i] * h.tap[i].x;
            pc.y = pc.y + wts[i] * h.tap[i].y;
            pc.z = pc.z + wts[i] * h.tap[i].z;
            tw = tw + wts[i];
        }
    }
    const double rtw = quotient_reciprocal(tw);   // pc / tw, IEEE bits; unguarded like bmfr.cl:962
    pc.x = quotient(pc.x, rtw);
    pc.y = quotient(pc.y, rtw);
    pc.z = quotient(pc.z, rtw);
    return rgb_to_ycocg(pc);
}

// The history colour `pyc` clamped to the pixel's 3x3 neighbourhood and
// blended with its tone-mapped colour `mine` (bmfr.cl:893-920, 964-973).
// `nbp` is the top-left texel of the neighbourhood in an LDS tile of row
// stride STRIDE (NaN outside the image).
template <int STRIDE>
__device__ __forceinline__ float3 taa_clamp_blend(const StageArgs &a, float3 mine, float3 pyc, const float4 *nbp) {
    float4 nb[3][3];
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) nb[dy][dx] = nbp[dy * STRIDE + dx];
    // cross: centre and its 4-neighbours; box: all 9
    float3 mnc, mxc, mnb, mxb;
    mnc.x = vmin3(vmin3(nb[0][1].x, nb[1][0].x, nb[1][1].x), nb[1][2].x, nb[2][1].x);
    mnc.y = vmin3(vmin3(nb[0][1].y, nb[1][0].y, nb[1][1].y), nb[1][2].y, nb[2][1].y);
    mnc.z = vmin3(vmin3(nb[0][1].z, nb[1][0].z, nb[1][1].z), nb[1][2].z, nb[2][1].z);
    mxc.x = vmax3(vmax3(nb[0][1].x, nb[1][0].x, nb[1][1].x), nb[1][2].x, nb[2][1].x);
    mxc.y = vmax3(vmax3(nb[0][1].y, nb[1][0].y, nb[1][1].y), nb[1][2].y, nb[2][1].y);
    mxc.z = vmax3(vmax3(nb[0][1].z, nb[1][0].z, nb[1][1].z), nb[1][2].z, nb[2][1].z);
    mnb.x = vmin3(vmin3(mnc.x, nb[0][0].x, nb[0][2].x), nb[2][0].x, nb[2][2].x);
    mnb.y = vmin3(vmin3(mnc.y, nb[0][0].y, nb[0][2].y), nb[2][0].y, nb[2][2].y);
    mnb.z = vmin3(vmin3(mnc.z, nb[0][0].z, nb[0][2].z), nb[2][0].z, nb[2][2].z);
    mxb.x = vmax3(vmax3(mxc.x, nb[0][0].x, nb[0][2].x), nb[2][0].x, nb[2][2].x);
    mxb.y = vmax3(vmax3(mxc.y, nb[0][0].y, nb[0][2].y), nb[2][0].y, nb[2][2].y);
    mxb.z = vmax3(vmax3(mxc.z, nb[0][0].z, nb[0][2].z), nb[2][0].z, nb[2][2].z);
    float3 cl;
    cl.x = fminf(fmaxf(pyc.x, (mnb.x + mnc.x) / 2.f), (mxb.x + mxc.x) / 2.f);
    cl.y = fminf(fmaxf(pyc.y, (mnb.y + mnc.y) / 2.f), (mxb.y + mxc.y) / 2.f);
    cl.z = fminf(fmaxf(pyc.z, (mnb.z + mnc.z) / 2.f), (mxb.z + mxc.z) / 2.f);
    const float3 prgb = ycocg_to_rgb(cl);
    const float ta = a.taa_blend_alpha, omt = 1.f - ta;
    float3 r;
    r.x = ta * mine.x + omt * prgb.x;
    r.y = ta * mine.y + omt * prgb.y;
    r.z = ta * mine.z + omt * prgb.z;
    return r;
}

// TAA of one pixel (bmfr.cl:884-973): `mine` is its tone-mapped colour, `nbp`
// as for taa_clamp_blend.
template <int STRIDE, typename GT>
__device__ __forceinline__ float3 taa_resolve(const StageArgs &a, float2 pf, float3 mine, const TaaHistory<GT> &h,
                                              const float4 *nbp) {
    if (!h.history) return mine;
    return taa_clamp_blend<STRIDE>(a, mine, taa_history_ycocg<GT>(pf, h), nbp);
}

// The frame's TAA result into the result history (result_buffer.current())
// and, when one is bound, the caller's output surface, which receives the
// stored value (rounded to the history's storage type).
template <typename GT>
__device__ __forceinline__ void store_result(const StageArgs &a, const StageBuffers &bufs, int64_t lp, int x, int y,
                                             float3 r) {
    store3_px((GT *)bufs.result_cur, (uint32_t)lp, r);
    if (bufs.surface) {
        if constexpr (!std::is_same<GT, float>::value) r = make_float3(round_half(r.x), round_half(r.y), round_half(r.z));
        store_surface(bufs.surface, a.surface_format, a.surface_pitch, x, y, r);
    }
}

__device__ __forceinline__ float4 ycocg_texel(float3 rgb) {
    const float3 yc = rgb_to_ycocg(rgb);
    return make_float4(yc.x, yc.y, yc.z, 0.f);
}

template <typename GT>
__device__ __forceinline__ void stage5_body(const StageArgs &a, const StageBuffers &bufs, int x, int y) {
    __shared__ float4 tile[kTaaTileH][kTaaTileW];
    const float2 *__restrict__ prev_pixel = (const float2 *)bufs.prev_pixel;
    const float *__restrict__ new_frame = (const float *)bufs.tone_mapped;
    const int W = a.width, H = a.height;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int x0 = x - tx, y0 = y - ty;
    // the pixel's own loads and its history taps are issued before the
    // tile's, so that they are in flight during the tile load and the barrier
    const bool inside = x < W && y < H;
    const int64_t lp = inside ? __umul24((unsigned)y, (unsigned)W) + (unsigned)x : 0;
    const float2 pf = gload_px(prev_pixel, (uint32_t)lp);
    const float3 mine = load3_px(new_frame, (uint32_t)lp);
    const TaaHistory<GT> h = taa_history<GT>(a, bufs, lp, footprint(pf, W, H));
    for (int e = threadIdx.x; e < kTaaTileW * kTaaTileH; e += 256) {
        const int ry = e / kTaaTileW, rx = e - ry * kTaaTileW;
        const int gx = x0 - 1 + rx, gy = y0 - 1 + ry;
        float4 t = make_float4(__builtin_nanf(""), __builtin_nanf(""), __builtin_nanf(""), 0.f);
        if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const int64_t ls = __umul24((unsigned)gy, (unsigned)W) + (unsigned)gx;
            t = ycocg_texel(load3_px(new_frame, (uint32_t)ls));
        }
        tile[ry][rx] = t;
    }
    __syncthreads();
    if (!inside) return;
    store_result<GT>(a, bufs, lp, x, y, taa_resolve<kTaaTileW, GT>(a, pf, mine, h, &tile[ty][tx]));
}

// ---------------------------------------------------------------------------
// Stages 3, 4 and 5 fused for the chained paths: one work-group per padded
// 32x32 block (whose image pixels share one set of weights and bounds).
// Phase 1 computes the filtered, accumulated and tone-mapped colour of the
// block's image pixels, 4 per thread, stores the accumulated colour and puts
// the tone-mapped colour, in YCoCg, into an LDS tile; it also recomputes the
// one-pixel ring around the block (pixels of the neighbouring blocks, whose
// weights and bounds are vector loads), which only goes into the tile.
// Phase 2 runs TAA on the block's pixels from that tile.  Neither `filtered`
// nor `tone_mapped` is written; each pixel's arithmetic is that of the
// separate stages, so every result has the same bits.
constexpr int kRing = kBlock + 2;                 // 34x34 tile with the ring
constexpr int kRingTexels = 4 * kBlock + 4;       // 132 ring texels

// The tone curve's tables read from a copy in LDS (the same values as the
// constant tables, so the same bits).  This kernel runs at 4 waves per SIMD
// and its pixels are processed one after another, so an L1/L2 round trip per
// pixel for the table loads is not hidden; an LDS read is short.
struct ToneTablesLds {
    const double2 *lg;   // {c_i, c_i^-y}
    const double *p2;    // 2^(-y k)
    __device__ __forceinline__ double c(int i) const { return lg[i].x; }
    __device__ __forceinline__ double cy(int i) const { return lg[i].y; }
    __device__ __forceinline__ double pow2(int k) const { return p2[k]; }
};

template <typename GT, bool CUSTOM = false>
__device__ __forceinline__ void stage345_body(const StageArgs &a, const StageBuffers &bufs, int blk) {
    __shared__ float4 tile[kRing][kRing];
    __shared__ double2 tone_log[256];
    __shared__ double tone_pow2[149];
    static_assert(sizeof(kToneLog) == sizeof(tone_log) && sizeof(kTonePow2) == sizeof(tone_pow2), "tables");
    const int W = a.width, H = a.height;
    const int2 off = block_offset(a.frame);
    const int by = blk / a.blocks_x, bx = blk - by * a.blocks_x;
    // image pixel of the block's padded-grid corner (bmfr.cl:707-710 inverted)
    const int x0 = bx * kBlock - kBlock / 2 + off.x;
    const int y0 = by * kBlock - kBlock / 2 + off.y;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const float4 nan4 = make_float4(__builtin_nanf(""), __builtin_nanf(""), __builtin_nanf(""), 0.f);

    // Pixel r (0..3) of this thread is (x0 + tx, y0 + ty + 8 r).  A pixel
    // outside the image reads the nearest image pixel and its results are
    // dropped, so that no load sits behind a branch: the loop below is a
    // software pipeline in which pixel r + 1's inputs are in flight while
    // pixel r's history taps are, and each pixel's TAA history taps are
    // issued with its stage-4 taps (the second round trip) instead of after
    // the tile barrier.
    uint32_t lp[4];
    bool inside[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int x = x0 + tx, y = y0 + ty + 8 * r;
        inside[r] = x >= 0 && y >= 0 && x < W && y < H;
        lp[r] = (uint32_t)__umul24((unsigned)min(max(y, 0), H - 1), (unsigned)W) + (uint32_t)min(max(x, 0), W - 1);
    }
    // The one-pixel ring around the block (pixels of the neighbouring blocks,
    // whose weights and bounds are vector loads), one texel per thread of the
    // first 132; it only goes into the tile.
    const bool ring = threadIdx.x < kRingTexels;
    int rx, ry;   // ring position in -1..32
    {
        const int h = threadIdx.x;
        if (h < kRing) { rx = h - 1; ry = -1; }
        else if (h < 2 * kRing) { rx = h - kRing - 1; ry = kBlock; }
        else if (h < 2 * kRing + kBlock) { rx = -1; ry = h - 2 * kRing; }
        else { rx = kBlock; ry = h - 2 * kRing - kBlock; }
    }
    const int rgx = x0 + rx, rgy = y0 + ry;
    const bool ring_inside = ring && rgx >= 0 && rgy >= 0 && rgx < W && rgy < H;
    const int rcx = min(max(rgx, 0), W - 1), rcy = min(max(rgy, 0), H - 1);
    const uint32_t ring_lp = (uint32_t)__umul24((unsigned)rcy, (unsigned)W) + (uint32_t)rcx;

    float3 mine[4], hist[4];
    uint32_t has_history = 0;   // bit r: pixel r has a TAA history (bmfr.cl:884-890)
    // The first pixel's inputs, the block's fit and the tone tables are one
    // memory round trip: issued together, before the tables' barrier.
    Stage4Inputs in = stage4_inputs(a, bufs, lp[0]);
    float3 wp = load3_px((const GT *)bufs.positions_cur, lp[0]);
    float3 nrm = load3_px((const GT *)bufs.normals_cur, lp[0]);
    // 256 threads: one log entry each, the first 149 one power of two each
    const double2 tl = make_double2(kToneLog[threadIdx.x][0], kToneLog[threadIdx.x][1]);
    const double tp = kTonePow2[min((int)threadIdx.x, 148)];
    ScaleRecips rec;
    BlockFit fit;
    if constexpr (!CUSTOM) {
        rec = scale_recips(bufs, blk);   // all lanes active; blk is the work-group's
        fit = block_fit(bufs, blk);
    }
    tone_log[threadIdx.x] = tl;
    if (threadIdx.x < 149) tone_pow2[threadIdx.x] = tp;
    __syncthreads();
    const ToneTablesLds tt{tone_log, tone_pow2};
    Stage4Inputs ring_in;
    float3 ring_wp, ring_n;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const Footprint fp = footprint(in.pf, W, H);
        const Stage4Gather<GT> gth = stage4_taps<GT>(a, bufs, lp[r], in, fp);
        const TaaHistory<GT> h = taa_history<GT>(a, bufs, lp[r], fp);
        Stage4Inputs nin;
        float3 nwp, nn;
        if (r < 3) {
            nin = stage4_inputs(a, bufs, lp[r + 1]);
            nwp = load3_px((const GT *)bufs.positions_cur, lp[r + 1]);
            nn = load3_px((const GT *)bufs.normals_cur, lp[r + 1]);
        } else if (ring) {
            ring_in = stage4_inputs(a, bufs, ring_lp);
            ring_wp = load3_px((const GT *)bufs.positions_cur, ring_lp);
            ring_n = load3_px((const GT *)bufs.normals_cur, ring_lp);
        }
        float3 fc;
        if constexpr (CUSTOM) fc = stage3_color_of<GT, true>(a, bufs, wp, nrm, blk);
        else fc = stage3_color_fit(wp, nrm, fit, &rec);
        const float3 ac = stage4_blend<GT>(a, fc, gth);
        if (inside[r]) store3_px((GT *)bufs.out_cur, lp[r], ac);
        mine[r] = tone_map(gth.albedo, ac, tt);
        hist[r] = taa_history_ycocg<GT>(in.pf, h);
        has_history |= inside[r] && h.history ? 1u << r : 0u;
        tile[ty + 8 * r + 1][tx + 1] = inside[r] ? ycocg_texel(mine[r]) : nan4;
        if (r < 3) {
            in = nin;
            wp = nwp;
            nrm = nn;
        }
    }
    if (ring) {
        float4 t = nan4;
        // the pixel's own block (bmfr.cl:715-718), from the clamped position
        // so that it stays a block of the frame
        const int64_t g = (int64_t)((rcy + kBlock / 2 - off.y) / kBlock) * a.blocks_x +
                          (rcx + kBlock / 2 - off.x) / kBlock;
        const Stage4Gather<GT> gth = stage4_taps<GT>(a, bufs, ring_lp, ring_in, footprint(ring_in.pf, W, H));
        const float3 ac = stage4_blend<GT>(a, stage3_color_of<GT, CUSTOM>(a, bufs, ring_wp, ring_n, g), gth);
        if (ring_inside) t = ycocg_texel(tone_map(gth.albedo, ac, tt));
        tile[ry + 1][rx + 1] = t;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (inside[r]) {
            const float3 res = (has_history >> r) & 1u
                                   ? taa_clamp_blend<kRing>(a, mine[r], hist[r], &tile[ty + 8 * r][tx])
                                   : mine[r];
            store_result<GT>(a, bufs, lp[r], x0 + tx, y0 + ty + 8 * r, res);
        }
    }
}

template <typename GT, bool CUSTOM>
__global__ __launch_bounds__(256) void k_filter_taa(const StageArgs a, const StageBuffers b) {
    stage345_body<GT, CUSTOM>(a, b, xcd_tile(blockIdx.x, gridDim.x));
}

template <typename GT>
__global__ __launch_bounds__(256) void k_taa(const StageArgs a, const StageBuffers b) {
    int x, y;
    image_tile(a, x, y);
    stage5_body<GT>(a, b, x, y);
}

// ---------------------------------------------------------------------------
// Batched kernels: one 1-D grid over the concatenated tiles of all streams
// of a batch.  Each work-group finds its stream from the prefix table
// (uniform binary search, batch_stream_of), then runs the same per-pixel
// body as the single-stream kernel.
__device__ __forceinline__ void resolve_batch(const BatchLaunch &L, int s, StageArgs &a, StageBuffers &b) {
    const BatchStream &bs = L.streams[s];
    const BatchStep &st = L.steps[s];
    a = bs.args;
    a.frame = st.frame;
#pragma unroll
    for (int i = 0; i < 16; ++i) a.matrix[i] = st.matrix[i];
    a.pixel_offset_x = st.pixel_offset_x;
    a.pixel_offset_y = st.pixel_offset_y;
    a.surface_format = st.surface_format;
    a.surface_pitch = st.surface_pitch;
    const int c = st.cur, p = 1 - st.cur;
    b.albedo = const_cast<void *>(st.albedo);
    b.normals_cur = const_cast<void *>(st.normals_cur);
    b.normals_prev = const_cast<void *>(st.normals_prev);
    b.positions_cur = const_cast<void *>(st.positions_cur);
    b.positions_prev = const_cast<void *>(st.positions_prev);
    b.noisy_in = const_cast<void *>(st.noisy_in);
    b.surface = st.surface;
    b.noisy_acc_cur = bs.noisy_acc[c];
    b.noisy_acc_prev = bs.noisy_acc[p];
    b.spp_cur = bs.spp[c];
    b.spp_prev = bs.spp[p];
    b.prev_pixel = bs.prev_pixel;
    b.accept = bs.accept;
    b.tmp_data = bs.tmp_data;
    b.weights = bs.weights;
    b.mins_maxs = bs.mins_maxs;
    b.filtered = bs.filtered;
    b.tone_mapped = bs.tone_mapped;
    b.out_cur = bs.out[c];
    b.out_prev = bs.out[p];
    b.result_cur = bs.result[c];
    b.result_prev = bs.result[p];
    b.noise = bs.noise;
}

template <typename GT, typename TT>
__global__ __launch_bounds__(256) void kb_accumulate_noisy_data(const BatchLaunch L) {
    const int32_t *starts = L.starts + kStartsPadded * (kMaxBatch + 1);
    const int unit = xcd_tile(blockIdx.x, gridDim.x);
    const int s = batch_stream_of(starts, L.n, unit);
    StageArgs a;
    StageBuffers b;
    resolve_batch(L, s, a, b);
    const int local = unit - starts[s];
    const int tiles_x = a.margin_width / 32;
    const int ty = local / tiles_x, tx = local - ty * tiles_x;
    stage1_body<GT, TT>(a, b, tx * 32 + (threadIdx.x & 31), ty * 8 + (threadIdx.x >> 5),
                        local * 256 + threadIdx.x, true);
}

// Stages 3 and 3+4 run over the padded tiles (block-uniform weights), 4 and
// 5 over the image tiles.
template <int STAGE, typename GT>
__global__ __launch_bounds__(256) void kb_image_stage(const BatchLaunch L) {
    constexpr bool kBlockTiles = STAGE == 3 || STAGE == 34;
    const int32_t *starts =
        L.starts + (STAGE == 345 ? kStartsBlocks : kBlockTiles ? kStartsPadded : kStartsImage) * (kMaxBatch + 1);
    const int unit = xcd_tile(blockIdx.x, gridDim.x);
    const int s = batch_stream_of(starts, L.n, unit);
    StageArgs a;
    StageBuffers b;
    resolve_batch(L, s, a, b);
    const int local = unit - starts[s];
    if constexpr (STAGE == 345) {
        stage345_body<GT>(a, b, local);
    } else if constexpr (kBlockTiles) {
        if constexpr (STAGE == 3) stage3_body<GT>(a, b, local);
        else stage34_body<GT>(a, b, local);
    } else {
        const int tiles_x = (a.width + 63) / 64;
        const int ty = local / tiles_x, tx = local - ty * tiles_x;
        const int x = tx * 64 + (threadIdx.x & 63), y = ty * 4 + (threadIdx.x >> 6);
        if constexpr (STAGE == 4) stage4_body<GT>(a, b, x, y);
        else stage5_body<GT>(a, b, x, y);
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers.
template <typename GT, typename TT>
static hipError_t launch_s1(const StageArgs &a, const StageBuffers &b, bool wn, hipStream_t s) {
    dim3 grid((a.margin_width / 32) * (a.margin_height / 8));
    if (a.features.count)
        hipLaunchKernelGGL((k_accumulate_noisy_data<GT, TT, true>), grid, dim3(256), 0, s, a, b, wn ? 1 : 0);
    else
        hipLaunchKernelGGL((k_accumulate_noisy_data<GT, TT, false>), grid, dim3(256), 0, s, a, b, wn ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_accumulate_noisy_data(const StageArgs &a, const StageBuffers &b, bool wn,
                                        hipStream_t s) {
    if (a.storage_f16) {
        return a.tmp_f16 ? launch_s1<__half, __half>(a, b, wn, s) : launch_s1<__half, float>(a, b, wn, s);
    }
    return a.tmp_f16 ? launch_s1<float, __half>(a, b, wn, s) : launch_s1<float, float>(a, b, wn, s);
}

// The noise table of N features: (N-1) x 1024 terms, buffer count N + 3.
__global__ __launch_bounds__(256) void k_noise_table(int frame, int features, float *__restrict__ noise) {
    const int lin = blockIdx.x * 256 + threadIdx.x;
    if (lin < (features - 1) * kBlockPixels)
        gstore(noise, lin, noise_random(lin & (kBlockPixels - 1), 1 + (lin >> 10), frame, features + 3));
}

hipError_t launch_noise_table(const StageArgs &a, const StageBuffers &b, hipStream_t s) {
    const int nf = a.features.count ? a.features.count : kFeatures;
    if (nf < 2) return hipSuccess;   // one feature: no noised column
    hipLaunchKernelGGL(k_noise_table, dim3((nf - 1) * kBlockPixels / 256), dim3(256), 0, s, a.frame, nf,
                       (float *)b.noise);
    return hipGetLastError();
}

static inline dim3 image_grid(const StageArgs &a) {
    return dim3(((a.width + 63) / 64) * ((a.height + 3) / 4));
}
static inline dim3 block_grid(const StageArgs &a) {
    return dim3((a.margin_width / 32) * (a.margin_height / 8));
}

// Launches kernel K over (G-buffer type, configured feature set or not).
enum { K_WEIGHTED_SUM, K_WEIGHTED_SUM_FILTERED, K_FILTER_TAA };
template <int K, typename GT, bool CUSTOM>
static void launch_one(const StageArgs &a, dim3 grid, const StageBuffers &b, hipStream_t s) {
    if constexpr (K == K_WEIGHTED_SUM)
        hipLaunchKernelGGL((k_weighted_sum<GT, CUSTOM>), grid, dim3(256), 0, s, a, b);
    else if constexpr (K == K_WEIGHTED_SUM_FILTERED)
        hipLaunchKernelGGL((k_weighted_sum_filtered<GT, CUSTOM>), grid, dim3(256), 0, s, a, b);
    else
        hipLaunchKernelGGL((k_filter_taa<GT, CUSTOM>), grid, dim3(256), 0, s, a, b);
}
template <int K>
static hipError_t launch_gt_set(const StageArgs &a, dim3 grid, const StageBuffers &b, hipStream_t s) {
    const bool custom = a.features.count != 0;
    if (a.storage_f16) {
        if (custom) launch_one<K, __half, true>(a, grid, b, s);
        else launch_one<K, __half, false>(a, grid, b, s);
    } else {
        if (custom) launch_one<K, float, true>(a, grid, b, s);
        else launch_one<K, float, false>(a, grid, b, s);
    }
    return hipGetLastError();
}

hipError_t launch_weighted_sum(const StageArgs &a, const StageBuffers &b, hipStream_t s) {
    return launch_gt_set<K_WEIGHTED_SUM>(a, block_grid(a), b, s);
}

hipError_t launch_accumulate_filtered_data(const StageArgs &a, const StageBuffers &b, hipStream_t s) {
    if (a.storage_f16)
        hipLaunchKernelGGL((k_accumulate_filtered_data<__half>), image_grid(a), dim3(256), 0, s, a, b);
    else
        hipLaunchKernelGGL((k_accumulate_filtered_data<float>), image_grid(a), dim3(256), 0, s, a, b);
    return hipGetLastError();
}

hipError_t launch_weighted_sum_filtered(const StageArgs &a, const StageBuffers &b, hipStream_t s) {
    return launch_gt_set<K_WEIGHTED_SUM_FILTERED>(a, block_grid(a), b, s);
}

hipError_t launch_filter_taa(const StageArgs &a, const StageBuffers &b, hipStream_t s) {
    return launch_gt_set<K_FILTER_TAA>(a, dim3(a.blocks), b, s);
}

hipError_t launch_taa(const StageArgs &a, const StageBuffers &b, hipStream_t s) {
    if (a.storage_f16)
        hipLaunchKernelGGL((k_taa<__half>), image_grid(a), dim3(256), 0, s, a, b);
    else
        hipLaunchKernelGGL((k_taa<float>), image_grid(a), dim3(256), 0, s, a, b);
    return hipGetLastError();
}

hipError_t launch_batch_accumulate_noisy_data(const BatchLaunch &L, int total, bool g16, bool t16,
                                              hipStream_t s) {
    if (total <= 0) return hipErrorInvalidValue;
    if (g16) {
        if (t16) hipLaunchKernelGGL((kb_accumulate_noisy_data<__half, __half>), dim3(total), dim3(256), 0, s, L);
        else hipLaunchKernelGGL((kb_accumulate_noisy_data<__half, float>), dim3(total), dim3(256), 0, s, L);
    } else {
        if (t16) hipLaunchKernelGGL((kb_accumulate_noisy_data<float, __half>), dim3(total), dim3(256), 0, s, L);
        else hipLaunchKernelGGL((kb_accumulate_noisy_data<float, float>), dim3(total), dim3(256), 0, s, L);
    }
    return hipGetLastError();
}

template <int STAGE>
static hipError_t launch_batch_image(const BatchLaunch &L, int total, bool g16, hipStream_t s) {
    if (total <= 0) return hipErrorInvalidValue;
    if (g16) hipLaunchKernelGGL((kb_image_stage<STAGE, __half>), dim3(total), dim3(256), 0, s, L);
    else hipLaunchKernelGGL((kb_image_stage<STAGE, float>), dim3(total), dim3(256), 0, s, L);
    return hipGetLastError();
}

hipError_t launch_batch_weighted_sum(const BatchLaunch &L, int total, bool g16, hipStream_t s) {
    return launch_batch_image<3>(L, total, g16, s);
}
hipError_t launch_batch_accumulate_filtered_data(const BatchLaunch &L, int total, bool g16,
                                                 hipStream_t s) {
    return launch_batch_image<4>(L, total, g16, s);
}
hipError_t launch_batch_weighted_sum_filtered(const BatchLaunch &L, int total, bool g16, hipStream_t s) {
    return launch_batch_image<34>(L, total, g16, s);
}
hipError_t launch_batch_taa(const BatchLaunch &L, int total, bool g16, hipStream_t s) {
    return launch_batch_image<5>(L, total, g16, s);
}
hipError_t launch_batch_filter_taa(const BatchLaunch &L, int total, bool g16, hipStream_t s) {
    return launch_batch_image<345>(L, total, g16, s);
}

}  // namespace bmfr
