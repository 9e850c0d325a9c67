/*
 * CPU oracle for the BMFR five-stage frame chain.  TEST INFRASTRUCTURE ONLY
 * (the checker, never the product): see bmfr_oracle.h for scope and the
 * semantic decisions taken where the reference is racy or implementation
 * defined.  Build with -O2 -ffp-contract=off (oracle/Makefile): every float
 * operation rounded separately, except the FMAs written out as fmaf (stage
 * 3's sum, which the reference's OpenCL build contracts).
 *
 * Pinning status (what checks each function; tests/test_oracle_*.py):
 *   orc_hash/orc_random      known-answer table computed independently in the tests
 *   orc_f32_to_f16           exhaustive/random comparison with numpy float16 (RNE)
 *   orc_mirror, block offsets closed form + bijection invariant
 *   orc_tree_*               compared with exact sums (integers) and with np.min/max
 *   orc_accumulate_noisy_data static-camera closed form, frame-0 pass-through,
 *                            brute-force per-pixel numpy restatement on tiny images
 *   orc_fitter               FP64 Householder (orc_lstsq_fp64) and numpy.linalg.lstsq,
 *                            exact-fit recovery of generating weights
 *   orc_weighted_sum         closed form from weights and mins_maxs
 *   orc_accumulate_filtered  frame-0 closed form, static-history closed form
 *   orc_taa                  frame-0 pass-through, constant image identity,
 *                            YCoCg round trip, edge renormalisation on 2x2
 */
#include "bmfr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

void orc_sizes_of(int32_t w, int32_t h, orc_sizes *o) {
    /* bmfr.cpp:107-118 */
    o->workset_w = ORC_BLOCK * ((w + ORC_BLOCK - 1) / ORC_BLOCK);
    o->workset_h = ORC_BLOCK * ((h + ORC_BLOCK - 1) / ORC_BLOCK);
    o->margin_w = o->workset_w + ORC_BLOCK;
    o->margin_h = o->workset_h + ORC_BLOCK;
    o->blocks_x = o->margin_w / ORC_BLOCK;
    o->blocks_y = o->margin_h / ORC_BLOCK;
    o->blocks = o->blocks_x * o->blocks_y;
    o->pixels = (int64_t)w * h;
    o->padded = (int64_t)o->margin_w * o->margin_h;
}

void orc_default_params(int32_t w, int32_t h, orc_params *p) {
    /* bmfr.cpp:58-62, 88; limits are project-chosen (the dataset header that
     * holds them is not shipped, bmfr.cpp:47). */
    p->width = w;
    p->height = h;
    p->position_limit_squared = 0.01f;
    p->normal_limit_squared = 0.1f;
    p->blend_alpha = 0.2f;
    p->second_blend_alpha = 0.1f;
    p->taa_blend_alpha = 0.2f;
    p->noise_amount = 1e-2;
    p->half_tmp = 1;
    p->feature_count = 0;
    p->features_not_scaled = 0;
    memset(p->feature_factors, 0, sizeof(p->feature_factors));
}

int32_t orc_feature_count(const orc_params *p) { return p->feature_count ? p->feature_count : ORC_FEATURES; }
int32_t orc_not_scaled(const orc_params *p) { return p->feature_count ? p->features_not_scaled : ORC_NOT_SCALED; }

static float orc_factor(int32_t id, const float n[3], const float wp[3]) {
    switch (id) {
        case 1: return n[0];
        case 2: return n[1];
        case 3: return n[2];
        case 4: return wp[0];
        case 5: return wp[1];
        case 6: return wp[2];
        default: return 1.f;
    }
}

float orc_feature(const orc_params *p, int32_t f, const float n[3], const float wp[3]) {
    if (p->feature_count == 0) {
        /* bmfr.cpp:65-77: 1.f, normal.xyz, world_position.xyz, its squares */
        switch (f) {
            case 0: return 1.f;
            case 1: case 2: case 3: return n[f - 1];
            case 4: case 5: case 6: return wp[f - 4];
            default: return wp[f - 7] * wp[f - 7];
        }
    }
    /* an expression `a*b` is one rounded float product; `a` alone is a*1 */
    return orc_factor(p->feature_factors[f][0], n, wp) * orc_factor(p->feature_factors[f][1], n, wp);
}

/* bmfr.cl:267-285: 16-entry jitter table, indexed by frame % 16. */
static const int32_t k_block_offsets[16][2] = {
    {-14, -14}, {4, -6},  {-8, 14},  {8, 0},   {-10, -8}, {2, 12},
    {12, -12},  {-10, 0}, {12, 14},  {-8, -16}, {6, 6},   {-2, -2},
    {6, -14},   {-16, 12}, {14, -4}, {-6, 4}};

void orc_block_offset(int32_t frame, int32_t out[2]) {
    int32_t k = frame % 16; /* OpenCL % truncates like C; frames are >= 0 */
    if (k < 0) k += 16;
    out[0] = k_block_offsets[k][0];
    out[1] = k_block_offsets[k][1];
}

int32_t orc_mirror(int32_t index, int32_t size) {
    /* bmfr.cl:209-216 */
    if (index < 0)
        index = abs(index) - 1;
    else if (index >= size)
        index = 2 * size - index - 1;
    return index;
}

uint32_t orc_hash(uint32_t a) {
    /* bmfr.cl:162-168 (32-bit wrapping arithmetic) */
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    return a;
}

float orc_random(uint32_t a) {
    /* bmfr.cl:170: convert_float(a) / convert_float(UINT_MAX); both
     * conversions round to nearest even, UINT_MAX becomes 2^32. */
    return (float)orc_hash(a) / (float)4294967295u;
}

float orc_add_random_n(float value, int32_t id, int32_t sub_vector, int32_t feature,
                       int32_t frame, double noise_amount, int32_t buffer_count) {
    /* bmfr.cl:173-182.  NOISE_AMOUNT is the double literal 1e-2, so
     * NOISE_AMOUNT * 2.f * (r - 0.5f) is evaluated in double, added to the
     * float value in double, and rounded to float on return. */
    uint32_t arg = (uint32_t)id + (uint32_t)sub_vector * ORC_LOCAL +
                   (uint32_t)feature * ORC_BLOCK_PIXELS +
                   (uint32_t)frame * (uint32_t)buffer_count * ORC_BLOCK_PIXELS;
    float r = orc_random(arg) - 0.5f;
    double noise = noise_amount * 2.0 * (double)r;
    return (float)((double)value + noise);
}

float orc_add_random(float value, int32_t id, int32_t sub_vector, int32_t feature,
                     int32_t frame, double noise_amount) {
    return orc_add_random_n(value, id, sub_vector, feature, frame, noise_amount, ORC_BUFFER_COUNT);
}

uint16_t orc_f32_to_f16(float f) {
    /* IEEE 754 binary32 -> binary16, round to nearest, ties to even. */
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xffu) { /* inf / nan */
        if (man) return (uint16_t)(sign | 0x7e00u | (man >> 13));
        return (uint16_t)(sign | 0x7c00u);
    }
    int32_t e = (int32_t)exp - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u); /* overflow -> inf */
    if (e <= 0) {
        /* subnormal half (or zero) */
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        uint32_t shift = (uint32_t)(14 - e);
        uint32_t half_man = man >> shift;
        uint32_t rem = man & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1u);
        if (rem > halfway || (rem == halfway && (half_man & 1u))) half_man++;
        return (uint16_t)(sign | half_man);
    }
    uint32_t half = sign | ((uint32_t)e << 10) | (man >> 13);
    uint32_t rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (half & 1u))) half++; /* may carry to inf */
    return (uint16_t)half;
}

float orc_f16_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t exp = ((uint32_t)h >> 10) & 0x1fu;
    uint32_t man = (uint32_t)h & 0x3ffu;
    uint32_t x;
    if (exp == 0) {
        if (man == 0) {
            x = sign;
        } else { /* normalise the subnormal */
            int32_t e = -1;
            do { man <<= 1; e++; } while (!(man & 0x400u));
            man &= 0x3ffu;
            x = sign | ((uint32_t)(127 - 15 - e) << 23) | (man << 13);
        }
    } else if (exp == 31) {
        x = sign | 0x7f800000u | (man << 13);
    } else {
        x = sign | ((exp - 15 + 127) << 23) | (man << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

float orc_round_half(float f) { return orc_f16_to_f32(orc_f32_to_f16(f)); }

float orc_scale(float value, float mn, float mx) {
    /* bmfr.cl:200-205 */
    if (fabsf(mx - mn) > 1.0f) return (value - mn) / (mx - mn);
    return value - mn;
}

static inline float dot3(const float a[3], float b0, float b1, float b2) {
    return (a[0] * b0 + a[1] * b1) + a[2] * b2;
}

void orc_rgb_to_ycocg(const float rgb[3], float out[3]) {
    /* bmfr.cl:184-190 */
    out[0] = dot3(rgb, 1.f, 2.f, 1.f);
    out[1] = dot3(rgb, 2.f, 0.f, -2.f);
    out[2] = dot3(rgb, -1.f, 2.f, -1.f);
}

void orc_ycocg_to_rgb(const float y[3], float out[3]) {
    /* bmfr.cl:192-198 */
    out[0] = dot3(y, 0.25f, 0.25f, -0.25f);
    out[1] = dot3(y, 0.25f, 0.f, 0.25f);
    out[2] = dot3(y, 0.25f, -0.25f, -0.25f);
}

float orc_tree_sum(float *s) {
    /* bmfr.cl:26-44: 256 -> 64 -> 8 -> 1 with this exact association. */
    for (int id = 0; id < 64; ++id)
        s[id] = s[id] + ((s[id + 64] + s[id + 128]) + s[id + 192]);
    for (int id = 0; id < 8; ++id)
        s[id] = s[id] + ((((((s[id + 8] + s[id + 16]) + s[id + 24]) + s[id + 32]) +
                           s[id + 40]) + s[id + 48]) + s[id + 56]);
    return ((((((s[0] + s[1]) + s[2]) + s[3]) + s[4]) + s[5]) + s[6]) + s[7];
}

float orc_tree_min(float *s) {
    /* bmfr.cl:46-66 */
    for (int id = 0; id < 64; ++id)
        s[id] = fminf(fminf(fminf(s[id], s[id + 64]), s[id + 128]), s[id + 192]);
    for (int id = 0; id < 8; ++id) {
        float v = s[id];
        for (int k = 1; k < 8; ++k) v = fminf(v, s[id + 8 * k]);
        s[id] = v;
    }
    float v = s[0];
    for (int k = 1; k < 8; ++k) v = fminf(v, s[k]);
    return v;
}

float orc_tree_max(float *s) {
    /* bmfr.cl:68-87 */
    for (int id = 0; id < 64; ++id)
        s[id] = fmaxf(fmaxf(fmaxf(s[id], s[id + 64]), s[id + 128]), s[id + 192]);
    for (int id = 0; id < 8; ++id) {
        float v = s[id];
        for (int k = 1; k < 8; ++k) v = fmaxf(v, s[id + 8 * k]);
        s[id] = v;
    }
    float v = s[0];
    for (int k = 1; k < 8; ++k) v = fmaxf(v, s[k]);
    return v;
}

/* convert_int_rtn with AMD hardware semantics: floor, saturate, NaN -> 0. */
static inline int32_t cvt_rtn_sat(float f) {
    if (f != f) return 0;
    float fl = floorf(f);
    if (fl >= 2147483648.0f) return 2147483647;
    if (fl < -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)fl;
}

static inline int32_t wrap_add(int32_t a, int32_t b) {
    return (int32_t)((uint32_t)a + (uint32_t)b);
}
static inline int32_t wrap_mad(int32_t a, int32_t b, int32_t c) {
    return (int32_t)((uint32_t)a * (uint32_t)b + (uint32_t)c);
}

int32_t orc_accumulate_noisy_data(
    const orc_params *p, const float *current_normals, const float *previous_normals,
    const float *current_positions, const float *previous_positions,
    const float *current_noisy, const float *previous_noisy, const uint8_t *previous_spp,
    float *accumulated_noisy, float *out_prev_frame_pixel, uint8_t *accept_bools,
    uint8_t *current_spp, float *tmp_data, const float m[16],
    const float pixel_offset[2], int32_t frame_number) {
    orc_sizes s;
    orc_sizes_of(p->width, p->height, &s);
    const int32_t W = p->width, H = p->height;
    if (W < ORC_MIN_SIZE || H < ORC_MIN_SIZE) return -1;
    int32_t off[2];
    orc_block_offset(frame_number, off);
    const float pos_lim = p->position_limit_squared;
    const float nrm_lim = p->normal_limit_squared;
    const int32_t NF = orc_feature_count(p), NB = NF + 3;

#pragma omp parallel for schedule(static)
    for (int32_t gy = 0; gy < s.margin_h; ++gy) {
        for (int32_t gx = 0; gx < s.margin_w; ++gx) {
            /* bmfr.cl:310-322 */
            const int32_t pwx = gx - ORC_BLOCK / 2 + off[0];
            const int32_t pwy = gy - ORC_BLOCK / 2 + off[1];
            const int32_t px = orc_mirror(pwx, W);
            const int32_t py = orc_mirror(pwy, H);
            const int64_t lp = (int64_t)py * W + px;
            const float wp[3] = {current_positions[lp * 3 + 0], current_positions[lp * 3 + 1],
                                 current_positions[lp * 3 + 2]};
            const float nrm[3] = {current_normals[lp * 3 + 0], current_normals[lp * 3 + 1],
                                  current_normals[lp * 3 + 2]};
            const float cur[3] = {current_noisy[lp * 3 + 0], current_noisy[lp * 3 + 1],
                                  current_noisy[lp * 3 + 2]};
            float ppf_x = (float)px, ppf_y = (float)py; /* bmfr.cl:325 */
            uint8_t store_accept = 0;
            float blend_alpha = 1.f;
            float prev_col[3] = {0.f, 0.f, 0.f};
            float sample_spp = 0.f;
            if (frame_number > 0) {
                /* bmfr.cl:343-356: row-vector matrix, dot() left to right, w = 1 */
                float uvx = ((m[0] * wp[0] + m[4] * wp[1]) + m[8] * wp[2]) + m[12] * 1.f;
                float uvy = ((m[1] * wp[0] + m[5] * wp[1]) + m[9] * wp[2]) + m[13] * 1.f;
                float ww = ((m[3] * wp[0] + m[7] * wp[1]) + m[11] * wp[2]) + m[15] * 1.f;
                uvx = uvx / ww;
                uvy = uvy / ww;
                uvx = uvx + 1.f;
                uvy = uvy + 1.f;
                uvx = uvx / 2.f;
                uvy = uvy / 2.f;
                ppf_x = uvx * (float)W;
                ppf_y = uvy * (float)H;
                ppf_x = ppf_x - pixel_offset[0];
                ppf_y = ppf_y - (1.f - pixel_offset[1]);
                const int32_t ipx = cvt_rtn_sat(ppf_x), ipy = cvt_rtn_sat(ppf_y);
                /* bmfr.cl:364-370 */
                const float fx = ppf_x - (float)ipx, fy = ppf_y - (float)ipy;
                const float omx = 1.f - fx, omy = 1.f - fy;
                const float wts[4] = {omx * omy, fx * omy, omx * fy, fx * fy};
                float total_weight = 0.f;
                for (int i = 0; i < 4; ++i) { /* bmfr.cl:374-419 */
                    const int32_t sx = wrap_add(ipx, i & 1);
                    const int32_t sy = wrap_add(ipy, i >> 1);
                    if (sx >= 0 && sy >= 0 && sx < W && sy < H) {
                        const int64_t ls = (int64_t)wrap_mad(sy, W, sx);
                        const float dx = previous_positions[ls * 3 + 0] - wp[0];
                        const float dy = previous_positions[ls * 3 + 1] - wp[1];
                        const float dz = previous_positions[ls * 3 + 2] - wp[2];
                        const float pd = (dx * dx + dy * dy) + dz * dz;
                        if (pd < pos_lim) {
                            const float nx = previous_normals[ls * 3 + 0] - nrm[0];
                            const float ny = previous_normals[ls * 3 + 1] - nrm[1];
                            const float nz = previous_normals[ls * 3 + 2] - nrm[2];
                            const float nd = (nx * nx + ny * ny) + nz * nz;
                            if (nd < nrm_lim) {
                                store_accept |= (uint8_t)(1u << i);
                                sample_spp = sample_spp + wts[i] * (float)previous_spp[ls];
                                prev_col[0] = prev_col[0] + wts[i] * previous_noisy[ls * 3 + 0];
                                prev_col[1] = prev_col[1] + wts[i] * previous_noisy[ls * 3 + 1];
                                prev_col[2] = prev_col[2] + wts[i] * previous_noisy[ls * 3 + 2];
                                total_weight = total_weight + wts[i];
                            }
                        }
                    }
                }
                if (total_weight > 0.f) { /* bmfr.cl:421-429 */
                    prev_col[0] = prev_col[0] / total_weight;
                    prev_col[1] = prev_col[1] / total_weight;
                    prev_col[2] = prev_col[2] / total_weight;
                    sample_spp = sample_spp / total_weight;
                    blend_alpha = 1.f / (sample_spp + 1.f);
                    blend_alpha = fmaxf(blend_alpha, p->blend_alpha);
                }
            }
            /* bmfr.cl:433-442 */
            uint8_t new_spp = 1;
            if (blend_alpha < 1.f) {
                if (sample_spp > 254.f) {
                    new_spp = 255;
                } else {
                    float r = nearbyintf(sample_spp); /* rte; sample_spp >= 0 here */
                    if (r > 255.f) r = 255.f;
                    if (!(r >= 0.f)) r = 0.f;
                    new_spp = (uint8_t)((uint8_t)r + 1);
                }
            }
            /* bmfr.cl:442 lets mirrored work-items write spp too; they compute
             * the same value as the owner (spp does not depend on colour), so
             * writing it from the owner only gives identical results. */
            const float omb = 1.f - blend_alpha;
            float nc[3];
            for (int c = 0; c < 3; ++c) nc[c] = blend_alpha * cur[c] + omb * prev_col[c];

            /* bmfr.cl:448-476: FEATURE_BUFFERS, then the colour */
            float feat[ORC_MAX_FEATURES + 3];
            for (int f = 0; f < NF; ++f) feat[f] = orc_feature(p, f, nrm, wp);
            feat[NF] = nc[0];
            feat[NF + 1] = nc[1];
            feat[NF + 2] = nc[2];
            const int32_t xin = gx % ORC_BLOCK, yin = gy % ORC_BLOCK;
            const int32_t xb = gx / ORC_BLOCK, yb = gy / ORC_BLOCK;
            const int64_t base = (int64_t)xin + (int64_t)yin * ORC_BLOCK +
                                 (int64_t)xb * ORC_BLOCK_PIXELS * NB +
                                 (int64_t)yb * s.blocks_x * ORC_BLOCK_PIXELS * NB;
            for (int f = 0; f < NB; ++f) {
                float v = feat[f];
                if (v != v) v = 0.f;
                if (p->half_tmp) {
                    v = fmaxf(fminf(v, 65504.f), -65504.f);
                    v = orc_round_half(v);
                }
                tmp_data[base + (int64_t)f * ORC_BLOCK_PIXELS] = v;
            }
            /* bmfr.cl:478-484: only the non-mirrored owner writes these */
            if (pwx >= 0 && pwx < W && pwy >= 0 && pwy < H) {
                accumulated_noisy[lp * 3 + 0] = nc[0];
                accumulated_noisy[lp * 3 + 1] = nc[1];
                accumulated_noisy[lp * 3 + 2] = nc[2];
                out_prev_frame_pixel[lp * 2 + 0] = ppf_x;
                out_prev_frame_pixel[lp * 2 + 1] = ppf_y;
                accept_bools[lp] = store_accept;
                current_spp[lp] = new_spp;
            }
        }
    }
    return 0;
}

/* LOAD/STORE of bmfr.cl:255-265: in half mode every store rounds to binary16. */
static inline float st_round(const orc_params *p, float v) {
    return p->half_tmp ? orc_round_half(v) : v;
}

static void fitter_block(const orc_params *p, float *d, float *mins_maxs, float *weights,
                         int32_t blk, int32_t frame) {
    const int NF = orc_feature_count(p), NNS = orc_not_scaled(p);
    const int NB = NF + 3, NS = NF - NNS, RE = NF + 1;   /* BUFFER_COUNT, FEATURES_SCALED, R_EDGE */
    float sum_vec[ORC_LOCAL];
    float u_vec[ORC_BLOCK_PIXELS];
    float R[ORC_MAX_FEATURES + 1][ORC_MAX_FEATURES + 1][3];
    memset(R, 0, sizeof(R));
#define IN(f, sv, id) d[(f) * ORC_BLOCK_PIXELS + (sv) * ORC_LOCAL + (id)]

    /* Scaling, bmfr.cl:511-542 */
    for (int f = NNS; f < NB - 3; ++f) {
        float tmax[ORC_LOCAL], tmin[ORC_LOCAL];
        for (int id = 0; id < ORC_LOCAL; ++id) {
            float mx = -INFINITY, mn = INFINITY;
            for (int sv = 0; sv < 4; ++sv) {
                float v = IN(f, sv, id);
                mx = fmaxf(v, mx);
                mn = fminf(v, mn);
            }
            tmax[id] = mx;
            tmin[id] = mn;
        }
        memcpy(sum_vec, tmax, sizeof(sum_vec));
        const float bmax = orc_tree_max(sum_vec);
        memcpy(sum_vec, tmin, sizeof(sum_vec));
        const float bmin = orc_tree_min(sum_vec);
        const int idx = (blk * NS + f - NNS) * 2;
        mins_maxs[idx + 0] = bmin;
        mins_maxs[idx + 1] = bmax;
        for (int i = 0; i < ORC_BLOCK_PIXELS; ++i)
            d[f * ORC_BLOCK_PIXELS + i] = st_round(p, orc_scale(d[f * ORC_BLOCK_PIXELS + i], bmin, bmax));
    }

    /* Householder QR, bmfr.cl:546-656 */
    for (int col = 0; col < NB; ++col) {
        const int cl = col < NB - 3 ? col : NB - 3;
        for (int id = 0; id < ORC_LOCAL; ++id) {
            float acc = 0.f;
            for (int sv = 0; sv < 4; ++sv) {
                const float t = IN(col, sv, id);
                const int index = id + sv * ORC_LOCAL;
                u_vec[index] = t;
                if (index >= cl + 1) acc = acc + t * t;
            }
            sum_vec[id] = acc;
        }
        float vec_length = orc_tree_sum(sum_vec);
        /* R entries above the pivot are the column's current values
         * (bmfr.cl:575-579); row N of the colour columns is never read
         * by the back substitution and is skipped. */
        if (col < NB - 3) {
            for (int r = 0; r < col; ++r)
                R[r][col][0] = R[r][col][1] = R[r][col][2] = u_vec[r];
        } else {
            for (int r = 0; r < RE - 1; ++r) R[r][RE - 1][col - NF] = u_vec[r];
        }
        /* work-item `col`, bmfr.cl:580-587 */
        float u_length_squared = vec_length;
        vec_length = sqrtf(vec_length + u_vec[cl] * u_vec[cl]);
        u_vec[cl] = u_vec[cl] - vec_length;
        u_length_squared = u_length_squared + u_vec[cl] * u_vec[cl];
        if (col < NB - 3) R[col][col][0] = R[col][col][1] = R[col][col][2] = vec_length;

        /* trailing columns, bmfr.cl:606-655 (CACHE_TMP_DATA=1) */
        for (int fb = cl + 1; fb < NB; ++fb) {
            float cache[ORC_LOCAL][4];
            for (int id = 0; id < ORC_LOCAL; ++id) {
                float acc = 0.f;
                for (int sv = 0; sv < 4; ++sv) {
                    const int index = id + sv * ORC_LOCAL;
                    if (index >= cl) {
                        float t = IN(fb, sv, id);
                        if (col == 0 && fb < NB - 3)
                            t = orc_add_random_n(t, id, sv, fb, frame, p->noise_amount, NB);
                        cache[id][sv] = t;
                        acc = acc + t * u_vec[index];
                    }
                }
                sum_vec[id] = acc;
            }
            const float dot = orc_tree_sum(sum_vec);
            for (int id = 0; id < ORC_LOCAL; ++id) {
                for (int sv = 0; sv < 4; ++sv) {
                    const int index = id + sv * ORC_LOCAL;
                    if (index >= cl) {
                        float v = cache[id][sv];
                        float t = (2.f * u_vec[index]) * dot;
                        t = t / u_length_squared;
                        v = v - t;
                        IN(fb, sv, id) = st_round(p, v);
                    }
                }
            }
        }
    }

    /* Back substitution, bmfr.cl:658-692, per colour channel. */
    for (int ch = 0; ch < 3; ++ch) {
        for (int i = RE - 2; i >= 0; --i) {
            const float divider = R[i][i][ch];
            for (int c = i; c < RE; ++c) R[i][c][ch] = R[i][c][ch] / divider;
            for (int j = i + 1; j < RE - 1; ++j)
                R[i][RE - 1][ch] = R[i][RE - 1][ch] - R[i][j][ch];
            const float wi = R[i][RE - 1][ch];
            for (int r = 0; r <= i; ++r) R[r][i][ch] = R[r][i][ch] * wi;
        }
        for (int f = 0; f < NF; ++f)
            weights[((int64_t)blk * NF + f) * 3 + ch] = R[f][RE - 1][ch];
    }
#undef IN
}

void orc_fitter(const orc_params *p, float *tmp_data, float *mins_maxs, float *weights,
                int32_t frame_number, int32_t blocks_first, int32_t blocks_count) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int32_t b = blocks_first; b < blocks_first + blocks_count; ++b)
        fitter_block(p, tmp_data + (int64_t)b * (orc_feature_count(p) + 3) * ORC_BLOCK_PIXELS, mins_maxs,
                     weights, b, frame_number);
}

void orc_fitter_inputs(const orc_params *p, const float *tmp_data, int32_t block,
                       int32_t frame, double *A, double *b, float *mm) {
    const int NF = orc_feature_count(p), NNS = orc_not_scaled(p), NB = NF + 3;
    const float *d = tmp_data + (int64_t)block * NB * ORC_BLOCK_PIXELS;
    float sum_vec[ORC_LOCAL];
    for (int f = 0; f < NB; ++f) {
        float bmin = 0.f, bmax = 0.f;
        if (f >= NNS && f < NF) {
            float tmax[ORC_LOCAL], tmin[ORC_LOCAL];
            for (int id = 0; id < ORC_LOCAL; ++id) {
                float mx = -INFINITY, mn = INFINITY;
                for (int sv = 0; sv < 4; ++sv) {
                    float v = d[f * ORC_BLOCK_PIXELS + sv * ORC_LOCAL + id];
                    mx = fmaxf(v, mx);
                    mn = fminf(v, mn);
                }
                tmax[id] = mx;
                tmin[id] = mn;
            }
            memcpy(sum_vec, tmax, sizeof(sum_vec));
            bmax = orc_tree_max(sum_vec);
            memcpy(sum_vec, tmin, sizeof(sum_vec));
            bmin = orc_tree_min(sum_vec);
            mm[(f - NNS) * 2 + 0] = bmin;
            mm[(f - NNS) * 2 + 1] = bmax;
        }
        for (int i = 0; i < ORC_BLOCK_PIXELS; ++i) {
            float v = d[f * ORC_BLOCK_PIXELS + i];
            if (f >= NNS && f < NF) v = st_round(p, orc_scale(v, bmin, bmax));
            if (f >= 1 && f < NF)
                v = orc_add_random_n(v, i % ORC_LOCAL, i / ORC_LOCAL, f, frame, p->noise_amount, NB);
            if (f < NF)
                A[(int64_t)i * NF + f] = (double)v;
            else
                b[(int64_t)i * 3 + (f - NF)] = (double)v;
        }
    }
}

double orc_lstsq_fp64(const double *A_in, const double *b_in, int32_t m, int32_t n, int32_t k,
                      double *w) {
    /* Householder QR (v = x + sign(x_0)||x|| e_0, Golub & Van Loan 5.1) on
     * [A | b] in double, then back substitution. */
    double *M = (double *)malloc(sizeof(double) * (size_t)m * (size_t)(n + k));
    for (int32_t i = 0; i < m; ++i) {
        for (int32_t j = 0; j < n; ++j) M[(size_t)i * (n + k) + j] = A_in[(size_t)i * n + j];
        for (int32_t j = 0; j < k; ++j) M[(size_t)i * (n + k) + n + j] = b_in[(size_t)i * k + j];
    }
    const int32_t ld = n + k;
    double *v = (double *)malloc(sizeof(double) * (size_t)m);
    for (int32_t c = 0; c < n; ++c) {
        double norm2 = 0.0;
        for (int32_t i = c; i < m; ++i) norm2 += M[(size_t)i * ld + c] * M[(size_t)i * ld + c];
        double norm = sqrt(norm2);
        double x0 = M[(size_t)c * ld + c];
        double alpha = x0 >= 0 ? -norm : norm;
        for (int32_t i = c; i < m; ++i) v[i] = M[(size_t)i * ld + c];
        v[c] -= alpha;
        double vv = 0.0;
        for (int32_t i = c; i < m; ++i) vv += v[i] * v[i];
        if (vv == 0.0) continue;
        for (int32_t j = c; j < ld; ++j) {
            double dd = 0.0;
            for (int32_t i = c; i < m; ++i) dd += v[i] * M[(size_t)i * ld + j];
            double f = 2.0 * dd / vv;
            for (int32_t i = c; i < m; ++i) M[(size_t)i * ld + j] -= f * v[i];
        }
    }
    double rmax = 0.0, rmin = INFINITY;
    for (int32_t c = 0; c < n; ++c) {
        double r = fabs(M[(size_t)c * ld + c]);
        if (r > rmax) rmax = r;
        if (r < rmin) rmin = r;
    }
    for (int32_t j = 0; j < k; ++j) {
        for (int32_t i = n - 1; i >= 0; --i) {
            double acc = M[(size_t)i * ld + n + j];
            for (int32_t c = i + 1; c < n; ++c) acc -= M[(size_t)i * ld + c] * w[(size_t)c * k + j];
            w[(size_t)i * k + j] = acc / M[(size_t)i * ld + i];
        }
    }
    free(v);
    free(M);
    return rmax > 0 ? rmin / rmax : 0.0;
}

void orc_weighted_sum(const orc_params *p, const float *weights, const float *mins_maxs,
                      float *output, const float *current_normals,
                      const float *current_positions, int32_t frame_number) {
    orc_sizes s;
    orc_sizes_of(p->width, p->height, &s);
    const int32_t W = p->width, H = p->height;
    const int32_t NF = orc_feature_count(p), NNS = orc_not_scaled(p);
    int32_t off[2];
    orc_block_offset(frame_number, off);
#pragma omp parallel for schedule(static)
    for (int32_t y = 0; y < H; ++y) {
        for (int32_t x = 0; x < W; ++x) {
            const int64_t lp = (int64_t)y * W + x;
            /* bmfr.cl:718-722 */
            const int32_t ox = x + ORC_BLOCK / 2 - off[0];
            const int32_t oy = y + ORC_BLOCK / 2 - off[1];
            const int64_t g = (int64_t)(ox / ORC_BLOCK) + (int64_t)(oy / ORC_BLOCK) * s.blocks_x;
            const float wp[3] = {current_positions[lp * 3 + 0], current_positions[lp * 3 + 1],
                                 current_positions[lp * 3 + 2]};
            const float n[3] = {current_normals[lp * 3 + 0], current_normals[lp * 3 + 1],
                                current_normals[lp * 3 + 2]};
            float col[3] = {0.f, 0.f, 0.f};
            for (int f = 0; f < NF; ++f) { /* bmfr.cl:727-747 */
                float v = orc_feature(p, f, n, wp);
                if (f >= NNS) {
                    const int64_t mi = (g * (NF - NNS) + f - NNS) * 2;
                    v = orc_scale(v, mins_maxs[mi + 0], mins_maxs[mi + 1]);
                }
                /* color += weight * feature (bmfr.cl:745): one expression,
                 * which OpenCL's default FP_CONTRACT ON fuses, as the
                 * reference's compiled kernel does (one FMA a term) */
                const float *wt = weights + (g * NF + f) * 3;
                col[0] = fmaf(wt[0], v, col[0]);
                col[1] = fmaf(wt[1], v, col[1]);
                col[2] = fmaf(wt[2], v, col[2]);
            }
            for (int c = 0; c < 3; ++c) output[lp * 3 + c] = col[c] < 0.f ? 0.f : col[c];
        }
    }
}

void orc_accumulate_filtered_data(const orc_params *p, const float *filtered,
                                  const float *ppix, const uint8_t *accept,
                                  const float *albedo, float *tone_mapped,
                                  const uint8_t *current_spp, const float *acc_prev,
                                  float *acc_out, int32_t frame_number) {
    const int32_t W = p->width, H = p->height;
#pragma omp parallel for schedule(static)
    for (int32_t y = 0; y < H; ++y) {
        for (int32_t x = 0; x < W; ++x) {
            const int64_t lp = (int64_t)y * W + x;
            const float fc[3] = {filtered[lp * 3 + 0], filtered[lp * 3 + 1], filtered[lp * 3 + 2]};
            float pc[3] = {0.f, 0.f, 0.f};
            float blend_alpha = 1.f;
            if (frame_number > 0) {
                const uint8_t a = accept[lp];
                if (a > 0) { /* bmfr.cl:789-843 */
                    const float pfx = ppix[lp * 2 + 0], pfy = ppix[lp * 2 + 1];
                    const int32_t ix = cvt_rtn_sat(pfx), iy = cvt_rtn_sat(pfy);
                    const float fx = pfx - (float)ix, fy = pfy - (float)iy;
                    const float omx = 1.f - fx, omy = 1.f - fy;
                    float total_weight = 0.f;
                    const float wts[4] = {omx * omy, fx * omy, omx * fy, fx * fy};
                    for (int i = 0; i < 4; ++i) {
                        const int64_t ls =
                            (int64_t)wrap_add(wrap_mad(wrap_add(iy, i >> 1), W, ix), i & 1);
                        /* stage 1 only sets bits of in-image taps; the range
                         * check covers inconsistent inputs */
                        if ((a & (1u << i)) && ls >= 0 && ls < (int64_t)W * H) {
                            total_weight = total_weight + wts[i];
                            pc[0] = pc[0] + wts[i] * acc_prev[ls * 3 + 0];
                            pc[1] = pc[1] + wts[i] * acc_prev[ls * 3 + 1];
                            pc[2] = pc[2] + wts[i] * acc_prev[ls * 3 + 2];
                        }
                    }
                    if (total_weight > 0.f) {
                        blend_alpha = 1.f / (float)current_spp[lp];
                        blend_alpha = fmaxf(blend_alpha, p->second_blend_alpha);
                        pc[0] = pc[0] / total_weight;
                        pc[1] = pc[1] / total_weight;
                        pc[2] = pc[2] / total_weight;
                    }
                }
            }
            const float omb = 1.f - blend_alpha;
            for (int c = 0; c < 3; ++c) {
                const float ac = blend_alpha * fc[c] + omb * pc[c];
                acc_out[lp * 3 + c] = ac;
                /* bmfr.cl:852-856 */
                const float v = fmaxf(0.f, albedo[lp * 3 + c] * ac);
                tone_mapped[lp * 3 + c] = fminf(fmaxf(powf(v, 0.454545f), 0.f), 1.f);
            }
        }
    }
}

void orc_taa(const orc_params *p, const float *ppix, const float *new_frame, float *result,
             const float *prev_frame, int32_t frame_number) {
    const int32_t W = p->width, H = p->height;
    const float ta = p->taa_blend_alpha;
#pragma omp parallel for schedule(static)
    for (int32_t y = 0; y < H; ++y) {
        for (int32_t x = 0; x < W; ++x) {
            const int64_t lp = (int64_t)y * W + x;
            const float mine[3] = {new_frame[lp * 3 + 0], new_frame[lp * 3 + 1],
                                   new_frame[lp * 3 + 2]};
            const float pfx = ppix[lp * 2 + 0], pfy = ppix[lp * 2 + 1];
            const int32_t ix = cvt_rtn_sat(pfx), iy = cvt_rtn_sat(pfy);
            if (frame_number == 0 || ix < -1 || iy < -1 || ix >= W || iy >= H) {
                for (int c = 0; c < 3; ++c) result[lp * 3 + c] = mine[c];
                continue;
            }
            float mnb[3] = {INFINITY, INFINITY, INFINITY}, mnc[3] = {INFINITY, INFINITY, INFINITY};
            float mxb[3] = {-INFINITY, -INFINITY, -INFINITY},
                  mxc[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int dy = -1; dy < 2; ++dy) { /* bmfr.cl:897-920 */
                for (int dx = -1; dx < 2; ++dx) {
                    const int32_t sx = x + dx, sy = y + dy;
                    if (sx >= 0 && sy >= 0 && sx < W && sy < H) {
                        float sc[3], yc[3];
                        if (dx == 0 && dy == 0) {
                            sc[0] = mine[0]; sc[1] = mine[1]; sc[2] = mine[2];
                        } else {
                            const int64_t ls = (int64_t)sx + (int64_t)sy * W;
                            sc[0] = new_frame[ls * 3 + 0];
                            sc[1] = new_frame[ls * 3 + 1];
                            sc[2] = new_frame[ls * 3 + 2];
                        }
                        orc_rgb_to_ycocg(sc, yc);
                        for (int c = 0; c < 3; ++c) {
                            if (dx == 0 || dy == 0) {
                                mnc[c] = fminf(mnc[c], yc[c]);
                                mxc[c] = fmaxf(mxc[c], yc[c]);
                            }
                            mnb[c] = fminf(mnb[c], yc[c]);
                            mxb[c] = fmaxf(mxb[c], yc[c]);
                        }
                    }
                }
            }
            /* bmfr.cl:924-962 */
            float pc[3] = {0.f, 0.f, 0.f};
            float tw = 0.f;
            const float fx = pfx - (float)ix, fy = pfy - (float)iy;
            const float omx = 1.f - fx, omy = 1.f - fy;
            const int64_t r0 = (int64_t)iy * W, r1 = (int64_t)(iy + 1) * W;
            if (iy >= 0) {
                if (ix >= 0) {
                    const float w = omx * omy;
                    for (int c = 0; c < 3; ++c) pc[c] = pc[c] + w * prev_frame[(r0 + ix) * 3 + c];
                    tw = tw + w;
                }
                if (ix < W - 1) {
                    const float w = fx * omy;
                    for (int c = 0; c < 3; ++c) pc[c] = pc[c] + w * prev_frame[(r0 + ix + 1) * 3 + c];
                    tw = tw + w;
                }
            }
            if (iy < H - 1) {
                if (ix >= 0) {
                    const float w = omx * fy;
                    for (int c = 0; c < 3; ++c) pc[c] = pc[c] + w * prev_frame[(r1 + ix) * 3 + c];
                    tw = tw + w;
                }
                if (ix < W - 1) {
                    const float w = fx * fy;
                    for (int c = 0; c < 3; ++c) pc[c] = pc[c] + w * prev_frame[(r1 + ix + 1) * 3 + c];
                    tw = tw + w;
                }
            }
            for (int c = 0; c < 3; ++c) pc[c] = pc[c] / tw; /* unguarded, bmfr.cl:962 */
            float pyc[3], lo[3], hi[3], cl[3], prgb[3];
            orc_rgb_to_ycocg(pc, pyc);
            for (int c = 0; c < 3; ++c) {
                lo[c] = (mnb[c] + mnc[c]) / 2.f;
                hi[c] = (mxb[c] + mxc[c]) / 2.f;
                cl[c] = fminf(fmaxf(pyc[c], lo[c]), hi[c]);
            }
            orc_ycocg_to_rgb(cl, prgb);
            const float omt = 1.f - ta;
            for (int c = 0; c < 3; ++c) result[lp * 3 + c] = ta * mine[c] + omt * prgb[c];
        }
    }
}
