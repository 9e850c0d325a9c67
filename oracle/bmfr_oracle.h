/*
 * CPU oracle for the BMFR five-stage frame chain.  TEST INFRASTRUCTURE ONLY:
 * this code is the checker that the HIP path is compared against.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it;
 * the product library never links or calls it.
 *
 * It is a scalar C restatement of the reference kernels in opencl/bmfr.cl,
 * following their operation order, with the host parameters of
 * opencl/bmfr.cpp:56-118.  Every function cites the reference lines it
 * restates.  Floating point is evaluated without contraction
 * (-ffp-contract=off) so that results are reproducible bit for bit.
 *
 * Deliberate semantic decisions (the reference is racy or implementation
 * defined at these points):
 *   - accumulate_noisy_data reads the uploaded noisy colour and writes the
 *     accumulated colour to a separate buffer (read-before-write); the
 *     reference overwrites current_noisy in place (bmfr.cl:322 vs 481).
 *   - float->int conversions of reprojected coordinates saturate and map
 *     NaN to 0 (the behaviour of AMD's v_cvt_i32_f32); integer sums wrap.
 *   - OpenCL dot() is evaluated left to right: ((a0*b0 + a1*b1) + a2*b2)...
 *   - max(0,x)/clamp() in tone mapping use fmax/fmin (NaN -> bound).
 */
#ifndef BMFR_ORACLE_H
#define BMFR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_BLOCK 32
#define ORC_BLOCK_PIXELS 1024
#define ORC_BUFFER_COUNT 13      /* 10 features + RGB, bmfr.cpp:193-202 */
#define ORC_FEATURES 10
#define ORC_NOT_SCALED 4         /* 1, n.x, n.y, n.z */
#define ORC_SCALED 6             /* p.xyz, p.xyz^2 */
#define ORC_R_EDGE 11            /* bmfr.cpp:221 */
#define ORC_MAX_FEATURES 16      /* longest configurable feature set */
#define ORC_LOCAL 256            /* bmfr.cpp:116 */

typedef struct {
    int32_t width;
    int32_t height;
    float position_limit_squared;
    float normal_limit_squared;
    float blend_alpha;          /* BLEND_ALPHA 0.2f */
    float second_blend_alpha;   /* SECOND_BLEND_ALPHA 0.1f */
    float taa_blend_alpha;      /* TAA_BLEND_ALPHA 0.2f */
    double noise_amount;        /* NOISE_AMOUNT 1e-2 (a double literal) */
    int32_t half_tmp;           /* USE_HALF_PRECISION_IN_TMP_DATA */
    /* The feature set (bmfr.cpp:63-77, 193-214).  feature_count 0 is the
     * reference's 10 (ORC_FEATURES/ORC_NOT_SCALED above).  Otherwise feature
     * i is factor[i][0] * factor[i][1] over {1, normal.xyz,
     * world_position.xyz} (ids 0..6), the first features_not_scaled unscaled:
     * BUFFER_COUNT = feature_count + 3, R_EDGE = feature_count + 1. */
    int32_t feature_count;
    int32_t features_not_scaled;
    uint8_t feature_factors[ORC_MAX_FEATURES][2];
} orc_params;

/* N, FEATURES_NOT_SCALED and BUFFER_COUNT of a params set. */
int32_t orc_feature_count(const orc_params *p);
int32_t orc_not_scaled(const orc_params *p);
/* Feature f of a pixel with normal n and world position wp, as the
 * reference's FEATURE_BUFFERS expression evaluates it (bmfr.cl:448-453). */
float orc_feature(const orc_params *p, int32_t f, const float n[3], const float wp[3]);

/* Derived sizes, bmfr.cpp:104-118. */
typedef struct {
    int32_t workset_w, workset_h;   /* WORKSET_WIDTH/HEIGHT */
    int32_t margin_w, margin_h;     /* WORKSET_WITH_MARGINS_WIDTH/HEIGHT */
    int32_t blocks_x, blocks_y;     /* margin_w/32, margin_h/32 */
    int32_t blocks;                 /* FITTER_GLOBAL / 256 */
    int64_t pixels;                 /* P = W*H */
    int64_t padded;                 /* Q = margin_w*margin_h */
} orc_sizes;

void orc_sizes_of(int32_t w, int32_t h, orc_sizes *out);
void orc_default_params(int32_t w, int32_t h, orc_params *out);

/* bmfr.cl:267-285 */
void orc_block_offset(int32_t frame, int32_t out[2]);
/* bmfr.cl:209-216 */
int32_t orc_mirror(int32_t index, int32_t size);
/* bmfr.cl:162-171: 32-bit integer hash, returned before the float divide */
uint32_t orc_hash(uint32_t a);
float orc_random(uint32_t a);
/* bmfr.cl:173-182 */
float orc_add_random(float value, int32_t id, int32_t sub_vector, int32_t feature,
                     int32_t frame, double noise_amount);
/* The same for BUFFER_COUNT buffer_count (orc_add_random: 13). */
float orc_add_random_n(float value, int32_t id, int32_t sub_vector, int32_t feature,
                       int32_t frame, double noise_amount, int32_t buffer_count);
/* IEEE binary16 conversion, round to nearest even (vstore_half). */
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);
float orc_round_half(float f);
/* bmfr.cl:200-205 */
float orc_scale(float value, float mn, float mx);
/* bmfr.cl:184-198 */
void orc_rgb_to_ycocg(const float rgb[3], float out[3]);
void orc_ycocg_to_rgb(const float ycc[3], float out[3]);
/* bmfr.cl:26-87: the reference's 256->64->8->1 tree over sum_vec[256] */
float orc_tree_sum(float *sum_vec);
float orc_tree_min(float *sum_vec);
float orc_tree_max(float *sum_vec);

/* Smallest W, H for which stage 1's mirroring (bmfr.cl:209-216, 312-316)
 * maps every padded-grid coordinate (up to WORKSET+29) into the image. */
#define ORC_MIN_SIZE 47

/* Stage 1, bmfr.cl:287-485.  Returns 0, or -1 (and does nothing) when W or
 * H is below ORC_MIN_SIZE.  Buffers are AoS float3 with row stride W over
 * the first P texels; tmp_data has (N+3)*Q floats in block-major order
 * [by][bx][feature][y_in][x_in] (bmfr.cl:455-464).  In half mode the values
 * written to tmp_data are binary16-representable floats. */
int32_t orc_accumulate_noisy_data(
    const orc_params *p,
    const float *current_normals, const float *previous_normals,
    const float *current_positions, const float *previous_positions,
    const float *current_noisy,          /* uploaded colour, read only */
    const float *previous_noisy,         /* accumulated colour of frame-1 */
    const uint8_t *previous_spp,
    float *accumulated_noisy,            /* out: accumulated colour (3P) */
    float *out_prev_frame_pixel,         /* out: float2[P] */
    uint8_t *accept_bools,               /* out: uchar[P] */
    uint8_t *current_spp,                /* out: uchar[P] */
    float *tmp_data,                     /* out: (N+3)*Q */
    const float prev_frame_camera_matrix[16],
    const float pixel_offset[2],
    int32_t frame_number);

/* Stage 2, bmfr.cl:487-700.  tmp_data is updated in place like the
 * reference (scaled features, then Householder updates).  mins_maxs gets
 * 2S floats per block (S scaled features, 6 in the reference's set),
 * weights 3N per block ([block][feature][rgb]).
 * blocks_first/blocks_count select a sub-range of blocks (for sampling). */
void orc_fitter(const orc_params *p, float *tmp_data, float *mins_maxs,
                float *weights, int32_t frame_number,
                int32_t blocks_first, int32_t blocks_count);

/* The fitter's regression input for one block, exactly as the QR of
 * bmfr.cl:549-656 sees it at column 0: the scaled features scaled
 * (bmfr.cl:538-541, half-rounded in half mode), features 1..N-1 plus noise
 * (bmfr.cl:623-627), colour unchanged.  A is 1024xN row-major, b 1024x3.
 * Also returns the block's 2S mins/maxs.  tmp_data is not modified. */
void orc_fitter_inputs(const orc_params *p, const float *tmp_data, int32_t block,
                       int32_t frame_number, double *A, double *b, float *mins_maxs);

/* FP64 least squares min ||A w - b|| by Householder QR with the stable
 * sign choice; A is m x n row-major (n <= 16), b m x k.  w is n x k.
 * Returns the smallest |R_ii| / max |R_jj|. */
double orc_lstsq_fp64(const double *A, const double *b, int32_t m, int32_t n,
                      int32_t k, double *w);

/* Stage 3, bmfr.cl:703-758 */
void orc_weighted_sum(const orc_params *p, const float *weights,
                      const float *mins_maxs, float *output,
                      const float *current_normals, const float *current_positions,
                      int32_t frame_number);

/* Stage 4, bmfr.cl:761-857 */
void orc_accumulate_filtered_data(const orc_params *p, const float *filtered_frame,
                                  const float *in_prev_frame_pixel,
                                  const uint8_t *accept_bools, const float *albedo,
                                  float *tone_mapped_frame, const uint8_t *current_spp,
                                  const float *accumulated_prev_frame,
                                  float *accumulated_frame, int32_t frame_number);

/* Stage 5, bmfr.cl:860-974 */
void orc_taa(const orc_params *p, const float *in_prev_frame_pixel,
             const float *new_frame, float *result_frame, const float *prev_frame,
             int32_t frame_number);

#ifdef __cplusplus
}
#endif
#endif
