/*
 * bmfr_hip.h -- C ABI of the MI355X (gfx950) BMFR denoiser.
 *
 * The five stage calls replace the five OpenCL kernels of the reference and
 * keep their names, order and arguments' meaning (opencl/bmfr.cl:290, 490,
 * 703, 761, 860, enqueued per frame at opencl/bmfr.cpp:429-476).  The
 * reference's compile-time -D configuration (opencl/bmfr.cpp:32-118,
 * 205-232) becomes the runtime struct bmfr_params, and the reference's
 * cl::Buffer set with its Double_buffer swaps (opencl/bmfr.cpp:120-135,
 * 315-347, 482-484) becomes a bmfr_stream that owns its device buffers.
 *
 * Conventions: plain C types only; every call returns a bmfr_status (0 = OK)
 * and never throws or exits; bmfr_last_error() gives a message for the last
 * failure on the calling thread.  All image buffers are AoS float3 (RGB or
 * XYZ) with row stride `width`, matching opencl/bmfr.cl:224-241.  Stage
 * calls are asynchronous on the stream's HIP stream and ordered like the
 * reference's in-order queue (opencl/bmfr.cpp:191).  A stream handle is
 * driven by one host thread at a time.
 */
#ifndef BMFR_HIP_H
#define BMFR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BMFR_API_VERSION 3

typedef enum bmfr_status {
    BMFR_OK = 0,
    BMFR_ERR_INVALID_ARGUMENT = 1, /* bad size, null pointer, unknown enum */
    BMFR_ERR_OUT_OF_BUDGET = 2,    /* admission refused by the HBM budget */
    BMFR_ERR_HIP = 3,              /* a HIP runtime call failed */
    BMFR_ERR_NO_DEVICE = 4,        /* no usable gfx950 device */
    BMFR_ERR_STATE = 5             /* call not valid in the current state */
} bmfr_status;

/* Storage precision of G-buffers and temporal histories. */
typedef enum bmfr_storage { BMFR_STORAGE_F32 = 0, BMFR_STORAGE_F16 = 1 } bmfr_storage;

/* Fitter arithmetic. */
typedef enum bmfr_qr_mode {
    BMFR_QR_FP32 = 0,         /* plain FP32 Householder, like bmfr.cl:549-656 */
    BMFR_QR_KAHAN = 1,        /* FP32 with compensated norm and dot sums */
    BMFR_QR_KAHAN_FP64 = 2    /* as KAHAN, plus FP64 re-factorisation of blocks
                                 whose min |R_ii| < refactor_threshold*max |R_ii| */
} bmfr_qr_mode;

/* Feature sets (opencl/bmfr.cpp:63-77, 193-214).  The reference configures
 * its fitted features as OpenCL expression text, NOT_SCALED_FEATURE_BUFFERS
 * then SCALED_FEATURE_BUFFERS, over the variables `normal` and
 * `world_position` that accumulate_noisy_data and weighted_sum both hold
 * (bmfr.cl:448-453, 725-729).  Here a feature is the product of two factors
 * of that pixel, feature = factor[0] * factor[1] (BMFR_FACTOR_ONE for a
 * single term: x * 1 is exact), which spans the reference's default set and
 * every polynomial of degree <= 2 in those six channels.  The first
 * `features_not_scaled` features enter the fit as they are; the others are
 * scaled to [0, 1] per block (bmfr.cl:511-542). */
#define BMFR_MAX_FEATURES 16
typedef enum bmfr_factor {
    BMFR_FACTOR_ONE = 0,
    BMFR_FACTOR_NORMAL_X = 1,
    BMFR_FACTOR_NORMAL_Y = 2,
    BMFR_FACTOR_NORMAL_Z = 3,
    BMFR_FACTOR_POSITION_X = 4,
    BMFR_FACTOR_POSITION_Y = 5,
    BMFR_FACTOR_POSITION_Z = 6
} bmfr_factor;

/* Runtime form of the reference's -D options (opencl/bmfr.cpp:56-118, 205-232). */
typedef struct bmfr_params {
    int32_t width;                  /* IMAGE_WIDTH */
    int32_t height;                 /* IMAGE_HEIGHT */
    float position_limit_squared;   /* POSITION_LIMIT_SQUARED */
    float normal_limit_squared;     /* NORMAL_LIMIT_SQUARED */
    float blend_alpha;              /* BLEND_ALPHA (0.2) */
    float second_blend_alpha;       /* SECOND_BLEND_ALPHA (0.1) */
    float taa_blend_alpha;          /* TAA_BLEND_ALPHA (0.2) */
    double noise_amount;            /* NOISE_AMOUNT (1e-2, applied in double) */
    int32_t storage;                /* bmfr_storage of normals, positions, noisy
                                       colour and the accumulation/TAA histories */
    int32_t tmp_storage;            /* bmfr_storage of the fitter input tmp_data;
                                       F16 is USE_HALF_PRECISION_IN_TMP_DATA=1 */
    int32_t qr_mode;                /* bmfr_qr_mode */
    float refactor_threshold;       /* used by BMFR_QR_KAHAN_FP64 (1e-6) */
    /* Fitted features, BUFFER_COUNT - 3 of bmfr.cpp:204 (1..BMFR_MAX_FEATURES);
     * 0 selects the reference's set: 1, normal.xyz not scaled, then
     * world_position.xyz and its squares scaled (bmfr.cpp:65-77).  Another set
     * runs with qr_mode BMFR_QR_FP32 on single streams (not in batches or
     * servers). */
    int32_t feature_count;
    int32_t features_not_scaled;    /* FEATURES_NOT_SCALED (0..feature_count) */
    uint8_t feature_factors[BMFR_MAX_FEATURES][2];   /* bmfr_factor pairs */
} bmfr_params;

/* Buffer ids for the debug/parity accessors.  Element types: F = float or
 * half per `storage`, T = float or half per `tmp_storage`, u8 = uint8.  With
 * N features (10 for the reference's set) of which S are scaled (6), a
 * block holds N + 3 tmp_data columns, 3N weights and 2S mins_maxs. */
typedef enum bmfr_buffer {
    BMFR_BUF_ALBEDO = 0,          /* float[3P]   albedo_buffer */
    BMFR_BUF_NORMALS_CUR = 1,     /* F[3P]       normals_buffer.current() */
    BMFR_BUF_NORMALS_PREV = 2,    /* F[3P]       normals_buffer.previous() */
    BMFR_BUF_POSITIONS_CUR = 3,   /* F[3P] */
    BMFR_BUF_POSITIONS_PREV = 4,  /* F[3P] */
    BMFR_BUF_NOISY_IN = 5,        /* F[3P]       uploaded noisy colour */
    BMFR_BUF_NOISY_ACC_CUR = 6,   /* F[3P]       accumulated noisy colour */
    BMFR_BUF_NOISY_ACC_PREV = 7,  /* F[3P]       previous_noisy of stage 1 */
    BMFR_BUF_SPP_CUR = 8,         /* u8[P] */
    BMFR_BUF_SPP_PREV = 9,        /* u8[P] */
    BMFR_BUF_PREV_PIXEL = 10,     /* float2[P]   prev_pixels_buffer */
    BMFR_BUF_ACCEPT = 11,         /* u8[P]       accept_buffer */
    BMFR_BUF_TMP_DATA = 12,       /* T[(N+3)Q]   in_buffer (block-major) */
    BMFR_BUF_WEIGHTS = 13,        /* float[3N nB] */
    BMFR_BUF_MINS_MAXS = 14,      /* float[2S nB] */
    BMFR_BUF_FILTERED = 15,       /* float[3P] */
    BMFR_BUF_TONE_MAPPED = 16,    /* float[3P] */
    BMFR_BUF_OUT_CUR = 17,        /* F[3P]       out_buffer.current() */
    BMFR_BUF_OUT_PREV = 18,       /* F[3P] */
    BMFR_BUF_RESULT_CUR = 19,     /* F[3P]       result_buffer.current() */
    BMFR_BUF_RESULT_PREV = 20,    /* F[3P] */
    BMFR_BUF_COUNT = 21
} bmfr_buffer;

typedef struct bmfr_stream_s *bmfr_stream;
typedef struct bmfr_server_s *bmfr_server;

/* Sizes derived exactly as opencl/bmfr.cpp:104-118. */
typedef struct bmfr_sizes {
    int32_t workset_width, workset_height;   /* WORKSET_WIDTH/HEIGHT */
    int32_t margin_width, margin_height;     /* WORKSET_WITH_MARGINS_* */
    int32_t blocks_x, blocks_y, blocks;      /* fitter work-groups */
    int64_t pixels;                          /* P = W*H */
    int64_t padded_pixels;                   /* Q = margin_width*margin_height */
    int64_t device_bytes;                    /* per-stream device footprint */
} bmfr_sizes;

/* ---- library ------------------------------------------------------------ */
int32_t bmfr_api_version(void);
const char *bmfr_status_string(int32_t status);
const char *bmfr_last_error(void);
/* Number of visible HIP devices (0 without a GPU); never fails. */
int32_t bmfr_device_count(void);
/* Measured LDS read bandwidth of the current device in bytes/s: 16-byte
 * conflict-free reads from a 32 KB tile by eight 256-wide work-groups per
 * CU, `steps` x 64 bytes per work-item, five timed launches.  No counterpart
 * in the reference; it is the ceiling against which the fitter's work is
 * stated (the reference's fitter streams its block matrix through LDS). */
int32_t bmfr_measure_lds_read_bandwidth(int32_t steps, double *bytes_per_second);

/* Reference defaults (opencl/bmfr.cpp:58-62, 88) for a W x H stream:
 * F32 storage, F16 tmp_data, plain FP32 QR, the reference's feature set. */
int32_t bmfr_default_params(int32_t width, int32_t height, bmfr_params *out);
/* Sets params' feature set from the reference's two feature strings
 * (NOT_SCALED_FEATURE_BUFFERS, SCALED_FEATURE_BUFFERS, opencl/bmfr.cpp:65-77;
 * counted like bmfr.cpp:193-201).  Comma-separated terms; a term is `1.f`
 * (also `1`, `1.0`, `1.0f`), `normal.x|y|z`, `world_position.x|y|z`, or the
 * product `a*b` of two of those; spaces are ignored and an empty term (the
 * reference's trailing comma) is skipped.  Anything else, or more than
 * BMFR_MAX_FEATURES terms, is BMFR_ERR_INVALID_ARGUMENT naming the term. */
int32_t bmfr_parse_feature_buffers(const char *not_scaled, const char *scaled, bmfr_params *params);
/* Validates params and fills sizes.  W and H must be >= 47: stage 1 mirrors
 * padded-grid coordinates up to WORKSET+29 back into the image
 * (opencl/bmfr.cl:209-216, 312-316), which needs WORKSET+30 <= 2*size. */
int32_t bmfr_compute_sizes(const bmfr_params *params, bmfr_sizes *out);

/* ---- stream lifecycle (replaces tasks() set-up, opencl/bmfr.cpp:179-384) */
int32_t bmfr_stream_open(const bmfr_params *params, bmfr_stream *out);
int32_t bmfr_stream_close(bmfr_stream stream);
int32_t bmfr_stream_sizes(bmfr_stream stream, bmfr_sizes *out);
/* Clears all temporal history so the next frame behaves like frame 0. */
int32_t bmfr_stream_reset(bmfr_stream stream);

/* Fitter launch shape: rows_per_lane 4, 8 or 16 (a 32x32 block spread over
 * 4, 2 or 1 wavefronts) and blocks_per_group 1, 2 or 4 (supported pairs:
 * 4x1, 4x2, 8x1, 8x2, 16x1, 16x2, 16x4; default 16x4, one wavefront per
 * block and four blocks per work-group, so that a frame with fewer blocks
 * than SIMDs gets one block per SIMD).  Shapes with the same rows_per_lane
 * give the same bits.  A stream with a configured feature set always runs
 * 16x2.  No counterpart in the reference, whose fitter is
 * always one 256-wide work-group per block (opencl/bmfr.cpp:117-118, 249). */
int32_t bmfr_set_fitter_config(bmfr_stream stream, int32_t rows_per_lane, int32_t blocks_per_group);

/* ---- per-frame inputs (opencl/bmfr.cpp:420-427) -------------------------- */
/* Host arrays of W*H*3 floats; converted to the stream's storage.  The call
 * returns once the data is on the device. */
int32_t bmfr_upload_frame(bmfr_stream stream, const float *albedo, const float *normals,
                          const float *positions, const float *noisy);

/* Device-to-device form of bmfr_upload_frame for a renderer whose G-buffers
 * are already in HBM: albedo as float[3P], the other three in the stream's
 * storage type (float or half [3P]); 16-byte aligned.  Enqueued on the
 * stream (asynchronous); the sources must stay valid until it has run. */
int32_t bmfr_copy_frame_inputs_device(bmfr_stream stream, const void *albedo, const void *normals,
                                      const void *positions, const void *noisy);

/* ---- renderer handoff (opencl/bmfr.cpp:415-416, 478-480) -------------------
 * The reference uploads every frame from the host and reads the result back,
 * and notes that "in real use case there would not be WriteBuffer and
 * ReadBuffer function calls because the input data comes from the path
 * tracer and output goes to the screen".  These calls are that case. */
/* Zero-copy inputs: the stream's next frame (its next
 * bmfr_accumulate_noisy_data, bmfr_run_frame, batch step or server step, up
 * to the bmfr_end_frame that closes it) reads these caller-owned device
 * buffers instead of the stream's own input buffers: albedo as float[3P],
 * normals, positions and noisy colour in the stream's storage type (float or
 * half [3P]), AoS with row stride `width`, 16-byte aligned.  No copy is made.
 * The binding covers one frame; after it the stream uses its own buffers
 * again unless bound anew.  The frame's normals and positions are also the
 * NEXT frame's previous G-buffers (read by its accumulate_noisy_data, as
 * normals_buffer.previous() in opencl/bmfr.cpp:431-434): keep them unchanged
 * until that work has completed (a renderer with two G-buffer sets used in
 * turn does this naturally).  All four null: drop the binding.  The debug
 * accessors (bmfr_read_buffer and friends) still name the stream's own
 * buffers. */
int32_t bmfr_bind_device_inputs(bmfr_stream stream, const void *albedo, const void *normals,
                                const void *positions, const void *noisy);

/* Output surfaces: every later frame's TAA result (the value stored in the
 * result history, i.e. what bmfr_download_result returns) is also written
 * into a caller-owned device surface, by the kernel that computes it: row y
 * of the image starts at surface + y * row_pitch_bytes. */
typedef enum bmfr_surface_format {
    BMFR_SURFACE_RGB_F32 = 0,   /* float[3] per pixel (the reference's out_data) */
    BMFR_SURFACE_RGBA_F32 = 1,  /* float[4], alpha 1 */
    BMFR_SURFACE_RGBA_F16 = 2,  /* half[4] (round to nearest even), alpha 1 */
    BMFR_SURFACE_RGBA_U8 = 3    /* uint8[4]: round(clamp(v, 0, 1) * 255) per channel, the
                                   8-bit conversion of the reference's PNG output
                                   (opencl/bmfr.cpp:519-547, bmfr_png_write_rgb), alpha 255 */
} bmfr_surface_format;
/* Binds (surface non-null) or unbinds (null) the stream's output surface;
 * the binding lasts until changed.  row_pitch_bytes must hold a row
 * (>= width x pixel size) and be a multiple of the pixel's alignment (4 for
 * U8 and F32 formats, 8 for RGBA_F16); the surface must be aligned likewise.
 * The surface is written on the stream's HIP stream (bmfr_stream_hip_stream)
 * or, for batch and server steps, on theirs, which the stream's later work
 * waits for. */
int32_t bmfr_bind_output_surface(bmfr_stream stream, void *surface, int32_t format,
                                 int64_t row_pitch_bytes);
/* The HIP stream (hipStream_t) the stream's work is enqueued on, so that a
 * renderer or compositor can order its own work against it with events. */
int32_t bmfr_stream_hip_stream(bmfr_stream stream, void **hip_stream);

/* ---- the five stages (opencl/bmfr.cpp:429-476) ---------------------------- */
/* prev_frame_camera_matrix: 16 floats, the previous frame's view-projection
 * in the reference's row-vector layout (camera_matrices[max(f-1,0)],
 * opencl/bmfr.cpp:440-442); pixel_offset: 2 floats (pixel_offsets[f]). */
int32_t bmfr_accumulate_noisy_data(bmfr_stream stream, int32_t frame_number,
                                   const float *prev_frame_camera_matrix,
                                   const float *pixel_offset);
int32_t bmfr_fitter(bmfr_stream stream, int32_t frame_number);
int32_t bmfr_weighted_sum(bmfr_stream stream, int32_t frame_number);
/* FP64 least-squares weights of every block of the last fitter call, from
 * the same scaled and noised matrix (W = float[30 nB]); for stating the
 * fitter's weight error against FP64.  No counterpart in the reference. */
int32_t bmfr_fitter_fp64_weights(bmfr_stream stream, float *weights_out);
/* The fitter over caller-owned buffers, allocating nothing.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Fits `blocks` 32x32 blocks of tmp_data (block-major, the
 * reference's 13 columns, float or half per tmp_storage, 16-byte aligned)
 * for frame `frame_number` exactly as bmfr_fitter does (same kernels, same
 * bits) and writes weights (float[30 blocks]) and mins_maxs (float[12
 * blocks]); enqueued on `hip_stream` (a hipStream_t; null: the null stream).
 * Its scratch (the frame's noise table and the FP64 re-factorisation flags,
 * 100 KB) is a fixed device array each GPU holds until the process ends:
 * at most 16384 blocks per call (larger is BMFR_ERR_INVALID_ARGUMENT), and
 * calls that overlap on one GPU race on it. */
int32_t bmfr_fit_blocks(const void *tmp_data, int32_t tmp_storage, int32_t blocks, int32_t frame_number,
                        double noise_amount, int32_t qr_mode, float refactor_threshold, float *weights,
                        float *mins_maxs, void *hip_stream);
/* Blocks the last BMFR_QR_KAHAN_FP64 fitter call re-factorised in FP64. */
int32_t bmfr_refactor_count(bmfr_stream stream, int32_t *count);
int32_t bmfr_accumulate_filtered_data(bmfr_stream stream, int32_t frame_number);
int32_t bmfr_taa(bmfr_stream stream, int32_t frame_number);
/* Swaps the double buffers (opencl/bmfr.cpp:482-484). */
int32_t bmfr_end_frame(bmfr_stream stream);
/* How bmfr_run_frame, batches and server steps run stages 3-5.  Every level
 * gives the same bits; a fused kernel does not write the intermediates it
 * keeps on chip (filtered, and for BMFR_FUSE_345 also tone_mapped), and
 * bmfr_stage_times reports 0 ms for the stages folded into it, with the
 * fused kernel's time under stage 4 (accumulate_filtered_data).  The five
 * separate entry points always run the five separate kernels. */
typedef enum bmfr_fuse_level {
    BMFR_FUSE_NONE = 0,  /* weighted_sum, accumulate_filtered_data, taa */
    BMFR_FUSE_34 = 1,    /* weighted_sum inside accumulate_filtered_data, then taa */
    BMFR_FUSE_345 = 2    /* all three in one kernel per 32x32 block, with the
                            one-pixel TAA ring recomputed (the default of
                            streams, batches and server steps) */
} bmfr_fuse_level;
int32_t bmfr_set_fused(bmfr_stream stream, int32_t level);
/* All five stages in order, then bmfr_end_frame (opencl/bmfr.cpp:429-484). */
int32_t bmfr_run_frame(bmfr_stream stream, int32_t frame_number,
                       const float *prev_frame_camera_matrix, const float *pixel_offset);

/* ---- outputs ---------------------------------------------------------------- */
/* The last completed frame's TAA result (result_buffer of the frame before
 * the most recent bmfr_end_frame), W*H*3 floats (opencl/bmfr.cpp:479-480). */
int32_t bmfr_download_result(bmfr_stream stream, float *out);
int32_t bmfr_synchronize(bmfr_stream stream);

/* Debug/parity access to every device buffer, in its native element type. */
int32_t bmfr_buffer_bytes(bmfr_stream stream, int32_t buffer, size_t *bytes);
int32_t bmfr_read_buffer(bmfr_stream stream, int32_t buffer, void *host, size_t bytes);
int32_t bmfr_write_buffer(bmfr_stream stream, int32_t buffer, const void *host, size_t bytes);
/* Device address of a buffer, for zero-copy producers (a renderer writing the
 * current G-buffers directly).  Valid until the next bmfr_end_frame. */
int32_t bmfr_buffer_device_ptr(bmfr_stream stream, int32_t buffer, void **dptr);

/* ---- timing (opencl/bmfr.cpp:386-412, 488-517) -------------------------------- */
/* Level 1: HIP events bracket every stage on the stream's HIP stream.
 * Level 2: events only before the first and after the last stage (nothing
 * recorded between kernels; per-stage values read -1).  bmfr_stage_times
 * waits for the last frame and returns 6 values in ms: accumulate_noisy_data,
 * fitter, weighted_sum, accumulate_filtered_data, taa, and the total from the
 * first stage's start to the last stage's end (the reference's "Total time in
 * all kernels"). */
int32_t bmfr_set_timing(bmfr_stream stream, int32_t enable);
int32_t bmfr_stage_times(bmfr_stream stream, float *ms6);

/* Markers 0..7: HIP events recorded on the stream's HIP stream, to time any
 * enqueued work; bmfr_marker_elapsed waits for m1 and returns ms. */
int32_t bmfr_record_marker(bmfr_stream stream, int32_t marker);
int32_t bmfr_marker_elapsed(bmfr_stream stream, int32_t m0, int32_t m1, float *ms);

/* ---- batches: many streams, one launch per stage ------------------------------
 * A batch runs one frame of each member stream per step: each of the five
 * stages is one launch over the concatenated tiles/blocks of every member
 * (one launch per storage type present), the fitter a persistent launch over
 * all members' blocks, edge blocks included.  No counterpart in the
 * reference, which runs one stream with five launches per frame
 * (opencl/bmfr.cpp:417-485); each member's results are those of running it
 * alone.  Members keep their own buffers and history; a stream belongs to at
 * most one batch at a time and must not be stepped on its own meanwhile. */
typedef struct bmfr_batch_s *bmfr_batch;
int32_t bmfr_batch_create(const bmfr_stream *streams, int32_t count, bmfr_batch *out);
int32_t bmfr_batch_destroy(bmfr_batch batch);
/* One step: frames[i], matrices[16*i..], offsets[2*i..] are member i's
 * bmfr_accumulate_noisy_data arguments.  Ordered after each member's pending
 * work; each member's later work waits for it; members' double buffers are
 * swapped.  timed != 0 records events: around every stage of an eager step
 * (run stage by stage), or around the whole replay when a graph is used. */
int32_t bmfr_batch_run(bmfr_batch batch, const int32_t *frames, const float *matrices,
                       const float *offsets, int32_t timed);
int32_t bmfr_batch_synchronize(bmfr_batch batch);
/* hipGraph capture of the five batched launches, replayed each step. */
int32_t bmfr_batch_set_graph(bmfr_batch batch, int32_t enable);
/* BMFR_FUSE_34 or BMFR_FUSE_345 (default) for the batch's stages 3-5. */
int32_t bmfr_batch_set_fused(bmfr_batch batch, int32_t level);
/* 2 (default) to 4: the members are split into that many lanes of about
 * equal pixel count and an untimed (or graph-replayed) step runs them as a
 * software pipeline on one HIP stream per lane, lane l's stage 1 starting
 * when lane l - 1's is done, so that each lane's fitter overlaps other
 * lanes' streaming stages; 1: stage by stage over all members.  At most one
 * lane per member.  A step timed per stage always runs stage by stage.
 * Waits for the batch's last step when the lane count changes. */
int32_t bmfr_batch_set_lanes(bmfr_batch batch, int32_t lanes);
/* Fitter shape for the batch; grid_cap > 0 bounds the persistent grid. */
int32_t bmfr_batch_set_fitter_config(bmfr_batch batch, int32_t rows_per_lane,
                                     int32_t blocks_per_group, int32_t grid_cap);
/* Per-stage ms of the last timed step (as bmfr_stage_times); after a timed
 * graph replay only the total (ms6[5]) is measured and ms6[0..4] are -1. */
int32_t bmfr_batch_stage_times(bmfr_batch batch, float *ms6);
/* Autotuner: times every supported fitter shape (rows per lane, blocks per
 * work-group, persistent grid cap) on the data of the last step with HIP
 * events, keeps the fastest and writes a log line into `log`. */
int32_t bmfr_batch_autotune(bmfr_batch batch, int32_t reps, char *log, size_t log_len);
/* Markers 0..4 on the batch's HIP stream (shared with the stage events). */
int32_t bmfr_batch_record_marker(bmfr_batch batch, int32_t marker);
int32_t bmfr_batch_marker_elapsed(bmfr_batch batch, int32_t m0, int32_t m1, float *ms);

/* ---- server: streams admitted by HBM budget, continuous batching ------------
 * No counterpart in the reference, which runs one stream with five launches
 * per frame (opencl/bmfr.cpp:417-485); a server runs many render streams of
 * any resolution and storage precision.  It owns the streams it admits (their
 * device footprint counts against its HBM budget) and runs batched steps: one
 * frame of each of a set of its streams, one launch per stage per launch
 * group (storage x tmp storage x QR mode), ordered after each member's
 * pending work (its inputs) and before its later work.  A stream's results
 * are the same in any step as when it runs alone.  Two ways to drive it:
 *  - explicit steps: bmfr_server_step launches one step now;
 *  - the request loop: bmfr_server_submit queues a frame of a stream, and a
 *    dispatcher thread inside the library forms steps continuously: whenever
 *    fewer than `max_in_flight` steps (default 2) are on the GPU it launches
 *    one holding the oldest queued frame of every stream that has one (at
 *    most `max_streams_per_step`, oldest arrivals first); completions come
 *    back through bmfr_server_poll with their timestamps.
 * Step ids are a monotonically increasing count (the first step is 1). */
int32_t bmfr_server_create(int64_t hbm_budget_bytes, bmfr_server *out);
/* Waits for the steps on the GPU, drops frames still queued, closes every
 * stream the server owns and frees the server. */
int32_t bmfr_server_destroy(bmfr_server server);
int32_t bmfr_server_usage(bmfr_server server, int64_t *used_bytes, int64_t *budget_bytes,
                          int32_t *open_streams);
/* Opens a stream now, or refuses with BMFR_ERR_OUT_OF_BUDGET. */
int32_t bmfr_server_open_stream(bmfr_server server, const bmfr_params *params, bmfr_stream *out);
/* Admission FIFO: opens the stream now if nothing is waiting and it fits
 * (*out = the stream); otherwise queues the request (*out = NULL) and the
 * server opens it, in request order, once closes have freed enough of the
 * budget, reporting it as a BMFR_EVENT_STREAM_ADMITTED event with `tag`.  A
 * stream larger than the whole budget is refused (BMFR_ERR_OUT_OF_BUDGET). */
int32_t bmfr_server_request_stream(bmfr_server server, const bmfr_params *params, uint64_t tag,
                                   bmfr_stream *out);
/* Withdraws a queued admission request (*found = 1 if it was still queued). */
int32_t bmfr_server_cancel_request(bmfr_server server, uint64_t tag, int32_t *found);
/* Waits until the stream's queued frames have run, then closes it and frees
 * its budget (which may admit queued requests). */
int32_t bmfr_server_close_stream(bmfr_server server, bmfr_stream stream);
/* One step now over `count` distinct streams of the server (none of which may
 * have frames queued through bmfr_server_submit); frames[i], matrices[16*i..]
 * and offsets[2*i..] are stream i's bmfr_accumulate_noisy_data arguments.
 * Ordering: a step runs after each member's earlier steps and after the work
 * that calls on the member enqueued; any call on a server stream runs after
 * the steps the stream took part in.  Work enqueued directly on a stream's
 * HIP stream is ordered the same way once bmfr_stream_hip_stream has been
 * called for that stream. */
int32_t bmfr_server_step(bmfr_server server, int32_t count, const bmfr_stream *streams,
                         const int32_t *frames, const float *matrices, const float *offsets,
                         int64_t *step_id);
/* *done = 1 and the step's GPU time once it has completed, 0 while it runs.
 * BMFR_ERR_STATE for a step whose record has been reused (the server keeps
 * the last few steps' records). */
int32_t bmfr_server_step_query(bmfr_server server, int64_t step_id, int32_t *done, float *gpu_ms);
int32_t bmfr_server_synchronize(bmfr_server server);
/* BMFR_FUSE_34 or BMFR_FUSE_345 (default) for later steps' stages 3-5. */
int32_t bmfr_server_set_fused(bmfr_server server, int32_t level);
/* Request-loop policy: steps on the GPU at once (1..4, default 2; 0 pauses
 * launching, during which bmfr_server_close_stream and
 * bmfr_server_synchronize wait for it to resume) and streams per step
 * (1..64, default 64). */
int32_t bmfr_server_set_policy(bmfr_server server, int32_t max_in_flight, int32_t max_streams_per_step);

/* Seconds on the clock the request loop's timestamps use (a monotonic clock,
 * CLOCK_MONOTONIC on Linux). */
double bmfr_clock_seconds(void);

/* One frame of a stream for the request loop. */
typedef struct bmfr_frame_request {
    bmfr_stream stream;
    int32_t frame_number;                   /* as bmfr_accumulate_noisy_data's */
    float prev_frame_camera_matrix[16];
    float pixel_offset[2];
    /* Device inputs, as bmfr_bind_device_inputs (the frame's normals and
     * positions stay in use until the stream's next frame has run); all null:
     * the stream's own input buffers, which hold one frame, so such a request
     * is refused (BMFR_ERR_STATE) while another frame of the stream waits. */
    const void *albedo, *normals, *positions, *noisy;
    /* Output surface for this frame (null: the stream's bound surface, if any). */
    void *surface;
    int32_t surface_format;                 /* bmfr_surface_format */
    int64_t surface_row_pitch;              /* bytes */
    /* hipStream_t on which the inputs are being written: the step waits for
     * the work enqueued on it before this call (null: the inputs are complete;
     * the stream's own HIP stream is always waited for). */
    void *producer_stream;
    double t_arrival;                       /* seconds, bmfr_clock_seconds() */
    uint64_t tag;                           /* returned with the completion */
} bmfr_frame_request;
/* Queues a frame; frames of one stream run in submission order, one per step.
 * The caller keeps the inputs and the surface valid until the frame's
 * BMFR_EVENT_FRAME_DONE (normals and positions: until the stream's next
 * frame's). */
int32_t bmfr_server_submit(bmfr_server server, const bmfr_frame_request *request);

typedef enum bmfr_event_kind {
    BMFR_EVENT_FRAME_DONE = 1,       /* a submitted frame's step has completed */
    BMFR_EVENT_FRAME_FAILED = 2,     /* its step could not be launched (status) */
    BMFR_EVENT_STREAM_ADMITTED = 3   /* a queued bmfr_server_request_stream was opened */
} bmfr_event_kind;
typedef struct bmfr_server_event {
    int32_t kind;                    /* bmfr_event_kind */
    int32_t status;                  /* bmfr_status (FRAME_FAILED) */
    bmfr_stream stream;
    uint64_t tag;                    /* the request's tag */
    int32_t frame_number;
    int32_t step_streams;            /* streams in the frame's step */
    int64_t step_id;
    double t_arrival;                /* the request's t_arrival (admission: request time) */
    double t_launch;                 /* when the step was enqueued */
    double t_done;                   /* when the dispatcher saw it complete (admission time) */
    float step_gpu_ms;               /* the step's GPU time, first launch to last end */
} bmfr_server_event;
/* Up to max_events events, oldest first; waits up to timeout_us for the
 * first one (0: does not wait). */
int32_t bmfr_server_poll(bmfr_server server, bmfr_server_event *events, int32_t max_events,
                         int32_t timeout_us, int32_t *count);

/* ---- procedural scene (replaces the EXR dataset, opencl/bmfr.cpp:145-172) ---- */
typedef struct bmfr_scene_s *bmfr_scene;
/* Plane y=0 (|x|,|z| <= 64) with `spheres` spheres, seeded; a pinhole camera
 * (60 deg vertical FOV) at (0,2,-10) looking at the origin that moves 0.05 in
 * x and yaws 0.3 deg per frame.  `static_camera` != 0 freezes camera and
 * pixel offset (for reprojection invariants). */
int32_t bmfr_scene_create(uint64_t seed, int32_t spheres, float firefly_probability,
                          int32_t width, int32_t height, int32_t static_camera,
                          bmfr_scene *out);
int32_t bmfr_scene_destroy(bmfr_scene scene);
/* Row-vector view-projection of `frame` (16 floats) and its pixel offset. */
int32_t bmfr_scene_camera(bmfr_scene scene, int32_t frame, float *view_projection,
                          float *pixel_offset);
/* CPU render of one frame into host arrays of W*H*3 floats (any may be null). */
int32_t bmfr_scene_render_host(bmfr_scene scene, int32_t frame, float *albedo,
                               float *normals, float *positions, float *noisy);
/* GPU render of one frame straight into the stream's current input buffers
 * (albedo, normals, positions, noisy colour), enqueued on the stream. */
int32_t bmfr_scene_render_to_stream(bmfr_scene scene, int32_t frame, bmfr_stream stream);
/* GPU render of one frame into caller-owned device buffers (a renderer's own
 * G-buffers, for bmfr_bind_device_inputs or bmfr_server_submit): albedo as
 * float[3P], the other three as float (storage BMFR_STORAGE_F32) or half
 * [3P]; enqueued on `hip_stream` (a hipStream_t; null: the null stream). */
int32_t bmfr_scene_render_device(bmfr_scene scene, int32_t frame, int32_t storage, void *albedo,
                                 void *normals, void *positions, void *noisy, void *hip_stream);

/* ---- dataset-format IO (the reference's OpenImageIO and camera header use) ----
 * Host-only helpers so that the published scenes run when a user supplies
 * them: the reference reads color/shading_normal/world_position/albedo NN.exr
 * as 3-channel float (opencl/bmfr.cpp:145-172), compiles the dataset's
 * camera_matrices.h in (opencl/bmfr.cpp:43-47, 226-227, 440-444) and writes
 * outputNN.png cropped to the image size (opencl/bmfr.cpp:519-547). */
/* Size and channel count of a single-part scanline OpenEXR file. */
int32_t bmfr_exr_info(const char *path, int32_t *width, int32_t *height, int32_t *channels);
/* Channels R, G, B (or a single Y) of a scanline OpenEXR file with NONE,
 * RLE, ZIPS or ZIP compression and HALF, FLOAT or UINT pixels, as
 * interleaved float RGB with row stride `width` (as read_image_file's
 * conversion to float, opencl/bmfr.cpp:145-164).  Other compressions and
 * tiled or multi-part files are refused with BMFR_ERR_INVALID_ARGUMENT. */
int32_t bmfr_exr_read_rgb(const char *path, int32_t width, int32_t height, float *rgb);
/* Writes interleaved float RGB (row stride `row_stride` pixels) as a FLOAT
 * B, G, R OpenEXR file; compression 0 (NONE) or 3 (ZIP). */
int32_t bmfr_exr_write_rgb(const char *path, int32_t width, int32_t height, const float *rgb,
                           int64_t row_stride, int32_t compression);
/* Writes interleaved float RGB (row stride `row_stride` pixels) as an 8-bit
 * RGB PNG, round(clamp(v, 0, 1) * 255) per channel (opencl/bmfr.cpp:519-547;
 * the reference passes a WORKSET_WIDTH stride for IMAGE_WIDTH rows, which
 * this does not reproduce). */
int32_t bmfr_png_write_rgb(const char *path, int32_t width, int32_t height, const float *rgb,
                           int64_t row_stride);
/* Parses the dataset's camera_matrices.h: camera_matrices[frames][4][4]
 * (row-vector view-projection, 16 floats per frame), pixel_offsets[frames][2]
 * and the scalars position_limit_squared and normal_limit_squared (C
 * declarations or #defines).  Outputs may be null. */
int32_t bmfr_camera_matrices_load(const char *path, int32_t frames, float *matrices, float *offsets,
                                  float *position_limit_squared, float *normal_limit_squared);

#ifdef __cplusplus
}
#endif
#endif /* BMFR_HIP_H */
