// Stage 2: fitter, opencl/bmfr.cl:487-700, for gfx950.
//
// Method kept from the reference: per-block min/max scaling of the six
// world-position features, feature noise on the first pass, Householder QR
// of the 1024x13 block matrix with the reference's sign choice
// (u_c = x_c - ||x||), and the reference's back-substitution sequence.
//
// Layout: the whole 1024x13 block matrix lives in VGPRs.  A block is
// handled by W = 16/RPL waves; each lane owns RPL rows (13*RPL registers),
// so the QR never touches memory after the initial load.  Column sums are
// 64-lane DPP reductions (quad_perm, row_half_mirror, row_mirror,
// row_bcast15, row_bcast31, then v_readlane of lane 63); with W > 1 the W
// wave partials meet in a few LDS words behind one barrier.  All trailing
// dot products of a column step are reduced together, so a block needs two
// reduction rounds per pivot column.
//
// Columns 10-12 (the RGB right-hand sides) are never pivots: the reference's
// steps col = 10..12 only reflect rows >= 10 of the colour columns, which
// the back substitution never reads (it uses rows 0-9 of column 10 of R),
// so skipping them leaves mins_maxs and weights unchanged.  The factorised
// matrix is not written back to tmp_data: nothing reads it after the fitter
// (opencl/bmfr.cpp:449-476).
#include "fitter_body.h"

namespace bmfr {

// ---- FP64 re-factorisation (BMFR_QR_KAHAN_FP64) ------------------------------
// One 256-thread work-group per flagged block (unflagged blocks exit at once).
// The block matrix is rebuilt exactly as the FP32 fitter saw it at column 0
// (the same scaled, noised float values), widened to double and kept in
// registers (4 rows x 13 columns per thread); Householder QR with the stable
// sign choice (v = x - alpha e_c, alpha = -sign(x_c) ||x||), then back
// substitution, all in FP64.
template <int NV>
__device__ __forceinline__ void block_sum_f64(double (&v)[NV], double (*red)[4]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[i] += __shfl_xor(v[i], o);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) red[i][wave] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = (red[i][0] + red[i][1]) + (red[i][2] + red[i][3]);
}

template <typename TT, bool HALF_ROUND>
__device__ __forceinline__ void refactor_block(const TT *__restrict__ src, const float *__restrict__ mm,
                                               NoiseTable noise, float *__restrict__ w_out) {
    __shared__ double red[16][4];
    __shared__ double rl[kFeatures][kBufferCount];
    const int t = threadIdx.x;
    double a[kBufferCount][4];
#pragma unroll
    for (int f = 0; f < kBufferCount; ++f) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int row = k * 256 + t;
            float v;
            if constexpr (sizeof(TT) == 2) v = __half2float(src[f * kBlockPixels + row]);
            else v = src[f * kBlockPixels + row];
            if (f >= kNotScaled && f < kFeatures) {
                const float mn = mm[(f - kNotScaled) * 2], mx = mm[(f - kNotScaled) * 2 + 1];
                const float d = mx - mn;
                float sc = v - mn;
                if (fabsf(d) > 1.0f) {
                    const float y = 1.0f / d;
                    const float q = sc * y;
                    const float r = __builtin_fmaf(-q, d, sc);
                    sc = __builtin_fmaf(r, y, q);
                }
                v = HALF_ROUND ? round_half_of_float(sc) : sc;
            }
            if (f >= 1 && f < kFeatures)
                v = (float)((double)v + noise.term((double)noise.r[(f - 1) * kBlockPixels + row]));
            a[f][k] = (double)v;
        }
    }
    double alpha[kFeatures];
#pragma unroll
    for (int c = 0; c < kFeatures; ++c) {
        double r2[2];
        r2[0] = 0.0;
        r2[1] = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int row = k * 256 + t;
            if (row >= c) r2[0] += a[c][k] * a[c][k];
            if (row == c) r2[1] = a[c][k];  // exact: all other threads add zeros
        }
        block_sum_f64<2>(r2, red);
        const double norm = sqrt(r2[0]), x0 = r2[1];
        const double al = x0 >= 0.0 ? -norm : norm;
        const double vc = x0 - al;
        const double vv = 2.0 * norm * (norm + fabs(x0));
        alpha[c] = al;
        double v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int row = k * 256 + t;
            v[k] = row > c ? a[c][k] : (row == c ? vc : 0.0);
        }
        double d[kBufferCount - 1];
#pragma unroll
        for (int j = 0; j < kBufferCount - 1; ++j) d[j] = 0.0;
#pragma unroll
        for (int j = c + 1; j < kBufferCount; ++j)
#pragma unroll
            for (int k = 0; k < 4; ++k) d[j - c - 1] += v[k] * a[j][k];
        block_sum_f64<kBufferCount - 1>(d, red);
        if (vv > 0.0) {
#pragma unroll
            for (int j = c + 1; j < kBufferCount; ++j) {
                const double f = 2.0 * d[j - c - 1] / vv;
#pragma unroll
                for (int k = 0; k < 4; ++k) a[j][k] -= f * v[k];
            }
        }
    }
    // R (rows 0..9) and Q^T b: thread i < 10 holds row i of every column
    if (t < kFeatures) {
#pragma unroll
        for (int j = 0; j < kBufferCount; ++j) rl[t][j] = a[j][0];
    }
    __syncthreads();
    if (t < 3) {
        double w[kFeatures];
#pragma unroll
        for (int i = kFeatures - 1; i >= 0; --i) {
            double acc = rl[i][kFeatures + t];
#pragma unroll
            for (int j = i + 1; j < kFeatures; ++j) acc -= rl[i][j] * w[j];
            w[i] = acc / alpha[i];
        }
#pragma unroll
        for (int i = 0; i < kFeatures; ++i) w_out[i * 3 + t] = (float)w[i];
    }
}

struct RefactorParams {
    const void *tmp;
    const float *mins_maxs;
    NoiseTable noise;
    const int32_t *flags;   // null: every block
    float *out;             // weights (in place) or a separate FP64-solution buffer
    int32_t blocks;
};

template <typename TT, bool HALF_ROUND>
__global__ __launch_bounds__(256) void k_refactor(const RefactorParams p) {
    const int blk = blockIdx.x;
    if (blk >= p.blocks || (p.flags && !p.flags[blk])) return;
    refactor_block<TT, HALF_ROUND>(reinterpret_cast<const TT *>(p.tmp) + (int64_t)blk * (kBufferCount * kBlockPixels),
                                   p.mins_maxs + (int64_t)blk * (kScaled * 2), p.noise,
                                   p.out + (int64_t)blk * (kFeatures * 3));
}

template <typename TT, bool HALF_ROUND>
__global__ __launch_bounds__(256) void kb_refactor(const BatchLaunch L) {
    const int32_t *starts = L.starts + kStartsBlocks * (kMaxBatch + 1);
    const int u = blockIdx.x;
    if (u >= starts[L.n]) return;
    const int s = batch_stream_of(starts, L.n, u);
    const int blk = u - starts[s];
    const BatchStream &bs = L.streams[s];
    if (!bs.flags || !((const int32_t *)bs.flags)[blk]) return;
    refactor_block<TT, HALF_ROUND>(reinterpret_cast<const TT *>(bs.tmp_data) + (int64_t)blk * (kBufferCount * kBlockPixels),
                                   (const float *)bs.mins_maxs + (int64_t)blk * (kScaled * 2),
                                   NoiseTable{(const float *)bs.noise, 2.0 * bs.args.noise_amount},
                                   (float *)bs.weights + (int64_t)blk * (kFeatures * 3));
}

static hipError_t launch_refactor_all(const StageArgs &a, const StageBuffers &b, const int32_t *flags,
                                      float *out, hipStream_t s) {
    RefactorParams rp;
    rp.tmp = b.tmp_data;
    rp.mins_maxs = (const float *)b.mins_maxs;
    rp.noise = NoiseTable{(const float *)b.noise, 2.0 * a.noise_amount};
    rp.flags = flags;
    rp.out = out;
    rp.blocks = a.blocks;
    if (a.tmp_f16) hipLaunchKernelGGL((k_refactor<__half, true>), dim3(a.blocks), dim3(256), 0, s, rp);
    else hipLaunchKernelGGL((k_refactor<float, false>), dim3(a.blocks), dim3(256), 0, s, rp);
    return hipGetLastError();
}

hipError_t fitter_single_plain(const FitterParams &fp, bool tmp_f16, const FitterConfig &cfg, hipStream_t s);
hipError_t fitter_single_comp(const FitterParams &fp, bool tmp_f16, const FitterConfig &cfg, hipStream_t s);
hipError_t fitter_batch_plain(const BatchLaunch &L, int total, bool tmp_f16, const FitterConfig &cfg, int grid_cap,
                              hipStream_t s);
hipError_t fitter_batch_comp(const BatchLaunch &L, int total, bool tmp_f16, const FitterConfig &cfg, int grid_cap,
                             hipStream_t s);

hipError_t fitter_dispatch_single(const FitterParams &fp, bool tmp_f16, bool comp, const FitterConfig &cfg,
                                  hipStream_t s) {
    return comp ? fitter_single_comp(fp, tmp_f16, cfg, s) : fitter_single_plain(fp, tmp_f16, cfg, s);
}

hipError_t fitter_dispatch_batch(const BatchLaunch &L, int total, bool tmp_f16, bool comp,
                                 const FitterConfig &cfg, int grid_cap, hipStream_t s) {
    return comp ? fitter_batch_comp(L, total, tmp_f16, cfg, grid_cap, s)
                : fitter_batch_plain(L, total, tmp_f16, cfg, grid_cap, s);
}

hipError_t launch_fitter(const StageArgs &a, const StageBuffers &b, const FitterConfig &cfg,
                         int32_t qr_mode, hipStream_t s) {
    FitterParams fp;
    fp.tmp = b.tmp_data;
    fp.mins_maxs = (float *)b.mins_maxs;
    fp.weights = (float *)b.weights;
    fp.noise = (const float *)b.noise;
    fp.noise_scale = 2.0 * a.noise_amount;
    fp.flags = qr_mode == 2 ? (int32_t *)b.flags : nullptr;
    fp.threshold = a.refactor_threshold;
    fp.blocks = a.blocks;
    fp.frame = a.frame;
    fp.not_scaled = a.features.not_scaled;
    const bool h = a.tmp_f16 != 0;
    if (a.features.count) {   // a configured feature set: plain FP32 QR only
        if (qr_mode != 0) return hipErrorInvalidValue;
        return a.features.count <= 8 ? fitter_set_lo(fp, h, a.features.count, s)
                                     : fitter_set_hi(fp, h, a.features.count, s);
    }
    hipError_t e = fitter_dispatch_single(fp, h, qr_mode >= 1, cfg, s);
    if (e != hipSuccess || qr_mode != 2) return e;
    return launch_refactor_all(a, b, (const int32_t *)b.flags, (float *)b.weights, s);
}

hipError_t launch_fitter_fp64_all(const StageArgs &a, const StageBuffers &b, float *out, hipStream_t s) {
    if (a.features.count) return hipErrorInvalidValue;   // the reference's set only
    return launch_refactor_all(a, b, nullptr, out, s);
}

hipError_t launch_batch_fitter(const BatchLaunch &L, int total, bool t16, const FitterConfig &cfg,
                               int grid_cap, int32_t qr_mode, hipStream_t s) {
    if (total <= 0) return hipErrorInvalidValue;
    hipError_t e = fitter_dispatch_batch(L, total, t16, qr_mode >= 1, cfg, grid_cap, s);
    if (e != hipSuccess || qr_mode != 2) return e;
    if (t16) hipLaunchKernelGGL((kb_refactor<__half, true>), dim3(total), dim3(256), 0, s, L);
    else hipLaunchKernelGGL((kb_refactor<float, false>), dim3(total), dim3(256), 0, s, L);
    return hipGetLastError();
}

}  // namespace bmfr
