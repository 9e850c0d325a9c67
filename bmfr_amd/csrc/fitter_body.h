// Stage 2 (fitter) device code shared by the fitter translation units.
#pragma once
//
// Method kept from the reference: per-block min/max scaling of the six
// world-position features, feature noise on the first pass, Householder QR
// of the 1024x13 block matrix with the reference's sign choice
// (u_c = x_c - ||x||), and the reference's back-substitution sequence.
//
// Layout: the whole 1024x13 block matrix lives in VGPRs.  A block is
// handled by W = 16/RPL waves; each lane owns RPL rows (13*RPL registers),
// so the QR never touches memory after the initial load.  With binary16
// tmp_data and one wave per block (RPL 16, the default) the columns are
// binary16 pairs from the first pivot's update on (PackedRows: 8 registers
// a column, v_dot2_f32_f16 dot products, half updates).  Column sums are
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

#include "device_common.h"
#include "kernels.h"

namespace bmfr {

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp(float old, float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL,
                                                      ROW_MASK, 0xf, false));
}
// Full-mask DPP move: every lane is written, so the compiler can fold it
// into the consuming VALU op (v_add_f32_dpp / v_min_f32_dpp).
template <int CTRL>
__device__ __forceinline__ float dpp_all(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, true));
}

__device__ __forceinline__ float readlane(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

struct OpAdd {
    static constexpr float identity = 0.f;
    __device__ __forceinline__ static float f(float a, float b) { return a + b; }
};
struct OpMin {
    static constexpr float identity = __builtin_huge_valf();
    __device__ __forceinline__ static float f(float a, float b) { return fminf(a, b); }
};

// 64-lane reduction; the result is wave-uniform.
template <typename Op>
__device__ __forceinline__ float wave_reduce(float v) {
    v = Op::f(v, dpp_all<0xB1>(v));    // quad_perm [1,0,3,2]
    v = Op::f(v, dpp_all<0x4E>(v));    // quad_perm [2,3,0,1]
    v = Op::f(v, dpp_all<0x141>(v));   // row_half_mirror
    v = Op::f(v, dpp_all<0x140>(v));   // row_mirror
    v = Op::f(v, dpp<0x142, 0xA>(Op::identity, v));  // row_bcast15 -> rows 1, 3
    v = Op::f(v, dpp<0x143, 0xC>(Op::identity, v));  // row_bcast31 -> rows 2, 3
    return readlane(v, 63);
}

// gfx950 cross-row swaps.  permlane32_swap(a, b): a <- [a lanes 0-31,
// b lanes 0-31], b <- [a lanes 32-63, b lanes 32-63]; permlane16_swap does
// the same with 16-lane rows: a <- [a.r0, b.r0, a.r2, b.r2],
// b <- [a.r1, b.r1, a.r3, b.r3].
__device__ __forceinline__ void swap32(float &a, float &b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void swap16(float &a, float &b) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}

// N independent 64-lane reductions; results are wave-uniform.  Values go
// in groups of four through one register: a permlane32 swap and an Op fold
// each pair (lanes l and l+32; the pair's first value stays in lanes 0-31,
// the second in 32-63), a permlane16 swap and an Op fold the two pairs
// (rows r and r+1), leaving value 4g in row 0, 4g+2 in row 1, 4g+1 in row 2
// and 4g+3 in row 3; four in-row DPP steps then finish all four at once.
// Each step runs over every group before the next (independent chains fill
// each other's hazard slots).  About 3.5 instructions per value, against 8
// for six DPP steps with the two row broadcasts.
template <typename Op, int N>
__device__ __forceinline__ void wave_reduce_n(float (&v)[N]) {
    constexpr int G = (N + 3) / 4;
    float w[4 * G];
#pragma unroll
    for (int i = 0; i < 4 * G; ++i) w[i] = i < N ? v[i] : Op::identity;
    float p[2 * G];
#pragma unroll
    for (int i = 0; i < 2 * G; ++i) {
        float a = w[2 * i], b = w[2 * i + 1];
        swap32(a, b);
        p[i] = Op::f(a, b);
    }
    float s[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        float a = p[2 * g], b = p[2 * g + 1];
        swap16(a, b);
        s[g] = Op::f(a, b);
    }
#pragma unroll
    for (int g = 0; g < G; ++g) s[g] = Op::f(s[g], dpp_all<0x140>(s[g]));   // row_mirror
#pragma unroll
    for (int g = 0; g < G; ++g) s[g] = Op::f(s[g], dpp_all<0x141>(s[g]));   // row_half_mirror
#pragma unroll
    for (int g = 0; g < G; ++g) s[g] = Op::f(s[g], dpp_all<0x4E>(s[g]));    // quad_perm [2,3,0,1]
#pragma unroll
    for (int g = 0; g < G; ++g) s[g] = Op::f(s[g], dpp_all<0xB1>(s[g]));    // quad_perm [1,0,3,2]
    constexpr int kLane[4] = {0, 32, 16, 48};
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = readlane(s[i / 4], kLane[i % 4]);
}

// Reduce NV values over the W waves of one block.  `red` points at this
// block's [W][16] slot of the round's LDS buffer.
template <int W, int NV, typename Op, int RS = 16>
__device__ __forceinline__ void block_reduce(float (&v)[NV], float *red, int wave, int lane) {
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wave_reduce<Op>(v[i]);
    if constexpr (W > 1) {
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < NV; ++i) red[wave * RS + i] = v[i];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            float acc = red[i];
#pragma unroll
            for (int w = 1; w < W; ++w) acc = Op::f(acc, red[w * RS + i]);
            v[i] = acc;
        }
    }
}

// Error-free transformations for the compensated (BMFR_QR_KAHAN) sums.
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    const float z = s - a;
    e = (a - (s - z)) + (b - z);
}
// s + c += x * y with the product's and the sum's rounding errors kept in c.
__device__ __forceinline__ void dot2_acc(float &s, float &c, float x, float y) {
    const float p = x * y;
    const float pe = __builtin_fmaf(x, y, -p);
    float t, e;
    two_sum(s, p, t, e);
    s = t;
    c = c + (e + pe);
}
// 64-lane reduction of a compensated pair; the result (s + c) is uniform.
__device__ __forceinline__ float wave_reduce_pair(float s, float c) {
    float t, e;
#define BMFR_PAIR_STEP(S2, C2)  \
    {                           \
        const float s2 = (S2);  \
        const float c2 = (C2);  \
        two_sum(s, s2, t, e);   \
        s = t;                  \
        c = c + (c2 + e);       \
    }
    BMFR_PAIR_STEP(dpp_all<0xB1>(s), dpp_all<0xB1>(c));
    BMFR_PAIR_STEP(dpp_all<0x4E>(s), dpp_all<0x4E>(c));
    BMFR_PAIR_STEP(dpp_all<0x141>(s), dpp_all<0x141>(c));
    BMFR_PAIR_STEP(dpp_all<0x140>(s), dpp_all<0x140>(c));
    BMFR_PAIR_STEP((dpp<0x142, 0xA>(0.f, s)), (dpp<0x142, 0xA>(0.f, c)));
    BMFR_PAIR_STEP((dpp<0x143, 0xC>(0.f, s)), (dpp<0x143, 0xC>(0.f, c)));
#undef BMFR_PAIR_STEP
    return readlane(s, 63) + readlane(c, 63);
}

template <typename TT, int V>
__device__ __forceinline__ void load_rows(const TT *p, float *out) {
    if constexpr (sizeof(TT) == 4) {
        static_assert(V == 4, "fp32 rows load as float4");
        const float4 x = *reinterpret_cast<const float4 *>(p);
        out[0] = x.x; out[1] = x.y; out[2] = x.z; out[3] = x.w;
    } else if constexpr (V == 4) {
        const uint2 x = *reinterpret_cast<const uint2 *>(p);
        const __half2 h0 = *reinterpret_cast<const __half2 *>(&x.x);
        const __half2 h1 = *reinterpret_cast<const __half2 *>(&x.y);
        out[0] = __low2float(h0); out[1] = __high2float(h0);
        out[2] = __low2float(h1); out[3] = __high2float(h1);
    } else {
        static_assert(V == 8, "fp16 rows load as 4 or 8 halves");
        const uint4 x = *reinterpret_cast<const uint4 *>(p);
        const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const __half2 h = *reinterpret_cast<const __half2 *>(&w[i]);
            out[2 * i] = __low2float(h);
            out[2 * i + 1] = __high2float(h);
        }
    }
}

struct FitterParams {
    const void *tmp;
    float *mins_maxs;
    float *weights;
    const float *noise;    // the frame's noise table (random() - 0.5f per term)
    double noise_scale;    // 2 * NOISE_AMOUNT
    int32_t *flags;        // per block: 1 = re-factor in FP64 (null: not wanted)
    float threshold;       // flag when min |R_ii| < threshold * max |R_ii|
    int32_t blocks;
    int32_t frame;
    int32_t not_scaled;    // configured feature sets: features [not_scaled, N) are scaled
};

// Per-block outputs of the FP32 fitter besides weights and mins_maxs.
struct BlockFlag {
    int32_t *flag;
    float threshold;
};

// Correctly rounded a / d from y = RN(1/d): q = RN(a*y), r = RN(a - q*d)
// (exact by FMA), RN(q + r*y) (Markstein's corrected quotient).
__device__ __forceinline__ float div_by(float a, float d, float y) {
    const float q = a * y;
    const float r = __builtin_fmaf(-q, d, a);
    return __builtin_fmaf(r, y, q);
}

// Back substitution, bmfr.cl:658-699, by one wave, for F features
// (R_EDGE = F + 1).  Lane L = 3*c + ch holds column c (0..F) of R for
// colour channel ch: R[r][c] = elem(c, r) (the factorised column c at row r,
// r < c, wave-uniform), R[c][c] = norms[c]; column F is Q^T b = rows 0..F-1
// of colour column F+ch.  Weights are written by lane 3F + ch.
template <int F = kFeatures, typename Elem>
__device__ __forceinline__ void back_substitution(Elem elem, const float (&norms)[F], float *__restrict__ weights,
                                                  bool active, int lane) {
    static_assert(3 * (F + 1) <= 64, "one lane per column and channel");
    constexpr int RE = F + 1;   // R_EDGE
    const int cl = lane / 3;
    const int ch = lane - 3 * cl;
    float rc[F];
#pragma unroll
    for (int r = 0; r < F; ++r) rc[r] = 0.f;
#pragma unroll
    for (int c = 0; c < F; ++c) {
#pragma unroll
        for (int r = 0; r < F; ++r) {
            float val;
            if (r < c) val = elem(c, r);
            else if (r == c) val = norms[c];
            else val = 0.f;
            if (cl == c) rc[r] = val;
        }
    }
#pragma unroll
    for (int r = 0; r < F; ++r) {
        const float v0 = elem(F, r);
        const float v1 = elem(F + 1, r);
        const float v2 = elem(F + 2, r);
        if (cl == F) rc[r] = ch == 0 ? v0 : (ch == 1 ? v1 : v2);
    }
#pragma unroll
    for (int i = RE - 2; i >= 0; --i) {
        const float divider = __shfl(rc[i], 3 * i + ch);
        if (cl >= i) rc[i] = rc[i] / divider;
        float acc = __shfl(rc[i], 3 * F + ch);
#pragma unroll
        for (int j = i + 1; j < RE - 1; ++j) acc = acc - __shfl(rc[i], 3 * j + ch);
        if (cl == F) rc[i] = acc;
        const float wi = __shfl(rc[i], 3 * F + ch);
        if (cl == i) {
#pragma unroll
            for (int r = 0; r <= i; ++r) rc[r] = rc[r] * wi;
        }
    }
    if (active && cl == F && ch < 3) {
#pragma unroll
        for (int r = 0; r < F; ++r) weights[r * 3 + ch] = rc[r];
    }
}

// One 32x32 block: RPL rows per lane, BPG blocks per work-group, TT tmp
// element type, HALF_ROUND: round every stored element to binary16 like the
// reference's USE_HALF_PRECISION_IN_TMP_DATA=1 (bmfr.cl:255-258, 540, 651).
// `red` is this work-group's LDS exchange area; `active` false recomputes a
// valid block (keeping barriers uniform) without storing.
// COMP: compensated norm and dot sums (BMFR_QR_KAHAN).  `flag` (optional)
// receives 1 when min |R_ii| < threshold * max |R_ii|.
constexpr int red_slots(bool comp) { return comp ? 32 : 16; }

template <int RPL, int BPG, typename TT, bool HALF_ROUND, bool COMP>
__device__ __forceinline__ void fitter_block(const TT *__restrict__ src, float *__restrict__ mins_maxs,
                                             float *__restrict__ weights, NoiseTable noise,
                                             BlockFlag flag, bool active,
                                             float (&red)[2][BPG][16 / RPL][red_slots(COMP)]) {
    constexpr int RS = red_slots(COMP);
    constexpr int W = 16 / RPL;            // waves per block
    constexpr int T = 64 * W;              // lanes per block
    constexpr int V = (RPL < (int)(16 / sizeof(TT))) ? RPL : (int)(16 / sizeof(TT));
    constexpr int Q = RPL / V;             // vector loads per feature

    const int tid = threadIdx.x;
    const int sub = tid / T;
    const int t = tid - sub * T;
    const int lane = tid & 63;
    const int wave = t >> 6;

    // Row held in element k of this lane: rows t*V..t*V+V-1 of each T*V strip.
    auto row_of = [&](int k) { return (k / V) * (T * V) + t * V + (k % V); };

    float a[kBufferCount][RPL];
#pragma unroll
    for (int f = 0; f < kBufferCount; ++f)
#pragma unroll
        for (int q = 0; q < Q; ++q) load_rows<TT, V>(src + f * kBlockPixels + q * (T * V) + t * V, &a[f][q * V]);

    int round = 0;
    // ---- scaling, bmfr.cl:511-542 (max taken as -min(-x): exact) ----
    float bmin[kScaled], bmax[kScaled];
    {
        float v[2 * kScaled];
#pragma unroll
        for (int i = 0; i < kScaled; ++i) {
            float mn = INFINITY, nmx = INFINITY;
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
                mn = fminf(mn, a[kNotScaled + i][k]);
                nmx = fminf(nmx, -a[kNotScaled + i][k]);
            }
            v[i] = mn;
            v[kScaled + i] = nmx;
        }
        block_reduce<W, 2 * kScaled, OpMin, RS>(v, &red[round & 1][sub][0][0], wave, lane);
        ++round;
#pragma unroll
        for (int i = 0; i < kScaled; ++i) {
            bmin[i] = v[i];
            bmax[i] = -v[kScaled + i];
        }
    }
    if (active && t == 0) {
        float *mm = mins_maxs;
#pragma unroll
        for (int i = 0; i < kScaled; ++i) {
            mm[2 * i + 0] = bmin[i];
            mm[2 * i + 1] = bmax[i];
        }
    }
    // scale (bmfr.cl:200-205).  (v-min)/(max-min) is evaluated as Markstein's
    // correctly rounded quotient: y = RN(1/d) once per feature, then
    // q = RN(a*y), r = RN(a - q*d) (exact), q' = RN(q + r*y) = RN(a/d).
#pragma unroll
    for (int i = 0; i < kScaled; ++i) {
        const float d = bmax[i] - bmin[i];
        const bool divide = fabsf(d) > 1.0f;
        const float y = 1.0f / d;
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
            const float num = a[kNotScaled + i][k] - bmin[i];
            float sc = num;
            if (divide) {
                const float q = num * y;
                const float r = __builtin_fmaf(-q, d, num);
                sc = __builtin_fmaf(r, y, q);
            }
            a[kNotScaled + i][k] = HALF_ROUND ? round_half_of_float(sc) : sc;
        }
    }
    // ---- feature noise of the first pass, bmfr.cl:623-627: value + noise in
    // FP64, rounded once to float (NOISE_AMOUNT is a double literal) ----
#pragma unroll
    for (int f = 1; f < kFeatures; ++f) {
        const float *nz = noise.r + (f - 1) * kBlockPixels;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            float n[V];
#pragma unroll
            for (int e = 0; e < V; e += 4) {
                const float4 x = *reinterpret_cast<const float4 *>(nz + q * (T * V) + t * V + e);
                n[e] = x.x;
                n[e + 1] = x.y;
                n[e + 2] = x.z;
                n[e + 3] = x.w;
            }
#pragma unroll
            for (int e = 0; e < V; ++e)
                a[f][q * V + e] = (float)((double)a[f][q * V + e] + noise.term((double)n[e]));
        }
    }

    // ---- Householder QR over the 10 feature columns, bmfr.cl:549-656 ----
    float norms[kFeatures];
#pragma unroll
    for (int col = 0; col < kFeatures; ++col) {
        // Pivot row `col` sits in lane col/V (wave 0), element col%V.
        float r1[1];
        float xc = readlane(a[col][col % V], col / V);
        if constexpr (!COMP) {
            float s = 0.f;
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
                const float x = row_of(k) > col ? a[col][k] : 0.f;
                s = __builtin_fmaf(x, x, s);
            }
            r1[0] = wave_reduce<OpAdd>(s);
        } else {
            float s = 0.f, c = 0.f;
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
                const float x = row_of(k) > col ? a[col][k] : 0.f;
                dot2_acc(s, c, x, x);
            }
            r1[0] = wave_reduce_pair(s, c);
        }
        if constexpr (W > 1) {
            float *rd = &red[round & 1][sub][0][0];
            if (lane == 0) {
                rd[wave * RS + 0] = r1[0];
                if (wave == 0) rd[1] = xc;
            }
            __syncthreads();
            float acc = rd[0], cc = 0.f;
#pragma unroll
            for (int w = 1; w < W; ++w) {
                if constexpr (COMP) {
                    float t, e;
                    two_sum(acc, rd[w * RS], t, e);
                    acc = t;
                    cc = cc + e;
                } else {
                    acc = acc + rd[w * RS];
                }
            }
            r1[0] = acc + cc;
            xc = rd[1];
        }
        ++round;
        const float sigma = r1[0];
        const float norm = sqrtf(sigma + xc * xc);
        const float uc = xc - norm;
        const float uu = sigma + uc * uc;
        norms[col] = norm;

        float u[RPL];
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
            const int row = row_of(k);
            u[k] = row > col ? a[col][k] : (row == col ? uc : 0.f);
        }
        constexpr int kMaxTrail = kBufferCount - 1;
        float d[kMaxTrail];
#pragma unroll
        for (int j = 0; j < kMaxTrail; ++j) d[j] = 0.f;
        {
            const int nv = kBufferCount - 1 - col;
            if constexpr (!COMP) {
#pragma unroll
                for (int j = col + 1; j < kBufferCount; ++j) {
                    float s = 0.f;
#pragma unroll
                    for (int k = 0; k < RPL; ++k) s = __builtin_fmaf(u[k], a[j][k], s);
                    d[j - col - 1] = wave_reduce<OpAdd>(s);
                }
            } else {
#pragma unroll
                for (int j = col + 1; j < kBufferCount; ++j) {
                    float s = 0.f, c = 0.f;
#pragma unroll
                    for (int k = 0; k < RPL; ++k) dot2_acc(s, c, u[k], a[j][k]);
                    d[j - col - 1] = wave_reduce_pair(s, c);
                }
            }
            if constexpr (W > 1) {
                float *rd = &red[round & 1][sub][0][0];
                if (lane == 0) {
#pragma unroll
                    for (int i = 0; i < kMaxTrail; ++i)
                        if (i < nv) rd[wave * RS + i] = d[i];
                }
                __syncthreads();
#pragma unroll
                for (int i = 0; i < kMaxTrail; ++i) {
                    if (i < nv) {
                        float acc = rd[i], cc = 0.f;
#pragma unroll
                        for (int w = 1; w < W; ++w) {
                            if constexpr (COMP) {
                                float t, e;
                                two_sum(acc, rd[w * RS + i], t, e);
                                acc = t;
                                cc = cc + e;
                            } else {
                                acc = acc + rd[w * RS + i];
                            }
                        }
                        d[i] = acc + cc;
                    }
                }
            }
            ++round;
        }
        // a_j -= u * (2 u.a_j / uu): one reciprocal per pivot and a fused
        // multiply-add per element (the reference divides per element,
        // bmfr.cl:650; the difference is within a couple of ulp).
        const float inv_uu = 1.f / uu;
#pragma unroll
        for (int j = col + 1; j < kBufferCount; ++j) {
            const float fct = (2.f * d[j - col - 1]) * inv_uu;
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
                const float v = __builtin_fmaf(-u[k], fct, a[j][k]);
                a[j][k] = HALF_ROUND ? round_half(v) : v;
            }
        }
    }

    // conditioning flag for the FP64 re-factorisation (BMFR_QR_KAHAN_FP64)
    // fminf/fmaxf skip a NaN operand, so a NaN pivot norm is tracked on its own
    if (flag.flag && active && t == 0) {
        float rmin = fabsf(norms[0]), rmax = fabsf(norms[0]);
        bool nan = !(norms[0] == norms[0]);
#pragma unroll
        for (int i = 1; i < kFeatures; ++i) {
            rmin = fminf(rmin, fabsf(norms[i]));
            rmax = fmaxf(rmax, fabsf(norms[i]));
            nan |= !(norms[i] == norms[i]);
        }
        *flag.flag = (rmin < flag.threshold * rmax || nan) ? 1 : 0;
    }

    // ---- back substitution, bmfr.cl:658-699, in wave 0 ----
    if (wave == 0)
        back_substitution([&](int c, int r) { return readlane(a[c][r % V], r / V); }, norms, weights, active,
                          lane);
}

// ---- Fast fitter body (plain FP32 sums): the default configuration ---------
// The reference's method with one reduction round per pivot column instead
// of two.  For pivot c, with x = column c below the pivot row and x_c the
// pivot-row value, one round reduces sigma = sum_{i>c} x_i^2 together with
// p_j = sum_{i>c} x_i a_ij for every trailing column j; then
// ||x|| = sqrt(sigma + x_c^2), u_c = x_c - ||x|| (the reference's sign,
// bmfr.cl:571-584), uu = sigma + u_c^2 and u.a_j = p_j + u_c a_cj, the
// reference's dot product (bmfr.cl:609-633) summed in another order.  The
// values of a round are reduced step-major (one DPP step for all of them
// before the next), so the independent chains fill each other's hazard
// slots.  A block is spread over W = 16/RPL waves of RPL rows per lane; with
// W > 1 the wave partials meet in LDS behind one barrier per round
// (ping-pong buffers).  The pivot rows 0..12 all live in wave 0.
// HALF_ROUND rounds every stored element to binary16 like the reference's
// USE_HALF_PRECISION_IN_TMP_DATA stores (bmfr.cl:255-258, 540, 651); the
// feature noise is added to the loaded value in FP64 and rounded once to
// float, and is never stored on its own (bmfr.cl:612-633).
template <int RPL, typename TT>
struct FastLayout {
    static constexpr int W = 16 / RPL;   // waves per block
    static constexpr int T = 64 * W;     // lanes per block (a power of two)
    static constexpr int V = RPL < (int)(16 / sizeof(TT)) ? RPL : (int)(16 / sizeof(TT));  // rows per load
    static constexpr int Q = RPL / V;    // loads per column
    // element k of block lane t holds row (k/V)*T*V + t*V + k%V
    __device__ static __forceinline__ int row(int k, int t) { return (k / V) * (T * V) + t * V + (k % V); }
};

constexpr int kFastSlots = 32;   // LDS words per wave and round: partials 0..12, wave-0 extras 16..28

template <int RPL, int BPG>
using FastRed = float[2][BPG][16 / RPL][kFastSlots];

// Combine the wave-reduced values `v` of the W waves of one block (Op order
// w = 0..W-1) and broadcast wave 0's uniform extras `e` through LDS.
template <int W, typename Op, int N, int NE>
__device__ __forceinline__ void fast_exchange(float (&v)[N], float (&e)[NE], float *rd, int wave, int lane) {
    if constexpr (W > 1) {
        static_assert(N <= 16 && NE <= 16, "slot layout");
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < N; ++i) rd[wave * kFastSlots + i] = v[i];
            if (wave == 0) {
#pragma unroll
                for (int i = 0; i < NE; ++i) rd[16 + i] = e[i];
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < N; ++i) {
            float acc = rd[i];
#pragma unroll
            for (int w = 1; w < W; ++w) acc = Op::f(acc, rd[w * kFastSlots + i]);
            v[i] = acc;
        }
#pragma unroll
        for (int i = 0; i < NE; ++i) e[i] = rd[16 + i];
    }
}

// Feature count of fitter template argument NF: 0 is the reference's set
// (10 features, 4 not scaled, compile-time), 1..16 a configured set.
template <int NF>
struct SetShape {
    static constexpr int F = NF ? NF : kFeatures;   // features
    static constexpr int C = F + 3;                  // BUFFER_COUNT
};

template <int RPL, typename TT, bool HALF_ROUND, int C = kBufferCount>
struct FastRows {
    using L = FastLayout<RPL, TT>;
    float a[C][RPL];
    // A Householder update (the compiler may round its FMA to binary16 once).
    __device__ __forceinline__ void put(int f, int k, float v) { a[f][k] = HALF_ROUND ? round_half(v) : v; }
    // A scaled feature: the float quotient, then binary16 (bmfr.cl:540).
    __device__ __forceinline__ void put_scaled(int f, int k, float v) {
        a[f][k] = HALF_ROUND ? round_half_of_float(v) : v;
    }
    // Row r (< 64 V) of column f, wave-uniform; valid in wave 0.
    __device__ __forceinline__ float pivot_row(int f, int r) const { return readlane(a[f][r % L::V], r / L::V); }
};

// A lane's share of a column dot product, sum over k of x[k] * y(k): four
// FMA chains joined pairwise (shorter chains round less than one running
// sum).
template <int RPL, typename Y>
__device__ __forceinline__ float lane_dot(const float (&x)[RPL], Y y) {
    static_assert(RPL % 4 == 0, "four chains");
    float s[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < RPL; ++k) s[k % 4] = __builtin_fmaf(x[k], y(k), s[k % 4]);
    return (s[0] + s[1]) + (s[2] + s[3]);
}

// Pivot step COL (round COL + 1): norm, u, the C - 1 - COL trailing dot
// products, update (C = F + 3 columns).
template <int COL, int RPL, int BPG, typename TT, bool HALF_ROUND, int F>
__device__ __forceinline__ void fast_pivot(FastRows<RPL, TT, HALF_ROUND, F + 3> &m, float (&norms)[F],
                                           FastRed<RPL, BPG> &red, int sub, int t, int wave, int lane) {
    using L = FastLayout<RPL, TT>;
    constexpr int NV = F + 3 - 1 - COL;
    constexpr int V = L::V;
    constexpr int KC = COL % V;   // element of lane COL / V holding the pivot row
    float x[RPL];
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
        // only elements k < V of lanes t <= COL / V hold rows <= COL
        if (k < V && k <= COL) x[k] = L::row(k, t) > COL ? m.a[COL][k] : 0.f;
        else x[k] = m.a[COL][k];
    }
    float v[NV + 1];
    v[0] = lane_dot<RPL>(x, [&](int k) { return x[k]; });
#pragma unroll
    for (int i = 0; i < NV; ++i) v[1 + i] = lane_dot<RPL>(x, [&](int k) { return m.a[COL + 1 + i][k]; });
    wave_reduce_n<OpAdd, NV + 1>(v);
    float e[NV + 1];   // x_c and the pivot row of the trailing columns
    e[0] = m.pivot_row(COL, COL);
#pragma unroll
    for (int i = 0; i < NV; ++i) e[1 + i] = m.pivot_row(COL + 1 + i, COL);
    fast_exchange<L::W, OpAdd>(v, e, &red[(COL + 1) & 1][sub][0][0], wave, lane);

    // v_sqrt_f32 and v_rcp_f32 (1 ulp) for the pivot's norm and 1/uu: the
    // reference's sqrt and per-element division are not correctly rounded
    // either, and the trailing update is checked by tolerance
    const float sigma = v[0], xc = e[0];
    const float norm = __builtin_amdgcn_sqrtf(sigma + xc * xc);
    const float uc = xc - norm;
    const float uu = sigma + uc * uc;
    norms[COL] = norm;
    // (2 d) / uu as d * (2 / uu): scaling by 2 is exact, so the same bits
    const float inv2_uu = 2.f * __builtin_amdgcn_rcpf(uu);
    const float u_kc = t == COL / V ? uc : x[KC];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int j = COL + 1 + i;
        const float d = v[1 + i] + uc * e[1 + i];
        const float fct = d * inv2_uu;
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
            const float u = k == KC ? u_kc : x[k];
            m.put(j, k, __builtin_fmaf(-u, fct, m.a[j][k]));
        }
    }
}

template <int COL, int RPL, int BPG, typename TT, bool HALF_ROUND, int F>
__device__ __forceinline__ void fast_pivots(FastRows<RPL, TT, HALF_ROUND, F + 3> &m, float (&norms)[F],
                                            FastRed<RPL, BPG> &red, int sub, int t, int wave, int lane) {
    if constexpr (COL < F) {
        fast_pivot<COL, RPL, BPG, TT, HALF_ROUND, F>(m, norms, red, sub, t, wave, lane);
        fast_pivots<COL + 1, RPL, BPG, TT, HALF_ROUND, F>(m, norms, red, sub, t, wave, lane);
    }
}

// ---- Binary16 pairs after the first pivot (one wave per block, binary16
// tmp_data) ----
// Every element the reference stores is a binary16 value, and after the
// first pivot's update (bmfr.cl:651) so is every element the later pivots
// read.  A lane's 16 rows of a column are then held as 8 pairs of binary16
// in 8 registers (element pair (2i, 2i+1), the layout the 16-byte loads
// deliver), the dot products are v_dot2_f32_f16 (two exact products and a
// sum a step) and the updates v_fma_mixlo_f16 / v_fma_mixhi_f16 (the FMA of
// the binary16 element and its binary16 u with the FP32 factor, rounded to
// binary16 once: the same single rounding as the unpacked update), written
// in asm because the compiler keeps each half in a register of its own and
// repacks the pairs for every dot product.
typedef _Float16 Half2 __attribute__((ext_vector_type(2)));

template <int C>
struct PackedRows {
    Half2 p[C][8];
    // Row r (< 512) of column f, wave-uniform: lane r / 8, element r % 8.
    __device__ __forceinline__ float pivot_row(int f, int r) const {
        const uint32_t s = (uint32_t)__builtin_amdgcn_readlane(__builtin_bit_cast(int, p[f][(r % 8) / 2]), r / 8);
        return (float)__builtin_bit_cast(_Float16, (uint16_t)((r & 1) ? s >> 16 : s));
    }
};

// a = RN16(a - u * f) in both halves, u binary16 pairs.  Each half is an
// asm statement of its own, so that the scheduler can put other work
// between the two writes of the register.
__device__ __forceinline__ void pair_update(Half2 &a, Half2 u, float f) {
    asm("v_fma_mixlo_f16 %0, -%1, %2, %0 op_sel_hi:[1,0,1]" : "+v"(a) : "v"(u), "v"(f));
    asm("v_fma_mixhi_f16 %0, -%1, %2, %0 op_sel:[1,0,1] op_sel_hi:[1,0,1]" : "+v"(a) : "v"(u), "v"(f));
}
// The same with u in FP32 (the pair holding the pivot row: u_c is not a
// binary16 value).
__device__ __forceinline__ void pair_update_f32u(Half2 &a, float ulo, float uhi, float f) {
    asm("v_fma_mixlo_f16 %0, -%1, %2, %0 op_sel_hi:[0,0,1]" : "+v"(a) : "v"(ulo), "v"(f));
    asm("v_fma_mixhi_f16 %0, -%1, %2, %0 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "+v"(a) : "v"(uhi), "v"(f));
}
// RN16(alo - ulo * f), RN16(ahi - uhi * f) from FP32 operands (the first
// pivot, whose noised features are FP32), as a pair in alo's register.
__device__ __forceinline__ Half2 pair_from_f32(float ulo, float uhi, float f, float alo, float ahi) {
    asm("v_fma_mixlo_f16 %0, -%1, %2, %0 op_sel_hi:[0,0,0]" : "+v"(alo) : "v"(ulo), "v"(f));
    asm("v_fma_mixhi_f16 %0, -%1, %2, %3 op_sel_hi:[0,0,0]" : "+v"(alo) : "v"(uhi), "v"(f), "v"(ahi));
    return __builtin_bit_cast(Half2, alo);
}

// A lane's share of sum_i x_i y_i over its 16 rows as pairs: two chains of
// v_dot2_f32_f16 (four measured slower: their accumulators start as moves).
__device__ __forceinline__ float pair_dot(const Half2 (&x)[8], const Half2 (&y)[8]) {
    float s[2] = {0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 8; ++i) s[i % 2] = __builtin_amdgcn_fdot2(x[i], y[i], s[i % 2], false);
    return s[0] + s[1];
}

// The first pivot: fast_pivot<0>'s arithmetic with the F feature columns in
// FP32 (the noised features are not binary16 values) and the three colour
// columns as binary16 pairs from the load; every updated column is written
// as pairs into `pk`.
template <int RPL, typename TT, int F>
__device__ __forceinline__ void packed_pivot0(FastRows<RPL, TT, true, F + 3> &m, PackedRows<F + 3> &pk,
                                              float (&norms)[F], int t) {
    static_assert(RPL == 16, "one wave per block");
    constexpr int NV = F + 3 - 1;
    float x[RPL];
#pragma unroll
    for (int k = 0; k < RPL; ++k) x[k] = (k == 0 && t == 0) ? 0.f : m.a[0][k];   // row 0 is the pivot
    Half2 xp[8];   // the same values as pairs (column 0 holds binary16 values)
#pragma unroll
    for (int q = 0; q < 8; ++q) xp[q] = Half2{(_Float16)x[2 * q], (_Float16)x[2 * q + 1]};
    float v[NV + 1];
    v[0] = lane_dot<RPL>(x, [&](int k) { return x[k]; });
#pragma unroll
    for (int i = 0; i < F - 1; ++i) v[1 + i] = lane_dot<RPL>(x, [&](int k) { return m.a[1 + i][k]; });
#pragma unroll
    for (int c = 0; c < 3; ++c) v[F + c] = pair_dot(xp, pk.p[F + c]);
    wave_reduce_n<OpAdd, NV + 1>(v);
    float e[NV + 1];
    e[0] = m.pivot_row(0, 0);
#pragma unroll
    for (int i = 0; i < F - 1; ++i) e[1 + i] = m.pivot_row(1 + i, 0);
#pragma unroll
    for (int c = 0; c < 3; ++c) e[F + c] = pk.pivot_row(F + c, 0);
    const float sigma = v[0], xc = e[0];
    const float norm = __builtin_amdgcn_sqrtf(sigma + xc * xc);
    const float uc = xc - norm;
    const float uu = sigma + uc * uc;
    norms[0] = norm;
    const float inv2_uu = 2.f * __builtin_amdgcn_rcpf(uu);
    const float u0 = t == 0 ? uc : x[0];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int j = 1 + i;
        const float fct = (v[1 + i] + uc * e[1 + i]) * inv2_uu;
        if (j < F) {
#pragma unroll
            for (int q = 0; q < 8; ++q)
                pk.p[j][q] = pair_from_f32(q == 0 ? u0 : x[2 * q], x[2 * q + 1], fct, m.a[j][2 * q],
                                           m.a[j][2 * q + 1]);
        } else {
            pair_update_f32u(pk.p[j][0], u0, x[1], fct);
#pragma unroll
            for (int q = 1; q < 8; ++q) pair_update(pk.p[j][q], xp[q], fct);
        }
    }
}

// Pivot COL >= 1 on binary16 pairs: fast_pivot's arithmetic (the dot
// products summed in another order).
template <int COL, int F>
__device__ __forceinline__ void packed_pivot(PackedRows<F + 3> &pk, float (&norms)[F], int t) {
    constexpr int NV = F + 3 - 1 - COL;
    constexpr int KC = COL % 8;   // element of lane COL / 8 holding the pivot row
    // x: the pivot column below the pivot row (rows <= COL zeroed; they sit
    // in elements 0..7 of lanes 0 and 1)
    Half2 x[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (i < 4 && 2 * i <= COL) {
            const int rlo = t * 8 + 2 * i;
            const uint32_t mask = (rlo > COL ? 0xFFFFu : 0u) | (rlo + 1 > COL ? 0xFFFF0000u : 0u);
            x[i] = __builtin_bit_cast(Half2, __builtin_bit_cast(uint32_t, pk.p[COL][i]) & mask);
        } else {
            x[i] = pk.p[COL][i];
        }
    }
    float v[NV + 1];
    v[0] = pair_dot(x, x);
#pragma unroll
    for (int i = 0; i < NV; ++i) v[1 + i] = pair_dot(x, pk.p[COL + 1 + i]);
    wave_reduce_n<OpAdd, NV + 1>(v);
    float e[NV + 1];
    e[0] = pk.pivot_row(COL, COL);
#pragma unroll
    for (int i = 0; i < NV; ++i) e[1 + i] = pk.pivot_row(COL + 1 + i, COL);
    const float sigma = v[0], xc = e[0];
    const float norm = __builtin_amdgcn_sqrtf(sigma + xc * xc);
    const float uc = xc - norm;
    const float uu = sigma + uc * uc;
    norms[COL] = norm;
    const float inv2_uu = 2.f * __builtin_amdgcn_rcpf(uu);
    // the pair holding the pivot row, with u_c in the pivot lane, in FP32
    const bool pl = t == COL / 8;
    const float ulo = (pl && (KC & 1) == 0) ? uc : (float)x[KC / 2].x;
    const float uhi = (pl && (KC & 1) == 1) ? uc : (float)x[KC / 2].y;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int j = COL + 1 + i;
        const float fct = (v[1 + i] + uc * e[1 + i]) * inv2_uu;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (q == KC / 2) pair_update_f32u(pk.p[j][q], ulo, uhi, fct);
            else pair_update(pk.p[j][q], x[q], fct);
        }
    }
}

template <int COL, int F>
__device__ __forceinline__ void packed_pivots(PackedRows<F + 3> &pk, float (&norms)[F], int t) {
    if constexpr (COL < F) {
        packed_pivot<COL, F>(pk, norms, t);
        packed_pivots<COL + 1, F>(pk, norms, t);
    }
}

// One 32x32 block by the W waves of work-group slice `sub`; `active` false
// recomputes a valid block (keeping barriers uniform) without storing.
// NF: the feature set (SetShape); a configured set (NF > 0) scales its
// features [not_scaled, NF) and runs one wave per block (RPL 16).
template <int RPL, int BPG, typename TT, bool HALF_ROUND, int NF = 0>
__device__ __forceinline__ void fitter_block_fast(const TT *__restrict__ src, float *__restrict__ mins_maxs,
                                                  float *__restrict__ weights, NoiseTable noise, bool active,
                                                  FastRed<RPL, BPG> &red, int not_scaled = kNotScaled) {
    using L = FastLayout<RPL, TT>;
    constexpr int F = SetShape<NF>::F, C = SetShape<NF>::C;
    static_assert(NF == 0 || RPL == 16, "a configured feature set runs one wave per block");
    constexpr int V = L::V, T = L::T;
    const int t = threadIdx.x & (T - 1);
    const int sub = threadIdx.x / T;
    const int lane = threadIdx.x & 63;
    const int wave = t >> 6;
    // binary16 pairs from the first pivot on (one wave per block, binary16
    // tmp_data): the colour columns stay pairs from the load
    constexpr bool kPairs = HALF_ROUND && RPL == 16;
    FastRows<RPL, TT, HALF_ROUND, C> m;
    PackedRows<C> pk;
#pragma unroll
    for (int f = 0; f < C; ++f) {
#pragma unroll
        for (int q = 0; q < L::Q; ++q) {
            const TT *p = src + f * kBlockPixels + q * (T * V) + t * V;
            if (kPairs && f >= F) {
                const uint4 w = *reinterpret_cast<const uint4 *>(p);
                pk.p[f][4 * q + 0] = __builtin_bit_cast(Half2, w.x);
                pk.p[f][4 * q + 1] = __builtin_bit_cast(Half2, w.y);
                pk.p[f][4 * q + 2] = __builtin_bit_cast(Half2, w.z);
                pk.p[f][4 * q + 3] = __builtin_bit_cast(Half2, w.w);
            } else {
                load_rows<TT, V>(p, &m.a[f][q * V]);
            }
        }
    }

    // ---- scaling, bmfr.cl:511-542 (max taken as -min(-x): exact) ----
    if constexpr (NF == 0) {
        float bmin[kScaled], bmax[kScaled];
        {
            float v[2 * kScaled];
#pragma unroll
            for (int i = 0; i < kScaled; ++i) {
                float mn = INFINITY, nmx = INFINITY;
#pragma unroll
                for (int k = 0; k < RPL; ++k) {
                    const float x = m.a[kNotScaled + i][k];
                    mn = fminf(mn, x);
                    nmx = fminf(nmx, -x);
                }
                v[i] = mn;
                v[kScaled + i] = nmx;
            }
            wave_reduce_n<OpMin, 2 * kScaled>(v);
            float none[1] = {0.f};
            fast_exchange<L::W, OpMin>(v, none, &red[0][sub][0][0], wave, lane);
#pragma unroll
            for (int i = 0; i < kScaled; ++i) {
                bmin[i] = v[i];
                bmax[i] = -v[kScaled + i];
            }
        }
        if (active && t == 0) {
#pragma unroll
            for (int i = 0; i < kScaled; ++i) {
                mins_maxs[2 * i + 0] = bmin[i];
                mins_maxs[2 * i + 1] = bmax[i];
            }
        }
        // scale (bmfr.cl:200-205).  (v-min)/(max-min) is evaluated as
        // Markstein's correctly rounded quotient: y = RN(1/d) once per
        // feature, then q = RN(a*y), r = RN(a - q*d) (exact),
        // q' = RN(q + r*y) = RN(a/d).  Without the division
        // (|max - min| <= 1, the same for the whole block) the numerator.
#pragma unroll
        for (int i = 0; i < kScaled; ++i) {
            const float dd = bmax[i] - bmin[i];
            if (fabsf(dd) > 1.0f) {
                const float y = 1.0f / dd;
#pragma unroll
                for (int k = 0; k < RPL; ++k)
                    m.put_scaled(kNotScaled + i, k, div_by(m.a[kNotScaled + i][k] - bmin[i], dd, y));
            } else {
#pragma unroll
                for (int k = 0; k < RPL; ++k) m.put_scaled(kNotScaled + i, k, m.a[kNotScaled + i][k] - bmin[i]);
            }
        }
    } else {
        // every feature's bounds are reduced (one round); those of the scaled
        // ones [not_scaled, F) are stored and applied, as above
        float v[2 * F];
#pragma unroll
        for (int f = 0; f < F; ++f) {
            float mn = INFINITY, nmx = INFINITY;
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
                mn = fminf(mn, m.a[f][k]);
                nmx = fminf(nmx, -m.a[f][k]);
            }
            v[f] = mn;
            v[F + f] = nmx;
        }
        wave_reduce_n<OpMin, 2 * F>(v);
#pragma unroll
        for (int f = 0; f < F; ++f) {
            if (f < not_scaled) continue;   // uniform: unscaled columns stay as loaded
            const float bmin = v[f], bmax = -v[F + f];
            if (active && t == 0) {
                mins_maxs[2 * (f - not_scaled) + 0] = bmin;
                mins_maxs[2 * (f - not_scaled) + 1] = bmax;
            }
            const float dd = bmax - bmin;
            if (fabsf(dd) > 1.0f) {
                const float y = 1.0f / dd;
#pragma unroll
                for (int k = 0; k < RPL; ++k) m.put_scaled(f, k, div_by(m.a[f][k] - bmin, dd, y));
            } else {
#pragma unroll
                for (int k = 0; k < RPL; ++k) m.put_scaled(f, k, m.a[f][k] - bmin);
            }
        }
    }
    // ---- feature noise of the first pass (bmfr.cl:623-627): the value plus
    // the FP64 noise term, rounded once to float (NOISE_AMOUNT is a double
    // literal); features 1..F-1 ----
#pragma unroll
    for (int f = 1; f < F; ++f) {
#pragma unroll
        for (int q = 0; q < L::Q; ++q) {
            const float *nz = noise.r + (f - 1) * kBlockPixels + q * (T * V) + t * V;
#pragma unroll
            for (int e = 0; e < V; e += 4) {
                const float4 x = *reinterpret_cast<const float4 *>(nz + e);
                const float r4[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    m.a[f][q * V + e + i] = (float)((double)m.a[f][q * V + e + i] + noise.term((double)r4[i]));
            }
        }
    }

    // ---- Householder QR over the F feature columns, bmfr.cl:549-656 ----
    float norms[F];
    if constexpr (kPairs) {
        packed_pivot0<RPL, TT, F>(m, pk, norms, t);
        packed_pivots<1, F>(pk, norms, t);
        back_substitution<F>([&](int c, int r) { return pk.pivot_row(c, r); }, norms, weights, active, lane);
    } else {
        fast_pivots<0, RPL, BPG, TT, HALF_ROUND, F>(m, norms, red, sub, t, wave, lane);
        // ---- back substitution, bmfr.cl:658-699, in wave 0 ----
        if (wave == 0)
            back_substitution<F>([&](int c, int r) { return m.pivot_row(c, r); }, norms, weights, active, lane);
    }
}


// The fast body serves every configuration with plain FP32 sums; the
// compensated sums (BMFR_QR_KAHAN) run fitter_block.  Both use the same LDS
// exchange area (red_slots(true) == kFastSlots words per wave and round).
static_assert(red_slots(true) == kFastSlots, "one LDS layout for both bodies");

template <int RPL, int BPG, typename TT, bool HALF_ROUND, bool COMP>
__global__ __launch_bounds__(64 * (16 / RPL) * BPG) void k_fitter(const FitterParams p) {
    __shared__ float red[2][BPG][16 / RPL][kFastSlots];
    const int sub = threadIdx.x / (64 * (16 / RPL));
    // Work-group g runs on XCD g % 8 and takes the blocks of that XCD's eighth
    // of the frame (raster order, as stage 1's XCD-contiguous tiles wrote
    // them), last-written first, so that its tmp_data can still be in that
    // XCD's L2.
    const int x = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
    const int b0 = (int)((int64_t)x * p.blocks / kXcds), b1 = (int)((int64_t)(x + 1) * p.blocks / kXcds);
    const int idx = j * BPG + sub;
    const bool active = idx < b1 - b0;
    const int blk = active ? b1 - 1 - idx : 0;
    if constexpr (!COMP) {
        fitter_block_fast<RPL, BPG, TT, HALF_ROUND>(
            reinterpret_cast<const TT *>(p.tmp) + (int64_t)blk * (kBufferCount * kBlockPixels),
            p.mins_maxs + (int64_t)blk * (kScaled * 2), p.weights + (int64_t)blk * (kFeatures * 3),
            NoiseTable{p.noise, p.noise_scale}, active, red);
    } else {
        fitter_block<RPL, BPG, TT, HALF_ROUND, COMP>(
            reinterpret_cast<const TT *>(p.tmp) + (int64_t)blk * (kBufferCount * kBlockPixels),
            p.mins_maxs + (int64_t)blk * (kScaled * 2), p.weights + (int64_t)blk * (kFeatures * 3),
            NoiseTable{p.noise, p.noise_scale}, BlockFlag{p.flags ? p.flags + blk : nullptr, p.threshold}, active,
            red);
    }
}

// A configured feature set of NF features (NF + 3 columns): one wave per
// block, two blocks per work-group, the XCD-matched order of k_fitter.
template <typename TT, bool HALF_ROUND, int NF>
__global__ __launch_bounds__(128) void k_fitter_set(const FitterParams p) {
    constexpr int BPG = 2, C = NF + 3;
    __shared__ float red[2][BPG][1][kFastSlots];
    const int sub = threadIdx.x >> 6;
    const int x = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
    const int b0 = (int)((int64_t)x * p.blocks / kXcds), b1 = (int)((int64_t)(x + 1) * p.blocks / kXcds);
    const int idx = j * BPG + sub;
    const bool active = idx < b1 - b0;
    const int blk = active ? b1 - 1 - idx : 0;
    fitter_block_fast<16, BPG, TT, HALF_ROUND, NF>(
        reinterpret_cast<const TT *>(p.tmp) + (int64_t)blk * (C * kBlockPixels),
        p.mins_maxs + (int64_t)blk * ((NF - p.not_scaled) * 2), p.weights + (int64_t)blk * (NF * 3),
        NoiseTable{p.noise, p.noise_scale}, active, red, p.not_scaled);
}

template <int NF>
inline hipError_t launch_set(const FitterParams &fp, bool tmp_f16, hipStream_t s) {
    const int per_xcd = (fp.blocks + kXcds - 1) / kXcds;
    const dim3 grid(kXcds * ((per_xcd + 1) / 2));
    if (tmp_f16) hipLaunchKernelGGL((k_fitter_set<__half, true, NF>), grid, dim3(128), 0, s, fp);
    else hipLaunchKernelGGL((k_fitter_set<float, false, NF>), grid, dim3(128), 0, s, fp);
    return hipGetLastError();
}

// Persistent batched fitter: the work-groups stride over the concatenated
// block list of every stream in the batch (edge blocks included).  The trip
// count depends only on blockIdx, so barriers stay uniform.
// The fast binary16-pair body is held to two waves per SIMD (256 registers;
// the batch loop's own registers would take it just past that).
template <int RPL, int BPG, typename TT, bool HALF_ROUND, bool COMP>
__global__ __launch_bounds__(64 * (16 / RPL) * BPG)
__attribute__((amdgpu_waves_per_eu(HALF_ROUND && RPL == 16 && !COMP ? 2 : 1))) void kb_fitter(const BatchLaunch L) {
    __shared__ float red[2][BPG][16 / RPL][kFastSlots];
    const int32_t *starts = L.starts + kStartsBlocks * (kMaxBatch + 1);
    const int total = starts[L.n];
    const int sub = threadIdx.x / (64 * (16 / RPL));
    for (int g = blockIdx.x; g * BPG < total; g += gridDim.x) {
        const int raw = g * BPG + sub;
        const bool active = raw < total;
        const int u = active ? raw : total - 1;
        const int s = batch_stream_of(starts, L.n, u);
        const int blk = u - starts[s];
        const BatchStream &bs = L.streams[s];
        // flags only for BMFR_QR_KAHAN_FP64 streams, as the single-stream launch
        int32_t *fl = bs.args.qr_mode == 2 ? (int32_t *)bs.flags : nullptr;
        if constexpr (!COMP) {
            fitter_block_fast<RPL, BPG, TT, HALF_ROUND>(
                reinterpret_cast<const TT *>(bs.tmp_data) + (int64_t)blk * (kBufferCount * kBlockPixels),
                (float *)bs.mins_maxs + (int64_t)blk * (kScaled * 2),
                (float *)bs.weights + (int64_t)blk * (kFeatures * 3),
                NoiseTable{(const float *)bs.noise, 2.0 * bs.args.noise_amount}, active, red);
        } else {
            fitter_block<RPL, BPG, TT, HALF_ROUND, COMP>(
                reinterpret_cast<const TT *>(bs.tmp_data) + (int64_t)blk * (kBufferCount * kBlockPixels),
                (float *)bs.mins_maxs + (int64_t)blk * (kScaled * 2),
                (float *)bs.weights + (int64_t)blk * (kFeatures * 3),
                NoiseTable{(const float *)bs.noise, 2.0 * bs.args.noise_amount},
                BlockFlag{fl ? fl + blk : nullptr, bs.args.refactor_threshold}, active, red);
        }
        // A block ends with its last round in LDS buffer 0 (round 20 of
        // fitter_block, round 10 of fitter_block_fast) and the next one starts
        // with round 0 in the same buffer: wait until every wave has read it.
        if constexpr (16 / RPL > 1) __syncthreads();
    }
}

template <int RPL, int BPG, bool COMP>
inline hipError_t launch_rpl(const FitterParams &fp, bool tmp_f16, hipStream_t s) {
    constexpr int threads = 64 * (16 / RPL) * BPG;
    const int per_xcd = (fp.blocks + kXcds - 1) / kXcds;   // XCD-matched block order (k_fitter)
    const dim3 grid(kXcds * ((per_xcd + BPG - 1) / BPG));
    if (tmp_f16)
        hipLaunchKernelGGL((k_fitter<RPL, BPG, __half, true, COMP>), grid, dim3(threads), 0, s, fp);
    else
        hipLaunchKernelGGL((k_fitter<RPL, BPG, float, false, COMP>), grid, dim3(threads), 0, s, fp);
    return hipGetLastError();
}

template <int RPL, int BPG, bool COMP>
inline hipError_t launch_batch_rpl(const BatchLaunch &L, int total, bool tmp_f16, int grid_cap,
                                   hipStream_t s) {
    constexpr int threads = 64 * (16 / RPL) * BPG;
    int groups = (total + BPG - 1) / BPG;
    if (grid_cap > 0 && groups > grid_cap) groups = grid_cap;
    if (tmp_f16)
        hipLaunchKernelGGL((kb_fitter<RPL, BPG, __half, true, COMP>), dim3(groups), dim3(threads), 0, s, L);
    else
        hipLaunchKernelGGL((kb_fitter<RPL, BPG, float, false, COMP>), dim3(groups), dim3(threads), 0, s, L);
    return hipGetLastError();
}

template <bool COMP>
inline hipError_t dispatch_single(const FitterParams &fp, bool h, const FitterConfig &cfg, hipStream_t s) {
    switch (cfg.rows_per_lane * 10 + cfg.blocks_per_group) {
        case 41: return launch_rpl<4, 1, COMP>(fp, h, s);
        case 42: return launch_rpl<4, 2, COMP>(fp, h, s);
        case 81: return launch_rpl<8, 1, COMP>(fp, h, s);
        case 82: return launch_rpl<8, 2, COMP>(fp, h, s);
        case 161: return launch_rpl<16, 1, COMP>(fp, h, s);
        case 162: return launch_rpl<16, 2, COMP>(fp, h, s);
        case 164: return launch_rpl<16, 4, COMP>(fp, h, s);
        default: return hipErrorInvalidValue;
    }
}

template <bool COMP>
inline hipError_t dispatch_batch(const BatchLaunch &L, int total, bool h, const FitterConfig &cfg,
                                 int grid_cap, hipStream_t s) {
    switch (cfg.rows_per_lane * 10 + cfg.blocks_per_group) {
        case 41: return launch_batch_rpl<4, 1, COMP>(L, total, h, grid_cap, s);
        case 42: return launch_batch_rpl<4, 2, COMP>(L, total, h, grid_cap, s);
        case 81: return launch_batch_rpl<8, 1, COMP>(L, total, h, grid_cap, s);
        case 82: return launch_batch_rpl<8, 2, COMP>(L, total, h, grid_cap, s);
        case 161: return launch_batch_rpl<16, 1, COMP>(L, total, h, grid_cap, s);
        case 162: return launch_batch_rpl<16, 2, COMP>(L, total, h, grid_cap, s);
        case 164: return launch_batch_rpl<16, 4, COMP>(L, total, h, grid_cap, s);
        default: return hipErrorInvalidValue;
    }
}

// The four launch families, each instantiated in a translation unit of its
// own (fitter_single.hip, fitter_single_comp.hip, fitter_batch.hip,
// fitter_batch_comp.hip) so that they compile in parallel.
hipError_t fitter_dispatch_single(const FitterParams &fp, bool tmp_f16, bool comp, const FitterConfig &cfg,
                                  hipStream_t s);
hipError_t fitter_dispatch_batch(const BatchLaunch &L, int total, bool tmp_f16, bool comp,
                                 const FitterConfig &cfg, int grid_cap, hipStream_t s);
// Configured feature sets of 1..8 and 9..16 features (fitter_set_lo.hip,
// fitter_set_hi.hip).
hipError_t fitter_set_lo(const FitterParams &fp, bool tmp_f16, int features, hipStream_t s);
hipError_t fitter_set_hi(const FitterParams &fp, bool tmp_f16, int features, hipStream_t s);

}  // namespace bmfr
