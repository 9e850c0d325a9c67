// Device-side helpers shared by the gfx950 BMFR kernels.
//
// The whole library is compiled with -ffp-contract=off: every float
// operation in the stage kernels is rounded separately, in the order of the
// reference kernels (opencl/bmfr.cl), so that integer decisions derived from
// float math (reprojection floor, accept bits, spp rounding) are
// reproducible bit for bit.  Kernels that want fused multiply-adds ask for
// them explicitly with __builtin_fmaf.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace bmfr {

constexpr int kBlock = 32;               // BLOCK_EDGE_LENGTH
constexpr int kBlockPixels = 1024;       // BLOCK_PIXELS
constexpr int kBufferCount = 13;         // BUFFER_COUNT: 10 features + RGB
constexpr int kFeatures = 10;
constexpr int kNotScaled = 4;            // FEATURES_NOT_SCALED: 1, n.xyz
constexpr int kScaled = 6;               // FEATURES_SCALED: p.xyz, p.xyz^2
constexpr int kREdge = 11;               // R_EDGE
// Factor `id` of a pixel: 1, n.x, n.y, n.z, p.x, p.y, p.z (ids are
// wave-uniform kernel arguments, so the selects are scalar-driven).
__device__ __forceinline__ float feature_factor(int id, float3 n, float3 p) {
    float v = 1.f;
    v = id == 1 ? n.x : v;
    v = id == 2 ? n.y : v;
    v = id == 3 ? n.z : v;
    v = id == 4 ? p.x : v;
    v = id == 5 ? p.y : v;
    v = id == 6 ? p.z : v;
    return v;
}

// Feature f of the set: the reference evaluates the expression text in
// float (bmfr.cl:448-453, 727-729); a single term is x * 1 = x exactly.
__device__ __forceinline__ float feature_value(const FeatureSet &fs, int f, float3 n, float3 p) {
    return feature_factor(fs.factor[f][0], n, p) * feature_factor(fs.factor[f][1], n, p);
}

// Per-frame jitter of the block grid, opencl/bmfr.cl:267-285.  A table of
// constants that has to be identical for the results to match.
static __constant__ int2 kBlockOffsets[16] = {
    {-14, -14}, {4, -6},  {-8, 14},  {8, 0},   {-10, -8}, {2, 12},
    {12, -12},  {-10, 0}, {12, 14},  {-8, -16}, {6, 6},   {-2, -2},
    {6, -14},   {-16, 12}, {14, -4}, {-6, 4}};

__device__ __forceinline__ int2 block_offset(int frame) {
    int k = frame % 16;
    if (k < 0) k += 16;
    return kBlockOffsets[k];
}

// opencl/bmfr.cl:209-216; valid when |index| is less than one size out.
__device__ __forceinline__ int mirror(int index, int size) {
    if (index < 0) return -index - 1;
    if (index >= size) return 2 * size - index - 1;
    return index;
}

// Global-memory access through an explicitly global pointer.  Every buffer
// the kernels touch is global device memory, but a pointer read from device
// memory (the batched launches' per-stream tables) is a generic pointer to
// the compiler, which then uses flat instructions: those also count against
// the LDS counter, so every wait on an LDS access waits for the loads in
// flight too, and block-uniform loads cannot become scalar loads.  The cast
// states the address space, so the accesses compile to global (or scalar)
// loads and stores whichever way the pointer was obtained.
// Class types (__half, the HIP vector types) go through a builtin type of
// their size: their constructors and assignment take generic references.
typedef unsigned int gword4 __attribute__((ext_vector_type(4)));
template <int N> struct GlobalRaw;
template <> struct GlobalRaw<1> { typedef unsigned char type; };
template <> struct GlobalRaw<2> { typedef unsigned short type; };
template <> struct GlobalRaw<4> { typedef unsigned int type; };
template <> struct GlobalRaw<8> { typedef unsigned long long type; };
template <> struct GlobalRaw<16> { typedef gword4 type; };
template <typename T>
__device__ __forceinline__ T gload(const T *p, int64_t i) {
    typedef typename GlobalRaw<sizeof(T)>::type U;
    const U u = ((const __attribute__((address_space(1))) U *)p)[i];
    T v;
    __builtin_memcpy(&v, &u, sizeof(T));
    return v;
}
template <typename T>
__device__ __forceinline__ void gstore(T *p, int64_t i, T v) {
    typedef typename GlobalRaw<sizeof(T)>::type U;
    U u;
    __builtin_memcpy(&u, &v, sizeof(T));
    ((__attribute__((address_space(1))) U *)p)[i] = u;
}

// Storage-type access for G-buffers and histories (float or half).
template <typename T> struct Store;
template <> struct Store<float> {
    __device__ __forceinline__ static float ld(const float *p, int64_t i) { return gload(p, i); }
    __device__ __forceinline__ static void st(float *p, int64_t i, float v) { gstore(p, i, v); }
};
template <> struct Store<__half> {
    __device__ __forceinline__ static float ld(const __half *p, int64_t i) {
        return __half2float(gload(p, i));
    }
    __device__ __forceinline__ static void st(__half *p, int64_t i, float v) {
        gstore(p, i, __float2half_rn(v));
    }
};

template <typename T>
__device__ __forceinline__ float3 load3(const T *p, int64_t pixel) {
    return make_float3(Store<T>::ld(p, pixel * 3 + 0), Store<T>::ld(p, pixel * 3 + 1),
                       Store<T>::ld(p, pixel * 3 + 2));
}
template <typename T>
__device__ __forceinline__ void store3(T *p, int64_t pixel, float3 v) {
    Store<T>::st(p, pixel * 3 + 0, v.x);
    Store<T>::st(p, pixel * 3 + 1, v.y);
    Store<T>::st(p, pixel * 3 + 2, v.z);
}

// Per-pixel buffers (one to three elements a pixel) of the largest frame the
// library accepts, 16384 x 16384, hold at most 3 GiB, so a byte offset into
// one fits in 32 bits.  Formed in 32 bits and added to the buffer's uniform
// base, it lets the compiler use the scalar-base form of the global
// instructions (one VGPR of offset) instead of a 64-bit VGPR address per
// access.  `px` is the element index (pixel * elements per pixel for the
// single-element forms, the pixel for the three-element ones).
template <typename T>
__device__ __forceinline__ T *px_ptr(T *p, uint32_t byte_offset) {
    return (T *)((char *)p + byte_offset);
}
template <typename T>
__device__ __forceinline__ const T *px_ptr(const T *p, uint32_t byte_offset) {
    return (const T *)((const char *)p + byte_offset);
}
template <typename T>
__device__ __forceinline__ T gload_px(const T *p, uint32_t px) {
    return gload(px_ptr(p, px * (uint32_t)sizeof(T)), 0);
}
template <typename T>
__device__ __forceinline__ void gstore_px(T *p, uint32_t px, T v) {
    gstore(px_ptr(p, px * (uint32_t)sizeof(T)), 0, v);
}
// p * 3 as one full-rate v_lshl_add_u32 ((p << 1) + p).  Written as asm
// because the compiler turns the shift-add back into a multiply, and a
// 32-bit v_mul_lo_u32 issues at a quarter of the rate.
__device__ __forceinline__ uint32_t times3(uint32_t p) {
    uint32_t r;
    asm("v_lshl_add_u32 %0, %1, 1, %1" : "=v"(r) : "v"(p));
    return r;
}
// Byte offset of pixel `pixel` in a three-element-per-pixel buffer of T.
template <typename T>
__device__ __forceinline__ uint32_t px3_bytes(uint32_t pixel) {
    return times3(pixel) * (uint32_t)sizeof(T);
}
template <typename T>
__device__ __forceinline__ float3 load3_at(const T *p, uint32_t byte_offset) {
    const T *q = px_ptr(p, byte_offset);
    return make_float3(Store<T>::ld(q, 0), Store<T>::ld(q, 1), Store<T>::ld(q, 2));
}
template <typename T>
__device__ __forceinline__ void store3_at(T *p, uint32_t byte_offset, float3 v) {
    T *q = px_ptr(p, byte_offset);
    Store<T>::st(q, 0, v.x);
    Store<T>::st(q, 1, v.y);
    Store<T>::st(q, 2, v.z);
}
template <typename T>
__device__ __forceinline__ float3 load3_px(const T *p, uint32_t pixel) {
    return load3_at(p, px3_bytes<T>(pixel));
}
template <typename T>
__device__ __forceinline__ void store3_px(T *p, uint32_t pixel, float3 v) {
    store3_at(p, px3_bytes<T>(pixel), v);
}


// convert_int_rtn (opencl/bmfr.cl:356, 794, 879) with an explicitly
// saturating conversion, NaN -> 0 (what v_cvt_i32_f32 does); written out so
// that no out-of-range float->int conversion is left undefined.
__device__ __forceinline__ int cvt_rtn(float f) {
    if (f != f) return 0;
    const float fl = floorf(f);
    if (fl >= 2147483648.0f) return 2147483647;
    if (fl < -2147483648.0f) return -2147483647 - 1;
    return (int)fl;
}

// Integer sums of reprojected coordinates wrap (two's complement).
__device__ __forceinline__ int wrap_add(int a, int b) {
    return (int)((unsigned)a + (unsigned)b);
}
__device__ __forceinline__ int wrap_mad(int a, int b, int c) {
    return (int)((unsigned)a * (unsigned)b + (unsigned)c);
}

// The 2x2 bilinear footprint at a reprojected position pf (bmfr.cl:356-360,
// 794-798, 879-890): top-left texel (ix, iy) = floor(pf); per texel i
// (ix + (i & 1), iy + (i >> 1)) whether it lies in the W x H image, and the
// pixel index of that texel clamped into the image, which is the texel
// itself when it lies inside and otherwise some valid address whose value
// is not used (so that no load needs a select or a branch).  Row products
// run on the 24-bit multiplier (clamped rows < H <= 16384, W <= 16384).
struct Footprint {
    int ix, iy;          // floor(pf), saturated (cvt_rtn)
    uint32_t idx[4];     // pixel index of texel i clamped into the image
    uint32_t in_image;   // bit i: texel i in the image
};
__device__ __forceinline__ Footprint footprint(float2 pf, int W, int H) {
    Footprint f;
    f.ix = cvt_rtn(pf.x);
    f.iy = cvt_rtn(pf.y);
    const int ix1 = wrap_add(f.ix, 1), iy1 = wrap_add(f.iy, 1);
    const uint32_t cx0 = (uint32_t)min(max(f.ix, 0), W - 1), cx1 = (uint32_t)min(max(ix1, 0), W - 1);
    const uint32_t r0 = __umul24((unsigned)min(max(f.iy, 0), H - 1), (unsigned)W);
    const uint32_t r1 = __umul24((unsigned)min(max(iy1, 0), H - 1), (unsigned)W);
    f.idx[0] = r0 + cx0;
    f.idx[1] = r0 + cx1;
    f.idx[2] = r1 + cx0;
    f.idx[3] = r1 + cx1;
    // bits 0 and 2 are column ix, 1 and 3 column ix + 1; bits 0, 1 row iy
    const uint32_t xm = ((uint32_t)f.ix < (uint32_t)W ? 5u : 0u) | ((uint32_t)ix1 < (uint32_t)W ? 10u : 0u);
    const uint32_t ym = ((uint32_t)f.iy < (uint32_t)H ? 3u : 0u) | ((uint32_t)iy1 < (uint32_t)H ? 12u : 0u);
    f.in_image = xm & ym;
    return f;
}

// opencl/bmfr.cl:162-171.  The hash argument of the feature noise
// (bmfr.cl:179-181), id + sub_vector*256, equals the row index in the block.
__device__ __forceinline__ uint32_t hash32(uint32_t a) {
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    return a;
}

// The feature noise of one frame depends only on (row, feature, frame)
// (bmfr.cl:179-181 has no block term), so one table of (N-1) x 1024 terms
// serves every block of a frame (N features, N + 3 buffers).  The table
// holds the float r = random(row + 1024 f + 1024 (N+3) frame) - 0.5f that
// the reference computes; the term added is NOISE_AMOUNT * 2 * r in double
// (NOISE_AMOUNT is a double literal, bmfr.cl:179), formed where it is used
// (NoiseTable::term), so the table is 4 bytes per term.
constexpr int kNoiseTable = (kFeatures - 1) * kBlockPixels;
constexpr int kMaxNoiseTable = (kMaxFeatures - 1) * kBlockPixels;
__device__ __forceinline__ float noise_random(int row, int feature, int frame, int buffer_count = kBufferCount) {
    const uint32_t arg = (uint32_t)row + (uint32_t)feature * kBlockPixels +
                         (uint32_t)frame * (uint32_t)(buffer_count * kBlockPixels);
    return (float)hash32(arg) / 4294967296.0f - 0.5f;
}

// A frame's noise table and 2 * NOISE_AMOUNT: term(i) = RN64(scale * r[i]),
// the reference's NOISE_AMOUNT * 2.f * (random(...) - 0.5f) (the doubling
// is exact).
struct NoiseTable {
    const float *r;
    double scale;
    __device__ __forceinline__ double term(double rv) const { return scale * rv; }
};

// Correctly rounded float quotients a / d for several numerators sharing one
// divisor d: one FP64 reciprocal y (v_rcp_f64, then two Newton steps, each
// with one rounding: relative error below 2^-52), then (float)(a * y).
// Why that is the IEEE quotient: for floats a and d, a / d is either a float
// or at least 2^-49 (relative) away from every rounding boundary of float
// (a - d m, for a boundary m, is a nonzero multiple of 2^-48 |a| / |m|-ish
// units: |a/d - m| >= 2^(e-48) with 2^e <= |a/d|; for subnormal quotients
// the distance is at least 2^-174 against an error below 2^-178), and the
// double product a * y is within 2^-51 of a / d, so both round to the same
// float.  Zero, infinite and NaN divisors take v_rcp_f64's own result (inf,
// 0, NaN), which gives a * y the IEEE quotient's value and sign.  Checked on
// the host against IEEE division (tests/native/quotient_check.c: 4e8 pairs,
// including quotients next to rounding boundaries, with the reciprocal's
// first approximation perturbed by up to 2^-14; one Newton step fails it).
// Three quotients cost 15 instructions against about 30.
__device__ __forceinline__ double quotient_reciprocal(float d) {
    const double dd = (double)d;
    const double y0 = __builtin_amdgcn_rcp(dd);
    double e = __builtin_fma(-dd, y0, 1.0);
    double y = __builtin_fma(y0, e, y0);
    e = __builtin_fma(-dd, y, 1.0);
    y = __builtin_fma(y, e, y);
    return (__builtin_isfinite(dd) && dd != 0.0) ? y : y0;
}
__device__ __forceinline__ float quotient(float a, double y) { return (float)((double)a * y); }

// opencl/bmfr.cl:200-205
__device__ __forceinline__ float scale_feature(float v, float mn, float mx) {
    if (fabsf(mx - mn) > 1.0f) return (v - mn) / (mx - mn);
    return v - mn;
}

// Work-groups are dispatched round-robin over the 8 XCDs (work-group b runs
// on XCD b % 8), and each XCD has its own L2.  Give each XCD a contiguous
// range of the n tiles so that the halo a tile shares with its neighbours
// (bilinear taps, 3x3 neighbourhoods, mirrored margins) is read through the
// same L2.  Bijection on [0, n): XCD x owns q + (x < r) tiles, q = n / 8,
// r = n % 8, starting at x*q + min(x, r).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int q = n / kXcds, r = n % kXcds;
    const int x = b % kXcds, j = b / kXcds;
    return x * q + (x < r ? x : r) + j;
}

// Member of a batch that owns work unit `unit`: the largest s < n with
// starts[s] <= unit (starts[0] = 0, non-decreasing), by binary search over
// the prefix table (wave-uniform when `unit` is).
__device__ __forceinline__ int batch_stream_of(const int32_t *starts, int n, int unit) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= unit) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ float round_half(float v) {
    return __half2float(__float2half_rn(v));
}

// round_half of a value rounded to float first: a float result stored with
// vstore_half.  Without the empty asm the compiler may fold the float
// operation that produced v into its binary16 conversion (v_fma_mixlo_f16
// and the like), which rounds the exact result to binary16 once: a different
// binary16 value for about 1 in 16,000 FMAs.
__device__ __forceinline__ float round_half_of_float(float v) {
    asm("" : "+v"(v));
    return round_half(v);
}

}  // namespace bmfr
