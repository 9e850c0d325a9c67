// The tone curve of stage 4 (opencl/bmfr.cl:851-856):
//   clamp(pow(max(0, v), 0.454545f), 0, 1)
// evaluated in FP64 by table reduction, our own method.
//
// Why not powf: the device library's single-precision pow carries the
// extended-precision log/exp it needs for arbitrary exponents and is about
// 160 VALU instructions; three of them were over half of the fused
// weighted_sum/accumulate_filtered kernel.  gfx950 issues FP64 FMAs at the
// FP32 (non-packed) rate, so a short FP64 evaluation is far cheaper.
//
// Method, for 0 < v < 1 (every other input clamps to 0 or 1), with
// y = (double)0.454545f:
//   v = m 2^e, m in [0.5, 1), e in [-148, 0]          (frexp, exact)
//   i = the 8 bits of m after its leading one; c_i ~ 1 / m with 12
//   significant bits, so r = m c_i - 1 is exact (one fma) and |r| < 2.3e-3
//   v^y = 2^(y e) * c_i^-y * (1 + r)^y
//   (1 + r)^y = 1 + y r + C(y,2) r^2 + ... + C(y,5) r^5  (remainder < 3e-18)
// 2^(-y k) (k = -e) and c_i^-y are correctly rounded doubles from two
// tables (tone_tables.h, gen_tone_tables.py), indexed by e and i straight
// after frexp, so both loads are in flight together; the product is rounded
// once to float.  The relative error before that rounding is below 5e-16,
// so the result is the correctly rounded power except when the exact power
// lies within that distance of a rounding boundary, and then it is one ulp
// away.
//
// The function uses only IEEE-exact operations (fma, frexp, table loads),
// so the host build of this header computes the same bits as the device
// build (both with contraction off); the exhaustive check over all floats
// in [0, 1] runs the host build.
#pragma once

#if defined(__HIPCC__)
#define BMFR_TONE_FN __host__ __device__ __forceinline__
#define BMFR_TONE_TABLE static __constant__
#else
#define BMFR_TONE_FN inline
#define BMFR_TONE_TABLE static
#endif
#define BMFR_TONE_ALIGN __attribute__((aligned(16)))

namespace bmfr {

#include "tone_tables.h"

// The two tables as the function reads them: from the constant tables
// above, or from a copy in LDS (a kernel that runs at low occupancy, where
// an L1/L2 round trip per pixel is not hidden, copies them once per
// work-group; same values, so the same bits).
struct ToneTables {
    BMFR_TONE_FN double c(int i) const { return kToneLog[i][0]; }
    BMFR_TONE_FN double cy(int i) const { return kToneLog[i][1]; }
    BMFR_TONE_FN double pow2(int k) const { return kTonePow2[k]; }
};

// The series is summed as acc += c_k * r^k from the smallest term up, with
// the powers formed by multiplication: the coefficient is then a multiplicand
// (an SGPR operand of v_fmac_f64) rather than the addend of a Horner step,
// which the compiler would first copy into VGPRs at every step.
//
// There is no branch: every input is evaluated (one outside (0, 1) at 0.5)
// and the clamps are applied at the end, so that the compiler can keep the
// table loads of several calls (the three channels of a pixel) in flight
// together instead of one exec-masked block after another.
template <typename Tables>
BMFR_TONE_FN float tone_curve(float v, const Tables &tab) {
    const bool inside = v > 0.f && v < 1.f;
    int e = 0;
    const double m = __builtin_frexp((double)(inside ? v : 0.5f), &e);   // at most 24 bits
    unsigned long long mb;
    __builtin_memcpy(&mb, &m, sizeof mb);
    const int i = (int)((mb >> 44) & 255);      // m in [0.5 + i/512, 0.5 + (i+1)/512)
    const double scale = tab.pow2(-e);          // 2^(y e); e in [-148, 0]
    const double c = tab.c(i);
    const double cy = tab.cy(i);                // c_i^-y
    const double r = __builtin_fma(m, c, -1.0);  // exact: a multiple of 2^-36 below 2^-8
    const double r2 = r * r, r3 = r2 * r, r4 = r2 * r2, r5 = r4 * r;
    double q = r5 * kToneBinom5;                 // (1 + r)^y - 1
    q = __builtin_fma(kToneBinom4, r4, q);
    q = __builtin_fma(kToneBinom3, r3, q);
    q = __builtin_fma(kToneBinom2, r2, q);
    q = __builtin_fma((double)0.454545f, r, q);
    const double sc = scale * cy;
    const float res = (float)__builtin_fma(sc, q, sc);   // a normal float in (0, 1]
    // res * 1 + 0 inside (0, 1); res * 0 + 1 for v >= 1 (inf included: pow
    // >= 1 clamps to 1); res * 0 + 0 = +0 for v <= 0 and NaN (max(0, v), and
    // pow(0, y) = 0).  res is finite, so each case is exact.  An fma rather
    // than a select: the compiler turns a select whose operand hangs off the
    // table loads back into a branch around them.
    return __builtin_fmaf(res, inside ? 1.f : 0.f, v >= 1.f ? 1.f : 0.f);
}

BMFR_TONE_FN float tone_curve(float v) { return tone_curve(v, ToneTables{}); }

}  // namespace bmfr

#undef BMFR_TONE_FN
#undef BMFR_TONE_TABLE
#undef BMFR_TONE_ALIGN
