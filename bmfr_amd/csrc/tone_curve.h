// The tone curve of stage 4 (opencl/bmfr.cl:851-856):
//   clamp(pow(max(0, v), 0.454545f), 0, 1)
// evaluated in FP64 from a logarithm and an exponential of our own.
//
// Why not powf: the device library's single-precision pow carries the
// extended-precision log/exp it needs for arbitrary exponents and is about
// 160 VALU instructions; three of them were over half of the fused
// weighted_sum/accumulate_filtered kernel.  gfx950 issues FP64 FMAs at the
// FP32 (non-packed) rate, so a short FP64 evaluation is about 3x cheaper.
//
// Method, for 0 < v < 1 (every other input clamps to 0 or 1):
//   v = m 2^e with m in [sqrt(1/2), sqrt(2))          (frexp, exact)
//   ln m = 2 atanh(s), s = (m-1)/(m+1), |s| <= 0.1716,
//        = 2s (1 + s^2/3 + ... + s^14/15)              (remainder < 4e-14)
//   t = y log2 v = y e + (y / ln 2) ln m,  y = (double)0.454545f
//   2^t = 2^n 2^f, n = rint(t), f = t - n (exact), |f| <= 1/2,
//   2^f = exp(f ln 2), Taylor to degree 10             (remainder < 1e-14)
// and one final rounding to float.  The relative error before that rounding
// is below 1e-13, so the result is the correctly rounded power except when
// the exact power lies within ~1e-13 (relative) of a rounding boundary, and
// then it is one ulp away.
//
// The function uses only IEEE-exact operations (division, fma, frexp,
// ldexp, rint), so the host build of this header computes the same bits as
// the device build (both with contraction off); the exhaustive check over
// all floats in [0, 1] runs the host build.
#pragma once

#if defined(__HIPCC__)
#define BMFR_TONE_FN __host__ __device__ __forceinline__
#else
#define BMFR_TONE_FN inline
#endif

namespace bmfr {

// Polynomials are summed as acc += c_k * x^k from the smallest term up, with
// the powers formed by multiplication: the coefficient is then a multiplicand
// (an SGPR operand of v_fmac_f64) rather than the addend of a Horner step,
// which the compiler would first copy into VGPRs at every step.
BMFR_TONE_FN float tone_curve(float v) {
    if (!(v > 0.f)) return 0.f;   // max(0, v), NaN -> 0; pow(0, y) = 0
    if (v >= 1.f) return 1.f;     // pow >= 1 (inf included) clamps to 1
    int e = 0;
    double m = __builtin_frexp((double)v, &e);   // m in [0.5, 1)
    if (m < 0.70710678118654752440) {
        m = m * 2.0;                            // exact
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0);      // m - 1, m + 1 exact (24-bit m)
    // ln m = 2s (1 + z/3 + ... + z^7/15), z = s^2 <= 0.0295: remainder < 4e-14
    const double z = s * s;
    const double z2 = z * z, z3 = z2 * z, z4 = z2 * z2, z5 = z4 * z, z6 = z3 * z3, z7 = z4 * z3;
    double p = z7 * (1.0 / 15.0);
    p = __builtin_fma(1.0 / 13.0, z6, p);
    p = __builtin_fma(1.0 / 11.0, z5, p);
    p = __builtin_fma(1.0 / 9.0, z4, p);
    p = __builtin_fma(1.0 / 7.0, z3, p);
    p = __builtin_fma(1.0 / 5.0, z2, p);
    p = __builtin_fma(1.0 / 3.0, z, p);
    p = p + 1.0;
    const double ln_m = (2.0 * s) * p;
    const double y = (double)0.454545f;
    const double y_log2e = y * 1.44269504088896340736;             // y / ln 2
    const double t = __builtin_fma(ln_m, y_log2e, y * (double)e);  // y*e exact
    const double n = __builtin_rint(t);
    // 2^f = exp(g), g = f ln 2, |g| <= 0.347: degree 10, remainder < 3e-13
    // relative to the last term's size, below 1e-14 overall
    const double g = (t - n) * 0.69314718055994530942;
    const double g2 = g * g, g3 = g2 * g, g4 = g2 * g2, g5 = g4 * g, g6 = g3 * g3, g7 = g4 * g3,
                 g8 = g4 * g4, g9 = g8 * g, g10 = g5 * g5;
    double q = g10 * (1.0 / 3628800.0);
    q = __builtin_fma(1.0 / 362880.0, g9, q);
    q = __builtin_fma(1.0 / 40320.0, g8, q);
    q = __builtin_fma(1.0 / 5040.0, g7, q);
    q = __builtin_fma(1.0 / 720.0, g6, q);
    q = __builtin_fma(1.0 / 120.0, g5, q);
    q = __builtin_fma(1.0 / 24.0, g4, q);
    q = __builtin_fma(1.0 / 6.0, g3, q);
    q = __builtin_fma(0.5, g2, q);
    q = q + g;
    q = q + 1.0;
    return (float)__builtin_ldexp(q, (int)n);   // t in (-68, 0]: n fits, result normal
}

}  // namespace bmfr

#undef BMFR_TONE_FN
