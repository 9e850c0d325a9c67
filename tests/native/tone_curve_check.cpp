// Exhaustive host check of bmfr::tone_curve (bmfr_amd/csrc/tone_curve.h):
// every float in [0, 1] plus special and out-of-range inputs, against
// (a) the power in double precision, pow((double)v, (double)0.454545f),
//     rounded once to float (correctly rounded up to that double rounding),
// (b) the C library's powf, the oracle's tone map (oracle/bmfr_oracle.c).
// Prints one line: checked, max ulp and mismatch counts against each.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "tone_curve.h"

static float clamp01(float x) { return std::fmin(std::fmax(x, 0.f), 1.f); }
static int32_t bits(float f) { int32_t i; std::memcpy(&i, &f, 4); return i; }

int main() {
    const float y = 0.454545f;
    long long n = 0, diff_dbl = 0, diff_powf = 0;
    long long max_dbl = 0, max_powf = 0;
#pragma omp parallel for schedule(static, 1 << 16) reduction(+ : n, diff_dbl, diff_powf) \
    reduction(max : max_dbl, max_powf)
    for (int64_t i = 0; i <= 0x3F800000LL; ++i) {
        float v;
        const int32_t b = (int32_t)i;
        std::memcpy(&v, &b, 4);
        const float got = bmfr::tone_curve(v);
        const float ref_d = clamp01((float)std::pow((double)std::fmax(0.f, v), (double)y));
        const float ref_f = clamp01(std::pow(std::fmax(0.f, v), y));
        const long long dd = llabs((long long)bits(got) - bits(ref_d));
        const long long df = llabs((long long)bits(got) - bits(ref_f));
        n += 1;
        diff_dbl += dd != 0;
        diff_powf += df != 0;
        if (dd > max_dbl) max_dbl = dd;
        if (df > max_powf) max_powf = df;
    }
    // special and out-of-range inputs: clamps and NaN -> 0
    const float sp[] = {-0.f, -1.f, -INFINITY, NAN, 1.f, 1.0000001f, 2.f, 3.0e38f, INFINITY};
    const float want[] = {0.f, 0.f, 0.f, 0.f, 1.f, 1.f, 1.f, 1.f, 1.f};
    int special_bad = 0;
    for (int k = 0; k < 9; ++k)
        if (bits(bmfr::tone_curve(sp[k])) != bits(want[k])) ++special_bad;
    std::printf("checked %lld max_ulp_vs_double_pow %lld mismatches_vs_double_pow %lld "
                "max_ulp_vs_powf %lld mismatches_vs_powf %lld special_bad %d\n",
                n, max_dbl, diff_dbl, max_powf, diff_powf, special_bad);
    return 0;
}
