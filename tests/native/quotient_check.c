// Host check of the shared-divisor quotient of bmfr_amd/csrc/device_common.h
// (quotient_reciprocal, quotient): (float)((double)a * y), y from a
// first reciprocal approximation (v_rcp_f64 on the device; here 1/d perturbed
// by up to 2^-14 relative) and two Newton steps, against IEEE float division
// bit for bit: random bit patterns, quotients next to float rounding
// boundaries, integer divisors and special values.  Prints one line.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t rnd(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static float bits2f(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static uint32_t f2bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }
static double recip(float d, double eps) {
    const double dd = d;
    const double y0 = (1.0 / dd) * (1.0 + eps);
    double e = fma(-dd, y0, 1.0), y = fma(y0, e, y0);
    e = fma(-dd, y, 1.0); y = fma(y, e, y);
    double raw = 1.0 / dd;   // what v_rcp gives for 0, inf, NaN
    return (isfinite(dd) && dd != 0.0) ? y : raw;
}
static long long bad = 0, n = 0;
static void one(float a, float d, double eps) {
    float want = a / d;
    float got = (float)((double)a * recip(d, eps));
    n++;
    if (f2bits(want) != f2bits(got) && !(isnan(want) && isnan(got))) {
        if (bad < 10) printf("BAD a=%a d=%a want=%a got=%a\n", a, d, want, got);
        bad++;
    }
}
int main(void) {
    const double eps_list[] = {0.0, 0x1p-20, -0x1p-20, 0x1p-14, -0x1p-14};
    for (int k = 0; k < 5; ++k) {
        double eps = eps_list[k];
        // random bit patterns (finite, both signs, incl. subnormals)
        for (long i = 0; i < 20000000; ++i) {
            float a = bits2f((uint32_t)rnd()), d = bits2f((uint32_t)rnd());
            one(a, d, eps);
        }
        // near-midpoint quotients: a = RN(m * d) for m a float midpoint
        for (long i = 0; i < 20000000; ++i) {
            float d = bits2f(0x3f800000u | ((uint32_t)rnd() & 0x7fffff)) * ldexpf(1.f, (int)(rnd() % 60) - 30);
            uint32_t q = 0x3f800000u | ((uint32_t)rnd() & 0x7fffff);
            double m = ((double)bits2f(q) + (double)bits2f(q + 1)) / 2.0 * ldexp(1.0, (int)(rnd() % 40) - 20);
            float a = (float)(m * d);
            one(a, d, eps);
            one(nextafterf(a, 0.f), d, eps);
            one(nextafterf(a, 1e30f), d, eps);
        }
        // integer divisors and small weights like the stages use
        for (int sp = 1; sp < 100000; ++sp) one(1.f, (float)sp, eps);
        const float sp_[] = {0.f, -0.f, INFINITY, -INFINITY, NAN, 1e-45f, 3e38f};
        for (int i = 0; i < 7; ++i)
            for (int j = 0; j < 7; ++j) one(sp_[i], sp_[j], eps), one(1.5f, sp_[j], eps), one(sp_[i], 3.f, eps);
    }
    printf("checked %lld bad %lld\n", n, bad);
    return bad != 0;
}
