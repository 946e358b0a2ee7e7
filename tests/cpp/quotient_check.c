/* The correctly rounded quotients the wave engine takes as Markstein-corrected products
 * (q0 = a * y, q = fma(fma(-q0, b, a), y, q0) with y = RN(1 / b)):
 *   - the weighted view cost tot / wn: wn a whole number of draws 1..32, tot 0 or in
 *     [2^-60, 64];
 *   - the view-sampling exponent x / -0.18: x 0 or in [2^-60, 1].
 * Every float numerator of those ranges is compared with the IEEE quotient when `stride` is 1
 * (about 3.4e10 quotients); a stride s > 1 checks every s-th float.  Exit status 0: all equal.
 * Usage: quotient_check [stride] */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float f_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static long long check(float b, float top, uint32_t stride) {
    volatile float one = 1.0f, vb = b;
    const float y = one / vb;
    long long bad = 0;
    const uint32_t lo = u_of(ldexpf(1.0f, -60)), hi = u_of(top);
    for (uint64_t u = lo; u <= hi; u += stride) {
        const float a = f_of((uint32_t)u);
        const float q0 = a * y;
        const float q = fmaf(fmaf(-q0, b, a), y, q0);
        if (u_of(q) != u_of(a / b)) ++bad;
    }
    const float q0 = 0.0f * y;
    if (u_of(fmaf(fmaf(-q0, b, 0.0f), y, q0)) != u_of(0.0f / b)) ++bad;
    return bad;
}

int main(int argc, char** argv) {
    const uint32_t stride = argc > 1 ? (uint32_t)strtoul(argv[1], NULL, 10) : 1u;
    long long bad = 0;
    for (int d = 1; d <= 32; ++d) {
        const long long n = check((float)d, 64.0f, stride);
        if (n) printf("divisor %d: %lld mismatches\n", d, n);
        bad += n;
    }
    const long long n = check(-0.18f, 1.0f, stride);
    if (n) printf("divisor -0.18: %lld mismatches\n", n);
    bad += n;
    printf("stride %u: %lld mismatches\n", stride, bad);
    return bad != 0;
}
