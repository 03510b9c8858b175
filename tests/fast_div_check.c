/* Exhaustive check of the explicit kernel's shortened division: for a divisor b and
 * rb = RN(1/b), q = RN(a rb) corrected once by the exact remainder,
 *     q' = fma(fma(-b, q, a), rb, q),
 * must equal the correctly rounded a / b.  Every significand of a is tried at the given
 * binade offsets; the divisors come from the command line (hexadecimal float bits).
 * Prints the number of mismatches; exit status 1 if any. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float from_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t to_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main(int argc, char **argv)
{
    const int exps[] = {0, -80, 79, -40, 40};
    long bad = 0, tried = 0;
    for (int k = 1; k < argc; ++k) {
        const float b = from_bits((uint32_t)strtoul(argv[k], NULL, 16));
        volatile float one = 1.0f;
        const float rb = one / b;
        for (unsigned e = 0; e < sizeof(exps) / sizeof(exps[0]); ++e)
            for (uint32_t m = 0; m < (1u << 23); ++m) {
                const float a = ldexpf(from_bits(0x3f800000u | m), exps[e]);
                for (int sgn = 0; sgn < 2; ++sgn) {
                    const float aa = sgn ? -a : a;
                    const float q = aa * rb;
                    const float q2 = fmaf(fmaf(-b, q, aa), rb, q);
                    const float t = aa / b;
                    bad += to_bits(q2) != to_bits(t);
                    ++tried;
                }
            }
    }
    printf("%ld of %ld differ\n", bad, tried);
    return bad != 0;
}
