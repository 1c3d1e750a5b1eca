/* The FMA-corrected division of the thermal kernels (div_fma in
 * geoflac_amd/csrc/gf_kernels.h) against IEEE division on random operands:
 * q0 = x rd, r = fma(-q0, d, x), q = fma(r, rd, q0) with rd = 1/d rounded.
 * Operands cover mantissas with all bits set and the constant divisor 6; exponents
 * stay in +-300 (no overflow or underflow).  Prints the number of mismatches. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t s = 88172645463325252ull;
static uint64_t rnd(void)
{
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
}
static double rd(void)
{
    uint64_t u = ((uint64_t)(1023 + (int)(rnd() % 600) - 300) << 52) | (rnd() & 0xfffffffffffffull);
    double d;
    memcpy(&d, &u, 8);
    return (rnd() & 1) ? -d : d;
}
static double ones(double v)
{
    uint64_t u;
    memcpy(&u, &v, 8);
    u |= 0xfffffffffffffull;
    memcpy(&v, &u, 8);
    return v;
}

int main(int argc, char **argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 10000000;
    long bad = 0;
    for (long k = 0; k < n; k++) {
        double x = rd(), d = fabs(rd());
        if (k % 4 == 1) d = 6.0;
        if (k % 4 == 2) d = ones(d);
        if (k % 4 == 3) x = ones(x);
        const double r = 1.0 / d;
        const double q0 = x * r;
        const double q = fma(fma(-q0, d, x), r, q0);
        const double e = x / d;
        if (memcmp(&q, &e, 8) != 0) bad++;
    }
    printf("%ld\n", bad);
    return 0;
}
