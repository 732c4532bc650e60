/*
 * tsu_quotient.c -- test support for the GPU path's shared-denominator
 * quotients (TEST INFRASTRUCTURE ONLY; not a restatement of the reference).
 *
 * The FP64 solver divides several numerators by one denominator d (velocity
 * and concentration by h + DBL_EPSILON, the HLL coefficients by S_R - S_L,
 * both momenta by 1 + c) with one IEEE reciprocal y = 1/d and, per
 * numerator x, q = x*y followed by q + (x - d*q)*y with both products fused.
 * Markstein's theorem says this is the correctly rounded x/d when y is the
 * correctly rounded reciprocal, q is within an ulp and the residual does not
 * underflow.  (The device computes y itself from the hardware reciprocal
 * estimate; that y equals 1.0 / d, all-ones significands included, is
 * checked on the device against numpy.)  orc_shared_quotient_mismatches() counts, over n seeded random
 * pairs spread over the ranges the solver meets (and near-exact quotients),
 * the pairs where that value differs from x / d.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

static uint64_t q_state;

static uint64_t q_next(void)
{
    q_state ^= q_state << 13;
    q_state ^= q_state >> 7;
    q_state ^= q_state << 17;
    return q_state;
}

/* random double with exponent in [emin, emax] and a random sign */
static double q_rand(int emin, int emax)
{
    uint64_t bits = ((uint64_t)(emin + (int)(q_next() % (uint64_t)(emax - emin + 1)) + 1023) << 52)
                    | (q_next() & ((1ull << 52) - 1));
    double v;
    memcpy(&v, &bits, sizeof v);
    return (q_next() & 1) ? -v : v;
}

long orc_shared_quotient_mismatches(long n, unsigned long long seed)
{
    long bad = 0, i;
    q_state = seed ? seed : 88172645463325252ull;
    for (i = 0; i < n; i++) {
        double d = fabs(q_rand(-30, 30)), x = q_rand(-60, 40);
        if (i % 3 == 0) d = fabs(q_rand(-20, 14)) + 2.2204460492503131e-16;   /* h + DBL_EPSILON */
        if (i % 5 == 0) d = 1.0 + fabs(q_rand(-50, 2));                        /* 1 + c */
        if (i % 7 == 0) x = q_rand(-60, 40) * d;                               /* near-exact quotients */
        if (i % 11 == 0) x = nextafter(d * (double)(q_next() % 1000), 0.0);
        {
            double y = 1.0 / d;
            double q = x * y;
            double r = fma(-d, q, x);
            if (fma(r, y, q) != x / d) bad++;
        }
    }
    return bad;
}


/* The same quotient for tiny numerators: |x| from the smallest subnormal up
 * to 2^-900 (exponent fields 0..123), d over the solver's denominators.
 * guarded = 0 counts where q differs from x / d (the residual underflows);
 * guarded = 1 takes x / d whenever x is nonzero and below 2^-968, as the
 * device does, and counts the remaining differences. */
long orc_shared_quotient_tiny_mismatches(long n, unsigned long long seed, int guarded)
{
    long bad = 0, i;
    q_state = seed ? seed : 88172645463325252ull;
    for (i = 0; i < n; i++) {
        double d = fabs(q_rand(-30, 30)), x, y, q, r, got;
        uint64_t bits = ((q_next() % 124) << 52) | (q_next() & ((1ull << 52) - 1)) | (q_next() & (1ull << 63));
        memcpy(&x, &bits, sizeof x);
        if (i % 3 == 0) d = fabs(q_rand(-20, 14)) + 2.2204460492503131e-16;
        if (i % 5 == 0) d = 1.0 + fabs(q_rand(-50, 2));
        y = 1.0 / d;
        q = x * y;
        r = fma(-d, q, x);
        got = fma(r, y, q);
        if (guarded && x != 0.0 && fabs(x) < 0x1p-968) got = x / d;
        if (got != x / d) bad++;
    }
    return bad;
}
