/*
 * Portable FP64 exp / log / pow.
 *
 * The DPD random force must come out bit-identical on the gfx950 device and on
 * the host.  The reference evaluates the dissipative/random weight with
 * fast::pow (DPDPairEvaluatorGeneralWeight.h:242), i.e. the platform pow, whose
 * last bits differ between libm and any GPU math library.  These routines are
 * built only from IEEE-754 operations that are correctly rounded everywhere
 * (+, -, *, /, sqrt, fma, rint, frexp, ldexp), evaluated in a fixed order with
 * no contraction, so gcc on x86-64 and hipcc on gfx950 produce the same bits.
 *
 * Methods (textbook, no vendor code):
 *   exp: Cody & Waite (1980) reduction x = k ln2 + r with ln2 split into a
 *        head with trailing zero bits and a tail, then the Taylor polynomial of
 *        degree 13 for |r| <= ln2/2 evaluated by Horner's rule in fma.
 *   log: x = 2^e m with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
 *        s = (m-1)/(m+1), summed as an odd series in s up to s^23.
 *   pow: exp(y log x) for x > 0; exact special cases for x == 0, y == 0;
 *        y = 1, 1/2, 1/4 as x, sqrt(x), sqrt(sqrt(x)) (the weight exponent s/2
 *        of s = 2, 1, 1/2), each within one ulp of the exact power.
 * Accuracy: within a few ulp of a correctly rounded result for the arguments
 * the DPD weight uses (x in (0,1], y in [0, 4]).
 *
 * The header is valid C99 and HIP C++.  Host compilers must be run with
 * -ffp-contract=off; HIP translation units get the clang pragma below.
 */
#ifndef AZP_PORTABLE_MATH_H
#define AZP_PORTABLE_MATH_H

#include <math.h>

#if defined(__HIPCC__)
#define AZP_PM_FN __host__ __device__ static inline
#else
#define AZP_PM_FN static inline
#endif

#if defined(__clang__)
#define AZP_PM_NOCONTRACT _Pragma("clang fp contract(off)")
#else
#define AZP_PM_NOCONTRACT
#endif

/* ln 2 split: head has 21 trailing zero bits, so k * AZP_LN2_HI is exact for |k| < 2^21. */
#define AZP_LN2_HI 6.93147180369123816490e-01
#define AZP_LN2_LO 1.90821492927058770002e-10
#define AZP_LOG2E 1.44269504088896338700e+00
#define AZP_SQRT_HALF 0.70710678118654752440

AZP_PM_FN double azp_pexp(double x)
{
    AZP_PM_NOCONTRACT
    if (x != x)
        return x;
    if (x > 709.782712893384)
        return HUGE_VAL;
    if (x < -745.1332191019412)
        return 0.0;
    const double k = rint(x * AZP_LOG2E);
    double r = fma(-k, AZP_LN2_HI, x);
    r = fma(-k, AZP_LN2_LO, r);
    /* Taylor coefficients 1/n!, n = 13 .. 0 */
    double p = 1.0 / 6227020800.0;
    p = fma(p, r, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)k);
}

AZP_PM_FN double azp_plog(double x)
{
    AZP_PM_NOCONTRACT
    if (x != x || x < 0.0)
        return NAN;
    if (x == 0.0)
        return -HUGE_VAL;
    if (x == HUGE_VAL)
        return x;
    int e;
    double m = frexp(x, &e); /* m in [0.5, 1) */
    if (m < AZP_SQRT_HALF)
        {
        m = m * 2.0;
        e = e - 1;
        }
    const double f = m - 1.0; /* exact (Sterbenz) */
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 23.0;
    p = fma(p, z, 1.0 / 21.0);
    p = fma(p, z, 1.0 / 19.0);
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    const double two_s = 2.0 * s;
    const double logm = fma(two_s * z, p, two_s);
    const double de = (double)e;
    return fma(de, AZP_LN2_HI, fma(de, AZP_LN2_LO, logm));
}

/* pow for x >= 0 (the only domain the DPD weight uses). */
AZP_PM_FN double azp_ppow(double x, double y)
{
    AZP_PM_NOCONTRACT
    if (y == 0.0)
        return 1.0;
    if (x == 0.0)
        return (y > 0.0) ? 0.0 : HUGE_VAL;
    if (x == 1.0)
        return 1.0;
    if (y == 1.0)
        return x;
    if (y == 0.5)
        return sqrt(x);
    if (y == 0.25)
        return sqrt(sqrt(x));
    return azp_pexp(y * azp_plog(x));
}

#endif /* AZP_PORTABLE_MATH_H */
