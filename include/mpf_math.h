/*
 * mpf_math.h -- double-precision tan, acos and tanh that give the same bits
 * on the host (gcc/clang x86-64) and on gfx950.
 *
 * Why: the reference evaluates b(theta) with math.acos/math.tan/math.tanh and
 * a(theta) with math.tan(a0) (src/gpu_kernels_3d.py:24-55).  Host libm and the
 * device math library may differ in the last bit, which would break the
 * bit-exact FP64 comparison between the GPU solver and the CPU oracle.  These
 * routines use only +, -, *, / and sqrt (all IEEE-754 correctly rounded on
 * both sides) in a fixed order, so both builds produce identical results as
 * long as floating-point contraction is disabled (-ffp-contract=off).
 *
 * Methods (textbook, no vendor code):
 *   exp/expm1 : Cody-Waite reduction x = k*ln2 + r, |r| <= ln2/2, then the
 *               Taylor series of expm1(r) to r^14 (truncation < 1e-18 rel).
 *   tanh      : tanh(x) = e/(e+2), e = expm1(2x) for small x;
 *               1 - 2/(exp(2x)+1) otherwise; +-1 for |x| > 22.
 *   tan       : reduction by pi/2 with a three-part pi/2, then sin/cos Taylor
 *               series on |r| <= pi/4 to r^21 / r^20.
 *   acos      : acos(c) = 2 asin(sqrt((1-c)/2))      for c >  0.5
 *                       = pi - 2 asin(sqrt((1+c)/2)) for c < -0.5
 *                       = pi/2 - asin(c)             otherwise,
 *               asin by its Maclaurin series on |s| <= 0.5 to s^49.
 * Accuracy is a few ulp (tests/test_math.py checks it against libm).
 */
#ifndef MPF_MATH_H
#define MPF_MATH_H

#include <math.h>

#if defined(__HIPCC__)
#define MPF_HD __host__ __device__
#else
#define MPF_HD
#endif

#ifdef __cplusplus
#define MPF_INLINE static inline
#else
#define MPF_INLINE static inline
#endif

/* 2^k for -1022 <= k <= 1023, built from the exponent bits. */
MPF_HD MPF_INLINE double mpf_pow2i(int k)
{
    union { unsigned long long u; double d; } v;
    v.u = ((unsigned long long)(k + 1023)) << 52;
    return v.d;
}

/* Taylor series of expm1 on |r| <= 0.35 (Horner, fixed order). */
MPF_HD MPF_INLINE double mpf_expm1_small(double r)
{
    double p = 1.1470745597729725e-11;      /* 1/14! */
    p = 1.6059043836821613e-10 + r * p;     /* 1/13! */
    p = 2.08767569878681e-09 + r * p;       /* 1/12! */
    p = 2.505210838544172e-08 + r * p;      /* 1/11! */
    p = 2.755731922398589e-07 + r * p;      /* 1/10! */
    p = 2.7557319223985893e-06 + r * p;     /* 1/9!  */
    p = 2.48015873015873e-05 + r * p;       /* 1/8!  */
    p = 0.0001984126984126984 + r * p;      /* 1/7!  */
    p = 0.001388888888888889 + r * p;       /* 1/6!  */
    p = 0.008333333333333333 + r * p;       /* 1/5!  */
    p = 0.041666666666666664 + r * p;       /* 1/4!  */
    p = 0.16666666666666666 + r * p;        /* 1/3!  */
    p = 0.5 + r * p;                        /* 1/2!  */
    p = 1.0 + r * p;                        /* 1/1!  */
    return r * p;
}

/* exp(x) for |x| < 700. */
MPF_HD MPF_INLINE double mpf_exp(double x)
{
    const double inv_ln2 = 1.4426950408889634;
    const double ln2_hi = 6.93147180369123816490e-01; /* 32 significant bits */
    const double ln2_lo = 1.90821492927058770002e-10; /* ln2 - ln2_hi        */
    double kd = x * inv_ln2 + 0.5;
    int k = (int)kd;
    if ((double)k > kd) k -= 1;                     /* floor */
    double r = (x - (double)k * ln2_hi) - (double)k * ln2_lo;
    double er = 1.0 + mpf_expm1_small(r);
    if (k > 1000) return er * mpf_pow2i(1000) * mpf_pow2i(k - 1000);
    if (k < -1000) return er * mpf_pow2i(-1000) * mpf_pow2i(k + 1000);
    return er * mpf_pow2i(k);
}

MPF_HD MPF_INLINE double mpf_tanh(double x)
{
    double ax = x < 0.0 ? -x : x;
    double t;
    if (!(ax == ax)) return x;                       /* NaN */
    if (ax > 22.0) {
        t = 1.0;
    } else if (ax < 0.17) {
        double e = mpf_expm1_small(2.0 * ax);
        t = e / (e + 2.0);
    } else {
        double e = mpf_exp(2.0 * ax);
        t = 1.0 - 2.0 / (e + 1.0);
    }
    return x < 0.0 ? -t : t;
}

/* sin(r), cos(r) on |r| <= pi/4 by Taylor series. */
MPF_HD MPF_INLINE double mpf_sin_small(double r)
{
    double z = r * r;
    double p = 1.9572941063391263e-20;       /* 1/21! */
    p = -8.22063524662433e-18 + z * p;       /* 1/19! */
    p = 2.8114572543455206e-15 + z * p;      /* 1/17! */
    p = -7.647163731819816e-13 + z * p;      /* 1/15! */
    p = 1.6059043836821613e-10 + z * p;      /* 1/13! */
    p = -2.505210838544172e-08 + z * p;      /* 1/11! */
    p = 2.7557319223985893e-06 + z * p;      /* 1/9!  */
    p = -0.0001984126984126984 + z * p;      /* 1/7!  */
    p = 0.008333333333333333 + z * p;        /* 1/5!  */
    p = -0.16666666666666666 + z * p;        /* 1/3!  */
    return r + r * (z * p);
}

MPF_HD MPF_INLINE double mpf_cos_small(double r)
{
    double z = r * r;
    double p = 4.110317623312165e-19;        /* 1/20! */
    p = -1.5619206968586225e-16 + z * p;     /* 1/18! */
    p = 4.779477332387385e-14 + z * p;       /* 1/16! */
    p = -1.1470745597729725e-11 + z * p;     /* 1/14! */
    p = 2.08767569878681e-09 + z * p;        /* 1/12! */
    p = -2.755731922398589e-07 + z * p;      /* 1/10! */
    p = 2.48015873015873e-05 + z * p;        /* 1/8!  */
    p = -0.001388888888888889 + z * p;       /* 1/6!  */
    p = 0.041666666666666664 + z * p;        /* 1/4!  */
    p = -0.5 + z * p;                        /* 1/2!  */
    return 1.0 + z * p;
}

/* tan(x) for |x| < 1e6. */
MPF_HD MPF_INLINE double mpf_tan(double x)
{
    const double two_over_pi = 6.36619772367581382433e-01;
    const double pio2_1 = 1.57079632673412561417e+00;  /* first 33 bits of pi/2 */
    const double pio2_2 = 6.07710050630396597660e-11;  /* next 33 bits          */
    const double pio2_3 = 2.02226624871116645580e-21;  /* next 33 bits          */
    double kd = x * two_over_pi;
    kd = kd < 0.0 ? kd - 0.5 : kd + 0.5;
    int k = (int)kd;                                    /* round half away */
    double kk = (double)k;
    double r = ((x - kk * pio2_1) - kk * pio2_2) - kk * pio2_3;
    double s = mpf_sin_small(r);
    double c = mpf_cos_small(r);
    if (k & 1) return -c / s;
    return s / c;
}

/* asin(s) for |s| <= 0.5 by its Maclaurin series. */
MPF_HD MPF_INLINE double mpf_asin_small(double s)
{
    double z = s * s;
    double p = 0.002338091892111975;         /* n = 24 */
    p = 0.0024894486782468836 + z * p;
    p = 0.00265787063820729 + z * p;
    p = 0.002846178401108942 + z * p;
    p = 0.0030578216492580306 + z * p;
    p = 0.003297059503473485 + z * p;
    p = 0.0035692053938259347 + z * p;
    p = 0.003880964558837669 + z * p;
    p = 0.004240907093679363 + z * p;
    p = 0.004660143486915096 + z * p;
    p = 0.005153309682319905 + z * p;
    p = 0.005740037670841924 + z * p;
    p = 0.006447210311889649 + z * p;
    p = 0.0073125258735988454 + z * p;
    p = 0.008390335809616815 + z * p;
    p = 0.009761609529194078 + z * p;
    p = 0.011551800896139705 + z * p;
    p = 0.01396484375 + z * p;
    p = 0.017352764423076924 + z * p;
    p = 0.022372159090909092 + z * p;
    p = 0.030381944444444444 + z * p;
    p = 0.044642857142857144 + z * p;
    p = 0.075 + z * p;
    p = 0.16666666666666666 + z * p;         /* n = 1 */
    return s + s * (z * p);
}

MPF_HD MPF_INLINE double mpf_acos(double c)
{
    const double pi_hi = 3.14159265358979311600e+00;
    const double pi_lo = 1.22464679914735317720e-16;
    const double pio2_hi = 1.57079632679489655800e+00;
    const double pio2_lo = 6.12323399573676588613e-17;
    if (!(c == c)) return c;
    if (c > 0.5) {
        double s = sqrt((1.0 - c) * 0.5);
        return 2.0 * mpf_asin_small(s);
    }
    if (c < -0.5) {
        double s = sqrt((1.0 + c) * 0.5);
        return pi_hi - (2.0 * mpf_asin_small(s) - pi_lo);
    }
    return pio2_hi - (mpf_asin_small(c) - pio2_lo);
}

/* Python's max(a, b): the first argument unless the second is larger. */
MPF_HD MPF_INLINE double mpf_pymax(double a, double b)
{
    return (b > a) ? b : a;
}

#endif /* MPF_MATH_H */
