/*
 * spac_math.h — FP64 elementary functions, Philox4x32-10 and Box–Muller normals,
 * written once and compiled both as C (gcc, CPU) and as HIP (device, gfx950).
 *
 * Why this file exists: the reference draws its Brownian increments from
 * boost::mt19937 + normal_distribution and evaluates exp/pow through glibc
 * (Solvers/FeynmanKacSolver.cpp:2-6, BVPs/Franke_Poisson.hpp:4-17).  Neither
 * stream nor libm bits can be reproduced on a GPU, so this project defines its
 * own: every function here uses only IEEE-754 operations that are correctly
 * rounded on both targets (+, -, *, /, fma, sqrt, rint) and plain integer
 * arithmetic, with floating-point contraction disabled at build time
 * (-ffp-contract=off).  Given the same inputs a CPU build and a GPU build
 * therefore produce the same bits, which is what lets per-path exit data be
 * compared exactly between the HIP kernels and the CPU restatement.
 *
 * Methods (all from the mathematics).  The polynomials are Chebyshev fits of
 * the remainder after each function's exact leading terms, the shortest whose
 * approximation error is below 0.02 ulp (derive_coefficients.py, mpmath at 60
 * digits, printed as the literals below); atan's series are Taylor, with
 * coefficients rounded from exact rationals:
 *   exp   : Cody–Waite reduction x = k ln2 + r with a 32-bit ln2 head,
 *           e^r = 1 + r + r^2 Q(r) on |r| <= ln2/2 (Q of degree 10), scaling by 2^k.
 *   log   : x = 2^e m, m in [sqrt(2)/2, sqrt(2)), f = (m-1)/(m+1),
 *           log m = 2f + f^3 R(f^2) (R of degree 7).
 *   sincospi : exact reduction a = q/2 + r, |r| <= 1/4, sin(pi r) = pi r + r^3 S(r^2)
 *           and cos(pi r) = 1 + r^2 C(r^2) (degree 6 each), quadrant rotation by q mod 4.
 *   sinpi, cospi : exact reduction a = n + r, |r| <= 1/2, one polynomial each,
 *           sin(pi t) = pi t + t^3 P(t^2) (P of degree 7; cos(pi r) = sin(pi (1/2 - |r|))).
 *   atan2 : octant reduction to t in [0,1], t > tan(pi/8) mapped by
 *           atan t = pi/4 + atan((t-1)/(t+1)), one half-angle step
 *           atan u = 2 atan(u / (1 + sqrt(1+u^2))), odd Taylor series.
 *   Philox4x32-10 : Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 *           easy as 1, 2, 3", SC'11 (multipliers, Weyl key increments, 10 rounds).
 *   Box–Muller : two 32-bit uniforms (two words of a Philox block) give two N(0,1); one
 *           block serves two path-steps.
 */
#ifndef SPAC_MATH_H
#define SPAC_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SPAC_HD __host__ __device__ __forceinline__
#define SPAC_RINT(x) __builtin_rint(x)
#define SPAC_FMA(a, b, c) __builtin_fma((a), (b), (c))
#define SPAC_SQRT(x) __builtin_sqrt(x)
#define SPAC_FABS(x) __builtin_fabs(x)
#define SPAC_COPYSIGN(a, b) __builtin_copysign((a), (b))
#else
#include <math.h>
#define SPAC_HD static inline
#define SPAC_RINT(x) rint(x)
#define SPAC_FMA(a, b, c) fma((a), (b), (c))
#define SPAC_SQRT(x) sqrt(x)
#define SPAC_FABS(x) fabs(x)
#define SPAC_COPYSIGN(a, b) copysign((a), (b))
#endif

#define SPAC_PI 3.141592653589793
#define SPAC_PI_2 1.5707963267948966
#define SPAC_PI_4 0.7853981633974483

/* SPAC_FMAK(p, x, c): a Horner step p x + c whose constant c the compiler keeps in a VGPR across
 * the path loop (the log and sincospi coefficients: the loop's scalar registers are full).  With
 * the builtin it picks the two-address v_fmac_f64 and copies c into the destination first, one
 * extra v_mov_b64 per step; the three-address v_fma_f64 written directly reads c in place.  Same
 * operation, same bits. */
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ double spac_fmak(double a, double b, double c) {
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
#define SPAC_FMAK(a, b, c) spac_fmak((a), (b), (c))
#else
#define SPAC_FMAK(a, b, c) SPAC_FMA(a, b, c)
#endif

/* ---------------------------------------------------------------- bits -- */
SPAC_HD double spac_u64_as_f64(uint64_t u) {
    union { uint64_t u; double d; } v;
    v.u = u;
    return v.d;
}
SPAC_HD uint64_t spac_f64_as_u64(double d) {
    union { uint64_t u; double d; } v;
    v.d = d;
    return v.u;
}
/* 2^k for k in [-1022, 1023] */
SPAC_HD double spac_pow2i(int k) {
    return spac_u64_as_f64((uint64_t)(k + 1023) << 52);
}

/* ----------------------------------------------------------------- exp -- */
/* 2^k p for p in [0.7, 1.5] and k in [-1076, 1025]: correctly rounded, also into the subnormal
 * range (one rounding). */
SPAC_HD double spac_scale2(double p, int k) {
#if defined(__HIPCC__)
    return __builtin_ldexp(p, k);
#else
    return ldexp(p, k);
#endif
}

/* e^x for x < 709.78 (finite results): spac_exp without its NaN and overflow selects, the same
 * bits for every such x (the same reduction and polynomial; 2^k applied in one correctly rounded
 * step instead of two exact-then-rounded products, which round to the same value).  Larger x
 * give +inf, as from spac_exp. */
SPAC_HD double spac_exp_finite(double x) {
    const double LOG2E = 1.4426950408889634;
    const double LN2_HI = 0.6931471803691238;     /* 32 significant bits */
    const double LN2_LO = 1.9082149292705877e-10; /* ln2 - LN2_HI */
    const double xc = (x < -746.0) ? -746.0 : ((x > 710.0) ? 710.0 : x);
    const double kd = SPAC_RINT(xc * LOG2E);
    double r = SPAC_FMA(-kd, LN2_HI, xc);
    r = SPAC_FMA(-kd, LN2_LO, r);
    double p = 2.0914679452495193e-09;
    p = SPAC_FMA(p, r, 2.5105206480259365e-08);
    p = SPAC_FMA(p, r, 2.755727366116628e-07);
    p = SPAC_FMA(p, r, 2.755725542549107e-06);
    p = SPAC_FMA(p, r, 2.480158732553351e-05);
    p = SPAC_FMA(p, r, 0.00019841269874800694);
    p = SPAC_FMA(p, r, 0.0013888888888883752);
    p = SPAC_FMA(p, r, 0.008333333333326141);
    p = SPAC_FMA(p, r, 0.04166666666666667);
    p = SPAC_FMA(p, r, 0.1666666666666667);
    p = SPAC_FMA(p, r, 0.5);
    p = SPAC_FMA(p, r, 1.0);
    p = SPAC_FMA(p, r, 1.0);
    const double res = spac_scale2(p, (int)kd);
    return (x < -745.1332191019412) ? 0.0 : res;
}

SPAC_HD double spac_exp(double x) {
    const double LOG2E = 1.4426950408889634;
    const double LN2_HI = 0.6931471803691238;     /* 32 significant bits */
    const double LN2_LO = 1.9082149292705877e-10; /* ln2 - LN2_HI */
    /* clamp keeps k in [-1076, 1025]; out-of-range inputs are fixed up at the end */
    double xc = (x < -746.0) ? -746.0 : ((x > 710.0) ? 710.0 : x);
    double kd = SPAC_RINT(xc * LOG2E);
    double r = SPAC_FMA(-kd, LN2_HI, xc); /* exact: |k| <= 1076 needs 11 bits */
    r = SPAC_FMA(-kd, LN2_LO, r);
    /* e^r = 1 + r + r^2 Q(r), |r| <= 0.3466, Q of degree 10 */
    double p = 2.0914679452495193e-09;
    p = SPAC_FMA(p, r, 2.5105206480259365e-08);
    p = SPAC_FMA(p, r, 2.755727366116628e-07);
    p = SPAC_FMA(p, r, 2.755725542549107e-06);
    p = SPAC_FMA(p, r, 2.480158732553351e-05);
    p = SPAC_FMA(p, r, 0.00019841269874800694);
    p = SPAC_FMA(p, r, 0.0013888888888883752);
    p = SPAC_FMA(p, r, 0.008333333333326141);
    p = SPAC_FMA(p, r, 0.04166666666666667);
    p = SPAC_FMA(p, r, 0.1666666666666667);
    p = SPAC_FMA(p, r, 0.5);
    p = SPAC_FMA(p, r, 1.0);
    p = SPAC_FMA(p, r, 1.0);
    /* p 2^k as (p 2^k1) 2^k2, k1 = floor(k/2): the first product is exact, the
     * second rounds once, also into the subnormal range */
    int k = (int)kd;
    int k1 = k >> 1;
    double res = (p * spac_pow2i(k1)) * spac_pow2i(k - k1);
    /* special cases as selects, not branches, so that a caller's independent
     * exponentials can be interleaved by the GPU scheduler */
    double big = (x != x) ? x : spac_u64_as_f64(0x7ff0000000000000ULL);
    res = (x < 709.782712893384) ? res : big;
    return (x < -745.1332191019412) ? 0.0 : res;
}

/* ----------------------------------------------------------------- log -- */
/* Natural log for normal x > 0 (x >= 2^-1022, finite): spac_log without its subnormal, zero and
 * NaN cases, the same bits for every such x (the uniforms of the normal and exponential draws
 * are at least 2^-53). */
SPAC_HD double spac_log_normal(double x) {
    const double LN2_HI = 0.6931471803691238;
    const double LN2_LO = 1.9082149292705877e-10;
    uint64_t bits = spac_f64_as_u64(x);
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    double m = spac_u64_as_f64((bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL); /* [1,2) */
    const int hi = m > 1.4142135623730951;
    m = hi ? m * 0.5 : m;
    e += hi;
    double f = (m - 1.0) / (m + 1.0);
    double s = f * f;
    double p = 0.13086626149248715;
    p = SPAC_FMAK(p, s, 0.1326877313852219);
    p = SPAC_FMAK(p, s, 0.15386239702819687);
    p = SPAC_FMAK(p, s, 0.18181795640132814);
    p = SPAC_FMAK(p, s, 0.2222222239171392);
    p = SPAC_FMAK(p, s, 0.28571428570803614);
    p = SPAC_FMAK(p, s, 0.4000000000000088);
    p = SPAC_FMAK(p, s, 0.6666666666666666);
    double logm = SPAC_FMA(f * s, p, 2.0 * f);
    double ed = (double)e;
    return SPAC_FMA(ed, LN2_HI, SPAC_FMA(ed, LN2_LO, logm));
}

/* Natural log for finite x > 0 (the only domain the solver uses). */
SPAC_HD double spac_log(double x) {
    const double LN2_HI = 0.6931471803691238;
    const double LN2_LO = 1.9082149292705877e-10;
    /* straight-line code (selects, no branches); x <= 0 and NaN are fixed up at the end */
    const int sub = x < 2.2250738585072014e-308;
    const double xs = sub ? x * 18014398509481984.0 : x; /* 2^54 */
    uint64_t bits = spac_f64_as_u64(xs);
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    double m = spac_u64_as_f64((bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL); /* [1,2) */
    const int hi = m > 1.4142135623730951;
    m = hi ? m * 0.5 : m;
    e += hi + (sub ? -54 : 0);
    double f = (m - 1.0) / (m + 1.0);
    double s = f * f;
    /* log m = 2f + f s R(s), R of degree 7 */
    double p = 0.13086626149248715;
    p = SPAC_FMA(p, s, 0.1326877313852219);
    p = SPAC_FMA(p, s, 0.15386239702819687);
    p = SPAC_FMA(p, s, 0.18181795640132814);
    p = SPAC_FMA(p, s, 0.2222222239171392);
    p = SPAC_FMA(p, s, 0.28571428570803614);
    p = SPAC_FMA(p, s, 0.4000000000000088);
    p = SPAC_FMA(p, s, 0.6666666666666666);
    double logm = SPAC_FMA(f * s, p, 2.0 * f);
    double ed = (double)e;
    double res = SPAC_FMA(ed, LN2_HI, SPAC_FMA(ed, LN2_LO, logm));
    double bad = (x == 0.0) ? -spac_u64_as_f64(0x7ff0000000000000ULL) : spac_u64_as_f64(0x7ff8000000000000ULL);
    return (x > 0.0) ? res : bad;
}

/* ------------------------------------------------------------ sincospi -- */
/* sin(pi a), cos(pi a) for |a| < 2^51. */
SPAC_HD void spac_sincospi(double a, double* s_out, double* c_out) {
    double q = SPAC_RINT(a * 2.0);
    double r = a - 0.5 * q; /* exact, |r| <= 1/4 */
    double r2 = r * r;
    double sp = -2.1742067342478005e-05;
    sp = SPAC_FMAK(sp, r2, 0.0004662879326147481);
    sp = SPAC_FMAK(sp, r2, -0.0073704303122948235);
    sp = SPAC_FMAK(sp, r2, 0.08214588659727676);
    sp = SPAC_FMAK(sp, r2, -0.5992645293206478);
    sp = SPAC_FMAK(sp, r2, 2.550164039877345);
    sp = SPAC_FMAK(sp, r2, -5.16771278004997);
    double sr = SPAC_FMA(r * r2, sp, r * 3.141592653589793);
    double cp = -0.00010370082971594627;
    cp = SPAC_FMAK(cp, r2, 0.001929493874894941);
    cp = SPAC_FMAK(cp, r2, -0.025806887964701878);
    cp = SPAC_FMAK(cp, r2, 0.23533063028399223);
    cp = SPAC_FMAK(cp, r2, -1.3352627688538095);
    cp = SPAC_FMAK(cp, r2, 4.058712126416765);
    cp = SPAC_FMAK(cp, r2, -4.934802200544679);
    double cr = SPAC_FMA(cp, r2, 1.0);
    int64_t qi = (int64_t)q;
    int quad = (int)(qi & 3);
    /* quadrant rotation: q = 0 (sr, cr), 1 (cr, -sr), 2 (-sr, -cr), 3 (-cr, sr) */
    double s = (quad & 1) ? cr : sr;
    double c = (quad & 1) ? -sr : cr;
    s = (quad & 2) ? -s : s;
    c = (quad & 2) ? -c : c;
    *s_out = s;
    *c_out = c;
}
/* sin(pi t) for |t| <= 1/2: pi t + t^3 P(t^2), P of degree 7 (approximation error 0.012 ulp). */
SPAC_HD double spac_sinpi_half(double t) {
    const double t2 = t * t;
    double p = 7.725743030789876e-07;
    p = SPAC_FMA(p, t2, -2.1906201655131958e-05);
    p = SPAC_FMA(p, t2, 0.00046630087411496363);
    p = SPAC_FMA(p, t2, -0.007370430719634586);
    p = SPAC_FMA(p, t2, 0.08214588659675232);
    p = SPAC_FMA(p, t2, -0.599264529320343);
    p = SPAC_FMA(p, t2, 2.55016403987734);
    p = SPAC_FMA(p, t2, -5.16771278004997);
    /* pi t as hi + lo exactly (an FMA gives the product's rounding error), plus pi's own tail:
       near |t| = 1/2 the rounding of pi t alone is an ulp of the result */
    const double hi = t * 3.141592653589793;
    const double lo = SPAC_FMA(t, 3.141592653589793, -hi);
    const double s = hi + SPAC_FMA(t * t2, p, SPAC_FMA(t, 1.2246467991473532e-16, lo));
    return (SPAC_FABS(t) == 0.5) ? SPAC_COPYSIGN(1.0, t) : s; /* sin(pi / 2) = 1 exactly */
}
/* sin(pi a), cos(pi a) for |a| < 2^51, one polynomial each: a = n + r with n = rint(a),
 * |r| <= 1/2 (exact); sin(pi a) = (-1)^n sin(pi r) and cos(pi a) = (-1)^n sin(pi (1/2 - |r|)). */
/* n odd, for an integer-valued n with |n| < 2^53 */
SPAC_HD int spac_odd(double n) {
    const double h = 0.5 * n;
    return h != SPAC_RINT(h);
}
SPAC_HD double spac_sinpi(double a) {
    const double n = SPAC_RINT(a);
    const double s = spac_sinpi_half(a - n);
    return spac_odd(n) ? -s : s;
}
SPAC_HD double spac_cospi(double a) {
    const double n = SPAC_RINT(a);
    const double s = spac_sinpi_half(0.5 - SPAC_FABS(a - n));
    return spac_odd(n) ? -s : s;
}

/* --------------------------------------------------------------- atan2 -- */
/* atan(u) for |u| <= 0.4143 */
SPAC_HD double spac_atan_small(double u) {
    /* half-angle step: atan u = 2 atan(v), v = u / (1 + sqrt(1 + u^2)), |v| <= 0.199 */
    double v = u / (1.0 + SPAC_SQRT(SPAC_FMA(u, u, 1.0)));
    double v2 = v * v;
    double p = -0.043478260869565216;
    p = SPAC_FMA(p, v2, 0.047619047619047616);
    p = SPAC_FMA(p, v2, -0.05263157894736842);
    p = SPAC_FMA(p, v2, 0.058823529411764705);
    p = SPAC_FMA(p, v2, -0.06666666666666667);
    p = SPAC_FMA(p, v2, 0.07692307692307693);
    p = SPAC_FMA(p, v2, -0.09090909090909091);
    p = SPAC_FMA(p, v2, 0.1111111111111111);
    p = SPAC_FMA(p, v2, -0.14285714285714285);
    p = SPAC_FMA(p, v2, 0.2);
    p = SPAC_FMA(p, v2, -0.3333333333333333);
    double at = SPAC_FMA(v * v2, p, v);
    return 2.0 * at;
}
SPAC_HD double spac_atan2(double y, double x) {
    double ax = SPAC_FABS(x), ay = SPAC_FABS(y);
    if (ax == 0.0 && ay == 0.0) {
        double base = (spac_f64_as_u64(x) >> 63) ? SPAC_PI : 0.0;
        return SPAC_COPYSIGN(base, y);
    }
    int swap = ay > ax;
    double t = swap ? (ax / ay) : (ay / ax); /* [0, 1] */
    double a;
    if (t > 0.41421356237309503) {
        a = SPAC_PI_4 + spac_atan_small((t - 1.0) / (t + 1.0));
    } else {
        a = spac_atan_small(t);
    }
    if (swap) a = SPAC_PI_2 - a;
    if (spac_f64_as_u64(x) >> 63) a = SPAC_PI - a;
    return SPAC_COPYSIGN(a, y);
}

/* -------------------------------------------------------------- Philox -- */
typedef struct { uint32_t v[4]; } spac_u32x4;

SPAC_HD void spac_mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

/* Philox4x32 with R rounds (R = 10 is the standard generator). */
SPAC_HD spac_u32x4 spac_philox4x32_R(int R, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                      uint32_t k0, uint32_t k1) {
    for (int r = 0; r < R; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0, lo0, hi1, lo1;
        spac_mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
        spac_mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    spac_u32x4 out;
    out.v[0] = c0; out.v[1] = c1; out.v[2] = c2; out.v[3] = c3;
    return out;
}
SPAC_HD spac_u32x4 spac_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       uint32_t k0, uint32_t k1) {
    return spac_philox4x32_R(10, c0, c1, c2, c3, k0, k1);
}

/* 53-bit uniform in [0, 1) from two 32-bit words (a supplies the top 27 bits). */
SPAC_HD double spac_u01_53(uint32_t a, uint32_t b) {
    uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return (double)m * 1.1102230246251565e-16; /* 2^-53 */
}

/*
 * Stream layout shared by every kernel and the CPU restatement:
 *   key     = (low 32 bits of the seed, stream id = solve/node index)
 *   counter = (step pair = step / 2, path index, high 32 bits of the seed, purpose tag)
 * One Philox block per two path-steps: words (v0, v1) give the two Brownian normals of the
 * even step, (v2, v3) those of the odd step, each pair by Box–Muller from two 32-bit uniforms
 * (the resolution of the reference's 32-bit mt19937 draws).  The path kernels keep the odd
 * step's two words in registers, so a path-step costs half a Philox block.
 */
#define SPAC_TAG_BROWNIAN 0u
#define SPAC_TAG_WORKLOAD 7u

SPAC_HD spac_u32x4 spac_brownian_block(uint64_t seed, uint32_t stream, uint32_t path, uint32_t pair) {
    return spac_philox4x32_10(pair, path, (uint32_t)(seed >> 32), SPAC_TAG_BROWNIAN, (uint32_t)seed, stream);
}
/* Box–Muller from two 32-bit words: u1 = 1 - a 2^-32 in (0, 1], u2 = b 2^-32 in [0, 1) (both
 * exact in double). */
SPAC_HD void spac_normal_words(uint32_t a, uint32_t b, double* n0, double* n1) {
    const double u1 = 1.0 - (double)a * 2.3283064365386963e-10;
    const double u2 = (double)b * 2.3283064365386963e-10;
    const double rad = SPAC_SQRT(-2.0 * spac_log_normal(u1));
    double s, c;
    spac_sincospi(2.0 * u2, &s, &c);
    *n0 = rad * c;
    *n1 = rad * s;
}
/* the normals of one path-step */
SPAC_HD void spac_normal_pair(uint64_t seed, uint32_t stream, uint32_t path, uint32_t step,
                              double* n0, double* n1) {
    const spac_u32x4 r = spac_brownian_block(seed, stream, path, step >> 1);
    if (step & 1u) spac_normal_words(r.v[2], r.v[3], n0, n1);
    else spac_normal_words(r.v[0], r.v[1], n0, n1);
}

/* Unit exponential E = -log U, U in (0, 1], for the reflecting step's bridge
 * maximum: its own Philox block per path-step (tag SPAC_TAG_REFLECT), so a path
 * that never reflects draws the same normals as under absorbing sides. */
#define SPAC_TAG_REFLECT 1u

SPAC_HD double spac_exponential(uint64_t seed, uint32_t stream, uint32_t path, uint32_t step) {
    spac_u32x4 r = spac_philox4x32_10(step, path, (uint32_t)(seed >> 32), SPAC_TAG_REFLECT,
                                      (uint32_t)seed, stream);
    double u = 1.0 - spac_u01_53(r.v[0], r.v[1]); /* (0, 1] */
    return -spac_log_normal(u);
}

/* ---------------------------------------------------------- stop test -- */
/* The Gobet–Menozzi stop test of Dentro (Solvers/FeynmanKacSolver.cpp:84,96), "inside" when
 * d < g |N^T sigma| sqrt(h), from k = g sqrt(h) and nrm2 = |N^T sigma|^2, compared in squares
 * (no square root).  With k <= 0 (the reference's g = -0.5826) both sides of d < k |.| are
 * non-positive: d < 0 and d^2 > k^2 nrm2.  With k > 0 the right side is non-negative: d < 0, or
 * d >= 0 and d^2 < k^2 nrm2. */
SPAC_HD int spac_gobet_inside(double d, double k, double nrm2) {
    const double d2 = d * d, r2 = (k * k) * nrm2;
    return (k <= 0.0) ? (d < 0.0 && d2 > r2) : (d < 0.0 || d2 < r2);
}

/* -------------------------------------------- compensated accumulation -- */
/* Knuth TwoSum: s + e == a + b exactly. */
typedef struct { double s, c; } spac_ksum;

SPAC_HD spac_ksum spac_ksum_zero(void) { spac_ksum k; k.s = 0.0; k.c = 0.0; return k; }
SPAC_HD spac_ksum spac_ksum_add(spac_ksum a, double x) {
    double t = a.s + x;
    double bp = t - a.s;
    double err = (a.s - (t - bp)) + (x - bp);
    spac_ksum r;
    r.s = t;
    r.c = a.c + err;
    return r;
}
SPAC_HD spac_ksum spac_ksum_merge(spac_ksum a, spac_ksum b) {
    double t = a.s + b.s;
    double bp = t - a.s;
    double err = (a.s - (t - bp)) + (b.s - bp);
    spac_ksum r;
    r.s = t;
    r.c = (a.c + b.c) + err;
    return r;
}
SPAC_HD double spac_ksum_value(spac_ksum a) { return a.s + a.c; }

#endif /* SPAC_MATH_H */
