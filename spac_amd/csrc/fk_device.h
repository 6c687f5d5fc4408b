/*
 * fk_device.h — device twins of the plugin surface used on the path kernel.
 *
 * The reference evaluates every coefficient through a host function pointer
 * once per use: sigma, f, psi, F, c, mu (twice), varphi and b per step and
 * sigma once more per boundary test (Solvers/FeynmanKacSolver.cpp:7-18,
 * 68-106).  A GPU cannot call host pointers, so each registered coefficient
 * set has a fused device functor here: the shared Franke Gaussians are
 * evaluated once per step and every derived quantity is formed with the same
 * IEEE operations, in the same order, as the per-plugin host functions
 * (BVPs/Franke_Poisson.hpp:4-73).  Terms that are identically zero for a set
 * (psi, varphi, mu; b and c for Franke) are dropped: adding an exact zero does
 * not change a finite IEEE value, so the path bits are unchanged.
 */
#ifndef SPAC_FK_DEVICE_H
#define SPAC_FK_DEVICE_H

#include "spac_cv.h"
#include "spac_math.h"
#include "spac_tab.h"

namespace spac {

/* Per-step coefficient values at X: sigma = s I, drift b, killing c, source f,
 * control-variate field F = -sigma^T grad u. */
struct StepCoef {
    double s;
    double bx, by;
    double c;
    double f;
    double Fx, Fy;
};

struct FrankeTerms {
    double a1, b1, a2, a3, b3, a4, b4;
    double s1, q2, s3, s4; /* the squared distances in the exponents */
    double F1, F2, F3, F4;
};

__device__ __forceinline__ FrankeTerms franke_terms(double x, double y) {
    FrankeTerms t;
    t.a1 = 9.0 * x - 2.0; t.b1 = 9.0 * y - 2.0;
    t.a2 = 9.0 * x + 1.0;
    t.a3 = 9.0 * x - 7.0; t.b3 = 9.0 * y - 3.0;
    t.a4 = 9.0 * x - 4.0; t.b4 = 9.0 * y - 7.0;
    t.s1 = t.a1 * t.a1 + t.b1 * t.b1;
    t.q2 = t.a2 * t.a2;
    t.s3 = t.a3 * t.a3 + t.b3 * t.b3;
    t.s4 = t.a4 * t.a4 + t.b4 * t.b4;
    t.F1 = spac_exp_finite(-0.25 * t.s1);
    t.F2 = spac_exp_finite((-1.0 / 49.0) * t.q2 - 0.1 * (9.0 * y + 1.0));
    t.F3 = spac_exp_finite(-0.25 * t.s3);
    t.F4 = spac_exp_finite(-1.0 * t.s4);
    return t;
}
__device__ __forceinline__ double franke_u(const FrankeTerms& t) {
    return 0.75 * t.F1 + 0.75 * t.F2 + 0.5 * t.F3 - 0.2 * t.F4;
}
__device__ __forceinline__ double franke_ux(const FrankeTerms& t) {
    return -3.375 * t.a1 * t.F1 - (27.0 / 98.0) * t.a2 * t.F2 - 2.25 * t.a3 * t.F3 + 3.6 * t.a4 * t.F4;
}
__device__ __forceinline__ double franke_uy(const FrankeTerms& t) {
    return -3.375 * t.b1 * t.F1 - 0.675 * t.F2 - 2.25 * t.b3 * t.F3 + 3.6 * t.b4 * t.F4;
}
/* u_xx + u_yy with one factor per Gaussian, linear in its squared distance: the sum of the
 * reference's u_xx and u_yy terms (BVPs/Franke_Poisson.hpp), e.g. for F1
 * -3.375 (9 - 4.5 a1^2) - 3.375 (9 - 4.5 b1^2) = 15.1875 s1 - 60.75; for F2
 * (243/2401) q2 + (243/400 - 243/98), the constants rounded once from the exact rationals. */
__device__ __forceinline__ double franke_lap(const FrankeTerms& t) {
    const double A1 = SPAC_FMA(15.1875, t.s1, -60.75);
    const double A2 = SPAC_FMA(0.10120783007080383, t.q2, -1.872091836734694);
    const double A3 = SPAC_FMA(10.125, t.s3, -40.5);
    const double A4 = SPAC_FMA(-64.8, t.s4, 64.8);
    return SPAC_FMA(A1, t.F1, SPAC_FMA(A2, t.F2, SPAC_FMA(A3, t.F3, A4 * t.F4)));
}

/* Franke Poisson (BVPs/Franke_Poisson.hpp): sigma = 1.41421356237 I, b = 0,
 * c = -0, f = -(u_xx + u_yy), F = -sigma^T grad u, g = u. */
struct CoefFranke {
    static constexpr bool kHasDrift = false;
    static constexpr bool kHasKilling = false;
    static constexpr bool kHasRobin = false; /* psi = varphi = 0 */
    static constexpr bool kConstSigma = true;
    static constexpr double kSigma = 1.41421356237;
    __device__ __forceinline__ static double sigma(double, double) { return kSigma; }
    __device__ __forceinline__ static StepCoef step_s(double x, double y, double) { return step(x, y); }
    __device__ __forceinline__ static StepCoef step(double x, double y) {
        FrankeTerms t = franke_terms(x, y);
        StepCoef c;
        c.s = kSigma;
        c.bx = 0.0; c.by = 0.0; c.c = 0.0;
        c.f = -franke_lap(t);
        c.Fx = -(kSigma * franke_ux(t));
        c.Fy = -(kSigma * franke_uy(t));
        return c;
    }
    __device__ __forceinline__ static double g(double x, double y) { return franke_u(franke_terms(x, y)); }
    __device__ __forceinline__ static double u(double x, double y) { return franke_u(franke_terms(x, y)); }
    __device__ __forceinline__ static void robin(double, double, double, double, double* psi, double* phi) {
        *psi = 0.0;
        *phi = 0.0;
    }
};

/* Franke Poisson with Robin data on reflecting sides: varphi = -0.5,
 * psi(W, N) = grad u(W).N + 0.5 u(W), so du/dN - varphi u = psi for the Franke u. */
struct CoefFrankeRobin : CoefFranke {
    static constexpr bool kHasRobin = true;
    __device__ __forceinline__ static void robin(double x, double y, double nx, double ny, double* psi,
                                                 double* phi) {
        FrankeTerms t = franke_terms(x, y);
        *psi = franke_ux(t) * nx + franke_uy(t) * ny + 0.5 * franke_u(t);
        *phi = -0.5;
    }
};

/* Variable-coefficient operator with the Franke manufactured solution:
 * sigma = s I, s = 1 + 0.5 sin(pi x) cos(pi y); b = 0.3 (y, -x);
 * c = -1 - x^2; f = -((1/2 s^2 lap u + b.grad u) + c u); F = -s grad u. */
struct CoefVarFranke {
    static constexpr bool kHasDrift = true;
    static constexpr bool kHasKilling = true;
    static constexpr bool kHasRobin = false;
    static constexpr bool kConstSigma = false;
    __device__ __forceinline__ static double sigma(double x, double y) {
        return 1.0 + 0.5 * spac_sinpi(x) * spac_cospi(y);
    }
    __device__ __forceinline__ static StepCoef step(double x, double y) { return step_s(x, y, sigma(x, y)); }
    /* s = sigma(x, y), already known to the caller */
    __device__ __forceinline__ static StepCoef step_s(double x, double y, double s) {
        FrankeTerms t = franke_terms(x, y);
        StepCoef c;
        c.s = s;
        c.bx = 0.3 * y;
        c.by = -0.3 * x;
        c.c = -1.0 - x * x;
        double ux = franke_ux(t), uy = franke_uy(t);
        double lap = franke_lap(t);
        double adv = c.bx * ux + c.by * uy;
        double reac = c.c * franke_u(t);
        double diff = (0.5 * (s * s)) * lap;
        c.f = -((diff + adv) + reac);
        c.Fx = -(s * ux);
        c.Fy = -(s * uy);
        return c;
    }
    __device__ __forceinline__ static double g(double x, double y) { return franke_u(franke_terms(x, y)); }
    __device__ __forceinline__ static double u(double x, double y) { return franke_u(franke_terms(x, y)); }
    __device__ __forceinline__ static void robin(double, double, double, double, double* psi, double* phi) {
        *psi = 0.0;
        *phi = 0.0;
    }
};

/* A coefficient set with tabulated coefficients (spac_tab.h) over the compiled set Base: the
 * coefficients in the table's mask come from the table, the others from Base evaluated as its
 * own plugins would be (Base's f and F with Base's sigma).  Unlike the compiled sets it carries
 * state (the table), so kernels call it through an instance. */
template <class Base>
struct CoefTab {
    static constexpr bool kHasDrift = true;
    static constexpr bool kHasKilling = true;
    static constexpr bool kHasRobin = false;
    static constexpr bool kConstSigma = false;
    spac_coef_table t;
    __device__ __forceinline__ double sigma(double x, double y) const {
        return (t.mask & SPAC_TAB_SIGMA) ? spac_tab_at(&t, x, y, SPAC_TABC_S) : Base::sigma(x, y);
    }
    __device__ __forceinline__ StepCoef step_s(double x, double y, double s) const {
        StepCoef c;
        const int32_t fbc = SPAC_TAB_SOURCE | SPAC_TAB_DRIFT | SPAC_TAB_KILLING;
        if ((t.mask & fbc) != fbc || t.need_F) {
            c = Base::step_s(x, y, (t.mask & SPAC_TAB_SIGMA) ? Base::sigma(x, y) : s);
        } else {
            c.Fx = 0.0;
            c.Fy = 0.0;
        }
        c.s = s;
        int32_t k00;
        double ax, ay;
        spac_tab_cell(&t, x, y, &k00, &ax, &ay);
        if (t.mask & SPAC_TAB_DRIFT) {
            c.bx = spac_tab_eval(&t, k00, ax, ay, SPAC_TABC_BX);
            c.by = spac_tab_eval(&t, k00, ax, ay, SPAC_TABC_BY);
        }
        if (t.mask & SPAC_TAB_KILLING) c.c = spac_tab_eval(&t, k00, ax, ay, SPAC_TABC_C);
        if (t.mask & SPAC_TAB_SOURCE) c.f = spac_tab_eval(&t, k00, ax, ay, SPAC_TABC_F);
        return c;
    }
    __device__ __forceinline__ double g(double x, double y) const {
        return (t.mask & SPAC_TAB_DIRICHLET) ? spac_tab_at(&t, x, y, SPAC_TABC_G) : Base::g(x, y);
    }
    __device__ __forceinline__ double u(double x, double y) const {
        return (t.mask & SPAC_TAB_SOLUTION) ? spac_tab_at(&t, x, y, SPAC_TABC_U) : Base::u(x, y);
    }
    __device__ __forceinline__ static void robin(double, double, double, double, double* psi, double* phi) {
        *psi = 0.0;
        *phi = 0.0;
    }
};

/* --------------------------------------------------------- boundaries -- */
struct BoundaryHit {
    double d;
    double ex, ey; /* closest boundary point (Npro) */
    double nx, ny; /* outward normal */
};

/* Rectangulo (Dominios/Rectangulo.hpp:7-229): quadrant tree, corner normals
 * from the float literal 0.70710678118f. */
__device__ __forceinline__ BoundaryHit rectangle_distance(const double* p, double x, double y) {
    const double K = (double)0.70710678118f;
    double c0 = 0.5 * (p[0] + p[2]), c1 = 0.5 * (p[1] + p[3]);
    double h0 = p[2] - c0, h1 = p[3] - c1;
    double d0 = __builtin_fabs(x - c0), d1 = __builtin_fabs(y - c1);
    double sx = (x > c0) ? 1.0 : -1.0;
    double sy = (y > c1) ? 1.0 : -1.0;
    /* the quadrant tree as selects (lanes of a wave take different sides): the
     * nearest side inside, the side crossed outside; corners below */
    const bool inside = (d0 <= h0) && (d1 <= h1);
    const bool xside = inside ? (d0 - h0 > d1 - h1) : (d0 > h0);
    BoundaryHit r;
    r.ex = xside ? sx * h0 + c0 : x;
    r.ey = xside ? y : sy * h1 + c1;
    r.nx = xside ? sx : 0.0;
    r.ny = xside ? 0.0 : sy;
    r.d = xside ? d0 - h0 : d1 - h1;
    if (!inside && d0 > h0 && d1 > h1) { /* outside a corner (rare) */
        double e0 = d0 - h0, e1 = d1 - h1;
        r.d = __builtin_sqrt(e0 * e0 + e1 * e1);
        r.ex = sx * h0 + c0; r.ey = sy * h1 + c1;
        r.nx = sx * K; r.ny = sy * K;
    }
    return r;
}

/* Frontera::Evalua_Absorbente == false (BVP/Frontera.hpp:23,34) on a rectangle
 * with the SPAC_REFLECT_* mask: the side or corner is read off the outward
 * normal; a corner reflects only when both of its sides do. */
__device__ __forceinline__ bool side_reflects(int32_t mask, double nx, double ny) {
    int32_t bits = (nx < 0.0 ? 1 : 0) | (ny < 0.0 ? 2 : 0) | (nx > 0.0 ? 4 : 0) | (ny > 0.0 ? 8 : 0);
    return bits != 0 && (bits & mask) == bits;
}

/* Circulo (Dominios/Circulo.hpp:7-22): params R, cx, cy; radius clamp 5e-4. */
__device__ __forceinline__ BoundaryHit circle_distance(double R, double cx, double cy, double x, double y) {
    BoundaryHit r;
    double nx = x - cx, ny = y - cy;
    double rr = __builtin_sqrt(nx * nx + ny * ny);
    rr = (5E-4 < rr) ? rr : 5E-4;
    const double ir = 1.0 / rr;
    nx = nx * ir;
    ny = ny * ir;
    r.nx = nx; r.ny = ny;
    r.ex = nx * R + cx;
    r.ey = ny * R + cy;
    r.d = rr - R;
    return r;
}

/* ----------------------------------------------------- FP32 stepping --
 * Mixed-precision twins: coefficients, increments and the per-step stopping
 * test in FP32 with the gfx950 transcendental instructions (v_exp_f32 = 2^x,
 * v_log_f32 = log2 x, v_sin_f32 / v_cos_f32 with the argument in revolutions);
 * the weight Y and the sums Z, xi stay FP64, and exits are projected in FP64. */
struct StepCoef32 {
    float s, bx, by, c, f, Fx, Fy;
};

/* The FP32 path is checked against the FP64 run within stated tolerances, not bit
 * for bit, so its arithmetic may contract a * b + c into FMAs (the build disables
 * contraction elsewhere, for the bit-reproducible FP64 path). */
#pragma clang fp contract(fast)

__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
/* sin(pi x), cos(pi x) */
__device__ __forceinline__ float fsinpi(float x) { return __builtin_amdgcn_sinf(0.5f * x); }
__device__ __forceinline__ float fcospi(float x) { return __builtin_amdgcn_cosf(0.5f * x); }

struct FrankeTerms32 {
    float a1, b1, a2, a3, b3, a4, b4;
    float F1, F2, F3, F4;
};
__device__ __forceinline__ FrankeTerms32 franke_terms32(float x, float y) {
    FrankeTerms32 t;
    t.a1 = 9.0f * x - 2.0f; t.b1 = 9.0f * y - 2.0f;
    t.a2 = 9.0f * x + 1.0f;
    t.a3 = 9.0f * x - 7.0f; t.b3 = 9.0f * y - 3.0f;
    t.a4 = 9.0f * x - 4.0f; t.b4 = 9.0f * y - 7.0f;
    t.F1 = fexp(-0.25f * (t.a1 * t.a1 + t.b1 * t.b1));
    t.F2 = fexp((-1.0f / 49.0f) * (t.a2 * t.a2) - 0.1f * (9.0f * y + 1.0f));
    t.F3 = fexp(-0.25f * (t.a3 * t.a3 + t.b3 * t.b3));
    t.F4 = fexp(-(t.a4 * t.a4 + t.b4 * t.b4));
    return t;
}
__device__ __forceinline__ float franke_ux32(const FrankeTerms32& t) {
    return -3.375f * t.a1 * t.F1 - (27.0f / 98.0f) * t.a2 * t.F2 - 2.25f * t.a3 * t.F3 + 3.6f * t.a4 * t.F4;
}
__device__ __forceinline__ float franke_uy32(const FrankeTerms32& t) {
    return -3.375f * t.b1 * t.F1 - 0.675f * t.F2 - 2.25f * t.b3 * t.F3 + 3.6f * t.b4 * t.F4;
}
/* the factors of franke_lap in FP32 */
__device__ __forceinline__ float franke_lap32(const FrankeTerms32& t) {
    const float A1 = 15.1875f * (t.a1 * t.a1 + t.b1 * t.b1) - 60.75f;
    const float A2 = 0.10120783f * (t.a2 * t.a2) - 1.8720918f;
    const float A3 = 10.125f * (t.a3 * t.a3 + t.b3 * t.b3) - 40.5f;
    const float A4 = 64.8f - 64.8f * (t.a4 * t.a4 + t.b4 * t.b4);
    return A1 * t.F1 + A2 * t.F2 + A3 * t.F3 + A4 * t.F4;
}
__device__ __forceinline__ float franke_u32(const FrankeTerms32& t) {
    return 0.75f * t.F1 + 0.75f * t.F2 + 0.5f * t.F3 - 0.2f * t.F4;
}

struct CoefFranke32 {
    static constexpr bool kHasDrift = false;
    static constexpr bool kHasKilling = false;
    static constexpr bool kConstSigma = true;
    using F64 = CoefFranke;
    __device__ __forceinline__ static float sigma(float, float) { return 1.41421356237f; }
    __device__ __forceinline__ static StepCoef32 step_s(float x, float y, float) { return step(x, y); }
    __device__ __forceinline__ static StepCoef32 step(float x, float y) {
        FrankeTerms32 t = franke_terms32(x, y);
        StepCoef32 c;
        c.s = 1.41421356237f;
        c.bx = c.by = c.c = 0.0f;
        c.f = -franke_lap32(t);
        c.Fx = -(c.s * franke_ux32(t));
        c.Fy = -(c.s * franke_uy32(t));
        return c;
    }
};

struct CoefVarFranke32 {
    static constexpr bool kHasDrift = true;
    static constexpr bool kHasKilling = true;
    static constexpr bool kConstSigma = false;
    using F64 = CoefVarFranke;
    __device__ __forceinline__ static float sigma(float x, float y) { return 1.0f + 0.5f * fsinpi(x) * fcospi(y); }
    __device__ __forceinline__ static StepCoef32 step(float x, float y) { return step_s(x, y, sigma(x, y)); }
    __device__ __forceinline__ static StepCoef32 step_s(float x, float y, float s) {
        FrankeTerms32 t = franke_terms32(x, y);
        StepCoef32 c;
        c.s = s;
        c.bx = 0.3f * y;
        c.by = -0.3f * x;
        c.c = -1.0f - x * x;
        float ux = franke_ux32(t), uy = franke_uy32(t);
        c.f = -(((0.5f * (s * s)) * franke_lap32(t) + (c.bx * ux + c.by * uy)) + c.c * franke_u32(t));
        c.Fx = -(s * ux);
        c.Fy = -(s * uy);
        return c;
    }
};

/* The FP32 twin of CoefTab: a float copy of the table, bilinear in FP32. */
struct CoefTable32 {
    const float* v; /* T * T * 8 floats */
    int32_t t, mask, need_F;
    float lox, loy, ix, iy, top;
};
__device__ __forceinline__ void tab_cell32(const CoefTable32& c, float x, float y, int32_t* k00, float* ax, float* ay) {
    float fx = (x - c.lox) * c.ix, fy = (y - c.loy) * c.iy;
    fx = __builtin_fminf(__builtin_fmaxf(fx, 0.0f), c.top);
    fy = __builtin_fminf(__builtin_fmaxf(fy, 0.0f), c.top);
    int32_t i = min((int32_t)fx, c.t - 2), j = min((int32_t)fy, c.t - 2);
    *ax = fx - (float)i;
    *ay = fy - (float)j;
    *k00 = j * c.t + i;
}
__device__ __forceinline__ float tab_eval32(const CoefTable32& c, int32_t k00, float ax, float ay, int comp) {
    const float* p = c.v + (int64_t)k00 * SPAC_TAB_STRIDE + comp;
    const int64_t row = (int64_t)c.t * SPAC_TAB_STRIDE;
    const float bx = 1.0f - ax, by = 1.0f - ay;
    return by * (bx * p[0] + ax * p[SPAC_TAB_STRIDE]) + ay * (bx * p[row] + ax * p[row + SPAC_TAB_STRIDE]);
}
template <class Base32>
struct CoefTab32 {
    static constexpr bool kHasDrift = true;
    static constexpr bool kHasKilling = true;
    static constexpr bool kConstSigma = false;
    using F64 = CoefTab<typename Base32::F64>;
    CoefTable32 t;
    __device__ __forceinline__ float sigma(float x, float y) const {
        if (!(t.mask & SPAC_TAB_SIGMA)) return Base32::sigma(x, y);
        int32_t k00;
        float ax, ay;
        tab_cell32(t, x, y, &k00, &ax, &ay);
        return tab_eval32(t, k00, ax, ay, SPAC_TABC_S);
    }
    __device__ __forceinline__ StepCoef32 step_s(float x, float y, float s) const {
        StepCoef32 c;
        const int32_t fbc = SPAC_TAB_SOURCE | SPAC_TAB_DRIFT | SPAC_TAB_KILLING;
        if ((t.mask & fbc) != fbc || t.need_F) {
            c = Base32::step_s(x, y, (t.mask & SPAC_TAB_SIGMA) ? Base32::sigma(x, y) : s);
        } else {
            c.Fx = 0.0f;
            c.Fy = 0.0f;
        }
        c.s = s;
        int32_t k00;
        float ax, ay;
        tab_cell32(t, x, y, &k00, &ax, &ay);
        if (t.mask & SPAC_TAB_DRIFT) {
            c.bx = tab_eval32(t, k00, ax, ay, SPAC_TABC_BX);
            c.by = tab_eval32(t, k00, ax, ay, SPAC_TABC_BY);
        }
        if (t.mask & SPAC_TAB_KILLING) c.c = tab_eval32(t, k00, ax, ay, SPAC_TABC_C);
        if (t.mask & SPAC_TAB_SOURCE) c.f = tab_eval32(t, k00, ax, ay, SPAC_TABC_F);
        return c;
    }
};

/* Box–Muller in FP32 from the same Philox block as the FP64 path (24-bit uniforms). */
/* FP32 Box–Muller from the same two words as spac_normal_words (their top 24 bits), so FP32 and
 * FP64 paths of one seed draw the same increments up to the uniforms' rounding. */
__device__ __forceinline__ void normal_words32(uint32_t a, uint32_t b, float* n0, float* n1) {
    float u1 = 1.0f - (float)(a >> 8) * 5.9604644775390625e-08f; /* (0, 1] */
    float u2 = (float)(b >> 8) * 5.9604644775390625e-08f;        /* [0, 1) */
    float rad = __builtin_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1)); /* -2 ln u1 = -2 ln2 log2 u1 */
    *n0 = rad * __builtin_amdgcn_cosf(u2);
    *n1 = rad * __builtin_amdgcn_sinf(u2);
}
/* step_normals (fk_kernels.hip) for the FP32 kernel */
__device__ __forceinline__ void step_normals32(uint64_t seed, uint32_t stream, uint32_t path, int32_t nstep,
                                               uint32_t* keep0, uint32_t* keep1, float* n0, float* n1) {
    uint32_t w0, w1;
    if ((nstep & 1) == 0) {
        const spac_u32x4 r = spac_brownian_block(seed, stream, path, (uint32_t)nstep >> 1);
        w0 = r.v[0]; w1 = r.v[1]; *keep0 = r.v[2]; *keep1 = r.v[3];
    } else {
        w0 = *keep0; w1 = *keep1;
    }
    normal_words32(w0, w1, n0, n1);
}

__device__ __forceinline__ float rect_distance32(const float* p, float x, float y, float* nx, float* ny, float* ex,
                                                 float* ey) {
    const float K = 0.70710678118f;
    float c0 = 0.5f * (p[0] + p[2]), c1 = 0.5f * (p[1] + p[3]);
    float h0 = p[2] - c0, h1 = p[3] - c1;
    float d0 = __builtin_fabsf(x - c0), d1 = __builtin_fabsf(y - c1);
    float sx = (x > c0) ? 1.0f : -1.0f, sy = (y > c1) ? 1.0f : -1.0f;
    /* the quadrant tree as selects, corners in a (rare) branch */
    const bool inside = d0 <= h0 && d1 <= h1;
    const bool xside = inside ? (d0 - h0 > d1 - h1) : (d0 > h0);
    *ex = xside ? sx * h0 + c0 : x;
    *ey = xside ? y : sy * h1 + c1;
    *nx = xside ? sx : 0.0f;
    *ny = xside ? 0.0f : sy;
    float d = xside ? d0 - h0 : d1 - h1;
    if (!inside && d0 > h0 && d1 > h1) {
        float e0 = d0 - h0, e1 = d1 - h1;
        *ex = sx * h0 + c0; *ey = sy * h1 + c1; *nx = sx * K; *ny = sy * K;
        d = __builtin_sqrtf(e0 * e0 + e1 * e1);
    }
    return d;
}

/* Tabulated control variate in FP32: bilinear d u_hat/dx, d u_hat/dy at (x, y),
 * the same cell rule as spac_cv_cell (clamped to the grid, NaN to the origin).  The
 * gradients are stored by cell: cell (i, j), i, j < T - 1, holds its four corners
 * (x, y gradient at (i, j), (i + 1, j), (i, j + 1), (i + 1, j + 1)) in 32 bytes, so a lookup
 * is two 16-byte loads from one cache line. */
struct CvTable32 {
    const float4* cells; /* (T - 1)^2 cells x 2 float4 */
    int32_t t;
    float lox, loy, ix, iy, top;
};
__device__ __forceinline__ void cv_grad32(const CvTable32& c, float x, float y, float* gx, float* gy) {
    float fx = (x - c.lox) * c.ix, fy = (y - c.loy) * c.iy;
    fx = __builtin_fminf(__builtin_fmaxf(fx, 0.0f), c.top);
    fy = __builtin_fminf(__builtin_fmaxf(fy, 0.0f), c.top);
    int32_t i = min((int32_t)fx, c.t - 2), j = min((int32_t)fy, c.t - 2);
    float ax = fx - (float)i, ay = fy - (float)j, bx = 1.0f - ax, by = 1.0f - ay;
    const float4* p = c.cells + 2 * (j * (c.t - 1) + i);
    const float4 lo = p[0], hi = p[1]; /* lo: corners (i, j), (i + 1, j); hi: (i, j + 1), (i + 1, j + 1) */
    *gx = by * (bx * lo.x + ax * lo.z) + ay * (bx * hi.x + ax * hi.z);
    *gy = by * (bx * lo.y + ax * lo.w) + ay * (bx * hi.y + ax * hi.w);
}

/* The learned control variate evaluated directly (SPAC_CV_NETWORK): the MLP 2 -> 64 -> 64 ->
 * 64 -> 3 of spac_mlp.h (tanh; parameters [W1 64x2, b1, W2 64x64, b2, W3 64x64, b3, W4 3x64,
 * b4], W row-major [out][in]; input z = 2 (x - lo) s - 1 per coordinate), in FP32 on one lane,
 * returning its gradient heads (outputs 1 and 2).  Layer 3 is accumulated while layer 2 is
 * produced, so one lane holds 128 activations.  The parameters are the same for every lane,
 * so they are read through the scalar cache. */
constexpr int kNetH = 64;
constexpr int kNetW1 = 0, kNetB1 = kNetW1 + 2 * kNetH, kNetW2 = kNetB1 + kNetH, kNetB2 = kNetW2 + kNetH * kNetH,
              kNetW3 = kNetB2 + kNetH, kNetB3 = kNetW3 + kNetH * kNetH, kNetW4 = kNetB3 + kNetH,
              kNetB4 = kNetW4 + 3 * kNetH, kNetParams = kNetB4 + 3;
static_assert(kNetParams == 8707, "network parameter layout of spac_mlp.h");
struct CvNet32 {
    const float* P;
    float lox, loy, sx, sy;
};
__device__ __forceinline__ float net_tanh(float x) {
    /* tanh x = 1 - 2 / (e^{2x} + 1), the hardware 2^x and reciprocal (the inference kernel's form) */
    const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * x);
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}
__device__ __forceinline__ void net_grad32(const CvNet32& n, float x, float y, float* gx, float* gy) {
    const float* __restrict__ P = n.P;
    const float z0 = 2.0f * (x - n.lox) * n.sx - 1.0f, z1 = 2.0f * (y - n.loy) * n.sy - 1.0f;
    float h1[kNetH], h3[kNetH];
#pragma unroll
    for (int f = 0; f < kNetH; ++f) {
        h1[f] = net_tanh(__builtin_fmaf(P[kNetW1 + 2 * f], z0, __builtin_fmaf(P[kNetW1 + 2 * f + 1], z1, P[kNetB1 + f])));
        h3[f] = P[kNetB3 + f];
    }
    for (int j = 0; j < kNetH; ++j) {
        float a = P[kNetB2 + j];
#pragma unroll
        for (int k = 0; k < kNetH; ++k) a = __builtin_fmaf(P[kNetW2 + j * kNetH + k], h1[k], a);
        const float h2 = net_tanh(a);
#pragma unroll
        for (int m = 0; m < kNetH; ++m) h3[m] = __builtin_fmaf(P[kNetW3 + m * kNetH + j], h2, h3[m]);
    }
    float ox = P[kNetB4 + 1], oy = P[kNetB4 + 2];
#pragma unroll
    for (int m = 0; m < kNetH; ++m) {
        const float t = net_tanh(h3[m]);
        ox = __builtin_fmaf(P[kNetW4 + kNetH + m], t, ox);
        oy = __builtin_fmaf(P[kNetW4 + 2 * kNetH + m], t, oy);
    }
    *gx = ox;
    *gy = oy;
}

#pragma clang fp contract(off)

} // namespace spac

#endif
