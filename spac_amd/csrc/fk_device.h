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

#include "spac_math.h"

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
    double F1, F2, F3, F4;
};

__device__ __forceinline__ FrankeTerms franke_terms(double x, double y) {
    FrankeTerms t;
    t.a1 = 9.0 * x - 2.0; t.b1 = 9.0 * y - 2.0;
    t.a2 = 9.0 * x + 1.0;
    t.a3 = 9.0 * x - 7.0; t.b3 = 9.0 * y - 3.0;
    t.a4 = 9.0 * x - 4.0; t.b4 = 9.0 * y - 7.0;
    t.F1 = spac_exp(-0.25 * (t.a1 * t.a1 + t.b1 * t.b1));
    t.F2 = spac_exp((-1.0 / 49.0) * (t.a2 * t.a2) - 0.1 * (9.0 * y + 1.0));
    t.F3 = spac_exp(-0.25 * (t.a3 * t.a3 + t.b3 * t.b3));
    t.F4 = spac_exp(-1.0 * (t.a4 * t.a4 + t.b4 * t.b4));
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
__device__ __forceinline__ double franke_uxx(const FrankeTerms& t) {
    return -3.375 * (9.0 - 4.5 * (t.a1 * t.a1)) * t.F1
           - (27.0 / 98.0) * (9.0 - (18.0 / 49.0) * (t.a2 * t.a2)) * t.F2
           - 2.25 * (9.0 - 4.5 * (t.a3 * t.a3)) * t.F3
           + 3.6 * (9.0 - 18.0 * (t.a4 * t.a4)) * t.F4;
}
__device__ __forceinline__ double franke_uyy(const FrankeTerms& t) {
    return -3.375 * (9.0 - 4.5 * (t.b1 * t.b1)) * t.F1
           + (243.0 / 400.0) * t.F2 - 2.25 * (9.0 - 4.5 * (t.b3 * t.b3)) * t.F3
           + 3.6 * (9.0 - 18.0 * (t.b4 * t.b4)) * t.F4;
}

/* Franke Poisson (BVPs/Franke_Poisson.hpp): sigma = 1.41421356237 I, b = 0,
 * c = -0, f = -(u_xx + u_yy), F = -sigma^T grad u, g = u. */
struct CoefFranke {
    static constexpr bool kHasDrift = false;
    static constexpr bool kHasKilling = false;
    static constexpr bool kConstSigma = true;
    static constexpr double kSigma = 1.41421356237;
    __device__ __forceinline__ static double sigma(double, double) { return kSigma; }
    __device__ __forceinline__ static StepCoef step(double x, double y) {
        FrankeTerms t = franke_terms(x, y);
        StepCoef c;
        c.s = kSigma;
        c.bx = 0.0; c.by = 0.0; c.c = 0.0;
        c.f = -franke_uxx(t) - franke_uyy(t);
        c.Fx = -(kSigma * franke_ux(t));
        c.Fy = -(kSigma * franke_uy(t));
        return c;
    }
    __device__ __forceinline__ static double g(double x, double y) { return franke_u(franke_terms(x, y)); }
    __device__ __forceinline__ static double u(double x, double y) { return franke_u(franke_terms(x, y)); }
};

/* Variable-coefficient operator with the Franke manufactured solution:
 * sigma = s I, s = 1 + 0.5 sin(pi x) cos(pi y); b = 0.3 (y, -x);
 * c = -1 - x^2; f = -((1/2 s^2 lap u + b.grad u) + c u); F = -s grad u. */
struct CoefVarFranke {
    static constexpr bool kHasDrift = true;
    static constexpr bool kHasKilling = true;
    static constexpr bool kConstSigma = false;
    __device__ __forceinline__ static double sigma(double x, double y) {
        return 1.0 + 0.5 * spac_sinpi(x) * spac_cospi(y);
    }
    __device__ __forceinline__ static StepCoef step(double x, double y) {
        FrankeTerms t = franke_terms(x, y);
        StepCoef c;
        double s = sigma(x, y);
        c.s = s;
        c.bx = 0.3 * y;
        c.by = -0.3 * x;
        c.c = -1.0 - x * x;
        double ux = franke_ux(t), uy = franke_uy(t);
        double lap = franke_uxx(t) + franke_uyy(t);
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
    BoundaryHit r;
    r.ex = x; r.ey = y;
    bool inside = (d0 <= h0) && (d1 <= h1);
    if (inside) {
        if (d0 - h0 > d1 - h1) {
            r.ex = sx * h0 + c0; r.nx = sx; r.ny = 0.0; r.d = d0 - h0;
        } else {
            r.ey = sy * h1 + c1; r.nx = 0.0; r.ny = sy; r.d = d1 - h1;
        }
        return r;
    }
    if (d0 > h0) {
        if (d1 > h1) {
            double e0 = d0 - h0, e1 = d1 - h1;
            r.d = __builtin_sqrt(e0 * e0 + e1 * e1);
            r.ex = sx * h0 + c0; r.ey = sy * h1 + c1;
            r.nx = sx * K; r.ny = sy * K;
            return r;
        }
        r.ex = sx * h0 + c0; r.nx = sx; r.ny = 0.0; r.d = d0 - h0;
        return r;
    }
    r.ey = sy * h1 + c1; r.nx = 0.0; r.ny = sy; r.d = d1 - h1;
    return r;
}

/* Circulo (Dominios/Circulo.hpp:7-22): params R, cx, cy; radius clamp 5e-4. */
__device__ __forceinline__ BoundaryHit circle_distance(double R, double cx, double cy, double x, double y) {
    BoundaryHit r;
    double nx = x - cx, ny = y - cy;
    double rr = __builtin_sqrt(nx * nx + ny * ny);
    rr = (5E-4 < rr) ? rr : 5E-4;
    nx = nx / rr;
    ny = ny / rr;
    r.nx = nx; r.ny = ny;
    r.ex = nx * R + cx;
    r.ey = ny * R + cy;
    r.d = rr - R;
    return r;
}

} // namespace spac

#endif
