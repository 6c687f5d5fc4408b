/*
 * spac_rbf.h — exit angle on an interface circle and the RBF row weights,
 * shared by the host table builder, the device kernels and the CPU
 * restatement (C and HIP).
 *
 * Exit angle (Mesh/Interfaz.cpp:222-251 with Mesh/Nudo.hpp:7-32): the angle of
 * X_tau - centre with the branch cut of the interface code: Atan180 for
 * -11, -22, -4; Atan0 for -33, -44, -2; Atan270 for -1; Atan90 for -3.  Full
 * circles (code 1) use the angle relative to the first node, in [0, 2pi).
 *
 * Global inverse-multiquadric rows (reference, arcs only):
 *   H_j(theta) = sum_i iPsi(i,j) / sqrt((theta - theta_i)^2 + c2).
 * Local Gaussian stencils (S nodes): the stencil is the S consecutive nodes
 * around the node nearest to theta (clamped to the arc; periodic on a full
 * circle); with t = (theta - theta_i) / L, L the stencil half-width,
 *   phi = exp(-(eps t)^2),  H_{s+j}(theta) = sum_i iPsiS(i,j) phi_i.
 */
#ifndef SPAC_RBF_H
#define SPAC_RBF_H

#include "spac_math.h"

#define SPAC_TWO_PI 6.283185307179586

/* 0 if the code is unknown */
SPAC_HD int spac_exit_angle(int code, double dx, double dy, double theta_first, double* theta) {
    double aux = spac_atan2(dy, dx);
    switch (code) {
    case -11: case -22: case -4:
        *theta = aux;
        return 1;
    case -33: case -44: case -2:
        *theta = (aux >= 0) ? aux : aux + 2.0 * SPAC_PI;
        return 1;
    case -1:
        *theta = (aux <= -0.5 * SPAC_PI) ? 2 * SPAC_PI + aux : aux;
        return 1;
    case -3:
        *theta = (aux <= 0.5 * SPAC_PI) ? 2.0 * SPAC_PI + aux : aux;
        return 1;
    case 1: {
        double r = aux - theta_first;
        while (r < 0.0) r += SPAC_TWO_PI;
        while (r >= SPAC_TWO_PI) r -= SPAC_TWO_PI;
        *theta = r;
        return 1;
    }
    default:
        return 0;
    }
}

SPAC_HD double spac_imq(double theta, double theta_j, double c2) {
    double d = theta - theta_j;
    return 1.0 / SPAC_SQRT(d * d + c2);
}

SPAC_HD double spac_gauss_phi(double t, double eps) {
    double e = eps * t;
    return spac_exp_finite(-(e * e));
}

/* Nearest node on an arc of sorted angles th[0..M-1] (ties: lower index). */
SPAC_HD int spac_arc_nearest(const double* th, int M, double theta) {
    int lo = 0, hi = M; /* first index with th >= theta */
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (th[mid] < theta) lo = mid + 1; else hi = mid;
    }
    if (lo == 0) return 0;
    if (lo == M) return M - 1;
    double dl = theta - th[lo - 1], dh = th[lo] - theta;
    return (dh < dl) ? lo : lo - 1;
}

/* Stencil start on an arc: S consecutive nodes centred on k, clamped. */
SPAC_HD int spac_arc_start(int k, int M, int S) {
    int s = k - S / 2;
    if (s > M - S) s = M - S;
    if (s < 0) s = 0;
    return s;
}

/* Arc stencil start s and half-width L for exit angle theta; phi_i = spac_gauss_phi((theta -
 * th[s + i]) / L, eps). */
SPAC_HD int spac_arc_stencil_start(const double* th, int M, int S, double theta, double* L) {
    int k = spac_arc_nearest(th, M, theta);
    int s = spac_arc_start(k, M, S);
    *L = (S > 1) ? (th[s + S - 1] - th[s]) / 2.0 : 1.0;
    return s;
}

/* Arc stencil: start s and the S values phi_i for exit angle theta. */
SPAC_HD int spac_arc_stencil(const double* th, int M, int S, double eps, double theta, double* phi) {
    double L;
    int s = spac_arc_stencil_start(th, M, S, theta, &L);
    for (int i = 0; i < S; ++i) phi[i] = spac_gauss_phi((theta - th[s + i]) / L, eps);
    return s;
}

/* phi_i = exp(-(a - i b)^2), i = 0..S-1, for equally spaced nodes: four exps
 * and the recurrence phi_{i+1} = phi_i (E F_i) with E = exp(2 a b) and
 * F_i = exp(-(2 i + 1) b^2) = F_{i-1} exp(-2 b^2).  Each product adds one
 * rounding, so phi_i carries at most ~2i ulp more than a direct exp. */
typedef struct {
    double p, E, F, G; /* p = the current phi_i */
} spac_gauss_seq;
SPAC_HD spac_gauss_seq spac_gauss_seq_start(double a, double b) {
    spac_gauss_seq q;
    const double b2 = b * b;
    q.p = spac_exp_finite(-(a * a));
    q.E = spac_exp_finite(2.0 * (a * b));
    q.F = spac_exp_finite(-b2);
    q.G = spac_exp_finite(-2.0 * b2);
    return q;
}
/* phi_i -> phi_{i+1} */
SPAC_HD void spac_gauss_seq_next(spac_gauss_seq* q) {
    q->p = q->p * (q->E * q->F);
    q->F = q->F * q->G;
}
SPAC_HD void spac_gauss_row(double a, double b, int S, double* phi) {
    spac_gauss_seq q = spac_gauss_seq_start(a, b);
    phi[0] = q.p;
    for (int i = 1; i < S; ++i) {
        spac_gauss_seq_next(&q);
        phi[i] = q.p;
    }
}

/* Fixed-point moments of Gaussian-stencil rows.  A row's G is linear in its paths:
 *   G_{s+r} = -(1/N) sum_p Y_p H_{s_p+r}(theta_p),  H_{s+r} = sum_i iPsi_s(i, r) phi_i,
 * so it is also -(1/N) sum_{s,r} sum_i iPsi_s(i, r) Phi[s][i] with the moments
 *   Phi[s][i] = sum over the row's circle-exit paths with stencil start s of Y_p phi_i(theta_p).
 * The moments are kept as 64-bit integers in units of 2^-q: each term Y_p phi_i is rounded
 * once to the nearest multiple, and integer sums are exact, so a moment does not depend on
 * the order its paths are added in.  q is the largest exponent with N max|Y| 2^q < 2^62 (no
 * sum of N terms, phi <= 1, overflows) and max|Y| 2^q < 2^50 (every term is below 2^50, so
 * adding 1.5 2^52 rounds it to the nearest integer, ties to even, and leaves that integer in
 * the low bits: a conversion of four instructions).  A term's rounding is at most 2^-q-1,
 * i.e. below 2^-63 N max|Y| and 2^-51 max|Y|. */
SPAC_HD int spac_exponent_above(double b) { /* e with b < 2^(e + 1), b > 0 finite */
    return (int)((spac_f64_as_u64(b) >> 52) & 0x7ff) - 1023;
}
SPAC_HD int spac_moment_scale(double n_paths, double ymax) {
    const double b = n_paths * ymax;
    if (!(b > 0.0) || !(b < 1e300)) return 0;
    int q = 61 - spac_exponent_above(b);
    const int qy = 49 - spac_exponent_above(ymax);
    if (qy < q) q = qy;
    if (q > 1000) q = 1000;
    if (q < -1000) q = -1000;
    return q;
}
#define SPAC_FIXED_MAGIC 6755399441055744.0 /* 1.5 2^52 */
SPAC_HD int64_t spac_to_fixed(double v, int q) {
    return (int64_t)(spac_f64_as_u64(v * spac_pow2i(q) + SPAC_FIXED_MAGIC) - spac_f64_as_u64(SPAC_FIXED_MAGIC));
}
SPAC_HD double spac_from_fixed(int64_t n, int q) { return (double)n * spac_pow2i(-q); }

/* Full-circle stencil start s (mod M) and the (a, b) of its Gaussians phi_i = exp(-(a - i b)^2). */
SPAC_HD int spac_circle_stencil_ab(int M, int S, double eps, double r, double* a, double* b) {
    double delta = SPAC_TWO_PI / (double)M;
    int m = S / 2;
    int k = (int)SPAC_RINT(r / delta);
    double L = (m > 0) ? (double)m * delta : 1.0;
    /* d_i / L = (r - (k - m) delta) / L - i delta / L */
    double d0 = r - (double)(k - m) * delta;
    *a = eps * (d0 / L);
    *b = eps * (delta / L);
    int s = (k - m) % M;
    if (s < 0) s += M;
    return s;
}

/* Full-circle stencil (M equally spaced nodes, relative angle r in [0, 2pi)): returns the
 * index of the first stencil node (mod M) and fills the S values phi_i. */
SPAC_HD int spac_circle_stencil(int M, int S, double eps, double r, double* phi) {
    double a, b;
    int s = spac_circle_stencil_ab(M, S, eps, r, &a, &b);
    spac_gauss_row(a, b, S, phi);
    return s;
}

#endif
