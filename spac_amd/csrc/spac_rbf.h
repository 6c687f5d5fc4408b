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
    return spac_exp(-(e * e));
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

/* Arc stencil: start s and the S values phi_i for exit angle theta. */
SPAC_HD int spac_arc_stencil(const double* th, int M, int S, double eps, double theta, double* phi) {
    int k = spac_arc_nearest(th, M, theta);
    int s = spac_arc_start(k, M, S);
    double L = (S > 1) ? (th[s + S - 1] - th[s]) / 2.0 : 1.0;
    for (int i = 0; i < S; ++i) phi[i] = spac_gauss_phi((theta - th[s + i]) / L, eps);
    return s;
}

/* Full-circle stencil (M equally spaced nodes, relative angle r in [0, 2pi)):
 * returns the index of the first stencil node (mod M). */
/* phi_i = exp(-(a - i b)^2), i = 0..S-1, for equally spaced nodes: four exps
 * and the recurrence phi_{i+1} = phi_i (E F_i) with E = exp(2 a b) and
 * F_i = exp(-(2 i + 1) b^2) = F_{i-1} exp(-2 b^2).  Each product adds one
 * rounding, so phi_i carries at most ~2i ulp more than a direct exp. */
SPAC_HD void spac_gauss_row(double a, double b, int S, double* phi) {
    const double b2 = b * b;
    double p = spac_exp(-(a * a));
    const double E = spac_exp(2.0 * (a * b));
    double F = spac_exp(-b2);
    const double G = spac_exp(-2.0 * b2);
    phi[0] = p;
    for (int i = 1; i < S; ++i) {
        p = p * (E * F);
        phi[i] = p;
        F = F * G;
    }
}

SPAC_HD int spac_circle_stencil(int M, int S, double eps, double r, double* phi) {
    double delta = SPAC_TWO_PI / (double)M;
    int m = S / 2;
    int k = (int)SPAC_RINT(r / delta);
    double L = (m > 0) ? (double)m * delta : 1.0;
    /* d_i / L = (r - (k - m) delta) / L - i delta / L */
    double d0 = r - (double)(k - m) * delta;
    spac_gauss_row(eps * (d0 / L), eps * (delta / L), S, phi);
    int s = (k - m) % M;
    if (s < 0) s += M;
    return s;
}

#endif
