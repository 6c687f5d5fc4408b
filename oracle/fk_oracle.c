/*
 * fk_oracle.c — CPU restatement of SpAc's Feynman–Kac hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP kernels in
 * spac_amd/csrc: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product never links or calls it.
 *
 * It keeps the reference's structure on purpose: every coefficient and
 * boundary is a plugin reached through a function pointer and evaluated once
 * per use, exactly as Paso_Absorbente and Dentro call them
 * (Solvers/FeynmanKacSolver.cpp:7-18, 68-106), so that it doubles as the
 * "reference OpenMP CPU path" timed beside the GPU.  Differences from the
 * reference, all deliberate:
 *   - Brownian increments come from Philox4x32-10 + Box–Muller keyed by
 *     (seed, stream) and counted by (step, path) instead of one boost::mt19937
 *     per OpenMP thread (Solvers/FeynmanKacSolver.cpp:2-6, 107-136), so a path
 *     is the same on any thread count and on the GPU;
 *   - exp/log/sin/cos are the project's own FP64 functions (spac_math.h) so
 *     that CPU and GPU bits agree;
 *   - the estimator sums follow one fixed compensated reduction order
 *     (64 lane sums over strided paths, folded pairwise) instead of
 *     per-thread partial sums merged under omp critical (:217-228, 276-290);
 *   - xi and xixi are the true moments (the reference's enum alias puts
 *     both into sums[8], Solvers/FeynmanKacSolver.hpp:16-23);
 *   - RNGC counts the normals actually drawn (2 per step; the reference
 *     counts 4, :111 and :275);
 *   - a step cap (max_steps) ends a path that never stops.
 *
 * Parity status of each function is stated beside it ("pinned by" = the test
 * that checks it independently of this file).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/spac_fk.h"
#include "../spac_amd/csrc/spac_math.h"
#include "../spac_amd/csrc/spac_rbf.h"
#include "../spac_amd/csrc/spac_cv.h"
#include "../spac_amd/csrc/spac_tab.h"

typedef struct { double x, y; } vec2;
typedef struct { double a00, a01, a10, a11; } mat2;

typedef double (*pfescalar)(vec2);
typedef double (*pfescalarN)(vec2, vec2);
typedef vec2 (*pfvector)(vec2);
typedef mat2 (*pfmatriz)(vec2);
typedef double (*pfdistancia)(const double* params, vec2 x, vec2* exitpoint, vec2* normal);

typedef struct {
    pfdistancia distancia;
    double params[4];
} frontera_t;

/* BVP container, BVP/BVP.hpp:17-28 */
typedef struct {
    pfescalar u, g, c, f;
    pfescalarN psi, varphi;
    pfvector b, gradiente_u, F, mu;
    pfmatriz sigma;
    frontera_t frontera_dominio, frontera_subdominio;
    int reflecting;  /* SPAC_REFLECT_* sides of the rectangle domain */
    double rho;      /* Lepingle band */
} bvp_t;

/* ------------------------------------------------------------ defaults --
 * BVP/Coeficiente_Escalar.hpp:7-9, Coeficiente_EscalarN.hpp:7-9,
 * Coeficiente_Vectorial.hpp:6-8, Coeficiente_Matricial.hpp:6-8 */
static double escalar_defecto(vec2 x) { (void)x; return 0.0; }
static double escalarN_defecto(vec2 x, vec2 n) { (void)x; (void)n; return 0.0; }
static vec2 vector_defecto(vec2 x) { vec2 r = {x.x * 0.0, x.y * 0.0}; return r; }

/* ------------------------------------------------------------- Franke --
 * BVPs/Franke_Poisson.hpp:4-73.  pow(a, 2) is written a*a (pow(a,2) is
 * exactly a*a under correct rounding).  Pinned by: tests/test_oracle.py
 * (values and derivatives against finite differences of u; f = -lap u). */
static double fr1(vec2 X) { double a = 9.0 * X.x - 2.0, b = 9.0 * X.y - 2.0; return spac_exp_finite(-0.25 * (a * a + b * b)); }
static double fr2(vec2 X) { double a = 9.0 * X.x + 1.0; return spac_exp_finite((-1.0 / 49.0) * (a * a) - 0.1 * (9.0 * X.y + 1.0)); }
static double fr3(vec2 X) { double a = 9.0 * X.x - 7.0, b = 9.0 * X.y - 3.0; return spac_exp_finite(-0.25 * (a * a + b * b)); }
static double fr4(vec2 X) { double a = 9.0 * X.x - 4.0, b = 9.0 * X.y - 7.0; return spac_exp_finite(-1.0 * (a * a + b * b)); }

static double franke_u(vec2 X) {
    return 0.75 * fr1(X) + 0.75 * fr2(X) + 0.5 * fr3(X) - 0.2 * fr4(X);
}
static double franke_dudx(vec2 X) {
    return -3.375 * (9.0 * X.x - 2.0) * fr1(X) - (27.0 / 98.0) * (9.0 * X.x + 1.0) * fr2(X)
           - 2.25 * (9.0 * X.x - 7.0) * fr3(X) + 3.6 * (9.0 * X.x - 4.0) * fr4(X);
}
static double franke_dudy(vec2 X) {
    return -3.375 * (9.0 * X.y - 2.0) * fr1(X) - 0.675 * fr2(X)
           - 2.25 * (9.0 * X.y - 3.0) * fr3(X) + 3.6 * (9.0 * X.y - 7.0) * fr4(X);
}
static double franke_d2udx2(vec2 X) {
    double a = 9.0 * X.x - 2.0, b = 9.0 * X.x + 1.0, c = 9.0 * X.x - 7.0, d = 9.0 * X.x - 4.0;
    return -3.375 * (9.0 - 4.5 * (a * a)) * fr1(X)
           - (27.0 / 98.0) * (9.0 - (18.0 / 49.0) * (b * b)) * fr2(X)
           - 2.25 * (9.0 - 4.5 * (c * c)) * fr3(X)
           + 3.6 * (9.0 - 18.0 * (d * d)) * fr4(X);
}
static double franke_d2udy2(vec2 X) {
    double a = 9.0 * X.y - 2.0, c = 9.0 * X.y - 3.0, d = 9.0 * X.y - 7.0;
    return -3.375 * (9.0 - 4.5 * (a * a)) * fr1(X)
           + (243.0 / 400.0) * fr2(X) - 2.25 * (9.0 - 4.5 * (c * c)) * fr3(X)
           + 3.6 * (9.0 - 18.0 * (d * d)) * fr4(X);
}
/* u_xx + u_yy as the device evaluates it: per Gaussian the sum of the u_xx and u_yy factors
 * above, linear in the squared distance of its exponent (F1: 15.1875 s1 - 60.75; F2:
 * (243/2401) q2 + 243/400 - 243/98, constants rounded once from the exact rationals). */
static double franke_lap(vec2 X) {
    double a1 = 9.0 * X.x - 2.0, b1 = 9.0 * X.y - 2.0, a2 = 9.0 * X.x + 1.0;
    double a3 = 9.0 * X.x - 7.0, b3 = 9.0 * X.y - 3.0, a4 = 9.0 * X.x - 4.0, b4 = 9.0 * X.y - 7.0;
    double A1 = fma(15.1875, a1 * a1 + b1 * b1, -60.75);
    double A2 = fma(0.10120783007080383, a2 * a2, -1.872091836734694);
    double A3 = fma(10.125, a3 * a3 + b3 * b3, -40.5);
    double A4 = fma(-64.8, a4 * a4 + b4 * b4, 64.8);
    return fma(A1, fr1(X), fma(A2, fr2(X), fma(A3, fr3(X), A4 * fr4(X))));
}
static double franke_c(vec2 X) { (void)X; return -0.0; }
/* Franke_b builds an uninitialised Vector2d times 0.0 (BVPs/Franke_Poisson.hpp:51-54);
 * restated as the intended zero vector. */
static vec2 franke_b(vec2 X) { (void)X; vec2 r = {0.0, 0.0}; return r; }
static mat2 franke_sigma(vec2 X) {
    (void)X;
    const double s = 1.41421356237; /* literal of BVPs/Franke_Poisson.hpp:56 */
    mat2 m = {1.0 * s, 0.0 * s, 0.0 * s, 1.0 * s};
    return m;
}
static double franke_f(vec2 X) { return -franke_lap(X); }
static vec2 franke_gradu(vec2 X) { vec2 r = {franke_dudx(X), franke_dudy(X)}; return r; }
/* F = -sigma^T grad u (BVPs/Franke_Poisson.hpp:67-73) */
static vec2 franke_F(vec2 X) {
    vec2 g = {franke_dudx(X), franke_dudy(X)};
    mat2 s = franke_sigma(X);
    vec2 r = {-(s.a00 * g.x + s.a10 * g.y), -(s.a01 * g.x + s.a11 * g.y)};
    return r;
}

/* Robin data for reflecting sides with the Franke solution: varphi = -0.5 and
 * psi(X, N) = grad u(X).N + 0.5 u(X), so that du/dN - varphi u = psi holds on
 * every side (the reflected Feynman–Kac representation needs exactly this
 * boundary relation).  Pinned by: tests/test_reflect.py (the estimator of
 * reflected paths converges to the Franke u). */
static double robin_psi(vec2 X, vec2 N) {
    return franke_dudx(X) * N.x + franke_dudy(X) * N.y + 0.5 * franke_u(X);
}
static double robin_varphi(vec2 X, vec2 N) { (void)X; (void)N; return -0.5; }

/* ----------------------------------------- variable-coefficient Franke --
 * Operator L = 1/2 tr(sigma sigma^T D2) + b.grad + c with sigma = s(x) I,
 * s = 1 + 0.5 sin(pi x) cos(pi y), b = 0.3 (y, -x), c = -1 - x^2; the
 * manufactured solution is Franke u, so f = -(L u) and g = u (the sign
 * convention L u + f = 0 of BVPs/Franke_Poisson.hpp:58-60,
 * BVPs/Monegros_Revamped.hpp:58).  Pinned by: tests/test_oracle.py (L u + f = 0
 * by finite differences). */
static double vc_s(vec2 X) { return 1.0 + 0.5 * spac_sinpi(X.x) * spac_cospi(X.y); }
static mat2 vc_sigma(vec2 X) {
    double s = vc_s(X);
    mat2 m = {1.0 * s, 0.0 * s, 0.0 * s, 1.0 * s};
    return m;
}
static vec2 vc_b(vec2 X) { vec2 r = {0.3 * X.y, -0.3 * X.x}; return r; }
static double vc_c(vec2 X) { return -1.0 - X.x * X.x; }
static double vc_f(vec2 X) {
    double s = vc_s(X);
    double lap = franke_lap(X);
    vec2 b = vc_b(X);
    double adv = b.x * franke_dudx(X) + b.y * franke_dudy(X);
    double reac = vc_c(X) * franke_u(X);
    double diff = (0.5 * (s * s)) * lap;
    return -((diff + adv) + reac);
}
static vec2 vc_F(vec2 X) {
    vec2 g = {franke_dudx(X), franke_dudy(X)};
    mat2 s = vc_sigma(X);
    vec2 r = {-(s.a00 * g.x + s.a10 * g.y), -(s.a01 * g.x + s.a11 * g.y)};
    return r;
}

/* ----------------------------------------------------------- boundaries --
 * Rectangulo: Dominios/Rectangulo.hpp:7-229, the quadrant tree restated
 * branch for branch (including the float literal 0.70710678118f of the corner
 * normals, :49-50,100-101,153-154,205-206).  Pinned by: tests/test_oracle.py
 * (brute-force signed distance to the rectangle outline). */
static double rectangulo(const double* params, vec2 position, vec2* exitpoint, vec2* normal) {
    const double K = (double)0.70710678118f;
    double center[2] = {0.5 * (params[0] + params[2]), 0.5 * (params[1] + params[3])};
    double halfside[2] = {params[2] - center[0], params[3] - center[1]};
    double distances[2] = {fabs(position.x - center[0]), fabs(position.y - center[1])};
    double r;
    *exitpoint = position;
    int inside = (distances[0] <= halfside[0]) && (distances[1] <= halfside[1]);
    int right = position.x > center[0];
    int top = position.y > center[1];
    double sx = right ? 1.0 : -1.0;
    double sy = top ? 1.0 : -1.0;
    if (inside) {
        if (distances[0] - halfside[0] > distances[1] - halfside[1]) {
            exitpoint->x = sx * halfside[0] + center[0];
            normal->x = sx; normal->y = 0.0;
            return distances[0] - halfside[0];
        }
        exitpoint->y = sy * halfside[1] + center[1];
        normal->x = 0.0; normal->y = sy;
        return distances[1] - halfside[1];
    }
    if (distances[0] > halfside[0]) {
        if (distances[1] > halfside[1]) {
            r = sqrt((distances[0] - halfside[0]) * (distances[0] - halfside[0]) +
                     (distances[1] - halfside[1]) * (distances[1] - halfside[1]));
            exitpoint->x = sx * halfside[0] + center[0];
            exitpoint->y = sy * halfside[1] + center[1];
            normal->x = sx * K; normal->y = sy * K;
            return r;
        }
        exitpoint->x = sx * halfside[0] + center[0];
        normal->x = sx; normal->y = 0.0;
        return distances[0] - halfside[0];
    }
    exitpoint->y = sy * halfside[1] + center[1];
    normal->x = 0.0; normal->y = sy;
    return distances[1] - halfside[1];
}

/* Circulo: Dominios/Circulo.hpp:7-22 (params R, cx, cy; radius clamp 5e-4). */
static double circulo(const double* p, vec2 posicion, vec2* exitpoint, vec2* normal) {
    normal->x = posicion.x - p[1];
    normal->y = posicion.y - p[2];
    double r = sqrt(normal->x * normal->x + normal->y * normal->y);
    r = (5E-4 < r) ? r : 5E-4; /* std::max(5E-4, r) */
    const double ir = 1.0 / r;   /* (X - C) / r, as one reciprocal and two products */
    normal->x = normal->x * ir;
    normal->y = normal->y * ir;
    exitpoint->x = normal->x * p[0] + p[1];
    exitpoint->y = normal->y * p[0] + p[2];
    return r - p[0];
}

/* No subdomain: never the nearest boundary (the reference's default frontier
 * returns distance 0 and would stop every path after one step,
 * BVP/Frontera.hpp:12-16). */
static double ninguna(const double* p, vec2 x, vec2* exitpoint, vec2* normal) {
    (void)p;
    *exitpoint = x;
    normal->x = 0.0; normal->y = 0.0;
    return -INFINITY;
}

static void frontera_init(frontera_t* f, int kind, const double* params) {
    switch (kind) {
    case SPAC_BOUNDARY_RECTANGLE: f->distancia = rectangulo; break;
    case SPAC_BOUNDARY_CIRCLE: f->distancia = circulo; break;
    default: f->distancia = ninguna; break;
    }
    for (int i = 0; i < 4; ++i) f->params[i] = params ? params[i] : 0.0;
}

/* ------------------------------------------------- tabulated coefficients --
 * The numeric coefficients of BVP/Coeficiente_*_Numerico.hpp: Evalua reads an interpolated
 * table (spac_tab.h, shared with the path kernels) instead of a formula.  Process-wide, set
 * with oracle_set_coef_tables; replaces, in the variable-coefficient set, the coefficients the
 * mask names.  Pinned by: tests/test_coef_tables.py (bilinear exactness on linear data, a
 * table of the compiled coefficients reproduces the compiled records within interpolation
 * error). */
static spac_coef_table g_tab = {NULL, 0, 0, 0.0, 0.0, 0.0, 0.0, 0, 0};

int oracle_set_coef_tables(int mask, const double* values, int t, const double* dom) {
    g_tab.mask = values ? mask : 0;
    g_tab.v = values;
    g_tab.t = t;
    if (values) {
        g_tab.lox = dom[0];
        g_tab.loy = dom[1];
        g_tab.ix = (double)(t - 1) / (dom[2] - dom[0]);
        g_tab.iy = (double)(t - 1) / (dom[3] - dom[1]);
    }
    return 0;
}
static mat2 tab_sigma(vec2 X) {
    double s = spac_tab_at(&g_tab, X.x, X.y, SPAC_TABC_S);
    mat2 m = {1.0 * s, 0.0 * s, 0.0 * s, 1.0 * s};
    return m;
}
static vec2 tab_b(vec2 X) {
    int32_t k00;
    double ax, ay;
    spac_tab_cell(&g_tab, X.x, X.y, &k00, &ax, &ay);
    vec2 r = {spac_tab_eval(&g_tab, k00, ax, ay, SPAC_TABC_BX), spac_tab_eval(&g_tab, k00, ax, ay, SPAC_TABC_BY)};
    return r;
}
static double tab_c(vec2 X) { return spac_tab_at(&g_tab, X.x, X.y, SPAC_TABC_C); }
static double tab_f(vec2 X) { return spac_tab_at(&g_tab, X.x, X.y, SPAC_TABC_F); }
static double tab_g(vec2 X) { return spac_tab_at(&g_tab, X.x, X.y, SPAC_TABC_G); }
static double tab_u(vec2 X) { return spac_tab_at(&g_tab, X.x, X.y, SPAC_TABC_U); }

static int bvp_init(bvp_t* bvp, const spac_problem_desc* d) {
    bvp->psi = escalarN_defecto;
    bvp->varphi = escalarN_defecto;
    bvp->mu = vector_defecto;
    switch (d->coef_kind) {
    case SPAC_COEF_FRANKE_POISSON:
        bvp->u = franke_u; bvp->g = franke_u; bvp->c = franke_c; bvp->f = franke_f;
        bvp->b = franke_b; bvp->gradiente_u = franke_gradu; bvp->F = franke_F;
        bvp->sigma = franke_sigma;
        break;
    case SPAC_COEF_VARCOEF_FRANKE:
        bvp->u = franke_u; bvp->g = franke_u; bvp->c = vc_c; bvp->f = vc_f;
        bvp->b = vc_b; bvp->gradiente_u = franke_gradu; bvp->F = vc_F;
        bvp->sigma = vc_sigma;
        break;
    case SPAC_COEF_FRANKE_ROBIN:
        bvp->u = franke_u; bvp->g = franke_u; bvp->c = franke_c; bvp->f = franke_f;
        bvp->b = franke_b; bvp->gradiente_u = franke_gradu; bvp->F = franke_F;
        bvp->sigma = franke_sigma;
        bvp->psi = robin_psi; bvp->varphi = robin_varphi;
        break;
    default:
        return SPAC_E_UNREGISTERED_PLUGIN;
    }
    if (g_tab.mask) {
        if (d->coef_kind != SPAC_COEF_VARCOEF_FRANKE || d->reflecting) return SPAC_E_INVALID_ARG;
        if (g_tab.mask & SPAC_TAB_SIGMA) bvp->sigma = tab_sigma;
        if (g_tab.mask & SPAC_TAB_DRIFT) bvp->b = tab_b;
        if (g_tab.mask & SPAC_TAB_KILLING) bvp->c = tab_c;
        if (g_tab.mask & SPAC_TAB_SOURCE) bvp->f = tab_f;
        if (g_tab.mask & SPAC_TAB_DIRICHLET) bvp->g = tab_g;
        if (g_tab.mask & SPAC_TAB_SOLUTION) bvp->u = tab_u;
    }
    if (d->reflecting & ~15) return SPAC_E_INVALID_ARG;
    if (d->reflecting && d->domain_kind != SPAC_BOUNDARY_RECTANGLE) return SPAC_E_INVALID_ARG;
    bvp->reflecting = d->reflecting;
    bvp->rho = d->reflect_rho;
    frontera_init(&bvp->frontera_dominio, d->domain_kind, d->domain_params);
    frontera_init(&bvp->frontera_subdominio, d->subdomain_kind, d->subdomain_params);
    (void)escalar_defecto;
    return SPAC_OK;
}

/* Frontera::Evalua_Absorbente (BVP/Frontera.hpp:23,34) for the rectangle with a
 * reflecting-side mask: the side, or the corner, is read off the outward normal
 * the distance plugin returns; a corner reflects only when both sides do. */
static int side_reflects(int mask, vec2 N) {
    int bits = 0;
    if (N.x < 0.0) bits |= SPAC_REFLECT_XMIN;
    if (N.x > 0.0) bits |= SPAC_REFLECT_XMAX;
    if (N.y < 0.0) bits |= SPAC_REFLECT_YMIN;
    if (N.y > 0.0) bits |= SPAC_REFLECT_YMAX;
    return bits != 0 && (bits & mask) == bits;
}

/* ----------------------------------------------------------- Dentro --
 * Solvers/FeynmanKacSolver.cpp:68-106.  Absorbing frontiers stop a path at the
 * Gobet-shifted band; a reflecting domain side (:88-90) keeps it, and a point
 * beyond it is reported as reflejada for the caller to project (:93-96).
 * *refl says whether the next step is the reflecting one (the reference's
 * ccabsorbentes == false). */
static int dentro(const bvp_t* bvp, vec2 X, double sqrth, double gobet, vec2* Npro, vec2* N,
                  double* distancia, int* refl) {
    vec2 Npro_aux, N_aux;
    double distancia_aux = bvp->frontera_dominio.distancia(bvp->frontera_dominio.params, X, &Npro_aux, &N_aux);
    double d = bvp->frontera_subdominio.distancia(bvp->frontera_subdominio.params, X, Npro, N);
    mat2 s;
    int status;
    *refl = 0;
    if (distancia_aux >= d) {
        d = distancia_aux;
        *Npro = Npro_aux;
        *N = N_aux;
        s = (d < 0.0) ? bvp->sigma(X) : bvp->sigma(*Npro);
        double r0 = N->x * s.a00 + N->y * s.a10;
        double r1 = N->x * s.a01 + N->y * s.a11;
        if (bvp->reflecting && side_reflects(bvp->reflecting, *N)) {
            *refl = 1;
            status = (d <= 0.0) ? SPAC_OUT_INSIDE : SPAC_OUT_REFLECTED;
        } else {
            status = spac_gobet_inside(d, gobet * sqrth, r0 * r0 + r1 * r1) ? SPAC_OUT_INSIDE : SPAC_OUT_DOMAIN;
        }
    } else {
        s = (d < 0.0) ? bvp->sigma(X) : bvp->sigma(*Npro);
        double r0 = N->x * s.a00 + N->y * s.a10;
        double r1 = N->x * s.a01 + N->y * s.a11;
        status = spac_gobet_inside(d, gobet * sqrth, r0 * r0 + r1 * r1) ? SPAC_OUT_INSIDE : SPAC_OUT_SUBDOMAIN;
    }
    *distancia = d;
    return status;
}

/* ------------------------------------------------- control variate --
 * The F of xi += Y F(X).dW (Solvers/FeynmanKacSolver.cpp:13): the plugin
 * (reference behaviour), none (plain Monte Carlo), or -sigma grad u_hat read
 * from a table with spac_cv.h (scalar diffusion sigma = s I).  In table mode
 * u_hat also replaces u in the subdomain-exit score, and the reduction may
 * drop xi from B where Var(score + xi) > max_ratio Var(score).
 * Pinned by: tests/test_oracle.py (table of the exact Franke u and gradient
 * reproduces the plugin's xi to interpolation accuracy; a zero table equals
 * plain Monte Carlo; the fallback rule on constructed records). */
static struct {
    int mode, fallback;
    double max_ratio;
    spac_cv_table tab;
} g_cv = {SPAC_CV_ANALYTIC, 0, 1.0, {NULL, 0, 0.0, 0.0, 0.0, 0.0}};

int oracle_set_cv(int mode, int fallback, double max_ratio, const float* table, int t, const double* dom) {
    if (mode == SPAC_CV_TABLE && (!table || t < 2 || !(dom[2] > dom[0]) || !(dom[3] > dom[1]))) return SPAC_E_INVALID_ARG;
    g_cv.mode = mode;
    g_cv.fallback = fallback;
    g_cv.max_ratio = max_ratio;
    if (mode == SPAC_CV_TABLE) {
        g_cv.tab.table = table;
        g_cv.tab.t = t;
        g_cv.tab.lox = dom[0];
        g_cv.tab.loy = dom[1];
        g_cv.tab.ix = (double)(t - 1) / (dom[2] - dom[0]);
        g_cv.tab.iy = (double)(t - 1) / (dom[3] - dom[1]);
    }
    return SPAC_OK;
}

/* xi += Y F(X).dW for the selected control variate */
static double cv_increment(const bvp_t* bvp, vec2 X, mat2 sig, double Y, vec2 dW) {
    if (g_cv.mode == SPAC_CV_ANALYTIC) {
        vec2 F = bvp->F(X);
        return Y * (F.x * dW.x + F.y * dW.y);
    }
    if (g_cv.mode == SPAC_CV_TABLE) {
        int32_t k00;
        double ax, ay;
        spac_cv_cell(&g_cv.tab, X.x, X.y, &k00, &ax, &ay);
        double Fx = -(sig.a00 * spac_cv_eval(&g_cv.tab, k00, ax, ay, 1));
        double Fy = -(sig.a00 * spac_cv_eval(&g_cv.tab, k00, ax, ay, 2));
        return Y * (Fx * dW.x + Fy * dW.y);
    }
    return 0.0;
}

/* --------------------------------------------------- Paso_Absorbente --
 * Solvers/FeynmanKacSolver.cpp:7-18, update order kept: Z, xi, Y, X, t.
 * ji_t is 0 for absorbing boundaries (:101-103). */
typedef struct { vec2 X, N; double Y, Z, xi, t; } path_state;

static void paso_absorbente(const bvp_t* bvp, path_state* st, double h, double sqrth, uint64_t seed,
                            uint32_t stream, uint32_t path, uint32_t step) {
    const double ji_t = 0.0;
    mat2 sig = bvp->sigma(st->X);
    double n0, n1;
    spac_normal_pair(seed, stream, path, step, &n0, &n1);
    vec2 dW = {sqrth * n0, sqrth * n1};
    double Y = st->Y;
    st->Z = st->Z + (bvp->f(st->X) * Y * h + bvp->psi(st->X, st->N) * Y * ji_t);
    if (g_cv.mode != SPAC_CV_NONE) st->xi = st->xi + cv_increment(bvp, st->X, sig, Y, dW);
    vec2 mu = bvp->mu(st->X);
    st->Y = Y + (bvp->c(st->X) * Y * h + Y * (mu.x * dW.x + mu.y * dW.y) +
                 bvp->varphi(st->X, st->N) * Y * ji_t);
    vec2 b = bvp->b(st->X);
    vec2 mu2 = bvp->mu(st->X);
    double smx = sig.a00 * mu2.x + sig.a01 * mu2.y;
    double smy = sig.a10 * mu2.x + sig.a11 * mu2.y;
    double sdx = sig.a00 * dW.x + sig.a01 * dW.y;
    double sdy = sig.a10 * dW.x + sig.a11 * dW.y;
    st->X.x = st->X.x + ((b.x - smx) * h + sdx);
    st->X.y = st->X.y + ((b.y - smy) * h + sdy);
    st->t = st->t + h;
}

/* -------------------------------------------------- Paso_Reflejante --
 * Solvers/FeynmanKacSolver.cpp:19-67, restated as the Lepingle scheme it
 * implements (D. Lepingle, Euler scheme for reflected stochastic differential
 * equations, Math. Comput. Simulation 38 (1995) 119-126).  With N the outward
 * normal of the nearest reflecting side and d <= 0 the signed distance of X:
 *   X' = X + (b - sigma mu) h + sigma dW;
 *   if d > -rho: the largest outward excursion of the step's Brownian bridge,
 *     nu = (uc + sqrt(|N^T sigma|^2 omega + uc^2)) / 2, uc = N.(X' - X),
 *     omega = 2 h E with E unit exponential (the reference's comment at :30:
 *     rate 1/(2h)), gives the local-time increment chi = max(0, nu + d) and
 *     X' -= chi N;
 *   a point still beyond a reflecting side is projected onto it and the
 *   overshoot is added to chi (:40-45, :51-56);
 *   Z += f Y h + psi(W, N) Y chi, xi += Y F.dW, Y += c Y h + Y mu.dW +
 *   varphi(W, N) Y chi, t += h (update order of :58-66), with the Robin
 *   data psi, varphi read at W = X - d N, the foot of X on the side: they are
 *   boundary data, and the reference's psi(X, N) at the interior point adds a
 *   bias of order rho.
 * The reference's defects are not reproduced: its rejection loop reads an
 * uninitialised point (:24,38), its d_k is the distance of an earlier point
 * (clamped to 0 when negative, :35; set from a stale value at :182), its
 * exponential has mean 1 instead of 2h, and it steps the clock by -h (:65).
 * chi enters with the overshoot of a reflejada Dentro, which the reference
 * drops.  Pinned by: tests/test_reflect.py (pure projection at rho = 0 against
 * a hand-computed walk; reflected estimator converges to the Franke u under
 * Robin data; unit-exponential moments of the draws). */
static void paso_reflejante(const bvp_t* bvp, path_state* st, double h, double sqrth, double d, double chi,
                            uint64_t seed, uint32_t stream, uint32_t path, uint32_t step) {
    mat2 sig = bvp->sigma(st->X);
    double n0, n1;
    spac_normal_pair(seed, stream, path, step, &n0, &n1);
    vec2 dW = {sqrth * n0, sqrth * n1};
    vec2 b = bvp->b(st->X);
    vec2 mu = bvp->mu(st->X);
    double smx = sig.a00 * mu.x + sig.a01 * mu.y;
    double smy = sig.a10 * mu.x + sig.a11 * mu.y;
    double sdx = sig.a00 * dW.x + sig.a01 * dW.y;
    double sdy = sig.a10 * dW.x + sig.a11 * dW.y;
    double drx = b.x - smx, dry = b.y - smy;
    vec2 Xp = {st->X.x + (drx * h + sdx), st->X.y + (dry * h + sdy)};
    if (d > -bvp->rho) {
        double omega = (2.0 * h) * spac_exponential(seed, stream, path, step);
        double r0 = st->N.x * sig.a00 + st->N.y * sig.a10;
        double r1 = st->N.x * sig.a01 + st->N.y * sig.a11;
        double uc = (r0 * dW.x + r1 * dW.y) + (st->N.x * drx + st->N.y * dry) * h;
        double nu = 0.5 * (uc + sqrt((r0 * r0 + r1 * r1) * omega + uc * uc));
        double v = nu + d;
        double chil = (v > 0.0) ? v : 0.0;
        Xp.x = Xp.x - chil * st->N.x;
        Xp.y = Xp.y - chil * st->N.y;
        chi = chi + chil;
    }
    vec2 Pp, Np;
    double dp = bvp->frontera_dominio.distancia(bvp->frontera_dominio.params, Xp, &Pp, &Np);
    if (dp > 0.0 && side_reflects(bvp->reflecting, Np)) {
        Xp = Pp;
        chi = chi + dp;
    }
    double Y = st->Y;
    vec2 W = {st->X.x - d * st->N.x, st->X.y - d * st->N.y}; /* foot of X on the side */
    st->Z = st->Z + (bvp->f(st->X) * Y * h + bvp->psi(W, st->N) * Y * chi);
    if (g_cv.mode != SPAC_CV_NONE) st->xi = st->xi + cv_increment(bvp, st->X, sig, Y, dW);
    st->Y = Y + (bvp->c(st->X) * Y * h + Y * (mu.x * dW.x + mu.y * dW.y) + bvp->varphi(W, st->N) * Y * chi);
    st->X = Xp;
    st->t = st->t + h;
}

/* ----------------------------------------------------- Simulacion_OMP --
 * Solvers/FeynmanKacSolver.cpp:137-197: initial Dentro, then do { step;
 * Dentro } while not stopped — at least one step even when X0 starts in the
 * stop band.  Records X_tau = Npro (the projected point). */
static void simulate_one(const bvp_t* bvp, vec2 X0, double h, double sqrth, double gobet, int max_steps,
                         uint64_t seed, uint32_t stream, uint32_t path, double* xt, double* yt,
                         double* Yo, double* Zo, double* xio, double* tauo, double* gvalo,
                         int32_t* stepso, int8_t* outo) {
    path_state st;
    st.X = X0; st.Y = 1.0; st.Z = 0.0; st.xi = 0.0; st.t = 0.0;
    vec2 Npro, N;
    double dist, chi = 0.0;
    int refl;
    int estado = dentro(bvp, st.X, sqrth, gobet, &Npro, &N, &dist, &refl);
    if (estado == SPAC_OUT_REFLECTED) { st.X = Npro; chi = dist; dist = 0.0; } /* :93-96 */
    st.N = N;
    int steps = 0;
    do {
        if (refl) paso_reflejante(bvp, &st, h, sqrth, dist, chi, seed, stream, path, (uint32_t)steps);
        else paso_absorbente(bvp, &st, h, sqrth, seed, stream, path, (uint32_t)steps);
        chi = 0.0;
        steps++;
        estado = dentro(bvp, st.X, sqrth, gobet, &Npro, &N, &dist, &refl);
        if (estado == SPAC_OUT_REFLECTED) { st.X = Npro; chi = dist; dist = 0.0; }
        st.N = N;
        if (estado != SPAC_OUT_DOMAIN && estado != SPAC_OUT_SUBDOMAIN && steps >= max_steps) {
            estado = SPAC_OUT_MAX_STEPS;
            break;
        }
    } while (estado != SPAC_OUT_DOMAIN && estado != SPAC_OUT_SUBDOMAIN);
    if (xt) *xt = Npro.x;
    if (yt) *yt = Npro.y;
    if (Yo) *Yo = st.Y;
    if (Zo) *Zo = st.Z;
    if (xio) *xio = st.xi;
    if (tauo) *tauo = st.t;
    if (gvalo) {
        if (estado != SPAC_OUT_SUBDOMAIN) {
            *gvalo = bvp->g(Npro);
        } else if (g_cv.mode == SPAC_CV_TABLE) {
            int32_t k00;
            double ax, ay;
            spac_cv_cell(&g_cv.tab, Npro.x, Npro.y, &k00, &ax, &ay);
            *gvalo = spac_cv_eval(&g_cv.tab, k00, ax, ay, 0);
        } else {
            *gvalo = bvp->u(Npro);
        }
    }
    if (stepso) *stepso = steps;
    if (outo) *outo = (int8_t)estado;
}

/* Simulate a batch of solves; records are laid out solve after solve. */
int oracle_simulate(const spac_problem_desc* desc, const spac_solve_batch* batch, uint64_t seed,
                    spac_path_records* rec, int nthreads) {
    bvp_t base;
    int rc = bvp_init(&base, desc);
    if (rc) return rc;
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(batch->n + 1));
    off[0] = 0;
    for (int k = 0; k < batch->n; ++k) off[k + 1] = off[k] + batch->paths[k];
    int max_steps = desc->max_steps > 0 ? desc->max_steps : 0x7fffffff;
    for (int k = 0; k < batch->n; ++k) {
        bvp_t bvp = base;
        if (batch->circle) {
            double cp[4] = {batch->circle[3 * k], batch->circle[3 * k + 1], batch->circle[3 * k + 2], 0.0};
            frontera_init(&bvp.frontera_subdominio, SPAC_BOUNDARY_CIRCLE, cp);
        }
        vec2 X0 = {batch->x0[k], batch->y0[k]};
        double h = batch->h[k];
        double sqrth = sqrt(h);
        int64_t np = batch->paths[k];
        int64_t o = off[k];
        uint32_t stream = batch->stream[k];
#ifdef _OPENMP
        if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
        for (int64_t p = 0; p < np; ++p) {
            int64_t g = o + p;
            simulate_one(&bvp, X0, h, sqrth, desc->gobet_constant, max_steps, seed, stream, (uint32_t)p,
                         rec->x_tau ? rec->x_tau + g : NULL, rec->y_tau ? rec->y_tau + g : NULL,
                         rec->Y ? rec->Y + g : NULL, rec->Z ? rec->Z + g : NULL,
                         rec->xi ? rec->xi + g : NULL, rec->tau ? rec->tau + g : NULL,
                         rec->gval ? rec->gval + g : NULL, rec->steps ? rec->steps + g : NULL,
                         rec->outcome ? rec->outcome + g : NULL);
        }
    }
    free(off);
    (void)nthreads;
    return SPAC_OK;
}

/* ------------------------------------------------------ Reduce_Analytic --
 * Solvers/FeynmanKacSolver.cpp:198-231 (point) and 233-292 (node) with the
 * fixed compensated order: 64 lane sums (lane l takes paths l, l+64, ... in
 * order) with TwoSum error terms, folded pairwise (lane i += lane i+off for
 * off = 32, 16, ..., 1).  Update(): :294-334. */
#define NSUMS 10
enum { S_SCORE, S_SCORE2, S_XI, S_XI2, S_VR, S_VR2, S_XISCORE, S_RNG, S_TAU, S_B };

void oracle_path_terms(int node_mode, double Y, double Z, double xi, double tau, double gval,
                       int32_t steps, int8_t outcome, double* v) {
    double score = Z + Y * gval;
    double vr = score + xi;
    v[S_SCORE] = score;
    v[S_SCORE2] = score * score;
    v[S_XI] = xi;
    v[S_XI2] = xi * xi;
    v[S_VR] = vr;
    v[S_VR2] = vr * vr;
    v[S_XISCORE] = xi * score;
    v[S_RNG] = 2.0 * (double)steps;
    v[S_TAU] = tau;
    if (node_mode) v[S_B] = (outcome == SPAC_OUT_SUBDOMAIN) ? (Z + xi) : vr;
    else v[S_B] = 0.0;
}

static void reduce_one(int node_mode, int64_t np, const double* Y, const double* Z, const double* xi,
                       const double* tau, const double* gval, const int32_t* steps, const int8_t* outcome,
                       spac_point_stats* st) {
    if (np == 0) { memset(st, 0, sizeof *st); return; }
    /* lane l (0..63) sums paths l, l + 64, ... in order, then the lane sums
     * are folded pairwise: lane i += lane i + off for off = 32, 16, ..., 1 */
    spac_ksum lane[NSUMS][64];
    for (int k = 0; k < NSUMS; ++k)
        for (int l = 0; l < 64; ++l) lane[k][l] = spac_ksum_zero();
    int64_t total_steps = 0, nmax = 0, ndom = 0;
    int has_nan = 0;
    for (int64_t p = 0; p < np; ++p) {
        double v[NSUMS];
        oracle_path_terms(node_mode, Y[p], Z[p], xi[p], tau[p], gval[p], steps[p], outcome[p], v);
        total_steps += steps[p];
        if (outcome[p] == SPAC_OUT_MAX_STEPS) nmax++;
        if (outcome[p] != SPAC_OUT_SUBDOMAIN) ndom++;
        for (int k = 0; k < NSUMS; ++k) lane[k][p % 64] = spac_ksum_add(lane[k][p % 64], v[k]);
    }
    for (int off = 32; off >= 1; off >>= 1)
        for (int l = 0; l < off; ++l)
            for (int k = 0; k < NSUMS; ++k) lane[k][l] = spac_ksum_merge(lane[k][l], lane[k][l + off]);
    spac_ksum tot[NSUMS];
    for (int k = 0; k < NSUMS; ++k) tot[k] = lane[k][0];
    double N = (double)np;
    double s[NSUMS];
    for (int k = 0; k < NSUMS; ++k) { s[k] = spac_ksum_value(tot[k]); if (s[k] != s[k]) has_nan = 1; }
    st->phi = s[S_SCORE] / N;
    st->phiphi = s[S_SCORE2] / N;
    st->xi = s[S_XI] / N;
    st->xixi = s[S_XI2] / N;
    st->phi_VR = s[S_VR] / N;
    st->phiphi_VR = s[S_VR2] / N;
    st->xiphi = s[S_XISCORE] / N;
    st->RNGC = s[S_RNG];
    st->APL = s[S_TAU] / N;
    st->B = node_mode ? s[S_B] / N : 0.0;
    st->paths = np;
    st->path_steps = total_steps;
    st->n_max_steps = nmax;
    st->n_domain_exits = (int32_t)ndom;
    st->status = has_nan ? SPAC_E_NAN : (nmax ? SPAC_E_MAX_STEPS : SPAC_OK);
    st->cv_fallback = 0;
    st->reserved = 0;
    if (g_cv.fallback && g_cv.mode == SPAC_CV_TABLE) {
        double var = st->phiphi - st->phi * st->phi;
        double var_vr = st->phiphi_VR - st->phi_VR * st->phi_VR;
        if (var_vr > g_cv.max_ratio * var) {
            st->cv_fallback = 1;
            if (node_mode) st->B = (s[S_B] - s[S_XI]) / N;
        }
    }
}

int oracle_reduce(int node_mode, const spac_solve_batch* batch, const spac_path_records* rec,
                  spac_point_stats* stats) {
    int64_t o = 0;
    for (int k = 0; k < batch->n; ++k) {
        int64_t np = batch->paths[k];
        reduce_one(node_mode, np, rec->Y + o, rec->Z + o, rec->xi + o, rec->tau + o, rec->gval + o,
                   rec->steps + o, rec->outcome + o, &stats[k]);
        o += np;
    }
    return SPAC_OK;
}

/* ------------------------------------------------------- math exports -- */
void oracle_math_eval(int fn, const double* x, const double* y, double* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        switch (fn) {
        case 0: out[i] = spac_exp(x[i]); break;
        case 1: out[i] = spac_log(x[i]); break;
        case 2: out[i] = spac_sinpi(x[i]); break;
        case 3: out[i] = spac_cospi(x[i]); break;
        case 4: out[i] = spac_atan2(x[i], y[i]); break;
        case 5: out[i] = sqrt(x[i]); break;
        case 6: out[i] = x[i] / y[i]; break;
        case 9: out[i] = spac_exp_finite(x[i]); break;
        case 10: out[i] = spac_log_normal(x[i]); break;
        case 11: out[i] = spac_gobet_inside(x[i], y[i], 1.7) ? 1.0 : 0.0; break; /* d, k; |N^T sigma|^2 = 1.7 */
        default: out[i] = 0.0; break;
        }
    }
}

void oracle_philox(int rounds, const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    spac_u32x4 r = spac_philox4x32_R(rounds, ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

void oracle_normals(uint64_t seed, uint32_t stream, uint32_t path, int32_t nsteps, double* out) {
    for (int32_t s = 0; s < nsteps; ++s) spac_normal_pair(seed, stream, path, (uint32_t)s, &out[2 * s], &out[2 * s + 1]);
}

void oracle_exponentials(uint64_t seed, uint32_t stream, uint32_t path, int32_t nsteps, double* out) {
    for (int32_t s = 0; s < nsteps; ++s) out[s] = spac_exponential(seed, stream, path, (uint32_t)s);
}

/* Coefficient probes for the plugin pins: which = 0 u, 1 f, 2 F.x, 3 F.y,
 * 4 sigma00, 5 b.x, 6 b.y, 7 c, 8 dudx, 9 dudy, 10 the Laplacian as the solver forms it,
 * 11 psi(X, (-1, 0)), 12 psi(X, (0, 1)), 13 varphi, 14 g, 15 d2udx2 + d2udy2 term by term */
int oracle_coef_eval(int coef_kind, int which, const double* x, const double* y, double* out, int64_t n) {
    spac_problem_desc d;
    memset(&d, 0, sizeof d);
    d.coef_kind = coef_kind;
    bvp_t bvp;
    if (bvp_init(&bvp, &d)) return SPAC_E_UNREGISTERED_PLUGIN;
    for (int64_t i = 0; i < n; ++i) {
        vec2 X = {x[i], y[i]};
        switch (which) {
        case 0: out[i] = bvp.u(X); break;
        case 1: out[i] = bvp.f(X); break;
        case 2: out[i] = bvp.F(X).x; break;
        case 3: out[i] = bvp.F(X).y; break;
        case 4: out[i] = bvp.sigma(X).a00; break;
        case 5: out[i] = bvp.b(X).x; break;
        case 6: out[i] = bvp.b(X).y; break;
        case 11: { vec2 N = {-1.0, 0.0}; out[i] = bvp.psi(X, N); break; }
        case 12: { vec2 N = {0.0, 1.0}; out[i] = bvp.psi(X, N); break; }
        case 13: { vec2 N = {1.0, 0.0}; out[i] = bvp.varphi(X, N); break; }
        case 7: out[i] = bvp.c(X); break;
        case 8: out[i] = franke_dudx(X); break;
        case 9: out[i] = franke_dudy(X); break;
        case 10: out[i] = franke_lap(X); break;
        case 15: out[i] = franke_d2udx2(X) + franke_d2udy2(X); break;
        case 14: out[i] = bvp.g(X); break;
        default: out[i] = 0.0; break;
        }
    }
    return SPAC_OK;
}

/* Boundary probes: kind SPAC_BOUNDARY_*; out = [dist, ex, ey, nx, ny] per point. */
void oracle_boundary_eval(int kind, const double* params, const double* x, const double* y, double* out, int64_t n) {
    frontera_t f;
    frontera_init(&f, kind, params);
    for (int64_t i = 0; i < n; ++i) {
        vec2 X = {x[i], y[i]}, e, nn;
        double d = f.distancia(f.params, X, &e, &nn);
        out[5 * i] = d; out[5 * i + 1] = e.x; out[5 * i + 2] = e.y; out[5 * i + 3] = nn.x; out[5 * i + 4] = nn.y;
    }
}

/* ------------------------------------------------------ interface rows --
 * Solve_OMP_Analytic(Nudo&, ...) for a batch of nodes
 * (Solvers/FeynmanKacSolver.cpp:233-292, 324-355).  G per node, column by
 * column, as a sequential TwoSum-compensated sum over the node's paths in
 * path order:
 *   IMQ (moment form):  S_i = sum_p Y_p phi(theta_p, theta_i),
 *                       G_j = -(sum_i iPsi(i,j) S_i) / N
 *   IMQ (direct form, direct != 0, Mesh/Interfaz.cpp:253-261):
 *                       G_j = (sum_p sum_i -Y_p iPsi(i,j) phi(theta_p, theta_i)) / N
 *   Gaussian stencils (moment form, spac_rbf.h): fixed-point moments
 *                       Phi[s][i] = sum_p Y_p phi_i(theta_p) by stencil start s, then
 *                       G_c = -(sum_{r, s = c - r} sum_i iPsi_s(i, r) Phi[s][i]) / N
 *   Gaussian stencils (direct form, direct != 0, Mesh/Interfaz.cpp:253-261):
 *                       G_j = -(sum_p Y_p H_j(theta_p)) / N over the paths whose
 *                       stencil holds column j, sixteen TwoSum partials.
 * Pinned by: tests/test_rows.py (cardinality of the interpolation tables,
 * consistency sum_j u(x_j) H_j(theta) ~ u(X) on the arcs, moment form = direct
 * form within cond(Psi) eps, G rows summing to -E[Y 1{circle exit}]). */
int oracle_assemble_rows(const spac_problem_desc* desc, const spac_interface_table* t, const spac_node_batch* nb,
                         uint64_t seed, int nthreads, int direct, double* B, double* G, spac_point_stats* stats) {
    int n = nb->n;
    uint32_t* paths = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    double* circle = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    off[0] = 0;
    for (int k = 0; k < n; ++k) {
        int f = nb->iface[k];
        paths[k] = nb->frontera[k] ? 0u : nb->paths[k];
        circle[3 * k] = t->radius[f]; circle[3 * k + 1] = t->cx[f]; circle[3 * k + 2] = t->cy[f];
        off[k + 1] = off[k] + paths[k];
    }
    int64_t tp = off[n] > 0 ? off[n] : 1;
    spac_path_records rec;
    rec.x_tau = (double*)malloc(8 * (size_t)tp); rec.y_tau = (double*)malloc(8 * (size_t)tp);
    rec.Y = (double*)malloc(8 * (size_t)tp); rec.Z = (double*)malloc(8 * (size_t)tp);
    rec.xi = (double*)malloc(8 * (size_t)tp); rec.tau = (double*)malloc(8 * (size_t)tp);
    rec.gval = (double*)malloc(8 * (size_t)tp); rec.steps = (int32_t*)malloc(4 * (size_t)tp);
    rec.outcome = (int8_t*)malloc((size_t)tp);
    spac_solve_batch b;
    memset(&b, 0, sizeof b);
    b.n = n; b.x0 = nb->x; b.y0 = nb->y; b.paths = paths; b.h = nb->h; b.stream = nb->stream; b.circle = circle;
    int rc = oracle_simulate(desc, &b, seed, &rec, nthreads);
    if (rc == 0) rc = oracle_reduce(1, &b, &rec, stats);
    bvp_t bvp;
    if (rc == 0) rc = bvp_init(&bvp, desc);
    int64_t goff = 0;
    for (int k = 0; k < n && rc == 0; ++k) {
        int f = nb->iface[k];
        int nbase = t->node_offset[f], M = t->node_offset[f + 1] - nbase;
        const double* th = t->node_theta + nbase;
        double* Gk = G + goff;
        goff += M;
        if (nb->frontera[k]) {
            vec2 X = {nb->x[k], nb->y[k]};
            B[k] = bvp.g(X);
            for (int j = 0; j < M; ++j) Gk[j] = 0.0;
            continue;
        }
        B[k] = stats[k].B;
        int code = t->code[f];
        double cx = t->cx[f], cy = t->cy[f], c2 = t->c2[f];
        const double* ipsi = t->ipsi + (t->ipsi_offset[f] >= 0 ? t->ipsi_offset[f] : 0);
        int64_t np = paths[k];
        int stencil = t->rbf_kind == SPAC_RBF_GAUSS_STENCIL;
        int S = stencil ? (t->stencil_size < M ? t->stencil_size : M) : 0;
        if (stencil && !direct) {
            /* moment form (spac_rbf.h): fixed-point Phi[s][i], then G_c = -(sum over r in
             * order, s = c - r, of the FMA chain over i of iPsi_s(i, r) Phi[s][i]) / N */
            double ymax = 0.0;
            for (int64_t p = 0; p < np; ++p) {
                int64_t g = off[k] + p;
                double ay = fabs(rec.Y[g]);
                if (rec.outcome[g] == SPAC_OUT_SUBDOMAIN && ay <= DBL_MAX && ay > ymax) ymax = ay;
            }
            const int q = spac_moment_scale((double)np, ymax);
            int64_t* mom = (int64_t*)calloc((size_t)M * S, sizeof(int64_t));
            double* momd = (double*)malloc(sizeof(double) * (size_t)M * S);
            for (int64_t p = 0; p < np; ++p) {
                int64_t g = off[k] + p;
                double Yp = rec.Y[g];
                if (rec.outcome[g] != SPAC_OUT_SUBDOMAIN || !(fabs(Yp) <= DBL_MAX)) continue;
                double theta = 0.0, phi[32];
                spac_exit_angle(code, rec.x_tau[g] - cx, rec.y_tau[g] - cy, th[0], &theta);
                int s0 = (code == 1) ? spac_circle_stencil(M, S, t->gauss_eps, theta, phi)
                                     : spac_arc_stencil(th, M, S, t->gauss_eps, theta, phi);
                for (int i = 0; i < S; ++i)
                    mom[(size_t)s0 * S + i] = (int64_t)((uint64_t)mom[(size_t)s0 * S + i] +
                                                        (uint64_t)spac_to_fixed(Yp * phi[i], q));
            }
            for (int i = 0; i < M * S; ++i) momd[i] = spac_from_fixed(mom[i], q);
            double N = (double)np;
            for (int c = 0; c < M; ++c) {
                double acc = 0.0;
                for (int r = 0; r < S; ++r) {
                    int s = c - r;
                    if (code == 1) { if (s < 0) s += M; }
                    else if (s < 0 || s > M - S) continue;
                    const double* blk = ipsi + ((code == 1) ? 0 : (int64_t)s * S * S);
                    for (int i = 0; i < S; ++i) acc = SPAC_FMA(blk[i * S + r], momd[(size_t)s * S + i], acc);
                }
                Gk[c] = (np > 0) ? -acc / N : 0.0;
            }
            free(mom);
            free(momd);
            continue;
        }
        /* Column sums in the fixed order of the GPU, paths in chunks of 256:
         * IMQ — partial w holds paths 64w..64w+63 of every chunk, four partials;
         * stencils — partial 4 (p % 256 / 64) + p % 4, sixteen partials; each
         * partial in path order, the partials merged in index order. */
        const int nparts = stencil ? 16 : 4;
        spac_ksum* part = (spac_ksum*)malloc(sizeof(spac_ksum) * 16 * (size_t)M);
        spac_ksum* acc4[16];
        for (int w = 0; w < nparts; ++w) acc4[w] = part + (size_t)w * M;
        double* dsum = (double*)calloc((size_t)M, sizeof(double));
        for (int j = 0; j < 16 * M; ++j) part[j] = spac_ksum_zero();
        for (int64_t p = 0; p < np; ++p) {
            int64_t g = off[k] + p;
            if (rec.outcome[g] != SPAC_OUT_SUBDOMAIN) continue;
            spac_ksum* acc = stencil ? acc4[4 * ((p % 256) / 64) + p % 4] : acc4[(p % 256) / 64];
            double theta = 0.0;
            spac_exit_angle(code, rec.x_tau[g] - cx, rec.y_tau[g] - cy, th[0], &theta);
            double Yp = rec.Y[g];
            if (stencil) {
                double phi[32];
                int s0 = (code == 1) ? spac_circle_stencil(M, S, t->gauss_eps, theta, phi)
                                     : spac_arc_stencil(th, M, S, t->gauss_eps, theta, phi);
                const double* blk = ipsi + ((code == 1) ? 0 : (int64_t)s0 * S * S);
                for (int r = 0; r < S; ++r) {
                    double w = 0.0;
                    for (int i = 0; i < S; ++i) w = SPAC_FMA(blk[i * S + r], phi[i], w);
                    int j = s0 + r;
                    if (j >= M) j -= M;
                    acc[j] = spac_ksum_add(acc[j], Yp * w);
                }
            } else if (direct) {
                for (int j = 0; j < M; ++j)
                    for (int i = 0; i < M; ++i) dsum[j] += -Yp * ipsi[(int64_t)i * M + j] * spac_imq(theta, th[i], c2);
            } else {
                for (int i = 0; i < M; ++i) acc[i] = spac_ksum_add(acc[i], Yp * spac_imq(theta, th[i], c2));
            }
        }
        double N = (double)np;
        spac_ksum* acc = acc4[0];
        for (int j = 0; j < M; ++j)
            for (int w = 1; w < nparts; ++w) acc[j] = spac_ksum_merge(acc[j], acc4[w][j]);
        if (stencil) {
            for (int j = 0; j < M; ++j) Gk[j] = -spac_ksum_value(acc[j]) / N;
        } else if (direct) {
            for (int j = 0; j < M; ++j) Gk[j] = dsum[j] / N;
        } else {
            for (int j = 0; j < M; ++j) {
                double w = 0.0;
                for (int i = 0; i < M; ++i) w = w + ipsi[(int64_t)i * M + j] * spac_ksum_value(acc[i]);
                Gk[j] = -w / N;
            }
        }
        free(part);
        free(dsum);
    }
    free(rec.x_tau); free(rec.y_tau); free(rec.Y); free(rec.Z); free(rec.xi); free(rec.tau); free(rec.gval);
    free(rec.steps); free(rec.outcome);
    free(paths); free(circle); free(off);
    return rc;
}

/* fixed-point moment probes (spac_rbf.h) */
int oracle_moment_scale(double n_paths, double ymax) { return spac_moment_scale(n_paths, ymax); }
int64_t oracle_to_fixed(double v, int q) { return spac_to_fixed(v, q); }
double oracle_from_fixed(int64_t n, int q) { return spac_from_fixed(n, q); }

/* exit angle probe (spac_exit_angle) */
int oracle_exit_angle(int code, double dx, double dy, double first, double* theta) {
    return spac_exit_angle(code, dx, dy, first, theta);
}

/* Row weights H_j(theta) of interface f for an exit angle theta (the table
 * the G rows apply per path); out has the interface's node count. */
int oracle_row_weights(const spac_interface_table* t, int f, double theta, double* out) {
    int nbase = t->node_offset[f], M = t->node_offset[f + 1] - nbase;
    const double* th = t->node_theta + nbase;
    const double* ipsi = t->ipsi + (t->ipsi_offset[f] >= 0 ? t->ipsi_offset[f] : 0);
    int code = t->code[f];
    for (int j = 0; j < M; ++j) out[j] = 0.0;
    if (t->rbf_kind == SPAC_RBF_GAUSS_STENCIL) {
        int S = t->stencil_size < M ? t->stencil_size : M;
        double phi[32];
        if (S > 32) return SPAC_E_INVALID_ARG;
        int s0 = (code == 1) ? spac_circle_stencil(M, S, t->gauss_eps, theta, phi)
                             : spac_arc_stencil(th, M, S, t->gauss_eps, theta, phi);
        const double* blk = ipsi + ((code == 1) ? 0 : (int64_t)s0 * S * S);
        for (int r = 0; r < S; ++r) {
            double w = 0.0;
            for (int i = 0; i < S; ++i) w = SPAC_FMA(blk[i * S + r], phi[i], w);
            out[(s0 + r) % M] = w;
        }
        return SPAC_OK;
    }
    if (t->ipsi_offset[f] < 0) return SPAC_E_BAD_INTERFACE;
    for (int j = 0; j < M; ++j) {
        double w = 0.0;
        for (int i = 0; i < M; ++i) w = w + ipsi[(int64_t)i * M + j] * spac_imq(theta, th[i], t->c2[f]);
        out[j] = w;
    }
    return SPAC_OK;
}
