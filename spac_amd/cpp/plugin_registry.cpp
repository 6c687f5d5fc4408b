/*
 * plugin_registry.cpp — host-pointer -> device-twin registry, the shipped
 * registrations, and the host boundary plugins.
 */
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/spac_fk.h"
#include "BVP.hpp"
#include "BVPs.hpp"
#include "Dominios.hpp"

namespace spac {

double Escalar_Defecto(Vec2) { return 0.0; }
double EscalarN_Defecto(Vec2, Vec2) { return 0.0; }
Vec2 Vector_Defecto(Vec2 x) { return x * 0.0; }
Mat2 Matriz_Defecto(Vec2) { return Mat2::Identity() * 0.0; }
double Distancia_Defecto(std::vector<double>, Vec2 x, Vec2& normal, Vec2& proyeccion) {
    normal = x * 0.0;
    proyeccion = normal;
    return 0.0;
}
bool Tipo_Defecto(Vec2) { return true; }

static std::mutex& reg_mutex() {
    static std::mutex m;
    return m;
}
static std::vector<CoefSetEntry>& coef_sets() {
    static std::vector<CoefSetEntry> v;
    return v;
}
static std::vector<BoundaryEntry>& boundaries() {
    static std::vector<BoundaryEntry> v;
    return v;
}

bool register_coefficient_set(const CoefSetEntry& e) {
    std::lock_guard<std::mutex> lk(reg_mutex());
    coef_sets().push_back(e);
    return true;
}
bool register_boundary(const BoundaryEntry& e) {
    std::lock_guard<std::mutex> lk(reg_mutex());
    boundaries().push_back(e);
    return true;
}

bool coefficient_set_of_kind(int kind, CoefSetEntry* out) {
    std::lock_guard<std::mutex> lk(reg_mutex());
    for (const CoefSetEntry& e : coef_sets())
        if (e.kind == kind) {
            *out = e;
            return true;
        }
    return false;
}

int resolve_coefficients(const BVP& bvp, int* kind) {
    /* a numeric coefficient has no compiled twin, whatever its function pointer says */
    if (bvp.u.Es_Numerico() || bvp.g.Es_Numerico() || bvp.c.Es_Numerico() || bvp.f.Es_Numerico() ||
        bvp.b.Es_Numerico() || bvp.F.Es_Numerico() || bvp.mu.Es_Numerico() || bvp.sigma.Es_Numerico() ||
        bvp.psi.Es_Numerico() || bvp.varphi.Es_Numerico())
        return SPAC_E_UNREGISTERED_PLUGIN;
    std::lock_guard<std::mutex> lk(reg_mutex());
    for (const CoefSetEntry& e : coef_sets()) {
        if (bvp.u.funcion == e.u && bvp.g.funcion == e.g && bvp.c.funcion == e.c && bvp.f.funcion == e.f &&
            bvp.b.funcion == e.b && bvp.F.funcion == e.F && bvp.mu.funcion == e.mu && bvp.sigma.funcion == e.sigma &&
            bvp.psi.funcion == e.psi && bvp.varphi.funcion == e.varphi) {
            *kind = e.kind;
            return SPAC_OK;
        }
    }
    return SPAC_E_UNREGISTERED_PLUGIN;
}

int resolve_numeric_coefficients(const BVP& bvp, int* kind, int* tab_mask, bool* F_zero) {
    *tab_mask = 0;
    *F_zero = false;
    const int mask = (bvp.sigma.Es_Numerico() ? SPAC_TAB_SIGMA : 0) | (bvp.b.Es_Numerico() ? SPAC_TAB_DRIFT : 0) |
                     (bvp.c.Es_Numerico() ? SPAC_TAB_KILLING : 0) | (bvp.f.Es_Numerico() ? SPAC_TAB_SOURCE : 0) |
                     (bvp.g.Es_Numerico() ? SPAC_TAB_DIRICHLET : 0) | (bvp.u.Es_Numerico() ? SPAC_TAB_SOLUTION : 0);
    if (bvp.F.Es_Numerico() || bvp.mu.Es_Numerico()) return SPAC_E_UNREGISTERED_PLUGIN;
    if (!mask && !bvp.psi.Es_Numerico() && !bvp.varphi.Es_Numerico()) return resolve_coefficients(bvp, kind);
    CoefSetEntry e;
    if (!coefficient_set_of_kind(SPAC_COEF_VARCOEF_FRANKE, &e)) return SPAC_E_UNREGISTERED_PLUGIN;
    const bool ok = (bvp.u.Es_Numerico() || bvp.u.funcion == e.u) && (bvp.g.Es_Numerico() || bvp.g.funcion == e.g) &&
                    (bvp.c.Es_Numerico() || bvp.c.funcion == e.c) && (bvp.f.Es_Numerico() || bvp.f.funcion == e.f) &&
                    (bvp.b.Es_Numerico() || bvp.b.funcion == e.b) &&
                    (bvp.sigma.Es_Numerico() || bvp.sigma.funcion == e.sigma) &&
                    (bvp.F.funcion == e.F || bvp.F.funcion == Vector_Defecto) && bvp.mu.funcion == e.mu &&
                    (bvp.psi.Es_Numerico() || bvp.psi.funcion == e.psi) &&
                    (bvp.varphi.Es_Numerico() || bvp.varphi.funcion == e.varphi);
    if (!ok) return SPAC_E_UNREGISTERED_PLUGIN;
    *kind = e.kind;
    *tab_mask = mask;
    *F_zero = bvp.F.funcion != e.F;
    return SPAC_OK;
}

int resolve_boundary(const Frontera& fr, int* kind) {
    if (fr.Distancia == Distancia_Defecto) { *kind = SPAC_BOUNDARY_NONE; return SPAC_OK; }
    std::lock_guard<std::mutex> lk(reg_mutex());
    for (const BoundaryEntry& e : boundaries())
        if (fr.Distancia == e.distancia) { *kind = e.kind; return SPAC_OK; }
    return SPAC_E_UNREGISTERED_PLUGIN;
}

int resolve_reflecting(const Frontera& fr, int* mask) {
    *mask = 0;
    if (fr.Absorbente == Tipo_Defecto || fr.Absorbente == Stopping) return SPAC_OK;
    if (fr.Distancia != Rectangulo || fr.parametros.size() < 4) return SPAC_E_UNREGISTERED_PLUGIN;
    const double x0 = fr.parametros[0], y0 = fr.parametros[1], x1 = fr.parametros[2], y1 = fr.parametros[3];
    const int bits[4] = {SPAC_REFLECT_XMIN, SPAC_REFLECT_YMIN, SPAC_REFLECT_XMAX, SPAC_REFLECT_YMAX};
    for (int side = 0; side < 4; ++side) {
        int absorbing = 0;
        for (int k = 1; k <= 3; ++k) {
            const double t = 0.25 * k;
            Vec2 P = side == 0 ? Vec2(x0, y0 + t * (y1 - y0))
                   : side == 1 ? Vec2(x0 + t * (x1 - x0), y0)
                   : side == 2 ? Vec2(x1, y0 + t * (y1 - y0))
                               : Vec2(x0 + t * (x1 - x0), y1);
            absorbing += fr.Evalua_Absorbente(P) ? 1 : 0;
        }
        if (absorbing != 0 && absorbing != 3) return SPAC_E_UNREGISTERED_PLUGIN;
        if (absorbing == 0) *mask |= bits[side];
    }
    return SPAC_OK;
}

double Inverse_Multiquadric_RBF(double theta, double theta_j, double c2) {
    return 1.0 / std::sqrt(std::pow(theta - theta_j, 2) + c2);
}
/* BVP/BVP.cpp:25-29 */
double Cardinal_Fourier(double theta, double theta_j, int n) {
    double aux = 0.0;
    for (int k = -n; k < n; k++) aux += std::cos(k * (theta - theta_j));
    return aux / (2 * n);
}

/* Dominios/Rectangulo.hpp:7-229 (the four quadrants share one branch tree) */
double Rectangulo(std::vector<double> params, Vec2 position, Vec2& exitpoint, Vec2& normal) {
    const double K = (double)0.70710678118f;
    double center[2] = {0.5 * (params[0] + params[2]), 0.5 * (params[1] + params[3])};
    double halfside[2] = {params[2] - center[0], params[3] - center[1]};
    double d0 = std::fabs(position(0) - center[0]), d1 = std::fabs(position(1) - center[1]);
    double sx = position(0) > center[0] ? 1.0 : -1.0, sy = position(1) > center[1] ? 1.0 : -1.0;
    exitpoint = position;
    if (d0 <= halfside[0] && d1 <= halfside[1]) {
        if (d0 - halfside[0] > d1 - halfside[1]) {
            exitpoint(0) = sx * halfside[0] + center[0];
            normal = Vec2(sx, 0.0);
            return d0 - halfside[0];
        }
        exitpoint(1) = sy * halfside[1] + center[1];
        normal = Vec2(0.0, sy);
        return d1 - halfside[1];
    }
    if (d0 > halfside[0]) {
        if (d1 > halfside[1]) {
            exitpoint = Vec2(sx * halfside[0] + center[0], sy * halfside[1] + center[1]);
            normal = Vec2(sx * K, sy * K);
            return std::sqrt((d0 - halfside[0]) * (d0 - halfside[0]) + (d1 - halfside[1]) * (d1 - halfside[1]));
        }
        exitpoint(0) = sx * halfside[0] + center[0];
        normal = Vec2(sx, 0.0);
        return d0 - halfside[0];
    }
    exitpoint(1) = sy * halfside[1] + center[1];
    normal = Vec2(0.0, sy);
    return d1 - halfside[1];
}

bool Stopping(Vec2) { return true; }

/* Dominios/Circulo.hpp:7-22 */
double Circulo(std::vector<double> p, Vec2 posicion, Vec2& exitpoint, Vec2& normal) {
    normal = Vec2(posicion(0) - p[1], posicion(1) - p[2]);
    double r = normal.norm();
    r = (5E-4 < r) ? r : 5E-4;
    normal = normal / r;
    exitpoint = Vec2(normal(0) * p[0] + p[1], normal(1) * p[0] + p[2]);
    return r - p[0];
}

/* shipped registrations */
SPAC_REGISTER_COEFFICIENT_SET(SPAC_COEF_FRANKE_POISSON, Franke_u, Franke_g, Franke_c, Franke_f, Franke_b, Franke_F,
                              Vector_Defecto, Franke_sigma, EscalarN_Defecto, EscalarN_Defecto);
SPAC_REGISTER_COEFFICIENT_SET(SPAC_COEF_VARCOEF_FRANKE, Franke_u, Franke_g, VarCoef_c, VarCoef_f, VarCoef_b,
                              VarCoef_F, Vector_Defecto, VarCoef_sigma, EscalarN_Defecto, EscalarN_Defecto);
SPAC_REGISTER_COEFFICIENT_SET(SPAC_COEF_FRANKE_ROBIN, Franke_u, Franke_g, Franke_c, Franke_f, Franke_b, Franke_F,
                              Vector_Defecto, Franke_sigma, Robin_psi, Robin_varphi);
SPAC_REGISTER_BOUNDARY(SPAC_BOUNDARY_RECTANGLE, Rectangulo);
SPAC_REGISTER_BOUNDARY(SPAC_BOUNDARY_CIRCLE, Circulo);

} // namespace spac
