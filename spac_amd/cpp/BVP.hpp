/*
 * BVP.hpp — the plugin surface: coefficients, frontiers and the BVP bundle
 * (BVP/BVP.hpp:17-28, BVP/Coeficiente_*.hpp, BVP/Frontera.hpp:6-31), with the
 * reference's function-pointer typedefs over Vec2/Mat2.
 *
 * A GPU cannot call host function pointers, so every coefficient set and
 * boundary function that should run on the device is registered against its
 * device twin (SPAC_REGISTER_COEFFICIENT_SET / SPAC_REGISTER_BOUNDARY, see
 * plugin_registry.cpp).  Resolving a BVP whose pointers are not registered
 * fails with SPAC_E_UNREGISTERED_PLUGIN: there is no CPU fallback.
 */
#ifndef SPAC_BVP_HPP
#define SPAC_BVP_HPP

#include <vector>

#include "spac_la.hpp"

namespace spac {

typedef double (*pfescalar)(Vec2);
typedef double (*pfescalarN)(Vec2, Vec2);
typedef Vec2 (*pfvector)(Vec2);
typedef Mat2 (*pfmatriz)(Vec2);
typedef double (*pfdistancia)(std::vector<double>, Vec2, Vec2&, Vec2&);
typedef bool (*pftipo)(Vec2);
typedef double (*pRBF)(double, double, double);
typedef double (*pCardinalFourier)(double, double, int);

/* defaults (BVP/Coeficiente_Escalar.hpp:7-9 and siblings, BVP/Frontera.hpp:12-19);
 * out of line so that every caller sees the same function address */
double Escalar_Defecto(Vec2);
double EscalarN_Defecto(Vec2, Vec2);
Vec2 Vector_Defecto(Vec2 x);
Mat2 Matriz_Defecto(Vec2);
double Distancia_Defecto(std::vector<double>, Vec2 x, Vec2& normal, Vec2& proyeccion);
bool Tipo_Defecto(Vec2);

class Coeficiente_Escalar {
public:
    pfescalar funcion = Escalar_Defecto;
    Coeficiente_Escalar() = default;
    explicit Coeficiente_Escalar(pfescalar f) : funcion(f) {}
    void Inicializa(pfescalar f) { funcion = f; }
    double Evalua(Vec2 x) const { return funcion(x); }
    bool Es_Numerico() const { return false; }
};

class Coeficiente_EscalarN {
public:
    pfescalarN funcion = EscalarN_Defecto;
    Coeficiente_EscalarN() = default;
    explicit Coeficiente_EscalarN(pfescalarN f) : funcion(f) {}
    void Inicializa(pfescalarN f) { funcion = f; }
    double Evalua(Vec2 x, Vec2 n) const { return funcion(x, n); }
    bool Es_Numerico() const { return false; }
};

class Coeficiente_Vectorial {
public:
    pfvector funcion = Vector_Defecto;
    Coeficiente_Vectorial() = default;
    explicit Coeficiente_Vectorial(pfvector f) : funcion(f) {}
    void Inicializa(pfvector f) { funcion = f; }
    Vec2 Evalua(Vec2 x) const { return funcion(x); }
    bool Es_Numerico() const { return false; }
};

class Coeficiente_Matricial {
public:
    pfmatriz funcion = Matriz_Defecto;
    Coeficiente_Matricial() = default;
    explicit Coeficiente_Matricial(pfmatriz f) : funcion(f) {}
    void Inicializa(pfmatriz f) { funcion = f; }
    Mat2 Evalua(Vec2 x) const { return funcion(x); }
    bool Es_Numerico() const { return false; }
};

class Frontera {
public:
    pfdistancia Distancia = Distancia_Defecto;
    pftipo Absorbente = Tipo_Defecto, Neumann = Tipo_Defecto;
    std::vector<double> parametros;
    Frontera() = default;
    Frontera(std::vector<double> p, pfdistancia d, pftipo a, pftipo n) { Inicializa(p, d, a, n); }
    void Inicializa(std::vector<double> p, pfdistancia d, pftipo a, pftipo n) {
        parametros = p;
        Distancia = d;
        Absorbente = a;
        Neumann = n;
    }
    double Evalua_Distancia(Vec2 x, Vec2& punto_mas_cercano, Vec2& normal) const {
        return Distancia(parametros, x, punto_mas_cercano, normal);
    }
    bool Evalua_Absorbente(Vec2 x) const { return Absorbente(x); }
    bool Evalua_Neumann(Vec2 x) const { return Neumann(x); }
};

double Inverse_Multiquadric_RBF(double theta, double theta_j, double c2);
double Cardinal_Fourier(double theta, double theta_j, int n);

class BVP {
public:
    Coeficiente_Escalar u, g, c, f;
    Coeficiente_EscalarN psi, varphi;
    Coeficiente_Vectorial b, gradiente_u, F, mu;
    Coeficiente_Matricial sigma;
    Frontera frontera_dominio, frontera_subdominio;
    pRBF RBF = Inverse_Multiquadric_RBF;
    pCardinalFourier fcardinal_Fourier = Cardinal_Fourier;
    BVP() = default;
    void Reset() { *this = BVP(); }
};

/* ------------------------------------------------------ registry -- */
struct CoefSetEntry {
    int kind; /* SPAC_COEF_* */
    pfescalar u, g, c, f;
    pfvector b, F, mu;
    pfmatriz sigma;
    pfescalarN psi, varphi;
};
struct BoundaryEntry {
    int kind; /* SPAC_BOUNDARY_* */
    pfdistancia distancia;
};
bool register_coefficient_set(const CoefSetEntry& e);
bool register_boundary(const BoundaryEntry& e);
/* The host functions registered for a device coefficient kind (false: none). */
bool coefficient_set_of_kind(int kind, CoefSetEntry* out);
/* SPAC_OK and the device coefficient kind, or SPAC_E_UNREGISTERED_PLUGIN. */
int resolve_coefficients(const BVP& bvp, int* kind);
/* SPAC_OK and the device boundary kind, or SPAC_E_UNREGISTERED_PLUGIN.  A
 * frontier with no distance function (the default) resolves to NONE. */
int resolve_boundary(const Frontera& fr, int* kind);
/* The SPAC_REFLECT_* mask of a rectangle frontier: a side reflects where
 * Evalua_Absorbente is false.  The predicate is sampled at three interior
 * points of each side; SPAC_E_UNREGISTERED_PLUGIN if it changes along a side
 * (the device has no mixed sides). */
int resolve_reflecting(const Frontera& fr, int* mask);

} // namespace spac

#define SPAC_CONCAT2(a, b) a##b
#define SPAC_CONCAT(a, b) SPAC_CONCAT2(a, b)
/* Pair a host coefficient set with its device twin (kind = SPAC_COEF_*). */
#define SPAC_REGISTER_COEFFICIENT_SET(kind, u, g, c, f, b, F, mu, sigma, psi, varphi)                      \
    static const bool SPAC_CONCAT(spac_coef_reg_, __LINE__) =                                            \
        ::spac::register_coefficient_set(::spac::CoefSetEntry{kind, u, g, c, f, b, F, mu, sigma, psi, varphi})
/* Pair a host distance function with its device twin (kind = SPAC_BOUNDARY_*). */
#define SPAC_REGISTER_BOUNDARY(kind, fn) \
    static const bool SPAC_CONCAT(spac_bnd_reg_, __LINE__) = ::spac::register_boundary(::spac::BoundaryEntry{kind, fn})

#endif
