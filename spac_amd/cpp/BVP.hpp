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

#include <memory>
#include <utility>
#include <vector>

#include "Interpolador.hpp"
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

/* A coefficient is either a function pointer (compiled plugin) or numeric: values of an
 * Interpolador (BVP/Coeficiente_Escalar_Numerico.hpp:6-19 and siblings).  The interpolator
 * is held by shared pointer in the base class, so a numeric coefficient assigned into a BVP
 * field (a copy of the base) stays numeric and keeps the derived interpolator, where the
 * reference's by-value Interpolador member would slice it to the base's NaN stubs.  In a
 * numeric vector or matrix an entry without an interpolator is 0.  FeynmanKacSolver samples
 * numeric coefficients onto the device tables (spac_fk_set_coef_tables). */
using InterpoladorPtr = std::shared_ptr<const Interpolador>;

class Coeficiente_Escalar {
public:
    pfescalar funcion = Escalar_Defecto;
    InterpoladorPtr interpolador;
    Coeficiente_Escalar() = default;
    explicit Coeficiente_Escalar(pfescalar f) : funcion(f) {}
    void Inicializa(pfescalar f) { funcion = f; interpolador.reset(); }
    double Evalua(Vec2 x) const { return interpolador ? interpolador->Interpola(x) : funcion(x); }
    bool Es_Numerico() const { return interpolador != nullptr; }
};

class Coeficiente_EscalarN {
public:
    pfescalarN funcion = EscalarN_Defecto;
    InterpoladorPtr interpolador; /* numeric: a value per point, the normal is not read */
    Coeficiente_EscalarN() = default;
    explicit Coeficiente_EscalarN(pfescalarN f) : funcion(f) {}
    void Inicializa(pfescalarN f) { funcion = f; interpolador.reset(); }
    double Evalua(Vec2 x, Vec2 n) const { return interpolador ? interpolador->Interpola(x) : funcion(x, n); }
    bool Es_Numerico() const { return interpolador != nullptr; }
};

class Coeficiente_Vectorial {
public:
    pfvector funcion = Vector_Defecto;
    InterpoladorPtr interpolador_x, interpolador_y;
    Coeficiente_Vectorial() = default;
    explicit Coeficiente_Vectorial(pfvector f) : funcion(f) {}
    void Inicializa(pfvector f) { funcion = f; interpolador_x.reset(); interpolador_y.reset(); }
    Vec2 Evalua(Vec2 x) const {
        if (!Es_Numerico()) return funcion(x);
        return Vec2(interpolador_x ? interpolador_x->Interpola(x) : 0.0,
                    interpolador_y ? interpolador_y->Interpola(x) : 0.0);
    }
    bool Es_Numerico() const { return interpolador_x || interpolador_y; }
};

class Coeficiente_Matricial {
public:
    pfmatriz funcion = Matriz_Defecto;
    InterpoladorPtr interpolador_11, interpolador_12, interpolador_21, interpolador_22;
    Coeficiente_Matricial() = default;
    explicit Coeficiente_Matricial(pfmatriz f) : funcion(f) {}
    void Inicializa(pfmatriz f) {
        funcion = f;
        interpolador_11.reset(); interpolador_12.reset(); interpolador_21.reset(); interpolador_22.reset();
    }
    Mat2 Evalua(Vec2 x) const {
        if (!Es_Numerico()) return funcion(x);
        Mat2 m;
        m(0, 0) = interpolador_11 ? interpolador_11->Interpola(x) : 0.0;
        m(0, 1) = interpolador_12 ? interpolador_12->Interpola(x) : 0.0;
        m(1, 0) = interpolador_21 ? interpolador_21->Interpola(x) : 0.0;
        m(1, 1) = interpolador_22 ? interpolador_22->Interpola(x) : 0.0;
        return m;
    }
    bool Es_Numerico() const { return interpolador_11 || interpolador_12 || interpolador_21 || interpolador_22; }
};

/* The reference's numeric classes (BVP/Coeficiente_*_Numerico.hpp): constructors that
 * install interpolators. */
class Coeficiente_Escalar_Numerico : public Coeficiente_Escalar {
public:
    Coeficiente_Escalar_Numerico() = default;
    explicit Coeficiente_Escalar_Numerico(InterpoladorPtr i) { Inicializa(std::move(i)); }
    void Inicializa(InterpoladorPtr i) { interpolador = std::move(i); }
};

class Coeficiente_EscalarN_Numerico : public Coeficiente_EscalarN {
public:
    Coeficiente_EscalarN_Numerico() = default;
    explicit Coeficiente_EscalarN_Numerico(InterpoladorPtr i) { Inicializa(std::move(i)); }
    void Inicializa(InterpoladorPtr i) { interpolador = std::move(i); }
};

class Coeficiente_Vectorial_Numerico : public Coeficiente_Vectorial {
public:
    Coeficiente_Vectorial_Numerico() = default;
    Coeficiente_Vectorial_Numerico(InterpoladorPtr ix, InterpoladorPtr iy) { Inicializa(std::move(ix), std::move(iy)); }
    void Inicializa(InterpoladorPtr ix, InterpoladorPtr iy) {
        interpolador_x = std::move(ix);
        interpolador_y = std::move(iy);
    }
};

class Coeficiente_Matricial_Numerico : public Coeficiente_Matricial {
public:
    Coeficiente_Matricial_Numerico() = default;
    Coeficiente_Matricial_Numerico(InterpoladorPtr i11, InterpoladorPtr i12, InterpoladorPtr i21,
                                   InterpoladorPtr i22) {
        Inicializa(std::move(i11), std::move(i12), std::move(i21), std::move(i22));
    }
    /* sigma = s(x) I: the form the device tables hold */
    explicit Coeficiente_Matricial_Numerico(const InterpoladorPtr& s) { Inicializa(s, nullptr, nullptr, s); }
    void Inicializa(InterpoladorPtr i11, InterpoladorPtr i12, InterpoladorPtr i21, InterpoladorPtr i22) {
        interpolador_11 = std::move(i11);
        interpolador_12 = std::move(i12);
        interpolador_21 = std::move(i21);
        interpolador_22 = std::move(i22);
    }
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
/* A BVP with numeric coefficients: the SPAC_TAB_* mask of its numeric sigma, b, c, f, g, u,
 * and the device kind of the compiled part, which must be a set the tables can replace
 * coefficients of (SPAC_COEF_VARCOEF_FRANKE): every coefficient that is not numeric must be
 * that set's, except F, which may also be the default (zero) -> *F_zero = true.  A numeric
 * F or mu has no device form (SPAC_E_UNREGISTERED_PLUGIN).  With no numeric coefficient this
 * is resolve_coefficients with *tab_mask = 0. */
int resolve_numeric_coefficients(const BVP& bvp, int* kind, int* tab_mask, bool* F_zero);
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
