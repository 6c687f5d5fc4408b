/*
 * Nudo.hpp — a node ("knot") of the interface mesh and one row of G u = B
 * (Mesh/Nudo.hpp:33-75), plus the angle helpers Atan0/90/180/270
 * (Mesh/Nudo.hpp:7-32) used for the arc branch cuts.
 */
#ifndef SPAC_NUDO_HPP
#define SPAC_NUDO_HPP

#include <cmath>
#include <map>
#include <vector>

#include "spac_la.hpp"

namespace spac {

/* atan2 shifted into [0, 2pi) */
inline double Atan0(const Vec2& p) {
    double aux = std::atan2(p(1), p(0));
    if (aux >= 0) return aux;
    return aux + 2.0 * M_PI;
}
/* atan2 shifted into (pi/2, 5pi/2] */
inline double Atan90(const Vec2& p) {
    double aux = std::atan2(p(1), p(0));
    if (aux <= 0.5 * M_PI) return 2.0 * M_PI + aux;
    return aux;
}
/* plain atan2, (-pi, pi] */
inline double Atan180(const Vec2& p) { return std::atan2(p(1), p(0)); }
/* atan2 shifted into (-pi/2, 3pi/2] */
inline double Atan270(const Vec2& p) {
    double aux = std::atan2(p(1), p(0));
    if (aux <= -0.5 * M_PI) return 2 * M_PI + aux;
    return aux;
}

class Nudo {
public:
    bool frontera = false;          /* node on the domain boundary: B = g(x) */
    int indice_global = 0;          /* row of G and B */
    int indice_local = 0;           /* index inside its interface */
    std::vector<int> indice_interfaz;
    std::vector<std::vector<int>> indice_subdominios;
    Vec2 posicion_cartesiana;
    double posicion_angular = 0.0;
    double B = 0.0, solucion = 0.0, error = 0.0, varianza = 0.0, coeficiente_Pearson = 0.0,
           error_estadistico = 0.0, bias = 0.0, c2 = 0.0;
    std::map<int, double> G;        /* COO row, sorted by global column */

    void Resetea() {
        indice_global = 0;
        indice_local = 0;
        indice_interfaz.clear();
        indice_subdominios.clear();
        posicion_cartesiana = Vec2();
        posicion_angular = 0.0;
        B = solucion = error = varianza = coeficiente_Pearson = error_estadistico = bias = c2 = 0.0;
        G.clear();
    }
};

} // namespace spac

#endif
