/*
 * Malla.hpp — the Nl x Nl lattice of overlapping circles on a square and its
 * interface system (Mesh/Malla.hpp:7-24, Mesh/Malla.cpp:12-112).
 */
#ifndef SPAC_MALLA_HPP
#define SPAC_MALLA_HPP

#include <map>
#include <string>
#include <vector>

#include "Interfaz.hpp"
#include "SistemaLineal.hpp"

namespace spac {

/* One FK row job: node `nudo` (a node of a neighbour circle) lies inside
 * interface `interfaz`; its row receives B and G over the interface's nodes. */
struct TrabajoFila {
    int interfaz;
    int nudo_global;
    Vec2 posicion;
    bool frontera;
};

class Malla {
public:
    std::vector<Interfaz> Interfaces;
    std::vector<double> parametros_dominio;
    double radio = 0, angulo_inicial = 0, superposicion = 1, distancia = 0;
    int nudos_por_circulo = 0, numero_subdominios_lado = 0, numero_total_nudos = 0;
    std::map<std::string, double> c2;
    std::map<std::string, int> vecinos;
    SistemaLineal sistema;

    /* Mesh/Malla.cpp:12-83; returns 0 or a negative status (bad parameters). */
    int Construye_Cuadrado(const std::vector<double>& PARAMETROS_DOMINIO, int NUMERO_SUBDOMINIOS_LADO,
                           double SUPERPOSICION, double ANGULO_INICIAL, int NUDOS_POR_CIRCULO,
                           const std::map<std::string, double>& C2);
    bool Es_vecino(int it_inter, const std::string& dir) const;
    /* FK rows in the reference's order (interfaces in lattice order, their interior
     * nodes in neighbour-scan order).  all_circles = false keeps the reference's
     * split: only perimeter circles, full circles being the pseudospectral solver's. */
    std::vector<TrabajoFila> Trabajos(bool all_circles) const;
    /* Output/knot_position.txt: "global x y u(x)" (Mesh/Malla.cpp:99-112) */
    bool Escribe_Posiciones(const std::string& path, double (*u)(double, double)) const;
};

} // namespace spac

#endif
