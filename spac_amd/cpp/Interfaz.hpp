/*
 * Interfaz.hpp — one circle of the PDDSparse lattice (Mesh/Interfaz.hpp:8-30).
 *
 * Besides the reference's global inverse-multiquadric interpolation over the
 * arc nodes (Calcula_Psi / Update_G, Mesh/Interfaz.cpp:206-264), an interface
 * can carry local Gaussian stencils (Calcula_Stencils): each exit angle is
 * interpolated from the stencil_size nodes nearest to it.  Full interior circles
 * (interior = 1), which the reference never interpolates (:248-250), are
 * supported in stencil mode with periodic angles.
 */
#ifndef SPAC_INTERFAZ_HPP
#define SPAC_INTERFAZ_HPP

#include <map>
#include <string>
#include <vector>

#include "Nudo.hpp"
#include "spac_la.hpp"

namespace spac {

class BVP;

/* Inverse multiquadric, BVP/BVP.cpp:22-24 */
inline double Inverse_Multiquadric(double theta, double theta_j, double c2) {
    double d = theta - theta_j;
    return 1.0 / std::sqrt(d * d + c2);
}

class Interfaz {
public:
    bool es_perimeter = false;
    std::vector<int> posicion_centro;
    int ubicacion = -1;
    int interior = -10; /* -11,-22,-33,-44 corners; -1,-2,-3,-4 sides; 1 full circle */
    Vec2 centro;
    double radio = 0.0, c2 = -1.0;
    std::vector<double> parametros_dominio;
    MatX Psi, iPsi;
    std::vector<Nudo> nudos_circunferencia; /* nodes on this circle */
    std::vector<Nudo> nudos_interior;       /* nodes of other circles inside this one */

    /* local Gaussian stencils */
    int stencil_size = 0;
    double gauss_eps = 0.0;
    std::vector<double> stencil_inverses; /* blocks of S*S, row-major, one per start (arcs) or one (circle) */

    /* Mesh/Interfaz.cpp:2-204 */
    int Inicia_Cuadrado(double RADIO, const std::vector<double>& PARAMETROS_DOMINIO, double theta_0,
                        const std::map<std::string, double>& c2map, Vec2 CENTRO, int indice_posicion_x,
                        int indice_posicion_y, int indice_maximo_x, int indice_maximo_y, int numero_nudos_total,
                        int primer_nudo);
    /* Mesh/Interfaz.cpp:206-217: Psi_ij = IMQ(theta_i, theta_j), iPsi = Psi^-1; returns |iPsi|_F |Psi|_F */
    double Calcula_Psi(double C2);
    /* Local Gaussian stencils; returns the largest |iPsi_s|_F |Psi_s|_F */
    double Calcula_Stencils(double eps, int S);
    /* Mesh/Interfaz.cpp:219-264 (direct form, host): G[col_j] += -Y sum_i iPsi(i,j) phi(theta, theta_i) */
    int Update_G(double Y, const Vec2& X, std::map<int, double>& G) const;
    /* the angle of a point on this circle in the interface's branch convention */
    bool Angulo(const Vec2& X, double* theta) const;
    void Resetea();
};

} // namespace spac

#endif
