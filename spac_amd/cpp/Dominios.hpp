/*
 * Dominios.hpp — host boundary plugins with the reference's pfdistancia
 * signature: Rectangulo (Dominios/Rectangulo.hpp:7-229, params xmin, ymin,
 * xmax, ymax), Circulo (Dominios/Circulo.hpp:7-22, params R, cx, cy) and the
 * absorbing-type predicate Stopping (Dominios/Rectangulo.hpp:230-232).  Each is
 * registered against its device twin in plugin_registry.cpp.
 */
#ifndef SPAC_DOMINIOS_HPP
#define SPAC_DOMINIOS_HPP

#include <vector>

#include "spac_la.hpp"

namespace spac {

double Rectangulo(std::vector<double> params, Vec2 position, Vec2& exitpoint, Vec2& normal);
double Circulo(std::vector<double> parametros, Vec2 posicion, Vec2& exitpoint, Vec2& normal);
bool Stopping(Vec2);

} // namespace spac

#endif
