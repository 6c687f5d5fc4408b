/*
 * Malla.cpp — Construye_Cuadrado (Mesh/Malla.cpp:12-83): circle centres on the
 * lattice, nodes per circle, node-in-neighbour-circle scan (8 neighbours, in
 * the sorted key order of the reference's std::map), CSR pattern.
 */
#include "Malla.hpp"

#include <cmath>
#include <cstdio>

namespace spac {

int Malla::Construye_Cuadrado(const std::vector<double>& P, int Nl, double sup, double theta0, int Nt,
                              const std::map<std::string, double>& C2) {
    if (P.size() < 4 || Nl < 2 || Nt < 4 || !(P[2] > P[0]) || !(P[3] > P[1])) return -1;
    if (sup < 1.0 || sup > 2.0) return -1; /* the reference prints FATAL WARNING (:42-47) */
    Interfaces.clear();
    parametros_dominio.assign(P.begin(), P.begin() + 4);
    superposicion = sup;
    angulo_inicial = theta0;
    nudos_por_circulo = Nt;
    numero_subdominios_lado = Nl;
    c2 = C2;
    distancia = (P[2] - P[0]) / (Nl - 1);
    radio = superposicion * distancia / std::sqrt(2.0);
    int nudo_centinela = 0;
    for (int x = 0; x <= Nl - 1; x++) {
        for (int y = 0; y <= Nl - 1; y++) {
            Interfaz it;
            Vec2 c(P[0] + x * distancia, P[1] + y * distancia);
            nudo_centinela = it.Inicia_Cuadrado(radio, parametros_dominio, angulo_inicial, c2, c, x, y, Nl - 1, Nl - 1,
                                                Nt, nudo_centinela);
            it.es_perimeter = (x * y == 0 || x == Nl - 1 || y == Nl - 1);
            Interfaces.push_back(it);
        }
    }
    numero_total_nudos = nudo_centinela;
    sistema.Reset_Lista_Adyacencia(numero_total_nudos);
    vecinos.clear();
    vecinos["SW"] = -Nl - 1;
    vecinos["W"] = -Nl;
    vecinos["NW"] = -Nl + 1;
    vecinos["S"] = -1;
    vecinos["N"] = +1;
    vecinos["SE"] = Nl - 1;
    vecinos["E"] = Nl;
    vecinos["NE"] = Nl + 1;
    for (int k = 0; k < (int)Interfaces.size(); ++k) {
        Interfaz& it = Interfaces[k];
        for (const auto& dir : vecinos) {
            if (!Es_vecino(k, dir.first)) continue;
            const Interfaz& nb = Interfaces[k + dir.second];
            for (const Nudo& n : nb.nudos_circunferencia)
                if ((n.posicion_cartesiana - it.centro).norm() < it.radio) it.nudos_interior.push_back(n);
        }
        sistema.Anadir_Lista_Adyacencia(it);
    }
    sistema.Purgar_Lista_Adyacencia();
    sistema.Inicializa_CSR();
    return 0;
}

bool Malla::Es_vecino(int k, const std::string& d) const {
    const Interfaz& it = Interfaces[k];
    if (d.find('N') != std::string::npos && it.posicion_centro[1] == numero_subdominios_lado - 1) return false;
    if (d.find('S') != std::string::npos && it.posicion_centro[1] == 0) return false;
    if (d.find('E') != std::string::npos && it.posicion_centro[0] == numero_subdominios_lado - 1) return false;
    if (d.find('W') != std::string::npos && it.posicion_centro[0] == 0) return false;
    return true;
}

std::vector<TrabajoFila> Malla::Trabajos(bool all_circles) const {
    std::vector<TrabajoFila> out;
    for (int k = 0; k < (int)Interfaces.size(); ++k) {
        const Interfaz& it = Interfaces[k];
        if (!all_circles && !it.es_perimeter) continue;
        for (const Nudo& n : it.nudos_interior)
            out.push_back({k, n.indice_global, n.posicion_cartesiana, n.frontera});
    }
    return out;
}

bool Malla::Escribe_Posiciones(const std::string& path, double (*u)(double, double)) const {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    for (const Interfaz& it : Interfaces)
        for (const Nudo& n : it.nudos_circunferencia)
            std::fprintf(f, "%d %.12f %.12f %.12f\n", n.indice_global, n.posicion_cartesiana(0),
                         n.posicion_cartesiana(1), u ? u(n.posicion_cartesiana(0), n.posicion_cartesiana(1)) : 0.0);
    std::fclose(f);
    return true;
}

} // namespace spac
