/*
 * Interfaz.cpp — circle set-up, RBF tables and the host form of Update_G.
 * Mesh/Interfaz.cpp:2-264 restated; angles and positions use libm exactly as
 * the reference does, so node angles are the reference's bits.
 */
#include "Interfaz.hpp"

#include <algorithm>
#include <cmath>

#include "spac_rbf.h"

namespace spac {

int Interfaz::Inicia_Cuadrado(double RADIO, const std::vector<double>& P, double theta_0,
                              const std::map<std::string, double>& c2map, Vec2 CENTRO, int ix, int iy, int imx,
                              int imy, int Nt, int primer_nudo) {
    int n_arc = 0;
    radio = RADIO;
    centro = CENTRO;
    posicion_centro = {ix, iy};
    parametros_dominio.assign(P.begin(), P.begin() + 4);
    nudos_circunferencia.clear();
    nudos_interior.clear();
    double inc = 2 * (M_PI) / Nt, first = 0.0, last = 0.0;
    Vec2 aux;
    const double r2 = radio * radio;
    auto count = [&](double a, double b) { return (int)(Nt * std::fabs(b - a) / (2 * M_PI)); };
    if (ix == 0) {
        if (iy == 0) {
            aux(1) = P[1] - centro[1]; aux(0) = std::sqrt(r2 - aux(1) * aux(1)); first = Atan180(aux);
            aux(0) = P[0] - centro[0]; aux(1) = std::sqrt(r2 - aux(0) * aux(0)); last = Atan180(aux);
            interior = -11;
        } else if (iy == imy) {
            aux(0) = P[0] - centro[0]; aux(1) = -std::sqrt(r2 - aux(0) * aux(0)); first = Atan180(aux);
            aux(1) = P[3] - centro[1]; aux(0) = std::sqrt(r2 - aux(1) * aux(1)); last = Atan180(aux);
            interior = -22;
        } else {
            aux(0) = P[0] - centro[0]; aux(1) = -std::sqrt(r2 - aux(0) * aux(0)); first = Atan180(aux);
            aux(1) = std::sqrt(r2 - aux(0) * aux(0)); last = Atan180(aux);
            interior = -4;
        }
        n_arc = count(first, last);
    } else if (ix == imx) {
        if (iy == 0) {
            aux(0) = P[2] - centro[0]; aux(1) = std::sqrt(r2 - aux(0) * aux(0)); first = Atan0(aux);
            aux(1) = P[1] - centro[1]; aux(0) = -std::sqrt(r2 - aux(1) * aux(1)); last = Atan0(aux);
            interior = -44;
        } else if (iy == imy) {
            aux(1) = P[3] - centro[1]; aux(0) = -std::sqrt(r2 - aux(1) * aux(1)); first = Atan0(aux);
            aux(0) = P[2] - centro[0]; aux(1) = -std::sqrt(r2 - aux(0) * aux(0)); last = Atan0(aux);
            interior = -33;
        } else {
            /* the reference reads the domain's ymax here (Mesh/Interfaz.cpp:79); equal to xmax on a square */
            aux(0) = P[3] - centro[0]; aux(1) = std::sqrt(r2 - aux(0) * aux(0)); first = Atan0(aux);
            aux(1) = -std::sqrt(r2 - aux(0) * aux(0)); last = Atan0(aux);
            interior = -2;
        }
        n_arc = count(first, last);
    } else {
        if (iy == 0) {
            aux(1) = P[1] - centro[1]; aux(0) = +std::sqrt(r2 - aux(1) * aux(1)); first = Atan270(aux);
            aux(0) = -std::sqrt(r2 - aux(1) * aux(1)); last = Atan270(aux);
            interior = -1;
            n_arc = count(first, last);
        } else if (iy == imy) {
            aux(1) = P[3] - centro[1]; aux(0) = -std::sqrt(r2 - aux(1) * aux(1)); first = Atan90(aux);
            aux(0) = +std::sqrt(r2 - aux(1) * aux(1)); last = Atan90(aux);
            interior = -3;
            n_arc = count(first, last);
        } else {
            nudos_circunferencia.resize(Nt);
            first = theta_0;
            interior = 1;
        }
    }
    int aux_index = primer_nudo;
    auto place = [&](Nudo& nd, double theta, int local, bool frontera) {
        nd.indice_local = local;
        nd.indice_global = aux_index++;
        nd.indice_interfaz = {posicion_centro[0], posicion_centro[1]};
        nd.posicion_angular = theta;
        Vec2 a(radio * std::cos(theta), radio * std::sin(theta));
        nd.posicion_cartesiana = centro + a;
        nd.frontera = frontera;
    };
    if (interior > 0) {
        for (int i = 0; i < (int)nudos_circunferencia.size(); i++)
            place(nudos_circunferencia[i], first + inc * i, i, false);
        es_perimeter = false;
    } else {
        nudos_circunferencia.resize(n_arc + 2);
        place(nudos_circunferencia[0], first, 0, true);
        for (int i = 0; i < n_arc; i++) place(nudos_circunferencia[i + 1], first + theta_0 + inc * i, i + 1, false);
        place(nudos_circunferencia[n_arc + 1], last, n_arc + 1, true);
        es_perimeter = true;
        auto it = c2map.find(interior < -10 ? "esquina" : "lado");
        Calcula_Psi(it != c2map.end() ? it->second : 1.0);
    }
    return aux_index;
}

double Interfaz::Calcula_Psi(double C2) {
    c2 = C2;
    const int M = (int)nudos_circunferencia.size();
    Psi.resize(M, M);
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++)
            Psi(i, j) = Inverse_Multiquadric(nudos_circunferencia[i].posicion_angular,
                                             nudos_circunferencia[j].posicion_angular, c2);
    iPsi = full_piv_lu_inverse(Psi);
    return iPsi.norm() * Psi.norm();
}

double Interfaz::Calcula_Stencils(double eps, int S) {
    const int M = (int)nudos_circunferencia.size();
    S = std::min(S, M);
    stencil_size = S;
    gauss_eps = eps;
    stencil_inverses.clear();
    double worst = 0.0;
    auto add_block = [&](const MatX& P) {
        MatX inv = full_piv_lu_inverse(P);
        worst = std::max(worst, inv.norm() * P.norm());
        stencil_inverses.insert(stencil_inverses.end(), inv.d.begin(), inv.d.end());
    };
    if (interior == 1) {
        const double delta = SPAC_TWO_PI / (double)M;
        const int m = S / 2;
        const double L = m > 0 ? (double)m * delta : 1.0;
        MatX P(S, S);
        for (int k = 0; k < S; ++k)
            for (int i = 0; i < S; ++i) P(k, i) = spac_gauss_phi(((double)(k - i) * delta) / L, eps);
        add_block(P);
    } else {
        std::vector<double> th(M);
        for (int i = 0; i < M; ++i) th[i] = nudos_circunferencia[i].posicion_angular;
        for (int s = 0; s + S <= M; ++s) {
            const double L = S > 1 ? (th[s + S - 1] - th[s]) / 2.0 : 1.0;
            MatX P(S, S);
            for (int k = 0; k < S; ++k)
                for (int i = 0; i < S; ++i) P(k, i) = spac_gauss_phi((th[s + k] - th[s + i]) / L, eps);
            add_block(P);
        }
    }
    return worst;
}

bool Interfaz::Angulo(const Vec2& X, double* theta) const {
    Vec2 d = X - centro;
    switch (interior) {
    case -11: case -22: case -4: *theta = Atan180(d); return true;
    case -33: case -44: case -2: *theta = Atan0(d); return true;
    case -1: *theta = Atan270(d); return true;
    case -3: *theta = Atan90(d); return true;
    default: return false;
    }
}

int Interfaz::Update_G(double Y, const Vec2& X, std::map<int, double>& G) const {
    double theta;
    if (!Angulo(X, &theta) || iPsi.rows != (int)nudos_circunferencia.size()) return -1;
    const int M = (int)nudos_circunferencia.size();
    for (int j = 0; j < M; j++)
        for (int i = 0; i < M; i++)
            G[nudos_circunferencia[j].indice_global] +=
                -Y * iPsi(i, j) * Inverse_Multiquadric(theta, nudos_circunferencia[i].posicion_angular, c2);
    return 0;
}

void Interfaz::Resetea() {
    posicion_centro = {-1, -1};
    ubicacion = -1;
    interior = -10;
    parametros_dominio.clear();
    Psi.resize(0, 0);
    iPsi.resize(0, 0);
    centro = Vec2();
    radio = 0.0;
    c2 = -1.0;
    nudos_circunferencia.clear();
    nudos_interior.clear();
    stencil_inverses.clear();
    stencil_size = 0;
}

} // namespace spac
