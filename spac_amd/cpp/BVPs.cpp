/*
 * BVPs.cpp — definitions of the shipped host plugin functions (see BVPs.hpp).
 */
#include "BVPs.hpp"

#include <cmath>

namespace spac {


double Franke_1(const Vec2& X) { return std::exp(-0.25 * (std::pow(9.0 * X(0) - 2.0, 2) + std::pow(9.0 * X(1) - 2.0, 2))); }
double Franke_2(const Vec2& X) { return std::exp((-1.0 / 49.0) * std::pow(9.0 * X(0) + 1.0, 2) - 0.1 * (9.0 * X(1) + 1.0)); }
double Franke_3(const Vec2& X) { return std::exp(-0.25 * (std::pow(9.0 * X(0) - 7.0, 2) + std::pow(9.0 * X(1) - 3.0, 2))); }
double Franke_4(const Vec2& X) { return std::exp(-1.0 * (std::pow(9.0 * X(0) - 4.0, 2) + std::pow(9.0 * X(1) - 7.0, 2))); }

double Franke_u(Vec2 X) { return 0.75 * Franke_1(X) + 0.75 * Franke_2(X) + 0.5 * Franke_3(X) - 0.2 * Franke_4(X); }
double Franke_dudx(const Vec2& X) {
    return -3.375 * (9.0 * X(0) - 2.0) * Franke_1(X) - (27.0 / 98.0) * (9.0 * X(0) + 1.0) * Franke_2(X)
           - 2.25 * (9.0 * X(0) - 7.0) * Franke_3(X) + 3.6 * (9.0 * X(0) - 4) * Franke_4(X);
}
double Franke_dudy(const Vec2& X) {
    return -3.375 * (9.0 * X(1) - 2.0) * Franke_1(X) - 0.675 * Franke_2(X)
           - 2.25 * (9.0 * X(1) - 3.0) * Franke_3(X) + 3.6 * (9.0 * X(1) - 7.0) * Franke_4(X);
}
double Franke_d2udx2(const Vec2& X) {
    return -3.375 * (9.0 - 4.5 * std::pow(9.0 * X(0) - 2.0, 2)) * Franke_1(X)
           - (27.0 / 98.0) * (9.0 - (18.0 / 49.0) * std::pow(9.0 * X(0) + 1.0, 2)) * Franke_2(X)
           - 2.25 * (9.0 - 4.5 * std::pow(9.0 * X(0) - 7.0, 2)) * Franke_3(X)
           + 3.6 * (9.0 - 18.0 * std::pow(9.0 * X(0) - 4.0, 2)) * Franke_4(X);
}
double Franke_d2udy2(const Vec2& X) {
    return -3.375 * (9.0 - 4.5 * std::pow((9.0 * X(1) - 2.0), 2)) * Franke_1(X) + (243.0 / 400.0) * Franke_2(X)
           - 2.25 * (9.0 - 4.5 * std::pow(9.0 * X(1) - 3.0, 2)) * Franke_3(X)
           + 3.6 * (9.0 - 18.0 * std::pow(9.0 * X(1) - 7.0, 2)) * Franke_4(X);
}
double Franke_c(Vec2) { return -0.0; }
double Franke_g(Vec2 X) { return Franke_u(X); }
Vec2 Franke_b(Vec2) { return Vec2(0.0, 0.0); }
Mat2 Franke_sigma(Vec2) { return Mat2::Identity() * 1.41421356237; }
double Franke_f(Vec2 X) { return -Franke_d2udx2(X) - Franke_d2udy2(X); }
Vec2 Franke_gradu(Vec2 X) { return Vec2(Franke_dudx(X), Franke_dudy(X)); }
Vec2 Franke_F(Vec2 X) {
    Vec2 g(Franke_dudx(X), Franke_dudy(X));
    return -(Franke_sigma(X).transpose() * g);
}

double VarCoef_s(const Vec2& X) { return 1.0 + 0.5 * std::sin(M_PI * X(0)) * std::cos(M_PI * X(1)); }
Mat2 VarCoef_sigma(Vec2 X) { return Mat2::Identity() * VarCoef_s(X); }
Vec2 VarCoef_b(Vec2 X) { return Vec2(0.3 * X(1), -0.3 * X(0)); }
double VarCoef_c(Vec2 X) { return -1.0 - X(0) * X(0); }
double VarCoef_f(Vec2 X) {
    double s = VarCoef_s(X);
    Vec2 b = VarCoef_b(X);
    double lap = Franke_d2udx2(X) + Franke_d2udy2(X);
    return -(((0.5 * (s * s)) * lap + (b(0) * Franke_dudx(X) + b(1) * Franke_dudy(X))) + VarCoef_c(X) * Franke_u(X));
}
Vec2 VarCoef_F(Vec2 X) {
    Vec2 g(Franke_dudx(X), Franke_dudy(X));
    return -(VarCoef_sigma(X).transpose() * g);
}

double Robin_psi(Vec2 X, Vec2 N) { return Franke_dudx(X) * N(0) + Franke_dudy(X) * N(1) + 0.5 * Franke_u(X); }
double Robin_varphi(Vec2, Vec2) { return -0.5; }

} // namespace spac
