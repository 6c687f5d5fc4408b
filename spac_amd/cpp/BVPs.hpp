/*
 * BVPs.hpp — host plugin functions of the shipped problems, with the
 * reference's plugin signatures (pfescalar, pfvector, pfmatriz) over Vec2/Mat2.
 *
 * Franke Poisson: BVPs/Franke_Poisson.hpp:4-73 (f = -lap u, sigma =
 * 1.41421356237 I, F = -sigma^T grad u).  Variable-coefficient Franke: the
 * operator L = 1/2 s^2 lap + b.grad + c with s = 1 + 0.5 sin(pi x) cos(pi y),
 * b = 0.3 (y, -x), c = -1 - x^2 and the manufactured solution u = Franke
 * (f = -(L u), g = u).  Each set is registered against its device twin
 * (SPAC_COEF_*), so a BVP built from these pointers runs on the GPU.
 */
#ifndef SPAC_BVPS_HPP
#define SPAC_BVPS_HPP

#include "spac_la.hpp"

namespace spac {

double Franke_1(const Vec2& X);
double Franke_2(const Vec2& X);
double Franke_3(const Vec2& X);
double Franke_4(const Vec2& X);
double Franke_u(Vec2 X);
double Franke_dudx(const Vec2& X);
double Franke_dudy(const Vec2& X);
double Franke_d2udx2(const Vec2& X);
double Franke_d2udy2(const Vec2& X);
double Franke_c(Vec2);
double Franke_g(Vec2 X);
Vec2 Franke_b(Vec2);
Mat2 Franke_sigma(Vec2);
double Franke_f(Vec2 X);
Vec2 Franke_gradu(Vec2 X);
Vec2 Franke_F(Vec2 X);
double VarCoef_s(const Vec2& X);
Mat2 VarCoef_sigma(Vec2 X);
Vec2 VarCoef_b(Vec2 X);
double VarCoef_c(Vec2 X);
double VarCoef_f(Vec2 X);
Vec2 VarCoef_F(Vec2 X);
/* Robin data of reflecting sides for the Franke u: du/dN - varphi u = psi with
 * varphi = -0.5 (the psi/varphi slots of BVP/BVP.hpp:17-28) */
double Robin_psi(Vec2 X, Vec2 N);
double Robin_varphi(Vec2 X, Vec2 N);

} // namespace spac

#endif
