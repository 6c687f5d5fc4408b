/*
 * Interpolador_RBF.hpp — scattered-data interpolation with radial basis functions, the
 * reference's Interpolador_RBF (BVP/Interpolador_RBF.hpp:5-33), whose methods are stubs
 * that print "Programar en clases derivadas" (BVP/Interpolador_RBF.cpp:5-52).
 *
 * Here the interpolant is
 *     s(x) = sum_j w_j phi(|x - c_j|) + a_0 + a_1 x + a_2 y,   phi(r) = 1 / sqrt(1 + (eps r)^2)
 * (inverse multiquadric, the kernel the reference's interface tables use), with the side
 * conditions sum_j w_j = sum_j w_j x_j = sum_j w_j y_j = 0, so that s reproduces every linear
 * function exactly.  The (n + 3) x (n + 3) saddle-point system is solved once, by Gaussian
 * elimination with partial pivoting, in Inicializa (O(n^3): at most kMaxCentros centres).
 * Every derivative the reference declares is analytic; the polar ones are taken about
 * `centro` (the centres' bounding-box centre unless set).
 */
#ifndef SPAC_INTERPOLADOR_RBF_HPP
#define SPAC_INTERPOLADOR_RBF_HPP

#include <vector>

#include "Interpolador.hpp"

namespace spac {

class Interpolador_RBF : public Interpolador {
public:
    static constexpr int kMaxCentros = 4096;
    Vec2 centro; /* pole of Deriva_r / Deriva_t / Deriva_r2 / Deriva_t2 / Deriva_tr */

    /* Centres = the table's T x T grid nodes, data = its z; eps from the grid spacing. */
    void Inicializa(const TablaDeValores& v) override;
    /* Scattered centres and data; eps <= 0 picks 1 / (mean spacing sqrt(area / n)).
     * Returns 0, or -1 (size mismatch, fewer than 3 centres, more than kMaxCentros,
     * non-finite data, or a singular system: repeated or collinear centres). */
    int Inicializa(const std::vector<Vec2>& centros, const std::vector<double>& z, double eps = 0.0);

    double Interpola(Vec2 p) const override;
    double Deriva_x(Vec2 p) const override;
    double Deriva_y(Vec2 p) const override;
    double Deriva_r(Vec2 p) const override;
    double Deriva_t(Vec2 p) const override;
    double Deriva_x2(Vec2 p) const override;
    double Deriva_y2(Vec2 p) const override;
    double Deriva_r2(Vec2 p) const override;
    double Deriva_t2(Vec2 p) const override;
    double Deriva_xy(Vec2 p) const override;
    double Deriva_tr(Vec2 p) const override;

    bool listo() const { return !w_.empty(); }
    double epsilon() const { return eps_; }
    int n_centros() const { return (int)c_.size(); }

private:
    std::vector<Vec2> c_;
    std::vector<double> w_;
    double a_[3] = {0.0, 0.0, 0.0};
    double eps_ = 0.0;
    /* s and its first and second Cartesian derivatives at p: {s, sx, sy, sxx, syy, sxy} */
    void todo(Vec2 p, double out[6]) const;
    /* polar frame of p about centro: r, cos t, sin t */
    void polar(Vec2 p, double& r, double& ct, double& st) const;
};

} // namespace spac

#endif
