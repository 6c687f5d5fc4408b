/*
 * Interpolador.hpp — the reference's tabulated-function pair TablaDeValores +
 * Interpolador (BVP/TablaDeValores.hpp:5-18, BVP/Interpolador.hpp:7-47) on the
 * drop-in surface, and Interpolador_Tabla, the interpolator the GPU path
 * kernels can read.
 *
 * The reference's Interpolador is a base class whose methods print "Programar
 * en clases derivadas" and return NaN; derived classes (RBF, pseudospectral)
 * implement them.  Here the base returns NaN without printing, and
 * Interpolador_Tabla interpolates bilinearly on a uniform T x T grid — the
 * lookup of spac_cv.h, the same bits the FP64 path kernel computes — so a
 * tabulated u_hat can serve as the control variate F = -sigma grad u_hat
 * (FeynmanKacSolver::Variable_Control).
 */
#ifndef SPAC_INTERPOLADOR_HPP
#define SPAC_INTERPOLADOR_HPP

#include <cmath>
#include <limits>
#include <utility>
#include <vector>

#include "../csrc/spac_cv.h"
#include "spac_la.hpp"

namespace spac {

/* Values z of a scalar on the uniform grid x_i = xmin + i (xmax - xmin) / (T - 1),
 * y_j likewise, stored at [j T + i]; optional tabulated derivatives (else
 * Interpolador_Tabla takes centred differences of z). */
class TablaDeValores {
public:
    double xmin = 0.0, ymin = 0.0, xmax = 1.0, ymax = 1.0;
    int T = 0;
    std::vector<double> z, dzdx, dzdy;

    TablaDeValores() = default;
    TablaDeValores(double x0, double y0, double x1, double y1, int t, std::vector<double> Z,
                   std::vector<double> DX = {}, std::vector<double> DY = {}) {
        Inicializa_TablaDeValores(x0, y0, x1, y1, t, std::move(Z), std::move(DX), std::move(DY));
    }
    void Inicializa_TablaDeValores(double x0, double y0, double x1, double y1, int t, std::vector<double> Z,
                                   std::vector<double> DX = {}, std::vector<double> DY = {}) {
        xmin = x0; ymin = y0; xmax = x1; ymax = y1; T = t;
        z = std::move(Z); dzdx = std::move(DX); dzdy = std::move(DY);
    }
    double x(int i) const { return xmin + (double)i * ((xmax - xmin) / (double)(T - 1)); }
    double y(int j) const { return ymin + (double)j * ((ymax - ymin) / (double)(T - 1)); }
    bool valida() const {
        const size_t n = (size_t)T * (size_t)T;
        return T >= 2 && xmax > xmin && ymax > ymin && z.size() == n && (dzdx.empty() || dzdx.size() == n) &&
               (dzdy.empty() || dzdy.size() == n);
    }
};

class Interpolador {
public:
    TablaDeValores valores;
    virtual ~Interpolador() = default;
    void Resetea() { valores = TablaDeValores(); }
    virtual void Inicializa(const TablaDeValores& v) { valores = v; }
    static double nan() { return std::numeric_limits<double>::quiet_NaN(); }
    virtual double Interpola(Vec2) const { return nan(); }
    virtual double Deriva_x(Vec2) const { return nan(); }
    virtual double Deriva_y(Vec2) const { return nan(); }
    virtual double Deriva_r(Vec2) const { return nan(); }
    virtual double Deriva_t(Vec2) const { return nan(); }
    virtual double Deriva_x2(Vec2) const { return nan(); }
    virtual double Deriva_y2(Vec2) const { return nan(); }
    virtual double Deriva_r2(Vec2) const { return nan(); }
    virtual double Deriva_t2(Vec2) const { return nan(); }
    virtual double Deriva_xy(Vec2) const { return nan(); }
    virtual double Deriva_tr(Vec2) const { return nan(); }
};

/* Bilinear interpolation of z, dz/dx, dz/dy on the table, with the nearest edge
 * value outside it; the values are held as floats, (z, dz/dx, dz/dy, 0) per node,
 * the layout spac_cv_config.table uploads to the device. */
class Interpolador_Tabla : public Interpolador {
public:
    void Inicializa(const TablaDeValores& v) override {
        valores = v;
        packed_.clear();
        if (!v.valida()) return;
        const int T = v.T;
        packed_.assign((size_t)T * T * 4, 0.0f);
        const double hx = (v.xmax - v.xmin) / (double)(T - 1), hy = (v.ymax - v.ymin) / (double)(T - 1);
        for (int j = 0; j < T; ++j)
            for (int i = 0; i < T; ++i) {
                const size_t k = (size_t)j * T + i;
                double gx, gy;
                if (!v.dzdx.empty()) gx = v.dzdx[k];
                else {
                    const int i0 = i > 0 ? i - 1 : i, i1 = i < T - 1 ? i + 1 : i;
                    gx = (v.z[(size_t)j * T + i1] - v.z[(size_t)j * T + i0]) / ((double)(i1 - i0) * hx);
                }
                if (!v.dzdy.empty()) gy = v.dzdy[k];
                else {
                    const int j0 = j > 0 ? j - 1 : j, j1 = j < T - 1 ? j + 1 : j;
                    gy = (v.z[(size_t)j1 * T + i] - v.z[(size_t)j0 * T + i]) / ((double)(j1 - j0) * hy);
                }
                packed_[4 * k] = (float)v.z[k];
                packed_[4 * k + 1] = (float)gx;
                packed_[4 * k + 2] = (float)gy;
            }
    }
    double Interpola(Vec2 p) const override { return eval(p, 0); }
    double Deriva_x(Vec2 p) const override { return eval(p, 1); }
    double Deriva_y(Vec2 p) const override { return eval(p, 2); }
    /* T * T * 4 floats, empty until initialised with a valid table */
    const std::vector<float>& tabla() const { return packed_; }

private:
    std::vector<float> packed_;
    double eval(Vec2 p, int comp) const {
        if (packed_.empty()) return nan();
        spac_cv_table t;
        t.table = packed_.data();
        t.t = valores.T;
        t.lox = valores.xmin;
        t.loy = valores.ymin;
        t.ix = (double)(valores.T - 1) / (valores.xmax - valores.xmin);
        t.iy = (double)(valores.T - 1) / (valores.ymax - valores.ymin);
        int32_t k00;
        double ax, ay;
        spac_cv_cell(&t, p(0), p(1), &k00, &ax, &ay);
        return spac_cv_eval(&t, k00, ax, ay, comp);
    }
};

} // namespace spac

#endif
