/*
 * Interpolador_RBF.cpp — inverse-multiquadric interpolation with linear polynomial
 * augmentation (the class the reference leaves as stubs, BVP/Interpolador_RBF.cpp:2-52).
 */
#include "Interpolador_RBF.hpp"

#include <algorithm>
#include <cmath>

namespace spac {

void Interpolador_RBF::Inicializa(const TablaDeValores& v) {
    valores = v;
    c_.clear();
    w_.clear();
    if (!v.valida() || (size_t)v.T * (size_t)v.T > (size_t)kMaxCentros) return;
    std::vector<Vec2> cs;
    cs.reserve((size_t)v.T * v.T);
    for (int j = 0; j < v.T; ++j)
        for (int i = 0; i < v.T; ++i) cs.emplace_back(v.x(i), v.y(j));
    (void)Inicializa(cs, v.z, 0.0);
}

int Interpolador_RBF::Inicializa(const std::vector<Vec2>& centros, const std::vector<double>& z, double eps) {
    c_.clear();
    w_.clear();
    const int n = (int)centros.size();
    if (n < 3 || n > kMaxCentros || z.size() != centros.size()) return -1;
    double x0 = centros[0](0), x1 = x0, y0 = centros[0](1), y1 = y0;
    for (int j = 0; j < n; ++j) {
        if (!std::isfinite(centros[j](0)) || !std::isfinite(centros[j](1)) || !std::isfinite(z[j])) return -1;
        x0 = std::min(x0, centros[j](0)); x1 = std::max(x1, centros[j](0));
        y0 = std::min(y0, centros[j](1)); y1 = std::max(y1, centros[j](1));
    }
    const double wdt = x1 - x0, hgt = y1 - y0;
    if (!(eps > 0.0)) {
        const double area = (wdt > 0.0 ? wdt : 1.0) * (hgt > 0.0 ? hgt : 1.0);
        eps = 1.0 / std::sqrt(area / (double)n);
    }
    centro = Vec2(0.5 * (x0 + x1), 0.5 * (y0 + y1));
    /* [Phi P; P^T 0] [w; a] = [z; 0], row-major, partial pivoting */
    const int m = n + 3;
    std::vector<double> A((size_t)m * m, 0.0), rhs(m, 0.0);
    auto at = [&](int i, int j) -> double& { return A[(size_t)i * m + j]; };
    const double e2 = eps * eps;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            const double dx = centros[i](0) - centros[j](0), dy = centros[i](1) - centros[j](1);
            at(i, j) = 1.0 / std::sqrt(1.0 + e2 * (dx * dx + dy * dy));
        }
        at(i, n) = at(n, i) = 1.0;
        at(i, n + 1) = at(n + 1, i) = centros[i](0);
        at(i, n + 2) = at(n + 2, i) = centros[i](1);
        rhs[i] = z[i];
    }
    for (int k = 0; k < m; ++k) {
        int p = k;
        double best = std::fabs(at(k, k));
        for (int i = k + 1; i < m; ++i)
            if (std::fabs(at(i, k)) > best) { best = std::fabs(at(i, k)); p = i; }
        if (!(best > 1e-300)) return -1;
        if (p != k) {
            for (int j = 0; j < m; ++j) std::swap(at(k, j), at(p, j));
            std::swap(rhs[k], rhs[p]);
        }
        const double inv = 1.0 / at(k, k);
        for (int i = k + 1; i < m; ++i) {
            const double f = at(i, k) * inv;
            if (f == 0.0) continue;
            for (int j = k + 1; j < m; ++j) at(i, j) -= f * at(k, j);
            rhs[i] -= f * rhs[k];
        }
    }
    std::vector<double> sol(m);
    for (int i = m - 1; i >= 0; --i) {
        double s = rhs[i];
        for (int j = i + 1; j < m; ++j) s -= at(i, j) * sol[j];
        sol[i] = s / at(i, i);
        if (!std::isfinite(sol[i])) return -1;
    }
    c_ = centros;
    w_.assign(sol.begin(), sol.begin() + n);
    for (int k = 0; k < 3; ++k) a_[k] = sol[n + k];
    eps_ = eps;
    return 0;
}

void Interpolador_RBF::todo(Vec2 p, double o[6]) const {
    const double e2 = eps_ * eps_;
    double s = 0.0, sx = 0.0, sy = 0.0, sxx = 0.0, syy = 0.0, sxy = 0.0;
    for (size_t j = 0; j < c_.size(); ++j) {
        const double dx = p(0) - c_[j](0), dy = p(1) - c_[j](1);
        const double q = 1.0 + e2 * (dx * dx + dy * dy);
        const double r1 = 1.0 / std::sqrt(q);     /* q^-1/2 */
        const double r3 = r1 / q, r5 = r3 / q;     /* q^-3/2, q^-5/2 */
        const double w = w_[j];
        s += w * r1;
        sx -= w * e2 * dx * r3;
        sy -= w * e2 * dy * r3;
        sxx += w * (3.0 * e2 * e2 * dx * dx * r5 - e2 * r3);
        syy += w * (3.0 * e2 * e2 * dy * dy * r5 - e2 * r3);
        sxy += w * 3.0 * e2 * e2 * dx * dy * r5;
    }
    o[0] = s + a_[0] + a_[1] * p(0) + a_[2] * p(1);
    o[1] = sx + a_[1];
    o[2] = sy + a_[2];
    o[3] = sxx;
    o[4] = syy;
    o[5] = sxy;
}

void Interpolador_RBF::polar(Vec2 p, double& r, double& ct, double& st) const {
    const double dx = p(0) - centro(0), dy = p(1) - centro(1);
    r = std::sqrt(dx * dx + dy * dy);
    if (r > 0.0) { ct = dx / r; st = dy / r; }
    else { ct = 1.0; st = 0.0; }
}

double Interpolador_RBF::Interpola(Vec2 p) const {
    if (!listo()) return nan();
    double o[6];
    todo(p, o);
    return o[0];
}
double Interpolador_RBF::Deriva_x(Vec2 p) const {
    if (!listo()) return nan();
    double o[6];
    todo(p, o);
    return o[1];
}
double Interpolador_RBF::Deriva_y(Vec2 p) const {
    if (!listo()) return nan();
    double o[6];
    todo(p, o);
    return o[2];
}
double Interpolador_RBF::Deriva_x2(Vec2 p) const {
    if (!listo()) return nan();
    double o[6];
    todo(p, o);
    return o[3];
}
double Interpolador_RBF::Deriva_y2(Vec2 p) const {
    if (!listo()) return nan();
    double o[6];
    todo(p, o);
    return o[4];
}
double Interpolador_RBF::Deriva_xy(Vec2 p) const {
    if (!listo()) return nan();
    double o[6];
    todo(p, o);
    return o[5];
}
/* x = cx + r cos t, y = cy + r sin t */
double Interpolador_RBF::Deriva_r(Vec2 p) const {
    if (!listo()) return nan();
    double o[6], r, c, s;
    todo(p, o);
    polar(p, r, c, s);
    return c * o[1] + s * o[2];
}
double Interpolador_RBF::Deriva_t(Vec2 p) const {
    if (!listo()) return nan();
    double o[6], r, c, s;
    todo(p, o);
    polar(p, r, c, s);
    return r * (c * o[2] - s * o[1]);
}
double Interpolador_RBF::Deriva_r2(Vec2 p) const {
    if (!listo()) return nan();
    double o[6], r, c, s;
    todo(p, o);
    polar(p, r, c, s);
    return c * c * o[3] + 2.0 * s * c * o[5] + s * s * o[4];
}
double Interpolador_RBF::Deriva_t2(Vec2 p) const {
    if (!listo()) return nan();
    double o[6], r, c, s;
    todo(p, o);
    polar(p, r, c, s);
    return r * r * (s * s * o[3] - 2.0 * s * c * o[5] + c * c * o[4]) - r * (c * o[1] + s * o[2]);
}
double Interpolador_RBF::Deriva_tr(Vec2 p) const {
    if (!listo()) return nan();
    double o[6], r, c, s;
    todo(p, o);
    polar(p, r, c, s);
    return -s * o[1] + c * o[2] + r * (s * c * (o[4] - o[3]) + (c * c - s * s) * o[5]);
}

} // namespace spac
