/*
 * spac_la.hpp — the small dense linear algebra the plugin surface needs.
 *
 * The reference passes Eigen::Vector2d / Eigen::Matrix2d through every plugin
 * signature (BVP/Coeficiente_*.hpp, BVP/Frontera.hpp) and inverts the RBF
 * matrix with Eigen::FullPivLU (Mesh/Interfaz.cpp:209-215).  Eigen is not a
 * dependency here: Vec2/Mat2 provide exactly the operations those signatures
 * use, and FullPivLU is a complete-pivoting Gaussian elimination.
 */
#ifndef SPAC_LA_HPP
#define SPAC_LA_HPP

#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <vector>

namespace spac {

struct Vec2 {
    double v[2] = {0.0, 0.0};
    Vec2() = default;
    Vec2(double a, double b) { v[0] = a; v[1] = b; }
    double& operator()(int i) { return v[i]; }
    double operator()(int i) const { return v[i]; }
    double& operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
    Vec2 operator+(const Vec2& o) const { return {v[0] + o.v[0], v[1] + o.v[1]}; }
    Vec2 operator-(const Vec2& o) const { return {v[0] - o.v[0], v[1] - o.v[1]}; }
    Vec2 operator-() const { return {-v[0], -v[1]}; }
    Vec2 operator*(double s) const { return {v[0] * s, v[1] * s}; }
    Vec2 operator/(double s) const { return {v[0] / s, v[1] / s}; }
    Vec2& operator+=(const Vec2& o) { v[0] += o.v[0]; v[1] += o.v[1]; return *this; }
    Vec2& operator-=(const Vec2& o) { v[0] -= o.v[0]; v[1] -= o.v[1]; return *this; }
    double dot(const Vec2& o) const { return v[0] * o.v[0] + v[1] * o.v[1]; }
    double squaredNorm() const { return v[0] * v[0] + v[1] * v[1]; }
    double norm() const { return std::sqrt(squaredNorm()); }
};
inline Vec2 operator*(double s, const Vec2& a) { return {s * a.v[0], s * a.v[1]}; }

struct Mat2 {
    double a[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    static Mat2 Identity() { Mat2 m; m.a[0][0] = 1.0; m.a[1][1] = 1.0; return m; }
    double& operator()(int i, int j) { return a[i][j]; }
    double operator()(int i, int j) const { return a[i][j]; }
    Mat2 operator*(double s) const {
        Mat2 m;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) m.a[i][j] = a[i][j] * s;
        return m;
    }
    Vec2 operator*(const Vec2& x) const {
        return {a[0][0] * x.v[0] + a[0][1] * x.v[1], a[1][0] * x.v[0] + a[1][1] * x.v[1]};
    }
    Mat2 transpose() const {
        Mat2 m;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) m.a[i][j] = a[j][i];
        return m;
    }
};

/* Dense row-major matrix. */
struct MatX {
    int rows = 0, cols = 0;
    std::vector<double> d;
    MatX() = default;
    MatX(int r, int c) : rows(r), cols(c), d((size_t)r * (size_t)c, 0.0) {}
    void resize(int r, int c) { rows = r; cols = c; d.assign((size_t)r * (size_t)c, 0.0); }
    double& operator()(int i, int j) { return d[(size_t)i * (size_t)cols + (size_t)j]; }
    double operator()(int i, int j) const { return d[(size_t)i * (size_t)cols + (size_t)j]; }
    /* Frobenius norm (Eigen's MatrixBase::norm()). */
    double norm() const {
        double s = 0.0;
        for (double x : d) s += x * x;
        return std::sqrt(s);
    }
};

/* Inverse by Gaussian elimination with complete pivoting (the algorithm of
 * Eigen::FullPivLU, written from the textbook: P A Q = L U). */
inline MatX full_piv_lu_inverse(const MatX& A) {
    if (A.rows != A.cols) throw std::invalid_argument("full_piv_lu_inverse: matrix not square");
    const int n = A.rows;
    MatX lu = A;
    std::vector<int> rp(n), cp(n);
    for (int i = 0; i < n; ++i) { rp[i] = i; cp[i] = i; }
    for (int k = 0; k < n; ++k) {
        int pi = k, pj = k;
        double best = -1.0;
        for (int i = k; i < n; ++i)
            for (int j = k; j < n; ++j) {
                double a = std::fabs(lu(i, j));
                if (a > best) { best = a; pi = i; pj = j; }
            }
        if (!(best > 0.0)) throw std::runtime_error("full_piv_lu_inverse: singular matrix");
        if (pi != k) { for (int j = 0; j < n; ++j) std::swap(lu(k, j), lu(pi, j)); std::swap(rp[k], rp[pi]); }
        if (pj != k) { for (int i = 0; i < n; ++i) std::swap(lu(i, k), lu(i, pj)); std::swap(cp[k], cp[pj]); }
        double piv = lu(k, k);
        for (int i = k + 1; i < n; ++i) {
            double l = lu(i, k) / piv;
            lu(i, k) = l;
            for (int j = k + 1; j < n; ++j) lu(i, j) -= l * lu(k, j);
        }
    }
    /* P A Q = L U  =>  A^-1 = Q U^-1 L^-1 P */
    MatX inv(n, n);
    std::vector<double> col(n);
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) col[i] = (rp[i] == c) ? 1.0 : 0.0; /* P e_c */
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) col[i] -= lu(i, j) * col[j];
        for (int i = n - 1; i >= 0; --i) {
            for (int j = i + 1; j < n; ++j) col[i] -= lu(i, j) * col[j];
            col[i] /= lu(i, i);
        }
        for (int i = 0; i < n; ++i) inv(cp[i], c) = col[i];
    }
    return inv;
}

} // namespace spac

#endif
