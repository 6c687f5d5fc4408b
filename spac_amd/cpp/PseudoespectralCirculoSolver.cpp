/*
 * PseudoespectralCirculoSolver.cpp — polar Chebyshev x Fourier collocation of
 * Laplacian + c on a circle and the interpolation of its grid solutions
 * (Solvers/PseudoespectralCirculoSolver.cpp:2-260, BVP/Interpolador_PS.cpp:18-61).
 *
 * The steps follow the reference (and, through it, the polar-coordinates Laplacian of
 * Trefethen's "Spectral Methods in MATLAB", program 29): the same grid, the same operator
 * blocks and the same interpolation formulas, so that the rows are the reference's up to the
 * factorisation.  The reference factors with Eigen's FullPivHouseholderQR; here the operator
 * is factored by Gaussian elimination with partial pivoting (our own), which agrees to the
 * operator's condition number times machine epsilon.
 */
#include "PseudoespectralCirculoSolver.hpp"

#include <cmath>
#include <cstring>

namespace spac {

void Interpolador_PS::Inicializa(const Vec2& c0, double R, const std::vector<double>& rr,
                                 const std::vector<double>& tt) {
    centro = c0;
    radio = R;
    r = rr;
    t = tt;
    N = (int)t.size();
    n = N / 2;
    normN = 1.0 / N;
    m = 2 * (int)r.size();
    /* the diameter through the centre, ascending in units of R: -r_0 .. -r_N2, r_N2 .. r_0 */
    diametro.assign(m, 0.0);
    for (int i = 0; i < m / 2; ++i) {
        diametro[i] = -r[i] / radio;
        diametro[m - 1 - i] = r[i] / radio;
    }
    /* the reference's barycentric weights (BVP/Interpolador_PS.cpp:28-31) */
    c.assign(m, 0.0);
    const double i_ts = -1.0 + 1.0 / m;
    for (int i = 0; i < m; ++i) c[i] = ((i % 2) ? -1.0 : 1.0) * std::sin(M_PI * (i_ts + (2.0 / m) * i + 1) / 2);
    z.assign((size_t)m / 2 * (size_t)N, 0.0);
}

double Interpolador_PS::Interpola(const Vec2& punto) const {
    const int nr = m / 2;
    const double dx = punto(0) - centro(0), dy = punto(1) - centro(1);
    const double r_punto = std::sqrt(dx * dx + dy * dy) / radio;
    double theta_punto = std::atan2(dy, dx);
    theta_punto = (theta_punto >= 0) ? theta_punto : theta_punto + 2.0 * M_PI;
    /* ring values at the point's angle (head) and at the opposite angle (tail):
       sum over angles l and modes k = -n .. n-1 of cos(k (theta - t_l)) z(l, i), the opposite
       angle with the factor cos(k pi) */
    std::vector<double> head(nr, 0.0), tail(nr, 0.0);
    for (int l = 0; l < N; ++l) {
        double w = 0.0, wo = 0.0;
        for (int k = -n; k < n; ++k) {
            const double ck = std::cos(k * (theta_punto - t[l]));
            w += ck;
            wo += std::cos(k * M_PI) * ck;
        }
        for (int i = 0; i < nr; ++i) {
            const double zi = z[(size_t)i * N + l];
            head[i] += w * zi;
            tail[i] += wo * zi;
        }
    }
    /* values along the diameter, ascending: opposite side from r = R inwards, then this side
       from the centre out */
    double num = 0.0, den = 0.0;
    for (int q = 0; q < m; ++q) {
        const double u = normN * (q < nr ? tail[q] : head[m - 1 - q]);
        const double tq = c[q] / (r_punto - diametro[q]);
        num += tq * u;
        den += tq;
    }
    return num / den;
}

int PseudoespectralCirculoSolver::Inicializa(const Vec2& centro, double radio, unsigned N_circunferencia,
                                             unsigned N_radio, double theta_0, const BVP& Problema) {
    if (N_radio % 2 != 1 || N_radio < 3 || N_circunferencia % 2 != 0 || N_circunferencia < 2 || !(radio > 0.0))
        return -1;
    const int Nr = (int)N_radio, M = (int)N_circunferencia, N2 = (Nr - 1) / 2, M2 = M / 2;
    /* Chebyshev points on the diameter and their differentiation matrix (:20-44) */
    std::vector<double> r(Nr + 1), cw(Nr + 1);
    for (int i = 0; i <= Nr; ++i) {
        r[i] = radio * std::cos(M_PI * i / Nr);
        cw[i] = ((i == 0 || i == Nr) ? 2.0 : 1.0) * ((i % 2) ? -1.0 : 1.0);
    }
    MatX D(Nr + 1, Nr + 1);
    for (int i = 0; i <= Nr; ++i) {
        double s = 0.0;
        for (int j = 0; j <= Nr; ++j) {
            D(i, j) = (cw[i] / cw[j]) / (r[i] - r[j] + (i == j ? 1.0 : 0.0));
            s += D(i, j);
        }
        D(i, i) -= s;
    }
    /* D^2 and D restricted to the r > 0 rows, split into the columns of r > 0 and their mirror
       images r -> -r (:47-62) */
    MatX D2f(Nr + 1, Nr + 1);
    for (int i = 0; i <= Nr; ++i)
        for (int k = 0; k <= Nr; ++k) {
            const double a = D(i, k);
            if (a == 0.0) continue;
            for (int j = 0; j <= Nr; ++j) D2f(i, j) += a * D(k, j);
        }
    MatX A1(N2 + 1, N2 + 1), A2(N2 + 1, N2 + 1); /* D1 + R E1, D2 + R E2 */
    for (int i = 0; i <= N2; ++i)
        for (int j = 0; j <= N2; ++j) {
            A1(i, j) = D2f(i, j) + D(i, j) / r[i];
            A2(i, j) = D2f(i, Nr - j) + D(i, Nr - j) / r[i];
        }
    /* periodic second derivative in theta (:63-81) */
    const double dtheta = 2 * M_PI / M;
    std::vector<double> t(M), col(M);
    for (int l = 0; l < M; ++l) {
        t[l] = l * dtheta + theta_0;
        col[l] = (l == 0) ? -(M_PI * M_PI / (3 * dtheta * dtheta)) - (1.0 / 6)
                          : 0.5 * ((l % 2) ? 1.0 : -1.0) / std::pow(std::sin(dtheta * l * 0.5), 2);
    }
    /* L = kron(A1, I) + kron(A2, IZZI) + kron(diag(1/r^2), D2t), unknowns i M + l (:82-126) */
    dim = (N2 + 1) * M;
    lu.assign((size_t)dim * (size_t)dim, 0.0);
    auto at = [&](int a, int b) -> double& { return lu[(size_t)a * (size_t)dim + (size_t)b]; };
    for (int i = 0; i <= N2; ++i)
        for (int l = 0; l < M; ++l) {
            const int row = i * M + l;
            for (int j = 0; j <= N2; ++j) {
                at(row, j * M + l) += A1(i, j);
                at(row, j * M + (l + M2) % M) += A2(i, j);
            }
            const double ir2 = 1.0 / (r[i] * r[i]);
            for (int q = 0; q < M; ++q) at(row, i * M + q) += ir2 * col[(q - l + M) % M];
        }
    interpolador.Inicializa(centro, radio, std::vector<double>(r.begin(), r.begin() + N2 + 1), t);
    /* c on the diagonal (:163-168), Dirichlet rows on r = R (:169-172) */
    cdiag.assign(dim, 0.0);
    for (int k = 0; k < dim; ++k) {
        cdiag[k] = Problema.c.Evalua(punto(k));
        at(k, k) += cdiag[k];
    }
    for (int l = 0; l < M; ++l) {
        std::memset(&at(l, 0), 0, sizeof(double) * (size_t)dim);
        at(l, l) = 1.0;
    }
    /* P L = LU, partial pivoting */
    piv.assign(dim, 0);
    for (int k = 0; k < dim; ++k) {
        int p = k;
        double best = std::fabs(at(k, k));
        for (int i = k + 1; i < dim; ++i)
            if (std::fabs(at(i, k)) > best) { best = std::fabs(at(i, k)); p = i; }
        if (!(best > 0.0)) return -1;
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < dim; ++j) std::swap(at(k, j), at(p, j));
        const double inv = 1.0 / at(k, k);
        double* rk = &at(k, 0);
        for (int i = k + 1; i < dim; ++i) {
            double* ri = &at(i, 0);
            const double f = ri[k] * inv;
            ri[k] = f;
            if (f == 0.0) continue;
            for (int j = k + 1; j < dim; ++j) ri[j] -= f * rk[j];
        }
    }
    rhs_pseudospectral.assign(dim, 0.0);
    return 0;
}

bool PseudoespectralCirculoSolver::Reubica(const Vec2& centro, const BVP& Problema) {
    if (dim == 0) return false;
    const Vec2 old = interpolador.centro;
    interpolador.centro = centro;
    for (int k = 0; k < dim; ++k)
        if (Problema.c.Evalua(punto(k)) != cdiag[k]) {
            interpolador.centro = old;
            return false;
        }
    return true;
}

Vec2 PseudoespectralCirculoSolver::punto(int k) const {
    const int M = interpolador.N, i = k / M, l = k % M;
    const double ri = interpolador.r[i], tl = interpolador.t[l];
    return Vec2(ri * std::cos(tl) + interpolador.centro(0), ri * std::sin(tl) + interpolador.centro(1));
}

void PseudoespectralCirculoSolver::solve(std::vector<double>& b) const {
    const int n = dim;
    for (int k = 0; k < n; ++k)
        if (piv[k] != k) std::swap(b[k], b[piv[k]]);
    for (int i = 0; i < n; ++i) {
        const double* ri = &lu[(size_t)i * (size_t)n];
        double s = b[i];
        for (int j = 0; j < i; ++j) s -= ri[j] * b[j];
        b[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        const double* ri = &lu[(size_t)i * (size_t)n];
        double s = b[i];
        for (int j = i + 1; j < n; ++j) s -= ri[j] * b[j];
        b[i] = s / ri[i];
    }
}

void PseudoespectralCirculoSolver::Resuelve(const BVP& Problema) {
    const int M = interpolador.N;
    for (int k = 0; k < dim; ++k) {
        const Vec2 x = punto(k);
        rhs_pseudospectral[k] = (k >= M) ? -Problema.f.Evalua(x) : Problema.g.Evalua(x);
    }
    interpolador.z = rhs_pseudospectral;
    solve(interpolador.z);
}

void PseudoespectralCirculoSolver::Resuelve(const BVP& Problema, std::vector<Nudo>& nudos_interior,
                                            const std::vector<Nudo>& nudos_periferia) {
    const int M = interpolador.N;
    /* B: zero boundary data, source f */
    for (int k = 0; k < dim; ++k) rhs_pseudospectral[k] = (k >= M) ? -Problema.f.Evalua(punto(k)) : 0.0;
    interpolador.z = rhs_pseudospectral;
    solve(interpolador.z);
    for (Nudo& nd : nudos_interior) nd.B = interpolador.Interpola(nd.posicion_cartesiana);
    /* G: boundary data the Fourier cardinal of each periphery node, no source */
    const int nh = (int)nudos_periferia.size() / 2;
    for (const Nudo& pj : nudos_periferia) {
        std::vector<double> z(dim, 0.0);
        for (int l = 0; l < M; ++l) z[l] = Problema.fcardinal_Fourier(interpolador.t[l], pj.posicion_angular, nh);
        solve(z);
        interpolador.z = z;
        for (Nudo& nd : nudos_interior) nd.G[pj.indice_global] = -interpolador.Interpola(nd.posicion_cartesiana);
    }
}

} // namespace spac
