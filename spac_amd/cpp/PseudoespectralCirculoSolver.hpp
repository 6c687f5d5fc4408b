/*
 * PseudoespectralCirculoSolver.hpp — the deterministic G/B source for nodes inside a full
 * interior circle (Solvers/PseudoespectralCirculoSolver.hpp:8-18, BVP/Interpolador_PS.hpp:5-38).
 *
 * On a circle of radius R the operator L = Laplacian + c(x) is discretised on a polar grid:
 * N_circunferencia equispaced angles (Fourier) times the r > 0 half of N_radio + 1 Chebyshev
 * points on a diameter, with Dirichlet rows on r = R.  B of an interior node is the solution
 * with zero boundary data and source f there; column j of its G row is minus the solution
 * whose boundary data is the Fourier cardinal function of circle node j, both interpolated to
 * the node (trigonometric along rings, barycentric along the diameter through the node).
 * L u + f = 0 with L = 1/2 tr(sigma sigma^T Hess) + b.grad + c reduces to this operator when
 * sigma = sqrt(2) I and b = 0, the only operators the reference applies it to
 * (main_test_serial.cpp:49-79 with BVPs/Franke_Poisson.hpp).
 */
#ifndef SPAC_PSEUDOESPECTRAL_CIRCULO_SOLVER_HPP
#define SPAC_PSEUDOESPECTRAL_CIRCULO_SOLVER_HPP

#include <vector>

#include "BVP.hpp"
#include "Nudo.hpp"
#include "spac_la.hpp"

namespace spac {

/* Interpolation of values on the polar grid (BVP/Interpolador_PS.cpp:18-61). */
class Interpolador_PS {
public:
    int m = 0;           /* points on a full diameter: 2 (N_radio + 1) / 2 */
    int N = 0;           /* angles */
    int n = 0;           /* N / 2: Fourier modes -n .. n-1 */
    double normN = 0.0;  /* 1 / N */
    std::vector<double> c, diametro; /* barycentric weights, diameter points / R (ascending) */
    Vec2 centro;
    double radio = 0.0;
    std::vector<double> r;  /* the r > 0 Chebyshev radii, r[0] = R */
    std::vector<double> t;  /* the angles */
    std::vector<double> z;  /* grid values, index i N + l (radius i, angle l) */
    /* the grid of Inicializa: radii r (descending from R), angles t */
    void Inicializa(const Vec2& centro, double radio, const std::vector<double>& r, const std::vector<double>& t);
    double Interpola(const Vec2& punto) const;
};

class PseudoespectralCirculoSolver {
public:
    Interpolador_PS interpolador;
    std::vector<double> rhs_pseudospectral;
    /* Solvers/PseudoespectralCirculoSolver.cpp:2-186.  N_circunferencia even, N_radio odd.
     * Returns 0, or -1 for bad sizes / a singular operator. */
    int Inicializa(const Vec2& centro, double radio, unsigned N_circunferencia, unsigned N_radio, double theta_0,
                   const BVP& Problema);
    /* Move the grid to another centre.  True when c takes the same values on the moved grid,
     * so that the operator and its factorisation are unchanged (then only the centre moves);
     * false leaves everything as it was (call Inicializa). */
    bool Reubica(const Vec2& centro, const BVP& Problema);
    /* :187-205: the grid solution with boundary data g and source f (into interpolador.z) */
    void Resuelve(const BVP& Problema);
    /* :206-260: B and G (columns: the periphery nodes' global indices) of the interior nodes */
    void Resuelve(const BVP& Problema, std::vector<Nudo>& nudos_interior, const std::vector<Nudo>& nudos_periferia);
    /* the unknowns: (N_radio + 1) / 2 radii times N_circunferencia angles */
    int size() const { return dim; }
    /* grid point k (index i N + l) */
    Vec2 punto(int k) const;
    /* solve L z = rhs with the factorisation of Inicializa */
    void solve(std::vector<double>& rhs) const;

private:
    int dim = 0;
    std::vector<double> lu;  /* L with its Dirichlet rows, LU-factored (partial pivoting), row-major */
    std::vector<int> piv;
    std::vector<double> cdiag; /* c at the grid points */
};

} // namespace spac

#endif
