"""CPU restatement (numpy, FP64) of the reference's pseudospectral interior-circle solver.

Test infrastructure only: the parity tests compare the library's solver
(spac_amd/cpp/PseudoespectralCirculoSolver.cpp) with this restatement, and this restatement
with the mathematics (Chebyshev / Fourier differentiation of polynomials, spectral accuracy on
the Franke problem, the row identity B + G u = u on a smooth solution).

Follows, step by step:
- PseudoespectralCirculoSolver::Inicializa (Solvers/PseudoespectralCirculoSolver.cpp:2-186):
  Chebyshev points r_i = R cos(pi i / N) on a diameter and their differentiation matrix (the
  cheb construction, :20-44); its square split into the r > 0 half and the mirrored half
  (:47-62); the periodic second-derivative matrix in theta (:63-81); the polar Laplacian
  kron(D1 + R E1, I) + kron(D2 + R E2, IZZI) + kron(R^2, D2t) (:82-126), unknowns ordered
  radius-major (index i M + l); c(x) added on the diagonal (:163-168); the M rows of r = R
  replaced by Dirichlet rows (:169-172);
- Resuelve(BVP, nudos_interior, nudos_periferia) (:206-260): B from zero boundary data and
  -f inside; column j of G from the Fourier cardinal of node j on the boundary, as -u_j at
  every interior node;
- Interpolador_PS::Interpola (BVP/Interpolador_PS.cpp:35-61): trigonometric interpolation along
  each ring at the point's angle and the opposite one, then barycentric interpolation along
  the diameter with the reference's weights c_i = (-1)^i sin(pi (2 i + 1) / (2 m))
  (:28-31);
- BVP::Cardinal_Fourier (BVP/BVP.cpp:25-29).
The reference factors with Eigen's FullPivHouseholderQR; this restatement solves with LAPACK's
LU (numpy.linalg.solve): the two agree to the conditioning of L times machine epsilon.
"""
from __future__ import annotations

import numpy as np


def cardinal_fourier(theta, theta_j, n):
    k = np.arange(-n, n)
    return np.cos(np.multiply.outer(np.asarray(theta) - theta_j, k)).sum(-1) / (2 * n)


def cheb(radius, N):
    """Chebyshev points r_i = R cos(pi i / N), i = 0..N, and their differentiation matrix."""
    i = np.arange(N + 1)
    r = radius * np.cos(np.pi * i / N)
    c = np.where((i == 0) | (i == N), 2.0, 1.0) * (-1.0) ** i
    R = np.repeat(r[:, None], N + 1, 1)
    C = np.outer(c, 1.0 / c)
    D = C / (R - R.T + np.eye(N + 1))
    D = D - np.diag(D.sum(1))
    return r, D


def fourier_d2(M, dtheta):
    """Periodic second-derivative matrix on M equispaced angles (Toeplitz; first column
    -pi^2 / (3 dtheta^2) - 1/6, then 0.5 (-1)^(k+1) / sin^2(k dtheta / 2))."""
    k = np.arange(M)
    col = np.empty(M)
    col[0] = -(np.pi ** 2) / (3 * dtheta ** 2) - 1.0 / 6.0
    col[1:] = 0.5 * (-1.0) ** (k[1:] + 1) / np.sin(dtheta * k[1:] * 0.5) ** 2
    return np.array([np.roll(col, i) for i in range(M)])


class PseudospectralCircle:
    """Inicializa: the operator L = Laplacian + c on the polar grid of a circle, with
    Dirichlet rows on r = R."""

    def __init__(self, centre, radius, M, N, theta0, c_fn=None):
        assert N % 2 == 1 and M % 2 == 0
        r, D = cheb(radius, N)
        N2, M2 = (N - 1) // 2, M // 2
        D2 = D @ D
        idx = np.arange(N2 + 1)
        D1, D2m = D2[np.ix_(idx, idx)], D2[np.ix_(idx, N - idx)]
        E1, E2 = D[np.ix_(idx, idx)], D[np.ix_(idx, N - idx)]
        dtheta = 2 * np.pi / M
        theta = np.arange(M) * dtheta + theta0
        D2t = fourier_d2(M, dtheta)
        Rd = np.diag(1.0 / r[: N2 + 1])
        IZZI = np.zeros((M, M))
        IZZI[:M2, M2:] = np.eye(M2)
        IZZI[M2:, :M2] = np.eye(M2)
        L = np.kron(D1 + Rd @ E1, np.eye(M)) + np.kron(D2m + Rd @ E2, IZZI) + np.kron(Rd @ Rd, D2t)
        self.centre = np.asarray(centre, float)
        self.radius = float(radius)
        self.M, self.N, self.N2 = M, N, N2
        self.r = r[: N2 + 1]
        self.theta = theta
        # grid points, radius-major: index i M + l
        self.x = (np.outer(self.r, np.cos(theta)) + self.centre[0]).ravel()
        self.y = (np.outer(self.r, np.sin(theta)) + self.centre[1]).ravel()
        if c_fn is not None:
            L[np.diag_indices_from(L)] += c_fn(self.x, self.y)
        L[:M, :] = 0.0
        L[np.arange(M), np.arange(M)] = 1.0
        self.L = L
        # barycentric interpolation along a diameter (the reference's weights)
        m = 2 * (N2 + 1)
        self.m = m
        self.diam = np.concatenate([-self.r, self.r[::-1]]) / self.radius
        i = np.arange(m)
        self.c = (-1.0) ** i * np.sin(np.pi * ((-1.0 + 1.0 / m) + (2.0 / m) * i + 1) / 2)

    def solve(self, rhs):
        return np.linalg.solve(self.L, rhs)

    def interpolate(self, z, px, py):
        """Interpolador_PS::Interpola of the grid values z (index i M + l) at (px, py)."""
        Z = np.asarray(z).reshape(self.N2 + 1, self.M)  # [radius i][angle l]
        dx, dy = px - self.centre[0], py - self.centre[1]
        rp = np.hypot(dx, dy) / self.radius
        tp = np.arctan2(dy, dx)
        tp = tp if tp >= 0 else tp + 2 * np.pi
        n = self.M // 2
        k = np.arange(-n, n)
        # the reference reads the ring angles from column 1 of its table (the same for every radius)
        w = np.cos(np.multiply.outer(tp - self.theta, k)).sum(1)
        wo = (np.cos(np.multiply.outer(tp - self.theta, k)) * np.cos(k * np.pi)).sum(1)
        head = Z @ w           # u(theta_p, r_i)
        tail = Z @ wo          # u(theta_p + pi, r_i)
        ud = np.concatenate([head, tail[::-1]])[::-1] / self.M
        t = self.c / (rp - self.diam)
        return float((t * ud).sum() / t.sum())

    def rows(self, f_fn, node_theta, interior_xy):
        """Resuelve(BVP, nudos_interior, nudos_periferia): B[k] and G[k, j] of the interior
        nodes over the circle's nodes (angles node_theta, in their order)."""
        M = self.M
        rhs = np.zeros(self.L.shape[0])
        rhs[M:] = -f_fn(self.x[M:], self.y[M:])
        zB = self.solve(rhs)
        B = np.array([self.interpolate(zB, x, y) for x, y in interior_xy])
        n = len(node_theta) // 2
        G = np.zeros((len(interior_xy), len(node_theta)))
        for j, tj in enumerate(node_theta):
            rhs = np.zeros(self.L.shape[0])
            rhs[:M] = cardinal_fourier(self.theta, tj, n)
            zj = self.solve(rhs)
            G[:, j] = [-self.interpolate(zj, x, y) for x, y in interior_xy]
        return B, G
