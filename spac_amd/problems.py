"""Exact solutions of the shipped problems (numpy), for error reports.

Franke's test function is the exact u of the Franke Poisson problem
(BVPs/Franke_Poisson.hpp:16-18) and the manufactured u of the
variable-coefficient operator.
"""
import numpy as np


def franke_u(x, y):
    return (0.75 * np.exp(-0.25 * ((9 * x - 2) ** 2 + (9 * y - 2) ** 2))
            + 0.75 * np.exp(-((9 * x + 1) ** 2) / 49 - 0.1 * (9 * y + 1))
            + 0.5 * np.exp(-0.25 * ((9 * x - 7) ** 2 + (9 * y - 3) ** 2))
            - 0.2 * np.exp(-((9 * x - 4) ** 2) - (9 * y - 7) ** 2))


def franke_grad(x, y):
    """(du/dx, du/dy) of franke_u."""
    e1 = np.exp(-0.25 * ((9 * x - 2) ** 2 + (9 * y - 2) ** 2))
    e2 = np.exp(-((9 * x + 1) ** 2) / 49 - 0.1 * (9 * y + 1))
    e3 = np.exp(-0.25 * ((9 * x - 7) ** 2 + (9 * y - 3) ** 2))
    e4 = np.exp(-((9 * x - 4) ** 2) - (9 * y - 7) ** 2)
    ux = (0.75 * e1 * (-4.5 * (9 * x - 2)) + 0.75 * e2 * (-18 * (9 * x + 1) / 49)
          + 0.5 * e3 * (-4.5 * (9 * x - 7)) - 0.2 * e4 * (-18 * (9 * x - 4)))
    uy = (0.75 * e1 * (-4.5 * (9 * y - 2)) + 0.75 * e2 * (-0.9)
          + 0.5 * e3 * (-4.5 * (9 * y - 3)) - 0.2 * e4 * (-18 * (9 * y - 7)))
    return ux, uy


def franke_table_fn(x, y):
    ux, uy = franke_grad(x, y)
    return franke_u(x, y), ux, uy
