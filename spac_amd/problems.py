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
