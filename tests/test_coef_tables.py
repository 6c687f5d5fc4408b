"""Tabulated coefficients (spac_fk_set_coef_tables, spac_tab.h): the reference's numeric
coefficients, Coeficiente_*_Numerico over an Interpolador (BVP/Coeficiente_Escalar_Numerico.hpp:6-19).

CPU (the restatement, pinned by the mathematics of bilinear interpolation):
- a table of a linear function is exact everywhere inside the grid, and takes the nearest
  edge value outside it;
- a table of the compiled variable-coefficient set reproduces every coefficient within the
  bilinear bound h^2/8 max|second derivative|, and the restatement's estimators with it stay
  within the interpolation error of those without it.
GPU (tests/test_coef_tables_gpu.py): the FP64 kernels with tables equal the restatement with
tables bit for bit.
"""
import numpy as np
import pytest

import oracle_lib as o
from spac_amd import abi
from spac_amd.fk import SolveBatchArrays, problem_desc

VC = abi.SPAC_COEF_VARCOEF_FRANKE
DOM = (0.0, 0.0, 1.0, 1.0)


def compiled_table(T, dom=DOM):
    """(T, T, 8): the compiled variable-coefficient set at the grid nodes, [j, i] = (x_i, y_j)."""
    o.set_coef_tables()
    gx, gy = np.linspace(dom[0], dom[2], T), np.linspace(dom[1], dom[3], T)
    X, Y = np.meshgrid(gx, gy, indexing="xy")
    x, y = X.ravel(), Y.ravel()
    cols = [o.coef_eval(VC, w, x, y) for w in (4, 5, 6, 7, 1, 14, 0)]
    cols.append(np.zeros_like(x))
    return np.stack(cols, -1).reshape(T, T, 8)


@pytest.fixture(autouse=True)
def _reset():
    yield
    o.set_coef_tables()


def test_linear_data_is_exact_and_edges_clamp():
    T = 17
    gx = np.linspace(-1.0, 2.0, T)
    X, Y = np.meshgrid(gx, gx, indexing="xy")
    vals = np.zeros((T, T, 8))
    for k in range(7):
        vals[..., k] = (k + 1) * 0.25 + (k - 3) * 0.5 * X - 0.125 * (k + 2) * Y
    o.set_coef_tables(abi.SPAC_TAB_ALL, vals, (-1.0, -1.0, 2.0, 2.0))
    rng = np.random.default_rng(0)
    x, y = rng.uniform(-1.0, 2.0, 500), rng.uniform(-1.0, 2.0, 500)
    for k, w in ((3, 7), (4, 1), (5, 14), (6, 0), (1, 5), (2, 6), (0, 4)):
        got = o.coef_eval(VC, w, x, y)
        exact = (k + 1) * 0.25 + (k - 3) * 0.5 * x - 0.125 * (k + 2) * y
        assert np.abs(got - exact).max() < 1e-13, k
    # outside: the nearest edge value
    got = o.coef_eval(VC, 7, np.array([-5.0, 7.0]), np.array([0.5, 3.0]))
    exact = [4 * 0.25 + 0 - 0.125 * 5 * 0.5, 4 * 0.25 + 0 * 2.0 - 0.125 * 5 * 2.0]
    assert np.abs(got - exact).max() < 1e-13


def _second_derivative_bounds(w, hh=2e-3):
    """max |d2/dx2| and max |d2/dy2| of coefficient w on [0, 1]^2, by central differences on a
    fine grid (times 1.25 for the grid's own error)."""
    g = np.linspace(hh, 1.0 - hh, 400)
    X, Y = np.meshgrid(g, g, indexing="xy")
    x, y = X.ravel(), Y.ravel()
    f0 = o.coef_eval(VC, w, x, y)
    fxx = (o.coef_eval(VC, w, x + hh, y) - 2 * f0 + o.coef_eval(VC, w, x - hh, y)) / hh ** 2
    fyy = (o.coef_eval(VC, w, x, y + hh) - 2 * f0 + o.coef_eval(VC, w, x, y - hh)) / hh ** 2
    return 1.25 * np.abs(fxx).max(), 1.25 * np.abs(fyy).max()


def test_compiled_coefficients_within_the_bilinear_bound():
    """|bilinear - f| <= hx^2/8 max|f_xx| + hy^2/8 max|f_yy| for every tabulated coefficient."""
    T = 257
    vals = compiled_table(T)
    rng = np.random.default_rng(1)
    x, y = rng.uniform(0, 1, 3000), rng.uniform(0, 1, 3000)
    exact = {w: o.coef_eval(VC, w, x, y) for w in (4, 5, 6, 7, 1, 0, 14)}
    bounds = {w: _second_derivative_bounds(w) for w in exact}
    o.set_coef_tables(abi.SPAC_TAB_ALL, vals, DOM)
    h = 1.0 / (T - 1)
    for w, (mxx, myy) in bounds.items():
        got = o.coef_eval(VC, w, x, y)
        bound = h * h / 8 * (mxx + myy) + 1e-12
        err = np.abs(got - exact[w]).max()
        assert err <= bound, (w, err, bound)
        if w in (5, 6):  # b is linear: exact
            assert err < 1e-13
    assert np.abs(exact[1]).max() > 1.0  # f is not trivially small


def test_restatement_estimators_with_a_table_of_the_compiled_set():
    """The restatement's estimators with the compiled set tabulated (every coefficient) differ
    from the compiled ones only by the interpolation error (same Philox paths)."""
    desc = problem_desc(coef=VC, domain=DOM)
    n = 6
    rng = np.random.default_rng(2)
    cx, cy = rng.uniform(0.3, 0.7, n), rng.uniform(0.3, 0.7, n)
    b = SolveBatchArrays(cx + 0.05, cy, 2000, 1e-3, np.arange(n), circle=np.stack([np.full(n, 0.15), cx, cy], 1))
    o.set_cv(abi.SPAC_CV_NONE)
    try:
        ref = o.reduce(b, o.simulate(desc, b, seed=5), node_mode=True)
        o.set_coef_tables(abi.SPAC_TAB_ALL, compiled_table(513), DOM)
        tab = o.reduce(b, o.simulate(desc, b, seed=5), node_mode=True)
    finally:
        o.set_cv()
    for r, t in zip(ref, tab):
        se = np.sqrt(max(r["phiphi"] - r["phi"] ** 2, 0.0) / r["paths"])
        assert abs(r["phi"] - t["phi"]) < 0.05 * se + 1e-5
        assert abs(r["B"] - t["B"]) < 0.05 * se + 1e-5
        assert abs(r["path_steps"] - t["path_steps"]) <= 0.001 * r["path_steps"]
