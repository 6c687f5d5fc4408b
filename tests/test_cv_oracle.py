"""Pins for the oracle's control-variate modes (CPU only).

What fixes each part, independently of the oracle's own formulas:
- plain mode (F = 0): same paths as the plugin mode (xi does not feed back),
  xi identically 0, B = mean of Z + Y g over domain exits and Z over circle
  exits, recomputed here from the records;
- table mode: a table of the exact Franke u and gradient (numpy, the
  separately pinned derivatives) reproduces the plugin's xi per path to
  interpolation accuracy and u at circle exits; a zero table equals plain MC;
- fallback: a deliberately wrong table (gradient of the wrong sign, 5x too
  large) adds variance, every solve falls back, and B equals plain MC's B;
  with the exact table nothing falls back;
- unbiasedness: with any table E[xi] = 0 (martingale), within 4 standard errors.
"""
import numpy as np
import pytest

import oracle_lib as o
from spac_amd import abi
from spac_amd.fk import SolveBatchArrays, problem_desc
from spac_amd.problems import franke_table_fn, franke_u

VC = abi.SPAC_COEF_VARCOEF_FRANKE
DOM = (0.0, 0.0, 1.0, 1.0)
NP = 1500


@pytest.fixture(autouse=True)
def reset_cv():
    o.set_cv()
    yield
    o.set_cv()


def _batch():
    """Node-mode solves: three starts inside a circle that pokes out of the unit square."""
    x0 = np.array([0.30, 0.55, 0.72])
    y0 = np.array([0.40, 0.50, 0.22])
    circle = np.tile([0.33, 0.55, 0.45], 3)
    return SolveBatchArrays(x0, y0, NP, 5e-4, [5, 6, 7], circle=circle)


def _run(mode, table=None, fallback=False):
    o.set_cv(mode, fallback=fallback, table=table, domain=DOM if table is not None else None)
    b = _batch()
    desc = problem_desc(coef=VC, domain=DOM)
    rec = o.simulate(desc, b, seed=3)
    st = o.reduce(b, rec, node_mode=True)
    return b, rec, st


@pytest.fixture(scope="module")
def exact_table():
    return o.tabulate(franke_table_fn, 2049, DOM)


def test_plain_mode_keeps_paths_and_zeroes_xi():
    _, ra, sa = _run(abi.SPAC_CV_ANALYTIC)
    _, rn, sn = _run(abi.SPAC_CV_NONE)
    for k in ("x_tau", "y_tau", "Y", "Z", "tau", "gval", "steps", "outcome"):
        assert np.array_equal(ra[k], rn[k]), k
    assert np.all(rn["xi"] == 0.0)
    assert np.any(ra["xi"] != 0.0)
    sub = rn["outcome"] == abi.SPAC_OUT_SUBDOMAIN
    terms = np.where(sub, rn["Z"], rn["Z"] + rn["Y"] * rn["gval"])
    for k in range(3):
        sl = slice(k * NP, (k + 1) * NP)
        assert abs(sn[k]["B"] - terms[sl].mean()) < 1e-13
        assert sn[k]["cv_fallback"] == 0


def test_exact_table_reproduces_plugin(exact_table):
    _, ra, _ = _run(abi.SPAC_CV_ANALYTIC)
    _, rt, _ = _run(abi.SPAC_CV_TABLE, exact_table)
    for k in ("x_tau", "y_tau", "Y", "Z", "steps", "outcome"):
        assert np.array_equal(ra[k], rt[k]), k
    dxi = np.abs(rt["xi"] - ra["xi"])
    assert dxi.max() < 1e-4, dxi.max()
    assert np.abs(ra["xi"]).max() > 1e-2
    sub = rt["outcome"] == abi.SPAC_OUT_SUBDOMAIN
    assert sub.sum() > 100
    assert np.abs(rt["gval"][sub] - franke_u(rt["x_tau"][sub], rt["y_tau"][sub])).max() < 2e-5
    assert np.array_equal(rt["gval"][~sub], ra["gval"][~sub])


def test_zero_table_is_plain_mc():
    zero = np.zeros((16, 16, 4), dtype=np.float32)
    _, rn, _ = _run(abi.SPAC_CV_NONE)
    _, rz, _ = _run(abi.SPAC_CV_TABLE, zero)
    assert np.all(rz["xi"] == 0.0)
    sub = rz["outcome"] == abi.SPAC_OUT_SUBDOMAIN
    assert np.all(rz["gval"][sub] == 0.0)
    assert np.array_equal(rz["gval"][~sub], rn["gval"][~sub])


def test_fallback_on_a_harmful_table(exact_table):
    bad = exact_table.copy()
    bad[..., 1:3] *= -5.0
    _, rn, sn = _run(abi.SPAC_CV_NONE)
    _, rb, sb = _run(abi.SPAC_CV_TABLE, bad, fallback=True)
    for k in range(3):
        var = sb[k]["phiphi"] - sb[k]["phi"] ** 2
        var_vr = sb[k]["phiphi_VR"] - sb[k]["phi_VR"] ** 2
        assert var_vr > var
        assert sb[k]["cv_fallback"] == 1
        # dropping xi leaves the plain estimator (circle-exit u_hat does not enter B)
        assert abs(sb[k]["B"] - sn[k]["B"]) < 1e-12
        # the martingale term has mean zero whatever the table
        xi = rb["xi"][k * NP:(k + 1) * NP]
        assert abs(xi.mean()) < 4 * xi.std() / np.sqrt(NP)


def test_no_fallback_with_the_exact_table(exact_table):
    _, _, s0 = _run(abi.SPAC_CV_TABLE, exact_table, fallback=False)
    _, _, s1 = _run(abi.SPAC_CV_TABLE, exact_table, fallback=True)
    for k in range(3):
        assert s1[k]["cv_fallback"] == 0
        assert s1[k]["B"] == s0[k]["B"]
        var = s1[k]["phiphi"] - s1[k]["phi"] ** 2
        var_vr = s1[k]["phiphi_VR"] - s1[k]["phi_VR"] ** 2
        assert var_vr < 0.2 * var
