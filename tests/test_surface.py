"""The drop-in C++ surface (BVP / Coeficiente_* / Frontera / FeynmanKacSolver /
Malla / SistemaLineal) driven the way the reference's serial driver drives it
(main_test_serial.cpp:11-83), from a separately compiled C++ program.
"""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as o
from spac_amd import abi
from spac_amd.fk import SolveBatchArrays, problem_desc
from spac_amd.mesh import Mesh, mesh_desc

DRIVER = Path(__file__).resolve().parent / "cpp" / "_bin" / "surface_driver"


def run_driver(mode):
    return run_driver_args(mode)


def run_driver_args(*args):
    assert DRIVER.exists(), "build() compiles tests/cpp/surface_driver.cpp"
    out = subprocess.run([str(DRIVER), *args], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout)


def test_surface_cpu():
    r = run_driver("cpu")
    assert r["mesh_rc"] == 0 and r["nodes"] == 408
    assert r["coef_rc"] == 0 and r["coef_kind"] == abi.SPAC_COEF_FRANKE_POISSON
    assert r["dom_kind"] == abi.SPAC_BOUNDARY_RECTANGLE and r["sub_kind"] == abi.SPAC_BOUNDARY_CIRCLE
    assert r["bad_rc"] == abi.SPAC_E_UNREGISTERED_PLUGIN
    # Update_G (host, direct form): the IMQ cardinal functions sum to ~1 on the arc
    assert r["update_g_rc"] == 0 and abs(r["update_g_sum"] + 1.0) < 1e-4
    m = Mesh(mesh_desc(domain=(0.0, 0.0, 1.25, 1.25), n_side=4, nodes_per_circle=44, overlap=1.0, theta0=0.1428,
                       rbf_kind=abi.SPAC_RBF_IMQ_GLOBAL, fk_all_circles=False))
    assert r["nnz"] == m.info.nnz
    # TablaDeValores + Interpolador_Tabla (513^2 grid of Franke u on [0, 1.25]^2): bilinear
    # errors of order h^2 |D2 u| / 8 for u and h^2 |D3 u| for the gradient; the base class is NaN
    assert r["interp_err_u"] < 2e-4
    assert r["interp_err_grad"] < 3e-3 * r["grad_max"]
    assert r["interp_err_grad_fd"] < 5e-3 * r["grad_max"]
    assert r["base_nan"] == 1
    # Frontera::Absorbente false on the left side -> SPAC_REFLECT_XMIN; a side whose type
    # changes along it has no device form; Stopping everywhere -> no reflecting side
    assert r["refl"] == [0, abi.SPAC_REFLECT_XMIN]
    assert r["refl_mixed_rc"] == abi.SPAC_E_UNREGISTERED_PLUGIN
    assert r["refl_stopping"] == [0, 0]


@pytest.mark.gpu
def test_surface_gpu_rows_match_restatement(gpu):
    print("surface driver gpu", flush=True)
    r = run_driver("gpu")
    assert r["worst_rc"] == 0
    m = Mesh(mesh_desc(domain=(0.0, 0.0, 1.25, 1.25), n_side=4, nodes_per_circle=44, overlap=1.0, theta0=0.1428,
                       rbf_kind=abi.SPAC_RBF_IMQ_GLOBAL, fk_all_circles=False))
    jobs = np.arange(m.info.n_jobs)
    assert np.array_equal(m.job_row[jobs], r["rows"])
    live = m.job_frontera == 0
    streams = np.cumsum(live) - 1  # one Philox stream per FK call, in call order
    desc = problem_desc(coef=abi.SPAC_COEF_FRANKE_POISSON, domain=(0.0, 0.0, 1.25, 1.25))
    B, G, st = o.assemble_rows(desc, m, 400, 1e-3, seed=0, jobs=jobs, streams=np.maximum(streams, 0))
    a = m.iface_arrays()
    M = a["node_offset"][1:] - a["node_offset"][:-1]
    off = np.concatenate([[0], np.cumsum(M[m.job_iface])])
    Bd, Gs = np.array(r["B"]), np.array(r["Gsum"])
    assert r["calls"] == int(live.sum())
    for k in range(jobs.size):
        if live[k]:
            assert Bd[k] == B[k]
            assert Gs[k] == pytest.approx(G[off[k]:off[k + 1]].sum(), rel=1e-12, abs=1e-14)
        else:  # B = g(x) with the host plugin (libm exp) vs the device twin: within 2 ulp
            assert Bd[k] == pytest.approx(B[k], rel=5e-16, abs=0) and Gs[k] == 0.0
    # point solve through Solve_OMP_Analytic(X0, ...)
    b = SolveBatchArrays([0.6], [0.7], 2000, 1e-3, [0])
    so = o.reduce(b, o.simulate(problem_desc(domain=(0.0, 0.0, 1.25, 1.25)), b, seed=0))[0]
    assert r["point_rc"] == 0 and r["point_phi"] == so["phi"] and r["point_phi_VR"] == so["phi_VR"]
    assert r["point_steps"] == so["path_steps"]


@pytest.mark.gpu
def test_surface_gpu_tabulated_control_variate(gpu):
    """FeynmanKacSolver with the plugin control variate, none, and a tabulated u_hat
    (Interpolador_Tabla of the exact Franke u): the same paths, so phi is identical; the
    table reproduces the plugin's variance reduction."""
    print("surface driver cv", flush=True)
    r = run_driver("cv")
    assert r["rc"] == [0, 0, 0]
    plugin, none, table = r["modes"]
    assert plugin["phi"] == none["phi"] == table["phi"]
    var = plugin["phiphi"] - plugin["phi"] ** 2
    var_plugin = plugin["phiphi_VR"] - plugin["phi_VR"] ** 2
    var_table = table["phiphi_VR"] - table["phi_VR"] ** 2
    assert none["xixi"] == 0.0 and none["phi_VR"] == none["phi"]
    assert var_plugin < 0.25 * var
    assert abs(var_table / var_plugin - 1.0) < 0.05
    assert abs(table["phi_VR"] - plugin["phi_VR"]) < 1e-3
    assert abs(table["phi_VR"] - r["u"]) < 4 * np.sqrt(var_table / 4000) + 5e-3


@pytest.mark.gpu
def test_surface_gpu_reflecting_side(gpu):
    """Solve_OMP_Analytic with the Robin plugins and Absorbente false on the left
    side: the surface resolves FRANKE_ROBIN with SPAC_REFLECT_XMIN and rho, and
    equals the restatement bit for bit."""
    print("surface driver reflect", flush=True)
    r = run_driver("reflect")
    assert r["coef_rc"] == 0 and r["coef_kind"] == abi.SPAC_COEF_FRANKE_ROBIN and r["rc"] == 0
    desc = problem_desc(coef=abi.SPAC_COEF_FRANKE_ROBIN, domain=(0.0, 0.0, 1.0, 1.0),
                        reflecting=abi.SPAC_REFLECT_XMIN, rho=0.08)
    b = SolveBatchArrays([0.05], [0.5], 4000, 1e-3, [0])
    so = o.reduce(b, o.simulate(desc, b, seed=0))[0]
    assert r["phi"] == so["phi"] and r["phi_VR"] == so["phi_VR"] and r["steps"] == so["path_steps"]


def test_surface_rbf_interpolator_and_numeric_classes():
    """Interpolador_RBF (the reference's stub class, BVP/Interpolador_RBF.cpp:2-52, made real) and
    the numeric coefficient classes (BVP/Coeficiente_*_Numerico.hpp), on the host.

    Pins: interpolation at the centres; linear functions reproduced exactly with their
    derivatives (the polynomial augmentation); analytic first and second derivatives, Cartesian
    and polar, equal central differences of the interpolant; the error on Franke's function falls
    as centres are added; repeated centres are refused."""
    r = run_driver("rbf")
    assert r["rc"] == 0
    assert r["bad_rc"] == -1 and r["bad_nan"] == 1
    assert r["at_centres"] < 1e-10
    assert r["lin_err"] < 1e-10 and r["lin_d1"] < 1e-9 and r["lin_d2"] < 1e-8
    assert r["fd_err"] < 1e-4 and r["polar_err"] < 1e-4
    assert r["e1600"] < 0.6 * r["e400"] and r["e1600"] < 1e-2
    # Es_Numerico: numeric scalar, a numeric scalar copied into a base-class field, matrix, vector;
    # compiled coefficients are not numeric
    assert r["numerico"] == [1, 1, 1, 1, 0, 0]
    assert r["sliced_value"] == r["direct_value"]
    s = r["direct_value"]
    assert r["matrix"] == [s, 0.0, 0.0, s] and r["vector"] == [s, 0.0]
    # numeric coefficients over the Franke-Poisson set have no device form; the exact-match
    # resolver refuses any numeric coefficient; over the variable-coefficient set they resolve
    # (sigma | c tabulated, F left at the default zero); a numeric F is refused
    assert r["num_rc_poisson"] == abi.SPAC_E_UNREGISTERED_PLUGIN
    assert r["plain_rc_numeric"] == abi.SPAC_E_UNREGISTERED_PLUGIN
    assert r["num_rc"] == 0 and r["num_kind"] == abi.SPAC_COEF_VARCOEF_FRANKE
    assert r["num_mask"] == abi.SPAC_TAB_SIGMA | abi.SPAC_TAB_KILLING and r["num_F_zero"] == 1
    assert r["numF_rc"] == abi.SPAC_E_UNREGISTERED_PLUGIN


@pytest.mark.gpu
def test_surface_gpu_numeric_coefficients(gpu, tmp_path):
    """FeynmanKacSolver with Coeficiente_*_Numerico for sigma, b, c, f, g, u (Interpolador_Tabla
    of the variable-coefficient set on the solver's 513 x 513 grid): the sampled table, fed to
    the restatement, gives the same estimators bit for bit; against the compiled set on the same
    paths the estimators move by the interpolation error only; a numeric c from scattered RBF
    data is sampled through Interpolador_RBF; the refusals."""
    print("surface driver numeric", flush=True)
    tab = tmp_path / "table.bin"
    r = run_driver_args("numeric", str(tab))
    assert r["rc"] == [0, 0, 0, 0, 0] and r["rbf_rc"] == 0
    assert r["mask"] == abi.SPAC_TAB_ALL and r["same_table"] == 1
    vals = np.fromfile(tab, dtype=np.float64).reshape(513, 513, 8)
    dom = (0.0, 0.0, 1.0, 1.0)
    desc = problem_desc(coef=abi.SPAC_COEF_VARCOEF_FRANKE, domain=dom)
    b = SolveBatchArrays([0.6], [0.7], 4000, 1e-3, [0])
    o.set_cv(abi.SPAC_CV_NONE)
    o.set_coef_tables(abi.SPAC_TAB_ALL, vals, dom)
    try:
        so = o.reduce(b, o.simulate(desc, b, seed=0))[0]
    finally:
        o.set_coef_tables()
        o.set_cv()
    n, c = r["numeric"], r["compiled"]
    assert n["phi"] == so["phi"] and n["phiphi"] == so["phiphi"] and n["steps"] == so["path_steps"]
    se = np.sqrt((c["phiphi"] - c["phi"] ** 2) / 4000)
    assert abs(n["phi"] - c["phi"]) < 0.05 * se + 1e-4
    assert abs(n["steps"] - c["steps"]) <= 0.001 * c["steps"]
    assert r["rbf"]["tab_err"] == 0.0  # the table holds Interpolador_RBF's own values at the nodes
    assert abs(r["rbf"]["phi"] - c["phi"]) < 4 * se
    assert r["bad_sigma"] == abi.SPAC_E_INVALID_ARG
    assert r["bad_F"] == abi.SPAC_E_UNREGISTERED_PLUGIN
    assert r["bad_refl"] == abi.SPAC_E_INVALID_ARG
    # F left at its default zero: the plugin control variate is xi = 0
    assert r["F_zero"]["phi_VR"] == r["F_zero"]["phi"] and r["F_zero"]["xixi"] == 0.0
