"""Pseudospectral rows against Feynman–Kac rows on the same full interior circles, and the
reference's serial pipeline (FK rows on perimeter circles, pseudospectral rows inside,
main_test_serial.cpp:51-81) solved end to end.

Bars (Franke Poisson on [0, 1.25]^2, 4 x 4 circles of 44 nodes, FP64 paths, 20,000 per row):
- every interior-circle row predicts u(x_k) = B_k - G_k u on the Franke solution; the FK
  prediction differs from the pseudospectral one by less than 4 standard errors of the FK
  estimator plus the stencil interpolation tolerance 5e-4;
- B (the source part) agrees within the same bound;
- the mixed system solves to the Franke solution as closely as the all-FK system does.
"""
import sys

import numpy as np
import pytest

from spac_amd import abi
from spac_amd.fk import FKProblem, problem_desc
from spac_amd.mesh import Mesh, gmres, mesh_desc, ps_assemble_rows
from spac_amd.problems import franke_u

pytestmark = pytest.mark.gpu

FP = abi.SPAC_COEF_FRANKE_POISSON
DOM = (0.0, 0.0, 1.25, 1.25)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_pseudospectral_rows_match_fk_rows(gpu):
    m = Mesh(mesh_desc(domain=DOM, n_side=4, nodes_per_circle=44, theta0=0.1428))
    a = m.iface_arrays()
    inner = (a["code"][m.job_iface] == 1) & (m.job_frontera == 0)
    jobs = np.nonzero(inner)[0]
    paths, h = 20000, 1e-4
    desc = problem_desc(coef=FP, domain=DOM)
    with FKProblem(desc) as p:
        p.set_control_variate(abi.SPAC_CV_NONE)
        fk = p.assemble_rows(m, paths, h, seed=3, jobs=jobs)
        fk_all = p.assemble_rows(m, paths, h, seed=3)
    ps, rep = ps_assemble_rows(m, FP, 43, jobs=jobs)
    say(f"{jobs.size} interior-circle rows; pseudospectral {rep.seconds:.2f} s, {rep.factorizations} factorisation")
    u = franke_u(m.node_x, m.node_y)
    goff = fk["G_offset"]
    assert np.array_equal(goff, ps["G_offset"]) and np.array_equal(fk["G_col"], ps["G_col"])

    def predict(r):
        return np.array([r["B"][i] - r["G"][goff[i]:goff[i + 1]] @ u[r["G_col"][goff[i]:goff[i + 1]]]
                         for i in range(jobs.size)])

    v_fk, v_ps = predict(fk), predict(ps)
    exact = franke_u(m.job_x[jobs], m.job_y[jobs])
    se = np.array([np.sqrt(max(s["phiphi"] - s["phi"] ** 2, 0.0) / s["paths"]) for s in fk["stats"]])
    d = np.abs(v_fk - v_ps)
    say(f"|v_FK - v_PS|: max {d.max():.2e} (max/SE {np.max(d / se):.2f}), |v_PS - u| max "
        f"{np.abs(v_ps - exact).max():.2e}; B: max |dB| {np.abs(fk['B'] - ps['B']).max():.2e}")
    assert np.abs(v_ps - exact).max() < 2e-6
    assert np.all(d < 4 * se + 5e-4)
    assert np.all(np.abs(fk["B"] - ps["B"]) < 4 * se + 5e-4)
    # the reference's serial pipeline: FK rows on perimeter circles, pseudospectral rows inside
    pos = {j: i for i, j in enumerate(jobs)}
    mixed = {k: v.copy() if hasattr(v, "copy") else v for k, v in fk_all.items()}
    go_all = fk_all["G_offset"]
    for j in jobs:
        i = pos[j]
        mixed["B"][j] = ps["B"][i]
        mixed["G"][go_all[j]:go_all[j + 1]] = ps["G"][goff[i]:goff[i + 1]]
    errs = {}
    for name, rows in (("all FK", fk_all), ("FK + pseudospectral", mixed)):
        m.accumulate(rows)
        Gi, Gj, Gv, B = m.csr()
        sol, info = gmres(Gi, Gj, Gv, B, tol=1e-10, max_iter=5000, restart=200, device=True)
        assert info.converged
        errs[name] = float(np.sqrt(np.mean((sol - u) ** 2)))
    say(f"solved u RMS error vs Franke: {errs}")
    assert errs["FK + pseudospectral"] < 1.5 * errs["all FK"] + 1e-4
