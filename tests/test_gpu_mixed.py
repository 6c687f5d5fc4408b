"""Mixed precision (FP32 stepping, FP64 Y/Z/xi sums and exit projection) against the
FP64 plain run.

The two modes draw the same Philox stream, so paths are coupled: a node's FP32
estimate differs from its FP64 estimate by the FP32 rounding of the path plus
the occasional flipped stop decision, far below the Monte-Carlo standard error.
The bar written here: per node |phi_32 - phi_64| < 0.25 standard errors + 2e-4,
and the mean difference over all nodes < 0.05 standard errors + 1e-5.
"""
import sys

import numpy as np
import pytest

from spac_amd import abi
from spac_amd.fk import FKProblem, problem_desc
from spac_amd.mesh import Mesh, gmres, mesh_desc
from spac_amd.problems import franke_u

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _grid16():
    t = -1.0 + 2.0 * (np.arange(16) + 1) / 17
    X, Y = np.meshgrid(t, t, indexing="ij")
    return X.ravel(), Y.ravel()


def _phi(st, key="phi"):
    v = np.array([s[key] for s in st])
    sq = np.array([s["phiphi" if key == "phi" else "phiphi_VR"] for s in st])
    n = np.array([s["paths"] for s in st])
    return v, np.sqrt(np.maximum(sq - v * v, 0) / n)


def test_fp32_points_track_fp64(gpu):
    say("fp32 vs fp64 points")
    x, y = _grid16()
    with FKProblem(problem_desc()) as p64, FKProblem(problem_desc(precision=abi.SPAC_FP32)) as p32:
        s64, _ = p64.solve_points(x, y, 10000, 1e-3, seed=0)
        s32, _ = p32.solve_points(x, y, 10000, 1e-3, seed=0)
    for key in ("phi", "phi_VR"):
        v64, se = _phi(s64, key)
        v32, _ = _phi(s32, key)
        d = v32 - v64
        say(f"{key}: max |d| {np.abs(d).max():.3e}, max |d|/se {np.abs(d / se).max():.3f}, mean d {d.mean():.2e}")
        assert np.all(np.abs(d) < 0.25 * se + 2e-4)
        assert abs(d.mean()) < 0.05 * se.mean() + 1e-5
    steps64 = sum(s["path_steps"] for s in s64)
    steps32 = sum(s["path_steps"] for s in s32)
    assert abs(steps32 / steps64 - 1) < 1e-3


def test_fp32_rows_track_fp64(gpu):
    say("fp32 vs fp64 rows (config 2)")
    m = Mesh(mesh_desc())
    sol = {}
    for prec in (abi.SPAC_FP64, abi.SPAC_FP32):
        with FKProblem(problem_desc(coef=abi.SPAC_COEF_VARCOEF_FRANKE, domain=(0.0, 0.0, 1.0, 1.0),
                                    precision=prec)) as p:
            rows = p.assemble_rows(m, 20000, 5e-4, seed=0)
        m.accumulate(rows)
        Gi, Gj, Gv, B = m.csr()
        u, info = gmres(Gi, Gj, Gv, B, tol=1e-10, max_iter=5000, restart=200)
        assert info.converged
        sol[prec] = (u, rows)
    du = sol[abi.SPAC_FP32][0] - sol[abi.SPAC_FP64][0]
    err64 = sol[abi.SPAC_FP64][0] - franke_u(m.node_x, m.node_y)
    say(f"rows: fp32-fp64 solution rms {np.sqrt(np.mean(du ** 2)):.3e} max {np.abs(du).max():.3e}; "
        f"fp64 rms err vs exact {np.sqrt(np.mean(err64 ** 2)):.3e}")
    assert np.sqrt(np.mean(du ** 2)) < 0.1 * np.sqrt(np.mean(err64 ** 2))
    dB = sol[abi.SPAC_FP32][1]["B"] - sol[abi.SPAC_FP64][1]["B"]
    assert np.abs(dB).max() < 2e-3
