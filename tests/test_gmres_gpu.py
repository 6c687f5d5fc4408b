"""GPU GMRES (spac_gmres_solve_gpu) against scipy's direct solve and the host GMRES.

Same method as the host solver (restarted GMRES(m), right Jacobi
preconditioning) with CGS2 instead of modified Gram–Schmidt, so iterates agree
to rounding and both reach the requested relative residual; the bar is the
residual itself and agreement with the direct solution.
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from spac_amd import abi
from spac_amd.fk import FKProblem, problem_desc
from spac_amd.mesh import Mesh, gmres, mesh_desc

GOLDEN = Path(__file__).resolve().parent / "golden"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_gpu_gmres_rejects_bad_csr():
    """Argument checks run before any device call."""
    Gi = np.array([0, 1, 2], np.int32)
    Gj = np.array([0, 5], np.int32)  # column 5 outside a 2x2 system
    Gv = np.ones(2)
    B = np.ones(2)
    with pytest.raises(abi.SpacError) as e:
        gmres(Gi, Gj, Gv, B, device=True)
    assert e.value.code == abi.SPAC_E_INVALID_ARG
    with pytest.raises(abi.SpacError):
        gmres(np.array([0, 2, 1], np.int32), np.array([0, 1], np.int32), Gv, B, device=True)


@pytest.mark.gpu
def test_gpu_gmres_fixture(gpu):
    say("gpu gmres: committed fixture")
    fx = json.loads((GOLDEN / "gmres_system.json").read_text())
    Gi, Gj, Gv, B = (np.array(fx[k], dtype=t) for k, t in (("G_i", np.int32), ("G_j", np.int32), ("G_ij", float),
                                                          ("B", float)))
    u_ref = np.array(fx["u_spsolve"])
    u, info = gmres(Gi, Gj, Gv, B, tol=1e-12, max_iter=5000, restart=200, device=True)
    assert info.converged and info.residual < 1e-11
    assert np.abs(u - u_ref).max() < 1e-9 * np.abs(u_ref).max()
    A = sp.csr_matrix((Gv, Gj, Gi), shape=(B.size, B.size))
    assert np.linalg.norm(A @ u - B) < 1e-11 * np.linalg.norm(B)


@pytest.mark.gpu
def test_gpu_gmres_random_nonsymmetric_with_restarts(gpu):
    say("gpu gmres: random nonsymmetric, restart 15")
    rng = np.random.default_rng(0)
    n = 3000
    A = (sp.random(n, n, density=0.002, random_state=1, format="csr") + sp.eye(n) * 2.0).tocsr()
    A.sort_indices()
    b = rng.normal(size=n)
    u, info = gmres(A.indptr, A.indices, A.data, b, tol=1e-12, max_iter=4000, restart=15, device=True)
    assert info.converged and info.restarts > 1
    assert np.allclose(u, spl.spsolve(A.tocsc(), b), rtol=1e-9, atol=1e-10)
    # warm start from the solution: converged at once
    u2, info2 = gmres(A.indptr, A.indices, A.data, b, tol=1e-10, restart=15, u0=u, device=True)
    assert info2.iterations == 0 and info2.converged


@pytest.mark.gpu
def test_gpu_gmres_matches_host_on_interface_system(gpu):
    say("gpu gmres vs host gmres: config-2 rows")
    m = Mesh(mesh_desc())
    with FKProblem(problem_desc(coef=abi.SPAC_COEF_VARCOEF_FRANKE, domain=(0.0, 0.0, 1.0, 1.0),
                                precision=abi.SPAC_FP32)) as p:
        rows = p.assemble_rows(m, 2000, 5e-4, seed=0)
    m.accumulate(rows)
    Gi, Gj, Gv, B = m.csr()
    uh, ih = gmres(Gi, Gj, Gv, B, tol=1e-10, max_iter=5000, restart=200)
    ud, idv = gmres(Gi, Gj, Gv, B, tol=1e-10, max_iter=5000, restart=200, device=True)
    say(f"host: {ih.iterations} it {ih.seconds * 1e3:.1f} ms; gpu: {idv.iterations} it {idv.seconds * 1e3:.1f} ms "
        f"(device {idv.device_ms:.1f} ms)")
    assert ih.converged and idv.converged
    assert abs(idv.iterations - ih.iterations) <= 2
    assert np.abs(ud - uh).max() < 1e-8 * np.abs(uh).max()
    A = sp.csr_matrix((Gv, Gj, Gi), shape=(B.size, B.size))
    assert np.linalg.norm(A @ ud - B) < 1.5e-10 * np.linalg.norm(B)


@pytest.mark.gpu
def test_gpu_gmres_breakdown_with_steps_queued_ahead(gpu):
    """A = I: the first Arnoldi step breaks down (||w|| = 0) while the next steps are already
    queued; the solve ends after one iteration with u = B and nothing of the unused steps."""
    say("gpu gmres: identity (Arnoldi breakdown at step 1)")
    n = 1000
    A = sp.eye(n, format="csr")
    b = np.random.default_rng(3).normal(size=n)
    u, info = gmres(A.indptr, A.indices, A.data, b, tol=1e-12, max_iter=100, restart=50, device=True)
    assert info.converged and info.iterations == 1
    assert np.all(np.isfinite(u)) and np.abs(u - b).max() < 1e-13 * np.abs(b).max()


@pytest.mark.gpu
def test_gpu_gmres_iteration_cap_and_short_restarts(gpu):
    """max_iter ending inside a restart cycle stops at exactly max_iter iterations with the
    cycle's partial correction applied; restart 2 (shorter than the queue-ahead depth) still
    converges to the direct solution."""
    say("gpu gmres: iteration cap mid-cycle, restart 2")
    rng = np.random.default_rng(5)
    n = 2000
    A = (sp.random(n, n, density=0.003, random_state=7, format="csr") + sp.eye(n) * 2.5).tocsr()
    A.sort_indices()
    b = rng.normal(size=n)
    u, info = gmres(A.indptr, A.indices, A.data, b, tol=1e-14, max_iter=7, restart=50, device=True)
    assert info.iterations == 7 and not info.converged
    r = np.linalg.norm(b - A @ u) / np.linalg.norm(b)
    assert np.all(np.isfinite(u)) and r < 0.5 and abs(r - info.residual) < 1e-6
    u2, info2 = gmres(A.indptr, A.indices, A.data, b, tol=1e-11, max_iter=5000, restart=2, device=True)
    assert info2.converged and info2.restarts > 3
    assert np.allclose(u2, spl.spsolve(A.tocsc(), b), rtol=1e-8, atol=1e-9)


@pytest.mark.gpu
def test_gpu_gmres_switches_to_two_passes_on_cancellation(gpu):
    """Eigenvalues in five tight clusters: after about five steps the Krylov space nearly holds
    A v_k, a one-pass step keeps far less than 1 % of ||w|| and the solve continues with two
    Gram-Schmidt passes; it still converges to the direct solution."""
    say("gpu gmres: clustered spectrum (one-pass -> two-pass switch)")
    rng = np.random.default_rng(9)
    n = 400
    centres = np.array([1.0, 2.0, 3.5, 5.0, 8.0])
    lam = centres[rng.integers(0, 5, n)] * (1.0 + 1e-7 * rng.standard_normal(n))
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = sp.csr_matrix((Q * lam) @ Q.T)
    A.sort_indices()
    b = rng.normal(size=n)
    u, info = gmres(A.indptr, A.indices, A.data, b, tol=1e-10, max_iter=2000, restart=60, device=True)
    assert info.converged
    x = np.linalg.solve(A.toarray(), b)
    assert np.abs(u - x).max() < 1e-7 * np.abs(x).max()
    assert np.linalg.norm(b - A @ u) < 2e-10 * np.linalg.norm(b)

