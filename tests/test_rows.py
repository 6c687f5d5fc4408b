"""Interface rows (G, B): CPU properties and GPU parity.

CPU pins of the restatement's rows:
- boundary nodes: B = g(x) exactly (Solvers/FeynmanKacSolver.cpp:346-347);
- with Y = 1 (c = 0), sum_j G_j = -(fraction of circle exits) up to the
  interpolation's reproduction error of constants;
- global IMQ: the moment form equals the reference's direct form
  (Mesh/Interfaz.cpp:253-261) within cond(Psi) * eps.
GPU: B, G and the node statistics are bit-identical to the restatement; the
full PDDSparse workload (config 2) solves to the manufactured solution.
"""
import sys

import numpy as np
import pytest

import oracle_lib as o
from spac_amd import abi
from spac_amd.fk import FKProblem, problem_desc
from spac_amd.mesh import Mesh, gmres, mesh_desc
from test_oracle import franke_published

FR = abi.SPAC_COEF_FRANKE_POISSON
VC = abi.SPAC_COEF_VARCOEF_FRANKE


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def serial_mesh():
    return Mesh(mesh_desc(domain=(0.0, 0.0, 1.25, 1.25), n_side=4, nodes_per_circle=44, overlap=1.0, theta0=0.1428,
                          c2_corner=1.5, c2_side=1.1, rbf_kind=abi.SPAC_RBF_IMQ_GLOBAL, fk_all_circles=False))


def _offsets(mesh, jobs):
    a = mesh.iface_arrays()
    M = a["node_offset"][1:] - a["node_offset"][:-1]
    return np.concatenate([[0], np.cumsum(M[mesh.job_iface[jobs]])])


def test_boundary_rows_and_row_sums():
    m = Mesh(mesh_desc())
    jobs = np.arange(0, m.info.n_jobs, 23)
    assert m.job_frontera[jobs].any()
    desc = problem_desc(coef=FR, domain=(0.0, 0.0, 1.0, 1.0))
    B, G, st = o.assemble_rows(desc, m, 600, 5e-4, seed=3, jobs=jobs)
    off = _offsets(m, jobs)
    for k, j in enumerate(jobs):
        Gk = G[off[k]:off[k + 1]]
        if m.job_frontera[j]:
            assert B[k] == o.coef_eval(FR, 0, m.job_x[j:j + 1], m.job_y[j:j + 1])[0]
            assert not Gk.any() and st[k]["paths"] == 0
        else:
            frac = 1.0 - st[k]["n_domain_exits"] / st[k]["paths"]
            assert abs(Gk.sum() + frac) < 2e-3, (j, Gk.sum(), frac)


def test_imq_moment_form_equals_direct_form():
    m = serial_mesh()
    jobs = np.arange(0, m.info.n_jobs, 5)
    desc = problem_desc(coef=FR, domain=(0.0, 0.0, 1.25, 1.25))
    B1, G1, _ = o.assemble_rows(desc, m, 400, 1e-3, seed=1, jobs=jobs)
    B2, G2, _ = o.assemble_rows(desc, m, 400, 1e-3, seed=1, jobs=jobs, direct=True)
    assert np.array_equal(B1, B2)
    # |iPsi| ~ cond(Psi) ~ 1e10: reassociating the sum moves G by <= cond * eps relative
    assert np.abs(G1 - G2).max() <= 1e-6 * max(1.0, np.abs(G2).max())


def test_stencil_moment_form_equals_direct_form():
    """Gaussian stencils: the fixed-point moment form (one stencil inverse per row) against the
    per-path direct form G_j += Y H_j(theta) (Mesh/Interfaz.cpp:253-261, sixteen TwoSum partials),
    on full circles and arcs, FP64 paths of the variable-coefficient operator.  Both compute
    the same linear map; they differ by the reassociation of sum_i iPsi(i, r) phi_i
    (|iPsi| up to ~4e3 here) and the 2^-63 N max|Y| rounding of each fixed-point term."""
    m = Mesh(mesh_desc())
    a = m.iface_arrays()
    jobs = np.arange(0, m.info.n_jobs, 7)
    assert (a["code"][m.job_iface[jobs]] == 1).any() and (a["code"][m.job_iface[jobs]] != 1).any()
    desc = problem_desc(coef=VC, domain=(0.0, 0.0, 1.0, 1.0))
    B1, G1, _ = o.assemble_rows(desc, m, 3000, 5e-4, seed=4, jobs=jobs)
    B2, G2, _ = o.assemble_rows(desc, m, 3000, 5e-4, seed=4, jobs=jobs, direct=True)
    assert np.array_equal(B1, B2)
    M = a["node_offset"][1:] - a["node_offset"][:-1]
    S = np.minimum(15, M)
    n_ipsi = int((a["ipsi_offset"] + np.where(a["code"] == 1, 1, M - S + 1) * S * S).max())
    scale = np.abs(np.ctypeslib.as_array(m.table.ipsi, (n_ipsi,))).max()
    err = np.abs(G1 - G2).max()
    assert err <= 64 * scale * 2.2e-16, (err, scale)
    assert np.abs(G2).max() > 1e-3  # rows are not trivially zero


def test_fixed_point_moment_helpers():
    """spac_moment_scale keeps N max|Y| 2^q below 2^62 and max|Y| 2^q below 2^50, and is the
    largest such exponent; to_fixed rounds to the nearest multiple of 2^-q with ties to even
    (numpy's rint), from_fixed maps back exactly."""
    for n, ym in ((1, 1.0), (1000, 0.37), (100000, 1.0), (7, 3.1e5), (2 ** 20, 2.0 ** -30), (2 ** 13, 1.0)):
        q = o.moment_scale(float(n), ym)
        assert n * ym * 2.0 ** q < 2.0 ** 62 and ym * 2.0 ** q < 2.0 ** 50
        assert n * ym * 2.0 ** (q + 1) >= 2.0 ** 62 or ym * 2.0 ** (q + 1) >= 2.0 ** 50
        for v in (ym * 0.734, -ym * 0.9999, ym * 1e-9, 0.0):
            t = o.to_fixed(v, q)
            assert t == int(np.rint(v * 2.0 ** q))
            assert o.from_fixed(t, q) == t * 2.0 ** (-q)
        for half in (2.5, 3.5, -2.5):  # ties go to even
            assert o.to_fixed(half * 2.0 ** (-q), q) == int(np.rint(half))
    assert o.moment_scale(100.0, 0.0) == 0


def _rows_equal(gpu, B, G, st, jobs, mesh):
    assert np.array_equal(gpu["B"].view(np.uint64), B.view(np.uint64))
    bad = np.nonzero(gpu["G"].view(np.uint64) != G.view(np.uint64))[0]
    assert bad.size == 0, (bad[:5], gpu["G"][bad[:5]], G[bad[:5]])
    off = _offsets(mesh, jobs)
    assert np.array_equal(gpu["G_offset"], off)
    a = mesh.iface_arrays()
    for k, j in enumerate(jobs):
        f = mesh.job_iface[j]
        assert np.array_equal(gpu["G_col"][off[k]:off[k + 1]], a["node_global"][a["node_offset"][f]:a["node_offset"][f + 1]])
    for s1, s2 in zip(gpu["stats"], st):
        for key in s1:
            assert s1[key] == s2[key], key


@pytest.mark.gpu
def test_gpu_rows_bit_identical_stencil(gpu):
    say("rows parity: Gaussian stencils, variable coefficients")
    m = Mesh(mesh_desc())
    jobs = np.concatenate([np.arange(0, m.info.n_jobs, 29), np.nonzero(m.job_frontera)[0][:3]])
    desc = problem_desc(coef=VC, domain=(0.0, 0.0, 1.0, 1.0))
    B, G, st = o.assemble_rows(desc, m, 700, 5e-4, seed=11, jobs=jobs)
    with FKProblem(desc) as p:
        rows = p.assemble_rows(m, 700, 5e-4, seed=11, jobs=jobs)
        p.set_launch(128, 3, 37)
        rows2 = p.assemble_rows(m, 700, 5e-4, seed=11, jobs=jobs)
    _rows_equal(rows, B, G, st, jobs, m)
    _rows_equal(rows2, B, G, st, jobs, m)


@pytest.mark.gpu
def test_gpu_rows_bit_identical_stencil_short_rows(gpu):
    """Batches whose rows have at most 1024 paths take the one-wave G kernel, longer ones the
    four-wave kernel.  Both equal the oracle bit for bit, with boundary rows (no paths) among
    the jobs, on a mesh of 111-node and one of 40-node circles."""
    say("rows parity: Gaussian stencils, rows of 100 / 1024 / 1025 paths")
    desc = problem_desc(coef=VC, domain=(0.0, 0.0, 1.0, 1.0))
    for m, h, step in ((Mesh(mesh_desc()), 5e-4, 31), (Mesh(mesh_desc(n_side=16, nodes_per_circle=40)), 2e-4, 97)):
        jobs = np.concatenate([np.arange(0, m.info.n_jobs, step), np.nonzero(m.job_frontera)[0][:2]])
        if jobs.size % 2 == 0:
            jobs = jobs[:-1]
        for paths in (100, 1024, 1025):
            B, G, st = o.assemble_rows(desc, m, paths, h, seed=13, jobs=jobs)
            with FKProblem(desc) as p:
                rows = p.assemble_rows(m, paths, h, seed=13, jobs=jobs)
            _rows_equal(rows, B, G, st, jobs, m)
        m.close()


@pytest.mark.gpu
def test_gpu_rows_bit_identical_imq(gpu):
    say("rows parity: global IMQ (reference mode), Franke")
    m = serial_mesh()
    jobs = np.arange(0, m.info.n_jobs, 3)
    desc = problem_desc(coef=FR, domain=(0.0, 0.0, 1.25, 1.25))
    B, G, st = o.assemble_rows(desc, m, 500, 1e-3, seed=5, jobs=jobs)
    with FKProblem(desc) as p:
        rows = p.assemble_rows(m, 500, 1e-3, seed=5, jobs=jobs)
    _rows_equal(rows, B, G, st, jobs, m)


@pytest.mark.gpu
def test_gpu_rows_reject_bad_input(gpu):
    say("rows: argument checks")
    m = serial_mesh()
    desc = problem_desc(coef=FR, domain=(0.0, 0.0, 1.25, 1.25))
    with FKProblem(desc) as p:
        rows = p.assemble_rows(m, 0, 1e-3, jobs=np.nonzero(m.job_frontera == 0)[0][:2], allow_status=True)
        assert rows["rc"] == abi.SPAC_E_INVALID_ARG


@pytest.mark.gpu
def test_config2_pddsparse_solve(gpu):
    """Unit square, 4x4 circles, 1016 nodes (1952 rows), 20k paths/row, h = 5e-4,
    variable-coefficient operator, Gaussian 15-point stencils; GMRES to 1e-10."""
    say("config 2: assemble + GMRES")
    m = Mesh(mesh_desc())
    desc = problem_desc(coef=VC, domain=(0.0, 0.0, 1.0, 1.0))
    with FKProblem(desc) as p:
        rows = p.assemble_rows(m, 20000, 5e-4, seed=0)
        rows_b = p.assemble_rows(m, 20000, 5e-4, seed=0)
    assert np.array_equal(rows["G"], rows_b["G"]) and np.array_equal(rows["B"], rows_b["B"])
    m.accumulate(rows)
    Gi, Gj, Gv, B = m.csr()
    u, info = gmres(Gi, Gj, Gv, B, tol=1e-10, max_iter=5000, restart=200)
    assert info.converged and info.residual < 1e-10
    exact = franke_published(m.node_x, m.node_y)
    err = u - exact
    rms, mx = np.sqrt(np.mean(err ** 2)), np.abs(err).max()
    say(f"config 2: GMRES iters {info.iterations}, rms err {rms:.3e}, max err {mx:.3e}")
    assert rms < 1e-2 and mx < 5e-2
    steps = sum(s["path_steps"] for s in rows["stats"])
    say(f"config 2: path-steps {steps}")
