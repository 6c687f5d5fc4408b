"""Mesh builder, interpolation tables, CSR assembly and GMRES (CPU).

Pins:
- the reference's own numbers: cond(Psi) = |iPsi|_F |Psi|_F of the serial
  driver's mesh is 1.02E+10 on corner arcs and 1.11E+10 on side arcs
  (main_test_serial.cpp:46-47, c2 = 1.5 / 1.1, Nt = 44, theta_0 = 0.1428);
- geometry by brute force: nodes on their circles, arc ends on the domain
  boundary, containing-circle counts (the CSR diagonal) and the CSR pattern;
- interpolation: cardinality H_j(theta_k) = delta_jk and the reference's
  consistency check sum_j u(x_j) H_j(theta) ~ u(X) (Mesh/Interfaz.cpp:252-262);
- GMRES against scipy.sparse.linalg.spsolve on a committed fixture system.
"""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle_lib as o
from spac_amd import abi
from spac_amd.mesh import Mesh, gmres, mesh_desc
from test_oracle import franke_published

GOLDEN = Path(__file__).resolve().parent / "golden"


def serial_mesh():
    return Mesh(mesh_desc(domain=(0.0, 0.0, 1.25, 1.25), n_side=4, nodes_per_circle=44, overlap=1.0, theta0=0.1428,
                          c2_corner=1.5, c2_side=1.1, rbf_kind=abi.SPAC_RBF_IMQ_GLOBAL, fk_all_circles=False))


@pytest.fixture(scope="module")
def mesh2():
    return Mesh(mesh_desc())


def test_reference_condition_numbers():
    m = serial_mesh()
    a = m.iface_arrays()
    corners = [k for k in range(16) if a["code"][k] < -10]
    sides = [k for k in range(16) if -10 < a["code"][k] < 0]
    assert len(corners) == 4 and len(sides) == 8
    for k in corners:
        assert float(f"{m.cond(k):.2e}") == 1.02e10
    for k in sides:
        assert float(f"{m.cond(k):.2e}") == 1.11e10


def test_geometry(mesh2):
    m = mesh2
    a = m.iface_arrays()
    d = 1.0 / 3.0
    assert m.info.radius == pytest.approx(1.2 * d / np.sqrt(2), rel=1e-15)
    for f in range(m.info.n_interfaces):
        sl = slice(a["node_offset"][f], a["node_offset"][f + 1])
        g = a["node_global"][sl]
        assert np.array_equal(g, np.arange(g[0], g[0] + g.size))  # contiguous global numbering
        x, y = m.node_x[g], m.node_y[g]
        assert np.allclose(np.hypot(x - a["cx"][f], y - a["cy"][f]), a["radius"][f], atol=1e-14)
        assert np.all(np.diff(a["theta"][sl]) > 0)
        fr = m.node_frontera[g]
        if a["code"][f] == 1:
            assert g.size == 111 and not fr.any()
        else:
            assert fr[0] == 1 and fr[-1] == 1 and not fr[1:-1].any()
            on_bd = lambda px, py: np.isclose(px, 0) | np.isclose(px, 1) | np.isclose(py, 0) | np.isclose(py, 1)
            assert on_bd(x[[0, -1]], y[[0, -1]]).all()
            assert np.all((x[1:-1] > 0) & (x[1:-1] < 1) & (y[1:-1] > 0) & (y[1:-1] < 1))
    assert m.info.n_nodes == 1016


def test_csr_pattern_and_diagonal_brute_force(mesh2):
    m = mesh2
    a = m.iface_arrays()
    K, n = m.info.n_interfaces, m.info.n_nodes
    owner = np.zeros(n, int)
    for f in range(K):
        owner[a["node_global"][a["node_offset"][f]:a["node_offset"][f + 1]]] = f
    cols = [set([i]) for i in range(n)]
    diag = np.zeros(n)
    for f in range(K):
        inside = (np.hypot(m.node_x - a["cx"][f], m.node_y - a["cy"][f]) < a["radius"][f]) & (owner != f)
        circ = a["node_global"][a["node_offset"][f]:a["node_offset"][f + 1]]
        for i in np.nonzero(inside)[0]:
            diag[i] += 1
            cols[i].update(circ.tolist())
            for j in circ:
                cols[j].add(i)
    Gi, Gj, Gv, B = m.csr()
    assert Gi[-1] == sum(len(c) for c in cols)
    for i in range(n):
        row = Gj[Gi[i]:Gi[i + 1]]
        assert row.tolist() == sorted(cols[i])
        dv = Gv[Gi[i]:Gi[i + 1]][row == i][0]
        assert dv == diag[i]
    assert np.all(Gv[Gj != np.repeat(np.arange(n), np.diff(Gi))] == 1e-23)
    assert np.all(diag >= 1)  # every node lies in some neighbour circle at overlap 1.2
    # jobs: one per (circle, node inside it)
    assert m.info.n_jobs == int(diag.sum())


@pytest.mark.parametrize("kind", ["stencil", "imq"])
def test_interpolation_cardinality_and_consistency(kind, mesh2):
    m = mesh2 if kind == "stencil" else serial_mesh()
    a = m.iface_arrays()
    worst_card, worst_cons = 0.0, 0.0
    for f in range(m.info.n_interfaces):
        if kind == "imq" and a["code"][f] == 1:
            continue
        sl = slice(a["node_offset"][f], a["node_offset"][f + 1])
        th = a["theta"][sl]
        code, cx, cy, R = a["code"][f], a["cx"][f], a["cy"][f], a["radius"][f]
        rel = (lambda t: t - th[0]) if code == 1 else (lambda t: t)
        for k in range(0, th.size, 3):
            e = np.zeros(th.size)
            e[k] = 1.0
            worst_card = max(worst_card, np.abs(o.row_weights(m, f, rel(th[k])) - e).max())
        uj = franke_published(cx + R * np.cos(th), cy + R * np.sin(th))
        for t in np.linspace(th[0], th[-1], 23)[1:-1]:
            w = o.row_weights(m, f, rel(t))
            worst_cons = max(worst_cons, abs(w @ uj - franke_published(cx + R * np.cos(t), cy + R * np.sin(t))))
    if kind == "stencil":
        assert worst_card < 1e-10
        assert worst_cons < 5e-3  # 15-point Gaussian stencils, eps = 3, node spacing 0.0566 rad
    else:
        assert worst_card < 1e-4  # cond(Psi) ~ 1e10
        assert worst_cons < 5e-2


def test_exit_angle_branch_cuts():
    # Atan0 / Atan90 / Atan180 / Atan270 of Mesh/Nudo.hpp:7-32
    assert o.exit_angle(-11, -1.0, -1e-300, 0) == pytest.approx(-np.pi)
    assert o.exit_angle(-33, 0.0, -1.0, 0) == pytest.approx(1.5 * np.pi)
    assert o.exit_angle(-1, -1.0, -1.0, 0) == pytest.approx(1.25 * np.pi)
    assert o.exit_angle(-1, 1.0, -1.0, 0) == pytest.approx(-0.25 * np.pi)
    assert o.exit_angle(-3, 1.0, 0.0, 0) == pytest.approx(2.0 * np.pi)
    assert o.exit_angle(-3, -1.0, 1.0, 0) == pytest.approx(0.75 * np.pi)
    assert o.exit_angle(1, 1.0, 0.0, 0.5) == pytest.approx(2 * np.pi - 0.5)


def test_gmres_against_spsolve_fixture():
    fx = json.loads((GOLDEN / "gmres_system.json").read_text())
    Gi, Gj, Gv, B = (np.array(fx[k], dtype=t) for k, t in (("G_i", np.int32), ("G_j", np.int32), ("G_ij", float),
                                                          ("B", float)))
    u_ref = np.array(fx["u_spsolve"])
    u, info = gmres(Gi, Gj, Gv, B, tol=1e-12, max_iter=5000, restart=200)
    assert info.converged and info.residual < 1e-11
    assert np.abs(u - u_ref).max() < 1e-9 * np.abs(u_ref).max()
    A = sp.csr_matrix((Gv, Gj, Gi), shape=(B.size, B.size))
    assert np.linalg.norm(A @ u - B) < 1e-11 * np.linalg.norm(B)


@pytest.mark.parametrize("max_iter,restart", [(7, 50), (13, 5), (18, 18)])
def test_gmres_iteration_cap_reports_true_residual(max_iter, restart):
    """A solve stopped by max_iter reports converged = 0 and the true relative residual of the
    u it returns (not the last Givens estimate), also when the cap falls inside a cycle."""
    fx = json.loads((GOLDEN / "gmres_system.json").read_text())
    Gi, Gj, Gv, B = (np.array(fx[k], dtype=t) for k, t in (("G_i", np.int32), ("G_j", np.int32), ("G_ij", float),
                                                          ("B", float)))
    u, info = gmres(Gi, Gj, Gv, B, tol=1e-15, max_iter=max_iter, restart=restart)
    A = sp.csr_matrix((Gv, Gj, Gi), shape=(B.size, B.size))
    true = np.linalg.norm(B - A @ u) / np.linalg.norm(B)
    assert info.iterations == max_iter and not info.converged
    assert info.residual == pytest.approx(true, rel=1e-9)
    # a tolerance just above what the cap reaches is met and reported as met
    u2, info2 = gmres(Gi, Gj, Gv, B, tol=2.0 * true, max_iter=max_iter, restart=restart)
    assert info2.converged and info2.residual < 2.0 * true
    assert info2.residual == pytest.approx(np.linalg.norm(B - A @ u2) / np.linalg.norm(B), rel=1e-9)


def test_gmres_random_nonsymmetric():
    rng = np.random.default_rng(0)
    n = 300
    A = sp.random(n, n, density=0.03, random_state=1, format="csr") + sp.eye(n) * 4.0
    A = A.tocsr()
    A.sort_indices()
    b = rng.normal(size=n)
    u, info = gmres(A.indptr, A.indices, A.data, b, tol=1e-12, restart=40)
    assert info.converged
    assert np.allclose(u, spl.spsolve(A.tocsc(), b), rtol=1e-9, atol=1e-10)


def test_coo_write_read_roundtrip(tmp_path):
    m = serial_mesh()
    lib = m.lib
    g, b = tmp_path / "G.txt", tmp_path / "B.txt"
    abi.check(lib.spac_mesh_write_coo(m.h, str(g).encode(), str(b).encode()), lib)
    Gi, Gj, Gv, B = m.csr()
    lines = g.read_text().splitlines()
    assert len(lines) == Gi[-1]
    # Lee_GB_COO reads a header "nnz n" then the "i j v" lines
    gin = tmp_path / "G_in.txt"
    gin.write_text(f"{Gi[-1]} {B.size}\n" + "\n".join(lines) + "\n")
    import ctypes as C

    n, nnz = C.c_int32(), C.c_int32()
    abi.check(lib.spac_read_coo(str(gin).encode(), C.byref(n), C.byref(nnz), None, None, None), lib)
    assert n.value == B.size and nnz.value == Gi[-1]
    Ri = np.zeros(n.value + 1, np.int32)
    Rj = np.zeros(nnz.value, np.int32)
    Rv = np.zeros(nnz.value)
    abi.check(lib.spac_read_coo(str(gin).encode(), C.byref(n), C.byref(nnz), abi.ptr(Ri, C.c_int32),
                                abi.ptr(Rj, C.c_int32), abi.ptr(Rv, C.c_double)), lib)
    assert np.array_equal(Ri, Gi) and np.array_equal(Rj, Gj)
    assert np.allclose(Rv, Gv, atol=5e-9)


def test_imq_full_circles_rejected():
    with pytest.raises(abi.SpacError) as e:
        Mesh(mesh_desc(rbf_kind=abi.SPAC_RBF_IMQ_GLOBAL, fk_all_circles=True))
    assert e.value.code == abi.SPAC_E_BAD_INTERFACE
