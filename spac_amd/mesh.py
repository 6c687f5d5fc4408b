"""Python handle over include/spac_mesh.h (plumbing for tests and bench.py).

Mirrors the reference's Malla / SistemaLineal call pattern
(main_test_serial.cpp:44-83): build the mesh, assemble one row per (circle,
interior node) job, accumulate into the CSR system, solve.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


class MeshDesc(C.Structure):
    _fields_ = [
        ("domain", C.c_double * 4),
        ("n_side", C.c_int32),
        ("nodes_per_circle", C.c_int32),
        ("overlap", C.c_double),
        ("theta0", C.c_double),
        ("c2_corner", C.c_double),
        ("c2_side", C.c_double),
        ("rbf_kind", C.c_int32),
        ("stencil_size", C.c_int32),
        ("gauss_eps", C.c_double),
        ("fk_all_circles", C.c_int32),
        ("reserved", C.c_int32),
    ]


class MeshInfo(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32),
        ("n_interfaces", C.c_int32),
        ("n_jobs", C.c_int32),
        ("nnz", C.c_int32),
        ("total_G", C.c_int64),
        ("radius", C.c_double),
        ("max_cond", C.c_double),
    ]


class GmresInfo(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("restarts", C.c_int32),
        ("residual", C.c_double),
        ("converged", C.c_int32),
        ("reserved", C.c_int32),
        ("seconds", C.c_double),
        ("device_ms", C.c_double),
    ]


def _setup(lib):
    if getattr(lib, "_mesh_ready", False):
        return lib
    vp = C.c_void_p
    lib.spac_mesh_build_square.argtypes = [C.POINTER(MeshDesc), C.POINTER(vp)]
    lib.spac_mesh_destroy.argtypes = [vp]
    lib.spac_mesh_get_info.argtypes = [vp, C.POINTER(MeshInfo)]
    lib.spac_mesh_interfaces.argtypes = [vp, C.POINTER(abi.InterfaceTable)]
    lib.spac_mesh_jobs.argtypes = [vp, abi.i32p, abi.dp, abi.dp, abi.i32p, abi.i32p]
    lib.spac_mesh_nodes.argtypes = [vp, abi.dp, abi.dp, abi.i32p]
    lib.spac_mesh_interface_cond.argtypes = [vp, C.c_int32, abi.dp]
    lib.spac_mesh_reset_system.argtypes = [vp]
    lib.spac_mesh_accumulate.argtypes = [vp, C.POINTER(abi.RowBlock)]
    lib.spac_mesh_csr.argtypes = [vp, abi.i32p, abi.i32p, abi.dp, abi.dp]
    lib.spac_mesh_write_coo.argtypes = [vp, C.c_char_p, C.c_char_p]
    lib.spac_mesh_write_csr.argtypes = [vp, C.c_char_p]
    lib.spac_mesh_write_positions.argtypes = [vp, C.c_int32, C.c_char_p]
    lib.spac_gmres_solve.argtypes = [C.c_int32, abi.i32p, abi.i32p, abi.dp, abi.dp, abi.dp, C.c_double, C.c_int32,
                                     C.c_int32, C.POINTER(GmresInfo)]
    lib.spac_gmres_solve_gpu.argtypes = lib.spac_gmres_solve.argtypes
    lib.spac_mesh_system_create.argtypes = [vp, C.POINTER(vp)]
    lib.spac_mesh_system_destroy.argtypes = [vp]
    lib.spac_mesh_system_read.argtypes = [vp, abi.i32p, abi.i32p, abi.dp, abi.dp]
    lib.spac_mesh_system_set_coarse.argtypes = [vp, vp, C.c_int32, C.c_int32]
    lib.spac_gmres_solve_system.argtypes = [vp, abi.dp, C.c_double, C.c_int32, C.c_int32, C.POINTER(GmresInfo)]
    lib.spac_read_coo.argtypes = [C.c_char_p, abi.i32p, abi.i32p, abi.i32p, abi.i32p, abi.dp]
    lib.spac_write_solution.argtypes = [C.c_char_p, C.c_int32, abi.dp]
    lib.spac_write_times_csv.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.c_char_p), abi.i32p, abi.dp, C.c_int32]
    lib._mesh_ready = True
    return lib


def mesh_desc(domain=(0.0, 0.0, 1.0, 1.0), n_side=4, nodes_per_circle=111, overlap=1.2, theta0=None,
              c2_corner=1.5, c2_side=1.1, rbf_kind=abi.SPAC_RBF_GAUSS_STENCIL, stencil_size=15, gauss_eps=3.0,
              fk_all_circles=True) -> MeshDesc:
    d = MeshDesc()
    for i, v in enumerate(domain):
        d.domain[i] = v
    d.n_side = n_side
    d.nodes_per_circle = nodes_per_circle
    d.overlap = overlap
    d.theta0 = np.pi / nodes_per_circle if theta0 is None else theta0
    d.c2_corner = c2_corner
    d.c2_side = c2_side
    d.rbf_kind = rbf_kind
    d.stencil_size = stencil_size
    d.gauss_eps = gauss_eps
    d.fk_all_circles = 1 if fk_all_circles else 0
    return d


class Mesh:
    def __init__(self, desc: MeshDesc | None = None, **kw):
        self.lib = _setup(abi.load())
        self.desc = desc if desc is not None else mesh_desc(**kw)
        h = C.c_void_p()
        abi.check(self.lib.spac_mesh_build_square(C.byref(self.desc), C.byref(h)), self.lib)
        self.h = h
        info = MeshInfo()
        abi.check(self.lib.spac_mesh_get_info(self.h, C.byref(info)), self.lib)
        self.info = info
        self.table = abi.InterfaceTable()
        abi.check(self.lib.spac_mesh_interfaces(self.h, C.byref(self.table)), self.lib)
        n = info.n_jobs
        self.job_row = np.zeros(n, np.int32)
        self.job_x = np.zeros(n)
        self.job_y = np.zeros(n)
        self.job_iface = np.zeros(n, np.int32)
        self.job_frontera = np.zeros(n, np.int32)
        abi.check(self.lib.spac_mesh_jobs(self.h, abi.ptr(self.job_row, C.c_int32), abi.ptr(self.job_x, C.c_double),
                                          abi.ptr(self.job_y, C.c_double), abi.ptr(self.job_iface, C.c_int32),
                                          abi.ptr(self.job_frontera, C.c_int32)), self.lib)
        m = info.n_nodes
        self.node_x = np.zeros(m)
        self.node_y = np.zeros(m)
        self.node_frontera = np.zeros(m, np.int32)
        abi.check(self.lib.spac_mesh_nodes(self.h, abi.ptr(self.node_x, C.c_double), abi.ptr(self.node_y, C.c_double),
                                           abi.ptr(self.node_frontera, C.c_int32)), self.lib)

    def close(self):
        if self.h:
            self.lib.spac_mesh_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # interface table views (valid while the mesh lives)
    def iface_arrays(self):
        t, K = self.table, self.info.n_interfaces
        off = np.ctypeslib.as_array(t.node_offset, (K + 1,)).copy()
        out = dict(
            cx=np.ctypeslib.as_array(t.cx, (K,)).copy(), cy=np.ctypeslib.as_array(t.cy, (K,)).copy(),
            radius=np.ctypeslib.as_array(t.radius, (K,)).copy(), code=np.ctypeslib.as_array(t.code, (K,)).copy(),
            c2=np.ctypeslib.as_array(t.c2, (K,)).copy(), node_offset=off,
            theta=np.ctypeslib.as_array(t.node_theta, (int(off[-1]),)).copy(),
            node_global=np.ctypeslib.as_array(t.node_global, (int(off[-1]),)).copy(),
            ipsi_offset=np.ctypeslib.as_array(t.ipsi_offset, (K,)).copy())
        return out

    def cond(self, k: int) -> float:
        c = C.c_double()
        abi.check(self.lib.spac_mesh_interface_cond(self.h, k, C.byref(c)), self.lib)
        return c.value

    def node_batch(self, paths, h, jobs=None) -> tuple[abi.NodeBatch, dict]:
        """spac_node_batch over all jobs or a subset; RNG stream id = job index.  The batch over
        all jobs at scalar (paths, h) is kept and handed out again for the same pair."""
        key = (int(paths), float(h)) if jobs is None and np.ndim(paths) == 0 and np.ndim(h) == 0 else None
        cached = getattr(self, "_batch_cache", None)
        if key is not None and cached is not None and cached[0] == key:
            return cached[1], cached[2]
        b, keep = self._node_batch(paths, h, jobs)
        if key is not None:
            self._batch_cache = (key, b, keep)
        return b, keep

    def _node_batch(self, paths, h, jobs):
        idx = np.arange(self.info.n_jobs) if jobs is None else np.asarray(jobs)
        n = idx.size
        keep = dict(x=np.ascontiguousarray(self.job_x[idx]), y=np.ascontiguousarray(self.job_y[idx]),
                    iface=np.ascontiguousarray(self.job_iface[idx]),
                    frontera=np.ascontiguousarray(self.job_frontera[idx]),
                    paths=np.ascontiguousarray(np.broadcast_to(paths, (n,)), dtype=np.uint32),
                    h=np.ascontiguousarray(np.broadcast_to(h, (n,)), dtype=np.float64),
                    stream=np.ascontiguousarray(idx, dtype=np.uint32))
        b = abi.NodeBatch()
        b.n = n
        b.x = abi.ptr(keep["x"], C.c_double)
        b.y = abi.ptr(keep["y"], C.c_double)
        b.iface = abi.ptr(keep["iface"], C.c_int32)
        b.frontera = abi.ptr(keep["frontera"], C.c_int32)
        b.paths = abi.ptr(keep["paths"], C.c_uint32)
        b.h = abi.ptr(keep["h"], C.c_double)
        b.stream = abi.ptr(keep["stream"], C.c_uint32)
        return b, keep

    def alloc_rows(self, jobs=None) -> dict:
        idx = np.arange(self.info.n_jobs) if jobs is None else np.asarray(jobs)
        a = self.iface_arrays()
        M = a["node_offset"][1:] - a["node_offset"][:-1]
        total = int(M[self.job_iface[idx]].sum())
        n = idx.size
        return dict(B=np.zeros(n), G=np.zeros(total), G_col=np.zeros(total, np.int32),
                    G_offset=np.zeros(n + 1, np.int64), stats=(abi.PointStats * n)())

    @staticmethod
    def row_block(rows: dict) -> abi.RowBlock:
        r = abi.RowBlock()
        r.B = abi.ptr(rows["B"], C.c_double)
        r.G = abi.ptr(rows["G"], C.c_double)
        r.G_col = abi.ptr(rows["G_col"], C.c_int32)
        r.G_offset = abi.ptr(rows["G_offset"], C.c_int64)
        st = rows.get("stats")
        r.stats = st if isinstance(st, C.Array) else None
        return r

    def accumulate(self, rows: dict, reset=True):
        if reset:
            abi.check(self.lib.spac_mesh_reset_system(self.h), self.lib)
        rb = self.row_block(rows)
        abi.check(self.lib.spac_mesh_accumulate(self.h, C.byref(rb)), self.lib)

    def csr(self):
        n, nnz = self.info.n_nodes, self.info.nnz
        Gi = np.zeros(n + 1, np.int32)
        Gj = np.zeros(nnz, np.int32)
        Gv = np.zeros(nnz)
        B = np.zeros(n)
        abi.check(self.lib.spac_mesh_csr(self.h, abi.ptr(Gi, C.c_int32), abi.ptr(Gj, C.c_int32),
                                         abi.ptr(Gv, C.c_double), abi.ptr(B, C.c_double)), self.lib)
        return Gi, Gj, Gv, B


class DeviceSystem:
    """A mesh's interface system resident on the device (spac_mesh_system_create): rows assembled
    by FKProblem.assemble_system land in it and solve() runs GPU GMRES on it in place."""

    def __init__(self, mesh: Mesh):
        self.lib = mesh.lib
        self.mesh = mesh
        h = C.c_void_p()
        abi.check(self.lib.spac_mesh_system_create(mesh.h, C.byref(h)), self.lib)
        self.h = h

    def close(self):
        if self.h:
            self.lib.spac_mesh_system_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_coarse(self, nx, ny=None):
        """spac_mesh_system_set_coarse: two-level GMRES preconditioner over an nx x ny grid of
        row aggregates (0: off)."""
        ny = nx if ny is None else ny
        abi.check(self.lib.spac_mesh_system_set_coarse(self.mesh.h, self.h, nx, ny), self.lib)

    def read(self):
        """(G_i, G_j, G_ij, B) copied out of the device."""
        n, nnz = self.mesh.info.n_nodes, self.mesh.info.nnz
        Gi = np.zeros(n + 1, np.int32)
        Gj = np.zeros(nnz, np.int32)
        Gv = np.zeros(nnz)
        B = np.zeros(n)
        abi.check(self.lib.spac_mesh_system_read(self.h, abi.ptr(Gi, C.c_int32), abi.ptr(Gj, C.c_int32),
                                                 abi.ptr(Gv, C.c_double), abi.ptr(B, C.c_double)), self.lib)
        return Gi, Gj, Gv, B

    def solve(self, tol=1e-10, max_iter=2000, restart=100, u0=None):
        """spac_gmres_solve_system; u0: initial guess (zeros by default)."""
        n = self.mesh.info.n_nodes
        u = np.zeros(n) if u0 is None else np.array(u0, dtype=np.float64)
        info = GmresInfo()
        abi.check(self.lib.spac_gmres_solve_system(self.h, abi.ptr(u, C.c_double), tol, max_iter, restart,
                                                   C.byref(info)), self.lib)
        return u, info


def gmres(Gi, Gj, Gv, B, tol=1e-10, max_iter=2000, restart=100, u0=None, device=False):
    """spac_gmres_solve (host) or, with device=True, spac_gmres_solve_gpu."""
    lib = _setup(abi.load())
    n = B.size
    u = np.zeros(n) if u0 is None else np.array(u0, dtype=np.float64)
    info = GmresInfo()
    fn = lib.spac_gmres_solve_gpu if device else lib.spac_gmres_solve
    abi.check(fn(n, abi.ptr(np.ascontiguousarray(Gi, np.int32), C.c_int32),
                                   abi.ptr(np.ascontiguousarray(Gj, np.int32), C.c_int32),
                                   abi.ptr(np.ascontiguousarray(Gv, np.float64), C.c_double),
                                   abi.ptr(np.ascontiguousarray(B, np.float64), C.c_double),
                                   abi.ptr(u, C.c_double), tol, max_iter, restart, C.byref(info)), lib)
    return u, info


def write_solution(path, u):
    """spac_write_solution: the reference's solution_GMRES.txt format."""
    lib = _setup(abi.load())
    u = np.ascontiguousarray(u, dtype=np.float64)
    abi.check(lib.spac_write_solution(str(path).encode(), u.size, abi.ptr(u, C.c_double)), lib)


def write_times_csv(path, rows, append=False):
    """spac_write_times_csv: rows of (phase, rank, seconds), the reference's times.csv format."""
    lib = _setup(abi.load())
    n = len(rows)
    names = (C.c_char_p * max(n, 1))(*[r[0].encode() for r in rows])
    ranks = np.array([r[1] for r in rows], dtype=np.int32)
    secs = np.array([r[2] for r in rows], dtype=np.float64)
    abi.check(lib.spac_write_times_csv(str(path).encode(), n, names, abi.ptr(ranks, C.c_int32),
                                       abi.ptr(secs, C.c_double), 1 if append else 0), lib)


def ps_assemble_rows(mesh, coef_kind=abi.SPAC_COEF_FRANKE_POISSON, n_radio=43, jobs=None):
    """spac_ps_assemble_rows over a mesh's jobs (or a subset): rows of nodes inside full interior
    circles from the pseudospectral solver.  Returns (rows, report)."""
    lib = _setup(abi.load())
    nb, keep = mesh.node_batch(0, 1.0, jobs)
    rows = mesh.alloc_rows(jobs)
    rb = mesh.row_block(rows)
    rep = abi.PsReport()
    abi.check(lib.spac_ps_assemble_rows(coef_kind, C.byref(mesh.table), C.byref(nb), n_radio, C.byref(rb),
                                        C.byref(rep)), lib)
    from .fk import stats_to_dicts

    rows["stats"] = stats_to_dicts(rows["stats"])
    return rows, rep
