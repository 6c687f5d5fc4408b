"""Python handle over the C ABI (tests, bench and smoke use it; no compute here).

Mirrors the reference's FeynmanKacSolver call pattern
(Solvers/FeynmanKacSolver.hpp:95-98): a problem (BVP + frontiers) is created
once, then point solves or batches of solves run on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


def problem_desc(coef=abi.SPAC_COEF_FRANKE_POISSON, domain=(-1.0, -1.0, 1.0, 1.0),
                 subdomain_kind=abi.SPAC_BOUNDARY_NONE, subdomain=(0.0, 0.0, 0.0),
                 gobet=abi.GOBET_CONSTANT, max_steps=0, precision=abi.SPAC_FP64, reflecting=0,
                 rho=0.0) -> abi.ProblemDesc:
    """reflecting: SPAC_REFLECT_* mask of rectangle sides that reflect (0: all
    absorbing, the reference's shipped setting); rho: the Lepingle band width."""
    d = abi.ProblemDesc()
    d.coef_kind = coef
    d.domain_kind = abi.SPAC_BOUNDARY_RECTANGLE
    for i, v in enumerate(domain):
        d.domain_params[i] = v
    d.subdomain_kind = subdomain_kind
    for i, v in enumerate(subdomain):
        d.subdomain_params[i] = v
    d.gobet_constant = gobet
    d.max_steps = max_steps
    d.precision = precision
    d.reflecting = reflecting
    d.reflect_rho = rho
    return d


RECORD_FIELDS = (("x_tau", np.float64), ("y_tau", np.float64), ("Y", np.float64), ("Z", np.float64),
                 ("xi", np.float64), ("tau", np.float64), ("gval", np.float64), ("steps", np.int32),
                 ("outcome", np.int8))
STATS_FIELDS = ("phi", "phiphi", "phi_VR", "phiphi_VR", "xi", "xixi", "xiphi", "APL", "RNGC", "B",
                "paths", "path_steps", "n_max_steps", "status", "n_domain_exits", "cv_fallback")


def alloc_records(total: int) -> dict:
    return {name: np.zeros(total, dtype=dt) for name, dt in RECORD_FIELDS}


def records_struct(rec: dict) -> abi.PathRecords:
    r = abi.PathRecords()
    ctmap = {np.float64: C.c_double, np.int32: C.c_int32, np.int8: C.c_int8}
    for name, dt in RECORD_FIELDS:
        setattr(r, name, abi.ptr(rec[name], ctmap[dt]))
    return r


def stats_to_dicts(arr) -> list[dict]:
    return [{f: getattr(s, f) for f in STATS_FIELDS} for s in arr]


class SolveBatchArrays:
    """Owns the numpy arrays behind a spac_solve_batch."""

    def __init__(self, x0, y0, paths, h, stream, circle=None):
        self.x0 = np.ascontiguousarray(x0, dtype=np.float64)
        self.y0 = np.ascontiguousarray(y0, dtype=np.float64)
        n = self.x0.size
        self.paths = np.ascontiguousarray(np.broadcast_to(paths, (n,)), dtype=np.uint32)
        self.h = np.ascontiguousarray(np.broadcast_to(h, (n,)), dtype=np.float64)
        self.stream = np.ascontiguousarray(np.broadcast_to(stream, (n,)) if np.ndim(stream) == 0 else stream,
                                           dtype=np.uint32)
        self.circle = None if circle is None else np.ascontiguousarray(circle, dtype=np.float64).reshape(-1)
        self.n = n

    @property
    def total_paths(self) -> int:
        return int(self.paths.astype(np.int64).sum())

    def struct(self) -> abi.SolveBatch:
        b = abi.SolveBatch()
        b.n = self.n
        b.x0 = abi.ptr(self.x0, C.c_double)
        b.y0 = abi.ptr(self.y0, C.c_double)
        b.paths = abi.ptr(self.paths, C.c_uint32)
        b.h = abi.ptr(self.h, C.c_double)
        b.stream = abi.ptr(self.stream, C.c_uint32)
        b.circle = abi.ptr(self.circle, C.c_double)
        return b


class FKProblem:
    """Owns one spac_problem handle (one HIP stream)."""

    def __init__(self, desc: abi.ProblemDesc | None = None, **kw):
        self.lib = abi.load()
        self.desc = desc if desc is not None else problem_desc(**kw)
        h = C.c_void_p()
        abi.check(self.lib.spac_fk_problem_create(C.byref(self.desc), C.byref(h)), self.lib)
        self.h = h

    def close(self):
        if self.h:
            self.lib.spac_fk_problem_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_launch(self, block_size=0, blocks_per_cu=0, chunk_paths=0, item_order=abi.SPAC_ORDER_AUTO):
        cfg = abi.LaunchConfig(block_size, blocks_per_cu, chunk_paths, item_order)
        abi.check(self.lib.spac_fk_set_launch(self.h, C.byref(cfg)), self.lib)

    def set_control_variate(self, mode, table=None, domain=None, fallback=False, max_ratio=1.0):
        """spac_fk_set_control_variate: mode SPAC_CV_*; table (T, T, 4) float32 for SPAC_CV_TABLE,
        indexed [j, i] = (u_hat, du/dx, du/dy, 0) at (xmin + i dx, ymin + j dy)."""
        cfg = abi.CvConfig()
        cfg.mode = mode
        cfg.fallback = 1 if fallback else 0
        cfg.max_ratio = max_ratio
        keep = None
        if table is not None:
            keep = np.ascontiguousarray(table, dtype=np.float32)
            assert keep.ndim == 3 and keep.shape[0] == keep.shape[1] and keep.shape[2] == 4
            cfg.table_size = keep.shape[0]
            for i, v in enumerate(domain):
                cfg.table_domain[i] = v
            cfg.table = abi.ptr(keep, C.c_float)
        abi.check(self.lib.spac_fk_set_control_variate(self.h, C.byref(cfg)), self.lib)

    def set_coef_tables(self, mask, values=None, domain=None):
        """spac_fk_set_coef_tables: values (T, T, 8) float64 indexed [j, i] = (s, b_x, b_y, c, f, g, u, 0)
        at (xmin + i dx, ymin + j dy); the coefficients in mask (SPAC_TAB_*) come from the table.
        mask 0 removes it."""
        cfg = abi.CoefTables()
        cfg.mask = mask
        if mask:
            keep = np.ascontiguousarray(values, dtype=np.float64)
            assert keep.ndim == 3 and keep.shape[0] == keep.shape[1] and keep.shape[2] == 8
            cfg.table_size = keep.shape[0]
            for i, v in enumerate(domain):
                cfg.table_domain[i] = v
            cfg.values = abi.ptr(keep, C.c_double)
        abi.check(self.lib.spac_fk_set_coef_tables(self.h, C.byref(cfg)), self.lib)

    def timing(self) -> abi.Timing:
        t = abi.Timing()
        abi.check(self.lib.spac_fk_last_timing(self.h, C.byref(t)), self.lib)
        return t

    def solve_points(self, x0, y0, paths, h, seed=0, allow_status=False, as_array=False):
        """spac_fk_solve_points; the estimators as a list of dicts, or (as_array) as a numpy
        structured array with the fields of spac_point_stats."""
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        n = x0.size
        stats = (abi.PointStats * n)()
        rc = self.lib.spac_fk_solve_points(self.h, abi.ptr(x0, C.c_double), abi.ptr(y0, C.c_double), n, paths, h,
                                           seed, stats)
        if not allow_status:
            abi.check(rc, self.lib)
        return (np.ctypeslib.as_array(stats) if as_array else stats_to_dicts(stats)), rc

    def simulate(self, batch: SolveBatchArrays, seed=0, want_records=True, want_stats=True):
        rec = alloc_records(batch.total_paths) if want_records else None
        stats = (abi.PointStats * batch.n)() if want_stats else None
        b = batch.struct()
        rc = self.lib.spac_fk_simulate(self.h, C.byref(b), seed, C.byref(records_struct(rec)) if rec else None,
                                       stats)
        abi.check(rc, self.lib)
        return rec, (stats_to_dicts(stats) if stats is not None else None)

    def assemble_rows(self, mesh, paths, h, seed=0, jobs=None, allow_status=False):
        """spac_fk_assemble_rows over a mesh's jobs (or a subset): B, G, G_col, G_offset, stats."""
        nb, keep = mesh.node_batch(paths, h, jobs)
        rows = mesh.alloc_rows(jobs)
        rb = mesh.row_block(rows)
        rc = self.lib.spac_fk_assemble_rows(self.h, C.byref(mesh.table), C.byref(nb), seed, C.byref(rb))
        if not allow_status:
            abi.check(rc, self.lib)
        rows["stats"] = stats_to_dicts(rows["stats"])
        rows["rc"] = rc
        return rows

    def assemble_rows_resumable(self, mesh, paths, h, directory, batch_nodes, seed=0, jobs=None):
        """spac_fk_assemble_rows_resumable: rows in node batches of batch_nodes, each read from
        directory/rows_<b>.bin when it holds these inputs' rows, else assembled and written.
        Returns (rows, report)."""
        nb, keep = mesh.node_batch(paths, h, jobs)
        rows = mesh.alloc_rows(jobs)
        rb = mesh.row_block(rows)
        rep = abi.ResumeReport()
        abi.check(self.lib.spac_fk_assemble_rows_resumable(self.h, C.byref(mesh.table), C.byref(nb), seed, batch_nodes,
                                                           str(directory).encode(), C.byref(rb), C.byref(rep)),
                  self.lib)
        rows["stats"] = stats_to_dicts(rows["stats"])
        return rows, rep

    def assemble_system(self, mesh, system, paths, h, seed=0, stats=True):
        """spac_fk_assemble_system: every job of the mesh, accumulated into the device-resident
        system (a mesh.DeviceSystem of the same mesh).  Returns the per-job estimator records as a
        numpy structured array (fields of spac_point_stats), or None with stats=False (no copy)."""
        nb, keep = mesh.node_batch(paths, h)
        st = (abi.PointStats * mesh.info.n_jobs)() if stats else None
        abi.check(self.lib.spac_fk_assemble_system(self.h, C.byref(mesh.table), C.byref(nb), seed, system.h, st),
                  self.lib)
        return np.ctypeslib.as_array(st) if stats else None

    def assemble_system_uniform(self, system, paths, h, seed=0, stats=True):
        """spac_fk_assemble_system_uniform: assemble_system over every job of the system's mesh at
        scalar (paths, h), with the solve table built on the device (no per-call upload).  Same
        results as assemble_system(system.mesh, system, paths, h, seed) bit for bit."""
        n = system.mesh.info.n_jobs
        st = (abi.PointStats * n)() if stats else None
        abi.check(self.lib.spac_fk_assemble_system_uniform(self.h, system.h, int(paths), float(h), seed, st),
                  self.lib)
        return np.ctypeslib.as_array(st) if stats else None

    def autotune_rows(self, mesh, paths, h, path_fraction=0.125, jobs=None):
        """spac_fk_autotune_rows: installs the fastest launch configuration; returns the report and
        the measured candidates [(block_size, blocks_per_cu, chunk_paths, item_order, ms), ...]."""
        nb, keep = mesh.node_batch(paths, h, jobs)
        cap = 64
        tried = (abi.LaunchConfig * cap)()
        ms = np.zeros(cap)
        rep = abi.AutotuneReport()
        abi.check(self.lib.spac_fk_autotune_rows(self.h, C.byref(mesh.table), C.byref(nb), path_fraction, tried,
                                                 abi.ptr(ms, C.c_double), cap, C.byref(rep)), self.lib)
        n = min(rep.n_candidates, cap)
        table = [(t.block_size, t.blocks_per_cu, t.chunk_paths, t.item_order, float(ms[i]))
                 for i, t in enumerate(tried[:n])]
        return rep, table


class RowsPlan:
    """spac_rows_plan: one node batch uploaded once, its assembly captured in a HIP graph."""

    def __init__(self, problem: FKProblem, mesh, paths, h, jobs=None):
        self.lib = problem.lib
        self.mesh = mesh
        self.jobs = jobs
        nb, self._keep = mesh.node_batch(paths, h, jobs)
        hp = C.c_void_p()
        abi.check(self.lib.spac_fk_plan_create(problem.h, C.byref(mesh.table), C.byref(nb), C.byref(hp)), self.lib)
        self.h = hp
        self.last_ms = None

    def execute(self, seed=0, allow_status=False):
        rows = self.mesh.alloc_rows(self.jobs)
        rb = self.mesh.row_block(rows)
        ms = C.c_double()
        rc = self.lib.spac_fk_plan_execute(self.h, seed, C.byref(rb), C.byref(ms))
        if not allow_status:
            abi.check(rc, self.lib)
        self.last_ms = ms.value
        rows["stats"] = stats_to_dicts(rows["stats"])
        rows["rc"] = rc
        return rows

    def close(self):
        if self.h:
            self.lib.spac_fk_plan_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
