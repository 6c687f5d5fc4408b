"""ctypes loader for the CPU restatement (oracle/_build/libfk_oracle.so).

Test infrastructure: only tests, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, always as the checker.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from spac_amd import abi
from spac_amd.fk import SolveBatchArrays, alloc_records, records_struct, stats_to_dicts

ROOT = Path(__file__).resolve().parent.parent
ORACLE_PATH = ROOT / "oracle" / "_build" / "libfk_oracle.so"

_LIB = None


def load() -> C.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not ORACLE_PATH.exists():
        raise RuntimeError(f"{ORACLE_PATH} missing: run __graft_entry__.build()")
    lib = C.CDLL(str(ORACLE_PATH))
    dp = C.POINTER(C.c_double)
    lib.oracle_simulate.argtypes = [C.POINTER(abi.ProblemDesc), C.POINTER(abi.SolveBatch), C.c_uint64,
                                    C.POINTER(abi.PathRecords), C.c_int]
    lib.oracle_simulate.restype = C.c_int
    lib.oracle_reduce.argtypes = [C.c_int, C.POINTER(abi.SolveBatch), C.POINTER(abi.PathRecords),
                                  C.POINTER(abi.PointStats)]
    lib.oracle_reduce.restype = C.c_int
    lib.oracle_math_eval.argtypes = [C.c_int, dp, dp, dp, C.c_int64]
    lib.oracle_math_eval.restype = None
    lib.oracle_philox.argtypes = [C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.oracle_philox.restype = None
    lib.oracle_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int32, dp]
    lib.oracle_normals.restype = None
    lib.oracle_exponentials.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int32, dp]
    lib.oracle_exponentials.restype = None
    lib.oracle_coef_eval.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.c_int64]
    lib.oracle_coef_eval.restype = C.c_int
    lib.oracle_boundary_eval.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int64]
    lib.oracle_boundary_eval.restype = None
    lib.oracle_path_terms.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_int32, C.c_int8, dp]
    lib.oracle_path_terms.restype = None
    _LIB = lib
    return lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def math_eval(fn: int, x, y=None) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, dtype=np.float64)
    out = np.empty_like(x)
    load().oracle_math_eval(fn, _p(x), _p(y), _p(out), x.size)
    return out


def philox(rounds: int, ctr, key) -> list[int]:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    load().oracle_philox(rounds, c, k, o)
    return list(o)


def normals(seed: int, stream: int, path: int, nsteps: int) -> np.ndarray:
    out = np.empty(2 * nsteps, dtype=np.float64)
    load().oracle_normals(seed, stream, path, nsteps, _p(out))
    return out


def exponentials(seed: int, stream: int, path: int, nsteps: int) -> np.ndarray:
    out = np.empty(nsteps, dtype=np.float64)
    load().oracle_exponentials(seed, stream, path, nsteps, _p(out))
    return out


def coef_eval(coef_kind: int, which: int, x, y) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty_like(x)
    rc = load().oracle_coef_eval(coef_kind, which, _p(x), _p(y), _p(out), x.size)
    assert rc == 0
    return out


def boundary_eval(kind: int, params, x, y) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    pr = np.zeros(4, dtype=np.float64)
    pr[: len(params)] = params
    out = np.empty((x.size, 5), dtype=np.float64)
    load().oracle_boundary_eval(kind, _p(pr), _p(x), _p(y), _p(out), x.size)
    return out


def simulate(desc: abi.ProblemDesc, batch: SolveBatchArrays, seed: int = 0, nthreads: int = 0) -> dict:
    rec = alloc_records(batch.total_paths)
    b = batch.struct()
    rs = records_struct(rec)
    rc = load().oracle_simulate(C.byref(desc), C.byref(b), seed, C.byref(rs), nthreads)
    assert rc == 0, rc
    return rec


def reduce(batch: SolveBatchArrays, rec: dict, node_mode: bool = False) -> list[dict]:
    stats = (abi.PointStats * batch.n)()
    b = batch.struct()
    rs = records_struct(rec)
    rc = load().oracle_reduce(1 if node_mode else 0, C.byref(b), C.byref(rs), stats)
    assert rc == 0
    return stats_to_dicts(stats)


def path_terms(node_mode, Y, Z, xi, tau, gval, steps, outcome) -> np.ndarray:
    out = np.empty(10, dtype=np.float64)
    load().oracle_path_terms(1 if node_mode else 0, Y, Z, xi, tau, gval, steps, outcome, _p(out))
    return out


def assemble_rows(desc: abi.ProblemDesc, mesh, paths, h, seed=0, jobs=None, direct=False, nthreads=0, streams=None):
    """CPU restatement of spac_fk_assemble_rows over a mesh's jobs (or a subset)."""
    lib = load()
    lib.oracle_assemble_rows.argtypes = [C.POINTER(abi.ProblemDesc), C.POINTER(abi.InterfaceTable),
                                         C.POINTER(abi.NodeBatch), C.c_uint64, C.c_int, C.c_int,
                                         C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(abi.PointStats)]
    lib.oracle_assemble_rows.restype = C.c_int
    nb, keep = node_batch(mesh, paths, h, jobs, streams)
    n = nb.n
    a = mesh.iface_arrays()
    M = a["node_offset"][1:] - a["node_offset"][:-1]
    total = int(M[keep["iface"]].sum())
    B = np.zeros(n)
    G = np.zeros(total)
    st = (abi.PointStats * n)()
    rc = lib.oracle_assemble_rows(C.byref(desc), C.byref(mesh.table), C.byref(nb), seed, nthreads, 1 if direct else 0,
                                  _p(B), _p(G), st)
    assert rc == 0, rc
    return B, G, stats_to_dicts(st)


def node_batch(mesh, paths, h, jobs=None, streams=None):
    """A spac_node_batch over all of a mesh's jobs or the selected ones (stream = job index)."""
    idx = np.arange(mesh.info.n_jobs) if jobs is None else np.asarray(jobs)
    n = idx.size
    keep = dict(x=np.ascontiguousarray(mesh.job_x[idx]), y=np.ascontiguousarray(mesh.job_y[idx]),
                iface=np.ascontiguousarray(mesh.job_iface[idx]), frontera=np.ascontiguousarray(mesh.job_frontera[idx]),
                paths=np.ascontiguousarray(np.broadcast_to(paths, (n,)), dtype=np.uint32),
                h=np.ascontiguousarray(np.broadcast_to(h, (n,)), dtype=np.float64),
                stream=np.ascontiguousarray(idx if streams is None else streams, dtype=np.uint32))
    nb = abi.NodeBatch()
    nb.n = n
    nb.x = abi.ptr(keep["x"], C.c_double)
    nb.y = abi.ptr(keep["y"], C.c_double)
    nb.iface = abi.ptr(keep["iface"], C.c_int32)
    nb.frontera = abi.ptr(keep["frontera"], C.c_int32)
    nb.paths = abi.ptr(keep["paths"], C.c_uint32)
    nb.h = abi.ptr(keep["h"], C.c_double)
    nb.stream = abi.ptr(keep["stream"], C.c_uint32)
    nb._keep = keep
    return nb, keep


def row_weights(mesh, f: int, theta: float) -> np.ndarray:
    lib = load()
    lib.oracle_row_weights.argtypes = [C.POINTER(abi.InterfaceTable), C.c_int, C.c_double, C.POINTER(C.c_double)]
    lib.oracle_row_weights.restype = C.c_int
    a = mesh.iface_arrays()
    M = int(a["node_offset"][f + 1] - a["node_offset"][f])
    out = np.zeros(M)
    rc = lib.oracle_row_weights(C.byref(mesh.table), f, theta, _p(out))
    assert rc == 0, rc
    return out


def moment_scale(n_paths: float, ymax: float) -> int:
    lib = load()
    lib.oracle_moment_scale.argtypes = [C.c_double, C.c_double]
    lib.oracle_moment_scale.restype = C.c_int
    return lib.oracle_moment_scale(n_paths, ymax)


def to_fixed(v: float, q: int) -> int:
    lib = load()
    lib.oracle_to_fixed.argtypes = [C.c_double, C.c_int]
    lib.oracle_to_fixed.restype = C.c_int64
    return lib.oracle_to_fixed(v, q)


def from_fixed(n: int, q: int) -> float:
    lib = load()
    lib.oracle_from_fixed.argtypes = [C.c_int64, C.c_int]
    lib.oracle_from_fixed.restype = C.c_double
    return lib.oracle_from_fixed(n, q)


def exit_angle(code: int, dx: float, dy: float, first: float) -> float:
    lib = load()
    lib.oracle_exit_angle.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]
    lib.oracle_exit_angle.restype = C.c_int
    t = C.c_double()
    assert lib.oracle_exit_angle(code, dx, dy, first, C.byref(t)) == 1
    return t.value


_CV_TABLE = None


def set_cv(mode=abi.SPAC_CV_ANALYTIC, fallback=False, max_ratio=1.0, table=None, domain=None):
    """oracle_set_cv: the control variate of later simulate/reduce/assemble calls
    (process-wide; reset with set_cv() when done)."""
    global _CV_TABLE
    lib = load()
    lib.oracle_set_cv.argtypes = [C.c_int, C.c_int, C.c_double, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_double)]
    lib.oracle_set_cv.restype = C.c_int
    t = 0
    tp = C.cast(None, C.POINTER(C.c_float))
    dom = np.zeros(4)
    if table is not None:
        _CV_TABLE = np.ascontiguousarray(table, dtype=np.float32)
        t = _CV_TABLE.shape[0]
        tp = _CV_TABLE.ctypes.data_as(C.POINTER(C.c_float))
        dom = np.asarray(domain, dtype=np.float64)
    rc = lib.oracle_set_cv(mode, 1 if fallback else 0, max_ratio, tp, t, _p(dom))
    assert rc == 0, rc


def tabulate(fn, t, domain):
    """(t, t, 4) float32 table of fn(x, y) -> (u, ux, uy) on the grid of spac_cv_config."""
    xs = domain[0] + np.arange(t) * ((domain[2] - domain[0]) / (t - 1))
    ys = domain[1] + np.arange(t) * ((domain[3] - domain[1]) / (t - 1))
    X, Y = np.meshgrid(xs, ys, indexing="xy")  # [j, i] = (x_i, y_j)
    u, ux, uy = fn(X, Y)
    out = np.zeros((t, t, 4), dtype=np.float32)
    out[..., 0], out[..., 1], out[..., 2] = u, ux, uy
    return out


_coef_tables_keep = None


def set_coef_tables(mask=0, values=None, domain=None):
    """oracle_set_coef_tables: tabulated coefficients (values (T, T, 8)) of later calls
    (process-wide; reset with set_coef_tables() when done)."""
    global _coef_tables_keep
    lib = load()
    lib.oracle_set_coef_tables.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double)]
    lib.oracle_set_coef_tables.restype = C.c_int
    if not mask:
        _coef_tables_keep = None
        lib.oracle_set_coef_tables(0, None, 0, None)
        return
    v = np.ascontiguousarray(values, dtype=np.float64)
    d = np.ascontiguousarray(domain, dtype=np.float64)
    _coef_tables_keep = (v, d)
    lib.oracle_set_coef_tables(mask, _p(v), v.shape[0], _p(d))
