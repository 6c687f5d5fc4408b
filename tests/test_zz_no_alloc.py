"""The no-allocation point solve (spac_fk_solve_points_device_noalloc) against the
allocating one (spac_fk_solve_points_device) through the C ABI.

Bar: the estimator records of both calls are bit-identical, on two calls in a row
(the second smaller, so it runs over the first call's scratch contents), in FP64
and FP32 stepping and with the plugin and table control variates; sizes beyond
the scratch maximum are refused with a status before anything runs.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from spac_amd import abi
from spac_amd.fk import FKProblem, problem_desc

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _points(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-0.9, 0.9, n), rng.uniform(-0.9, 0.9, n)


def _solve(lib, prob, fn, x, y, paths, h, seed):
    """One call of a device-resident point solve; device memory through the library's own
    runtime (no second HIP runtime in the process)."""
    n = x.size
    dx = abi.DeviceArray.from_numpy(np.ascontiguousarray(x, np.float64), lib)
    dy = abi.DeviceArray.from_numpy(np.ascontiguousarray(y, np.float64), lib)
    dst = abi.DeviceArray(n * C.sizeof(abi.PointStats), lib)
    dst.zero()
    abi.check(fn(prob.h, dx.ptr, dy.ptr, n, paths, h, seed, dst.ptr), lib)
    raw = dst.to_bytes()
    for d in (dx, dy, dst):
        d.free()
    return raw, (abi.PointStats * n).from_buffer_copy(raw)


@pytest.mark.parametrize("precision", [abi.SPAC_FP64, abi.SPAC_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("cv", ["analytic", "table"])
def test_noalloc_equals_allocating_solve_twice(gpu, precision, cv):
    lib = abi.load()
    with FKProblem(problem_desc(precision=precision)) as prob:
        if cv == "table":
            from spac_amd.problems import franke_table_fn

            g = np.linspace(-1.0, 1.0, 64)
            X, Y = np.meshgrid(g, g, indexing="xy")  # [j, i] = (x_i, y_j)
            u, ux, uy = franke_table_fn(X, Y)
            table = np.stack([u, ux, uy, np.zeros_like(u)], -1).astype(np.float32)
            prob.set_control_variate(abi.SPAC_CV_TABLE, table=table, domain=(-1.0, -1.0, 1.0, 1.0), fallback=True)
        # first call near the scratch maximum, then a smaller one over its leftovers
        for n, paths, h, seed in [(256, 2048, 1e-3, 11), (37, 1000, 2e-3, 12)]:
            assert n * paths <= abi.SPAC_NOALLOC_MAX_PATHS
            say(f"noalloc {precision=} {cv=} n={n} paths={paths} h={h} seed={seed}")
            x, y = _points(n, seed)
            ref_raw, ref = _solve(lib, prob, lib.spac_fk_solve_points_device, x, y, paths, h, seed)
            na_raw, na = _solve(lib, prob, lib.spac_fk_solve_points_device_noalloc, x, y, paths, h, seed)
            assert all(s.status == 0 for s in na)
            assert all(s.paths == paths for s in na)
            assert sum(s.path_steps for s in na) > n * paths
            assert na_raw == ref_raw


def test_noalloc_refuses_sizes_beyond_scratch(gpu):
    lib = abi.load()
    with FKProblem(problem_desc()) as prob:
        x, y = _points(4, 0)
        dx = (C.c_double * 4)(*x)
        st = (abi.PointStats * 4)()
        # checked on the host before any launch: the pointers are never dereferenced
        big = abi.SPAC_NOALLOC_MAX_PATHS // 4 + 1
        rc = lib.spac_fk_solve_points_device_noalloc(prob.h, C.addressof(dx), C.addressof(dx), 4, big, 1e-3, 0,
                                                     C.addressof(st))
        assert rc == abi.SPAC_E_INVALID_ARG
        assert "scratch maximum" in abi.last_error(lib)
        rc = lib.spac_fk_solve_points_device_noalloc(prob.h, C.addressof(dx), C.addressof(dx),
                                                     abi.SPAC_NOALLOC_MAX_SOLVES + 1, 1, 1e-3, 0, C.addressof(st))
        assert rc == abi.SPAC_E_INVALID_ARG
        assert "scratch maximum" in abi.last_error(lib)
