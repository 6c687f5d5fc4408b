"""Asynchronous calls (spac_fk_solve_points_device_async, spac_fk_assemble_system_async,
spac_fk_wait, spac_fk_query) through the C ABI.

Bar: the results after spac_fk_wait equal the synchronous calls' bit for bit; one call is in
flight per handle and every other call on the handle is refused with a status until the wait;
waiting with nothing in flight is a no-op.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from spac_amd import abi
from spac_amd.fk import FKProblem, problem_desc
from spac_amd.mesh import DeviceSystem, Mesh, mesh_desc

pytestmark = pytest.mark.gpu

VC = abi.SPAC_COEF_VARCOEF_FRANKE


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_async_point_solve_equals_sync(gpu):
    lib = abi.load()
    rng = np.random.default_rng(4)
    n, paths, h = 128, 4000, 1e-3
    x, y = rng.uniform(-0.9, 0.9, n), rng.uniform(-0.9, 0.9, n)
    dx, dy = abi.DeviceArray.from_numpy(x, lib), abi.DeviceArray.from_numpy(y, lib)
    sync_st, async_st = (abi.DeviceArray(n * C.sizeof(abi.PointStats), lib) for _ in range(2))
    sync_st.zero()
    async_st.zero()
    with FKProblem(problem_desc()) as p:
        say("async points")
        abi.check(lib.spac_fk_solve_points_device(p.h, dx.ptr, dy.ptr, n, paths, h, 9, sync_st.ptr), lib)
        abi.check(lib.spac_fk_solve_points_device_async(p.h, dx.ptr, dy.ptr, n, paths, h, 9, async_st.ptr), lib)
        # one call in flight: everything else on the handle is refused until the wait
        assert lib.spac_fk_solve_points_device_async(p.h, dx.ptr, dy.ptr, n, paths, h, 9, async_st.ptr) == \
            abi.SPAC_E_INVALID_ARG
        assert "spac_fk_wait" in abi.last_error(lib)
        assert lib.spac_fk_solve_points_device(p.h, dx.ptr, dy.ptr, n, paths, h, 9, sync_st.ptr) == \
            abi.SPAC_E_INVALID_ARG
        with pytest.raises(abi.SpacError):
            p.set_launch(block_size=128)
        done = C.c_int32(-1)
        abi.check(lib.spac_fk_query(p.h, C.byref(done)), lib)
        assert done.value in (0, 1)
        abi.check(lib.spac_fk_wait(p.h), lib)
        abi.check(lib.spac_fk_query(p.h, C.byref(done)), lib)
        assert done.value == 1
        abi.check(lib.spac_fk_wait(p.h), lib)  # nothing in flight
        t = p.timing()
        assert t.path_kernel_ms > 0 and t.total_ms >= t.path_kernel_ms
        assert async_st.to_bytes() == sync_st.to_bytes()
        p.set_launch(block_size=128)  # usable again
    for d in (dx, dy, sync_st, async_st):
        d.free()


@pytest.mark.parametrize("precision", [abi.SPAC_FP64, abi.SPAC_FP32])
def test_async_system_assembly_equals_sync(gpu, precision):
    lib = abi.load()
    m = Mesh(mesh_desc(n_side=6, nodes_per_circle=30))
    paths, h = 300, 5e-4
    with FKProblem(problem_desc(coef=VC, domain=(0.0, 0.0, 1.0, 1.0), precision=precision)) as p, \
            DeviceSystem(m) as s:
        say(f"async system, precision {precision}: {m.info.n_jobs} jobs")
        st_sync = p.assemble_system(m, s, paths, h, seed=13)
        Gi, Gj, Gv, B = s.read()
        ms_sync = p.timing().total_ms
        nb, keep = m.node_batch(paths, h)
        st = (abi.PointStats * m.info.n_jobs)()
        abi.check(lib.spac_fk_assemble_system_async(p.h, C.byref(m.table), C.byref(nb), 13, s.h, st), lib)
        with pytest.raises(abi.SpacError):
            p.assemble_system(m, s, paths, h, seed=13)
        abi.check(lib.spac_fk_wait(p.h), lib)
        Gi2, Gj2, Gv2, B2 = s.read()
        assert np.array_equal(Gv.view(np.int64), Gv2.view(np.int64))
        assert np.array_equal(B.view(np.int64), B2.view(np.int64))
        assert bytes(st) == bytes(np.ascontiguousarray(st_sync).tobytes())
        say(f"sync {ms_sync:.2f} ms, async {p.timing().total_ms:.2f} ms on the device")
        u, info = s.solve(tol=1e-10, max_iter=2000, restart=100)
        assert info.converged
