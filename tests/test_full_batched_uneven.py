"""The batched uneven workload at its full size, through the graph-captured plan.

16 x 16 circles (110 nodes per circle, 24,868 nodes, 48,268 rows), per-row paths log-uniform
in [1e3, 1e5] and h log-uniform in [1e-4, 1e-3] (the draws of spac_amd.workloads), ~1.03e9
paths in one replayed graph per precision.

Bars:
- FP64 plan: 50 rows spread over the batch equal the CPU restatement bit for bit (B, every G
  weight, path steps), each at its own paths and h (Solvers/FeynmanKacSolver.cpp:344-355 plus
  Mesh/Interfaz.cpp:219-264 per row);
- FP32 plan against the FP64 plan with the same seed (same Philox stream): every row's B
  within 0.5 of its FP64 standard error, or off by no more than one path's exit value (|g|/N,
  |g| <= 1.5 for the Franke u on the unit square: a path that leaves a circle in one precision
  and the domain in the other, where the two boundaries are equally near, moves g/N between
  B and the row's G), for at most 5 rows; the RMS over the other rows within 0.01 of the
  standard error (a path whose exit step moves under FP32 rounding changes its row by O(1/N)
  of a path's spread; measured: max 0.33-0.55, RMS 0.0036-0.004, one or two exit-kind rows per
  batch); the solved interface values within 2e-6 (RMS) of the FP64 solve.
"""
import sys

import numpy as np
import pytest

import oracle_lib as o
from spac_amd import abi
from spac_amd.fk import FKProblem, RowsPlan, problem_desc
from spac_amd.mesh import gmres
from spac_amd.workloads import UNIT_SQUARE, batched_uneven

pytestmark = pytest.mark.gpu

VC = abi.SPAC_COEF_VARCOEF_FRANKE
SEED = 0


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def full_rows():
    mesh, paths, h = batched_uneven()
    live = mesh.job_frontera == 0
    say(f"batched uneven: {mesh.info.n_jobs} rows, {int(paths[live].astype(np.int64).sum())} paths")
    out = {}
    for prec in (abi.SPAC_FP64, abi.SPAC_FP32):
        with FKProblem(problem_desc(coef=VC, domain=UNIT_SQUARE, precision=prec)) as p:
            with RowsPlan(p, mesh, paths, h) as plan:
                out[prec] = plan.execute(seed=SEED)
                say(f"plan precision {prec}: {plan.last_ms:.1f} ms, "
                    f"{sum(s['path_steps'] for s in out[prec]['stats'])} path-steps")
    return mesh, paths, h, out


def test_fp64_plan_rows_equal_oracle(gpu, full_rows):
    mesh, paths, h, out = full_rows
    rows = out[abi.SPAC_FP64]
    live = np.nonzero(mesh.job_frontera == 0)[0]
    jobs = live[np.linspace(0, live.size - 1, 50).astype(np.int64)]
    say(f"oracle on 50 rows: paths {int(paths[jobs].min())}..{int(paths[jobs].max())}, "
        f"h {h[jobs].min():.2e}..{h[jobs].max():.2e}")
    desc = problem_desc(coef=VC, domain=UNIT_SQUARE)
    B, G, st = o.assemble_rows(desc, mesh, paths[jobs], h[jobs], seed=SEED, jobs=jobs, nthreads=16)
    goff = rows["G_offset"]
    gpu_G = np.concatenate([rows["G"][goff[j]:goff[j + 1]] for j in jobs])
    assert np.array_equal(rows["B"][jobs].view(np.uint64), B.view(np.uint64))
    assert np.array_equal(gpu_G.view(np.uint64), G.view(np.uint64))
    assert [rows["stats"][j]["path_steps"] for j in jobs] == [s["path_steps"] for s in st]
    assert all(rows["stats"][j]["status"] == 0 for j in jobs)


def _solve(mesh, rows):
    mesh.accumulate(rows)
    Gi, Gj, Gv, B = mesh.csr()
    u, info = gmres(Gi, Gj, Gv, B, tol=1e-10, max_iter=20000, restart=200, device=True)
    assert info.converged, info.residual
    return u


def test_fp32_plan_tracks_fp64_plan(gpu, full_rows):
    mesh, paths, h, out = full_rows
    r64, r32 = out[abi.SPAC_FP64], out[abi.SPAC_FP32]
    live = mesh.job_frontera == 0
    assert all(s["status"] == 0 for s in r32["stats"])
    s64 = r64["stats"]
    var = np.array([max(s["phiphi"] - s["phi"] ** 2, 0.0) for s in s64])
    se = np.sqrt(var / np.maximum(paths, 1))
    diff = np.abs(r32["B"] - r64["B"])[live]
    dB = diff / se[live]
    one_exit = (dB >= 0.5) & (diff <= 1.5 / paths[live])  # one path's exit kind differs
    say(f"FP32 - FP64 row B: max {dB.max():.3e} SE, rms {np.sqrt(np.mean(dB ** 2)):.3e} SE, "
        f"{int(one_exit.sum())} rows off by one path's exit value")
    assert ((dB < 0.5) | one_exit).all()
    assert one_exit.sum() <= 5
    assert np.sqrt(np.mean(dB[~one_exit] ** 2)) < 0.01
    u64 = _solve(mesh, r64)
    u32 = _solve(mesh, r32)
    d = np.sqrt(np.mean((u32 - u64) ** 2))
    # the FP64 solve's statistical error at these path counts, from the row variances
    stat = float(np.sqrt(np.mean(se[live] ** 2)))
    say(f"solved u FP32 - FP64: rms {d:.3e}, max {np.abs(u32 - u64).max():.3e}; row SE rms {stat:.3e}")
    assert d < 2e-6
