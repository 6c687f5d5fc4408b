"""The full-scale interface system: 64 x 64 circles, 99,604 nodes, 186,920 rows, assembled on
the device (spac_fk_assemble_system) and solved there by two-level GPU GMRES
(spac_gmres_solve_system), in plain Monte Carlo and with the learned control variate.

Bars (all at 256 paths per row, the system's h):
- every assembly ends with status 0 and every solve converges (true residual < 1e-8);
- FP32 plain MC against FP64 plain MC with the same seed (same Philox stream): solved u within
  0.1 x the statistical RMS of the plain-MC solve at this path count (measured over 6 FP32
  replicas);
- learned control variate (network trained by the library's kernels, BF16-MFMA table): at most
  5 % of the rows fall back to plain MC, its statistical RMS over 4 replicas is below 1/3 of
  plain MC's at equal paths, and its solved u differs from the FP64 plain-MC solve by no more
  than their combined statistical error (1.5 x the plain RMS, which dominates).
"""
import sys

import numpy as np
import pytest

from spac_amd import abi
from spac_amd.fk import FKProblem, problem_desc
from spac_amd.mesh import DeviceSystem
from spac_amd.workloads import LEARNED_CV, UNIT_SQUARE, full_scale, train_learned_cv

pytestmark = pytest.mark.gpu

VC = abi.SPAC_COEF_VARCOEF_FRANKE
PATHS = 256
TOL = 1e-8


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _replica(prob, mesh, system, h, seed, stats=True):
    st = prob.assemble_system(mesh, system, PATHS, h, seed=seed, stats=stats)
    u, info = system.solve(tol=TOL, max_iter=20000, restart=200)
    assert info.converged and info.residual < TOL, (info.iterations, info.residual)
    if st is not None:
        assert (st["status"] == 0).all()
    return u, st, info


def _stat_rms(us):
    U = np.array(us)
    return float(np.sqrt(U.var(0, ddof=1).mean()))


@pytest.fixture(scope="module")
def system():
    mesh, h = full_scale()
    say(f"full scale: {mesh.info.n_nodes} nodes, {mesh.info.n_jobs} rows, nnz {mesh.info.nnz}, h {h:.3e}")
    s = DeviceSystem(mesh)
    s.set_coarse(12, 12)
    yield mesh, h, s
    s.close()


def test_full_scale_fp32_and_learned_against_fp64(gpu, system):
    mesh, h, s = system
    seed = 7
    with FKProblem(problem_desc(coef=VC, domain=UNIT_SQUARE)) as p64:
        p64.set_control_variate(abi.SPAC_CV_NONE)
        u64, st64, info = _replica(p64, mesh, s, h, seed)
        say(f"FP64 plain: {int(st64['path_steps'].sum())} path-steps, GMRES {info.iterations} it.")
    with FKProblem(problem_desc(coef=VC, domain=UNIT_SQUARE, precision=abi.SPAC_FP32)) as p32:
        p32.set_control_variate(abi.SPAC_CV_NONE)
        u32, st32, _ = _replica(p32, mesh, s, h, seed)
        plain = [u32] + [_replica(p32, mesh, s, h, seed + 100 + k, stats=False)[0] for k in range(5)]
    rms_plain = _stat_rms(plain)
    d32 = float(np.sqrt(np.mean((u32 - u64) ** 2)))
    steps_diff = int(np.abs(st32["path_steps"].astype(np.int64) - st64["path_steps"].astype(np.int64)).sum())
    say(f"plain statistical RMS at {PATHS} paths: {rms_plain:.3e}; FP32 - FP64 solve: rms {d32:.3e}, "
        f"max {np.abs(u32 - u64).max():.3e}; path-step count differs by {steps_diff} "
        f"of {int(st64['path_steps'].sum())}")
    assert d32 < 0.1 * rms_plain

    from spac_amd.mlp import MLP

    with MLP(UNIT_SQUARE, seed=LEARNED_CV["seed"], first_layer_scale=LEARNED_CV["first_layer_scale"]) as net, \
            FKProblem(problem_desc(coef=VC, domain=UNIT_SQUARE, precision=abi.SPAC_FP32)) as pl:
        cfg = {k: v for k, v in LEARNED_CV.items() if k not in ("seed", "first_layer_scale")}
        setup = train_learned_cv(net, **cfg)
        say(f"learned CV setup: {setup}")
        net.attach(pl, fallback=True)
        uL, stL, _ = _replica(pl, mesh, s, h, seed)
        learned = [uL] + [_replica(pl, mesh, s, h, seed + 100 + k, stats=False)[0] for k in range(3)]
    share = float(stL["cv_fallback"].sum()) / stL.size
    rms_learned = _stat_rms(learned)
    dL = float(np.sqrt(np.mean((uL - u64) ** 2)))
    say(f"learned: fallback share {share:.4f}, statistical RMS {rms_learned:.3e} (plain {rms_plain:.3e}), "
        f"learned - FP64 plain: rms {dL:.3e}")
    assert share < 0.05
    assert rms_learned < rms_plain / 3.0
    assert dL < 1.5 * rms_plain
