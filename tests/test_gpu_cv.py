"""Control-variate modes on the GPU against the CPU restatement, and what the
learned control variate gains.

Bar for the FP64 kernel: per-path records and estimators bit-identical to the
oracle in every mode (plugin F, plain MC, tabulated F, fallback).  FP32 table
mode tracks the FP64 table mode within a fraction of the standard error.
The learned control variate is trained on plain-MC point estimates by the
library's own kernels and must cut the per-node variance by a large factor
while its estimates stay consistent with plain MC.
"""
import sys

import numpy as np
import pytest

import oracle_lib as o
from spac_amd import abi
from spac_amd.fk import FKProblem, SolveBatchArrays, problem_desc
from spac_amd.problems import franke_table_fn, franke_u

pytestmark = pytest.mark.gpu

VC = abi.SPAC_COEF_VARCOEF_FRANKE
FR = abi.SPAC_COEF_FRANKE_POISSON
DOM = (0.0, 0.0, 1.0, 1.0)
NP = 1500


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def reset_oracle_cv():
    o.set_cv()
    yield
    o.set_cv()


def _batch():
    x0 = np.array([0.30, 0.55, 0.72, 0.45])
    y0 = np.array([0.40, 0.50, 0.22, 0.75])
    circle = np.tile([0.33, 0.55, 0.45], 4)
    return SolveBatchArrays(x0, y0, NP, 5e-4, [5, 6, 7, 8], circle=circle)


@pytest.fixture(scope="module")
def exact_table():
    return o.tabulate(franke_table_fn, 1025, DOM)


def _assert_same(rd, ro, sd, so):
    for k in ro:
        a, b = rd[k], ro[k]
        if a.dtype.kind == "f":
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), k
        else:
            assert np.array_equal(a, b), k
    for a, b in zip(sd, so):
        for k in a:
            if isinstance(a[k], float):
                assert np.float64(a[k]).view(np.uint64) == np.float64(b[k]).view(np.uint64), k
            else:
                assert a[k] == b[k], k


@pytest.mark.parametrize("mode,table,fallback", [("none", None, False), ("table", "exact", False),
                                                 ("table", "bad", True), ("table", "exact", True)])
def test_fp64_cv_modes_bit_identical(gpu, exact_table, mode, table, fallback):
    say(f"cv parity {mode} {table} fallback={fallback}")
    m = abi.SPAC_CV_NONE if mode == "none" else abi.SPAC_CV_TABLE
    tab = None
    if table == "exact":
        tab = exact_table
    elif table == "bad":
        tab = exact_table.copy()
        tab[..., 1:3] *= -5.0
    b = _batch()
    desc = problem_desc(coef=VC, domain=DOM)
    o.set_cv(m, fallback=fallback, table=tab, domain=DOM if tab is not None else None)
    ro = o.simulate(desc, b, seed=3)
    so = o.reduce(b, ro, node_mode=True)
    with FKProblem(desc) as p:
        p.set_control_variate(m, tab, DOM, fallback=fallback)
        rd, sd = p.simulate(b, seed=3)
    _assert_same(rd, ro, sd, so)
    if table == "bad":
        assert all(s["cv_fallback"] == 1 for s in sd)


def test_fp64_table_points_bit_identical(gpu):
    say("cv parity: Franke points, table over [-1,1]^2")
    dom = (-1.0, -1.0, 1.0, 1.0)
    tab = o.tabulate(franke_table_fn, 513, dom)
    b = SolveBatchArrays([0.1, -0.4, 0.6], [0.2, 0.5, -0.7], 800, 1e-3, [1, 2, 3])
    desc = problem_desc()
    o.set_cv(abi.SPAC_CV_TABLE, table=tab, domain=dom)
    ro = o.simulate(desc, b, seed=0)
    so = o.reduce(b, ro)
    with FKProblem(desc) as p:
        p.set_control_variate(abi.SPAC_CV_TABLE, tab, dom)
        rd, sd = p.simulate(b, seed=0)
    _assert_same(rd, ro, sd, so)


def test_fp32_table_tracks_fp64(gpu, exact_table):
    say("fp32 table mode vs fp64 table mode")
    b = _batch()
    out = {}
    for prec in (abi.SPAC_FP64, abi.SPAC_FP32):
        with FKProblem(problem_desc(coef=VC, domain=DOM, precision=prec)) as p:
            p.set_control_variate(abi.SPAC_CV_TABLE, exact_table, DOM)
            _, st = p.simulate(b, seed=11, want_records=False)
        out[prec] = st
    for a, c in zip(out[abi.SPAC_FP32], out[abi.SPAC_FP64]):
        se = np.sqrt(max(c["phiphi_VR"] - c["phi_VR"] ** 2, 0) / c["paths"])
        assert abs(a["phi_VR"] - c["phi_VR"]) < 0.25 * se + 2e-4
        assert abs(a["B"] - c["B"]) < 2e-3


def test_invalid_cv_config_rejected(gpu):
    with FKProblem(problem_desc(coef=VC, domain=DOM)) as p:
        with pytest.raises(abi.SpacError):
            p.set_control_variate(7)
        with pytest.raises(abi.SpacError):
            p.set_control_variate(abi.SPAC_CV_TABLE, np.zeros((1, 1, 4), np.float32), DOM)
        with pytest.raises(abi.SpacError):
            p.set_control_variate(abi.SPAC_CV_TABLE, np.zeros((4, 4, 4), np.float32), (1.0, 0.0, 0.0, 1.0))


def test_learned_cv_gain(gpu):
    """Train the MLP on plain-MC point estimates of the variable-coefficient
    operator (4096 points), tabulate it with the MFMA kernel, and compare node
    variances with and without it on circle solves of the same operator."""
    from spac_amd.mlp import MLP

    say("learned CV: training data")
    rng = np.random.default_rng(0)
    g = (np.arange(64) + 0.5) / 64
    X, Y = np.meshgrid(g, g, indexing="xy")
    xs = np.clip(X.ravel() + (rng.random(4096) - 0.5) / 64, 1e-3, 1 - 1e-3)
    ys = np.clip(Y.ravel() + (rng.random(4096) - 0.5) / 64, 1e-3, 1 - 1e-3)
    with FKProblem(problem_desc(coef=VC, domain=DOM, precision=abi.SPAC_FP32)) as p:
        p.set_control_variate(abi.SPAC_CV_NONE)
        st, _ = p.solve_points(xs, ys, 2000, 5e-4, seed=1)
    u_mc = np.array([s["phi"] for s in st])
    noise = np.sqrt(np.mean([(s["phiphi"] - s["phi"] ** 2) / s["paths"] for s in st]))
    say(f"training set: 4096 points, MC noise {noise:.2e}, rms vs exact {np.sqrt(np.mean((u_mc - franke_u(xs, ys)) ** 2)):.2e}")
    with MLP(DOM, seed=0) as m:
        rep = m.train(xs, ys, u_mc, steps=6000, lr=3e-3, lr_final=1e-5, grad_weight=0.03)
        t_ms = m.build_table(1024)
        say(f"trained: loss {rep['final_loss']:.3e} in {rep['train_ms']:.0f} ms; table {t_ms:.2f} ms")
        rng = np.random.default_rng(5)
        n = 64
        cx, cy = rng.uniform(0.25, 0.75, n), rng.uniform(0.25, 0.75, n)
        r = rng.uniform(0.0, 0.2, n)
        a = rng.uniform(0, 2 * np.pi, n)
        b = SolveBatchArrays(cx + r * np.cos(a), cy + r * np.sin(a), 4000, 5e-4, np.arange(n),
                             circle=np.stack([np.full(n, 0.25), cx, cy], 1))
        res = {}
        for mode in ("plain", "learned"):
            with FKProblem(problem_desc(coef=VC, domain=DOM, precision=abi.SPAC_FP32)) as p:
                if mode == "plain":
                    p.set_control_variate(abi.SPAC_CV_NONE)
                else:
                    m.attach(p, fallback=True)
                _, res[mode] = p.simulate(b, seed=2, want_records=False)
    var_plain = np.array([s["phiphi"] - s["phi"] ** 2 for s in res["plain"]])
    var_cv = np.array([s["phiphi_VR"] - s["phi_VR"] ** 2 for s in res["learned"]])
    ratio = var_cv / var_plain
    fb = sum(s["cv_fallback"] for s in res["learned"])
    say(f"variance ratio learned/plain: median {np.median(ratio):.3f}, mean {ratio.mean():.3f}, "
        f"max {ratio.max():.3f}; fallbacks {fb}/{n}")
    assert np.mean(var_cv) < 0.25 * np.mean(var_plain)
    # same target: B with and without the control variate agree within the plain error
    Bp = np.array([s["B"] for s in res["plain"]])
    Bl = np.array([s["B"] for s in res["learned"]])
    se = np.sqrt(var_plain / 4000)
    assert np.all(np.abs(Bl - Bp) < 5 * se + 1e-3)


def test_network_cv_per_step_against_table(gpu):
    """SPAC_CV_NETWORK (the trained network's gradient heads evaluated at every step) against
    its table (SPAC_CV_TABLE) on the same Philox paths: the plain score phi and the path-step
    counts are identical (the control variate does not move a path), xi is a martingale (mean
    0 within 4 standard errors), and the variances of phi_VR agree within 10 %: the table's
    interpolation changes the control variate by less than the network's own error.  A plan of
    a network-mode problem replays the eager assembly bit for bit; FP64 problems and the
    generic entry point refuse the mode."""
    from spac_amd.mlp import MLP
    from spac_amd.workloads import LEARNED_CV, train_learned_cv
    from spac_amd.fk import RowsPlan
    from spac_amd.mesh import Mesh, mesh_desc

    say("network CV: train")
    cfg = {k: v for k, v in LEARNED_CV.items() if k not in ("seed", "first_layer_scale")}
    desc32 = problem_desc(coef=VC, domain=DOM, precision=abi.SPAC_FP32)
    rng = np.random.default_rng(8)
    n = 48
    cx, cy = rng.uniform(0.25, 0.75, n), rng.uniform(0.25, 0.75, n)
    r = rng.uniform(0.0, 0.05, n)
    a = rng.uniform(0, 2 * np.pi, n)
    b = SolveBatchArrays(cx + r * np.cos(a), cy + r * np.sin(a), 2000, 1e-4, np.arange(n),
                         circle=np.stack([np.full(n, 0.08), cx, cy], 1))
    res = {}
    with MLP(DOM, seed=0, first_layer_scale=LEARNED_CV["first_layer_scale"]) as m, FKProblem(desc32) as dp:
        train_learned_cv(m, **cfg, eager=False, data_problem=dp)
        for mode in ("table", "network"):
            with FKProblem(desc32) as p:
                m.attach(p, fallback=False, network=(mode == "network"))
                res[mode] = p.simulate(b, seed=4, want_records=True)
        with FKProblem(problem_desc(coef=VC, domain=DOM)) as p64:
            with pytest.raises(abi.SpacError) as e:
                m.attach(p64, network=True)
            assert e.value.code == abi.SPAC_E_INVALID_ARG
        mesh = Mesh(mesh_desc())
        jobs = np.arange(0, mesh.info.n_jobs, 41)
        with FKProblem(desc32) as p:
            m.attach(p, fallback=True, network=True)
            rows = p.assemble_rows(mesh, 300, 5e-4, seed=9, jobs=jobs)
            with RowsPlan(p, mesh, 300, 5e-4, jobs=jobs) as plan:
                again = plan.execute(seed=9)
        assert np.array_equal(rows["B"].view(np.uint64), again["B"].view(np.uint64))
        assert np.array_equal(rows["G"].view(np.uint64), again["G"].view(np.uint64))
    with FKProblem(desc32) as p:
        with pytest.raises(abi.SpacError):
            p.set_control_variate(abi.SPAC_CV_NETWORK, np.zeros((4, 4, 4), np.float32), DOM)
    (rt, st), (rn, sn) = res["table"], res["network"]
    assert np.array_equal(np.asarray(rt["steps"]), np.asarray(rn["steps"]))
    vt = np.array([s["phiphi_VR"] - s["phi_VR"] ** 2 for s in st])
    vn = np.array([s["phiphi_VR"] - s["phi_VR"] ** 2 for s in sn])
    vp = np.array([s["phiphi"] - s["phi"] ** 2 for s in sn])
    for s1, s2 in zip(st, sn):
        assert s1["phi"] == s2["phi"]
    xi = np.asarray(rn["xi"])
    se = xi.std() / np.sqrt(xi.size)
    say(f"variance plain/table {vp.sum() / vt.sum():.1f}, plain/network {vp.sum() / vn.sum():.1f}; "
        f"mean xi {xi.mean():.2e} (SE {se:.1e})")
    assert abs(xi.mean()) < 4 * se
    assert abs(vn.sum() / vt.sum() - 1.0) < 0.1
    assert vp.sum() > 5 * vn.sum()
