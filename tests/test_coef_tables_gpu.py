"""Tabulated coefficients on the device (spac_fk_set_coef_tables) against the CPU restatement.

Bars: FP64 per-path records and estimators equal the restatement's bit for bit for several
masks and control-variate modes, and so do FP64 rows; FP32 stepping stays within a quarter
standard error of FP64; a table of the compiled coefficients reproduces the compiled records
within interpolation error; a captured plan replays the tabulated assembly bit for bit.
"""
import sys

import numpy as np
import pytest

import oracle_lib as o
from spac_amd import abi
from spac_amd.fk import FKProblem, RowsPlan, SolveBatchArrays, problem_desc
from spac_amd.mesh import Mesh, mesh_desc
from test_coef_tables import compiled_table

pytestmark = pytest.mark.gpu

VC = abi.SPAC_COEF_VARCOEF_FRANKE
DOM = (0.0, 0.0, 1.0, 1.0)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _reset():
    yield
    o.set_coef_tables()
    o.set_cv()


def _perturbed_table(T=65, seed=0):
    """A table unlike any compiled set: the compiled values plus smooth perturbations."""
    v = compiled_table(T)
    g = np.linspace(0, 1, T)
    X, Y = np.meshgrid(g, g, indexing="xy")
    v[..., 0] *= 1.0 + 0.2 * np.sin(3 * X) * np.cos(2 * Y)
    v[..., 1] += 0.1 * np.cos(4 * Y)
    v[..., 2] -= 0.1 * np.sin(5 * X)
    v[..., 3] += 0.3 * X * Y
    v[..., 4] *= 0.9 + 0.2 * X
    v[..., 5] += 0.05 * Y
    v[..., 6] -= 0.03 * X
    return v


def _batch(n=12, paths=400):
    rng = np.random.default_rng(4)
    cx, cy = rng.uniform(0.25, 0.75, n), rng.uniform(0.25, 0.75, n)
    a = rng.uniform(0, 2 * np.pi, n)
    R = rng.uniform(0.1, 0.3, n)
    x0 = np.clip(cx + 0.5 * R * np.cos(a), 0.02, 0.98)
    y0 = np.clip(cy + 0.5 * R * np.sin(a), 0.02, 0.98)
    return SolveBatchArrays(x0, y0, paths, 5e-4, np.arange(n), circle=np.stack([R, cx, cy], 1))


def _same(rg, sg, ro, so):
    for k in ro:
        assert np.array_equal(np.asarray(rg[k]), np.asarray(ro[k])), k
    for a, c in zip(sg, so):
        for k in a:
            assert a[k] == c[k], k


@pytest.mark.parametrize("mask", [abi.SPAC_TAB_ALL, abi.SPAC_TAB_SIGMA, abi.SPAC_TAB_SOURCE | abi.SPAC_TAB_KILLING,
                                  abi.SPAC_TAB_DRIFT, abi.SPAC_TAB_DIRICHLET | abi.SPAC_TAB_SOLUTION])
@pytest.mark.parametrize("cv", [abi.SPAC_CV_NONE, abi.SPAC_CV_ANALYTIC])
def test_records_equal_restatement(gpu, mask, cv):
    say(f"tables: mask {mask}, cv {cv}")
    tab = _perturbed_table()
    desc = problem_desc(coef=VC, domain=DOM)
    b = _batch()
    with FKProblem(desc) as p:
        p.set_control_variate(cv)
        p.set_coef_tables(mask, tab, DOM)
        rg, sg = p.simulate(b, seed=7)
    o.set_cv(cv)
    o.set_coef_tables(mask, tab, DOM)
    ro = o.simulate(desc, b, seed=7)
    so = o.reduce(b, ro, node_mode=True)
    _same(rg, sg, ro, so)


def test_rows_and_plan_equal_restatement(gpu):
    m = Mesh(mesh_desc())
    tab = _perturbed_table(129)
    jobs = np.nonzero(m.job_frontera == 0)[0][::53][:20]
    jobs = np.concatenate([jobs, np.nonzero(m.job_frontera)[0][:3]])
    desc = problem_desc(coef=VC, domain=DOM)
    with FKProblem(desc) as p:
        p.set_coef_tables(abi.SPAC_TAB_ALL, tab, DOM)
        rows = p.assemble_rows(m, 300, 5e-4, seed=9, jobs=jobs)
        with RowsPlan(p, m, 300, 5e-4, jobs=jobs) as plan:
            for _ in range(2):
                again = plan.execute(seed=9)
                assert np.array_equal(again["B"].view(np.uint64), rows["B"].view(np.uint64))
                assert np.array_equal(again["G"].view(np.uint64), rows["G"].view(np.uint64))
    o.set_coef_tables(abi.SPAC_TAB_ALL, tab, DOM)
    B, G, st = o.assemble_rows(desc, m, 300, 5e-4, seed=9, jobs=jobs, nthreads=8)
    assert np.array_equal(rows["B"].view(np.uint64), B.view(np.uint64))
    assert np.array_equal(rows["G"].view(np.uint64), G.view(np.uint64))
    # boundary nodes take the tabulated g
    front = m.job_frontera[jobs] != 0
    assert np.all(rows["B"][front] == B[front])


def test_fp32_tables_track_fp64(gpu):
    tab = _perturbed_table()
    b = _batch(paths=4000)
    out = {}
    for prec in (abi.SPAC_FP64, abi.SPAC_FP32):
        with FKProblem(problem_desc(coef=VC, domain=DOM, precision=prec)) as p:
            p.set_control_variate(abi.SPAC_CV_NONE)
            p.set_coef_tables(abi.SPAC_TAB_ALL, tab, DOM)
            _, out[prec] = p.simulate(b, seed=3, want_records=False)
    for a, c in zip(out[abi.SPAC_FP32], out[abi.SPAC_FP64]):
        se = np.sqrt(max(c["phiphi"] - c["phi"] ** 2, 0) / c["paths"])
        assert abs(a["phi"] - c["phi"]) < 0.25 * se + 2e-4
        assert abs(a["B"] - c["B"]) < 0.25 * se + 2e-4


def test_table_of_compiled_set_reproduces_compiled_records(gpu):
    """Coeficiente_*_Numerico of the compiled set (T = 513, h = 1/512): the estimators move by the
    interpolation error only (same Philox paths)."""
    b = _batch(paths=4000)
    desc = problem_desc(coef=VC, domain=DOM)
    with FKProblem(desc) as p:
        p.set_control_variate(abi.SPAC_CV_NONE)
        _, ref = p.simulate(b, seed=2, want_records=False)
        p.set_coef_tables(abi.SPAC_TAB_ALL, compiled_table(513), DOM)
        _, tab = p.simulate(b, seed=2, want_records=False)
        p.set_coef_tables(0)
        _, back = p.simulate(b, seed=2, want_records=False)
    for r, t, k in zip(ref, tab, back):
        se = np.sqrt(max(r["phiphi"] - r["phi"] ** 2, 0.0) / r["paths"])
        assert abs(r["phi"] - t["phi"]) < 0.05 * se + 1e-5
        assert abs(r["B"] - t["B"]) < 0.05 * se + 1e-5
        assert r == k  # removing the table restores the compiled set bit for bit


def test_refusals(gpu):
    with FKProblem(problem_desc()) as p:  # Franke Poisson: not the tabulated base set
        with pytest.raises(abi.SpacError):
            p.set_coef_tables(abi.SPAC_TAB_SOURCE, np.zeros((4, 4, 8)), DOM)
    with FKProblem(problem_desc(coef=VC, domain=DOM)) as p:
        with pytest.raises(abi.SpacError):
            p.set_coef_tables(128, np.zeros((4, 4, 8)), DOM)
        bad = np.zeros((4, 4, 8))
        bad[1, 2, 4] = np.nan
        with pytest.raises(abi.SpacError) as e:
            p.set_coef_tables(abi.SPAC_TAB_SOURCE, bad, DOM)
        assert e.value.code == abi.SPAC_E_NAN
