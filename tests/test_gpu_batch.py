"""Batched uneven workloads: per-node paths and h in one persistent launch, the
graph-captured plan and the autotuner.

Bar: the launch shape, the work-item order and graph replay never change a
bit — every variant equals the eager spac_fk_assemble_rows with the same seed,
and the FP64 rows equal the CPU restatement on a subset of nodes.
"""
import sys

import numpy as np
import pytest

import oracle_lib as o
from spac_amd import abi
from spac_amd.fk import FKProblem, RowsPlan, problem_desc
from spac_amd.mesh import Mesh, mesh_desc

pytestmark = pytest.mark.gpu

VC = abi.SPAC_COEF_VARCOEF_FRANKE
DOM = (0.0, 0.0, 1.0, 1.0)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def batch():
    m = Mesh(mesh_desc())
    rng = np.random.default_rng(3)
    n = m.info.n_jobs
    paths = np.round(10 ** rng.uniform(np.log10(200), np.log10(2000), n)).astype(np.uint32)
    h = 10 ** rng.uniform(np.log10(2e-4), -3, n)
    return m, paths, h


def _same_rows(a, b):
    for k in ("B", "G", "G_col", "G_offset"):
        x, y = a[k], b[k]
        if x.dtype.kind == "f":
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), k
        else:
            assert np.array_equal(x, y), k
    for sa, sb in zip(a["stats"], b["stats"]):
        for k in sa:
            if isinstance(sa[k], float):
                assert np.float64(sa[k]).view(np.uint64) == np.float64(sb[k]).view(np.uint64), k
            else:
                assert sa[k] == sb[k], k


def test_uneven_rows_match_oracle(gpu, batch):
    m, paths, h = batch
    say("uneven batch: fp64 rows vs oracle on 48 nodes")
    jobs = np.nonzero(m.job_frontera == 0)[0][::40][:48]
    desc = problem_desc(coef=VC, domain=DOM)
    with FKProblem(desc) as p:
        rows = p.assemble_rows(m, paths[jobs], h[jobs], seed=4, jobs=jobs)
    B, G, st = o.assemble_rows(desc, m, paths[jobs], h[jobs], seed=4, jobs=jobs, nthreads=8)
    assert np.array_equal(rows["B"].view(np.uint64), B.view(np.uint64))
    assert np.array_equal(rows["G"].view(np.uint64), G.view(np.uint64))
    assert [s["path_steps"] for s in rows["stats"]] == [s["path_steps"] for s in st]


@pytest.mark.parametrize("precision", [abi.SPAC_FP64, abi.SPAC_FP32])
def test_plan_replays_equal_eager(gpu, batch, precision):
    m, paths, h = batch
    say(f"plan replays, precision {precision}")
    desc = problem_desc(coef=VC, domain=DOM, precision=precision)
    with FKProblem(desc) as p:
        eager5 = p.assemble_rows(m, paths, h, seed=5)
        eager6 = p.assemble_rows(m, paths, h, seed=6)
        with RowsPlan(p, m, paths, h) as plan:
            for rep in range(3):
                r5 = plan.execute(seed=5)
                _same_rows(r5, eager5)
                # eager work on the parent problem between replays
                p.assemble_rows(m, paths[:100], h[:100], seed=9, jobs=np.arange(100))
                r6 = plan.execute(seed=6)
                _same_rows(r6, eager6)
            say(f"graph replay {plan.last_ms:.2f} ms")


def test_item_order_and_launch_shape_do_not_change_results(gpu, batch):
    m, paths, h = batch
    desc = problem_desc(coef=VC, domain=DOM, precision=abi.SPAC_FP32)
    with FKProblem(desc) as p:
        ref = p.assemble_rows(m, paths, h, seed=7)
        for cfg in ((64, 2, 32, abi.SPAC_ORDER_COST), (128, 16, 1024, abi.SPAC_ORDER_NODE),
                    (256, 4, 100, abi.SPAC_ORDER_COST), (64, 4, 16, abi.SPAC_ORDER_TAIL),
                    (256, 8, 64, abi.SPAC_ORDER_NODE)):
            say(f"launch {cfg}")
            p.set_launch(*cfg)
            _same_rows(p.assemble_rows(m, paths, h, seed=7), ref)


def test_autotuner_installs_a_faster_equal_config(gpu, batch):
    m, paths, h = batch
    desc = problem_desc(coef=VC, domain=DOM, precision=abi.SPAC_FP32)
    with FKProblem(desc) as p:
        ref = p.assemble_rows(m, paths, h, seed=8)
        rep, table = p.autotune_rows(m, paths, h, path_fraction=0.25)
        c = rep.chosen
        say(f"autotune: {rep.n_candidates} candidates, default {rep.default_ms:.3f} ms -> best {rep.best_ms:.3f} ms "
            f"with block {c.block_size} bpc {c.blocks_per_cu} chunk {c.chunk_paths} order {c.item_order}")
        assert rep.n_candidates == len(table) and rep.n_candidates >= 10
        assert rep.best_ms <= rep.default_ms
        assert min(t[-1] for t in table) == pytest.approx(rep.best_ms)
        _same_rows(p.assemble_rows(m, paths, h, seed=8), ref)
        with RowsPlan(p, m, paths, h) as plan:
            _same_rows(plan.execute(seed=8), ref)
