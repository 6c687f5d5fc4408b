"""Resumable row assembly (spac_fk_assemble_rows_resumable) on the device.

Bar: bit for bit the rows of one spac_fk_assemble_rows over the whole batch, whether every
batch is assembled, every batch is read back, or a mix after files were deleted, damaged or
left over from other inputs (rows depend only on seed, node stream id and path id).
"""
import sys

import numpy as np
import pytest

from spac_amd import abi
from spac_amd.fk import FKProblem, problem_desc
from spac_amd.mesh import Mesh, mesh_desc

pytestmark = pytest.mark.gpu

VC = abi.SPAC_COEF_VARCOEF_FRANKE


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same(a, b):
    for k in ("B", "G", "G_col", "G_offset"):
        x, y = a[k], b[k]
        assert np.array_equal(x.view(np.uint64) if x.dtype.kind == "f" else x,
                              y.view(np.uint64) if y.dtype.kind == "f" else y), k
    for sa, sb in zip(a["stats"], b["stats"]):
        assert sa == sb


@pytest.mark.parametrize("precision", [abi.SPAC_FP64, abi.SPAC_FP32])
def test_resumed_assembly_equals_whole(gpu, tmp_path, precision):
    m = Mesh(mesh_desc(n_side=4, nodes_per_circle=40))
    rng = np.random.default_rng(2)
    n = m.info.n_jobs
    paths = rng.integers(100, 600, n).astype(np.uint32)
    h = 10 ** rng.uniform(-3.5, -3, n)
    say(f"resume: {n} rows, precision {precision}")
    with FKProblem(problem_desc(coef=VC, domain=(0.0, 0.0, 1.0, 1.0), precision=precision)) as p:
        whole = p.assemble_rows(m, paths, h, seed=21)
        first, rep = p.assemble_rows_resumable(m, paths, h, tmp_path, 100, seed=21)
        nb = (n + 99) // 100
        assert nb >= 4
        assert (rep.batches, rep.computed, rep.loaded, rep.rejected) == (nb, nb, 0, 0)
        _same(first, whole)
        again, rep = p.assemble_rows_resumable(m, paths, h, tmp_path, 100, seed=21)
        assert (rep.computed, rep.loaded) == (0, nb)
        _same(again, whole)
        # a run that stopped after batch 2 (later files missing), one damaged file, one stale file
        for b in range(3, nb):
            (tmp_path / f"rows_{b:06d}.bin").unlink()
        f1 = tmp_path / "rows_000001.bin"
        data = bytearray(f1.read_bytes())
        data[-9] ^= 0x40
        f1.write_bytes(bytes(data))
        (tmp_path / "other").mkdir()
        p.assemble_rows_resumable(m, paths, h, tmp_path / "other", 100, seed=22)
        (tmp_path / "rows_000000.bin").write_bytes((tmp_path / "other" / "rows_000000.bin").read_bytes())
        resumed, rep = p.assemble_rows_resumable(m, paths, h, tmp_path, 100, seed=21)
        say(f"resumed: {rep.computed} computed, {rep.loaded} loaded, {rep.rejected} rejected")
        assert rep.loaded == 1 and rep.rejected == 2 and rep.computed == nb - 1
        _same(resumed, whole)
