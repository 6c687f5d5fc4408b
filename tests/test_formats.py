"""On-disk formats (host only, no GPU): the reference's text files and the per-batch row files.

- solution_GMRES.txt: one value per line, fixed with 12 decimals
  (Solvers/SistemaLinealSolverGMRES_OMP.hpp:606-611);
- times.csv: "phase,rank,seconds" lines (main_test_mpi.cpp:62,73,114-187);
- Escribe_GB_CSR: G_i (first n row pointers), G_j, G_ij (10 decimals), B (10 decimals)
  (Mesh/SistemaLineal.cpp:124-138);
- row files of a node batch (the resume format): a write/read round trip returns every byte,
  and a file for other inputs, a damaged or a truncated one is reported, never read.
"""
import ctypes as C

import numpy as np
import pytest

from spac_amd import abi
from spac_amd.mesh import Mesh, mesh_desc, write_solution, write_times_csv


def test_solution_file_format(tmp_path):
    u = np.array([0.0, 1.0 / 3.0, -2.5, 1234.56789012345678, -1e-13, 7.0])
    p = tmp_path / "solution_GMRES.txt"
    write_solution(p, u)
    lines = p.read_text().splitlines()
    assert lines == [f"{v:.12f}" for v in u]
    assert lines[1] == "0.333333333333" and lines[4] == "-0.000000000000"
    assert np.allclose([float(x) for x in lines], u, rtol=0, atol=5e-13)


def test_times_csv_format_and_append(tmp_path):
    p = tmp_path / "times.csv"
    write_times_csv(p, [("Construye_Cuadrado", 0, 0.4123456789), ("G_and_B_computation", 0, 12.5)])
    write_times_csv(p, [("Solving_LS", 0, 1.25e-3), ("Total", 0, 13.0)], append=True)
    assert p.read_text().splitlines() == ["Construye_Cuadrado,0,0.412346", "G_and_B_computation,0,12.5",
                                          "Solving_LS,0,0.00125", "Total,0,13"]
    write_times_csv(p, [("Total", 3, 2.0)])  # no append: replaced
    assert p.read_text() == "Total,3,2\n"
    with pytest.raises(abi.SpacError):
        write_times_csv(p, [("bad,name", 0, 1.0)])


def test_escribe_gb_csr_files(tmp_path):
    m = Mesh(mesh_desc(n_side=4, nodes_per_circle=24))
    lib = m.lib
    abi.check(lib.spac_mesh_write_csr(m.h, str(tmp_path).encode()), lib)
    Gi, Gj, Gv, B = m.csr()
    n = B.size
    gi = [int(x) for x in (tmp_path / "G_i.txt").read_text().split()]
    gj = [int(x) for x in (tmp_path / "G_j.txt").read_text().split()]
    gv = (tmp_path / "G_ij.txt").read_text().splitlines()
    b = (tmp_path / "B.txt").read_text().splitlines()
    # the reference writes the first n row pointers (its loop runs over rows)
    assert gi == list(Gi[:n]) and gj == list(Gj)
    assert gv == [f"{v:.10f}" for v in Gv] and b == [f"{v:.10f}" for v in B]
    # diagonal = number of containing circles, 1e-23 elsewhere (prints as 0)
    diag = [Gv[k] for r in range(n) for k in range(Gi[r], Gi[r + 1]) if Gj[k] == r]
    assert min(diag) >= 1.0 and "0.0000000000" in gv


def _rows(n, rng, M=5):
    goff = np.arange(n + 1, dtype=np.int64) * M
    rows = dict(B=rng.standard_normal(n), G=rng.standard_normal(n * M), G_col=rng.integers(0, 99, n * M).astype(np.int32),
                G_offset=goff, stats=(abi.PointStats * n)())
    for k in range(n):
        rows["stats"][k].phi = float(k) + 0.25
        rows["stats"][k].paths = 100 + k
        rows["stats"][k].status = 0
    return rows


def _block(rows):
    rb = abi.RowBlock()
    rb.B = abi.ptr(rows["B"], C.c_double)
    rb.G = abi.ptr(rows["G"], C.c_double)
    rb.G_col = abi.ptr(rows["G_col"], C.c_int32)
    rb.G_offset = abi.ptr(rows["G_offset"], C.c_int64)
    rb.stats = rows["stats"]
    return rb


def _read(lib, path, batch, seed, ih, n, entries):
    out = dict(B=np.zeros(n), G=np.zeros(entries), G_col=np.zeros(entries, np.int32),
               G_offset=np.zeros(n + 1, np.int64), stats=(abi.PointStats * n)())
    found = C.c_int32()
    abi.check(lib.spac_rows_file_read(str(path).encode(), batch, seed, ih, n, entries, C.byref(_block(out)),
                                      C.byref(found)), lib)
    return found.value, out


def test_row_file_roundtrip_and_rejections(tmp_path):
    lib = abi.load()
    rng = np.random.default_rng(0)
    n, M = 7, 5
    rows = _rows(n, rng, M)
    p = tmp_path / "rows_000003.bin"
    abi.check(lib.spac_rows_file_write(str(p).encode(), 3, 11, 0xABCDEF, n, C.byref(_block(rows))), lib)
    assert not (tmp_path / "rows_000003.bin.tmp").exists()
    found, out = _read(lib, p, 3, 11, 0xABCDEF, n, n * M)
    assert found == 1
    for k in ("B", "G", "G_col", "G_offset"):
        assert np.array_equal(out[k], rows[k]), k
    assert bytes(out["stats"]) == bytes(rows["stats"])
    # other inputs: batch index, seed, input hash, node count, entries
    for args in ((4, 11, 0xABCDEF, n, n * M), (3, 12, 0xABCDEF, n, n * M), (3, 11, 0xABCDE0, n, n * M),
                 (3, 11, 0xABCDEF, n - 1, (n - 1) * M), (3, 11, 0xABCDEF, n, n * M + 1)):
        assert _read(lib, p, *args)[0] == -1, args
    assert _read(lib, tmp_path / "missing.bin", 3, 11, 0xABCDEF, n, n * M)[0] == 0
    data = bytearray(p.read_bytes())
    bad = tmp_path / "bad.bin"
    data[64 + 200] ^= 0x01  # one payload bit
    bad.write_bytes(bytes(data))
    assert _read(lib, bad, 3, 11, 0xABCDEF, n, n * M)[0] == -1
    bad.write_bytes(p.read_bytes()[:-3])  # truncated
    assert _read(lib, bad, 3, 11, 0xABCDEF, n, n * M)[0] == -1
    bad.write_bytes(p.read_bytes() + b"\0")  # trailing bytes
    assert _read(lib, bad, 3, 11, 0xABCDEF, n, n * M)[0] == -1
    # rows taken from the middle of a larger block: offsets are stored from 0
    big = _rows(12, rng, M)
    sub = dict(B=big["B"][4:9], G=big["G"], G_col=big["G_col"], G_offset=big["G_offset"][4:10].copy(),
               stats=(abi.PointStats * 5).from_buffer_copy(bytes(big["stats"])[4 * C.sizeof(abi.PointStats):
                                                                                9 * C.sizeof(abi.PointStats)]))
    q = tmp_path / "mid.bin"
    abi.check(lib.spac_rows_file_write(str(q).encode(), 0, 1, 2, 5, C.byref(_block(sub))), lib)
    found, out = _read(lib, q, 0, 1, 2, 5, 5 * M)
    assert found == 1 and np.array_equal(out["G_offset"], np.arange(6) * M)
    assert np.array_equal(out["G"], big["G"][4 * M:9 * M]) and np.array_equal(out["B"], big["B"][4:9])
