"""Writes tests/golden/gmres_system.json: a small assembled interface system and
its direct solution by scipy.sparse.linalg.spsolve.

System: unit square, 3x3 circles, overlap 1.2, 24 nodes per circle, Gaussian
15-point stencils (eps 3), every row from the CPU restatement with 400 paths,
h = 1e-3, variable-coefficient Franke operator, seed 0.
"""
import json
import sys
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))

import oracle_lib as o  # noqa: E402
from spac_amd import abi  # noqa: E402
from spac_amd.fk import problem_desc  # noqa: E402
from spac_amd.mesh import Mesh, mesh_desc  # noqa: E402


def main():
    m = Mesh(mesh_desc(n_side=3, nodes_per_circle=24))
    desc = problem_desc(coef=abi.SPAC_COEF_VARCOEF_FRANKE, domain=(0.0, 0.0, 1.0, 1.0))
    B, G, st = o.assemble_rows(desc, m, 400, 1e-3, seed=0)
    rows = m.alloc_rows()
    rows["B"][:] = B
    rows["G"][:] = G
    a = m.iface_arrays()
    off = 0
    for k in range(m.info.n_jobs):
        f = m.job_iface[k]
        cols = a["node_global"][a["node_offset"][f]:a["node_offset"][f + 1]]
        rows["G_offset"][k] = off
        rows["G_col"][off:off + cols.size] = cols
        off += cols.size
    rows["G_offset"][-1] = off
    m.accumulate(rows)
    Gi, Gj, Gv, Bv = m.csr()
    A = sp.csr_matrix((Gv, Gj, Gi), shape=(Bv.size, Bv.size))
    u = spl.spsolve(A.tocsc(), Bv)
    out = dict(description=__doc__.strip(), G_i=Gi.tolist(), G_j=Gj.tolist(), G_ij=Gv.tolist(), B=Bv.tolist(),
               u_spsolve=u.tolist())
    (HERE / "gmres_system.json").write_text(json.dumps(out))
    print("rows", Bv.size, "nnz", Gi[-1])


if __name__ == "__main__":
    main()
