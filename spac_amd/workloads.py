"""The synthetic workloads of the benchmark and of the full-size parity tests, defined once.

- batched uneven rows: 16 x 16 circles on the unit square (overlap 1.2, 110 nodes per circle,
  ~25k nodes), per-row paths log-uniform in [1e3, 1e5] and h log-uniform in [1e-4, 1e-3],
  drawn from numpy's default_rng(0) in job order;
- full scale: 64 x 64 circles (25 nodes per circle, 99,604 nodes, 186,920 rows), h scaled with
  the circle radius squared so that paths take as many steps as at 4 x 4 circles and h = 5e-4;
- the learned control variate's recipe: plain-MC point estimates at 45^2 jittered grid points
  as training data, full-batch Adam on the MLP, a T x T table from BF16 MFMA inference.

Both meshes use the variable-coefficient operator with the manufactured Franke u
(SPAC_COEF_VARCOEF_FRANKE) and the reference's mesh construction (Mesh/Malla.cpp:12-98).
"""
from __future__ import annotations

import time

import numpy as np

from . import abi
from .fk import FKProblem, problem_desc
from .mesh import Mesh, mesh_desc

UNIT_SQUARE = (0.0, 0.0, 1.0, 1.0)


def batched_uneven():
    """(mesh, paths[n_jobs] uint32, h[n_jobs]) of the batched uneven workload."""
    mesh = Mesh(mesh_desc(n_side=16, nodes_per_circle=110))
    n = mesh.info.n_jobs
    rng = np.random.default_rng(0)
    paths = np.round(10 ** rng.uniform(3, 5, n)).astype(np.uint32)
    h = 10 ** rng.uniform(-4, -3, n)
    return mesh, paths, h


def full_scale():
    """(mesh, h) of the 64 x 64-circle interface system."""
    mesh = Mesh(mesh_desc(n_side=64, nodes_per_circle=25))
    h = 5e-4 * (mesh.info.radius / (1.2 / (3 * np.sqrt(2.0)))) ** 2
    return mesh, h


def training_points(n_side=64, seed=0):
    """n_side^2 jittered points of an n_side x n_side grid on the unit square (the training inputs)."""
    g = (np.arange(n_side) + 0.5) / n_side
    X, Y = np.meshgrid(g, g, indexing="xy")
    rng = np.random.default_rng(seed)
    xs = np.clip(X.ravel() + (rng.random(X.size) - 0.5) / n_side, 1e-3, 1 - 1e-3)
    ys = np.clip(Y.ravel() + (rng.random(X.size) - 0.5) / n_side, 1e-3, 1 - 1e-3)
    return xs, ys


# settings of the learned control variate, from sweeps of the end-to-end time at 64 x 64 circles
# (training time against the row-variance ratio it buys): the gradient-head weight, the training
# set (points x paths at equal data cost: 45^2 x 1000 trains in 15.7 ms per 1000 steps against
# 18.9 ms for 64^2 x 500), the first layer's initial scale (W1, b1 uniform in (-2, 2) reach a
# variance ratio of 80 in 800 steps, where Glorot's W1 and zero b1 reach 57 in 1000), and the
# schedule: 500 steps from 2e-2 to 1e-3 with beta2 = 0.9 reach 70 (62-83 over five
# initialisations) in 7.8 ms, 800 steps to 1e-4 with beta2 = 0.95 reach 76 in 12.5 ms
LEARNED_CV = dict(points_side=45, data_paths=1000, data_h=2e-3, steps=500, lr=(2e-2, 1e-3), beta2=0.9,
                  grad_weight=0.003, table=256, seed=0, first_layer_scale=2.0)


def train_learned_cv(net, points_side=64, data_paths=500, data_h=2e-3, steps=1000, lr=(2e-2, 1e-4), beta2=0.95,
                     grad_weight=0.003, table=256, eager=False, data_problem=None):
    """Train `net` (an mlp.MLP on the unit square) on plain-MC point estimates of the
    variable-coefficient operator (FP32 stepping) and tabulate it.  data_problem: an FP32
    FKProblem of that operator to run the point solves on (one is made when None).  Returns the
    device times of the three parts (ms) and the training report."""
    xs, ys = training_points(points_side)
    own = data_problem is None
    pp = FKProblem(problem_desc(coef=abi.SPAC_COEF_VARCOEF_FRANKE, domain=UNIT_SQUARE,
                                precision=abi.SPAC_FP32)) if own else data_problem
    try:
        pp.set_control_variate(abi.SPAC_CV_NONE)
        st, _ = pp.solve_points(xs, ys, data_paths, data_h, seed=1, as_array=True)
        data_ms = pp.timing().total_ms
    finally:
        if own:
            pp.close()
    u_mc = np.array(st["phi"])
    t0 = time.perf_counter()
    rep = net.train(xs, ys, u_mc, steps=steps, lr=lr[0], lr_final=lr[1], grad_weight=grad_weight, eager=eager,
                    beta2=beta2)
    train_wall = (time.perf_counter() - t0) * 1e3
    table_ms = net.build_table(table)
    return {"training_data_ms": data_ms, "train_ms": rep["train_ms"], "train_wall_ms": train_wall,
            "table_ms": table_ms, "train_loss": rep["final_loss"],
            "training_data": {"points": int(xs.size), "paths": data_paths, "h": data_h}}
