"""Benchmark of the Feynman–Kac hot path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload point|rows]

One step = one pass of the hot path over the whole workload with inputs resident
in HBM.  Prints one JSON line (metric path-steps/s) with a roofline object for
the path kernel (HIP events on the kernel's stream, measured inside the library)
and the CPU restatement timed on the host cores over a bounded sample.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from spac_amd import abi  # noqa: E402
from spac_amd.fk import FKProblem, SolveBatchArrays, problem_desc  # noqa: E402
from spac_amd.flops import FLOPS_PER_STEP, PEAK_FP64_TFLOPS  # noqa: E402
from spac_amd.mesh import DeviceSystem, Mesh, gmres, mesh_desc  # noqa: E402


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# HBM bytes per launch of the path kernel, from two rocprofv3 --pmc passes of this
# script (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) scaled by what the counters
# read for plain device copies of 512 MiB in the same runs (--copy-mib 512):
# FETCH_SIZE x 2.0, WRITE_SIZE x 1.0.  Commands and figures:
# profiles/fk_path_kernel_fp64_franke_points_256x10k.md and
# profiles/pddsparse_rows_fp64_4x4_1952rows_20k.md.
PATH_KERNEL_TRAFFIC_BYTES = {
    "points": (85.5 * 2.0 + 507894.5) * 1024,
    "rows": (7177.5 * 2.0 + 7598712.3) * 1024,
}
RECORD_BYTES_PER_PATH = 6 * 8 + 4 + 1  # x_tau, y_tau, Y, Z, xi, tau; steps; outcome
# The path kernel is compute-bound on FP64 vector FMAs, not on matrix cores: "mfma" names the
# compute roof, whose FP64 figure (78.6 TFLOP/s) is the same for the vector units and dense MFMA.
COMPUTE_UNIT = "VALU FP64 (no matrix-core work on this path; FP64 vector peak = dense FP64 MFMA peak = 78.6 TFLOP/s)"


def config1_points():
    t = -1.0 + 2.0 * (np.arange(16) + 1) / 17
    X, Y = np.meshgrid(t, t, indexing="ij")
    return X.ravel().copy(), Y.ravel().copy()


def cpu_baseline_points(x, y, paths, h, seed, budget_s=20.0):
    """The CPU restatement (OpenMP over paths) on a bounded sample of the workload."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as o

    threads = min(16, os.cpu_count() or 1)
    desc = problem_desc()
    # calibrate on 2 points x 500 paths, then size the sample to ~budget_s
    idx = np.arange(0, x.size, max(1, x.size // 2))[:2]
    b = SolveBatchArrays(x[idx], y[idx], 500, h, idx)
    t0 = time.perf_counter()
    rec = o.simulate(desc, b, seed=seed, nthreads=threads)
    dt = time.perf_counter() - t0
    rate = rec["steps"].astype(np.int64).sum() / dt
    per_point = rate and (rec["steps"].mean() * paths) / rate
    npts = int(max(1, min(x.size, budget_s / max(per_point, 1e-9))))
    idx = np.linspace(0, x.size - 1, npts).astype(np.int64)
    b = SolveBatchArrays(x[idx], y[idx], paths, h, idx)
    t0 = time.perf_counter()
    rec = o.simulate(desc, b, seed=seed, nthreads=threads)
    dt = time.perf_counter() - t0
    steps = int(rec["steps"].astype(np.int64).sum())
    return {"value": steps / dt, "unit": "path-steps/s", "cores": threads, "kind": "port",
            "sample": f"{npts} of the 256 grid points x {paths} paths, h={h} ({steps} path-steps, {dt:.1f} s)"}


def bench_points(args):
    import torch

    x, y = config1_points()
    n, paths, h, seed = x.size, 10000, 1e-3, 0
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(x).to(dev)
    dy = torch.from_numpy(y).to(dev)
    dstats = torch.zeros(n * abi.C.sizeof(abi.PointStats), dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    prob = FKProblem(problem_desc())
    lib = prob.lib

    def step():
        abi.check(lib.spac_fk_solve_points_device(prob.h, dx.data_ptr(), dy.data_ptr(), n, paths, h, seed,
                                                  dstats.data_ptr()), lib)

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    kern_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        kern_ms.append(prob.timing().path_kernel_ms)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {elapsed:.3f} s")
    host = dstats.cpu().numpy().view(np.uint8)
    stats = (abi.PointStats * n).from_buffer_copy(host.tobytes())
    path_steps = int(sum(s.path_steps for s in stats))
    assert all(s.status == 0 for s in stats)
    value = path_steps * args.steps / elapsed
    avg_ms = float(np.mean(kern_ms))
    fps = FLOPS_PER_STEP["franke_point_fp64"]
    achieved = path_steps * fps / (avg_ms * 1e-3) / 1e12
    out = {
        "metric": "path-steps/s",
        "value": value,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain Feynman-Kac FP64, 2D Poisson with Franke data on [-1,1]^2, 16x16 interior grid, "
                               "10000 paths/point, h=1e-3, sigma=sqrt(2) I, Philox seed 0",
                   "points": n, "paths_per_point": paths, "h": h, "path_steps_per_step": path_steps},
        "roofline": {"bound": "mfma", "compute_unit": COMPUTE_UNIT, "achieved": achieved,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": PATH_KERNEL_TRAFFIC_BYTES["points"],
                     "traffic_unit": "B per launch", "algorithmic_bytes": RECORD_BYTES_PER_PATH * n * paths,
                     "kernel": "fk_path_kernel_fp64<CoefFranke, SUB_NONE>", "kernel_avg_ms": avg_ms,
                     "flops_per_path_step": fps},
    }
    prob.close()
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline_points(x, y, paths, h, seed, budget_s=args.cpu_seconds)
    return out


def cpu_baseline_rows(mesh, desc, paths, h, seed, budget_s=20.0):
    """The CPU restatement's row assembly (OpenMP over paths) on a bounded sample of the rows."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as o

    threads = min(16, os.cpu_count() or 1)
    live = np.nonzero(mesh.job_frontera == 0)[0]
    probe = live[:: max(1, live.size // 4)][:4]
    t0 = time.perf_counter()
    _, _, st = o.assemble_rows(desc, mesh, 1000, h, seed=seed, jobs=probe, nthreads=threads)
    dt = time.perf_counter() - t0
    per_job = dt / probe.size * (paths / 1000.0)
    njobs = int(max(1, min(live.size, budget_s / max(per_job, 1e-9))))
    jobs = live[np.linspace(0, live.size - 1, njobs).astype(np.int64)]
    t0 = time.perf_counter()
    _, _, st = o.assemble_rows(desc, mesh, paths, h, seed=seed, jobs=jobs, nthreads=threads)
    dt = time.perf_counter() - t0
    steps = int(sum(x["path_steps"] for x in st))
    return {"value": steps / dt, "unit": "path-steps/s", "cores": threads, "kind": "port",
            "sample": f"{njobs} of the {mesh.info.n_jobs} rows x {paths} paths, h={h}, G and B included "
                      f"({steps} path-steps, {dt:.1f} s)"}


def bench_rows(args):
    """PDDSparse interface assembly: 4x4 circles on the unit square, 1016 nodes, 1952 rows."""
    paths, h, seed = 20000, 5e-4, 0
    mesh = Mesh(mesh_desc())
    desc = problem_desc(coef=abi.SPAC_COEF_VARCOEF_FRANKE, domain=(0.0, 0.0, 1.0, 1.0))
    prob = FKProblem(desc)
    for i in range(args.warmup):
        prob.assemble_rows(mesh, paths, h, seed=seed)
        log(f"warmup {i + 1}/{args.warmup} done")
    dev_ms, path_ms, wall = [], [], []
    rows = None
    for i in range(args.steps):
        t0 = time.perf_counter()
        rows = prob.assemble_rows(mesh, paths, h, seed=seed)
        wall.append(time.perf_counter() - t0)
        t = prob.timing()
        dev_ms.append(t.total_ms)
        path_ms.append(t.path_kernel_ms)
    log(f"timed {args.steps} steps")
    path_steps = int(sum(x["path_steps"] for x in rows["stats"]))
    n_paths = int(sum(x["paths"] for x in rows["stats"]))
    mesh.accumulate(rows)
    Gi, Gj, Gv, B = mesh.csr()
    t0 = time.perf_counter()
    u, info = gmres(Gi, Gj, Gv, B, tol=1e-10, max_iter=5000, restart=200)
    t_gmres = time.perf_counter() - t0
    from spac_amd.problems import franke_u

    err = u - franke_u(mesh.node_x, mesh.node_y)
    dev_s = float(np.mean(dev_ms)) * 1e-3
    avg_path_ms = float(np.mean(path_ms))
    fps = FLOPS_PER_STEP["varcoef_circle_fp64"]
    achieved = path_steps * fps / (avg_path_ms * 1e-3) / 1e12
    out = {
        "metric": "path-steps/s",
        "value": path_steps / dev_s,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dev_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "PDDSparse interface assembly FP64: unit square, 4x4 overlapping circles (overlap 1.2), "
                               "1016 nodes / 1952 rows, 20000 paths/row, h=5e-4, sigma=(1+0.5 sin(pi x) cos(pi y)) I, "
                               "b=0.3(y,-x), c=-1-x^2, manufactured Franke u, Gaussian RBF eps=3 with 15-node stencils",
                   "rows": int(mesh.info.n_jobs), "nodes": int(mesh.info.n_nodes), "paths_per_row": paths, "h": h,
                   "path_steps_per_step": path_steps},
        "roofline": {"bound": "mfma", "compute_unit": COMPUTE_UNIT, "achieved": achieved,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": PATH_KERNEL_TRAFFIC_BYTES["rows"],
                     "traffic_unit": "B per launch", "algorithmic_bytes": RECORD_BYTES_PER_PATH * n_paths,
                     "kernel": "fk_path_kernel_fp64<CoefVarFranke, SUB_CIRCLE>", "kernel_avg_ms": avg_path_ms,
                     "flops_per_path_step": fps},
        "pcie_inclusive_path_steps_per_s": path_steps / float(np.mean(wall)),
        "gmres": {"tol": 1e-10, "iterations": int(info.iterations), "residual": float(info.residual),
                  "seconds_host": t_gmres, "rms_err_vs_exact": float(np.sqrt(np.mean(err ** 2))),
                  "max_err_vs_exact": float(np.abs(err).max())},
    }
    prob.close()
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline_rows(mesh, desc, paths, h, seed, budget_s=args.cpu_seconds)
    return out


def bench_batched(steps=2, seed=0, fraction=1.0 / 32):
    """Uneven batch on one GPU: 16x16 circles (Nt = 110, ~25k nodes), per-row paths log-uniform in
    [1e3, 1e5] and h log-uniform in [1e-4, 1e-3] (numpy default_rng(0)), FP32 stepping; the launch
    shape and refill order come from the autotuner (on the batch with paths x fraction), then the
    assembly runs as one captured graph (plan) per step."""
    from spac_amd.fk import RowsPlan

    mesh = Mesh(mesh_desc(n_side=16, nodes_per_circle=110))
    n = mesh.info.n_jobs
    rng = np.random.default_rng(0)
    paths = np.round(10 ** rng.uniform(3, 5, n)).astype(np.uint32)
    h = 10 ** rng.uniform(-4, -3, n)
    live = mesh.job_frontera == 0
    desc = problem_desc(coef=abi.SPAC_COEF_VARCOEF_FRANKE, domain=(0.0, 0.0, 1.0, 1.0), precision=abi.SPAC_FP32)
    out = {"workload": "16x16 circles (overlap 1.2, 110 nodes per circle), per-row paths log-uniform [1e3, 1e5], "
                       "h log-uniform [1e-4, 1e-3], FP32 stepping, FP64 compensated sums",
           "rows": int(n), "nodes": int(mesh.info.n_nodes), "paths_total": int(paths[live].astype(np.int64).sum())}
    with FKProblem(desc) as prob:
        t0 = time.perf_counter()
        rep, table = prob.autotune_rows(mesh, paths, h, path_fraction=fraction)
        c = rep.chosen
        out["autotune"] = {"seconds": time.perf_counter() - t0, "candidates": int(rep.n_candidates),
                           "chosen": {"block_size": c.block_size, "blocks_per_cu": c.blocks_per_cu,
                                      "chunk_paths": c.chunk_paths, "item_order": c.item_order},
                           "sample_fraction": fraction, "default_ms": rep.default_ms, "best_ms": rep.best_ms}
        log(f"batched: autotune {out['autotune']}")
        with RowsPlan(prob, mesh, paths, h) as plan:
            plan.execute(seed=seed)  # warm-up
            ms, steps_total = [], 0
            for i in range(steps):
                rows = plan.execute(seed=seed + 1 + i)
                ms.append(plan.last_ms)
                steps_total = int(sum(x["path_steps"] for x in rows["stats"]))
        out["graph_ms"] = float(np.mean(ms))
        out["path_steps"] = steps_total
        out["path_steps_per_s"] = steps_total / (np.mean(ms) * 1e-3)
    log(f"batched: {out}")
    mesh.close()
    return out


def _ttr_mesh():
    """Config-5 mesh: 64x64 circles on the unit square, 25 nodes per circle (~100k nodes);
    h scaled with the circle radius squared so paths take as many steps as at 4x4 circles."""
    mesh = Mesh(mesh_desc(n_side=64, nodes_per_circle=25))
    h = 5e-4 * (mesh.info.radius / (1.2 / (3 * np.sqrt(2.0)))) ** 2
    return mesh, h


# GMRES relative-residual tolerance of the time-to-RMS solves: the reference's default
# toleranciaGMRES (main_test_mpi.cpp:32).
GMRES_TOL = 1e-8
# Coarse grid of the two-level GMRES preconditioner in the time-to-RMS solves (aggregates per
# side); 0 gives the Jacobi-only solver.  At 64x64 circles: 309 -> 67 iterations from a cold start.
COARSE_GRID = 12


def _replicas(mesh, system, prob, paths, h, seeds, u0=None):
    """Independent assemblies (one per seed) at `paths` per row into the device-resident system,
    each solved there by GPU GMRES.

    Returns per-replica timings and two error measures: the statistical RMS of the solved
    node values (sample variance across replicas, averaged over nodes) and the RMS of the
    row residuals G_k u_bar - B_k around the replica-mean solution (the per-row Monte-Carlo
    noise, independent across rows).  The system is copied out for these statistics outside
    the timed calls."""
    import scipy.sparse as sp

    out, us, mats = [], [], []
    n = mesh.info.n_nodes
    for k, s in enumerate(seeds):
        t0 = time.perf_counter()
        # the estimator records (path steps, fallbacks) of the set's first replica only
        st = prob.assemble_system(mesh, system, paths, h, seed=s, stats=k == 0)
        t1 = time.perf_counter()
        t = prob.timing()
        u, info = system.solve(tol=GMRES_TOL, max_iter=20000, restart=200, u0=u0)
        t2 = time.perf_counter()
        Gi, Gj, Gv, B = system.read()
        u0 = u
        us.append(u)
        mats.append((sp.csr_matrix((Gv, Gj, Gi), shape=(n, n)), B))
        out.append(dict(ms=t.total_ms, path_ms=t.path_kernel_ms,
                        steps=int(st["path_steps"].sum()) if st is not None else None,
                        fallbacks=int(st["cv_fallback"].sum()) if st is not None else None,
                        gmres_ms=info.seconds * 1e3,
                        gmres_device_ms=info.device_ms, gmres_it=int(info.iterations), converged=bool(info.converged),
                        assemble_wall_ms=(t1 - t0) * 1e3, wall_ms=(t2 - t0) * 1e3))
    U = np.array(us)
    ubar = U.mean(0)
    rms_u = float(np.sqrt(U.var(0, ddof=1).mean()))
    R = np.array([A @ ubar - B for A, B in mats])
    rms_row = float(np.sqrt(R.var(0, ddof=1).mean()))
    return out, rms_u, rms_row, ubar


def bench_ttr(target=1e-3, probe_paths=256, train_steps=1000, lr=(2e-2, 1e-4), n_rep=8, seed0=100, eager=False,
              beta2=0.95, data_paths=500, data_h=2e-3, max_refine=3, refine_rep=24, grad_weight=0.003):
    """Time to a statistical RMS error `target` of the solved interface values, learned control
    variate against plain Monte Carlo, both FP32 stepping on this GPU.

    Statistical RMS at a path count: independent replicas (seeds), sample variance per node,
    averaged over nodes.  The node errors are dominated by a few smooth modes of the solve, so
    an RMS from R replicas scatters by roughly +-(1 / R)^(1/2); no single measurement decides
    the path count.  Instead: a probe at probe_paths (n_rep replicas), then max_refine sets of
    refine_rep replicas, each at the path count where RMS^2 N = a + b / N, fitted by weighted
    least squares over every set so far, gives RMS = 0.9 x target (b: the higher-order term the
    noise of the assembled matrix adds at small N).  The last set is the timed one.  `rms` is the
    fit over every set (the last included) at the timed count; the last set's own measurement is
    beside it.  The same rule serves both methods.  The learned-CV time
    includes its training data (plain-MC point solves at 4096 points, data_paths paths each at
    step data_h: the control variate needs grad u, not an unbiased u), training and
    tabulation; GMRES (GPU, FP64, relative residual GMRES_TOL) is reported beside the
    Monte-Carlo time."""
    from spac_amd.mlp import MLP
    from spac_amd.problems import franke_u

    mesh, h = _ttr_mesh()
    dom = (0.0, 0.0, 1.0, 1.0)
    desc = problem_desc(coef=abi.SPAC_COEF_VARCOEF_FRANKE, domain=dom, precision=abi.SPAC_FP32)
    exact = franke_u(mesh.node_x, mesh.node_y)
    res = {"target_rms": target, "replicas": {"probe": n_rep, "refinement": refine_rep},
           "mesh": {"nodes": int(mesh.info.n_nodes), "rows": int(mesh.info.n_jobs),
                    "circles": int(mesh.info.n_interfaces), "h": h, "nnz": int(mesh.info.nnz)}}

    aim = 0.9 * target

    def fit_paths(sets):
        """Path count at which RMS reaches `aim` under RMS^2 N = a + b / N, fitted by least
        squares over every replica set so far (weights = replicas); b >= 0."""
        N = np.array([x[0] for x in sets], float)
        y = np.array([x[1] ** 2 * x[0] for x in sets])
        w = np.array([x[2] for x in sets], float)
        A = np.stack([np.ones_like(N), 1.0 / N], 1) * np.sqrt(w)[:, None]
        a, b = np.linalg.lstsq(A, y * np.sqrt(w), rcond=None)[0] if len(sets) > 1 else (y[0], 0.0)
        if b < 0 or a <= 0:
            a, b = float(np.sum(w * y) / np.sum(w)), 0.0
        T2 = aim ** 2
        return int(np.ceil((a + np.sqrt(a * a + 4.0 * b * T2)) / (2.0 * T2))), float(a), float(b)

    def measure(prob, name, extra):
        paths, hist, sets = probe_paths, [], []
        for it in range(max_refine + 1):
            reps = n_rep if it == 0 else refine_rep
            runs, rms, row, ubar = _replicas(mesh, system, prob, paths, h,
                                             [seed0 + 1000 * it + k for k in range(reps)])
            hist.append({"paths_per_row": paths, "rms": rms, "row_residual_rms": row, "replicas": reps})
            sets.append((paths, rms, reps))
            if it == 0:
                rms0, row0 = rms, row
            if it == max_refine:
                break
            paths, fa, fb = fit_paths(sets)
            log(f"ttr {name}: rms {rms:.3e} (row {row:.3e}) at {hist[-1]['paths_per_row']} paths; "
                f"fit a {fa:.3e} b {fb:.3e} -> {paths} paths")
        # the RMS at the timed count from the fit over every set (the last included): a pooled
        # estimate, steadier than the last set's own measurement
        _, fa, fb = fit_paths(sets)
        rms_fit = float(np.sqrt(fa / paths + fb / paths ** 2))
        # the count where the fit reaches the target itself (both methods compared at equal RMS)
        at_target = float((fa + np.sqrt(fa * fa + 4.0 * fb * target ** 2)) / (2.0 * target ** 2))
        r = {"paths_per_row": paths, "rms": rms_fit, "rms_last_set": rms, "reached_target": bool(rms_fit <= target),
             "row_residual_rms": row, "fit": {"a": fa, "b": fb, "aim_rms": aim,
                                             "replicas": int(sum(x[2] for x in sets))},
             "probe_rms": rms0, "probe_row_residual_rms": row0, "path_count_history": hist,
             "mc_ms": float(np.mean([x["ms"] for x in runs])),
             "paths_at_target": at_target,
             "path_kernel_ms": float(np.mean([x["path_ms"] for x in runs])),
             "gmres_ms": float(np.mean([x["gmres_ms"] for x in runs])),
             "gmres_device_ms": float(np.mean([x["gmres_device_ms"] for x in runs])),
             "assemble_wall_ms": float(np.mean([x["assemble_wall_ms"] for x in runs])),
             "wall_ms": float(np.mean([x["wall_ms"] for x in runs])),
             "gmres_iterations": runs[0]["gmres_it"], "path_steps": runs[0]["steps"],
             "fallback_rows": runs[0]["fallbacks"], "rms_vs_exact": float(np.sqrt(np.mean((ubar - exact) ** 2)))}
        r.update(extra)
        log(f"ttr {name}: {r}")
        return r

    system = DeviceSystem(mesh)
    # two-level GMRES: 12 x 12 aggregates of rows (~5 x 5 circles each) as coarse space, both methods
    system.set_coarse(COARSE_GRID, COARSE_GRID)
    with FKProblem(desc) as prob:
        prob.set_control_variate(abi.SPAC_CV_NONE)
        prob.assemble_system(mesh, system, 16, h, seed=1)  # warm-up
        res["plain"] = measure(prob, "plain", {})
    # learned control variate: training data, training, tabulation
    g = (np.arange(64) + 0.5) / 64
    X, Y = np.meshgrid(g, g, indexing="xy")
    rng = np.random.default_rng(0)
    xs = np.clip(X.ravel() + (rng.random(4096) - 0.5) / 64, 1e-3, 1 - 1e-3)
    ys = np.clip(Y.ravel() + (rng.random(4096) - 0.5) / 64, 1e-3, 1 - 1e-3)
    with FKProblem(problem_desc(coef=abi.SPAC_COEF_VARCOEF_FRANKE, domain=dom, precision=abi.SPAC_FP32)) as pp:
        pp.set_control_variate(abi.SPAC_CV_NONE)
        st, _ = pp.solve_points(xs, ys, data_paths, data_h, seed=1)
        data_ms = pp.timing().total_ms
    u_mc = np.array([x["phi"] for x in st])
    with MLP(dom, seed=0) as net, FKProblem(desc) as prob:
        rep = net.train(xs, ys, u_mc, steps=train_steps, lr=lr[0], lr_final=lr[1], grad_weight=grad_weight,
                        eager=eager, beta2=beta2)
        table_ms = net.build_table(256)
        net.attach(prob, fallback=True)
        res["learned"] = measure(prob, "learned", {"training_data_ms": data_ms, "training_data": {
                                                       "points": int(xs.size), "paths": data_paths, "h": data_h},
                                                   "train_ms": rep["train_ms"],
                                                   "train_steps": train_steps, "train_lr": list(lr), "train_beta2": beta2 or 0.999,
                                                   "train_grad_weight": grad_weight,
                                                   "train_loss": rep["final_loss"],
                                                   "table_ms": table_ms})
    P, L = res["plain"], res["learned"]
    for M in (P, L):
        # Monte-Carlo time projected linearly in N to the count where the fitted RMS is the target
        M["mc_ms_at_target"] = M["mc_ms"] * M["paths_at_target"] / M["paths_per_row"]
    setup = L["training_data_ms"] + L["train_ms"] + L["table_ms"]
    res["time_to_rms_ms"] = {"plain_mc": P["mc_ms"], "learned_mc": L["mc_ms"], "learned_setup": setup,
                             "plain_with_gmres": P["mc_ms"] + P["gmres_ms"],
                             "learned_with_setup_and_gmres": setup + L["mc_ms"] + L["gmres_ms"],
                             "plain_wall": P["wall_ms"], "learned_wall_with_setup": setup + L["wall_ms"]}
    res["row_variance_ratio_at_probe"] = (P["probe_row_residual_rms"] / L["probe_row_residual_rms"]) ** 2
    res["speedup"] = {"mc_only": P["mc_ms"] / L["mc_ms"],
                      "including_training": P["mc_ms"] / (setup + L["mc_ms"]),
                      "end_to_end_with_gmres": (P["mc_ms"] + P["gmres_ms"]) / (setup + L["mc_ms"] + L["gmres_ms"]),
                      "end_to_end_wall": P["wall_ms"] / (setup + L["wall_ms"])}
    Pt, Lt = P["mc_ms_at_target"], L["mc_ms_at_target"]
    res["speedup_at_target"] = {"mc_only": Pt / Lt, "including_training": Pt / (setup + Lt),
                                "end_to_end_with_gmres": (Pt + P["gmres_ms"]) / (setup + Lt + L["gmres_ms"])}
    system.close()
    mesh.close()
    return res


def bench_train(steps=1000, lr=(2e-2, 1e-4), beta2=0.95, eager=False):
    """The learned control variate's training alone (for kernel profiles): 4096 jittered grid
    points on the unit square, Franke values plus N(0, 0.01^2) noise as targets, `steps`
    full-batch Adam steps after 20 warm-up steps."""
    from spac_amd.mlp import MLP
    from spac_amd.problems import franke_u

    g = (np.arange(64) + 0.5) / 64
    X, Y = np.meshgrid(g, g, indexing="xy")
    rng = np.random.default_rng(0)
    xs = np.clip(X.ravel() + (rng.random(4096) - 0.5) / 64, 1e-3, 1 - 1e-3)
    ys = np.clip(Y.ravel() + (rng.random(4096) - 0.5) / 64, 1e-3, 1 - 1e-3)
    u = franke_u(xs, ys) + 0.01 * rng.standard_normal(xs.size)
    with MLP((0.0, 0.0, 1.0, 1.0), seed=0) as net:
        net.train(xs, ys, u, steps=20, lr=lr[0], lr_final=lr[1], grad_weight=0.003, eager=eager, beta2=beta2)
        rep = net.train(xs, ys, u, steps=steps, lr=lr[0], lr_final=lr[1], grad_weight=0.003, eager=eager,
                        beta2=beta2)
    return {"samples": int(xs.size), "steps": steps, "eager": bool(eager), "train_ms": rep["train_ms"],
            "us_per_step": rep["train_ms"] * 1e3 / steps, "final_loss": rep["final_loss"]}


def copy_calibration(mib: int, reps: int = 3) -> None:
    """Plain device copies of known size (torch's copy kernel, `mib` MiB read and
    written per launch): the reference dispatches that scale rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy calibration: {reps} copies of {mib} MiB")
    del a, b


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rows", choices=["rows", "points", "batched", "ttr", "train"],
                    help="rows (default line) | points | batched, ttr or train alone (for profiling)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-ttr", action="store_true", help="skip the time-to-RMS comparison (config 5)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched uneven workload (config 3)")
    ap.add_argument("--train-steps", type=int, default=1000)
    ap.add_argument("--train-lr", type=float, nargs=2, default=(2e-2, 1e-4), metavar=("LR0", "LR1"))
    ap.add_argument("--train-beta2", type=float, default=0.95, help="Adam beta2 (0: 0.999)")
    ap.add_argument("--train-grad-weight", type=float, default=0.003,
                    help="loss weight of the gradient heads (lambda)")
    ap.add_argument("--ttr-replicas", type=int, default=8, help="replicas of the probe")
    ap.add_argument("--ttr-refine-replicas", type=int, default=24, help="replicas of each path-count refinement")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run plain device copies of this many MiB (counter calibration for profiles)")
    ap.add_argument("--train-eager", action="store_true",
                    help="launch training steps one by one instead of replaying captured graphs (for kernel profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path shards by node batches "
              "(rows of one circle stay together), one GPU per share", file=sys.stderr)
        sys.exit(2)
    if args.copy_mib > 0:
        copy_calibration(args.copy_mib)
    if args.workload == "batched":
        print(json.dumps({"batched": bench_batched(steps=args.steps)}), flush=True)
        return
    if args.workload == "train":
        print(json.dumps({"train": bench_train(steps=args.train_steps, lr=tuple(args.train_lr), beta2=args.train_beta2,
                                               eager=args.train_eager)}), flush=True)
        return
    if args.workload == "ttr":
        print(json.dumps({"time_to_rms": bench_ttr(train_steps=args.train_steps, n_rep=args.ttr_replicas, refine_rep=args.ttr_refine_replicas,
                                     eager=args.train_eager,
                                     lr=tuple(args.train_lr), beta2=args.train_beta2,
                                     grad_weight=args.train_grad_weight)}), flush=True)
        return
    out = bench_rows(args) if args.workload == "rows" else bench_points(args)
    if not args.no_batched:
        log("batched uneven workload (config 3) ...")
        out["batched"] = bench_batched()
    if not args.no_ttr:
        log("time to RMS 1e-3 (config 5) ...")
        out["time_to_rms"] = bench_ttr(train_steps=args.train_steps, n_rep=args.ttr_replicas, refine_rep=args.ttr_refine_replicas,
                                     eager=args.train_eager,
                                     lr=tuple(args.train_lr), beta2=args.train_beta2,
                                     grad_weight=args.train_grad_weight)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
