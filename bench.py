"""Benchmark of the Feynman–Kac hot path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ttr|rows|points|batched|train]

Default (--workload ttr): the headline line, time to a statistical RMS error of 1e-3 of the
solved interface values on the full-scale system (64 x 64 circles, 99,604 nodes, 186,920
rows), the learned control variate end to end against plain Monte Carlo on the same GPU.
One step = one learned-CV solve from scratch: training data, training, table, assembly of
every row at the calibrated path count into the device-resident system, and GMRES from the
network's prediction at the nodes (plain MC: from zero).  Inputs (mesh, interface tables, job
arrays) are resident on the device before
the timed region.  Beside it: plain MC at its own calibrated count, bootstrap confidence
intervals of the speed-ups, path-steps/s, the FP32-against-FP64 difference, the roofline of
the dominant kernel (HIP events on its stream, measured inside the library), the PDDSparse
rows and batched workloads, and the CPU restatement timed on the host cores over a bounded
sample.
"""
import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from spac_amd import abi  # noqa: E402
from spac_amd.fk import FKProblem, SolveBatchArrays, problem_desc  # noqa: E402
from spac_amd.flops import FLOPS_PER_STEP, PEAK_FP32_TFLOPS, PEAK_FP64_TFLOPS  # noqa: E402
from spac_amd.mesh import DeviceSystem, Mesh, gmres, mesh_desc  # noqa: E402
from spac_amd.workloads import LEARNED_CV, UNIT_SQUARE, batched_uneven, full_scale, train_learned_cv  # noqa: E402

VC = abi.SPAC_COEF_VARCOEF_FRANKE


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# HBM bytes per launch of the path kernels, from two rocprofv3 --pmc passes of this script
# (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) scaled by what the counters read for plain
# device copies of 512 MiB in the same runs (--copy-mib 512): FETCH_SIZE x 2.0, WRITE_SIZE x 1.0.
# Commands and figures: profiles/fk_path_kernel_fp64_franke_points_256x10k.md,
# profiles/pddsparse_rows_fp64_4x4_1952rows_20k.md, profiles/learned_cv_time_to_rms_64x64_fp32.md.
PATH_KERNEL_TRAFFIC_BYTES = {
    "points": (92.7 * 2.0 + 564166.6) * 1024,
    "rows": (24920.6 * 2.0 + 7580208.8) * 1024,
}
# The learned-CV path kernel of the full-scale system, per path of the launch: measured at 94
# paths per row (17,570,480 paths; FETCH_SIZE 240,166.6 KiB x 2.0, WRITE_SIZE 4,360,928.6 KiB
# x 1.0, mean of two dispatches of `--workload assemble --paths-per-row 94`); its launches in
# the bench run at the calibrated count, so the bytes per launch are this figure times the
# launch's paths.
LEARNED_PATH_KERNEL_BYTES_PER_PATH = (240166.6 * 2.0 + 4360928.6) * 1024 / (186920 * 94)
RECORD_BYTES_PER_PATH = 6 * 8 + 4 + 1  # x_tau, y_tau, Y, Z, xi, tau; steps; outcome
COMPUTE_UNIT_64 = "VALU FP64 (no matrix-core work on this path; FP64 vector peak = dense FP64 MFMA peak = 78.6 TFLOP/s)"
COMPUTE_UNIT_32 = "VALU FP32 (no matrix-core work in the path kernel; MI355X FP32 vector peak 157.3 TFLOP/s)"


# ---------------------------------------------------------------- host --
def cpu_threads():
    """Threads for the CPU restatement: every CPU this process may run on (nproc), capped by
    OMP_NUM_THREADS where the environment sets it (the GPU box grants each GPU a CPU share and
    sets it)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def host_info():
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"cpu_model": model, "nproc": nproc, "os_cpu_count": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


# ------------------------------------------------------ config 1 points --
def config1_points():
    t = -1.0 + 2.0 * (np.arange(16) + 1) / 17
    X, Y = np.meshgrid(t, t, indexing="ij")
    return X.ravel().copy(), Y.ravel().copy()


def cpu_baseline_points(x, y, paths, h, seed, budget_s=20.0):
    """The CPU restatement (OpenMP over paths) on a bounded sample of the workload."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as o

    threads = cpu_threads()
    desc = problem_desc()
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
    return {"value": steps / dt, "unit": "path-steps/s", "cores": threads, "kind": "port", **host_info(),
            "sample": f"{npts} of the 256 grid points x {paths} paths, h={h} ({steps} path-steps, {dt:.1f} s)"}


def bench_points(args):
    """Config 1 on device-resident inputs (spac_fk_solve_points_device)."""
    x, y = config1_points()
    n, paths, h, seed = x.size, 10000, 1e-3, 0
    prob = FKProblem(problem_desc())
    lib = prob.lib
    dx, dy = abi.DeviceArray.from_numpy(x, lib), abi.DeviceArray.from_numpy(y, lib)
    dstats = abi.DeviceArray(n * abi.C.sizeof(abi.PointStats), lib)

    def step():
        abi.check(lib.spac_fk_solve_points_device(prob.h, dx.ptr, dy.ptr, n, paths, h, seed, dstats.ptr), lib)

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup} done")
    kern_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()  # synchronous
        kern_ms.append(prob.timing().path_kernel_ms)
    elapsed = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {elapsed:.3f} s")
    stats = (abi.PointStats * n).from_buffer_copy(dstats.to_bytes())
    path_steps = int(sum(s.path_steps for s in stats))
    assert all(s.status == 0 for s in stats)
    avg_ms = float(np.mean(kern_ms))
    fps = FLOPS_PER_STEP["franke_point_fp64"]
    achieved = path_steps * fps / (avg_ms * 1e-3) / 1e12
    out = {
        "metric": "path-steps/s", "value": path_steps * args.steps / elapsed, "unit": "path-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "plain Feynman-Kac FP64, 2D Poisson with Franke data on [-1,1]^2, 16x16 interior grid, "
                               "10000 paths/point, h=1e-3, sigma=sqrt(2) I, Philox seed 0",
                   "points": n, "paths_per_point": paths, "h": h, "path_steps_per_step": path_steps},
        "roofline": {"bound": "mfma", "compute_unit": COMPUTE_UNIT_64, "achieved": achieved,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": PATH_KERNEL_TRAFFIC_BYTES["points"], "traffic_unit": "B per launch",
                     "algorithmic_bytes": RECORD_BYTES_PER_PATH * n * paths,
                     "kernel": "fk_path_kernel_fp64<CoefFranke, SUB_NONE>", "kernel_avg_ms": avg_ms,
                     "flops_per_path_step": fps},
    }
    for d in (dx, dy, dstats):
        d.free()
    prob.close()
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline_points(x, y, paths, h, seed, budget_s=args.cpu_seconds)
    return out


# -------------------------------------------------------- config 2 rows --
def cpu_baseline_rows(mesh, desc, paths, h, seed, budget_s=20.0, cv=None):
    """The CPU restatement's row assembly (OpenMP over paths) on a bounded sample of the rows;
    cv: oracle control-variate mode (None: the plugin F, the reference's behaviour)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as o

    threads = cpu_threads()
    if cv is not None:
        o.set_cv(cv)
    try:
        live = np.nonzero(mesh.job_frontera == 0)[0]
        probe = live[:: max(1, live.size // 4)][:4]
        probe_paths = min(1000, paths)
        t0 = time.perf_counter()
        o.assemble_rows(desc, mesh, probe_paths, h, seed=seed, jobs=probe, nthreads=threads)
        dt = time.perf_counter() - t0
        per_job = dt / probe.size * (paths / probe_paths)
        njobs = int(max(1, min(live.size, budget_s / max(per_job, 1e-9))))
        jobs = live[np.linspace(0, live.size - 1, njobs).astype(np.int64)]
        t0 = time.perf_counter()
        _, _, st = o.assemble_rows(desc, mesh, paths, h, seed=seed, jobs=jobs, nthreads=threads)
        dt = time.perf_counter() - t0
    finally:
        if cv is not None:
            o.set_cv()
    steps = int(sum(x["path_steps"] for x in st))
    return {"value": steps / dt, "unit": "path-steps/s", "cores": threads, "kind": "port", **host_info(),
            "sample": f"{njobs} of the {mesh.info.n_jobs} rows x {paths} paths, h={h:.4g}, G and B included "
                      f"({steps} path-steps, {dt:.1f} s)"}


def bench_rows(args, steps=None, warmup=None, with_cpu=True):
    """PDDSparse interface assembly FP64: 4x4 circles on the unit square, 1016 nodes, 1952 rows."""
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    paths, h, seed = 20000, 5e-4, 0
    mesh = Mesh(mesh_desc())
    desc = problem_desc(coef=VC, domain=UNIT_SQUARE)
    prob = FKProblem(desc)
    for i in range(warmup):
        prob.assemble_rows(mesh, paths, h, seed=seed)
    dev_ms, path_ms, wall = [], [], []
    rows = None
    for i in range(steps):
        t0 = time.perf_counter()
        rows = prob.assemble_rows(mesh, paths, h, seed=seed)
        wall.append(time.perf_counter() - t0)
        t = prob.timing()
        dev_ms.append(t.total_ms)
        path_ms.append(t.path_kernel_ms)
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
        "metric": "path-steps/s", "value": path_steps / dev_s, "unit": "path-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": dev_s * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "PDDSparse interface assembly FP64: unit square, 4x4 overlapping circles (overlap 1.2), "
                               "1016 nodes / 1952 rows, 20000 paths/row, h=5e-4, sigma=(1+0.5 sin(pi x) cos(pi y)) I, "
                               "b=0.3(y,-x), c=-1-x^2, manufactured Franke u, Gaussian RBF eps=3 with 15-node stencils",
                   "rows": int(mesh.info.n_jobs), "nodes": int(mesh.info.n_nodes), "paths_per_row": paths, "h": h,
                   "path_steps_per_step": path_steps},
        "roofline": {"bound": "mfma", "compute_unit": COMPUTE_UNIT_64, "achieved": achieved,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": PATH_KERNEL_TRAFFIC_BYTES["rows"], "traffic_unit": "B per launch",
                     "algorithmic_bytes": RECORD_BYTES_PER_PATH * n_paths,
                     "kernel": "fk_path_kernel_fp64<CoefVarFranke, SUB_CIRCLE>", "kernel_avg_ms": avg_path_ms,
                     "flops_per_path_step": fps},
        "pcie_inclusive_path_steps_per_s": path_steps / float(np.mean(wall)),
        "gmres": {"tol": 1e-10, "iterations": int(info.iterations), "residual": float(info.residual),
                  "converged": bool(info.converged), "seconds_host": t_gmres,
                  "rms_err_vs_exact": float(np.sqrt(np.mean(err ** 2))), "max_err_vs_exact": float(np.abs(err).max())},
    }
    prob.close()
    if with_cpu and not args.no_cpu:
        log("cpu baseline (rows) ...")
        out["cpu_baseline"] = cpu_baseline_rows(mesh, desc, paths, h, seed, budget_s=args.cpu_seconds)
    mesh.close()
    return out


# ------------------------------------------------ config 3 batched rows --
def bench_batched(steps=2, seed=0, fraction=1.0 / 32):
    """Uneven batch on one GPU (spac_amd.workloads.batched_uneven), FP32 stepping; the launch
    shape and refill order come from the autotuner (on the batch with paths x fraction), then the
    assembly runs as one captured graph (plan) per step."""
    from spac_amd.fk import RowsPlan

    mesh, paths, h = batched_uneven()
    n = mesh.info.n_jobs
    live = mesh.job_frontera == 0
    desc = problem_desc(coef=VC, domain=UNIT_SQUARE, precision=abi.SPAC_FP32)
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


# ------------------------------------------- config 5 time to RMS 1e-3 --
# GMRES relative-residual tolerance of the time-to-RMS solves: the reference's default
# toleranciaGMRES (main_test_mpi.cpp:32).
GMRES_TOL = 1e-8
# Coarse grid of the two-level GMRES preconditioner (aggregates per side); 0: Jacobi only.
COARSE_GRID = 12


def _solve_replica(prob, mesh, system, paths, h, seed, stats=False, guess=None):
    """One assembly of every row into the device-resident system and its GPU GMRES solve, from a
    zero start or from guess() (the learned method: its network's u at the nodes, evaluated inside
    the timed solve); returns (u, record).  All calls are synchronous, so the host clock brackets
    the device work."""
    t0 = time.perf_counter()
    st = prob.assemble_system_uniform(system, paths, h, seed=seed, stats=stats)
    t1 = time.perf_counter()
    t = prob.timing()
    u0 = guess() if guess is not None else None
    u, info = system.solve(tol=GMRES_TOL, max_iter=20000, restart=200, u0=u0)
    t2 = time.perf_counter()
    if not info.converged:
        raise RuntimeError(f"GMRES did not converge (seed {seed}): residual {info.residual:.3e}")
    rec = dict(mc_ms=t.total_ms, path_ms=t.path_kernel_ms, assemble_wall_ms=(t1 - t0) * 1e3,
               gmres_ms=(t2 - t1) * 1e3, gmres_device_ms=info.device_ms, gmres_it=int(info.iterations),
               wall_ms=(t2 - t0) * 1e3)
    if st is not None:
        if not (st["status"] == 0).all():
            raise RuntimeError(f"assembly status {int(st['status'].min())} (seed {seed})")
        rec["path_steps"] = int(st["path_steps"].astype(np.int64).sum())
        rec["fallback_rows"] = int(st["cv_fallback"].sum())
    return u, rec


def _rms(U):
    """Statistical RMS of the solved node values: sample variance across replicas, averaged over
    nodes."""
    return float(np.sqrt(np.asarray(U).var(0, ddof=1).mean()))


def _boot_rms(U, n_boot, rng):
    """Bootstrap distribution of _rms over replicas (resampled with replacement), from the
    R x R Gram matrix of the centred replicas: with counts c, sum_i c_i |d_i - m|^2 =
    sum_i c_i K_ii - c^T K c / R."""
    D = np.asarray(U) - np.asarray(U).mean(0)
    R, n = D.shape
    K = D @ D.T
    c = rng.multinomial(R, np.full(R, 1.0 / R), size=n_boot).astype(np.float64)
    ss = c @ np.diag(K) - np.einsum("bi,ij,bj->b", c, K, c) / R
    return np.sqrt(np.maximum(ss, 0.0) / ((R - 1) * n))


def _fit_paths(sets, aim):
    """Path count where RMS reaches `aim` under RMS^2 N = a + b / N, fitted by weighted least
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


def _calibrate(name, prob, mesh, system, h, target, probe_paths, probe_rep, rep, seed0, max_extra=2, guess=None):
    """Path count per row at which this method's own replica set measures RMS <= target.

    A probe set at probe_paths, then two sets, each at the count where the fit over every set
    so far gives 0.95 x target, then further sets (at most max_extra) while the last set's own
    RMS is above the target, each at the larger of that fit's count and 1.05 x the last count.
    The last set (its replicas and timings) is what the bootstrap and the projection use."""
    sets, hist = [], []
    paths, reps, it = probe_paths, probe_rep, 0
    while True:
        us, recs = [], []
        for k in range(reps):
            u, r = _solve_replica(prob, mesh, system, paths, h, seed0 + 1000 * it + k, stats=(k == 0), guess=guess)
            us.append(u)
            recs.append(r)
        U = np.array(us)
        rms = _rms(U)
        sets.append((paths, rms, reps))
        hist.append({"paths_per_row": int(paths), "rms": rms, "replicas": reps})
        log(f"ttr {name}: set {it}: rms {rms:.3e} at {paths} paths per row ({reps} replicas)")
        if it >= 2 and rms <= target:
            break
        if it >= 2 + max_extra:
            break
        if it <= 1:
            paths = _fit_paths(sets, 0.95 * target)[0]
        else:
            # the fit over every set so far (one set's RMS scatters by ~10 % at 48 replicas), and at least 5 % more
            paths = max(_fit_paths(sets, 0.95 * target)[0], int(np.ceil(paths * 1.05)))
        reps, it = rep, it + 1
    return {"paths": int(paths), "U": U, "recs": recs, "rms": rms, "history": hist,
            "reached_target": bool(rms <= target)}


def _ci(x):
    q = np.percentile(x, [2.5, 50.0, 97.5])
    return {"lo": float(q[0]), "median": float(q[1]), "hi": float(q[2])}


def bench_ttr(args):
    """The headline: time to RMS 1e-3 on the full-scale system, learned CV against plain MC."""
    from spac_amd.mlp import MLP
    from spac_amd.problems import franke_u

    target = 1e-3
    mesh, h = full_scale()
    exact = franke_u(mesh.node_x, mesh.node_y)
    system = DeviceSystem(mesh)
    system.set_coarse(COARSE_GRID, COARSE_GRID)
    desc32 = problem_desc(coef=VC, domain=UNIT_SQUARE, precision=abi.SPAC_FP32)
    res = {"mesh": {"nodes": int(mesh.info.n_nodes), "rows": int(mesh.info.n_jobs),
                    "circles": int(mesh.info.n_interfaces), "h": h, "nnz": int(mesh.info.nnz)}}

    # ---- plain Monte Carlo: calibrate, then time
    plain_prob = FKProblem(desc32)
    plain_prob.set_control_variate(abi.SPAC_CV_NONE)
    _solve_replica(plain_prob, mesh, system, 16, h, 1)  # warm-up (buffers, code objects)
    cal_p = _calibrate("plain", plain_prob, mesh, system, h, target, args.ttr_probe_paths, args.ttr_replicas,
                       args.ttr_refine_replicas, 100)
    Np = cal_p["paths"]
    _, first = _solve_replica(plain_prob, mesh, system, Np, h, 49999, stats=True)  # untimed: path-steps
    plain_runs = []
    for k in range(args.plain_steps):
        _, r = _solve_replica(plain_prob, mesh, system, Np, h, 50000 + k)
        r["path_steps"], r["fallback_rows"] = first["path_steps"], first["fallback_rows"]
        plain_runs.append(r)
    log(f"ttr plain: {Np} paths per row, {np.mean([r['wall_ms'] for r in plain_runs]):.1f} ms end to end")

    # ---- learned control variate: calibrate with the trained network (training is deterministic,
    # so every timed step retrains exactly this network), then time the whole method per step
    cfg = {k: v for k, v in LEARNED_CV.items() if k not in ("seed", "first_layer_scale")}
    learned_prob = FKProblem(desc32)
    data_prob = FKProblem(desc32)  # the training data's point solves
    net = MLP(UNIT_SQUARE, seed=LEARNED_CV["seed"], first_layer_scale=LEARNED_CV["first_layer_scale"])
    theta0 = net.params
    train_learned_cv(net, **cfg, eager=args.train_eager, data_problem=data_prob)
    net.attach(learned_prob, fallback=True)

    def guess():
        return net.predict_system(system)

    _solve_replica(learned_prob, mesh, system, 16, h, 1, guess=guess)  # warm-up
    cal_l = _calibrate("learned", learned_prob, mesh, system, h, target, args.ttr_probe_paths, args.ttr_replicas,
                       args.ttr_refine_replicas, 200, guess=guess)
    Nl = cal_l["paths"]

    def learned_step(seed, stats=False):
        t0 = time.perf_counter()
        net.params = theta0  # fresh network: initial weights, Adam state and step counter reset
        setup = train_learned_cv(net, **cfg, eager=args.train_eager, data_problem=data_prob)
        net.attach(learned_prob, fallback=True)
        t1 = time.perf_counter()
        u, r = _solve_replica(learned_prob, mesh, system, Nl, h, seed, stats=stats, guess=guess)
        t2 = time.perf_counter()
        r.update(setup)
        r["setup_wall_ms"] = (t1 - t0) * 1e3
        r["step_wall_ms"] = (t2 - t0) * 1e3
        return u, r

    # warm-up steps (the first also reads the estimator records: path-steps, fallback rows)
    _, first = learned_step(60000, stats=True)
    for i in range(1, args.warmup):
        learned_step(60000 + i)
    learned_runs, u_learned = [], None
    for k in range(args.steps):
        u_learned, r = learned_step(70000 + k)
        learned_runs.append(r)
    for r in learned_runs:
        r["path_steps"], r["fallback_rows"] = first["path_steps"], first["fallback_rows"]
    log(f"ttr learned: {Nl} paths per row, {np.mean([r['step_wall_ms'] for r in learned_runs]):.1f} ms end to end")

    # ---- the learned control variate evaluated per step (SPAC_CV_NETWORK) against its table, same
    # paths: variance ratios from the per-row estimators of one assembly each, and path-kernel times
    def row_variances(prob, seed):
        st = prob.assemble_system_uniform(system, Nl, h, seed=seed, stats=True)
        t = prob.timing()
        vp = float(np.sum(st["phiphi"] - st["phi"] ** 2))
        vr = float(np.sum(st["phiphi_VR"] - st["phi_VR"] ** 2))
        return vp, vr, t.path_kernel_ms

    cvnet = {"paths_per_row": Nl, "seed": 95000}
    with FKProblem(desc32) as net_prob:
        net.attach(net_prob, fallback=True, network=True)
        row_variances(net_prob, 94999)  # warm-up (code objects)
        vp_t, vr_t, ms_t = row_variances(learned_prob, 95000)
        vp_n, vr_n, ms_n = row_variances(net_prob, 95000)
    cvnet.update({"variance_ratio_table": vp_t / vr_t, "variance_ratio_network": vp_n / vr_n,
                  "path_kernel_ms_table": ms_t, "path_kernel_ms_network": ms_n,
                  "note": "plain/control-variate variance summed over every row of one assembly at the learned "
                          "count (same Philox paths); network = the MLP's gradient heads evaluated in FP32 at "
                          "every step instead of read from the 256^2 table"})
    log(f"ttr learned: table vs per-step network: {cvnet}")

    # ---- FP64 plain-MC reference run at the plain count (same seed as an FP32 run)
    with FKProblem(problem_desc(coef=VC, domain=UNIT_SQUARE)) as p64:
        p64.set_control_variate(abi.SPAC_CV_NONE)
        u64, r64 = _solve_replica(p64, mesh, system, Np, h, 90000, stats=True)
    u32, r32 = _solve_replica(plain_prob, mesh, system, Np, h, 90000, stats=True)
    fp = {"paths_per_row": Np, "seed": 90000,
          "plain_fp32_minus_fp64_rms": float(np.sqrt(np.mean((u32 - u64) ** 2))),
          "plain_fp32_minus_fp64_max": float(np.abs(u32 - u64).max()),
          "learned_minus_fp64_plain_rms": float(np.sqrt(np.mean((u_learned - u64) ** 2))),
          "expected_if_independent": float(np.hypot(cal_p["rms"], cal_l["rms"])),
          "fp64_path_steps": r64["path_steps"], "fp32_path_steps": r32["path_steps"],
          "fp64_mc_ms": r64["mc_ms"]}
    log(f"ttr fp32 vs fp64: {fp}")

    # ---- summaries, projections at RMS = target and bootstrap intervals
    def summary(cal, runs, learned):
        R = cal["recs"]
        mc = float(np.mean([r["mc_ms"] for r in runs]))
        gm = float(np.mean([r["gmres_ms"] for r in runs]))
        wall = float(np.mean([r["step_wall_ms" if learned else "wall_ms"] for r in runs]))
        setup = float(np.mean([r["setup_wall_ms"] for r in runs])) if learned else 0.0
        ubar = cal["U"].mean(0)
        s = {"paths_per_row": cal["paths"], "rms_last_set": cal["rms"], "reached_target": cal["reached_target"],
             "replicas_last_set": int(cal["U"].shape[0]), "path_count_history": cal["history"],
             "end_to_end_wall_ms": wall, "mc_device_ms": mc,
             "path_kernel_ms": float(np.mean([r["path_ms"] for r in runs])),
             "assemble_wall_ms": float(np.mean([r["assemble_wall_ms"] for r in runs])),
             "gmres_wall_ms": gm, "gmres_device_ms": float(np.mean([r["gmres_device_ms"] for r in runs])),
             "gmres_iterations": runs[0]["gmres_it"], "path_steps": runs[0]["path_steps"],
             "path_steps_per_s_path_kernel": runs[0]["path_steps"] / (np.mean([r["path_ms"] for r in runs]) * 1e-3),
             "fallback_rows": runs[0]["fallback_rows"],
             "rms_vs_exact": float(np.sqrt(np.mean((ubar - exact) ** 2))),
             "timed_runs": len(runs)}
        if learned:
            s["setup_wall_ms"] = setup
            for k in ("training_data_ms", "train_ms", "train_wall_ms", "table_ms", "train_loss"):
                s[k] = float(np.mean([r[k] for r in runs]))
            s["training_data"] = runs[0]["training_data"]
            s["train"] = {"points": cfg["points_side"] ** 2, "steps": cfg["steps"], "lr": list(cfg["lr"]), "beta2": cfg["beta2"],
                          "grad_weight": cfg["grad_weight"], "table": cfg["table"]}
        # the part of the wall time that scales with paths per row (assembly), and the rest
        s["_scaling_ms"] = s["assemble_wall_ms"]
        s["_fixed_ms"] = wall - s["assemble_wall_ms"]
        return s

    P = summary(cal_p, plain_runs, False)
    L = summary(cal_l, learned_runs, True)
    rng = np.random.default_rng(12345)
    nb = args.bootstrap
    bp = _boot_rms(cal_p["U"], nb, rng)
    bl = _boot_rms(cal_l["U"], nb, rng)

    def at_target(S, rms):
        """wall time projected to RMS = target: the assembly scales with the path count."""
        return S["_fixed_ms"] + S["_scaling_ms"] * (rms / target) ** 2

    def mc_at_target(S, rms):
        return S["mc_device_ms"] * (rms / target) ** 2

    Tp, Tl = at_target(P, P["rms_last_set"]), at_target(L, L["rms_last_set"])
    setup = L["setup_wall_ms"]
    speed = {"end_to_end_wall_at_timed_counts": P["end_to_end_wall_ms"] / L["end_to_end_wall_ms"],
             "mc_only_at_timed_counts": P["mc_device_ms"] / L["mc_device_ms"],
             "end_to_end_wall_at_equal_rms": Tp / Tl,
             "mc_only_at_equal_rms": mc_at_target(P, P["rms_last_set"]) / mc_at_target(L, L["rms_last_set"]),
             "including_training_no_gmres_at_equal_rms":
                 mc_at_target(P, P["rms_last_set"]) / (setup + mc_at_target(L, L["rms_last_set"]))}
    ci = {"end_to_end_wall_at_equal_rms": _ci(at_target(P, bp) / at_target(L, bl)),
          "mc_only_at_equal_rms": _ci(mc_at_target(P, bp) / mc_at_target(L, bl)),
          "including_training_no_gmres_at_equal_rms": _ci(mc_at_target(P, bp) / (setup + mc_at_target(L, bl))),
          "learned_end_to_end_ms_at_target": _ci(at_target(L, bl)),
          "plain_end_to_end_ms_at_target": _ci(at_target(P, bp)),
          "rms_plain_last_set": _ci(bp), "rms_learned_last_set": _ci(bl),
          "bootstrap": f"{nb} resamples of each method's last replica set (replicas drawn with replacement); "
                       "time at RMS = 1e-3 = fixed part + assembly x (RMS / 1e-3)^2"}
    for S in (P, L):
        S["end_to_end_ms_at_target"] = at_target(S, S["rms_last_set"])
        del S["_scaling_ms"], S["_fixed_ms"]
    res.update({"plain": P, "learned": L, "speedup": speed, "speedup_ci95": ci, "fp32_vs_fp64": fp,
                "cv_table_vs_network": cvnet,
                "row_variance_ratio_note": "plain/learned variance at equal paths = "
                                           f"{(P['rms_last_set'] ** 2 * P['paths_per_row']) / (L['rms_last_set'] ** 2 * L['paths_per_row']):.1f}"})

    # roofline of the dominant kernel of the learned step: its path kernel (FP32, table CV)
    fps = FLOPS_PER_STEP["varcoef_circle_fp32_table"]
    kms = L["path_kernel_ms"]
    achieved = L["path_steps"] * fps / (kms * 1e-3) / 1e12
    n_paths = mesh.info.n_jobs * Nl
    roof = {"bound": "mfma", "compute_unit": COMPUTE_UNIT_32, "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": LEARNED_PATH_KERNEL_BYTES_PER_PATH * n_paths,
            "traffic_unit": "B per launch", "algorithmic_bytes": RECORD_BYTES_PER_PATH * n_paths,
            "kernel": "fk_path_kernel_fp32<CoefVarFranke32, SUB_CIRCLE, CV_TABLE>", "kernel_avg_ms": kms,
            "flops_per_path_step": fps, "path_steps_per_launch": L["path_steps"]}
    plain_prob.close()
    learned_prob.close()
    data_prob.close()
    net.close()
    system.close()
    return res, roof, mesh


def bench_assemble(args):
    """Assemblies of the full-scale system at a fixed path count, for kernel profiles of the FP32
    path kernel: the learned control variate trained by the recipe (CV_TABLE with the fallback),
    or plain MC with --plain; --warmup untimed and --steps timed assemblies of every row into the
    device-resident system at --paths-per-row paths and the workload's h."""
    from spac_amd.mlp import MLP

    mesh, h = full_scale()
    system = DeviceSystem(mesh)
    prob = FKProblem(problem_desc(coef=VC, domain=UNIT_SQUARE, precision=abi.SPAC_FP32))
    net = None
    if args.plain:
        prob.set_control_variate(abi.SPAC_CV_NONE)
    else:
        cfg = {k: v for k, v in LEARNED_CV.items() if k not in ("seed", "first_layer_scale")}
        net = MLP(UNIT_SQUARE, seed=LEARNED_CV["seed"], first_layer_scale=LEARNED_CV["first_layer_scale"])
        train_learned_cv(net, **cfg, eager=args.train_eager)
        net.attach(prob, fallback=True)
    N = args.paths_per_row
    st = prob.assemble_system_uniform(system, N, h, seed=1, stats=True)
    if not (st["status"] == 0).all():
        raise RuntimeError(f"assembly status {int(st['status'].min())}")
    path_steps = int(st["path_steps"].astype(np.int64).sum())
    for k in range(args.warmup):
        prob.assemble_system_uniform(system, N, h, seed=2 + k)
    wall, dev, kern = [], [], []
    for k in range(args.steps):
        t0 = time.perf_counter()
        prob.assemble_system_uniform(system, N, h, seed=100 + k)
        wall.append((time.perf_counter() - t0) * 1e3)
        t = prob.timing()
        dev.append(t.total_ms)
        kern.append(t.path_kernel_ms)
    fps = FLOPS_PER_STEP["varcoef_circle_fp32_plain" if args.plain else "varcoef_circle_fp32_table"]
    kms = float(np.mean(kern))
    achieved = path_steps * fps / (kms * 1e-3) / 1e12
    out = {"workload": "full-scale system assembly, FP32 stepping: 64x64 circles, 186,920 rows, "
                       f"{N} paths per row, h={h:.4g}, " + ("plain MC" if args.plain else "learned CV (table)"),
           "paths_per_row": N, "path_steps_per_assembly": path_steps, "assembly_wall_ms": float(np.mean(wall)),
           "assembly_device_ms": float(np.mean(dev)), "path_kernel_ms": kms,
           "path_steps_per_s_path_kernel": path_steps / (kms * 1e-3),
           "roofline": {"bound": "mfma", "compute_unit": COMPUTE_UNIT_32, "achieved": achieved,
                        "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                        "flops_per_path_step": fps},
           "fallback_rows": int(st["cv_fallback"].sum())}
    prob.close()
    if net is not None:
        net.close()
    system.close()
    return out


def cpu_baseline_ttr(mesh, h, plain, budget_s):
    """The CPU restatement (FP64, plain Monte Carlo, OpenMP over paths) on a sample of the
    full-scale rows at the plain GPU count; its time to RMS 1e-3 is projected from the measured
    path-steps/s and the path-steps plain MC needs there (same estimator, same variance)."""
    desc = problem_desc(coef=VC, domain=UNIT_SQUARE)
    c = cpu_baseline_rows(mesh, desc, plain["paths_per_row"], h, 0, budget_s=budget_s, cv=abi.SPAC_CV_NONE)
    steps_needed = plain["path_steps"] * (plain["rms_last_set"] / 1e-3) ** 2
    rate = c["value"]
    c.update({"value": steps_needed / rate * 1e3, "unit": "ms (projected time to RMS 1e-3, plain MC)",
              "path_steps_per_s": rate, "path_steps_to_target": steps_needed})
    return c


def bench_train(steps=1000, lr=(2e-2, 1e-4), beta2=0.95, eager=False):
    """The learned control variate's training alone (for kernel profiles): 4096 jittered grid
    points on the unit square, Franke values plus N(0, 0.01^2) noise as targets, `steps`
    full-batch Adam steps after 20 warm-up steps."""
    from spac_amd.mlp import MLP
    from spac_amd.problems import franke_u
    from spac_amd.workloads import training_points

    xs, ys = training_points()
    rng = np.random.default_rng(0)
    rng.random(xs.size)
    rng.random(xs.size)
    u = franke_u(xs, ys) + 0.01 * rng.standard_normal(xs.size)
    with MLP(UNIT_SQUARE, seed=0) as net:
        net.train(xs, ys, u, steps=20, lr=lr[0], lr_final=lr[1], grad_weight=0.003, eager=eager, beta2=beta2)
        rep = net.train(xs, ys, u, steps=steps, lr=lr[0], lr_final=lr[1], grad_weight=0.003, eager=eager,
                        beta2=beta2)
    return {"samples": int(xs.size), "steps": steps, "eager": bool(eager), "train_ms": rep["train_ms"],
            "us_per_step": rep["train_ms"] * 1e3 / steps, "final_loss": rep["final_loss"]}


def copy_calibration(mib: int, reps: int = 3) -> None:
    """Plain device-to-device copies of known size (`mib` MiB read and written per dispatch):
    the reference dispatches that scale rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes."""
    lib = abi.load()
    a = abi.DeviceArray(mib << 20, lib)
    b = abi.DeviceArray(mib << 20, lib)
    a.zero()
    for _ in range(reps):
        abi.check(lib.spac_debug_copy_device(b.ptr, a.ptr, mib << 20), lib)
    log(f"copy calibration: {reps} copies of {mib} MiB")
    a.free()
    b.free()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed steps (ttr: learned-CV solves end to end)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="ttr", choices=["ttr", "rows", "points", "batched", "train", "assemble"],
                    help="ttr (default: the headline line) | rows | points | batched | train | assemble")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-extras", action="store_true", help="ttr: skip the rows and batched workloads")
    ap.add_argument("--plain-steps", type=int, default=2, help="ttr: timed plain-MC solves at its count")
    ap.add_argument("--ttr-probe-paths", type=int, default=256)
    ap.add_argument("--ttr-replicas", type=int, default=16, help="replicas of the probe set")
    ap.add_argument("--ttr-refine-replicas", type=int, default=48, help="replicas of each later set")
    ap.add_argument("--bootstrap", type=int, default=4000)
    ap.add_argument("--train-steps", type=int, default=LEARNED_CV["steps"])
    ap.add_argument("--paths-per-row", type=int, default=90, help="assemble: paths per row")
    ap.add_argument("--plain", action="store_true", help="assemble: plain MC instead of the learned CV")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run plain device copies of this many MiB (counter calibration for profiles)")
    ap.add_argument("--train-eager", action="store_true",
                    help="launch training steps one by one instead of replaying captured graphs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path shards by node batches "
              "(rows of one circle stay together), one GPU per share", file=sys.stderr)
        sys.exit(2)
    LEARNED_CV["steps"] = args.train_steps
    if args.copy_mib > 0:
        copy_calibration(args.copy_mib)
    if args.workload == "batched":
        print(json.dumps({"batched": bench_batched(steps=args.steps)}), flush=True)
        return
    if args.workload == "assemble":
        print(json.dumps({"assemble": bench_assemble(args)}), flush=True)
        return
    if args.workload == "train":
        print(json.dumps({"train": bench_train(steps=args.train_steps, eager=args.train_eager)}), flush=True)
        return
    if args.workload in ("rows", "points"):
        out = bench_rows(args) if args.workload == "rows" else bench_points(args)
        print(json.dumps(out), flush=True)
        return
    t_start = time.perf_counter()
    res, roof, mesh = bench_ttr(args)
    P, L = res["plain"], res["learned"]
    out = {
        "metric": "time to RMS err 1e-3 (learned-CV vs plain MC, same GPU)",
        "value": L["end_to_end_wall_ms"],
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": L["end_to_end_wall_ms"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "full-scale PDDSparse interface system: unit square, 64x64 circles (overlap 1.2, "
                               "25 nodes per circle), 99,604 nodes / 186,920 rows, sigma=(1+0.5 sin(pi x) cos(pi y)) I, "
                               "b=0.3(y,-x), c=-1-x^2, manufactured Franke u, Gaussian RBF eps=3 on 15-node stencils; "
                               "learned control variate (MLP 2-64-64-64-3 trained on the GPU) + FP32 stepping with "
                               "FP64 sums + two-level FP64 GMRES to 1e-8 (learned: from the network's u at the "
                               "nodes; plain: from zero); one step = one solve end to end",
                   "target_rms": 1e-3, "rms": "statistical RMS of the solved node values over independent replicas",
                   "h": res["mesh"]["h"], "nodes": res["mesh"]["nodes"], "rows": res["mesh"]["rows"],
                   "paths_per_row_learned": L["paths_per_row"], "paths_per_row_plain": P["paths_per_row"],
                   "precision": "FP32 stepping; FP64 exit projection, compensated sums, rows and GMRES"},
        "plain_mc_ms": P["end_to_end_wall_ms"],
        "speedup": res["speedup"],
        "speedup_ci95": res["speedup_ci95"],
        "path_steps_per_s": {"learned_path_kernel": L["path_steps_per_s_path_kernel"],
                             "plain_path_kernel": P["path_steps_per_s_path_kernel"]},
        "rms_vs_exact": {"learned": L["rms_vs_exact"], "plain": P["rms_vs_exact"],
                         "note": "replica-mean solution against the exact Franke u: the bias of the time step and "
                                 "the RBF interpolation, shared by both methods and outside the statistical target"},
        "fp32_vs_fp64": res["fp32_vs_fp64"],
        "cv_table_vs_network": res["cv_table_vs_network"],
        "roofline": roof,
        "time_to_rms": res,
    }
    if not args.no_cpu:
        log("cpu baseline (full-scale rows, plain MC) ...")
        out["cpu_baseline"] = cpu_baseline_ttr(mesh, res["mesh"]["h"], P, args.cpu_seconds)
        out["cpu_baseline"]["gpu_learned_speedup"] = out["cpu_baseline"]["value"] / L["end_to_end_wall_ms"]
    mesh.close()
    if not args.no_extras:
        log("PDDSparse rows FP64 (4x4 circles) ...")
        out["rows_fp64"] = bench_rows(args, steps=5, warmup=1, with_cpu=False)
        log("batched uneven workload (16x16 circles) ...")
        out["batched"] = bench_batched()
    out["bench_seconds"] = time.perf_counter() - t_start
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
